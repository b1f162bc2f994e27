// Sampled accuracy check of a finished GEMM: for `samples` entries (i, j)
// chosen by a counter-based hash of `seed`, recompute
//   C*_ij = beta*C0_ij + alpha*sum_p A_ip B_pj
// in double-double (Dot2 of Ogita, Rump and Oishi, SIAM J. Sci. Comput. 26(6)
// 2005, Algorithm 5.3: TwoProduct by fma, TwoSum by Knuth), one wavefront per
// entry with the k range split over the 64 lanes and the lane partials merged
// by error-free TwoSum. Reports
//   max |C_ij - C*_ij| / (|alpha| (|A||B|)_ij + |beta| |C0_ij|)   (scaled)
//   max |C_ij - C*_ij| / |C*_ij|                                  (relative)
// This is how every result of the engine states its error against an FP64
// oracle; the reference's own check is the residual of src/util.rs:55-85.
#include <math.h>

#include <cstring>

#include "mfma.hpp"
#include "status.hpp"

namespace momms {

// SplitMix64 (Steele, Lea, Flood 2014), the sample index generator.
__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double z = s - a;
  e = (a - (s - z)) + (b - z);
}

__global__ void __launch_bounds__(256)
sampled_check(int64_t m, int64_t n, int64_t k, double alpha, const double* __restrict__ A,
              int64_t rsA, int64_t csA, const double* __restrict__ B, int64_t rsB, int64_t csB,
              double beta, const double* __restrict__ C0, const double* __restrict__ C,
              int64_t rsC, int64_t csC, int64_t samples, uint64_t seed, double* __restrict__ ref,
              unsigned long long* __restrict__ maxes) {
  const int64_t e = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (e >= samples) return;
  const int64_t i = (int64_t)(splitmix64(seed + 2 * (uint64_t)e) % (uint64_t)m);
  const int64_t j = (int64_t)(splitmix64(seed + 2 * (uint64_t)e + 1) % (uint64_t)n);
  double p = 0.0, s = 0.0, mag = 0.0;
  for (int64_t q = lane; q < k; q += 64) {
    const double x = A[i * rsA + q * csA], y = B[q * rsB + j * csB];
    const double h = x * y;
    const double r = __builtin_fma(x, y, -h);
    double t, u;
    two_sum(p, h, t, u);
    p = t;
    s += u + r;
    mag += fabs(h);
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double po = __shfl_xor(p, off), so = __shfl_xor(s, off), mo = __shfl_xor(mag, off);
    double t, u;
    two_sum(p, po, t, u);
    p = t;
    s = s + so + u;
    mag += mo;
  }
  if (lane != 0) return;
  // alpha*(p + s) + beta*c0 in double-double
  const double ph = alpha * p;
  const double pl = __builtin_fma(alpha, p, -ph) + alpha * s;
  double bh = 0.0, bl = 0.0;
  const double c0 = (beta != 0.0) ? C0[i * rsC + j * csC] : 0.0;
  if (beta != 0.0) {
    bh = beta * c0;
    bl = __builtin_fma(beta, c0, -bh);
  }
  double t, te;
  two_sum(ph, bh, t, te);
  const double exact = t + (te + pl + bl);
  const double scale = fabs(alpha) * mag + fabs(beta) * fabs(c0);
  const double c = C[i * rsC + j * csC];
  const double err = fabs(c - exact);
  if (ref) ref[e] = exact;
  const double scaled = scale > 0.0 ? err / scale : (err > 0.0 ? INFINITY : 0.0);
  const double rel = exact != 0.0 ? err / fabs(exact) : (err > 0.0 ? INFINITY : 0.0);
  // non-negative doubles order like their bit patterns; NaN sorts highest
  atomicMax(&maxes[0], (unsigned long long)__double_as_longlong(scaled));
  atomicMax(&maxes[1], (unsigned long long)__double_as_longlong(rel));
}

__global__ void zero2(unsigned long long* p) {
  if (threadIdx.x < 2) p[threadIdx.x] = 0;
}

}  // namespace momms

using namespace momms;

extern "C" int momms_hip_gemm_check_sampled(int64_t m, int64_t n, int64_t k,
                                            const double* alpha, const double* A, int64_t rsA,
                                            int64_t csA, const double* B, int64_t rsB,
                                            int64_t csB, const double* beta, const double* C0,
                                            const double* C, int64_t rsC, int64_t csC,
                                            int64_t samples, uint64_t seed, double* ref_out,
                                            double* max_scaled, double* max_rel, void* stream) {
  static const char* fn = "momms_hip_gemm_check_sampled";
  if (m < 1) return bad_arg(fn, 1, "m < 1");
  if (n < 1) return bad_arg(fn, 2, "n < 1");
  if (k < 0) return bad_arg(fn, 3, "k < 0");
  if (alpha == nullptr) return bad_arg(fn, 4, "alpha is null");
  int st;
  if ((st = check_operand(fn, 5, A, m, k, rsA, csA))) return st;
  if ((st = check_operand(fn, 8, B, k, n, rsB, csB))) return st;
  if (beta == nullptr) return bad_arg(fn, 11, "beta is null");
  if (*beta != 0.0 && C0 == nullptr) return bad_arg(fn, 12, "C0 is null");
  if ((st = check_operand(fn, 13, C, m, n, rsC, csC))) return st;
  if (samples < 1) return bad_arg(fn, 16, "samples < 1");
  if (max_scaled == nullptr) return bad_arg(fn, 19, "max_scaled is null");
  if (max_rel == nullptr) return bad_arg(fn, 20, "max_rel is null");
  hipStream_t s = (hipStream_t)stream;
  unsigned long long* d = nullptr;
  MOMMS_HIP_TRY(fn, hipMallocAsync((void**)&d, 16, s));
  hipLaunchKernelGGL(zero2, dim3(1), dim3(64), 0, s, d);
  hipLaunchKernelGGL(sampled_check, dim3((unsigned)ceil_div(samples, 4)), dim3(256), 0, s, m, n,
                     k, *alpha, A, rsA, csA, B, rsB, csB, *beta, C0, C, rsC, csC, samples, seed,
                     ref_out, d);
  MOMMS_HIP_TRY(fn, hipGetLastError());
  unsigned long long h[2] = {0, 0};
  MOMMS_HIP_TRY(fn, hipMemcpyAsync(h, d, 16, hipMemcpyDeviceToHost, s));
  MOMMS_HIP_TRY(fn, hipFreeAsync(d, s));
  MOMMS_HIP_TRY(fn, hipStreamSynchronize(s));
  std::memcpy(max_scaled, &h[0], 8);
  std::memcpy(max_rel, &h[1], 8);
  return 0;
}

extern "C" uint64_t momms_sample_index(uint64_t seed, uint64_t e, int which, uint64_t extent) {
  return splitmix64(seed + 2 * e + (which ? 1 : 0)) % (extent ? extent : 1);
}
