// FP64-accurate GEMM from BF16 slices (Ozaki scheme I) on BF16 matrix cores.
//
// No counterpart in the reference; it is the precision addition the engine
// offers beside the plain FP64 path. Exact arithmetic (restated on the CPU
// for bit-exact comparison):
//   row i of A:  E_i with max_p |a_ip| < 2^E_i;  r = a * 2^-E_i;
//                q_t = rint(r * 2^(8t)), r -= q_t * 2^(-8t)     (t = 1..s)
//   column j of B likewise (exponent F_j).  |q_1| <= 256, |q_t| <= 128:
//   every slice is an integer that bf16 holds exactly.
//   C_ij = beta*C0_ij + alpha * 2^(E_i+F_j) * sum_{d=2}^{s+1} 2^(-8d) S_d,
//   S_d = sum_{t+u=d} sum_p qa_t[i,p] qb_u[p,j], the level-d slice products.
// The k range is cut into chunks of kch(s) values (a multiple of 16). Within
// a chunk every S_d is an integer below 2^24 in magnitude, so the fp32
// accumulation of the matrix cores is exact in any order; after each chunk
// the levels are folded into an fp64 accumulator, d = s+1 down to 2, by
// acc = fma(S_d, 2^(-8d), acc).
//
// Data movement (the Goto tree on the device):
//   slicing writes A and B slices straight into packed panels, the device
//   form of PackA<RowPanelMatrix> / PackB<ColumnPanelMatrix>
//   (src/composables/pack.rs:51-133): for slice t, row-tile tm and 16-deep
//   k-slab ks, one contiguous block [kh][PANEL][8] of bf16, kh = k half.
//   The multiply kernel copies one block per slice per slab HBM -> LDS with
//   16-byte loads (double-buffered), and each wavefront reads 16 bytes per
//   lane conflict-free straight into the MFMA operand layout.
#include <math.h>

#include <cstdlib>

#include "mfma.hpp"
#include "split_bf16.hpp"
#include "split_tile.hpp"
#include "status.hpp"

namespace momms {

int split_chunk(int s) {
  int kch = 16 * (1024 / (16 * (s + 2)));
  return kch > 256 ? 256 : kch;
}

static inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

const SplitConfig kSplitConfig[kSplitConfigs] = {
    {4, 2},  // 0: 128 x 64 tile, 8 wavefronts
    {2, 2},  // 1: 64 x 64 tile, 4 wavefronts, 2-slab ring, two workgroups per CU
    {2, 4},  // 2: 64 x 128 tile, 8 wavefronts
};

int split_config(int64_t m, int64_t n, int64_t /*k*/, int /*s*/) {
  static int forced = [] {
    const char* e = std::getenv("MOMMS_SPLIT_CFG");
    const int v = e ? std::atoi(e) : -1;
    return (v >= 0 && v < kSplitConfigs) ? v : -1;
  }();
  if (forced >= 0) return forced;
  // small problems need more, smaller tiles to fill 256 CUs
  if (ceil_div(m, 128) * ceil_div(n, 64) < 512) return 1;
  return 0;
}

SplitLayout split_layout(int64_t m, int64_t n, int64_t k, int s) {
  SplitLayout L;
  L.s = s;
  L.cfg = split_config(m, n, k, s);
  L.pa = 32 * kSplitConfig[L.cfg].wm;
  L.pb = 32 * kSplitConfig[L.cfg].wn;
  L.mpad = ceil_div(m, L.pa) * L.pa;
  L.npad = ceil_div(n, L.pb) * L.pb;
  L.kpad = ceil_div(k > 0 ? k : 1, kSlab) * kSlab;
  size_t off = 0;
  L.off_aq = off;
  off = align256(off + (size_t)s * L.mpad * L.kpad * 2);
  L.off_bq = off;
  off = align256(off + (size_t)s * L.npad * L.kpad * 2);
  L.off_ea = off;
  off = align256(off + (size_t)L.mpad * 4);
  L.off_eb = off;
  off = align256(off + (size_t)L.npad * 4);
  L.off_maxa = off;
  off = align256(off + (size_t)L.mpad * 8);
  L.off_maxb = off;
  off = align256(off + (size_t)L.npad * 8);
  L.bytes = off;
  return L;
}

// ---------------------------------------------------------------------------
// 1. max |x| per vector, as the bit pattern of a non-negative double (which
//    orders like the value; NaN patterns sort above +Inf).

__global__ void zero_u64(uint64_t* p, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
}

// One thread per vector (coalesced when vectors are adjacent, stride_vec == 1),
// a grid.y split of the length, and an atomic max.
__global__ void absmax_strided(const double* __restrict__ X, int64_t count, int64_t len,
                               int64_t stride_vec, int64_t stride_elem, int64_t per_block,
                               uint64_t* mx) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= count) return;
  const int64_t p0 = (int64_t)blockIdx.y * per_block;
  const int64_t p1 = p0 + per_block < len ? p0 + per_block : len;
  uint64_t best = 0;
  // eight loads in flight per step (indices past p1 read element p1-1 again)
  for (int64_t pb = p0; pb < p1; pb += 8) {
    uint64_t b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int64_t p = pb + u < p1 ? pb + u : p1 - 1;
      b[u] = (uint64_t)__double_as_longlong(X[v * stride_vec + p * stride_elem]) & 0x7fffffffffffffffull;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) best = b[u] > best ? b[u] : best;
  }
  if (p1 > p0) atomicMax((unsigned long long*)&mx[v], (unsigned long long)best);
}

// Elements of a vector are adjacent (stride_elem == 1): one wavefront per vector.
__global__ void absmax_contig(const double* __restrict__ X, int64_t count, int64_t len,
                              int64_t stride_vec, uint64_t* mx) {
  const int64_t v = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (v >= count) return;
  const double* x = X + v * stride_vec;
  uint64_t best = 0;
  // eight loads in flight per step (indices past len read element len-1 again)
  for (int64_t pb = lane; pb < len; pb += 512) {
    uint64_t b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int64_t p = pb + 64 * u < len ? pb + 64 * u : len - 1;
      b[u] = (uint64_t)__double_as_longlong(x[p]) & 0x7fffffffffffffffull;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) best = b[u] > best ? b[u] : best;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t o = __shfl_xor(best, off);
    best = o > best ? o : best;
  }
  if (lane == 0) mx[v] = best;
}

// Exponent E with max < 2^E (frexp convention); 0 for an all-zero vector;
// kNonFinite for a vector holding Inf or NaN.
__global__ void exponents(const uint64_t* __restrict__ mx, int64_t count, int64_t padded,
                          int* __restrict__ e) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= padded) return;
  if (v >= count) {
    e[v] = 0;
    return;
  }
  const uint64_t b = mx[v];
  if (b >= 0x7ff0000000000000ull) {
    e[v] = kNonFinite;
  } else if (b == 0) {
    e[v] = 0;
  } else {
    int E;
    frexp(__longlong_as_double((long long)b), &E);
    e[v] = E;
  }
}

template <int S, int PANEL>
__global__ void __launch_bounds__(256)
slice_pack(const double* __restrict__ X, int64_t count, int64_t len, int64_t sv, int64_t sp,
           const int* __restrict__ ex, int64_t vpad, int64_t lpad, uint16_t* __restrict__ Q) {
  slice_block<S, PANEL>(X, count, len, sv, sp, ex, vpad, lpad, Q, (int64_t)blockIdx.x * 64,
                        (int64_t)blockIdx.y * 64);
}

template <int S, int CFG>
__global__ void __launch_bounds__(64 * (CFG == 0 ? 8 : CFG == 1 ? 4 : 8), CFG == 1 ? 2 : 1)
split_multiply(int64_t m, int64_t n, int64_t kpad, double alpha, double beta,
               double* __restrict__ C, int64_t rsC, int64_t csC,
               const uint16_t* __restrict__ Aq, const uint16_t* __restrict__ Bq,
               const int* __restrict__ EA, const int* __restrict__ EB, int64_t mpad, int64_t npad,
               int slabs_per_chunk) {
  // Ring depth and raster measured at 8192^3: with five or more slices a
  // 2-slab ring beats 3 and 4 slabs (about 15% at s = 7); with fewer, the
  // smaller slabs want 4. Row-by-row raster (group 1) for the 128 x 64 tile.
  constexpr int WM = CFG == 0 ? 4 : 2, WN = CFG == 2 ? 4 : 2;
  constexpr int ST = (CFG == 1 || S >= 5) ? 2 : 4;
  constexpr int GROUP = CFG == 0 ? 1 : 8;
  const int64_t tiles_m = mpad / (32 * WM);
  const int64_t tiles_n = npad / (32 * WN);
  const int64_t t = xcd_remap(blockIdx.x, (int64_t)gridDim.x);
  int64_t tm, tn;
  grouped_tile(t, tiles_m, tiles_n, GROUP, tm, tn);
  split_tile<S, WM, WN, ST>(m, n, kpad, alpha, beta, C, rsC, csC, Aq, Bq, EA, EB, tiles_m,
                            tiles_n, slabs_per_chunk, tm, tn);
}

// ---------------------------------------------------------------------------
// host side

// Vector v, element p at X[v*sv + p*se]. The strided form needs vectors one
// element apart (sv == 1); any other layout with se != 1 is gathered by it too,
// only less coalesced.
static hipError_t absmax(const double* X, int64_t count, int64_t len, int64_t sv, int64_t se,
                         uint64_t* mx, hipStream_t st) {
  if (count == 0) return hipSuccess;
  if (se == 1 && sv != 1) {
    hipLaunchKernelGGL(absmax_contig, dim3((unsigned)ceil_div(count, 4)), dim3(256), 0, st, X,
                       count, len, sv, mx);
  } else {
    hipLaunchKernelGGL(zero_u64, dim3((unsigned)ceil_div(count, 256)), dim3(256), 0, st, mx,
                       count);
    // enough blocks to fill the device: split the length in pieces of >= 64
    const int64_t vblocks = ceil_div(count, 256);
    int64_t ysplit = ceil_div(2048, vblocks);
    if (ysplit > ceil_div(len, 64)) ysplit = ceil_div(len, 64);
    if (ysplit < 1) ysplit = 1;
    const int64_t per = ceil_div(len, ysplit);
    hipLaunchKernelGGL(absmax_strided, dim3((unsigned)vblocks, (unsigned)ysplit), dim3(256), 0,
                       st, X, count, len, sv, se, per, mx);
  }
  return hipGetLastError();
}

hipError_t vector_exponents(const double* X, int64_t count, int64_t len, int64_t sv, int64_t se,
                            int64_t padded, uint64_t* mx, int* e, hipStream_t st) {
  hipError_t err = absmax(X, count, len, sv, se, mx, st);
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(exponents, dim3((unsigned)ceil_div(padded, 256)), dim3(256), 0, st, mx,
                     count, padded, e);
  return hipGetLastError();
}

template <int PANEL>
static hipError_t slice_launch(int s, const double* X, int64_t count, int64_t len, int64_t sv,
                               int64_t sp, const int* ex, int64_t vpad, int64_t lpad,
                               uint16_t* Q, hipStream_t st) {
  dim3 grid((unsigned)ceil_div(vpad, 64), (unsigned)ceil_div(lpad, 64)), block(256);
#define MOMMS_SL(S_) hipLaunchKernelGGL((slice_pack<S_, PANEL>), grid, block, 0, st, X, count, len, sv, sp, ex, vpad, lpad, Q)
  switch (s) {
    case 1: MOMMS_SL(1); break;
    case 2: MOMMS_SL(2); break;
    case 3: MOMMS_SL(3); break;
    case 4: MOMMS_SL(4); break;
    case 5: MOMMS_SL(5); break;
    case 6: MOMMS_SL(6); break;
    case 7: MOMMS_SL(7); break;
    default: MOMMS_SL(8); break;
  }
#undef MOMMS_SL
  return hipGetLastError();
}

hipError_t split_prepare(const SplitLayout& L, int64_t m, int64_t n, int64_t k, const double* A,
                         int64_t rsA, int64_t csA, const double* B, int64_t rsB, int64_t csB,
                         char* ws, hipStream_t st) {
  uint16_t* aq = (uint16_t*)(ws + L.off_aq);
  uint16_t* bq = (uint16_t*)(ws + L.off_bq);
  int* ea = (int*)(ws + L.off_ea);
  int* eb = (int*)(ws + L.off_eb);
  uint64_t* ma = (uint64_t*)(ws + L.off_maxa);
  uint64_t* mb = (uint64_t*)(ws + L.off_maxb);
  hipError_t e;
  // rows of A: vector stride rsA, element stride csA
  if ((e = vector_exponents(A, m, k, rsA, csA, L.mpad, ma, ea, st)) != hipSuccess) return e;
  // columns of B: vector stride csB, element stride rsB
  if ((e = vector_exponents(B, n, k, csB, rsB, L.npad, mb, eb, st)) != hipSuccess) return e;
  auto slice = [&](int panel, const double* X, int64_t cnt, int64_t sv, int64_t sp, const int* ex,
                   int64_t vpad, uint16_t* Q) {
    return panel == 128 ? slice_launch<128>(L.s, X, cnt, k, sv, sp, ex, vpad, L.kpad, Q, st)
                        : slice_launch<64>(L.s, X, cnt, k, sv, sp, ex, vpad, L.kpad, Q, st);
  };
  if ((e = slice(L.pa, A, m, rsA, csA, ea, L.mpad, aq)) != hipSuccess) return e;
  return slice(L.pb, B, n, csB, rsB, eb, L.npad, bq);
}

hipError_t split_multiply_launch(const SplitLayout& L, int64_t m, int64_t n, int64_t k,
                                 double alpha, double beta, double* C, int64_t rsC, int64_t csC,
                                 const char* ws, hipStream_t st) {
  const uint16_t* aq = (const uint16_t*)(ws + L.off_aq);
  const uint16_t* bq = (const uint16_t*)(ws + L.off_bq);
  const int* ea = (const int*)(ws + L.off_ea);
  const int* eb = (const int*)(ws + L.off_eb);
  const int64_t tiles = (L.mpad / L.pa) * (L.npad / L.pb);
  const int spc = split_chunk(L.s) / kSlab;
  const int64_t kp = (k > 0) ? L.kpad : 0;
  const int threads = 64 * kSplitConfig[L.cfg].wm * kSplitConfig[L.cfg].wn;
  dim3 grid((unsigned)tiles), block(threads);
#define MOMMS_MU2(S_, C_) hipLaunchKernelGGL((split_multiply<S_, C_>), grid, block, 0, st, m, n, kp, alpha, beta, C, rsC, csC, aq, bq, ea, eb, L.mpad, L.npad, spc)
#define MOMMS_MU(S_)                                  \
  do {                                                \
    if (L.cfg == 1) MOMMS_MU2(S_, 1);                 \
    else if (L.cfg == 2) MOMMS_MU2(S_, 2);            \
    else MOMMS_MU2(S_, 0);                            \
  } while (0)
  switch (L.s) {
    case 1: MOMMS_MU(1); break;
    case 2: MOMMS_MU(2); break;
    case 3: MOMMS_MU(3); break;
    case 4: MOMMS_MU(4); break;
    case 5: MOMMS_MU(5); break;
    case 6: MOMMS_MU(6); break;
    case 7: MOMMS_MU(7); break;
    default: MOMMS_MU(8); break;
  }
#undef MOMMS_MU2
#undef MOMMS_MU
  return hipGetLastError();
}

}  // namespace momms

// ---------------------------------------------------------------------------
// C ABI

using namespace momms;

extern "C" size_t momms_hip_split_bf16_workspace_size(int64_t m, int64_t n, int64_t k,
                                                      int slices) {
  if (m < 0 || n < 0 || k < 0 || slices < 1 || slices > kMaxSlices) return 0;
  return split_layout(m, n, k, slices).bytes;
}

extern "C" int momms_hip_split_bf16_prepare(int64_t m, int64_t n, int64_t k, const double* A,
                                            int64_t rsA, int64_t csA, const double* B,
                                            int64_t rsB, int64_t csB, int slices,
                                            void* workspace, size_t workspace_bytes,
                                            void* stream) {
  static const char* fn = "momms_hip_split_bf16_prepare";
  // argument numbering of this entry point: m n k A rsA csA B rsB csB slices ws bytes stream
  if (m < 0) return bad_arg(fn, 1, "m < 0");
  if (n < 0) return bad_arg(fn, 2, "n < 0");
  if (k < 0) return bad_arg(fn, 3, "k < 0");
  int st;
  if ((st = check_operand(fn, 4, A, m, k, rsA, csA))) return st;
  if ((st = check_operand(fn, 7, B, k, n, rsB, csB))) return st;
  if (slices < 1 || slices > kMaxSlices) return bad_arg(fn, 10, "slices must be 1..8");
  const SplitLayout L = split_layout(m, n, k, slices);
  if (workspace == nullptr) return bad_arg(fn, 11, "workspace is null");
  if (workspace_bytes < L.bytes) return bad_arg(fn, 12, "workspace too small");
  if (m == 0 && n == 0) return 0;
  return hip_status(fn, split_prepare(L, m, n, k, A, rsA, csA, B, rsB, csB, (char*)workspace,
                                      (hipStream_t)stream));
}

extern "C" int momms_hip_split_bf16_multiply(int64_t m, int64_t n, int64_t k,
                                             const double* alpha, const double* beta, double* C,
                                             int64_t rsC, int64_t csC, int slices,
                                             const void* workspace, size_t workspace_bytes,
                                             void* stream) {
  static const char* fn = "momms_hip_split_bf16_multiply";
  if (m < 0) return bad_arg(fn, 1, "m < 0");
  if (n < 0) return bad_arg(fn, 2, "n < 0");
  if (k < 0) return bad_arg(fn, 3, "k < 0");
  if (alpha == nullptr) return bad_arg(fn, 4, "alpha is null");
  if (beta == nullptr) return bad_arg(fn, 5, "beta is null");
  int st;
  if ((st = check_operand(fn, 6, C, m, n, rsC, csC))) return st;
  if (slices < 1 || slices > kMaxSlices) return bad_arg(fn, 9, "slices must be 1..8");
  const SplitLayout L = split_layout(m, n, k, slices);
  if (workspace == nullptr) return bad_arg(fn, 10, "workspace is null");
  if (workspace_bytes < L.bytes) return bad_arg(fn, 11, "workspace too small");
  if (m == 0 || n == 0) return 0;
  return hip_status(fn, split_multiply_launch(L, m, n, k, *alpha, *beta, C, rsC, csC,
                                              (const char*)workspace, (hipStream_t)stream));
}

extern "C" int momms_hip_dgemm_split_bf16(int64_t m, int64_t n, int64_t k, const double* alpha,
                                          const double* A, int64_t rsA, int64_t csA,
                                          const double* B, int64_t rsB, int64_t csB,
                                          const double* beta, double* C, int64_t rsC,
                                          int64_t csC, int slices, void* workspace,
                                          size_t workspace_bytes, void* stream) {
  static const char* fn = "momms_hip_dgemm_split_bf16";
  if (m < 0) return bad_arg(fn, 1, "m < 0");
  if (n < 0) return bad_arg(fn, 2, "n < 0");
  if (k < 0) return bad_arg(fn, 3, "k < 0");
  if (alpha == nullptr) return bad_arg(fn, 4, "alpha is null");
  int st;
  if ((st = check_operand(fn, 5, A, m, k, rsA, csA))) return st;
  if ((st = check_operand(fn, 8, B, k, n, rsB, csB))) return st;
  if (beta == nullptr) return bad_arg(fn, 11, "beta is null");
  if ((st = check_operand(fn, 12, C, m, n, rsC, csC))) return st;
  if (slices < 1 || slices > kMaxSlices) return bad_arg(fn, 15, "slices must be 1..8");
  const SplitLayout L = split_layout(m, n, k, slices);
  if (workspace == nullptr) return bad_arg(fn, 16, "workspace is null");
  if (workspace_bytes < L.bytes) return bad_arg(fn, 17, "workspace too small");
  if (m == 0 || n == 0) return 0;
  hipError_t e = split_prepare(L, m, n, k, A, rsA, csA, B, rsB, csB, (char*)workspace,
                               (hipStream_t)stream);
  if (e != hipSuccess) return hip_status(fn, e);
  return hip_status(fn, split_multiply_launch(L, m, n, k, *alpha, *beta, C, rsC, csC,
                                              (const char*)workspace, (hipStream_t)stream));
}
