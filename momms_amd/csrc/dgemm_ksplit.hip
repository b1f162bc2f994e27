// FP64 GEMM with the k range split inside one workgroup: the form for
// problems with too few output tiles to fill the chip (512^3: 16 tiles of
// 128 x 128, 64 of 64 x 64).
//
// The reference spreads a small problem over its threads with ParallelN /
// ParallelM and ThreadComm (src/composables/parallel_range.rs:98-163,
// src/thread_comm.rs:91-107); here one workgroup owns one output tile and
// its four wavefronts each multiply a quarter of the k range into a full
// tile of accumulators, straight from global memory into MFMA fragment
// registers (no LDS, no barrier in the k loop). The four partial tiles then
// meet in LDS and are added in wavefront order, p0 + p1 + p2 + p3, so the
// result is deterministic (the same bits on every run) and needs no second
// kernel or workspace: what stream-K spends on its fix-up launch (about
// 8 us) this form spends on one LDS pass.
//
// Operands: A and B column-major with leading dimensions at most 2^21 (the
// lean tiles' 32-bit offsets, batch_tiles.hpp); C any strides.
#include <hip/hip_runtime.h>

#include "batch_tiles.hpp"
#include "dgemm.hpp"
#include "dgemm_ksplit.hpp"
#include "mfma.hpp"

namespace momms {

// A (16 FM) x (16 FN) tile (row0, col0) over k in [kb, ke), by one
// wavefront, accumulated into acc (lane l holds C[16a + l%16][16b + l/16 +
// 4r] in acc[a][b][r]). 8-deep chunks with WD chunks of loads in flight;
// within a chunk lane group q = l/16 holds k-values k0 + 2q and k0 + 2q + 1,
// so MFMA step s takes {k0 + s, k0 + 2 + s, k0 + 4 + s, k0 + 6 + s}.
template <int FM, int FN, int WD>
__device__ __forceinline__ void ksplit_wave(int64_t m, int64_t n, const double* A, int64_t lda,
                                            const double* B, int64_t ldb, int64_t row0,
                                            int64_t col0, int64_t kb, int64_t ke,
                                            v4f64 (&acc)[FM][FN]) {
  constexpr int KC = 8, BM = 16 * FM, BN = 16 * FN;
  const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
  const int mr = (int)(m - row0 < BM ? m - row0 : BM), nr = (int)(n - col0 < BN ? n - col0 : BN);
  const uint32_t la = (uint32_t)lda, lb = (uint32_t)ldb, lda8 = la * 8u;
  uint32_t va[FM], vb[FN];
#pragma unroll
  for (int a = 0; a < FM; ++a) va[a] = 16 * a + lr < mr ? (16 * a + lr + 2 * lq * la) * 8u : kOob;
#pragma unroll
  for (int b = 0; b < FN; ++b) vb[b] = 16 * b + lr < nr ? (2 * lq + (16 * b + lr) * lb) * 8u : kOob;
  // records end at column ke-1 of A (so k >= ke reads zero) and at the last
  // element of B's block (k >= ke is masked per lane)
  const int64_t a_end = (ke - 1) * lda + m, b_end = (n - 1) * ldb + ke;
  const int64_t nchunks = ke > kb ? ceil_div(ke - kb, KC) : 0;

  auto load = [&](int64_t c, double (&xa)[FM][2], double (&xb)[FN][2]) {
    const int64_t k0 = kb + c * KC;
    const int64_t aoff = row0 + k0 * lda, boff = k0 + col0 * ldb;
    const auto ra = block_rsrc(A + aoff, a_end - aoff);
    const auto rb = block_rsrc(B + boff, b_end - boff);
    const int64_t left = ke - k0;
    const int kr = (int)(left < KC ? (left > 0 ? left : 0) : KC) - 2 * lq;  // k left for this lane group
#pragma unroll
    for (int s = 0; s < 2; ++s) {
#pragma unroll
      for (int a = 0; a < FM; ++a) xa[a][s] = bload(ra, va[a] + s * lda8);
#pragma unroll
      for (int b = 0; b < FN; ++b) xb[b][s] = bload(rb, s < kr ? vb[b] + 8 * s : kOob);
    }
  };
  auto mma = [&](const double (&xa)[FM][2], const double (&xb)[FN][2]) {
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) acc[a][b] = mfma_f64_16x16x4(xb[b][s], xa[a][s], acc[a][b]);
  };
  if (nchunks > 0) {
    double xa[WD][FM][2], xb[WD][FN][2];
#pragma unroll
    for (int d = 0; d < WD; ++d) load(d, xa[d], xb[d]);
    int64_t c = 0;
    // whole groups of WD chunks with no branch inside, so the compiler's
    // count of loads in flight stays exact
#pragma unroll 1
    for (; c + WD <= nchunks; c += WD) {
#pragma unroll
      for (int d = 0; d < WD; ++d) {
        mma(xa[d], xb[d]);
        __builtin_amdgcn_sched_barrier(0);
        load(c + d + WD, xa[d], xb[d]);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
#pragma unroll
    for (int d = 0; d + 1 < WD; ++d)
      if (c + d < nchunks) mma(xa[d], xb[d]);
  }
}

// One workgroup of NW wavefronts per (16 FM) x (16 FN) tile of C; wave w
// takes k-values [w kq, (w+1) kq), kq = k/NW rounded up to whole chunks.
template <int FM, int FN, int WD, int NW, int MINB>
__global__ void __launch_bounds__(64 * NW, MINB)
dgemm_ksplit_kernel(int64_t m, int64_t n, int64_t k, double alpha, const double* __restrict__ A,
                    int64_t lda, const double* __restrict__ B, int64_t ldb, double beta,
                    double* __restrict__ C, int64_t rsC, int64_t csC) {
  constexpr int BM = 16 * FM, BN = 16 * FN;
  __shared__ double part[NW][BN][BM];  // wave w's partial tile, column-major
  const int64_t tiles_m = ceil_div(m, BM);
  const int64_t t = xcd_remap(blockIdx.x, (int64_t)gridDim.x);
  const int64_t row0 = (t % tiles_m) * BM, col0 = (t / tiles_m) * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t kq = ceil_div(ceil_div(k, NW), 8) * 8;
  const int64_t kb = uniform((int64_t)wave * kq);
  const int64_t ke = uniform(kb + kq < k ? kb + kq : k);

  v4f64 acc[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b) acc[a][b] = v4f64{0.0, 0.0, 0.0, 0.0};
  ksplit_wave<FM, FN, WD>(m, n, A, lda, B, ldb, row0, col0, kb, ke, acc);

  const int lr = lane & 15, lq = lane >> 4;
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) part[wave][16 * b + lq + 4 * r][16 * a + lr] = acc[a][b][r];
  __syncthreads();

  // entry e = tid + 64 NW u (row e % BM, column e / BM): consecutive lanes
  // take consecutive rows of one column; the partials are added in wave order
  constexpr int E = BM * BN / (64 * NW);
  const int64_t mr = m - row0, nr = n - col0;
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int e = tid + 64 * NW * u, il = e % BM, jl = e / BM;
    if (il < mr && jl < nr) {
      double s = part[0][jl][il];
#pragma unroll
      for (int w = 1; w < NW; ++w) s += part[w][jl][il];
      const double v = alpha * s;
      double* c = C + (row0 + il) * rsC + (col0 + jl) * csC;
      *c = beta == 0.0 ? v : __builtin_fma(beta, *c, v);
    }
  }
}

bool ksplit_ok(const DgemmArgs& g) {
  return g.rsA == 1 && g.rsB == 1 && g.csA >= (g.m > 1 ? g.m : 1) && g.csB >= (g.k > 1 ? g.k : 1) &&
         g.csA <= kBatchLeanLd && g.csB <= kBatchLeanLd;
}

hipError_t ksplit_launch(const DgemmArgs& g, int form) {
  if (form == 0) {
    const int64_t tiles = ceil_div(g.m, 64) * ceil_div(g.n, 64);
    hipLaunchKernelGGL((dgemm_ksplit_kernel<4, 4, 3, 4, 1>), dim3((unsigned)tiles), dim3(256), 0,
                       g.stream, g.m, g.n, g.k, g.alpha, g.A, g.csA, g.B, g.csB, g.beta, g.C,
                       g.rsC, g.csC);
  } else {
    const int64_t tiles = ceil_div(g.m, 32) * ceil_div(g.n, 32);
    // (eight or sixteen wavefronts, each with a smaller share of k, measured
    // the same or slower: 512^3 19.8 / 18.7 / 16.2 TF)
    hipLaunchKernelGGL((dgemm_ksplit_kernel<2, 2, 4, 4, 2>), dim3((unsigned)tiles), dim3(256), 0,
                       g.stream, g.m, g.n, g.k, g.alpha, g.A, g.csA, g.B, g.csB, g.beta, g.C,
                       g.rsC, g.csC);
  }
  return hipGetLastError();
}

}  // namespace momms
