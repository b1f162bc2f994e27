// Lean FP64 tiles of the batched kernels: buffer loads and stores with 32-bit
// offsets instead of 64-bit addresses, so a workgroup needs few registers
// and four of them fit on a CU (the tiles of small problems are short and
// bound by memory latency, which only more resident waves hide).
//
// Every problem is a column-major block of A (m x k, lda), B (k x n, ldb) and
// C (m x n, ldc). A tile reads through a buffer resource whose base is the
// tile's corner in the current k-slab and whose record count is what is
// left of the problem's block from there, so:
//   - an element past column k-1 of A lies past the records (lda >= m) and
//     reads as 0, as does everything once the slab is past k;
//   - rows of A at or past m, columns of B or C at or past n, and k-values of
//     B at or past k get the offset kOob, past every record count: they read
//     as 0 and are never written.
// So no loaded value needs a mask, nothing reads outside the problem, and a
// problem whose k is 0 reads nothing. Offsets stay below 2^30 when lda, ldb
// and ldc are at most kBatchLeanLd (the plan sends other problems to
// dgemm_tile_deep).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dgemm_tile.hpp"
#include "mfma.hpp"

namespace momms {

typedef unsigned int v2u32 __attribute__((ext_vector_type(2)));

constexpr uint32_t kOob = 0x80000000u;          // past every record count used here
constexpr int64_t kBatchLeanLd = int64_t(1) << 21;

// A value equal on all lanes, moved to scalar registers.
__device__ __forceinline__ int64_t uniform(int64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Buffer over `left` doubles from `base` (none if left <= 0).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t block_rsrc(const double* base, int64_t left) {
  const int64_t bytes = left > 0 ? left * 8 : 0;
  return make_rsrc(base, (uint32_t)(bytes < 0x7FFFFFFF ? bytes : 0x7FFFFFFF));
}
__device__ __forceinline__ double bload(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, uint32_t off, double v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u32, v), r, (int)off, 0, 0);
}

// C := beta*C + v(t) for the N outputs t of one lane, at byte offsets off(t)
// of buffer rc (kOob: not in the problem). Every read of C is issued before
// the first write; offsets and values are recomputed at the write rather than
// kept, which keeps the registers live across the reads to the N values read.
template <int N, typename Off, typename Val>
__device__ __forceinline__ void lean_update_c(__amdgpu_buffer_rsrc_t rc, Off&& off, Val&& v,
                                              double beta) {
  if (beta == 0.0) {
#pragma unroll
    for (int t = 0; t < N; ++t) bstore(rc, off(t), v(t));
  } else {
    double c[N];
#pragma unroll
    for (int t = 0; t < N; ++t) c[t] = bload(rc, off(t));
#pragma unroll
    for (int t = 0; t < N; ++t) bstore(rc, off(t), __builtin_fma(beta, c[t], v(t)));
  }
}

// BM x BN tile (row0, col0) of one problem by a workgroup of four waves
// (2 x 2 of BM/2 x BN/2), 16-deep slabs through LDS with D slabs of loads in
// flight. All arguments are uniform over the workgroup. Each element sums its
// k-values in order, four per MFMA step.
template <int BM, int BN, int D>
__device__ __forceinline__ void batch_tile(int64_t m, int64_t n, int64_t k, double alpha,
                                           const double* A, int64_t lda, const double* B,
                                           int64_t ldb, double beta, double* C, int64_t ldc,
                                           int64_t row0, int64_t col0) {
  constexpr int BK = 16, NT = 256, WTM = BM / 2, WTN = BN / 2, FM = WTM / 16, FN = WTN / 16;
  constexpr int A_PER = BM * BK / NT, B_PER = BK * BN / NT, AK = NT / BM, BJ = NT / BK;
  static_assert(NT % BM == 0 && A_PER * AK == BK && B_PER * BJ == BN, "slab split");
  // A slab: k-rows of BM doubles; B slab: n-rows of BK doubles (k contiguous
  // as in memory) with a stride of BK + 2, so the 16 rows a fragment read or
  // a store touches fall on distinct bank pairs
  constexpr int LDA = BM + 16, LDBK = BK + 2;
  __shared__ double As[2][BK][LDA];
  __shared__ double Bs[2][BN][LDBK];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % 2, wn = wave / 2;
  const int ai = tid % BM, ak = tid / BM;
  const int bk = tid % BK, bj = tid / BK;
  // lane arithmetic in 32 bits: rows / columns of the tile inside the
  // problem, leading dimensions <= kBatchLeanLd
  const int mr = (int)(m - row0 < BM ? m - row0 : BM), nr = (int)(n - col0 < BN ? n - col0 : BN);
  const uint32_t la = (uint32_t)lda, lb = (uint32_t)ldb, lc = (uint32_t)ldc;
  const uint32_t va = ai < mr ? (ai + ak * la) * 8u : kOob;
  uint32_t vb[B_PER];
#pragma unroll
  for (int q = 0; q < B_PER; ++q) {
    const int jl = bj + BJ * q;
    vb[q] = jl < nr ? (bk + jl * lb) * 8u : kOob;
  }
  const uint32_t astep = AK * la * 8u;
  const int64_t a_end = (k - 1) * lda + m, b_end = (n - 1) * ldb + k;
  const int64_t ktiles = ceil_div(k, BK);

  auto load_slab = [&](int64_t kt, double (&xa)[A_PER], double (&xb)[B_PER]) {
    const int64_t k0 = kt * BK;
    const int64_t aoff = row0 + k0 * lda, boff = k0 + col0 * ldb;
    const auto ra = block_rsrc(A + aoff, a_end - aoff);
    const auto rb = block_rsrc(B + boff, b_end - boff);
#pragma unroll
    for (int q = 0; q < A_PER; ++q) xa[q] = bload(ra, va + q * astep);
    const bool kok = bk < (int)(k - k0 < BK ? k - k0 : BK);
#pragma unroll
    for (int q = 0; q < B_PER; ++q) xb[q] = bload(rb, kok ? vb[q] : kOob);
  };

  v4f64 acc[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b) acc[a][b] = v4f64{0.0, 0.0, 0.0, 0.0};
  const int lr = lane & 15, lk = lane >> 4;
  const int arow = wm * WTM + lr, bcol = wn * WTN + lr;

  auto step = [&](int64_t kt, double (&xa)[A_PER], double (&xb)[B_PER], bool reload) {
    const int buf = (int)(kt & 1);
#pragma unroll
    for (int q = 0; q < A_PER; ++q) As[buf][ak + AK * q][ai] = xa[q];
#pragma unroll
    for (int q = 0; q < B_PER; ++q) Bs[buf][bj + BJ * q][bk] = xb[q];
    __syncthreads();  // slab kt visible; everyone is done with slab kt-1
    if (reload) load_slab(kt + D, xa, xb);
    // keep the multiplies below the reload, so the wait before the next set
    // is stored leaves this set's loads in flight
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k4 = 0; k4 < BK / 4; ++k4) {
      double af[FM], bf[FN];
#pragma unroll
      for (int a = 0; a < FM; ++a) af[a] = As[buf][k4 * 4 + lk][arow + a * 16];
#pragma unroll
      for (int b = 0; b < FN; ++b) bf[b] = Bs[buf][bcol + b * 16][k4 * 4 + lk];
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) acc[a][b] = mfma_f64_16x16x4(bf[b], af[a], acc[a][b]);
    }
  };

  if (ktiles > 0) {
    double ra[D][A_PER], rb[D][B_PER];
#pragma unroll
    for (int d = 0; d < D; ++d) load_slab(d, ra[d], rb[d]);
    // whole groups of D slabs with no branch inside, so the compiler's
    // accounting of loads in flight stays exact across the loop
    int64_t kt = 0;
#pragma unroll 1
    for (; kt + D <= ktiles; kt += D) {
#pragma unroll
      for (int d = 0; d < D; ++d) step(kt + d, ra[d], rb[d], true);
    }
#pragma unroll
    for (int d = 0; d + 1 < D; ++d)
      if (kt + d < ktiles) step(kt + d, ra[d], rb[d], false);
  }
  __syncthreads();  // LDS free for the next tile of a persistent caller

  // lane l holds C[i][j], i = l%16, j = l/16 + 4r of each 16x16 block
  const int64_t coff = row0 + col0 * ldc;
  const auto rc = block_rsrc(C + coff, (n - 1) * ldc + m - coff);
  // outputs t = (a*FN + b)*4 + r in groups of G (fewer for the wider tiles,
  // whose accumulators leave fewer registers for values in flight)
  constexpr int NO = FM * FN * 4, G = NO <= 16 ? NO : 8;
#pragma unroll
  for (int t0 = 0; t0 < NO; t0 += G) {
    lean_update_c<G>(
        rc,
        [&](int u) {
          const int t = t0 + u, a = t / (4 * FN), b = (t / 4) % FN, r = t % 4;
          const int il = wm * WTM + a * 16 + lr, jl = wn * WTN + b * 16 + lk + 4 * r;
          return (il < mr && jl < nr) ? (il + jl * lc) * 8u : kOob;
        },
        [&](int u) {
          const int t = t0 + u;
          return alpha * acc[t / (4 * FN)][(t / 4) % FN][t % 4];
        },
        beta);
  }
}

// 32 x 32 tile (row0, col0) of one problem by a single wavefront, no LDS and
// no barrier; 8-deep chunks with WD chunks of loads in flight. All arguments
// are uniform over the wavefront. k is permuted inside a chunk: lane group
// q = l/16 holds k-values k0 + 2q and k0 + 2q + 1, so MFMA step s takes
// {k0 + s, k0 + 2 + s, k0 + 4 + s, k0 + 6 + s}.
template <int WD>
__device__ __forceinline__ void batch_wave_tile32(int64_t m, int64_t n, int64_t k, double alpha,
                                                  const double* A, int64_t lda, const double* B,
                                                  int64_t ldb, double beta, double* C,
                                                  int64_t ldc, int64_t row0, int64_t col0) {
  constexpr int KC = 8;
  const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
  // lane arithmetic in 32 bits (see batch_tile)
  const int mr = (int)(m - row0 < 32 ? m - row0 : 32), nr = (int)(n - col0 < 32 ? n - col0 : 32);
  const uint32_t la = (uint32_t)lda, lb = (uint32_t)ldb, lc = (uint32_t)ldc;
  const uint32_t lda8 = la * 8u;
  const uint32_t va0 = lr < mr ? (lr + 2 * lq * la) * 8u : kOob;
  const uint32_t va1 = 16 + lr < mr ? (16 + lr + 2 * lq * la) * 8u : kOob;
  const uint32_t vb0 = lr < nr ? (2 * lq + lr * lb) * 8u : kOob;
  const uint32_t vb1 = 16 + lr < nr ? (2 * lq + (16 + lr) * lb) * 8u : kOob;
  const int64_t a_end = (k - 1) * lda + m, b_end = (n - 1) * ldb + k;
  const int64_t nchunks = ceil_div(k, KC);

  auto load = [&](int64_t c, double (&x)[8]) {
    const int64_t k0 = c * KC;
    const int64_t aoff = row0 + k0 * lda, boff = k0 + col0 * ldb;
    const auto ra = block_rsrc(A + aoff, a_end - aoff);
    const auto rb = block_rsrc(B + boff, b_end - boff);
    const int kr = (int)(k - k0 < KC ? k - k0 : KC) - 2 * lq;  // k-values left for this lane group
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      x[s] = bload(ra, va0 + s * lda8);
      x[2 + s] = bload(ra, va1 + s * lda8);
      x[4 + s] = bload(rb, s < kr ? vb0 + 8 * s : kOob);
      x[6 + s] = bload(rb, s < kr ? vb1 + 8 * s : kOob);
    }
  };
  v4f64 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = v4f64{0.0, 0.0, 0.0, 0.0};
  auto mma = [&](const double (&x)[8]) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      acc[0][0] = mfma_f64_16x16x4(x[4 + s], x[s], acc[0][0]);
      acc[0][1] = mfma_f64_16x16x4(x[6 + s], x[s], acc[0][1]);
      acc[1][0] = mfma_f64_16x16x4(x[4 + s], x[2 + s], acc[1][0]);
      acc[1][1] = mfma_f64_16x16x4(x[6 + s], x[2 + s], acc[1][1]);
    }
  };
  if (nchunks > 0) {
    double x[WD][8];
#pragma unroll
    for (int d = 0; d < WD; ++d) load(d, x[d]);
    int64_t c = 0;
#pragma unroll 1
    for (; c + WD <= nchunks; c += WD) {
#pragma unroll
      for (int d = 0; d < WD; ++d) {
        mma(x[d]);
        __builtin_amdgcn_sched_barrier(0);
        load(c + d + WD, x[d]);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
#pragma unroll
    for (int d = 0; d + 1 < WD; ++d)
      if (c + d < nchunks) mma(x[d]);
  }
  // lane l holds C[i][j], i = l%16, j = l/16 + 4r of each 16x16 block
  const int64_t coff = row0 + col0 * ldc;
  const auto rc = block_rsrc(C + coff, (n - 1) * ldc + m - coff);
  // output t = (a*2 + b)*4 + r
  lean_update_c<16>(
      rc,
      [&](int t) {
        const int a = t / 8, b = (t / 4) % 2, r = t % 4;
        const int il = 16 * a + lr, jl = 16 * b + lq + 4 * r;
        return (il < mr && jl < nr) ? (il + jl * lc) * 8u : kOob;
      },
      [&](int t) { return alpha * acc[t / 8][(t / 4) % 2][t % 4]; }, beta);
}

}  // namespace momms
