// The FP64 matrix-core tile: one BM x BN tile of C := beta*C + alpha*A*B,
// used by the one-problem kernel (dgemm.hip) and the persistent batched
// kernel (batched.hip). See dgemm.hip for how it maps the Goto tree.
#pragma once
#include "mfma.hpp"

namespace momms {

// One BM x BN tile of C at (row0, col0): the whole k loop and the epilogue.
// Shared by the one-problem kernel below and the persistent batched kernel.
// Ends after a workgroup barrier, so the LDS slabs may be reused at once.
template <int BM, int BN, int BK, int WAVES_M, int WAVES_N, int LA, int LB>
__device__ __forceinline__ void dgemm_tile(int64_t m, int64_t n, int64_t k, double alpha,
                                           const double* __restrict__ A, int64_t rsA,
                                           int64_t csA, const double* __restrict__ B,
                                           int64_t rsB, int64_t csB, double beta,
                                           double* __restrict__ C, int64_t rsC, int64_t csC,
                                           int64_t row0, int64_t col0) {
  constexpr int NT = 64 * WAVES_M * WAVES_N;
  constexpr int WTM = BM / WAVES_M;  // wave tile rows
  constexpr int WTN = BN / WAVES_N;  // wave tile cols
  constexpr int FM = WTM / 16;       // 16x16 blocks per wave, M direction
  constexpr int FN = WTN / 16;
  constexpr int PAD = 16;            // row pitch = 128 B mod 256 B: conflict-free reads
  constexpr int LDA = BM + PAD;
  constexpr int LDB = BN + PAD;
  constexpr int A_PER = BM * BK / NT;
  constexpr int B_PER = BK * BN / NT;
  static_assert(BM * BK % NT == 0 && BK * BN % NT == 0, "slab must split evenly");
  static_assert(WTM % 16 == 0 && WTN % 16 == 0 && BK % 4 == 0, "tile shape");

  __shared__ double As[2][BK][LDA];
  __shared__ double Bs[2][BK][LDB];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave % WAVES_M;
  const int wn = wave / WAVES_M;

  const int64_t ktiles = ceil_div(k, BK);
  const bool m_full = row0 + BM <= m;
  const bool n_full = col0 + BN <= n;

  double ra[A_PER];
  double rb[B_PER];

  // Global -> registers for slab kt. Element (i, kk) of the A slab and
  // (kk, j) of the B slab; lanes run along the unit-stride dimension.
  auto load_slab = [&](int64_t kt) {
    const int64_t k0 = kt * BK;
    const bool k_full = k0 + BK <= k;
#pragma unroll
    for (int p = 0; p < A_PER; ++p) {
      const int e = tid + NT * p;
      const int i = (LA == kRow) ? e / BK : e % BM;
      const int kk = (LA == kRow) ? e % BK : e / BM;
      const int64_t gi = row0 + i, gk = k0 + kk;
      const int64_t idx = (LA == kCol) ? gi + gk * csA : (LA == kRow) ? gi * rsA + gk
                                                                     : gi * rsA + gk * csA;
      if (m_full && k_full) {
        ra[p] = A[idx];
      } else {
        // out-of-range elements load A[0] (always valid) and are replaced by 0
        const bool ok = gi < m && gk < k;
        const double v = A[ok ? idx : 0];
        ra[p] = ok ? v : 0.0;
      }
    }
#pragma unroll
    for (int p = 0; p < B_PER; ++p) {
      const int e = tid + NT * p;
      const int j = (LB == kRow) ? e % BN : e / BK;
      const int kk = (LB == kRow) ? e / BN : e % BK;
      const int64_t gj = col0 + j, gk = k0 + kk;
      const int64_t idx = (LB == kCol) ? gk + gj * csB : (LB == kRow) ? gk * rsB + gj
                                                                     : gk * rsB + gj * csB;
      if (n_full && k_full) {
        rb[p] = B[idx];
      } else {
        const bool ok = gj < n && gk < k;
        const double v = B[ok ? idx : 0];
        rb[p] = ok ? v : 0.0;
      }
    }
  };
  auto store_slab = [&](int buf) {
#pragma unroll
    for (int p = 0; p < A_PER; ++p) {
      const int e = tid + NT * p;
      const int i = (LA == kRow) ? e / BK : e % BM;
      const int kk = (LA == kRow) ? e % BK : e / BM;
      As[buf][kk][i] = ra[p];
    }
#pragma unroll
    for (int p = 0; p < B_PER; ++p) {
      const int e = tid + NT * p;
      const int j = (LB == kRow) ? e % BN : e / BK;
      const int kk = (LB == kRow) ? e / BN : e % BK;
      Bs[buf][kk][j] = rb[p];
    }
  };

  v4f64 acc[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b) acc[a][b] = v4f64{0.0, 0.0, 0.0, 0.0};

  const int lr = lane & 15;   // row (A) / column (B) inside a 16-block
  const int lk = lane >> 4;   // k inside a 4-deep step
  const int arow = wm * WTM + lr;
  const int bcol = wn * WTN + lr;

  if (ktiles > 0) {
    load_slab(0);
    store_slab(0);
  }
  __syncthreads();

  for (int64_t kt = 0; kt < ktiles; ++kt) {
    const int cur = (int)(kt & 1);
    if (kt + 1 < ktiles) load_slab(kt + 1);
#pragma unroll
    for (int k4 = 0; k4 < BK / 4; ++k4) {
      double af[FM], bf[FN];
#pragma unroll
      for (int a = 0; a < FM; ++a) af[a] = As[cur][k4 * 4 + lk][arow + a * 16];
#pragma unroll
      for (int b = 0; b < FN; ++b) bf[b] = Bs[cur][k4 * 4 + lk][bcol + b * 16];
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) acc[a][b] = mfma_f64_16x16x4(bf[b], af[a], acc[a][b]);
    }
    if (kt + 1 < ktiles) store_slab(cur ^ 1);
    __syncthreads();
  }

  // Epilogue. In the C^T block layout lane l holds C[i][j] with i = l%16 and
  // j = l/16 + 4r (r < 4) of each 16x16 block.
#pragma unroll
  for (int a = 0; a < FM; ++a) {
    const int64_t gi = row0 + wm * WTM + a * 16 + lr;
#pragma unroll
    for (int b = 0; b < FN; ++b) {
      const int64_t gj0 = col0 + wn * WTN + b * 16 + lk;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gj = gj0 + 4 * r;
        if (gi < m && gj < n) {
          double* cp = C + gi * rsC + gj * csC;
          const double s = alpha * acc[a][b][r];
          *cp = (beta == 0.0) ? s : __builtin_fma(beta, *cp, s);
        }
      }
    }
  }
}

}  // namespace momms
