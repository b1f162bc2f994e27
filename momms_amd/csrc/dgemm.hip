// Plain FP64 GEMM on FP64 matrix cores: C := beta*C + alpha*A*B.
//
// This kernel is the device form of the reference's Goto control tree
// (src/exper/l3_packing.rs:58-66):
//   PartN<Nc> / PartM<Mc>  -> the grid: one workgroup per BM x BN tile of C
//   PartK<Kc>              -> the k-loop over BK-deep slabs
//   PackB / PackA          -> global -> LDS staging of the A and B slabs into
//                             k-major panels, double-buffered so the next slab
//                             loads while the current one is multiplied
//   ParallelN + KernelNM   -> 4 wavefronts, each owning a WTM x WTN sub-tile
//   bli_dgemm_*_6x8        -> v_mfma_f64_16x16x4_f64 on 16x16 register blocks
// The reference's ThreadComm barriers (src/thread_comm.rs:50-66) become the
// workgroup barrier between slabs.
//
// The MFMA is issued with the operands swapped (B slab as the "A" operand),
// so it computes the 16x16 block of C^T. In that form lane l holds 4
// consecutive columns of one row, and the 16 lanes of a row group cover 16
// consecutive rows: stores to column-major C are 128-byte coalesced.
#include "dgemm.hpp"
#include "mfma.hpp"
#include "status.hpp"

namespace momms {

template <int BM, int BN, int BK, int WAVES_M, int WAVES_N, int LA, int LB>
__global__ void __launch_bounds__(64 * WAVES_M * WAVES_N)
dgemm_mfma_kernel(int64_t m, int64_t n, int64_t k, double alpha,
                  const double* __restrict__ A, int64_t rsA, int64_t csA,
                  const double* __restrict__ B, int64_t rsB, int64_t csB,
                  double beta, double* __restrict__ C, int64_t rsC, int64_t csC,
                  int64_t group_m) {
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

  const int64_t tiles_m = ceil_div(m, BM);
  const int64_t tiles_n = ceil_div(n, BN);
  const int64_t t = xcd_remap(blockIdx.x, (int64_t)gridDim.x);
  int64_t tm, tn;
  grouped_tile(t, tiles_m, tiles_n, group_m, tm, tn);
  const int64_t row0 = tm * BM;
  const int64_t col0 = tn * BN;

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
      double v = 0.0;
      if ((m_full && k_full) || (gi < m && gk < k)) {
        if (LA == kCol) v = A[gi + gk * csA];
        else if (LA == kRow) v = A[gi * rsA + gk];
        else v = A[gi * rsA + gk * csA];
      }
      ra[p] = v;
    }
#pragma unroll
    for (int p = 0; p < B_PER; ++p) {
      const int e = tid + NT * p;
      const int j = (LB == kRow) ? e % BN : e / BK;
      const int kk = (LB == kRow) ? e / BN : e % BK;
      const int64_t gj = col0 + j, gk = k0 + kk;
      double v = 0.0;
      if ((n_full && k_full) || (gj < n && gk < k)) {
        if (LB == kCol) v = B[gk + gj * csB];
        else if (LB == kRow) v = B[gk * rsB + gj];
        else v = B[gk * rsB + gj * csB];
      }
      rb[p] = v;
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

  // Epilogue. Lane holds C[i][j0+r] for r<4 of each block (C^T block layout).
#pragma unroll
  for (int a = 0; a < FM; ++a) {
    const int64_t gi = row0 + wm * WTM + a * 16 + lr;
#pragma unroll
    for (int b = 0; b < FN; ++b) {
      const int64_t gj0 = col0 + wn * WTN + b * 16 + 4 * lk;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gj = gj0 + r;
        if (gi < m && gj < n) {
          double* cp = C + gi * rsC + gj * csC;
          const double s = alpha * acc[a][b][r];
          *cp = (beta == 0.0) ? s : __builtin_fma(beta, *cp, s);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Host side: tile selection and dispatch over operand storage orders.

template <int BM, int BN, int BK, int WM, int WN>
static hipError_t launch_cfg(const DgemmArgs& g, int la, int lb, int64_t group_m) {
  const int64_t tiles = ceil_div(g.m, BM) * ceil_div(g.n, BN);
  dim3 grid((unsigned)tiles), block(64 * WM * WN);
#define MOMMS_DG(LA_, LB_)                                                                  \
  hipLaunchKernelGGL((dgemm_mfma_kernel<BM, BN, BK, WM, WN, LA_, LB_>), grid, block, 0,    \
                     g.stream, g.m, g.n, g.k, g.alpha, g.A, g.rsA, g.csA, g.B, g.rsB, g.csB, \
                     g.beta, g.C, g.rsC, g.csC, group_m)
  switch (la * 3 + lb) {
    case 0: MOMMS_DG(kCol, kCol); break;
    case 1: MOMMS_DG(kCol, kRow); break;
    case 2: MOMMS_DG(kCol, kGen); break;
    case 3: MOMMS_DG(kRow, kCol); break;
    case 4: MOMMS_DG(kRow, kRow); break;
    case 5: MOMMS_DG(kRow, kGen); break;
    case 6: MOMMS_DG(kGen, kCol); break;
    case 7: MOMMS_DG(kGen, kRow); break;
    default: MOMMS_DG(kGen, kGen); break;
  }
#undef MOMMS_DG
  return hipGetLastError();
}

static int layout_of(int64_t rs, int64_t cs, int64_t h, int64_t w) {
  if (rs == 1 || w == 1) return kCol;
  if (cs == 1 || h == 1) return kRow;
  return kGen;
}

hipError_t dgemm_launch(const DgemmArgs& g, int tile_cfg) {
  // A single column (or row) operand is unit-stride along its only index, so
  // it is treated as column (row) major whatever its other stride.
  const int la = layout_of(g.rsA, g.csA, g.m, g.k);
  const int lb = layout_of(g.rsB, g.csB, g.k, g.n);
  if (tile_cfg < 0) tile_cfg = dgemm_pick_tile(g.m, g.n, g.k);
  switch (tile_cfg) {
    case 1: return launch_cfg<64, 64, 16, 2, 2>(g, la, lb, 8);
    case 2: return launch_cfg<128, 64, 16, 2, 2>(g, la, lb, 8);
    default: return launch_cfg<128, 128, 16, 2, 2>(g, la, lb, 8);
  }
}

int dgemm_pick_tile(int64_t m, int64_t n, int64_t /*k*/) {
  // Enough workgroups to cover the 256 CUs at least once with the big tile;
  // otherwise fall back to smaller tiles.
  const int64_t big = ceil_div(m, 128) * ceil_div(n, 128);
  if (big >= 256) return 0;
  const int64_t mid = ceil_div(m, 128) * ceil_div(n, 64);
  if (mid >= 256) return 2;
  return 1;
}

}  // namespace momms

extern "C" int momms_hip_dgemm(int64_t m, int64_t n, int64_t k, const double* alpha,
                               const double* A, int64_t rsA, int64_t csA, const double* B,
                               int64_t rsB, int64_t csB, const double* beta, double* C,
                               int64_t rsC, int64_t csC, void* stream) {
  using namespace momms;
  static const char* fn = "momms_hip_dgemm";
  if (m < 0) return bad_arg(fn, 1, "m < 0");
  if (n < 0) return bad_arg(fn, 2, "n < 0");
  if (k < 0) return bad_arg(fn, 3, "k < 0");
  if (alpha == nullptr) return bad_arg(fn, 4, "alpha is null");
  int st;
  if ((st = check_operand(fn, 5, A, m, k, rsA, csA))) return st;
  if ((st = check_operand(fn, 8, B, k, n, rsB, csB))) return st;
  if (beta == nullptr) return bad_arg(fn, 11, "beta is null");
  if ((st = check_operand(fn, 12, C, m, n, rsC, csC))) return st;
  if (m == 0 || n == 0) return 0;
  DgemmArgs g{m, n, k, *alpha, A, rsA, csA, B, rsB, csB, *beta, C, rsC, csC,
              (hipStream_t)stream};
  return hip_status(fn, dgemm_launch(g, -1));
}

extern "C" int momms_hip_dgemm_tiled(int64_t m, int64_t n, int64_t k, const double* alpha,
                                     const double* A, int64_t rsA, int64_t csA,
                                     const double* B, int64_t rsB, int64_t csB,
                                     const double* beta, double* C, int64_t rsC, int64_t csC,
                                     int tile_cfg, void* stream) {
  using namespace momms;
  static const char* fn = "momms_hip_dgemm_tiled";
  if (tile_cfg < 0 || tile_cfg >= kDgemmTileConfigs) return bad_arg(fn, 15, "tile_cfg");
  if (m < 0) return bad_arg(fn, 1, "m < 0");
  if (n < 0) return bad_arg(fn, 2, "n < 0");
  if (k < 0) return bad_arg(fn, 3, "k < 0");
  if (alpha == nullptr) return bad_arg(fn, 4, "alpha is null");
  int st;
  if ((st = check_operand(fn, 5, A, m, k, rsA, csA))) return st;
  if ((st = check_operand(fn, 8, B, k, n, rsB, csB))) return st;
  if (beta == nullptr) return bad_arg(fn, 11, "beta is null");
  if ((st = check_operand(fn, 12, C, m, n, rsC, csC))) return st;
  if (m == 0 || n == 0) return 0;
  DgemmArgs g{m, n, k, *alpha, A, rsA, csA, B, rsB, csB, *beta, C, rsC, csC,
              (hipStream_t)stream};
  return hip_status(fn, dgemm_launch(g, tile_cfg));
}
