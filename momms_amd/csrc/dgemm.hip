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
// so it computes the 16x16 block of C^T. In that form lane l holds row l%16
// of the block in 4 columns l/16 + 4r, and each group of 16 lanes covers 16
// consecutive rows of one column: stores to column-major C are 128-byte
// coalesced.
#include "dgemm.hpp"
#include "dgemm_ksplit.hpp"
#include "dgemm_streamk.hpp"
#include "dgemm_tile.hpp"
#include "mfma.hpp"
#include "scratch.hpp"
#include "status.hpp"
#include "tile_order.hpp"
#include "momms_hip.h"

#include <mutex>

namespace momms {

template <int BM, int BN, int BK, int WAVES_M, int WAVES_N, int LA, int LB>
__global__ void __launch_bounds__(64 * WAVES_M * WAVES_N)
dgemm_mfma_kernel(int64_t m, int64_t n, int64_t k, double alpha,
                  const double* __restrict__ A, int64_t rsA, int64_t csA,
                  const double* __restrict__ B, int64_t rsB, int64_t csB,
                  double beta, double* __restrict__ C, int64_t rsC, int64_t csC,
                  int64_t group_m) {
  const int64_t tiles_m = ceil_div(m, BM);
  const int64_t tiles_n = ceil_div(n, BN);
  const int64_t t = xcd_remap(blockIdx.x, (int64_t)gridDim.x);
  int64_t tm, tn;
  grouped_tile(t, tiles_m, tiles_n, group_m, tm, tn);
  dgemm_tile<BM, BN, BK, WAVES_M, WAVES_N, LA, LB>(m, n, k, alpha, A, rsA, csA, B, rsB, csB,
                                                   beta, C, rsC, csC, tm * BM, tn * BN);
}

template <int STAGES, int MINB>
__global__ void __launch_bounds__(256, MINB)
dgemm_dma_kernel(int64_t m, int64_t n, int64_t k, double alpha, const double* __restrict__ A,
                 int64_t lda, uint32_t a_bytes, const double* __restrict__ B, int64_t ldb,
                 uint32_t b_bytes, double beta, double* __restrict__ C, int64_t ldc,
                 int64_t group_m) {
  __shared__ DgemmDmaSmem<STAGES> sm;
  const int64_t tiles_m = ceil_div(m, 128);
  const int64_t tiles_n = ceil_div(n, 128);
  const int64_t t = xcd_remap(blockIdx.x, (int64_t)gridDim.x);
  int64_t tm, tn;
  grouped_tile(t, tiles_m, tiles_n, group_m, tm, tn);
  dgemm_tile_dma<STAGES>(sm, m, n, k, alpha, make_rsrc(A, a_bytes), lda, make_rsrc(B, b_bytes),
                         ldb, beta, C, ldc, tm * 128, tn * 128);
}

// The data-parallel kernel with the tiles taken in the order of a control
// tree's M / N loops (tile_order.hpp) instead of the grouped raster: same
// tiles, same k loop, so the same bits as dgemm_dma_kernel.
template <int STAGES, int MINB>
__global__ void __launch_bounds__(256, MINB)
dgemm_dma_tree_kernel(int64_t m, int64_t n, int64_t k, double alpha, const double* __restrict__ A,
                      int64_t lda, uint32_t a_bytes, const double* __restrict__ B, int64_t ldb,
                      uint32_t b_bytes, double beta, double* __restrict__ C, int64_t ldc,
                      TileOrder order) {
  __shared__ DgemmDmaSmem<STAGES> sm;
  const int64_t tiles_m = ceil_div(m, 128);
  const int64_t tiles_n = ceil_div(n, 128);
  const int64_t t = xcd_remap(blockIdx.x, (int64_t)gridDim.x);  // consecutive tiles on one XCD
  int64_t tm, tn;
  tile_order_decode(t, tiles_m, tiles_n, order, tm, tn);
  dgemm_tile_dma<STAGES>(sm, m, n, k, alpha, make_rsrc(A, a_bytes), lda, make_rsrc(B, b_bytes),
                         ldb, beta, C, ldc, tm * 128, tn * 128);
}

// Split-k form for shapes with too few tiles to fill the device: workgroup
// (part, tile) multiplies k-range [part*kper, (part+1)*kper) of the tile into
// its own slab of the workspace W (m x n column-major per part, alpha = 1,
// beta = 0); dgemm_splitk_reduce then forms C := beta*C + alpha*(W_0 + ... ),
// the parts added in order, so the result does not depend on timing.
template <int STAGES, int MINB>
__global__ void __launch_bounds__(256, MINB)
dgemm_dma_splitk_kernel(int64_t m, int64_t n, int64_t k, int64_t kper, const double* __restrict__ A,
                        int64_t lda, uint64_t a_bytes, const double* __restrict__ B, int64_t ldb,
                        uint64_t b_bytes, double* __restrict__ W, int64_t group_m) {
  __shared__ DgemmDmaSmem<STAGES> sm;
  const int64_t tiles_m = ceil_div(m, 128);
  const int64_t tiles_n = ceil_div(n, 128);
  const int64_t tiles = tiles_m * tiles_n;
  const int64_t t = xcd_remap(blockIdx.x, (int64_t)gridDim.x);
  const int64_t part = t / tiles;
  int64_t tm, tn;
  grouped_tile(t - part * tiles, tiles_m, tiles_n, group_m, tm, tn);
  const int64_t kb = part * kper;
  const int64_t kl = (k - kb) < kper ? (k - kb) : kper;
  // the part's operands start at column kb of A and row kb of B; the buffer
  // extents shrink by what lies before them
  dgemm_tile_dma<STAGES>(sm, m, n, kl, 1.0, make_rsrc(A + kb * lda, (uint32_t)(a_bytes - kb * lda * 8)),
                         lda, make_rsrc(B + kb, (uint32_t)(b_bytes - kb * 8)), ldb, 0.0,
                         W + part * m * n, m, tm * 128, tn * 128);
}

__global__ void __launch_bounds__(256)
dgemm_splitk_reduce(int64_t m, int64_t n, int64_t parts, const double* __restrict__ W,
                    double alpha, double beta, double* __restrict__ C, int64_t rsC, int64_t csC) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  for (int64_t j = blockIdx.y; j < n; j += gridDim.y) {
    double s = W[i + j * m];
    for (int64_t p = 1; p < parts; ++p) s += W[p * m * n + i + j * m];
    double* cp = C + i * rsC + j * csC;
    const double v = alpha * s;
    *cp = (beta == 0.0) ? v : __builtin_fma(beta, *cp, v);
  }
}

// The DMA path needs column-major A, B and C, 16-byte aligned columns (for
// the 16-byte loads) and operand extents a 32-bit buffer offset can address.
static bool dma_ok(const DgemmArgs& g, uint32_t& a_bytes, uint32_t& b_bytes) {
  if (g.rsA != 1 || g.rsB != 1 || g.rsC != 1 || g.k == 0) return false;
  if ((g.csA & 1) || (g.csB & 1)) return false;
  if ((reinterpret_cast<uintptr_t>(g.A) & 15) || (reinterpret_cast<uintptr_t>(g.B) & 15)) return false;
  const uint64_t ea = (uint64_t)((g.k - 1) * g.csA + g.m) * 8;
  const uint64_t eb = (uint64_t)((g.n - 1) * g.csB + g.k) * 8;
  const uint64_t lim = (1ull << 32) - (1ull << 20);
  if (ea > lim || eb > lim) return false;
  a_bytes = (uint32_t)ea;
  b_bytes = (uint32_t)eb;
  return true;
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

// Number of k-parts for the 128 x 128 DMA tile: enough workgroups for two per
// CU, each part at least 256 deep; then, for small outputs, parts down to 64
// deep while the partial sums stay within 32 MiB (their traffic is what
// limits more parts: 512^3 42 -> 28 us, 1024 x 1024 x 512 47 -> 36 us, while
// eight parts at 1024^3 would be slower than four); 1 = no split.
int dgemm_splitk_parts(int64_t m, int64_t n, int64_t k) {
  const int64_t tiles = ceil_div(m, 128) * ceil_div(n, 128);
  int64_t parts = 1;
  while (tiles * parts < 512 && k / (parts * 2) >= 256 && parts < 16) parts *= 2;
  while (tiles * parts < 512 && k / (parts * 2) >= 64 && parts < 16 &&
         (double)(parts * 2) * (double)m * (double)n * 8.0 <= 32.0 * 1024 * 1024)
    parts *= 2;
  return (int)parts;
}

// k per part and the number of parts that hold any k, for the split-k form.
static void splitk_plan(const DgemmArgs& g, int64_t& kper, int64_t& used) {
  const int64_t parts = dgemm_splitk_parts(g.m, g.n, g.k);
  kper = ceil_div(ceil_div(g.k, parts), 16) * 16;
  used = ceil_div(g.k, kper);
}

// The configuration a launch of tile_cfg (-1: built-in) actually runs.
static int resolve_cfg(const DgemmArgs& g, int tile_cfg, uint32_t& a_bytes, uint32_t& b_bytes) {
  if (tile_cfg < 0) tile_cfg = dgemm_pick_tile(g.m, g.n, g.k);
  const bool dma = dma_ok(g, a_bytes, b_bytes);
  if (tile_cfg >= 10 && ksplit_ok(g)) return tile_cfg;
  if (tile_cfg >= 5 && (!dma || tile_cfg >= 10)) {
    // operands the buffer->LDS path cannot take: the register-staged tiles
    const int64_t big = ceil_div(g.m, 128) * ceil_div(g.n, 128);
    const int64_t mid = ceil_div(g.m, 128) * ceil_div(g.n, 64);
    tile_cfg = big >= 256 ? 0 : mid >= 256 ? 2 : 1;
  }
  return tile_cfg;
}

hipError_t dgemm_workspace_bytes(const DgemmArgs& g, int tile_cfg, size_t& parts_bytes) {
  parts_bytes = 0;
  if (g.m <= 0 || g.n <= 0) return hipSuccess;
  uint32_t a_bytes = 0, b_bytes = 0;
  tile_cfg = resolve_cfg(g, tile_cfg, a_bytes, b_bytes);
  if (tile_cfg == 5) {
    int64_t kper, used;
    splitk_plan(g, kper, used);
    parts_bytes = (size_t)used * g.m * g.n * 8;
  } else if (tile_cfg >= 6 && tile_cfg <= 9) {
    hipError_t e = hipSuccess;
    const int64_t grid = streamk_grid(g, tile_cfg - 6, kStreamkMinSlabs, e);
    if (e != hipSuccess) return e;
    parts_bytes = streamk_workspace_bytes(grid, tile_cfg - 6);
  }
  return hipSuccess;
}

hipError_t dgemm_launch(const DgemmArgs& g, int tile_cfg, const DgemmWorkspace* ws) {
  // A single column (or row) operand is unit-stride along its only index, so
  // it is treated as column (row) major whatever its other stride.
  const int la = layout_of(g.rsA, g.csA, g.m, g.k);
  const int lb = layout_of(g.rsB, g.csB, g.k, g.n);
  uint32_t a_bytes = 0, b_bytes = 0;
  tile_cfg = resolve_cfg(g, tile_cfg, a_bytes, b_bytes);
  if (tile_cfg >= 10) return ksplit_launch(g, tile_cfg - 10);
  if (tile_cfg >= 6 && tile_cfg <= 9) {
    const int cfg = tile_cfg - 6;
    hipError_t e = hipSuccess;
    const int64_t grid = streamk_grid(g, cfg, kStreamkMinSlabs, e);
    if (e != hipSuccess) return e;
    const size_t pb = streamk_workspace_bytes(grid, cfg);
    double* parts = nullptr;
    if (ws == nullptr) {
      if ((e = stream_scratch(g.stream, pb, (void**)&parts)) != hipSuccess) return e;
    } else {
      if (ws->parts_bytes < pb) return hipErrorInvalidValue;
      parts = ws->parts;
    }
    return streamk_launch(g, cfg, grid, a_bytes, b_bytes, parts);
  }
  if (tile_cfg == 5) {
    int64_t kper, used;
    splitk_plan(g, kper, used);
    const size_t need = (size_t)used * g.m * g.n * 8;
    double* W = nullptr;
    hipError_t e = hipSuccess;
    if (ws == nullptr) {
      if ((e = stream_scratch(g.stream, need, (void**)&W)) != hipSuccess) return e;
    } else {
      if (ws->parts_bytes < need) return hipErrorInvalidValue;
      W = ws->parts;
    }
    const int64_t tiles = ceil_div(g.m, 128) * ceil_div(g.n, 128);
    hipLaunchKernelGGL((dgemm_dma_splitk_kernel<2, 2>), dim3((unsigned)(tiles * used)), dim3(256),
                       0, g.stream, g.m, g.n, g.k, kper, g.A, g.csA, (uint64_t)a_bytes, g.B,
                       g.csB, (uint64_t)b_bytes, W, (int64_t)8);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  dim3 grid((unsigned)ceil_div(g.m, 256), (unsigned)(g.n < 65535 ? g.n : 65535));
    hipLaunchKernelGGL(dgemm_splitk_reduce, grid, dim3(256), 0, g.stream, g.m, g.n, used, W,
                       g.alpha, g.beta, g.C, g.rsC, g.csC);
    return hipGetLastError();
  }
  if (tile_cfg == 0 && dma_ok(g, a_bytes, b_bytes)) {
    const int64_t tiles = ceil_div(g.m, 128) * ceil_div(g.n, 128);
    hipLaunchKernelGGL((dgemm_dma_kernel<2, 2>), dim3((unsigned)tiles), dim3(256), 0, g.stream,
                       g.m, g.n, g.k, g.alpha, g.A, g.csA, a_bytes, g.B, g.csB, b_bytes, g.beta,
                       g.C, g.csC, (int64_t)8);
    return hipGetLastError();
  }
  if (tile_cfg == 4 && dma_ok(g, a_bytes, b_bytes)) {
    const int64_t tiles = ceil_div(g.m, 128) * ceil_div(g.n, 128);
    hipLaunchKernelGGL((dgemm_dma_kernel<4, 1>), dim3((unsigned)tiles), dim3(256), 0, g.stream,
                       g.m, g.n, g.k, g.alpha, g.A, g.csA, a_bytes, g.B, g.csB, b_bytes, g.beta,
                       g.C, g.csC, (int64_t)8);
    return hipGetLastError();
  }
  switch (tile_cfg) {
    case 1: return launch_cfg<64, 64, 16, 2, 2>(g, la, lb, 8);
    case 2: return launch_cfg<128, 64, 16, 2, 2>(g, la, lb, 8);
    default: return launch_cfg<128, 128, 16, 2, 2>(g, la, lb, 8);  // 0/3/4 without DMA
  }
}

int dgemm_pick_tile(int64_t m, int64_t n, int64_t k) {
  // The data-parallel grid of 128 x 128 tiles when they come in nearly whole
  // rounds of one per CU (one such workgroup keeps a CU's four SIMDs
  // multiplying); otherwise the persistent stream-K form, which shares the
  // (tile, k-slab) iterations out evenly whatever the tile count: 64 x 64
  // tiles at three workgroups per CU when there are fewer big tiles than
  // CUs, 128 x 64 tiles at two per CU when the last round of big tiles
  // would be less than 85 % full (3000^3: 47 -> 55 TF).
  // Problems of at most 16 such tiles take 32 x 32 tiles with k split over
  // the four wavefronts of each workgroup (512^3: 12.9 -> 19.8 TF, 256^3:
  // 2.3 -> 3.9 TF, 256 x 1024 x 1024: 14.3 -> 23.0 TF; from 640^3 on the
  // stream-K form is faster).
  (void)k;
  constexpr int64_t kCus = 256;
  const int64_t big = ceil_div(m, 128) * ceil_div(n, 128);
  if (big <= 16) return 11;
  if (big < kCus) return 6;
  const int64_t rounds = ceil_div(big, kCus);
  if (big * 100 >= rounds * kCus * 85) return 0;
  return 7;
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

extern "C" int momms_hip_dgemm_tree(int64_t m, int64_t n, int64_t k, const double* alpha,
                                    const double* A, int64_t rsA, int64_t csA, const double* B,
                                    int64_t rsB, int64_t csB, const double* beta, double* C,
                                    int64_t rsC, int64_t csC, const momms_algo_step* steps,
                                    int64_t nsteps, void* stream) {
  using namespace momms;
  static const char* fn = "momms_hip_dgemm_tree";
  if (m < 0) return bad_arg(fn, 1, "m < 0");
  if (n < 0) return bad_arg(fn, 2, "n < 0");
  if (k < 0) return bad_arg(fn, 3, "k < 0");
  if (alpha == nullptr) return bad_arg(fn, 4, "alpha is null");
  int st;
  if ((st = check_operand(fn, 5, A, m, k, rsA, csA))) return st;
  if ((st = check_operand(fn, 8, B, k, n, rsB, csB))) return st;
  if (beta == nullptr) return bad_arg(fn, 11, "beta is null");
  if ((st = check_operand(fn, 12, C, m, n, rsC, csC))) return st;
  if (nsteps < 0 || nsteps > 64 || (nsteps > 0 && steps == nullptr))
    return bad_arg(fn, 16, "nsteps");
  int32_t kinds[64];
  int64_t bszs[64];
  for (int64_t i = 0; i < nsteps; ++i) {
    kinds[i] = steps[i].kind;
    bszs[i] = steps[i].bsz;
  }
  TileOrder order;
  if (!tile_order_from_steps(kinds, bszs, nsteps, 128, 128, order))
    return bad_arg(fn, 15, "steps: a kind other than M, N, K, a block size below 1, or more "
                           "than 8 M / N steps");
  if (m == 0 || n == 0) return 0;
  DgemmArgs g{m, n, k, *alpha, A, rsA, csA, B, rsB, csB, *beta, C, rsC, csC,
              (hipStream_t)stream};
  uint32_t a_bytes = 0, b_bytes = 0;
  if (!dma_ok(g, a_bytes, b_bytes)) return hip_status(fn, dgemm_launch(g, 0));
  const int64_t tiles = ceil_div(m, 128) * ceil_div(n, 128);
  hipLaunchKernelGGL((dgemm_dma_tree_kernel<2, 2>), dim3((unsigned)tiles), dim3(256), 0, g.stream,
                     g.m, g.n, g.k, g.alpha, g.A, g.csA, a_bytes, g.B, g.csB, b_bytes, g.beta,
                     g.C, g.csC, order);
  return hip_status(fn, hipGetLastError());
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

// ---------------------------------------------------------------------------
// Allocation-free entry point. The workspace of the stream-K form (parts of
// split tiles) lives in a fixed device array instead of library scratch, so
// no call allocates. The built-in choice takes stream-K tiles of at most
// 128 x 64 doubles, two workspace slots per resident workgroup: 512
// workgroups x 2 x 64 KiB = 64 MiB on a 256-CU MI355X (768 x 2 x 32 KiB =
// 48 MiB for 64 x 64 tiles), so 80 MiB covers it with room to spare; a
// device that would need more is refused with an error.
constexpr size_t kNoAllocScratchBytes = size_t(80) << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

namespace momms {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices];  // device address of `scratch` on each device

// The device address of the current device's copy of `scratch`, looked up
// once per device with that device current.
hipError_t noalloc_scratch(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[dev] == nullptr) {
    void* p = nullptr;
    if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess) return e;
    if (p == nullptr) return hipErrorInvalidSymbol;
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[dev];
  return hipSuccess;
}

}  // namespace
}  // namespace momms

extern "C" int momms_hip_dgemm_noalloc(int64_t m, int64_t n, int64_t k, const double* alpha,
                                       const double* A, int64_t rsA, int64_t csA,
                                       const double* B, int64_t rsB, int64_t csB,
                                       const double* beta, double* C, int64_t rsC, int64_t csC,
                                       void* stream) {
  using namespace momms;
  static const char* fn = "momms_hip_dgemm_noalloc";
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
  size_t parts_bytes = 0;
  hipError_t e = dgemm_workspace_bytes(g, -1, parts_bytes);
  if (e != hipSuccess) return hip_status(fn, e);
  if (parts_bytes > kNoAllocScratchBytes)
    return bad_arg(fn, 1, "the workspace this shape needs exceeds the 80 MiB static scratch");
  DgemmWorkspace ws{nullptr, 0};
  if (parts_bytes > 0) {
    if ((e = noalloc_scratch(&ws.parts)) != hipSuccess) return hip_status(fn, e);
    ws.parts_bytes = kNoAllocScratchBytes;
  }
  return hip_status(fn, dgemm_launch(g, -1, &ws));
}
