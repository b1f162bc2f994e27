// Internal launch interface of the FP64 matrix-core GEMM (used by the C ABI
// entry points and by the batched and split-precision paths).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace momms {

struct DgemmArgs {
  int64_t m, n, k;
  double alpha;
  const double* A;
  int64_t rsA, csA;
  const double* B;
  int64_t rsB, csB;
  double beta;
  double* C;
  int64_t rsC, csC;
  hipStream_t stream;
};

// Tile configurations (BM x BN, 4 wavefronts): 0 = 128x128 (buffer->LDS
// with a 2-slab ring and two workgroups per CU when the operands allow it),
// 1 = 64x64, 2 = 128x64, 3 = 128x128 register-staged form only,
// 4 = 128x128 buffer->LDS with a 4-slab ring, one workgroup per CU,
// 5 = 128x128 buffer->LDS with k split into parts (dgemm_splitk_parts) and a
// fixed-order reduction (falls back to 0 when the DMA path does not apply).
constexpr int kDgemmTileConfigs = 6;

int dgemm_splitk_parts(int64_t m, int64_t n, int64_t k);

int dgemm_pick_tile(int64_t m, int64_t n, int64_t k);
hipError_t dgemm_launch(const DgemmArgs& g, int tile_cfg);

}  // namespace momms
