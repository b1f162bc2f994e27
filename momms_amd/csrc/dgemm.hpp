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
// fixed-order reduction, 6-9 = persistent stream-K with the tile forms of
// dgemm_streamk.hpp (6: 64x64 three per CU, 7: 128x64, 8: 64x64 two per CU,
// 9: 64x64 one per CU) (5-9 fall back to the built-in register-staged
// choice when the buffer->LDS path does not apply), 10-11 = k split over the
// four wavefronts of a workgroup, partial tiles added in LDS in wavefront
// order (10: 64x64, one workgroup per CU; 11: 32x32, two per CU;
// dgemm_ksplit.hpp; the register-staged choice when A or B is not
// column-major with leading dimension <= 2^21).
constexpr int kDgemmTileConfigs = 12;

int dgemm_splitk_parts(int64_t m, int64_t n, int64_t k);

int dgemm_pick_tile(int64_t m, int64_t n, int64_t k);

// Device workspace of one launch: the split-k form's partial products, or
// the stream-K form's parts of split tiles.
struct DgemmWorkspace {
  double* parts;
  size_t parts_bytes;
};

// Bytes of workspace the launch of `tile_cfg` (-1: the built-in choice)
// needs for g (0 for the forms that need none).
hipError_t dgemm_workspace_bytes(const DgemmArgs& g, int tile_cfg, size_t& parts_bytes);

// Enqueues the GEMM. The split-k and stream-K forms take their workspace
// from `ws` when given (sized by dgemm_workspace_bytes), else from the
// library's per-stream scratch.
hipError_t dgemm_launch(const DgemmArgs& g, int tile_cfg, const DgemmWorkspace* ws = nullptr);

// Minimum k-slabs (16 deep) per stream-K workgroup.
constexpr int kStreamkMinSlabs = 8;

}  // namespace momms
