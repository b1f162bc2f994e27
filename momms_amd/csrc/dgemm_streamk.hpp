// Launch interface of the persistent stream-K FP64 GEMM (dgemm_streamk.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "dgemm.hpp"

namespace momms {

// Tile forms: 0 = 64 x 64 (three-slab ring, three workgroups per CU),
// 1 = 128 x 64 (three-slab ring, two per CU), 2 = 64 x 64 (four-slab ring,
// two per CU), 3 = 64 x 64 (six-slab ring, one per CU).
constexpr int kStreamkConfigs = 4;

// Workgroups of one launch: the device's resident count for the form, cut
// so that each takes at least min_slabs k-slabs, in whole rounds of 8 (one
// per XCD), at most one per (tile, k-slab) iteration.
int64_t streamk_grid(const DgemmArgs& g, int cfg, int min_slabs, hipError_t& err);

// Workspace for the parts of split tiles: two slots per workgroup.
size_t streamk_workspace_bytes(int64_t grid, int cfg);

// Column-major A, B, C that the buffer->LDS path accepts (dma_ok). Enqueues
// the main kernel and, when a tile is split, the fix-up kernel.
hipError_t streamk_launch(const DgemmArgs& g, int cfg, int64_t grid, uint32_t a_bytes,
                          uint32_t b_bytes, double* parts);

}  // namespace momms
