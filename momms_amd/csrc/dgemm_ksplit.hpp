// The FP64 GEMM with k split across the wavefronts of one workgroup
// (dgemm_ksplit.hip): tile configurations 10 (64 x 64) and 11 (32 x 32).
#pragma once
#include <hip/hip_runtime.h>

#include "dgemm.hpp"

namespace momms {

// Whether the operands suit the form: A and B column-major with leading
// dimensions at most 2^21 (32-bit lane offsets); C any strides.
bool ksplit_ok(const DgemmArgs& g);

// Enqueues the GEMM: form 0 = 64 x 64 tiles (one workgroup per CU), 1 =
// 32 x 32 tiles (two per CU). Requires ksplit_ok(g).
hipError_t ksplit_launch(const DgemmArgs& g, int form);

}  // namespace momms
