// Library-owned device scratch, one buffer per (device, stream, slot), grown
// on demand and kept for the life of the process. Used by entry points whose
// ABI has no workspace argument (the split-k FP64 form, the packed BF16
// GEMM, the hierarchical offset tables: slot 0; the column-major operands of
// the GEMM on hierarchical matrices, which must outlive the slot-0 users it
// calls: slot 1). Work that uses it is ordered on its stream, so consecutive
// calls on one stream can share it; growing synchronizes the stream before
// the old buffer is freed. A call that needs to grow it cannot be captured
// in a HIP graph (it allocates); once large enough it allocates nothing.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace momms {

hipError_t stream_scratch(hipStream_t st, size_t bytes, void** out, int slot = 0);

}  // namespace momms
