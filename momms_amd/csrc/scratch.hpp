// Library-owned device scratch, one buffer per (device, stream, calling
// thread, slot), grown on demand and kept for the life of the process. Used
// by entry points whose ABI has no workspace argument (the split-k FP64 form,
// the packed BF16 GEMM, the hierarchical offset tables: slot 0; the
// column-major operands of the GEMM on hierarchical matrices, which must
// outlive the slot-0 users it calls: slot 1). Work that uses a buffer is
// ordered on its stream, so consecutive calls of one thread on one stream
// share it. Host threads that launch onto one stream (the NULL stream of
// several sub-teams, hipStreamPerThread) get buffers of their own, since
// their launches interleave on that stream. Growing synchronizes the stream
// before the old buffer is freed. A call that needs to grow it cannot be
// captured in a HIP graph (it allocates); once large enough it allocates
// nothing.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace momms {

hipError_t stream_scratch(hipStream_t st, size_t bytes, void** out, int slot = 0);

}  // namespace momms
