// Batched FP64 GEMM, problems whose lda, ldb or ldc exceed the lean tiles'
// 32-bit offsets: one persistent kernel of 64 x 64 tiles (dgemm_tile_deep,
// 64-bit addressing), one queue of tiles sorted by k.
#include <hip/hip_runtime.h>

#include "batched.hpp"
#include "dgemm_tile.hpp"

namespace momms {

__global__ void __launch_bounds__(256)
batched_wide_kernel(const momms_gemm_desc* __restrict__ d, const BatchItem* __restrict__ items,
                    int nitems, unsigned* counter, double alpha, const double* __restrict__ A,
                    const double* __restrict__ B, double beta, double* __restrict__ C) {
  run_queue(counter, nitems, [&](int it) {
    const BatchItem w = items[it];
    const momms_gemm_desc p = d[w.prob];
    dgemm_tile_deep<64, 64, 3>(p.m, p.n, p.k, alpha, A + p.a_off, p.lda, B + p.b_off, p.ldb, beta,
                               C + p.c_off, p.ldc, (int64_t)w.tm * 64, (int64_t)w.tn * 64);
  });
}

const void* batched_wide_kernel_ptr() { return reinterpret_cast<const void*>(&batched_wide_kernel); }

void launch_batched_wide(int grid, hipStream_t st, const momms_gemm_desc* d, const BatchItem* items,
                         int nitems, unsigned* counter, double alpha, const double* A,
                         const double* B, double beta, double* C) {
  hipLaunchKernelGGL(batched_wide_kernel, dim3((unsigned)grid), dim3(256), 0, st, d, items, nitems,
                     counter, alpha, A, B, beta, C);
}

}  // namespace momms
