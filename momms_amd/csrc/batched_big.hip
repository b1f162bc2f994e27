// Batched FP64 GEMM, large problems: one persistent kernel of 128 x 128
// lean tiles over per-XCD queues. Its own translation unit, so that its
// register allocation (64 accumulators a lane for a 64 x 64 wave tile) does
// not weigh on the small-tile kernel's.
#include <hip/hip_runtime.h>

#include "batch_tiles.hpp"
#include "batched.hpp"

namespace momms {

// Two workgroups per CU (74 KB of LDS each).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
batched_big_kernel(const momms_gemm_desc* __restrict__ d, const BatchItem* __restrict__ items,
                   QueueBounds qb, unsigned* counters, double alpha, const double* __restrict__ A,
                   const double* __restrict__ B, double beta, double* __restrict__ C) {
  run_queues(qb, counters, [&](int i) {
    const BatchItem w = items[uniform(i)];
    const momms_gemm_desc p = d[uniform(w.prob)];
    batch_tile<128, 128, 2>(p.m, p.n, p.k, alpha, A + p.a_off, p.lda, B + p.b_off, p.ldb, beta,
                            C + p.c_off, p.ldc, (int64_t)uniform(w.tm) * 128,
                            (int64_t)uniform(w.tn) * 128);
  });
}

const void* batched_big_kernel_ptr() { return reinterpret_cast<const void*>(&batched_big_kernel); }

void launch_batched_big(int grid, hipStream_t st, const momms_gemm_desc* d, const BatchItem* items,
                        QueueBounds qb, unsigned* counters, double alpha, const double* A,
                        const double* B, double beta, double* C) {
  hipLaunchKernelGGL(batched_big_kernel, dim3((unsigned)grid), dim3(256), 0, st, d, items, qb,
                     counters, alpha, A, B, beta, C);
}

}  // namespace momms
