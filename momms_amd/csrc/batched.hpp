// Pieces of the batched engine shared by its translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "momms_hip.h"

namespace momms {

struct BatchItem {
  int32_t prob, tm, tn, pad;
};

// Per-problem workspace of the split and residue modes (byte offsets into
// the plan's arena): slices of A and B (split mode only), the exponents of
// the rows of A and columns of B (mpad and npad entries), and in the residue
// mode the problem's bits beta of the integer images.
struct SplitProb {
  int64_t aq, bq, ea, eb;
  int32_t mpad, npad, kpad, beta;
};

// A persistent workgroup's loop over one queue of `count` items: body(it)
// for every item index it claims from `counter`. The claim of the next item
// (a device-scope atomic, a long round trip) is issued when an item starts
// and read when it ends, so it overlaps the item's work.
template <typename Body>
__device__ __forceinline__ void run_queue(unsigned* counter, int count, Body&& body) {
  __shared__ int s_item;
  if (threadIdx.x == 0) s_item = (int)atomicAdd(counter, 1u);
  __syncthreads();
  int it = s_item;
  __syncthreads();
  while (it < count) {
    unsigned next = 0;
    if (threadIdx.x == 0) next = atomicAdd(counter, 1u);
    body(it);
    if (threadIdx.x == 0) s_item = (int)next;
    __syncthreads();
    it = s_item;
    __syncthreads();
  }
}

// Per-XCD work queues of a persistent launch: workgroup b serves queue
// b % kBatchQueues first (workgroups are dispatched to the 8 XCDs in turn),
// then takes items from the next queues.
constexpr int kBatchQueues = 8;
struct QueueBounds {
  int off[kBatchQueues + 1];  // queue q holds items [off[q], off[q+1])
};

// body(q, it) for every item of every queue, home queue first.
template <typename Body>
__device__ __forceinline__ void run_queues(const QueueBounds& qb, unsigned* counters, Body&& body) {
  const int home = (int)(blockIdx.x % kBatchQueues);
  for (int s = 0; s < kBatchQueues; ++s) {
    const int q = (home + s) % kBatchQueues;
    const int base = qb.off[q];
    run_queue(counters + q, qb.off[q + 1] - base, [&](int it) { body(base + it); });
  }
}

// Problems whose smaller side reaches the plan's second threshold: 128 x 128
// lean tiles (batch_tile<128, 128>), two workgroups per CU, per-XCD queues
// (batched_big.hip).
const void* batched_big_kernel_ptr();
void launch_batched_big(int grid, hipStream_t st, const momms_gemm_desc* d, const BatchItem* items,
                        QueueBounds qb, unsigned* counters, double alpha, const double* A,
                        const double* B, double beta, double* C);

// Problems of a batch with a leading dimension too large for the 32-bit
// offsets of the lean tiles: 64 x 64 tiles with 64-bit addressing, one
// queue (batched_wide.hip; its own translation unit, so that its register
// allocation does not weigh on the lean kernel's).
const void* batched_wide_kernel_ptr();
void launch_batched_wide(int grid, hipStream_t st, const momms_gemm_desc* d, const BatchItem* items,
                         int nitems, unsigned* counter, double alpha, const double* A,
                         const double* B, double beta, double* C);

// Residue mode (batched_crt.hip): one persistent kernel of 64 x 64 tiles
// over per-XCD queues for N moduli (1 <= N <= 16).
const void* batched_crt_kernel_ptr(int N);
void launch_batched_crt(int N, int grid, hipStream_t st, const momms_gemm_desc* d,
                        const SplitProb* sp, const BatchItem* items, QueueBounds qb,
                        unsigned* counters, double alpha, const double* A, const double* B,
                        double beta, double* C, const char* ws);

}  // namespace momms
