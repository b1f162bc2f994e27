// The matrix cores' sustained rate on this device with the operands in
// registers: every CU runs four wavefronts (one per SIMD), each issuing
// independent chains of one MFMA form back to back, with no memory traffic
// in the loop. Run with random operands (the bits toggle as a GEMM's do) and
// with zero operands, it separates the clock the chip holds under full
// matrix-core load at its power limit from the dense peak's 2.4 GHz: a GEMM
// kernel cannot exceed the random-operand rate, whatever its data movement.
//
// Forms: v_mfma_f32_32x32x16_bf16 (the BF16 residue and slice GEMMs) and
// v_mfma_i32_32x32x32_i8 (the INT8 residue GEMMs). Prints one JSON line:
// per form and operand kind, the median over REPS launches (after one
// warm-up launch) of the launch time and the rate, in TFLOP/s or TOP/s.
//
// Usage: mfma_ceiling [iterations]   (default 900000: about 0.1-0.2 s a launch)
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <vector>

#define CHECK(x)                                                                      \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) {                                                           \
      fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      exit(1);                                                                        \
    }                                                                                 \
  } while (0)

typedef float v16f32 __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf16 __attribute__((ext_vector_type(8)));
typedef int v4i32 __attribute__((ext_vector_type(4)));
typedef int v16i32 __attribute__((ext_vector_type(16)));

constexpr int kChains = 8;    // independent accumulators per wavefront
constexpr int kThreads = 256;  // four wavefronts per workgroup
constexpr int kReps = 3;

__device__ inline uint32_t mix(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

// two A and two B operands per lane: random values in [-1, 1) or zeros
template <bool I8>
__global__ void __launch_bounds__(kThreads) mfma_chains(int iters, int zero, float* out) {
  const uint32_t gid = blockIdx.x * kThreads + threadIdx.x;
  float r = 0.f;
  if constexpr (!I8) {
    v8bf16 op[4];
    for (int q = 0; q < 4; ++q)
      for (int e = 0; e < 8; ++e) {
        const float x = (float)(mix(gid * 32 + q * 8 + e + 1) >> 8) * (1.0f / 8388608.0f) - 1.0f;
        op[q][e] = (__bf16)(zero ? 0.0f : x);
      }
    v16f32 acc[kChains];
    for (int c = 0; c < kChains; ++c) acc[c] = v16f32{};
    for (int it = 0; it < iters; ++it) {
#pragma unroll
      for (int c = 0; c < kChains; ++c)
        acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(op[c & 1], op[2 + ((c >> 1) & 1)], acc[c],
                                                         0, 0, 0);
    }
    for (int c = 0; c < kChains; ++c)
      for (int e = 0; e < 16; ++e) r += acc[c][e];
  } else {
    v4i32 op[4];
    for (int q = 0; q < 4; ++q)
      for (int e = 0; e < 4; ++e) op[q][e] = zero ? 0 : (int)mix(gid * 16 + q * 4 + e + 1);
    v16i32 acc[kChains];
    for (int c = 0; c < kChains; ++c) acc[c] = v16i32{};
    for (int it = 0; it < iters; ++it) {
#pragma unroll
      for (int c = 0; c < kChains; ++c)
        acc[c] = __builtin_amdgcn_mfma_i32_32x32x32_i8(op[c & 1], op[2 + ((c >> 1) & 1)], acc[c], 0,
                                                       0, 0);
    }
    for (int c = 0; c < kChains; ++c)
      for (int e = 0; e < 16; ++e) r += (float)acc[c][e];
  }
  out[gid] = r;  // keeps every chain live
}

struct Rate {
  double ms, rate;
};

template <bool I8>
Rate measure(int cus, int iters, int zero, float* out) {
  // flops (ops) per MFMA: 2 * 32 * 32 * 16 for BF16, 2 * 32 * 32 * 32 for INT8
  const double per_mfma = I8 ? 65536.0 : 32768.0;
  const double work = (double)cus * (kThreads / 64) * iters * kChains * per_mfma;
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  std::vector<float> ms;
  for (int rep = 0; rep <= kReps; ++rep) {
    CHECK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL(mfma_chains<I8>, dim3(cus), dim3(kThreads), 0, 0, iters, zero, out);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1, 0));
    CHECK(hipEventSynchronize(e1));
    float t = 0.f;
    CHECK(hipEventElapsedTime(&t, e0, e1));
    if (rep > 0) ms.push_back(t);  // rep 0 is the warm-up
  }
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  std::sort(ms.begin(), ms.end());
  const double med = ms[ms.size() / 2];
  return {med, work / med / 1e9};
}

int main(int argc, char** argv) {
  const int iters = argc > 1 ? atoi(argv[1]) : 900000;
  if (iters < 1) {
    fprintf(stderr, "iterations must be positive\n");
    return 2;
  }
  int dev = 0, cus = 0;
  CHECK(hipGetDevice(&dev));
  CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  float* out = nullptr;
  CHECK(hipMalloc(&out, (size_t)cus * kThreads * sizeof(float)));
  const Rate br = measure<false>(cus, iters, 0, out);
  const Rate bz = measure<false>(cus, iters, 1, out);
  const Rate ir = measure<true>(cus, iters, 0, out);
  const Rate iz = measure<true>(cus, iters, 1, out);
  CHECK(hipFree(out));
  printf("{\"cus\": %d, \"iterations\": %d, \"reps\": %d, "
         "\"bf16_32x32x16\": {\"random_tflops\": %.1f, \"random_ms\": %.3f, "
         "\"zero_tflops\": %.1f, \"zero_ms\": %.3f}, "
         "\"i8_32x32x32\": {\"random_tops\": %.1f, \"random_ms\": %.3f, "
         "\"zero_tops\": %.1f, \"zero_ms\": %.3f}}\n",
         cus, iters, kReps, br.rate, br.ms, bz.rate, bz.ms, ir.rate, ir.ms, iz.rate, iz.ms);
  return 0;
}
