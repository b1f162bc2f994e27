// CDNA4 (gfx950) matrix-core vocabulary shared by every kernel in this library.
//
// Operand and result lane layouts (64-wide wavefront, lane l, result register r):
//   v_mfma_f64_16x16x4_f64   A[p][k]: p=l%16, k=l/16      B[k][q]: q=l%16, k=l/16
//                            D[p][q]: p=l/16+4*r, q=l%16                   (r<4)
//   v_mfma_f32_32x32x16_bf16 A[p][k]: p=l%32, k=8*(l/32)+e  B[k][q]: q=l%32, k=8*(l/32)+e (e<8)
//                            D[p][q]: p=8*(r/4)+4*(l/32)+r%4, q=l%32       (r<16)
//   v_mfma_f32_32x32x2_f32   A[p][k]: p=l%32, k=l/32      B[k][q]: q=l%32, k=l/32
//                            D as the 32x32 bf16 form
//   v_mfma_i32_32x32x32_i8   A[p][k]: p=l%32, k=16*(l/32)+e B[k][q]: q=l%32, k=16*(l/32)+e (e<16)
//                            D as the 32x32 bf16 form
// These layouts are checked on the device by tests/test_gpu_mfma_layout.py.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace momms {

typedef double v4f64 __attribute__((ext_vector_type(4)));
typedef double v2f64 __attribute__((ext_vector_type(2)));
typedef float v4f32 __attribute__((ext_vector_type(4)));
typedef float v16f32 __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf16 __attribute__((ext_vector_type(8)));
typedef int v4i32 __attribute__((ext_vector_type(4)));
typedef int v16i32 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ v4f64 mfma_f64_16x16x4(double a, double b, v4f64 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v16f32 mfma_bf16_32x32x16(v8bf16 a, v8bf16 b, v16f32 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v16f32 mfma_f32_32x32x2(float a, float b, v16f32 c) {
  return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v16i32 mfma_i8_32x32x32(v4i32 a, v4i32 b, v16i32 c) {
  return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
}

// Storage order of a matrix operand in global memory.
//   kCol: unit row stride (column-major, the Matrix<T> default)
//   kRow: unit column stride (row-major, a transposed Matrix<T>)
//   kGen: neither stride is one (a strided view); element-wise gather
enum Layout : int { kCol = 0, kRow = 1, kGen = 2 };

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// MI355X places consecutive workgroup ids on its 8 XCDs round-robin; each XCD
// has its own L2. Remapping so that consecutive *logical* tile ids land on one
// XCD keeps the A row-panels and B column-panels a group of tiles shares in
// one L2. Only a bijection when the grid is a multiple of 8, so the tail is
// left in dispatch order.
__device__ __forceinline__ int64_t xcd_remap(int64_t bid, int64_t nblocks) {
  const int64_t kXcd = 8;
  const int64_t full = (nblocks / kXcd) * kXcd;
  if (bid >= full) return bid;
  const int64_t per = full / kXcd;
  return (bid % kXcd) * per + bid / kXcd;
}

// Grouped raster: walks GROUP tile-rows at a time down the N direction so that
// a group of concurrently resident tiles reuses the same B column-panels.
__device__ __forceinline__ void grouped_tile(int64_t t, int64_t tiles_m, int64_t tiles_n,
                                             int64_t group, int64_t& tm, int64_t& tn) {
  const int64_t per_group = group * tiles_n;
  const int64_t g = t / per_group;
  const int64_t first_m = g * group;
  const int64_t gsize = (tiles_m - first_m) < group ? (tiles_m - first_m) : group;
  const int64_t in = t - g * per_group;
  tm = first_m + in % gsize;
  tn = in / gsize;
}

}  // namespace momms
