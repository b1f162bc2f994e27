// Single-call small FP64 GEMM: the device counterpart of the reference's
// libxsmm leaves, Xsmm and KernelXsmmA2 (src/kern/kernel_xsmm.rs:8-139),
// which multiply a small block straight from its operands, without packing,
// when every operand is column-major (or, by transposition, row-major).
//
// One launch, one wavefront per 32 x 32 tile of C (four per workgroup), no
// LDS, no barrier and no workspace: the register-only one-wavefront tile of
// the batched kernels (batch_wave_tile32, 8-deep k chunks, four chunks of
// loads in flight), which reads A, B and C through buffer resources bounded
// by each operand's extent. Shapes it cannot take (other strides, leading
// dimensions over 2^21, operands over 2^28 doubles) run momms_hip_dgemm.
#include "batch_tiles.hpp"
#include "dgemm.hpp"
#include "momms_hip.h"
#include "status.hpp"

namespace momms {

__global__ void __launch_bounds__(256)
dgemm_small_kernel(int64_t m, int64_t n, int64_t k, double alpha, const double* __restrict__ A,
                   int64_t lda, const double* __restrict__ B, int64_t ldb, double beta,
                   double* __restrict__ C, int64_t ldc) {
  const int64_t tiles_m = ceil_div(m, 32), tiles = tiles_m * ceil_div(n, 32);
  const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= tiles) return;  // a whole wavefront; the tile has no barrier
  const int64_t tm = t % tiles_m, tn = t / tiles_m;
  batch_wave_tile32<4>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tm * 32, tn * 32);
}

// An h x w operand (strides rs, cs) read as column-major with leading
// dimension ld >= h: unit row stride (or a single row), and columns that do
// not overlap (or a single column).
static bool colmajor(int64_t h, int64_t w, int64_t rs, int64_t cs, int64_t& ld) {
  if (h > 1 && rs != 1) return false;
  if (w > 1 && cs < h) return false;
  ld = w > 1 ? cs : (h > 1 ? h : 1);
  return true;
}

// Column-major operands the lean tile addresses with 32-bit offsets.
static bool small_ok(int64_t m, int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc) {
  if (lda > kBatchLeanLd || ldb > kBatchLeanLd || ldc > kBatchLeanLd) return false;
  const int64_t lim = int64_t(1) << 28;
  return (k > 0 ? (k - 1) * lda + m : 0) <= lim && (n - 1) * ldb + k <= lim &&
         (n - 1) * ldc + m <= lim;
}

}  // namespace momms

extern "C" int momms_hip_dgemm_small(int64_t m, int64_t n, int64_t k, const double* alpha,
                                     const double* A, int64_t rsA, int64_t csA, const double* B,
                                     int64_t rsB, int64_t csB, const double* beta, double* C,
                                     int64_t rsC, int64_t csC, void* stream) {
  using namespace momms;
  static const char* fn = "momms_hip_dgemm_small";
  if (m < 0) return bad_arg(fn, 1, "m < 0");
  if (n < 0) return bad_arg(fn, 2, "n < 0");
  if (k < 0) return bad_arg(fn, 3, "k < 0");
  if (alpha == nullptr) return bad_arg(fn, 4, "alpha is null");
  int st;
  if ((st = check_operand(fn, 5, A, m, k, rsA, csA))) return st;
  if ((st = check_operand(fn, 8, B, k, n, rsB, csB))) return st;
  if (beta == nullptr) return bad_arg(fn, 11, "beta is null");
  if ((st = check_operand(fn, 12, C, m, n, rsC, csC))) return st;
  if (m == 0 || n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  int64_t lda, ldb, ldc;
  if (colmajor(m, k, rsA, csA, lda) && colmajor(k, n, rsB, csB, ldb) &&
      colmajor(m, n, rsC, csC, ldc) && small_ok(m, n, k, lda, ldb, ldc)) {
    const int64_t tiles = ceil_div(m, 32) * ceil_div(n, 32);
    hipLaunchKernelGGL(dgemm_small_kernel, dim3((unsigned)ceil_div(tiles, 4)), dim3(256), 0, s,
                       m, n, k, *alpha, A, lda, B, ldb, *beta, C, ldc);
    return hip_status(fn, hipGetLastError());
  }
  // row-major operands: C^T = B^T A^T, whose operands are column-major
  if (colmajor(k, m, csA, rsA, lda) && colmajor(n, k, csB, rsB, ldb) &&
      colmajor(n, m, csC, rsC, ldc) && small_ok(n, m, k, ldb, lda, ldc)) {
    const int64_t tiles = ceil_div(n, 32) * ceil_div(m, 32);
    hipLaunchKernelGGL(dgemm_small_kernel, dim3((unsigned)ceil_div(tiles, 4)), dim3(256), 0, s,
                       n, m, k, *alpha, B, ldb, A, lda, *beta, C, ldc);
    return hip_status(fn, hipGetLastError());
  }
  DgemmArgs g{m, n, k, *alpha, A, rsA, csA, B, rsB, csB, *beta, C, rsC, csC, s};
  return hip_status(fn, dgemm_launch(g, -1));
}
