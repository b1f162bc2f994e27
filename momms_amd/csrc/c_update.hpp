// The last step of every GEMM tile: C := beta*C + v for the outputs one lane
// holds (alpha already applied to v).
//
// Written element by element (`if (ok) *cp = fma(beta, *cp, v)`), each read
// of C sits behind the write before it and behind a branch, so the compiler
// waits for every read on its own: N outputs cost N memory round trips. Here
// all N reads are issued first, unconditionally, and share one round trip;
// an output outside the matrix (ok false) reads C[0] instead, which exists
// for any tile with m, n >= 1, and is not written. beta = 0 reads nothing,
// so C may hold NaN then (BLAS semantics).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace momms {

template <int N>
__device__ __forceinline__ void update_c(double* __restrict__ C, const int64_t (&idx)[N],
                                         const bool (&ok)[N], const double (&v)[N], double beta) {
  if (beta == 0.0) {
#pragma unroll
    for (int t = 0; t < N; ++t)
      if (ok[t]) C[idx[t]] = v[t];
  } else {
    double c[N];
#pragma unroll
    for (int t = 0; t < N; ++t) c[t] = C[ok[t] ? idx[t] : 0];
#pragma unroll
    for (int t = 0; t < N; ++t)
      if (ok[t]) C[idx[t]] = __builtin_fma(beta, c[t], v[t]);
  }
}

}  // namespace momms
