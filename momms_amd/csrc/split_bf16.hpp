// Internal interface of the split-BF16 FP64-accurate GEMM (used by the C ABI
// entry points and the batched path).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace momms {

// Workspace layout for one (m, n, k, s) problem: packed slices of A and B,
// per-row / per-column exponents and the max |x| scratch they come from.
struct SplitLayout {
  int s;
  int cfg;         // multiply configuration (split_config), fixes the panel sizes
  int pa, pb;      // rows of A / columns of B per packed panel (= tile BM, BN)
  int64_t mpad, npad, kpad;
  size_t off_aq, off_bq, off_ea, off_eb, off_maxa, off_maxb, bytes;
};

// Multiply configurations: wavefront grid WM x WN (tile 32*WM x 32*WN),
// ring depth, and workgroups per CU the kernel is compiled for.
struct SplitConfig {
  int wm, wn;  // wavefronts along m and n; the tile is 32*wm x 32*wn
};
constexpr int kSplitConfigs = 3;
extern const SplitConfig kSplitConfig[kSplitConfigs];
// The configuration used for a problem (MOMMS_SPLIT_CFG overrides the default).
int split_config(int64_t m, int64_t n, int64_t k, int s);

// Exponent E of each of count vectors of length len (vector v, element p at
// X[v*sv + p*se]): max |x| < 2^E (frexp), 0 for a zero vector, kNonFinite for
// a vector holding Inf or NaN; e[v] = 0 for count <= v < padded. mx is scratch
// of count words.
hipError_t vector_exponents(const double* X, int64_t count, int64_t len, int64_t sv, int64_t se,
                            int64_t padded, uint64_t* mx, int* e, hipStream_t st);

SplitLayout split_layout(int64_t m, int64_t n, int64_t k, int s);
int split_chunk(int s);

hipError_t split_prepare(const SplitLayout& L, int64_t m, int64_t n, int64_t k, const double* A,
                         int64_t rsA, int64_t csA, const double* B, int64_t rsB, int64_t csB,
                         char* ws, hipStream_t st);
hipError_t split_multiply_launch(const SplitLayout& L, int64_t m, int64_t n, int64_t k,
                                 double alpha, double beta, double* C, int64_t rsC, int64_t csC,
                                 const char* ws, hipStream_t st);

}  // namespace momms
