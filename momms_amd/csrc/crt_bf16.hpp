// Internal interface of the Chinese-remainder split-BF16 FP64-accurate GEMM
// (crt_bf16.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "crt_recon.hpp"

namespace momms {

constexpr int kCrtBM = 256, kCrtBN = 256;  // residue GEMM tile = panel sizes
constexpr int kCrtKAlign = 32;             // k padding: whole two-slab stages

// Workspace layout for one (m, n, k, N) problem: packed residues of A and B,
// the int16 residue matrices of C' (one per modulus), exponents and scratch.
struct CrtLayout {
  int nmod, beta;
  int64_t mpad, npad, kpad;
  size_t off_ra, off_rb, off_rc, off_ea, off_eb, off_maxa, off_maxb, bytes;
};

int crt_modulus(int l);
// Bits of the exponent-aligned integers for k and N moduli; -1 if N moduli
// cannot hold a k-term product.
int crt_beta(int64_t k, int N);
CrtLayout crt_layout(int64_t m, int64_t n, int64_t k, int N);

}  // namespace momms
