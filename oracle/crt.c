/*
 * ORACLE — test infrastructure only (tests/, the smoke check and bench.py's
 * CPU-baseline leg); the product never links it.
 *
 * CPU restatement of the Chinese-remainder split-BF16 FP64-accurate GEMM the
 * library computes (Ozaki scheme II: Ozaki, Uchino, Imamura, "Ozaki Scheme II:
 * A GEMM-oriented emulation of floating-point matrix multiplication using an
 * integer modular technique", 2025, here with BF16 residues). The reference
 * has no counterpart. This file fixes the exact result so the device output
 * can be compared bit for bit; tests/test_oracle.py pins it against
 * exact-rational golden vectors (tests/golden/gen_golden.py, Python integers),
 * the double-double oracle and the moduli/exponent invariants.
 *
 * Algorithm, for N moduli p_1..p_N (the N largest pairwise-coprime odd
 * integers <= 511, largest first; P = p_1 ... p_N):
 *   kb   = ceil(log2(max(k, 1)))
 *   beta = min(53, floor((bitlen(P) - 2 - kb) / 2), floor((125 - kb) / 2))
 *   row i of A: E_i with max_p |a_ip| < 2^E_i (frexp; 0 for a zero row);
 *     A'_ip = rint(a_ip * 2^(beta - E_i)), an integer with |A'_ip| <= 2^beta.
 *   column j of B likewise: F_j, B'_pj.
 *   C'_ij = sum_p A'_ip B'_pj, exactly; |C'_ij| <= k 2^(2 beta) < P/2 and
 *     <= 2^125, so C' is fixed by its residues modulo p_1..p_N.
 *   v = alpha * ldexp((double)C'_ij, E_i + F_j - 2 beta), the conversion
 *     rounding to nearest (ties to even);
 *   C_ij = beta_C == 0 ? v : fma(beta_C, C0_ij, v).
 * A row of A or column of B holding Inf or NaN is not converted: its entries
 * of C take v = alpha * (fma chain over p = 0..k-1 from 0 of a_ip b_pj).
 *
 * The device computes each residue C' mod p_l with one BF16 matrix-core GEMM
 * of the balanced residues of A' and B' and recovers C' by the Chinese
 * remainder theorem; here C' comes straight from 128-bit integer sums, so
 * equality with the device checks the whole residue path. The INT8 form
 * (oracle_crt_i8_*) takes the largest pairwise-coprime odd moduli <= 255 and
 * the same beta rule over their product; its result is the same function of
 * beta.
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

static const int kModuli[ORACLE_CRT_MAX_MODULI] = {511, 509, 507, 505, 503, 499, 493, 491,
                                                   487, 479, 473, 467, 463, 461, 457, 449};
/* The INT8 residue form: the largest pairwise-coprime odd integers <= 255,
 * whose balanced residues (|r| <= 127) are int8 values. */
static const int kModuliI8[ORACLE_CRT_MAX_MODULI] = {255, 253, 251, 247, 241, 239, 233, 229,
                                                     227, 223, 217, 211, 199, 197, 193, 191};

int oracle_crt_modulus(int l) { return (l >= 0 && l < ORACLE_CRT_MAX_MODULI) ? kModuli[l] : 0; }
int oracle_crt_i8_modulus(int l) {
  return (l >= 0 && l < ORACLE_CRT_MAX_MODULI) ? kModuliI8[l] : 0;
}

/* Bit length of prod_{l<N} p_l, exactly, in 32-bit limbs (16 moduli: 143 bits). */
static int product_bitlen(int N, const int* mods) {
  uint32_t w[8] = {1, 0, 0, 0, 0, 0, 0, 0};
  for (int l = 0; l < N; ++l) {
    uint64_t carry = 0;
    for (int i = 0; i < 8; ++i) {
      const uint64_t t = (uint64_t)w[i] * (uint64_t)mods[l] + carry;
      w[i] = (uint32_t)t;
      carry = t >> 32;
    }
  }
  for (int i = 7; i >= 0; --i)
    if (w[i]) return 32 * i + (32 - __builtin_clz(w[i]));
  return 0;
}

static int beta_for(int64_t k, int N, const int* mods) {
  if (N < 1 || N > ORACLE_CRT_MAX_MODULI) return -1;
  int kb = 0;
  while (((int64_t)1 << kb) < k) ++kb;
  const int room = product_bitlen(N, mods) - 2 - kb;
  int b = room >= 0 ? room / 2 : -1;
  const int cap = (125 - kb) / 2;
  if (b > cap) b = cap;
  if (b > 53) b = 53;
  return b;
}

int oracle_crt_beta(int64_t k, int N) { return beta_for(k, N, kModuli); }

/* The INT8 form's INT32 sums are exact while k 127^2 < 2^31. */
int oracle_crt_i8_beta(int64_t k, int N) {
  return k > 131072 ? -1 : beta_for(k, N, kModuliI8);
}

/* Integer images of count vectors of length k: vector v starts at X + v*sv,
 * elements se apart. out[v*k + p] = rint(x * 2^(beta - E_v)); E[v] is the
 * exponent, or ORACLE_NONFINITE for a vector holding Inf or NaN (its image is 0). */
void oracle_crt_integers(int64_t count, int64_t k, const double* X, int64_t sv, int64_t se,
                         int beta, int64_t* out, int* E) {
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < count; ++v) {
    const double* x = X + v * sv;
    double mu = 0.0;
    int finite = 1;
    for (int64_t p = 0; p < k; ++p) {
      const double a = fabs(x[p * se]);
      if (!(a <= 1.7976931348623157e308)) finite = 0;
      else if (a > mu) mu = a;
    }
    int e = 0;
    if (finite && mu > 0.0) frexp(mu, &e);
    E[v] = finite ? e : ORACLE_NONFINITE;
    for (int64_t p = 0; p < k; ++p)
      out[v * k + p] = finite ? (int64_t)rint(ldexp(x[p * se], beta - e)) : 0;
  }
}

/* Balanced residue of x modulo odd p: in [-(p-1)/2, (p-1)/2]. */
int oracle_crt_balanced(int64_t x, int p) {
  int64_t r = x % p;
  if (r > (p - 1) / 2) r -= p;
  if (r < -(p - 1) / 2) r += p;
  return (int)r;
}

static double int128_to_double(__int128 x) { return (double)x; }

/* The result of either residue form is fixed by beta alone (C' is the exact
 * product of the beta-bit integer images): both entry points below end here. */
static int crt_gemm_bits(int64_t m, int64_t n, int64_t k, const double* alpha, const double* A,
                         int64_t rsA, int64_t csA, const double* B, int64_t rsB, int64_t csB,
                         const double* betaC, double* C, int64_t rsC, int64_t csC, int beta) {
  if (m <= 0 || n <= 0) return 0;
  if (beta < 1) return -1;
  const int64_t kk = k > 0 ? k : 1;
  int64_t* ia = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m * kk));
  int64_t* ib = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n * kk));
  int* ea = (int*)malloc(sizeof(int) * (size_t)m);
  int* eb = (int*)malloc(sizeof(int) * (size_t)n);
  if (!ia || !ib || !ea || !eb) {
    free(ia);
    free(ib);
    free(ea);
    free(eb);
    return -2;
  }
  oracle_crt_integers(m, k, A, rsA, csA, beta, ia, ea); /* rows of A    */
  oracle_crt_integers(n, k, B, csB, rsB, beta, ib, eb); /* columns of B */
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < m; ++i) {
      double v;
      if (ea[i] == ORACLE_NONFINITE || eb[j] == ORACLE_NONFINITE) {
        double acc = 0.0;
        for (int64_t p = 0; p < k; ++p) acc = __builtin_fma(A[i * rsA + p * csA], B[p * rsB + j * csB], acc);
        v = *alpha * acc;
      } else {
        __int128 acc = 0;
        const int64_t* a = ia + i * kk;
        const int64_t* b = ib + j * kk;
        for (int64_t p = 0; p < k; ++p) acc += (__int128)a[p] * b[p];
        v = *alpha * ldexp(int128_to_double(acc), ea[i] + eb[j] - 2 * beta);
      }
      double* c = &C[i * rsC + j * csC];
      *c = (*betaC == 0.0) ? v : __builtin_fma(*betaC, *c, v);
    }
  free(ia);
  free(ib);
  free(ea);
  free(eb);
  return 0;
}

int oracle_crt_gemm(int64_t m, int64_t n, int64_t k, const double* alpha, const double* A,
                    int64_t rsA, int64_t csA, const double* B, int64_t rsB, int64_t csB,
                    const double* betaC, double* C, int64_t rsC, int64_t csC, int N) {
  return crt_gemm_bits(m, n, k, alpha, A, rsA, csA, B, rsB, csB, betaC, C, rsC, csC,
                       oracle_crt_beta(k, N));
}

/* The INT8 residue form (N moduli of kModuliI8). */
int oracle_crt_i8_gemm(int64_t m, int64_t n, int64_t k, const double* alpha, const double* A,
                       int64_t rsA, int64_t csA, const double* B, int64_t rsB, int64_t csB,
                       const double* betaC, double* C, int64_t rsC, int64_t csC, int N) {
  return crt_gemm_bits(m, n, k, alpha, A, rsA, csA, B, rsB, csB, betaC, C, rsC, csC,
                       oracle_crt_i8_beta(k, N));
}
