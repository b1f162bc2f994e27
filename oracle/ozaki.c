/*
 * ORACLE — test infrastructure only (tests/, the smoke check and bench.py's
 * CPU-baseline leg); the product never links it.
 *
 * CPU restatement of the split-BF16 FP64-accurate GEMM the library computes
 * (Ozaki scheme I: Ozaki, Ogita, Oishi, Rump, "Error-free transformations of
 * matrix multiplication by using fast routines of matrix multiplication and
 * its applications", Numer. Algorithms 59, 2012). The reference has no
 * counterpart; this file fixes the exact arithmetic so the device result can
 * be compared bit for bit, and tests/test_oracle.py pins it against the
 * double-double oracle (accuracy) and against exact-rational and structural
 * invariants (slices reconstruct the scaled input exactly, every slice is an
 * integer that bf16 holds exactly, chunk sums stay below 2^24).
 *
 * Algorithm, for s slices:
 *   row i of A: mu = max_p |a_ip|, E_i with mu < 2^E_i (frexp); r = a * 2^-E_i;
 *     for t = 1..s: q_t = rint(r * 2^(8t)); r -= q_t * 2^(-8t).
 *     |q_1| <= 256 and |q_t| <= 128 for t > 1: each is exact in bf16.
 *   column j of B likewise with exponent F_j.
 *   C_ij = beta*C0_ij + alpha * 2^(E_i+F_j) * sum_{d=2}^{s+1} 2^(-8d) S_d,
 *     S_d = sum_{t+u=d} sum_p qa_t[i,p] qb_u[p,j].
 *   The k range is cut into chunks of kch(s) values; per chunk each S_d is an
 *   integer below 2^24 in magnitude (so an fp32 sum of it is exact in any
 *   order), and it is folded into an fp64 accumulator for d = s+1 down to 2:
 *     acc = fma(S_d, 2^(-8d), acc).
 *   Finally v = (alpha * acc) * 2^(E_i+F_j), C = beta == 0 ? v : fma(beta, C0, v).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>

int oracle_ozaki_chunk(int s) {
  /* Per k, |sum_{t+u=d} q_t q_u| <= 2*2^8*2^7 + (d-3)*2^14 = (d+1)*2^14 for
   * d >= 3 and 2^16 for d = 2; the largest level is d = s+1. */
  int kch = 16 * (1024 / (16 * (s + 2)));
  if (kch > 256) kch = 256;
  return kch;
}

/* Slices one vector of length k (stride inc): q[t*k + p] for t < s. */
static void split_vector(int64_t k, const double* v, int64_t inc, int s, double* q, int* e) {
  double mu = 0.0;
  int finite = 1;
  for (int64_t p = 0; p < k; ++p) {
    const double a = fabs(v[p * inc]);
    if (!(a <= 1.7976931348623157e308)) finite = 0;
    else if (a > mu) mu = a;
  }
  int E = 0;
  if (finite && mu > 0.0) frexp(mu, &E);
  *e = E;
  for (int64_t p = 0; p < k; ++p) {
    double r = ldexp(v[p * inc], -E);
    for (int t = 1; t <= s; ++t) {
      double qt;
      if (!finite) {
        /* A vector holding Inf or NaN is passed through unsliced: its
         * non-finite values reach every product they touch. */
        qt = (t == 1) ? v[p * inc] : 0.0;
      } else {
        qt = rint(ldexp(r, 8 * t));
        r -= ldexp(qt, -8 * t);
      }
      q[(int64_t)(t - 1) * k + p] = qt;
    }
  }
}

void oracle_ozaki_split(int64_t count, int64_t k, const double* X, int64_t stride_vec,
                        int64_t stride_elem, int s, double* q, int* exps) {
  /* count vectors of length k; vector v starts at X + v*stride_vec and its
   * elements are stride_elem apart. q is [v][t][p] (count x s x k). */
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < count; ++v)
    split_vector(k, X + v * stride_vec, stride_elem, s, q + v * (int64_t)s * k, &exps[v]);
}

void oracle_ozaki_gemm(int64_t m, int64_t n, int64_t k, const double* alpha, const double* A,
                       int64_t rsA, int64_t csA, const double* B, int64_t rsB, int64_t csB,
                       const double* beta, double* C, int64_t rsC, int64_t csC, int s) {
  if (m <= 0 || n <= 0) return;
  double* qa = (double*)malloc(sizeof(double) * (size_t)(m * s * (k > 0 ? k : 1)));
  double* qb = (double*)malloc(sizeof(double) * (size_t)(n * s * (k > 0 ? k : 1)));
  int* ea = (int*)malloc(sizeof(int) * (size_t)m);
  int* eb = (int*)malloc(sizeof(int) * (size_t)n);
  oracle_ozaki_split(m, k, A, rsA, csA, s, qa, ea);  /* rows of A    */
  oracle_ozaki_split(n, k, B, csB, rsB, s, qb, eb);  /* columns of B */
  const int kch = oracle_ozaki_chunk(s);
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < m; ++i) {
      double acc = 0.0;
      for (int64_t c0 = 0; c0 < k; c0 += kch) {
        const int64_t c1 = c0 + kch < k ? c0 + kch : k;
        for (int d = s + 1; d >= 2; --d) {
          double S = 0.0; /* integer below 2^24: exact */
          for (int t = 1; t < d; ++t) {
            const double* a = qa + (i * s + (t - 1)) * k;
            const double* b = qb + (j * s + (d - t - 1)) * k;
            for (int64_t p = c0; p < c1; ++p) S += a[p] * b[p];
          }
          acc = __builtin_fma(S, ldexp(1.0, -8 * d), acc);
        }
      }
      const double v = ldexp(*alpha * acc, ea[i] + eb[j]);
      double* c = &C[i * rsC + j * csC];
      *c = (*beta == 0.0) ? v : __builtin_fma(*beta, *c, v);
    }
  free(qa);
  free(qb);
  free(ea);
  free(eb);
}
