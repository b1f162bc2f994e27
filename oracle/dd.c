/*
 * ORACLE — test infrastructure only (tests/, the smoke check and bench.py's
 * CPU-baseline leg); the product never links it.
 *
 * Accurate reference values for C* = beta*C0 + alpha*A*B, entry by entry, in
 * double-double arithmetic: Dot2 of Ogita, Rump and Oishi, "Accurate sum and
 * dot product", SIAM J. Sci. Comput. 26(6), 2005, Algorithm 5.3, built from
 * TwoProduct (one fma) and TwoSum (Knuth). The result is as accurate as if
 * the dot product were computed in twice the working precision and then
 * rounded: |res - x'y| <= u|x'y| + gamma_n^2 |x|'|y|.
 *
 * The reference has no accuracy oracle of its own; its only check is the
 * randomized residual of src/util.rs:55-85 (restated in oracle_residual).
 * Pins: exact rational arithmetic on small inputs (tests/golden), and closed
 * forms (integer-valued inputs, whose products are exact in doubles).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>

static inline void two_sum(double a, double b, double* s, double* e) {
  const double x = a + b;
  const double z = x - a;
  *e = (a - (x - z)) + (b - z);
  *s = x;
}

static inline void two_prod(double a, double b, double* p, double* e) {
  const double x = a * b;
  *e = __builtin_fma(a, b, -x);
  *p = x;
}

/* Accurate beta*c0 + alpha*sum_p x_p y_p, plus the magnitude
 * |alpha|*sum|x_p||y_p| + |beta||c0| used to scale componentwise errors. */
static void dd_entry(int64_t k, double alpha, const double* x, int64_t incx, const double* y,
                     int64_t incy, double beta, double c0, double* res, double* mag) {
  double p = 0.0, s = 0.0, absum = 0.0;
  for (int64_t i = 0; i < k; ++i) {
    double h, r, q;
    two_prod(x[i * incx], y[i * incy], &h, &r);
    two_sum(p, h, &p, &q);
    s += q + r;
    absum += fabs(x[i * incx] * y[i * incy]);
  }
  /* (p + s) * alpha in double-double, then + beta*c0. */
  double ph, pl;
  two_prod(alpha, p, &ph, &pl);
  pl += alpha * s;
  double bh = 0.0, bl = 0.0;
  if (beta != 0.0) two_prod(beta, c0, &bh, &bl);
  double t, te;
  two_sum(ph, bh, &t, &te);
  te += pl + bl;
  *res = t + te;
  *mag = fabs(alpha) * absum + (beta != 0.0 ? fabs(beta) * fabs(c0) : 0.0);
}

void oracle_dd_entries(int64_t k, const double* alpha, const double* A, int64_t rsA,
                       int64_t csA, const double* B, int64_t rsB, int64_t csB,
                       const double* beta, const double* C0, int64_t rsC, int64_t csC,
                       int64_t count, const int64_t* rows, const int64_t* cols, double* out,
                       double* mag) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t e = 0; e < count; ++e) {
    const int64_t i = rows[e], j = cols[e];
    const double c0 = (*beta != 0.0) ? C0[i * rsC + j * csC] : 0.0;
    dd_entry(k, *alpha, A + i * rsA, csA, B + j * csB, rsB, *beta, c0, &out[e], &mag[e]);
  }
}

void oracle_dd_gemm(int64_t m, int64_t n, int64_t k, const double* alpha, const double* A,
                    int64_t rsA, int64_t csA, const double* B, int64_t rsB, int64_t csB,
                    const double* beta, const double* C0, int64_t rsC, int64_t csC,
                    double* out, double* mag) {
  /* out / mag are column-major m x n. */
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < m; ++i) {
      const double c0 = (*beta != 0.0) ? C0[i * rsC + j * csC] : 0.0;
      dd_entry(k, *alpha, A + i * rsA, csA, B + j * csB, rsB, *beta, c0, &out[i + j * m],
               &mag[i + j * m]);
    }
}

/* Randomized residual of src/util.rs:55-85: ||C w - A (B w)||_2^2 with
 * TripleLoop-order products, extended by - beta*C0 w because the configs use
 * a random C0 (the reference's harness zero-fills C, l3_packing.rs:34). */
double oracle_residual(int64_t m, int64_t n, int64_t k, const double* A, int64_t rsA,
                       int64_t csA, const double* B, int64_t rsB, int64_t csB,
                       const double* C, const double* C0, int64_t rsC, int64_t csC,
                       double alpha, double beta, const double* w) {
  double* bw = (double*)calloc((size_t)(k > 0 ? k : 1), sizeof(double));
  double* abw = (double*)calloc((size_t)(m > 0 ? m : 1), sizeof(double));
  double* cw = (double*)calloc((size_t)(m > 0 ? m : 1), sizeof(double));
  for (int64_t z = 0; z < n; ++z)
    for (int64_t y = 0; y < k; ++y) bw[y] += B[y * rsB + z * csB] * w[z];
  for (int64_t z = 0; z < k; ++z)
    for (int64_t y = 0; y < m; ++y) abw[y] += A[y * rsA + z * csA] * bw[z];
  for (int64_t z = 0; z < n; ++z)
    for (int64_t y = 0; y < m; ++y) {
      double c = C[y * rsC + z * csC];
      if (C0 && beta != 0.0) c -= beta * C0[y * rsC + z * csC];
      cw[y] += c * w[z];
    }
  double acc = 0.0;
  for (int64_t y = 0; y < m; ++y) {
    const double d = cw[y] - alpha * abw[y];
    acc += d * d;
  }
  free(bw);
  free(abw);
  free(cw);
  return acc;
}
