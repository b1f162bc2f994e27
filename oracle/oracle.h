/*
 * ORACLE — test infrastructure only. Only tests/, the smoke check and
 * bench.py's CPU-baseline leg may load liboracle.so; the product never does.
 *
 * Functions and what pins each (see the file headers for details):
 *   goto.c   oracle_goto_dgemm     restated Goto tree + 6x8 kernel (reference path)
 *            oracle_goto_scalar    same rounding order, scalar, no packing
 *            oracle_ukernel_6x8    the micro-kernel alone
 *            oracle_triple_loop    TripleLoop node
 *   dd.c     oracle_dd_entries / oracle_dd_gemm   double-double (Dot2) values
 *            oracle_residual       randomized residual check
 *   ozaki.c  oracle_ozaki_*        split-BF16 algorithm restatement (slices)
 *   crt.c    oracle_crt_*          split-BF16 algorithm restatement (residues)
 *   hier.c   oracle_hier_*         hierarchical matrix layout, pack, unpack-add
 */
#ifndef MOMMS_ORACLE_H
#define MOMMS_ORACLE_H
#include <stdint.h>

int oracle_goto_dgemm(int64_t m, int64_t n, int64_t k, const double* alpha, const double* A,
                      int64_t rsA, int64_t csA, const double* B, int64_t rsB, int64_t csB,
                      const double* beta, double* C, int64_t rsC, int64_t csC, int nthreads);
/* oracle_goto_dgemm with thread t pinned to core cpus[t] (NULL: unpinned);
 * the caller's own affinity is restored before returning. */
int oracle_goto_dgemm_pinned(int64_t m, int64_t n, int64_t k, const double* alpha,
                             const double* A, int64_t rsA, int64_t csA, const double* B,
                             int64_t rsB, int64_t csB, const double* beta, double* C,
                             int64_t rsC, int64_t csC, int nthreads, const int* cpus);
void oracle_goto_scalar(int64_t m, int64_t n, int64_t k, const double* alpha, const double* A,
                        int64_t rsA, int64_t csA, const double* B, int64_t rsB, int64_t csB,
                        const double* beta, double* C, int64_t rsC, int64_t csC);
void oracle_ukernel_6x8(int64_t k, const double* alpha, const double* a, const double* b,
                        const double* beta, double* c, int64_t rs_c, int64_t cs_c);
void oracle_triple_loop(int64_t m, int64_t n, int64_t k, const double* A, int64_t rsA,
                        int64_t csA, const double* B, int64_t rsB, int64_t csB, double* C,
                        int64_t rsC, int64_t csC);

void oracle_dd_entries(int64_t k, const double* alpha, const double* A, int64_t rsA,
                       int64_t csA, const double* B, int64_t rsB, int64_t csB,
                       const double* beta, const double* C0, int64_t rsC, int64_t csC,
                       int64_t count, const int64_t* rows, const int64_t* cols, double* out,
                       double* mag);
void oracle_dd_gemm(int64_t m, int64_t n, int64_t k, const double* alpha, const double* A,
                    int64_t rsA, int64_t csA, const double* B, int64_t rsB, int64_t csB,
                    const double* beta, const double* C0, int64_t rsC, int64_t csC,
                    double* out, double* mag);
double oracle_residual(int64_t m, int64_t n, int64_t k, const double* A, int64_t rsA,
                       int64_t csA, const double* B, int64_t rsB, int64_t csB,
                       const double* C, const double* C0, int64_t rsC, int64_t csC,
                       double alpha, double beta, const double* w);

int oracle_ozaki_chunk(int s);
void oracle_ozaki_split(int64_t count, int64_t k, const double* X, int64_t stride_vec,
                        int64_t stride_elem, int s, double* q, int* exps);
void oracle_ozaki_gemm(int64_t m, int64_t n, int64_t k, const double* alpha, const double* A,
                       int64_t rsA, int64_t csA, const double* B, int64_t rsB, int64_t csB,
                       const double* beta, double* C, int64_t rsC, int64_t csC, int s);

/* Exponent flag of a vector holding Inf or NaN. */
#define ORACLE_NONFINITE 0x7fffffff
#define ORACLE_CRT_MAX_MODULI 16
int oracle_crt_modulus(int l);
int oracle_crt_beta(int64_t k, int N);
int oracle_crt_balanced(int64_t x, int p);
void oracle_crt_integers(int64_t count, int64_t k, const double* X, int64_t sv, int64_t se,
                         int beta, int64_t* out, int* E);
int oracle_crt_gemm(int64_t m, int64_t n, int64_t k, const double* alpha, const double* A,
                    int64_t rsA, int64_t csA, const double* B, int64_t rsB, int64_t csB,
                    const double* betaC, double* C, int64_t rsC, int64_t csC, int N);
int oracle_crt_i8_modulus(int l);
int oracle_crt_i8_beta(int64_t k, int N);
int oracle_crt_i8_gemm(int64_t m, int64_t n, int64_t k, const double* alpha, const double* A,
                       int64_t rsA, int64_t csA, const double* B, int64_t rsB, int64_t csB,
                       const double* betaC, double* C, int64_t rsC, int64_t csC, int N);

/* hierarchical layout: a layout is int64[9 + 4 * ORACLE_HIER_LEVELS] =
   [h, w, h_pad, w_pad, ny, nx, capacity, lrs, lcs, y_blksz[], y_stride[],
    x_blksz[], x_stride[]] (levels outermost first) */
#define ORACLE_HIER_LEVELS 8
int oracle_hier_layout(int64_t h, int64_t w, const int32_t* kind, const int64_t* bsz,
                       int64_t nsteps, int32_t ydim, int32_t xdim, int64_t lh, int64_t lw,
                       int64_t lrs, int64_t lcs, int64_t* out);
int64_t oracle_hier_offset(const int64_t* L, int64_t y, int64_t x);
/* out[y * w_pad + x] = offset(y, x) for y < h_pad, x < w_pad */
void oracle_hier_offsets(const int64_t* L, int64_t* out);
void oracle_hier_pack(const int64_t* L, const double* A, int64_t rs, int64_t cs, double* P);
void oracle_hier_unpack_add(const int64_t* L, const double* P, double* C, int64_t rs,
                            int64_t cs);

#endif
