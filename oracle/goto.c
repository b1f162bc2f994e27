/*
 * ORACLE — test infrastructure only. Nothing in the product links or calls
 * this file; tests/, the smoke check and bench.py's CPU-baseline leg do.
 *
 * CPU restatement, in plain C, of the reference's Goto GEMM path:
 *   SpawnThreads -> PartN<3000> -> PartK<192> -> PackB<ColumnPanelMatrix<8>>
 *   -> PartM<120> -> PackA<RowPanelMatrix<6>> -> ParallelN<8,TheRest>
 *   -> KernelNM<8,6> -> 6x8 micro-kernel
 * (composition: src/exper/l3_packing.rs:58-66 with MR=6, NR=8 as in
 * src/kern/ukernel_wrapper.rs:73-90).
 *
 * Parity pins (tests/test_oracle.py):
 *   - the 6x8 micro-kernel equals a scalar fma chain in p order from zero,
 *     then fma(beta, c, alpha*acc) (ukernels/bli_gemm_haswell_asm_d6x8.c:691,
 *     :725-888, :899-911, :939-1075; beta == 0 never reads C, :930);
 *     checked bitwise against exact rational arithmetic on small cases;
 *   - the whole tree equals a scalar restatement of the same rounding order
 *     (per-Kc partial sums, beta := 1 after the first k block,
 *     src/composables/part.rs:96-110) bitwise, and the double-double oracle
 *     within gamma_{k+1}.
 */
#define _GNU_SOURCE /* pthread_setaffinity_np, CPU_SET */
#include "oracle.h"

#include <immintrin.h>
#include <pthread.h>
#include <sched.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#define NC 3000
#define KC 192
#define MC 120
#define MR 6
#define NR 8

/* ---- 6x8 micro-kernel -------------------------------------------------
 * a: packed row panel, MR contiguous per k (src/matrix/row_panel.rs:72)
 * b: packed column panel, NR contiguous per k (src/matrix/column_panel.rs:73)
 * C[6x8] := beta*C + alpha*(sum_p a_p b_p^T), element-wise fma chain. */
static void ukernel_6x8(int64_t k, double alpha, const double* a, const double* b,
                        double beta, double* c, int64_t rs_c, int64_t cs_c) {
  /* twelve accumulators held in registers (the loops over the six rows are
   * written out: at -O2 gcc keeps a looped array of them on the stack), the
   * k loop unrolled by four with the panels prefetched a few steps ahead, as
   * the reference kernel does (d6x8.c:725-843); each element is still one
   * fma chain in p order */
  __m256d c00 = _mm256_setzero_pd(), c01 = c00, c10 = c00, c11 = c00, c20 = c00, c21 = c00;
  __m256d c30 = c00, c31 = c00, c40 = c00, c41 = c00, c50 = c00, c51 = c00;
#define MOMMS_STEP(p)                                              \
  do {                                                             \
    const __m256d b0 = _mm256_loadu_pd(b + (p) * NR);              \
    const __m256d b1 = _mm256_loadu_pd(b + (p) * NR + 4);          \
    __m256d ai = _mm256_broadcast_sd(a + (p) * MR + 0);            \
    c00 = _mm256_fmadd_pd(ai, b0, c00);                            \
    c01 = _mm256_fmadd_pd(ai, b1, c01);                            \
    ai = _mm256_broadcast_sd(a + (p) * MR + 1);                    \
    c10 = _mm256_fmadd_pd(ai, b0, c10);                            \
    c11 = _mm256_fmadd_pd(ai, b1, c11);                            \
    ai = _mm256_broadcast_sd(a + (p) * MR + 2);                    \
    c20 = _mm256_fmadd_pd(ai, b0, c20);                            \
    c21 = _mm256_fmadd_pd(ai, b1, c21);                            \
    ai = _mm256_broadcast_sd(a + (p) * MR + 3);                    \
    c30 = _mm256_fmadd_pd(ai, b0, c30);                            \
    c31 = _mm256_fmadd_pd(ai, b1, c31);                            \
    ai = _mm256_broadcast_sd(a + (p) * MR + 4);                    \
    c40 = _mm256_fmadd_pd(ai, b0, c40);                            \
    c41 = _mm256_fmadd_pd(ai, b1, c41);                            \
    ai = _mm256_broadcast_sd(a + (p) * MR + 5);                    \
    c50 = _mm256_fmadd_pd(ai, b0, c50);                            \
    c51 = _mm256_fmadd_pd(ai, b1, c51);                            \
  } while (0)
  int64_t p = 0;
  for (; p + 4 <= k; p += 4) {
    _mm_prefetch((const char*)(a + (p + 8) * MR), _MM_HINT_T0);
    _mm_prefetch((const char*)(b + (p + 8) * NR), _MM_HINT_T0);
    _mm_prefetch((const char*)(b + (p + 10) * NR), _MM_HINT_T0);
    MOMMS_STEP(p);
    MOMMS_STEP(p + 1);
    MOMMS_STEP(p + 2);
    MOMMS_STEP(p + 3);
  }
  for (; p < k; ++p) MOMMS_STEP(p);
#undef MOMMS_STEP
  __m256d acc[MR][2] = {{c00, c01}, {c10, c11}, {c20, c21}, {c30, c31}, {c40, c41}, {c50, c51}};
  const __m256d va = _mm256_set1_pd(alpha);
  double t[MR][NR];
  for (int i = 0; i < MR; ++i) {
    _mm256_storeu_pd(&t[i][0], _mm256_mul_pd(va, acc[i][0]));
    _mm256_storeu_pd(&t[i][4], _mm256_mul_pd(va, acc[i][1]));
  }
  if (beta == 0.0) {
    for (int i = 0; i < MR; ++i)
      for (int j = 0; j < NR; ++j) c[i * rs_c + j * cs_c] = t[i][j];
  } else {
    for (int i = 0; i < MR; ++i)
      for (int j = 0; j < NR; ++j) {
        double* cp = &c[i * rs_c + j * cs_c];
        *cp = __builtin_fma(beta, *cp, t[i][j]);
      }
  }
}

void oracle_ukernel_6x8(int64_t k, const double* alpha, const double* a, const double* b,
                        const double* beta, double* c, int64_t rs_c, int64_t cs_c) {
  ukernel_6x8(k, *alpha, a, b, *beta, c, rs_c, cs_c);
}

/* ---- thread team: the reference's ThreadComm barrier (src/thread_comm.rs:50-66),
 * a sense-reversing spin barrier, with acquire/release ordering instead of
 * Relaxed. (A futex-based pthread barrier costs tens of microseconds per
 * episode at 16 threads, and the tree passes one about 18,000 times at
 * m = k = 8192, n = 1024.) */
typedef struct {
  int n;
  _Atomic int arrived;
  _Atomic int sense;
} team_t;

typedef struct {
  team_t* team;
  int tid;
  int64_t m, n, k;
  double alpha, beta;
  const double* A;
  int64_t rsA, csA;
  const double* B;
  int64_t rsB, csB;
  double* C;
  int64_t rsC, csC;
  double* bpack; /* shared (ceil(NC/NR)+1)*NR*KC */
  double* apack; /* shared (ceil(MC/MR)+1)*MR*KC */
  int cpu;       /* core to pin this thread to, or -1 */
} job_t;

/* One thread per core, as SpawnThreads pins its pool with hwloc
 * (src/composables/spawn.rs:20-35). */
static void pin_to(int cpu) {
  if (cpu < 0) return;
  cpu_set_t set;
  CPU_ZERO(&set);
  CPU_SET(cpu, &set);
  pthread_setaffinity_np(pthread_self(), sizeof(set), &set);
}

static void barrier(team_t* t) {
  if (t->n <= 1) return;
  const int my = !atomic_load_explicit(&t->sense, memory_order_relaxed);
  if (atomic_fetch_add_explicit(&t->arrived, 1, memory_order_acq_rel) == t->n - 1) {
    atomic_store_explicit(&t->arrived, 0, memory_order_relaxed);
    atomic_store_explicit(&t->sense, my, memory_order_release);
  } else {
    while (atomic_load_explicit(&t->sense, memory_order_acquire) != my) _mm_pause();
  }
}

/* Packer<Matrix, ColumnPanelMatrix<8>> (src/composables/pack.rs:51-90), with
 * zero padding of the last panel instead of reading past the matrix edge. */
static void pack_b(const job_t* j, int64_t k0, int64_t kc, int64_t n0, int64_t nc) {
  const int64_t panels = (nc + NR - 1) / NR;
  const int nt = j->team->n;
  const int64_t per = (panels + nt - 1) / nt;
  const int64_t p0 = per * j->tid, p1 = p0 + per < panels ? p0 + per : panels;
  for (int64_t p = p0; p < p1; ++p) {
    double* dst = j->bpack + p * NR * kc;
    for (int64_t y = 0; y < kc; ++y)
      for (int i = 0; i < NR; ++i) {
        const int64_t x = p * NR + i;
        dst[y * NR + i] = x < nc ? j->B[(k0 + y) * j->rsB + (n0 + x) * j->csB] : 0.0;
      }
  }
}

/* Packer<Matrix, RowPanelMatrix<6>> (src/composables/pack.rs:94-133). */
static void pack_a(const job_t* j, int64_t m0, int64_t mc, int64_t k0, int64_t kc) {
  const int64_t panels = (mc + MR - 1) / MR;
  const int nt = j->team->n;
  const int64_t per = (panels + nt - 1) / nt;
  const int64_t p0 = per * j->tid, p1 = p0 + per < panels ? p0 + per : panels;
  for (int64_t p = p0; p < p1; ++p) {
    double* dst = j->apack + p * MR * kc;
    for (int64_t x = 0; x < kc; ++x)
      for (int i = 0; i < MR; ++i) {
        const int64_t y = p * MR + i;
        dst[x * MR + i] = y < mc ? j->A[(m0 + y) * j->rsA + (k0 + x) * j->csA] : 0.0;
      }
  }
}

/* KernelNM<8,6> (src/kern/kernel_nm.rs:31-106) over this thread's NR-aligned
 * share of the nc columns (ParallelN<8,TheRest>, src/composables/parallel_range.rs:130-154). */
static void kernel_nm(const job_t* j, int64_t m0, int64_t mc, int64_t n0, int64_t nc,
                      int64_t kc, double beta) {
  const int64_t panels = (nc + NR - 1) / NR;
  const int nt = j->team->n;
  const int64_t per = (panels + nt - 1) / nt;
  const int64_t p0 = per * j->tid, p1 = p0 + per < panels ? p0 + per : panels;
  double tmp[MR * NR];
  for (int64_t jp = p0; jp < p1; ++jp) {
    const int64_t jr = jp * NR;
    const double* bp = j->bpack + jp * NR * kc;
    for (int64_t ir = 0; ir < mc; ir += MR) {
      const double* ap = j->apack + (ir / MR) * MR * kc;
      double* cp = j->C + (m0 + ir) * j->rsC + (n0 + jr) * j->csC;
      if (nc - jr >= NR && mc - ir >= MR) {
        ukernel_6x8(kc, j->alpha, ap, bp, beta, cp, j->rsC, j->csC);
      } else {
        /* Edge tile: beta = 0 into a row-major temp, then c = tau + beta*c
         * unfused (kernel_nm.rs:77-95). */
        ukernel_6x8(kc, j->alpha, ap, bp, 0.0, tmp, NR, 1);
        const int64_t lm = mc - ir < MR ? mc - ir : MR;
        const int64_t ln = nc - jr < NR ? nc - jr : NR;
        for (int64_t ii = 0; ii < lm; ++ii)
          for (int64_t jj = 0; jj < ln; ++jj) {
            double* c = cp + ii * j->rsC + jj * j->csC;
            const double bc = beta * *c;
            *c = tmp[ii * NR + jj] + bc;
          }
      }
    }
  }
}

static void* goto_thread(void* arg) {
  job_t* j = (job_t*)arg;
  if (j->tid > 0) pin_to(j->cpu);
  for (int64_t n0 = 0; n0 < j->n; n0 += NC) {          /* PartN<3000> */
    const int64_t nc = j->n - n0 < NC ? j->n - n0 : NC;
    double beta = j->beta;
    for (int64_t k0 = 0; k0 < j->k; k0 += KC) {        /* PartK<192> */
      const int64_t kc = j->k - k0 < KC ? j->k - k0 : KC;
      barrier(j->team);
      pack_b(j, k0, kc, n0, nc);                       /* PackB */
      barrier(j->team);
      for (int64_t m0 = 0; m0 < j->m; m0 += MC) {      /* PartM<120> */
        const int64_t mc = j->m - m0 < MC ? j->m - m0 : MC;
        barrier(j->team);
        pack_a(j, m0, mc, k0, kc);                     /* PackA */
        barrier(j->team);
        kernel_nm(j, m0, mc, n0, nc, kc, beta);
      }
      beta = 1.0;                                      /* part.rs:105 */
    }
    if (j->k == 0) {
      /* No k block ran: the reference leaves C untouched (PartK never calls
       * its child). Scale by beta here, as BLAS does, only if beta != 1. */
      if (j->beta != 1.0 && j->tid == 0)
        for (int64_t x = 0; x < nc; ++x)
          for (int64_t y = 0; y < j->m; ++y) {
            double* c = j->C + y * j->rsC + (n0 + x) * j->csC;
            *c = j->beta == 0.0 ? 0.0 : j->beta * *c;
          }
    }
  }
  barrier(j->team);
  return NULL;
}

int oracle_goto_dgemm(int64_t m, int64_t n, int64_t k, const double* alpha, const double* A,
                      int64_t rsA, int64_t csA, const double* B, int64_t rsB, int64_t csB,
                      const double* beta, double* C, int64_t rsC, int64_t csC, int nthreads) {
  return oracle_goto_dgemm_pinned(m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C, rsC, csC,
                                  nthreads, NULL);
}

int oracle_goto_dgemm_pinned(int64_t m, int64_t n, int64_t k, const double* alpha,
                             const double* A, int64_t rsA, int64_t csA, const double* B,
                             int64_t rsB, int64_t csB, const double* beta, double* C,
                             int64_t rsC, int64_t csC, int nthreads, const int* cpus) {
  if (m <= 0 || n <= 0) return 0;
  if (nthreads < 1) nthreads = 1;
  team_t team;
  team.n = nthreads;
  atomic_init(&team.arrived, 0);
  atomic_init(&team.sense, 0);
  const size_t bcap = (size_t)((NC + NR - 1) / NR + 1) * NR * KC;
  const size_t acap = (size_t)((MC + MR - 1) / MR + 1) * MR * KC;
  double* bpack = (double*)aligned_alloc(4096, bcap * sizeof(double));
  double* apack = (double*)aligned_alloc(4096, acap * sizeof(double));
  if (!bpack || !apack) {
    free(bpack);
    free(apack);
    return 1;
  }
  job_t* jobs = (job_t*)calloc((size_t)nthreads, sizeof(job_t));
  pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
  for (int t = 0; t < nthreads; ++t) {
    job_t* j = &jobs[t];
    j->team = &team;
    j->tid = t;
    j->m = m; j->n = n; j->k = k;
    j->alpha = *alpha; j->beta = *beta;
    j->A = A; j->rsA = rsA; j->csA = csA;
    j->B = B; j->rsB = rsB; j->csB = csB;
    j->C = C; j->rsC = rsC; j->csC = csC;
    j->bpack = bpack; j->apack = apack;
    j->cpu = cpus ? cpus[t] : -1;
  }
  cpu_set_t saved;
  const int restore = cpus && pthread_getaffinity_np(pthread_self(), sizeof(saved), &saved) == 0;
  for (int t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, goto_thread, &jobs[t]);
  if (cpus) pin_to(cpus[0]);
  goto_thread(&jobs[0]);
  for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
  if (restore) pthread_setaffinity_np(pthread_self(), sizeof(saved), &saved);
  free(jobs);
  free(th);
  free(bpack);
  free(apack);
  return 0;
}

/* Scalar restatement of the same rounding order, with no packing or vector
 * code: for each k block of KC, acc = fma chain from zero in p order,
 * s = alpha*acc, c = fma(beta_blk, c, s) for full 6x8 tiles of the Goto
 * partition and c = s + beta_blk*c for edge tiles; beta_blk = beta for the
 * first block and 1 afterwards. Independent of goto_thread's data movement. */
void oracle_goto_scalar(int64_t m, int64_t n, int64_t k, const double* alpha, const double* A,
                        int64_t rsA, int64_t csA, const double* B, int64_t rsB, int64_t csB,
                        const double* beta, double* C, int64_t rsC, int64_t csC) {
  for (int64_t x = 0; x < n; ++x) {
    const int64_t n0 = (x / NC) * NC;
    const int64_t nc = n - n0 < NC ? n - n0 : NC;
    const int64_t jr = ((x - n0) / NR) * NR;
    const int full_n = nc - jr >= NR;
    for (int64_t y = 0; y < m; ++y) {
      const int64_t m0 = (y / MC) * MC;
      const int64_t mc = m - m0 < MC ? m - m0 : MC;
      const int64_t ir = ((y - m0) / MR) * MR;
      const int full = full_n && (mc - ir >= MR);
      double* c = &C[y * rsC + x * csC];
      double bb = *beta;
      for (int64_t k0 = 0; k0 < k; k0 += KC) {
        const int64_t kc = k - k0 < KC ? k - k0 : KC;
        double acc = 0.0;
        for (int64_t p = 0; p < kc; ++p)
          acc = __builtin_fma(A[y * rsA + (k0 + p) * csA], B[(k0 + p) * rsB + x * csB], acc);
        const double s = *alpha * acc;
        if (full) *c = bb == 0.0 ? s : __builtin_fma(bb, *c, s);
        else { const double bc = bb * *c; *c = s + bc; }
        bb = 1.0;
      }
    }
  }
}

/* TripleLoop (src/composables/triple_loop.rs:8-17): x, z, y loop order,
 * c = a*b + c with separate rounding of the product (no contraction). */
void oracle_triple_loop(int64_t m, int64_t n, int64_t k, const double* A, int64_t rsA,
                        int64_t csA, const double* B, int64_t rsB, int64_t csB, double* C,
                        int64_t rsC, int64_t csC) {
  for (int64_t x = 0; x < n; ++x)
    for (int64_t z = 0; z < k; ++z)
      for (int64_t y = 0; y < m; ++y) {
        volatile double prod = A[y * rsA + z * csA] * B[z * rsB + x * csB];
        C[y * rsC + x * csC] = prod + C[y * rsC + x * csC];
      }
}
