/*
 * momms_hip.h — C ABI of the MI355X (gfx950) GEMM engine.
 *
 * This is the boundary a momms GemmNode calls into. In the reference the
 * hot path ends in a C micro-kernel reached through Rust FFI,
 *   bli_dgemm_haswell_asm_6x8(k, alpha*, a*, b*, beta*, c*, rs_c, cs_c, ...)
 *   (ukernels/bli_gemm_haswell_asm_d6x8.c:666-676,
 *    bound at src/kern/ukernel_wrapper.rs:40-53 and :73-90),
 * called once per 6x8 tile by KernelNM (src/kern/kernel_nm.rs:31-106) under
 * the Goto tree (src/exper/l3_packing.rs:58-66). A per-tile GPU launch is not
 * viable, so these entry points replace the whole subtree below
 * GemmNode::run (src/composables/mod.rs:31-36) for Matrix<T> operands.
 * They keep that FFI's conventions:
 *   - scalars by pointer, C := beta*C + alpha*A*B, beta == 0 never reads C
 *     (d6x8.c:929-931, :1076-1193);
 *   - every operand has a row stride and a column stride (Matrix<T>'s
 *     get_row_stride/get_column_stride, src/matrix/general_stride.rs:42-43);
 *     column-major (rs == 1), row-major (cs == 1) and general strides work;
 *   - all pointers are borrowed; matrix pointers are DEVICE pointers
 *     (hipMalloc'd or otherwise GPU-visible), scalars are host pointers.
 * Return value: 0 on success, -i when argument i (1-based) is invalid (as
 * xerbla reports for the Fortran dgemm_ of src/util.rs:14-22), and a
 * positive hipError_t for a HIP runtime failure. Nothing aborts. The text of
 * the last failure on the calling thread is momms_last_error_string().
 * `stream` is a hipStream_t (NULL = the default stream); every entry point
 * only enqueues work on it and returns without synchronizing.
 */
#ifndef MOMMS_HIP_H
#define MOMMS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Text of the last error reported on this thread ("" if none). */
const char* momms_last_error_string(void);

/* Plain FP64 GEMM on FP64 matrix cores: C := beta*C + alpha*A*B, A m x k,
 * B k x n, C m x n. Replaces the Goto subtree SpawnThreads/PartN ... KernelNM
 * -> bli_dgemm_haswell_asm_6x8 (src/exper/l3_packing.rs:58-66,
 * d6x8.c:666-1221). k == 0 gives C := beta*C. */
int momms_hip_dgemm(int64_t m, int64_t n, int64_t k, const double* alpha, const double* A,
                    int64_t rsA, int64_t csA, const double* B, int64_t rsB, int64_t csB,
                    const double* beta, double* C, int64_t rsC, int64_t csC, void* stream);

/* momms_hip_dgemm with an explicit tile configuration instead of the
 * built-in choice, for tuning: 0 = 128x128 per workgroup (buffer->LDS with a
 * 2-slab ring, two workgroups per CU, when A, B, C are column-major with even
 * leading dimensions and 16-byte aligned A, B), 1 = 64x64, 2 = 128x64,
 * 3 = 128x128 register-staged, 4 = 128x128 buffer->LDS with a 4-slab ring,
 * 5 = 128x128 buffer->LDS with k split into up to 16 parts whose partial
 * products a second kernel adds in fixed order, 6-9 = persistent stream-K:
 * as many workgroups as stay resident share the (tile, 16-deep k-slab)
 * iterations evenly, and the parts of tiles split between workgroups are
 * added in k order by a second kernel (6: 64x64 tiles, three workgroups per
 * CU; 7: 128x64, two per CU; 8: 64x64, two per CU with a deeper slab ring;
 * 9: 64x64, one per CU with a six-slab ring). 5-9 use a device workspace the
 * library allocates once per stream and keeps. 10-11 = k split over the four
 * wavefronts of a workgroup, each multiplying a quarter of k into the whole
 * tile straight from global memory, the four partial tiles added in LDS in
 * wavefront order (10: 64x64 tiles, one workgroup per CU; 11: 32x32 tiles,
 * two per CU); no workspace; A and B column-major with leading dimensions
 * <= 2^21 (otherwise the register-staged tiles run). Argument 15 is tile_cfg.
 * Same semantics as momms_hip_dgemm. */
int momms_hip_dgemm_tiled(int64_t m, int64_t n, int64_t k, const double* alpha,
                          const double* A, int64_t rsA, int64_t csA, const double* B,
                          int64_t rsB, int64_t csB, const double* beta, double* C,
                          int64_t rsC, int64_t csC, int tile_cfg, void* stream);

/* momms_hip_dgemm for callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * No counterpart in the reference. Same arguments, kernels and bits as
 * momms_hip_dgemm; the stream-K form's workspace (the parts of output tiles
 * split between workgroups, the only device memory beyond A, B and C) lives
 * in a fixed 80 MiB array that each GPU the library runs on holds until the
 * process ends, instead of library scratch. That covers every shape on a
 * 256-CU MI355X; a shape that would need more returns -1. No call allocates
 * or frees memory. Nothing in the array needs to start at zero: every part
 * a call reads was written earlier in the same call. Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams,
 * other host threads) race and must be ordered by the caller. */
int momms_hip_dgemm_noalloc(int64_t m, int64_t n, int64_t k, const double* alpha,
                            const double* A, int64_t rsA, int64_t csA, const double* B,
                            int64_t rsB, int64_t csB, const double* beta, double* C,
                            int64_t rsC, int64_t csC, void* stream);

/* Small FP64 GEMM in one launch, without packing or workspace: the device
 * counterpart of the reference's libxsmm leaves Xsmm and KernelXsmmA2
 * (src/kern/kernel_xsmm.rs:8-139), which multiply a small block straight
 * from its operands when all are column-major, or all row-major (then
 * C^T = B^T A^T). One wavefront per 32 x 32 tile of C, registers only. Other
 * strides, leading dimensions over 2^21 or operands over 2^28 doubles run
 * momms_hip_dgemm. Same arguments and conventions as momms_hip_dgemm; the
 * result is within the same gamma_{k+1} bound (another summation order). */
int momms_hip_dgemm_small(int64_t m, int64_t n, int64_t k, const double* alpha, const double* A,
                          int64_t rsA, int64_t csA, const double* B, int64_t rsB, int64_t csB,
                          const double* beta, double* C, int64_t rsC, int64_t csC, void* stream);

/* ---- FP32-accumulating GEMMs ------------------------------------------------
 * C := beta*C + alpha*A*B with float C, alpha, beta; arguments as
 * momms_hip_dgemm. momms_hip_sgemm takes float A, B (FP32 matrix cores,
 * v_mfma_f32_32x32x2_f32): the counterpart of the reference's KNM sgemm leaf
 * (src/kern/knm_kernel.rs:8-81 -> sgemm_knm_asm_16x24,
 * ukernels/sgemm_knm_asm_24x16.c:30-38). momms_hip_gemm_bf16 takes A, B as
 * bfloat16 bit patterns (uint16_t) and accumulates in FP32 on BF16 matrix
 * cores (no reference counterpart). Any strides, as for momms_hip_dgemm. */
int momms_hip_sgemm(int64_t m, int64_t n, int64_t k, const float* alpha, const float* A,
                    int64_t rsA, int64_t csA, const float* B, int64_t rsB, int64_t csB,
                    const float* beta, float* C, int64_t rsC, int64_t csC, void* stream);
int momms_hip_gemm_bf16(int64_t m, int64_t n, int64_t k, const float* alpha, const uint16_t* A,
                        int64_t rsA, int64_t csA, const uint16_t* B, int64_t rsB, int64_t csB,
                        const float* beta, float* C, int64_t rsC, int64_t csC, void* stream);

/* ---- FP64-accurate GEMM from BF16 slices (Ozaki scheme I) -----------------
 * No reference counterpart: the precision addition beside the Goto path.
 * C := beta*C + alpha*A*B where each row of A and column of B is scaled by a
 * power of two and cut into `slices` (1..8) BF16 integer slices, the slice
 * products with t+u <= slices+1 run on BF16 matrix cores with exact FP32
 * accumulation per k-chunk, and the levels are folded and rescaled in FP64
 * on the GPU. The result equals, bit for bit, the CPU restatement in
 * oracle/ozaki.c. A row or column holding Inf/NaN yields non-finite outputs
 * in the entries it touches.
 * `workspace` is caller-owned device memory of at least
 * momms_hip_split_bf16_workspace_size(m, n, k, slices) bytes; it holds the
 * packed slices and exponents between the prepare and multiply stages.
 * Calls that share a workspace must be ordered on one stream. */
size_t momms_hip_split_bf16_workspace_size(int64_t m, int64_t n, int64_t k, int slices);

/* Whole split GEMM (prepare + multiply). Arguments 1-14 as momms_hip_dgemm,
 * then slices (15), workspace (16), workspace_bytes (17), stream. */
int momms_hip_dgemm_split_bf16(int64_t m, int64_t n, int64_t k, const double* alpha,
                               const double* A, int64_t rsA, int64_t csA, const double* B,
                               int64_t rsB, int64_t csB, const double* beta, double* C,
                               int64_t rsC, int64_t csC, int slices, void* workspace,
                               size_t workspace_bytes, void* stream);

/* Stage 1: exponents and packed BF16 slices of A (rows) and B (columns) into
 * the workspace: the device form of PackA/PackB (src/composables/pack.rs:311-419). */
int momms_hip_split_bf16_prepare(int64_t m, int64_t n, int64_t k, const double* A, int64_t rsA,
                                 int64_t csA, const double* B, int64_t rsB, int64_t csB,
                                 int slices, void* workspace, size_t workspace_bytes,
                                 void* stream);

/* Stage 2: the fused slice-product multiply and FP64 reconstruction into C,
 * from a workspace filled by stage 1 with the same m, n, k, slices. */
int momms_hip_split_bf16_multiply(int64_t m, int64_t n, int64_t k, const double* alpha,
                                  const double* beta, double* C, int64_t rsC, int64_t csC,
                                  int slices, const void* workspace, size_t workspace_bytes,
                                  void* stream);

/* ---- FP64-accurate GEMM from BF16 residues (Ozaki scheme II) ---------------
 * No reference counterpart: the second form of the precision addition. Each
 * row of A and column of B is scaled by a power of two and rounded to an
 * integer of beta bits (momms_hip_crt_bf16_beta); the exact integer product
 * is recovered from its residues modulo `moduli` (1..16) pairwise-coprime odd
 * numbers <= 511 (momms_hip_crt_bf16_modulus), one BF16 matrix-core GEMM with
 * exact FP32 accumulation per modulus, by the Chinese remainder theorem on
 * the GPU; one correctly rounded conversion, the rescaling and
 * C := beta*C + alpha*(...) follow in FP64. The result equals, bit for bit,
 * the CPU restatement in oracle/crt.c. A row or column holding Inf/NaN is
 * computed by an FP64 fma chain in the entries it touches.
 * Returns -15 when `moduli` is outside 1..16 or too few for k (beta < 1). */
int momms_hip_crt_bf16_beta(int64_t k, int moduli);     /* -1 if invalid */
int momms_hip_crt_bf16_modulus(int index);              /* 0 outside 0..15 */
size_t momms_hip_crt_bf16_workspace_size(int64_t m, int64_t n, int64_t k, int moduli);

/* Arguments 1-14 as momms_hip_dgemm, then moduli (15), workspace (16),
 * workspace_bytes (17), stream. The workspace is caller-owned device memory
 * of momms_hip_crt_bf16_workspace_size bytes; calls sharing one must be
 * ordered on one stream. */
int momms_hip_dgemm_crt_bf16(int64_t m, int64_t n, int64_t k, const double* alpha,
                             const double* A, int64_t rsA, int64_t csA, const double* B,
                             int64_t rsB, int64_t csB, const double* beta, double* C,
                             int64_t rsC, int64_t csC, int moduli, void* workspace,
                             size_t workspace_bytes, void* stream);

/* Staged form of momms_hip_dgemm_crt_bf16, for callers that time or overlap
 * the steps; calls on one workspace with the same m, n, k, moduli must be
 * ordered on one stream. Stage 1: exponents and packed BF16 residues of A
 * (rows) and B (columns), the device form of PackA/PackB
 * (src/composables/pack.rs:311-419). Arguments: m n k A rsA csA B rsB csB
 * moduli(10) workspace(11) workspace_bytes(12) stream. */
int momms_hip_crt_bf16_prepare(int64_t m, int64_t n, int64_t k, const double* A, int64_t rsA,
                               int64_t csA, const double* B, int64_t rsB, int64_t csB,
                               int moduli, void* workspace, size_t workspace_bytes,
                               void* stream);
/* Stage 2: the residue GEMMs, one per modulus, into the workspace.
 * Arguments: m n k moduli(4) workspace(5) workspace_bytes(6) stream. */
int momms_hip_crt_bf16_products(int64_t m, int64_t n, int64_t k, int moduli, void* workspace,
                                size_t workspace_bytes, void* stream);
/* Stage 3: Chinese remainder reconstruction and C := beta*C + alpha*(...).
 * A and B are read only for rows / columns holding Inf or NaN. Arguments as
 * momms_hip_dgemm_crt_bf16 (moduli 15, workspace 16, workspace_bytes 17). */
int momms_hip_crt_bf16_reconstruct(int64_t m, int64_t n, int64_t k, const double* alpha,
                                   const double* A, int64_t rsA, int64_t csA, const double* B,
                                   int64_t rsB, int64_t csB, const double* beta, double* C,
                                   int64_t rsC, int64_t csC, int moduli, const void* workspace,
                                   size_t workspace_bytes, void* stream);

/* ---- FP64-accurate GEMM from INT8 residues (Ozaki scheme II on INT8) ----------
 * No reference counterpart: the residue form above on INT8 matrix cores
 * (v_mfma_i32_32x32x32_i8, twice the BF16 rate). The moduli are the N
 * (1..16) largest pairwise-coprime odd integers <= 255
 * (momms_hip_crt_i8_modulus), whose balanced residues are int8 values; the
 * integer width beta follows the same rule over their product
 * (momms_hip_crt_i8_beta: 15 moduli give the 51 bits of 13 BF16 moduli at
 * k = 8192), and the INT32 sums are exact without reduction for
 * k <= 131072. The result is the same function of beta as the BF16 form's,
 * and equals, bit for bit, oracle/crt.c's INT8 restatement. Arguments,
 * workspace rules and return codes as the BF16 residue form; k > 131072
 * returns -3. */
int momms_hip_crt_i8_beta(int64_t k, int moduli);     /* -1 if invalid */
int momms_hip_crt_i8_modulus(int index);              /* 0 outside 0..15 */
size_t momms_hip_crt_i8_workspace_size(int64_t m, int64_t n, int64_t k, int moduli);
int momms_hip_dgemm_crt_i8(int64_t m, int64_t n, int64_t k, const double* alpha,
                           const double* A, int64_t rsA, int64_t csA, const double* B,
                           int64_t rsB, int64_t csB, const double* beta, double* C,
                           int64_t rsC, int64_t csC, int moduli, void* workspace,
                           size_t workspace_bytes, void* stream);
/* Staged form, as momms_hip_crt_bf16_prepare / _products / _reconstruct. */
int momms_hip_crt_i8_prepare(int64_t m, int64_t n, int64_t k, const double* A, int64_t rsA,
                             int64_t csA, const double* B, int64_t rsB, int64_t csB,
                             int moduli, void* workspace, size_t workspace_bytes, void* stream);
int momms_hip_crt_i8_products(int64_t m, int64_t n, int64_t k, int moduli, void* workspace,
                              size_t workspace_bytes, void* stream);
int momms_hip_crt_i8_reconstruct(int64_t m, int64_t n, int64_t k, const double* alpha,
                                 const double* A, int64_t rsA, int64_t csA, const double* B,
                                 int64_t rsB, int64_t csB, const double* beta, double* C,
                                 int64_t rsC, int64_t csC, int moduli, const void* workspace,
                                 size_t workspace_bytes, void* stream);

/* ---- batched uneven GEMMs -----------------------------------------------------
 * No reference counterpart: the batching addition. Problem i of a batch is
 * C_i := beta*C_i + alpha*A_i*B_i with column-major operands at element
 * offsets into three device arenas A, B, C (leading dimensions lda >= m,
 * ldb >= k, ldc >= m). Output regions must not overlap. */
typedef struct momms_gemm_desc {
  int64_t m, n, k;
  int64_t a_off, b_off, c_off;
  int64_t lda, ldb, ldc;
} momms_gemm_desc;

typedef struct momms_batch_plan momms_batch_plan;

/* Builds a plan for `count` problems (descs is a HOST array, copied).
 * mode 0: plain FP64 matrix cores; mode 1: split-BF16 FP64-accurate with
 * `slices` slices (0 = chosen by the autotuner: the smallest count whose
 * sampled error is within 2x plain FP64's, 7 without autotuning);
 * mode 2: the BF16 residue form with `slices` moduli (0 = chosen the same
 * way, 13 without autotuning; at least the count momms_hip_crt_bf16_beta
 * needs for the largest k of the batch), every problem's result bitwise
 * equal to momms_hip_dgemm_crt_bf16 with that many moduli, in one fused
 * launch (residues, products and reconstruction on chip per tile) after an
 * exponent pre-pass; lda and ldb of every problem at most 2^21 (-1 otherwise).
 * autotune != 0 times the candidate tile shapes per size bucket on this
 * batch (reading A and B, writing only plan-owned scratch) and remembers the
 * choice for the process. A and B may be NULL when autotune == 0.
 * The plan owns device tables and, in mode 1, the slice workspace. */
int momms_hip_batch_plan_create(const momms_gemm_desc* descs, int64_t count, int mode,
                                int slices, int autotune, const double* A, const double* B,
                                momms_batch_plan** out, void* stream);

/* Enqueues the whole batch on `stream` (reset kernel, slicing in mode 1,
 * persistent multiply kernels). Every run of a plan resets and uses the
 * plan's own device queue counters and workspace, so runs and graph launches
 * of one plan must be ordered on one stream; overlapping ones would skip or
 * repeat tiles. */
int momms_hip_batch_run(momms_batch_plan* plan, const double* alpha, const double* A,
                        const double* B, const double* beta, double* C, void* stream);

/* Captures the same sequence once into a hipGraph owned by the plan (on a
 * plan-owned stream; the captured graph holds kernel nodes only), for
 * momms_hip_batch_launch. Re-capturing replaces the graph. */
int momms_hip_batch_capture(momms_batch_plan* plan, const double* alpha, const double* A,
                            const double* B, const double* beta, double* C);

/* One hipGraphLaunch of the captured batch on `stream`. */
int momms_hip_batch_launch(momms_batch_plan* plan, void* stream);

/* JSON text describing the plan's choices (tile per bucket, slices). */
const char* momms_hip_batch_plan_describe(const momms_batch_plan* plan);

int momms_hip_batch_plan_destroy(momms_batch_plan* plan);

/* ---- the autotuner's table ---------------------------------------------------
 * No reference counterpart: the reference fixes its blocking at compile time
 * (Cargo features, Cargo.toml:59-66, and typenum block sizes,
 * src/exper/l3.rs:78-93); the batched planner chooses at run time and keeps
 * its choices in a per-process table, one entry per (kind, batch class).
 * Exporting the table and importing it into a later process lets that
 * process plan without timing candidates again; an imported entry is used by
 * momms_hip_batch_plan_create with or without autotune.
 *
 * The class of a batch (momms_hip_batch_class) is
 * 16 * min(15, floor(log2(1 + P))) + min(8, floor(log2(1 + K))), P being the
 * number of problems with m, n > 0 and K their mean k.
 *
 * Entry values by kind:
 *   MOMMS_TUNE_BATCH_FP64:  {T, T2, 0, 0}: problems whose smaller side is at
 *     least T2 take 128 x 128 tiles, at least T 64 x 64 tiles, the others
 *     32 x 32 one-wavefront tiles; 16 <= T <= T2 <= 2^30 (2^30: none).
 *   MOMMS_TUNE_BATCH_SPLIT: {slices, 0, 0, 0}, 1 <= slices <= 8.
 *   MOMMS_TUNE_BATCH_CRT:   {moduli, 0, 0, 0}, within the residue form's range.
 * A table whose version differs from MOMMS_TUNING_VERSION, with count above
 * MOMMS_TUNING_MAX, an unknown kind, a key outside 0..255 or a value outside
 * its range is refused as a whole (-1, nothing imported). */
#define MOMMS_TUNING_VERSION 1u
#define MOMMS_TUNING_MAX 256
enum { MOMMS_TUNE_BATCH_FP64 = 1, MOMMS_TUNE_BATCH_SPLIT = 2, MOMMS_TUNE_BATCH_CRT = 3 };

typedef struct momms_tuning_entry {
  int32_t kind, key;
  int32_t value[4];
} momms_tuning_entry;

typedef struct momms_tuning {
  uint32_t version; /* MOMMS_TUNING_VERSION */
  uint32_t count;   /* entries in use */
  momms_tuning_entry entry[MOMMS_TUNING_MAX];
} momms_tuning;

/* The class key of a batch table (host array), for entries written by hand. */
int momms_hip_batch_class(const momms_gemm_desc* descs, int64_t count, int32_t* key);

/* Copies the process's table into *out (sorted by kind, then key). */
int momms_hip_tuning_export(momms_tuning* out);

/* Adds the entries of *in to the process's table, an entry replacing one of
 * the same kind and key; replace != 0 empties the table first. */
int momms_hip_tuning_import(const momms_tuning* in, int replace);

/* The comparison baseline: one momms_hip_dgemm launch per problem. */
int momms_hip_dgemm_batched_loop(const momms_gemm_desc* descs, int64_t count,
                                 const double* alpha, const double* A, const double* B,
                                 const double* beta, double* C, void* stream);

/* ---- hierarchical (multi-level blocked) matrices --------------------------------
 * The device-resident counterpart of momms' Hierarch<T, LH, LW, LRS, LCS>
 * (src/matrix/hierarch.rs:18-184): a matrix stored as nested blocks whose
 * sizes come from a control tree's hierarchy description (the
 * AlgorithmStep list GemmNode::hierarchy_description returns, innermost
 * first, src/composables/mod.rs:24-36), with LH x LW leaves of strides
 * LRS, LCS. The offset of element (y, x) is the sum of a y part and an x
 * part (get_offset_y/_x, hierarch.rs:153-184): walking the levels outermost
 * first, b = y / y_blksz[i] adds b * y_stride[i] and leaves y - b * y_blksz[i]
 * for the next level; what is left at the leaf adds y * lrs; likewise x with
 * x_blksz, x_stride and lcs. The size is padded to whole top-level blocks. */
#define MOMMS_HIER_MAX_LEVELS 8
enum { MOMMS_STEP_M = 0, MOMMS_STEP_N = 1, MOMMS_STEP_K = 2 };

/* AlgorithmStep::{M, N, K}{bsz} (src/composables/mod.rs:25-29). */
typedef struct momms_algo_step {
  int32_t kind; /* MOMMS_STEP_M, _N or _K */
  int32_t pad;
  int64_t bsz;
} momms_algo_step;

typedef struct momms_hier_layout {
  int64_t h, w;         /* logical height and width */
  int64_t h_pad, w_pad; /* padded to whole top-level blocks */
  int64_t ny, nx;       /* block levels of y and x, outermost first */
  int64_t capacity;     /* elements: h_pad * w_pad */
  int64_t lrs, lcs;     /* leaf row and column stride */
  int64_t y_blksz[MOMMS_HIER_MAX_LEVELS], y_stride[MOMMS_HIER_MAX_LEVELS];
  int64_t x_blksz[MOMMS_HIER_MAX_LEVELS], x_stride[MOMMS_HIER_MAX_LEVELS];
} momms_hier_layout;

/* momms_hip_dgemm with the output tiles taken in the order of a control
 * tree's loops instead of the built-in raster: the device form of the
 * reference's variant control trees (Goto, src/exper/l3_packing.rs:58-66;
 * resident A / B / C, src/exper/l3_shapes.rs:69-100; L4C,
 * src/exper/l4.rs:71-78), which differ in the loop order and block sizes
 * that decide which operand stays in which cache. steps is the tree's
 * hierarchy description (GemmNode::hierarchy_description,
 * src/composables/mod.rs:24-36: innermost step first, block sizes in
 * elements, as for momms_hier_layout_init). Its M and N steps, outermost
 * first and rounded down to whole 128 x 128 tiles, split the tile grid into
 * nested blocks taken in order (what the last step leaves is walked with m
 * fastest), consecutive tiles going to one XCD, whose L2 then holds the
 * panels they share; K steps leave the order alone (each tile keeps its
 * whole k loop). Every tile is the 128 x 128 buffer->LDS tile of
 * momms_hip_dgemm_tiled's configuration 0, so the result is bit-identical
 * to it whatever the tree; operands that tile cannot take run configuration
 * 0's register-staged form in its own order. Returns -15 for a step kind
 * other than M, N, K, a block size below 1 or more than 8 M / N steps, -16
 * for a bad nsteps. */
int momms_hip_dgemm_tree(int64_t m, int64_t n, int64_t k, const double* alpha, const double* A,
                         int64_t rsA, int64_t csA, const double* B, int64_t rsB, int64_t csB,
                         const double* beta, double* C, int64_t rsC, int64_t csC,
                         const momms_algo_step* steps, int64_t nsteps, void* stream);

/* Hierarch::new (hierarch.rs:90-140): the layout of an h x w matrix whose
 * rows follow steps of kind ydim and columns steps of kind xdim; lh, lw give
 * the top-level block when no step of that kind exists. Host only. Returns
 * -1 for a bad argument (non-positive block size or negative leaf stride),
 * -2 for more than MOMMS_HIER_MAX_LEVELS levels in one dimension, -3 for a
 * layout in which two elements would share a place: a block that does not
 * split evenly into the next level's, or a leaf (innermost y block x
 * innermost x block) not stored densely by columns (lrs 1, lcs its height)
 * or by rows (lcs 1, lrs its width). The reference accepts the last kind
 * and overlaps elements. */
int momms_hier_layout_init(momms_hier_layout* L, int64_t h, int64_t w, const momms_algo_step* steps,
                           int64_t nsteps, int32_t ydim, int32_t xdim, int64_t lh, int64_t lw,
                           int64_t lrs, int64_t lcs);

/* Packer<Matrix, Hierarch>::pack (src/composables/pack.rs:153-309): P (device,
 * L->capacity doubles) := A (h x w, strides rsA, csA) in layout L, padding
 * zero (also whole padding leaves, which the reference leaves unwritten).
 * L is one momms_hier_layout_init filled; the entry points check its fields
 * for consistency, and a place outside [0, capacity) of a hand-made layout is
 * neither written here nor read by momms_hip_hier_unpack_add (it unpacks as
 * 0). */
int momms_hip_hier_pack(const momms_hier_layout* L, const double* A, int64_t rsA, int64_t csA,
                        double* P, void* stream);

/* Unpacker<Matrix, Hierarch>::add (src/composables/unpack.rs:61-98, :174-213),
 * the last step of UnpackC (unpack.rs:214-270): C[y][x] += P[offset(y, x)]
 * for every element of the h x w matrix. */
int momms_hip_hier_unpack_add(const momms_hier_layout* L, const double* P, double* C,
                              int64_t rsC, int64_t csC, void* stream);

/* momms_hip_hier_pack and momms_hip_hier_unpack_add for Hierarch<f32>
 * buffers (the KNM sgemm tree's packed operands, src/exper/knm.rs:89-97),
 * and the copy form of the unpack: C := P at the matrix's places (C not
 * read), which gives a column-major copy of a hierarchical operand. */
int momms_hip_hier_pack_f32(const momms_hier_layout* L, const float* A, int64_t rsA, int64_t csA,
                            float* P, void* stream);
int momms_hip_hier_unpack_add_f32(const momms_hier_layout* L, const float* P, float* C,
                                  int64_t rsC, int64_t csC, void* stream);
int momms_hip_hier_unpack_f32(const momms_hier_layout* L, const float* P, float* C, int64_t rsC,
                              int64_t csC, void* stream);

/* GemmNode::run over Hierarch<f64> operands, as the reference's l3
 * experiment runs its control trees on them (src/exper/l3.rs:46-72): C :=
 * beta*C + alpha*A*B on the logical elements of the hierarchical buffers PA
 * (LA: m x k), PB (LB: k x n) and PC (LC: m x n); C's padding places are
 * neither read nor written, C is not read when beta == 0. The operands are
 * copied column-major into library scratch (8(mk + kn + mn) bytes per
 * stream, kept), multiplied by momms_hip_dgemm (same kernels, same bits as
 * on the column-major matrices), and the result written back into PC. */
int momms_hip_dgemm_hier(const momms_hier_layout* LA, const double* PA,
                         const momms_hier_layout* LB, const double* PB, const double* alpha,
                         const double* beta, const momms_hier_layout* LC, double* PC,
                         void* stream);

/* ---- accuracy statement -----------------------------------------------------
 * Recomputes `samples` entries of beta*C0 + alpha*A*B in double-double on the
 * GPU (Dot2, Ogita-Rump-Oishi 2005) and returns, on the host, the largest
 * scaled error |C - C*| / (|alpha||A||B| + |beta||C0|) and relative error
 * |C - C*| / |C*| of C against them. Entry e is (i, j) =
 * (momms_sample_index(seed, e, 0, m), momms_sample_index(seed, e, 1, n)).
 * ref_out (optional, device, `samples` doubles) receives C*. C0 may be NULL
 * when beta == 0. Synchronizes `stream` before returning. The reference's
 * counterpart is the unasserted residual of src/util.rs:55-85. */
int momms_hip_gemm_check_sampled(int64_t m, int64_t n, int64_t k, const double* alpha,
                                 const double* A, int64_t rsA, int64_t csA, const double* B,
                                 int64_t rsB, int64_t csB, const double* beta, const double* C0,
                                 const double* C, int64_t rsC, int64_t csC, int64_t samples,
                                 uint64_t seed, double* ref_out, double* max_scaled,
                                 double* max_rel, void* stream);

/* Sample position used by momms_hip_gemm_check_sampled: SplitMix64 of
 * seed + 2e (+1 for the column, which != 0) modulo extent. Host only. */
uint64_t momms_sample_index(uint64_t seed, uint64_t e, int which, uint64_t extent);

#ifdef __cplusplus
}
#endif

#endif /* MOMMS_HIP_H */
