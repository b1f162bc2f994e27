// FP64-accurate GEMM from BF16 residues (Ozaki scheme II: Ozaki, Uchino,
// Imamura, "Ozaki Scheme II: A GEMM-oriented emulation of floating-point
// matrix multiplication using an integer modular technique", 2025) on BF16
// matrix cores.
//
// No counterpart in the reference; it is the second form of the precision
// addition beside the slice form of split_bf16.hip, and needs one BF16 GEMM
// per modulus instead of s(s+1)/2 slice products. Exact arithmetic (restated
// on the CPU for bit-exact comparison):
//   moduli p_1..p_N: the N largest pairwise-coprime odd integers <= 511,
//   P = p_1...p_N, kb = ceil(log2 k),
//   beta = min(53, floor((bitlen(P) - 2 - kb) / 2), floor((125 - kb) / 2)).
//   row i of A: E_i with max_p |a_ip| < 2^E_i; A'_ip = rint(a_ip 2^(beta - E_i));
//   column j of B likewise (F_j, B'_pj). |A'|, |B'| <= 2^beta.
//   C' = A'B' exactly: |C'| <= k 2^(2 beta) < P/2, so the residues of C'
//   modulo p_1..p_N fix it.
//   C_ij = beta_C C0_ij + alpha * 2^(E_i + F_j - 2 beta) * double(C'_ij)
//   (double() rounds to nearest, ties to even; fma for the beta_C update,
//   C not read when beta_C == 0). A row of A or column of B holding Inf or
//   NaN is not converted: its entries take alpha * (fma chain over p).
//
// Device steps:
//   1. exponents of the rows of A and columns of B (vector_exponents);
//   2. crt_pack: the balanced residues of A' and B' modulo each p_l
//      (|r| <= 255, integers bf16 holds exactly) in packed panels: for
//      modulus l, panel pv and 16-deep k-slab ks, one contiguous block
//      [kh][PANEL][8] (kh = which half of the slab), panels one unused slab
//      apart (panel_stride);
//   3. crt_multiply: one BF16 GEMM per modulus with FP32 accumulation. Any
//      256 consecutive k-values add at most 256 * 255^2 to an accumulator
//      that starts below 0.51 p in magnitude, so every partial sum is an
//      integer below 2^24 and exact in any order; each accumulator is reduced
//      modulo p_l (x - p rint(x / p), exact by fma) once per 256 k-values (at
//      a point staggered over the accumulators, panel_gemm.hpp) and at the
//      end. The tile ends as the balanced residue of C' modulo p_l (int16);
//   4. crt_reconstruct: per entry, the Chinese remainder sum of the N
//      residues in 128-bit integers (the quotient by P from an fp64
//      estimate) gives C', then one correctly rounded conversion, the scaling
//      and the alpha / beta update.
#include <math.h>

#include "crt_bf16.hpp"
#include "c_update.hpp"
#include "crt_common.hpp"
#include "crt_recon.hpp"
#include "panel_gemm.hpp"
#include "mfma.hpp"
#include "split_bf16.hpp"
#include "split_tile.hpp"
#include "status.hpp"

namespace momms {

// ---------------------------------------------------------------------------
// moduli and the constants derived from them

int crt_modulus(int l) { return (l >= 0 && l < kCrtMaxModuli) ? kCrtModuli[l] : 0; }

int crt_beta_for(int64_t k, int N, const int* mods) {
  if (N < 1 || N > kCrtMaxModuli || k < 0) return -1;
  uint32_t w[8] = {1, 0, 0, 0, 0, 0, 0, 0};  // P, exactly, in 32-bit limbs
  for (int l = 0; l < N; ++l) {
    uint64_t carry = 0;
    for (int i = 0; i < 8; ++i) {
      const uint64_t t = (uint64_t)w[i] * (uint64_t)mods[l] + carry;
      w[i] = (uint32_t)t;
      carry = t >> 32;
    }
  }
  int bitlen = 0;
  for (int i = 7; i >= 0; --i)
    if (w[i]) {
      bitlen = 32 * i + (32 - __builtin_clz(w[i]));
      break;
    }
  int kb = 0;
  while (((int64_t)1 << kb) < k) ++kb;
  const int room = bitlen - 2 - kb;
  int b = room >= 0 ? room / 2 : -1;
  const int cap = (125 - kb) / 2;
  if (b > cap) b = cap;
  if (b > 53) b = 53;
  return b;
}

int crt_beta(int64_t k, int N) { return crt_beta_for(k, N, kCrtModuli); }

static inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

CrtLayout crt_layout(int64_t m, int64_t n, int64_t k, int N) {
  CrtLayout L;
  L.nmod = N;
  L.beta = crt_beta(k, N);
  L.mpad = ceil_div(m > 0 ? m : 1, kCrtBM) * kCrtBM;
  L.npad = ceil_div(n > 0 ? n : 1, kCrtBN) * kCrtBN;
  L.kpad = ceil_div(k > 0 ? k : 1, kCrtKAlign) * kCrtKAlign;
  size_t off = 0;
  L.off_ra = off;
  const int64_t kstride = panel_stride(L.kpad / kSlab) * kSlab;  // k-values per panel
  off = align256(off + (size_t)N * L.mpad * kstride * 2);
  L.off_rb = off;
  off = align256(off + (size_t)N * L.npad * kstride * 2);
  L.off_rc = off;
  off = align256(off + (size_t)N * L.mpad * L.npad * 2);
  L.off_ea = off;
  off = align256(off + (size_t)L.mpad * 4);
  L.off_eb = off;
  off = align256(off + (size_t)L.npad * 4);
  L.off_maxa = off;
  off = align256(off + (size_t)L.mpad * 8);
  L.off_maxb = off;
  off = align256(off + (size_t)L.npad * 8);
  L.bytes = off;
  return L;
}

// ---------------------------------------------------------------------------
// 3. one residue GEMM per (modulus, tile): the panel GEMM (panel_gemm.hpp)
// with the modular reduction every 256 k-values, ending as the balanced
// residue of C' modulo p_l (int16, column-major, leading dimension mpad).

template <int FM, int FN, int WM, int WN, int SPS>
__global__ void __launch_bounds__(64 * WM * WN, 1)
crt_multiply(int64_t tiles_m, int64_t tiles_n, int64_t nslabs, int64_t mpad, int64_t npad,
             const uint16_t* __restrict__ RA, const uint16_t* __restrict__ RB,
             int16_t* __restrict__ RC) {
  typedef PanelGemm<FM, FN, WM, WN, SPS> G;
  __shared__ __attribute__((aligned(16))) uint16_t S0[G::SLOT];
  __shared__ __attribute__((aligned(16))) uint16_t S1[G::SLOT];
  __shared__ __attribute__((aligned(16))) uint16_t S2[G::SLOT];
  __shared__ __attribute__((aligned(16))) uint16_t S3[G::SLOT];

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave % WM, wn = wave / WM, kh = lane >> 5;
  const int64_t tiles = tiles_m * tiles_n;
  const int64_t t = xcd_remap(blockIdx.x, (int64_t)gridDim.x);
  const int l = (int)(t / tiles);
  int64_t tm, tn;
  grouped_tile(t - (int64_t)l * tiles, tiles_m, tiles_n, 8, tm, tn);
  const float p = kCrtBf16Tab.pf[l], invp = kCrtBf16Tab.invpf[l];

  v16f32 acc[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.0f;
  panel_mainloop<FM, FN, WM, WN, SPS, true>(
      acc, RA + ((int64_t)l * tiles_m + tm) * panel_stride(nslabs) * G::ABLK,
      RB + ((int64_t)l * tiles_n + tn) * panel_stride(nslabs) * G::BBLK, nslabs, p, invp, S0, S1,
      S2, S3);
  panel_reduce_mod<FM, FN>(acc, p, invp);

  // balanced residue in [-(p-1)/2, (p-1)/2]
  const float h = 0.5f * (p - 1.0f);
  int16_t* out = RC + (int64_t)l * mpad * npad;
#pragma unroll
  for (int a = 0; a < FM; ++a) {
    const int64_t gi = tm * G::BM + wm * 32 * FM + a * 32 + (lane & 31);
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t gj = tn * G::BN + wn * 32 * FN + b * 32 + 8 * (r >> 2) + 4 * kh + (r & 3);
        float x = acc[a][b][r];
        x = x > h ? x - p : x;
        x = x < -h ? x + p : x;
        out[gj * mpad + gi] = (int16_t)(int)x;
      }
  }
}

// ---------------------------------------------------------------------------
// host side

constexpr int kCrtFM = 2, kCrtFN = 4, kCrtWM = 4, kCrtWN = 2, kCrtSPS = 2;
static_assert(32 * kCrtFM * kCrtWM == kCrtBM && 32 * kCrtFN * kCrtWN == kCrtBN, "tile");
static_assert(kCrtKAlign % (kSlab * kCrtSPS) == 0, "panels are padded to whole stages");

hipError_t crt_prepare(const CrtLayout& L, int64_t m, int64_t n, int64_t k, const double* A,
                       int64_t rsA, int64_t csA, const double* B, int64_t rsB, int64_t csB,
                       char* ws, hipStream_t st) {
  uint16_t* ra = (uint16_t*)(ws + L.off_ra);
  uint16_t* rb = (uint16_t*)(ws + L.off_rb);
  int* ea = (int*)(ws + L.off_ea);
  int* eb = (int*)(ws + L.off_eb);
  uint64_t* ma = (uint64_t*)(ws + L.off_maxa);
  uint64_t* mb = (uint64_t*)(ws + L.off_maxb);
  hipError_t e;
  if ((e = vector_exponents(A, m, k, rsA, csA, L.mpad, ma, ea, st)) != hipSuccess) return e;
  if ((e = vector_exponents(B, n, k, csB, rsB, L.npad, mb, eb, st)) != hipSuccess) return e;
  dim3 ga((unsigned)ceil_div(L.mpad, 64), (unsigned)ceil_div(L.kpad, 64)), block(256);
  hipLaunchKernelGGL((crt_pack<kCrtBM, false>), ga, block, 0, st, A, m, k, rsA, csA, ea, L.mpad, L.kpad,
                     L.nmod, L.beta, ra);
  dim3 gb((unsigned)ceil_div(L.npad, 64), (unsigned)ceil_div(L.kpad, 64));
  hipLaunchKernelGGL((crt_pack<kCrtBN, false>), gb, block, 0, st, B, n, k, csB, rsB, eb, L.npad, L.kpad,
                     L.nmod, L.beta, rb);
  return hipGetLastError();
}

hipError_t crt_multiply_launch(const CrtLayout& L, const char* ws, hipStream_t st) {
  const int64_t tm = L.mpad / kCrtBM, tn = L.npad / kCrtBN;
  dim3 grid((unsigned)(tm * tn * L.nmod)), block(64 * kCrtWM * kCrtWN);
  hipLaunchKernelGGL((crt_multiply<kCrtFM, kCrtFN, kCrtWM, kCrtWN, kCrtSPS>), grid, block, 0,
                     st, tm, tn, L.kpad / kSlab, L.mpad, L.npad,
                     (const uint16_t*)(ws + L.off_ra), (const uint16_t*)(ws + L.off_rb),
                     (int16_t*)(ws + L.off_rc));
  return hipGetLastError();
}

hipError_t crt_reconstruct_launch(const CrtLayout& L, int64_t m, int64_t n, int64_t k,
                                  double alpha, double beta, double* C, int64_t rsC, int64_t csC,
                                  const double* A, int64_t rsA, int64_t csA, const double* B,
                                  int64_t rsB, int64_t csB, const char* ws, hipStream_t st) {
  const void* rc = ws + L.off_rc;
  const int* ea = (const int*)(ws + L.off_ea);
  const int* eb = (const int*)(ws + L.off_eb);
  dim3 grid((unsigned)ceil_div(m, 256 * kCrtRows), (unsigned)(n < 65535 ? n : 65535)), block(256);
#define MOMMS_RC(N_) hipLaunchKernelGGL((crt_reconstruct<N_, false>), grid, block, 0, st, m, n, k, L.beta, alpha, beta, C, rsC, csC, rc, L.mpad, L.npad, ea, eb, A, rsA, csA, B, rsB, csB)
  switch (L.nmod) {
    case 1: MOMMS_RC(1); break;
    case 2: MOMMS_RC(2); break;
    case 3: MOMMS_RC(3); break;
    case 4: MOMMS_RC(4); break;
    case 5: MOMMS_RC(5); break;
    case 6: MOMMS_RC(6); break;
    case 7: MOMMS_RC(7); break;
    case 8: MOMMS_RC(8); break;
    case 9: MOMMS_RC(9); break;
    case 10: MOMMS_RC(10); break;
    case 11: MOMMS_RC(11); break;
    case 12: MOMMS_RC(12); break;
    case 13: MOMMS_RC(13); break;
    case 14: MOMMS_RC(14); break;
    case 15: MOMMS_RC(15); break;
    default: MOMMS_RC(16); break;
  }
#undef MOMMS_RC
  return hipGetLastError();
}

}  // namespace momms

// ---------------------------------------------------------------------------
// C ABI

using namespace momms;

extern "C" int momms_hip_crt_bf16_beta(int64_t k, int moduli) {
  return (k < 0 || moduli < 1 || moduli > kCrtMaxModuli) ? -1 : crt_beta(k, moduli);
}

extern "C" int momms_hip_crt_bf16_modulus(int index) { return crt_modulus(index); }

extern "C" size_t momms_hip_crt_bf16_workspace_size(int64_t m, int64_t n, int64_t k,
                                                    int moduli) {
  if (m < 0 || n < 0 || k < 0 || moduli < 1 || moduli > kCrtMaxModuli) return 0;
  if (crt_beta(k, moduli) < 1) return 0;
  return crt_layout(m, n, k, moduli).bytes;
}

extern "C" int momms_hip_dgemm_crt_bf16(int64_t m, int64_t n, int64_t k, const double* alpha,
                                        const double* A, int64_t rsA, int64_t csA,
                                        const double* B, int64_t rsB, int64_t csB,
                                        const double* beta, double* C, int64_t rsC,
                                        int64_t csC, int moduli, void* workspace,
                                        size_t workspace_bytes, void* stream) {
  static const char* fn = "momms_hip_dgemm_crt_bf16";
  if (m < 0) return bad_arg(fn, 1, "m < 0");
  if (n < 0) return bad_arg(fn, 2, "n < 0");
  if (k < 0) return bad_arg(fn, 3, "k < 0");
  if (alpha == nullptr) return bad_arg(fn, 4, "alpha is null");
  int st;
  if ((st = check_operand(fn, 5, A, m, k, rsA, csA))) return st;
  if ((st = check_operand(fn, 8, B, k, n, rsB, csB))) return st;
  if (beta == nullptr) return bad_arg(fn, 11, "beta is null");
  if ((st = check_operand(fn, 12, C, m, n, rsC, csC))) return st;
  if (moduli < 1 || moduli > kCrtMaxModuli) return bad_arg(fn, 15, "moduli must be 1..16");
  if (crt_beta(k, moduli) < 1) return bad_arg(fn, 15, "too few moduli for this k");
  if (m == 0 || n == 0) return 0;
  const CrtLayout L = crt_layout(m, n, k, moduli);
  if (workspace == nullptr) return bad_arg(fn, 16, "workspace is null");
  if (workspace_bytes < L.bytes) return bad_arg(fn, 17, "workspace too small");
  char* ws = (char*)workspace;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = crt_prepare(L, m, n, k, A, rsA, csA, B, rsB, csB, ws, s);
  if (e == hipSuccess) e = crt_multiply_launch(L, ws, s);
  if (e == hipSuccess)
    e = crt_reconstruct_launch(L, m, n, k, *alpha, *beta, C, rsC, csC, A, rsA, csA, B, rsB, csB,
                               ws, s);
  return hip_status(fn, e);
}

// Staged form: the same three steps as momms_hip_dgemm_crt_bf16, one call each.
static int crt_check_common(const char* fn, int64_t m, int64_t n, int64_t k, int moduli,
                            int moduli_arg, const void* workspace, size_t workspace_bytes,
                            int ws_arg) {
  if (m < 0) return bad_arg(fn, 1, "m < 0");
  if (n < 0) return bad_arg(fn, 2, "n < 0");
  if (k < 0) return bad_arg(fn, 3, "k < 0");
  if (moduli < 1 || moduli > kCrtMaxModuli) return bad_arg(fn, moduli_arg, "moduli must be 1..16");
  if (crt_beta(k, moduli) < 1) return bad_arg(fn, moduli_arg, "too few moduli for this k");
  if (m == 0 || n == 0) return 0;
  if (workspace == nullptr) return bad_arg(fn, ws_arg, "workspace is null");
  if (workspace_bytes < crt_layout(m, n, k, moduli).bytes)
    return bad_arg(fn, ws_arg + 1, "workspace too small");
  return 0;
}

extern "C" int momms_hip_crt_bf16_prepare(int64_t m, int64_t n, int64_t k, const double* A,
                                          int64_t rsA, int64_t csA, const double* B,
                                          int64_t rsB, int64_t csB, int moduli, void* workspace,
                                          size_t workspace_bytes, void* stream) {
  static const char* fn = "momms_hip_crt_bf16_prepare";
  int st;
  if ((st = crt_check_common(fn, m, n, k, moduli, 10, workspace, workspace_bytes, 11))) return st;
  if ((st = check_operand(fn, 4, A, m, k, rsA, csA))) return st;
  if ((st = check_operand(fn, 7, B, k, n, rsB, csB))) return st;
  if (m == 0 || n == 0) return 0;
  return hip_status(fn, crt_prepare(crt_layout(m, n, k, moduli), m, n, k, A, rsA, csA, B, rsB,
                                    csB, (char*)workspace, (hipStream_t)stream));
}

extern "C" int momms_hip_crt_bf16_products(int64_t m, int64_t n, int64_t k, int moduli,
                                           void* workspace, size_t workspace_bytes,
                                           void* stream) {
  static const char* fn = "momms_hip_crt_bf16_products";
  int st;
  if ((st = crt_check_common(fn, m, n, k, moduli, 4, workspace, workspace_bytes, 5))) return st;
  if (m == 0 || n == 0) return 0;
  return hip_status(fn, crt_multiply_launch(crt_layout(m, n, k, moduli), (const char*)workspace,
                                            (hipStream_t)stream));
}

extern "C" int momms_hip_crt_bf16_reconstruct(int64_t m, int64_t n, int64_t k,
                                              const double* alpha, const double* A, int64_t rsA,
                                              int64_t csA, const double* B, int64_t rsB,
                                              int64_t csB, const double* beta, double* C,
                                              int64_t rsC, int64_t csC, int moduli,
                                              const void* workspace, size_t workspace_bytes,
                                              void* stream) {
  static const char* fn = "momms_hip_crt_bf16_reconstruct";
  int st;
  if ((st = crt_check_common(fn, m, n, k, moduli, 15, workspace, workspace_bytes, 16))) return st;
  if (alpha == nullptr) return bad_arg(fn, 4, "alpha is null");
  if ((st = check_operand(fn, 5, A, m, k, rsA, csA))) return st;
  if ((st = check_operand(fn, 8, B, k, n, rsB, csB))) return st;
  if (beta == nullptr) return bad_arg(fn, 11, "beta is null");
  if ((st = check_operand(fn, 12, C, m, n, rsC, csC))) return st;
  if (m == 0 || n == 0) return 0;
  return hip_status(fn, crt_reconstruct_launch(crt_layout(m, n, k, moduli), m, n, k, *alpha,
                                               *beta, C, rsC, csC, A, rsA, csA, B, rsB, csB,
                                               (const char*)workspace, (hipStream_t)stream));
}
