// FP64-accurate GEMM from INT8 residues: the residue form of crt_bf16.hip
// (Ozaki scheme II: Ozaki, Uchino, Imamura 2025) on INT8 matrix cores, which
// run at twice the BF16 rate on this chip.
//
// No counterpart in the reference; the third form of the precision addition.
// The exact arithmetic is that of the BF16 residue form with another moduli
// table, so the result is fixed by the integer width beta alone:
//   moduli p_1..p_N: the N largest pairwise-coprime odd integers <= 255
//   (kCrtI8Moduli), P = p_1...p_N, kb = ceil(log2 k),
//   beta = min(53, floor((bitlen(P) - 2 - kb) / 2), floor((125 - kb) / 2));
//   A'_ip = rint(a_ip 2^(beta - E_i)), B'_pj likewise, C' = A'B' exactly,
//   C_ij = beta_C C0_ij + alpha 2^(E_i + F_j - 2 beta) double(C'_ij).
// 15 INT8 moduli give beta = 51 at k = 8192, as 13 BF16 moduli do.
//
// Device steps:
//   1. exponents of the rows of A and columns of B;
//   2. crt_pack<PANEL, true>: balanced residues (|r| <= 127, int8) in the
//      panel layout, 16 k-values per 16-byte group, 32-deep slabs;
//   3. crt_i8_multiply: one INT8 GEMM per modulus (v_mfma_i32_32x32x32_i8)
//      over the panel main loop; the INT32 sums are exact (|sum| <= k 127^2
//      < 2^31 for k <= 131072), so no reduction runs inside the k loop; the
//      tile ends as the balanced residue of C' modulo p_l (int8);
//   4. crt_reconstruct<N, true>: the Chinese remainder sum of the N int8
//      residues, one correctly rounded conversion, scaling, alpha / beta.
#include "crt_bf16.hpp"
#include "crt_common.hpp"
#include "split_bf16.hpp"
#include "status.hpp"

namespace momms {

constexpr int kI8BM = 256, kI8BN = 256;  // residue GEMM tile = panel sizes
constexpr int kI8Slab = 32;              // k per slab (one v_mfma_i32_32x32x32_i8 step)
constexpr int kI8KAlign = 64;            // k padding: whole two-slab stages
constexpr int64_t kI8MaxK = 131072;      // k 127^2 < 2^31

struct CrtI8Layout {
  int nmod, beta;
  int64_t mpad, npad, kpad;
  size_t off_ra, off_rb, off_rc, off_ea, off_eb, off_maxa, off_maxb, bytes;
};

static int crt_i8_beta(int64_t k, int N) {
  if (k > kI8MaxK) return -1;
  return crt_beta_for(k, N, kCrtI8Moduli);
}

static inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

static CrtI8Layout crt_i8_layout(int64_t m, int64_t n, int64_t k, int N) {
  CrtI8Layout L;
  L.nmod = N;
  L.beta = crt_i8_beta(k, N);
  L.mpad = ceil_div(m > 0 ? m : 1, kI8BM) * kI8BM;
  L.npad = ceil_div(n > 0 ? n : 1, kI8BN) * kI8BN;
  L.kpad = ceil_div(k > 0 ? k : 1, kI8KAlign) * kI8KAlign;
  const int64_t kstride = panel_stride(L.kpad / kI8Slab) * kI8Slab;  // bytes per panel row
  size_t off = 0;
  L.off_ra = off;
  off = align256(off + (size_t)N * L.mpad * kstride);
  L.off_rb = off;
  off = align256(off + (size_t)N * L.npad * kstride);
  L.off_rc = off;
  off = align256(off + (size_t)N * L.mpad * L.npad);
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

// One INT8 residue GEMM per (modulus, tile), ending as the balanced residue
// of C' modulo p_l (int8, column-major, leading dimension mpad).
template <int FM, int FN, int WM, int WN, int SPS>
__global__ void __launch_bounds__(64 * WM * WN, 1)
crt_i8_multiply(int64_t tiles_m, int64_t tiles_n, int64_t nslabs, int64_t mpad, int64_t npad,
                const uint16_t* __restrict__ RA, const uint16_t* __restrict__ RB,
                int8_t* __restrict__ RC) {
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

  v16i32 acc[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][b][r] = 0;
  panel_mainloop<FM, FN, WM, WN, SPS, false, true>(
      acc, RA + ((int64_t)l * tiles_m + tm) * panel_stride(nslabs) * G::ABLK,
      RB + ((int64_t)l * tiles_n + tn) * panel_stride(nslabs) * G::BBLK, nslabs, 0.0f, 0.0f, S0,
      S1, S2, S3);

  // x mod p, balanced: q = rint(x / p) from the float of x is off by at most
  // one (|x| < 2^31: x / p is known to 2^7 / p < 1), so x - q p (exact in
  // int32) lies within 1.5 p of 0 and one correction each way ends in
  // [-(p-1)/2, (p-1)/2]
  const int p = (int)kCrtI8Tab.pf[l];
  const float invp = kCrtI8Tab.invpf[l];
  const int h = (p - 1) / 2;
  int8_t* out = RC + (int64_t)l * mpad * npad;
#pragma unroll
  for (int a = 0; a < FM; ++a) {
    const int64_t gi = tm * G::BM + wm * 32 * FM + a * 32 + (lane & 31);
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t gj = tn * G::BN + wn * 32 * FN + b * 32 + 8 * (r >> 2) + 4 * kh + (r & 3);
        const int x = acc[a][b][r];
        const int q = (int)__builtin_rintf((float)x * invp);
        int y = x - q * p;
        y = y > h ? y - p : y;
        y = y < -h ? y + p : y;
        out[gj * mpad + gi] = (int8_t)y;
      }
  }
}

constexpr int kI8FM = 2, kI8FN = 4, kI8WM = 4, kI8WN = 2, kI8SPS = 2;
static_assert(32 * kI8FM * kI8WM == kI8BM && 32 * kI8FN * kI8WN == kI8BN, "tile");
static_assert(kI8KAlign % (kI8Slab * kI8SPS) == 0, "panels are padded to whole stages");

static hipError_t crt_i8_prepare(const CrtI8Layout& L, int64_t m, int64_t n, int64_t k,
                                 const double* A, int64_t rsA, int64_t csA, const double* B,
                                 int64_t rsB, int64_t csB, char* ws, hipStream_t st) {
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
  hipLaunchKernelGGL((crt_pack<kI8BM, true>), ga, block, 0, st, A, m, k, rsA, csA, ea, L.mpad,
                     L.kpad, L.nmod, L.beta, ra);
  dim3 gb((unsigned)ceil_div(L.npad, 64), (unsigned)ceil_div(L.kpad, 64));
  hipLaunchKernelGGL((crt_pack<kI8BN, true>), gb, block, 0, st, B, n, k, csB, rsB, eb, L.npad,
                     L.kpad, L.nmod, L.beta, rb);
  return hipGetLastError();
}

static hipError_t crt_i8_multiply_launch(const CrtI8Layout& L, const char* ws, hipStream_t st) {
  const int64_t tm = L.mpad / kI8BM, tn = L.npad / kI8BN;
  dim3 grid((unsigned)(tm * tn * L.nmod)), block(64 * kI8WM * kI8WN);
  hipLaunchKernelGGL((crt_i8_multiply<kI8FM, kI8FN, kI8WM, kI8WN, kI8SPS>), grid, block, 0, st,
                     tm, tn, L.kpad / kI8Slab, L.mpad, L.npad, (const uint16_t*)(ws + L.off_ra),
                     (const uint16_t*)(ws + L.off_rb), (int8_t*)(ws + L.off_rc));
  return hipGetLastError();
}

static hipError_t crt_i8_reconstruct_launch(const CrtI8Layout& L, int64_t m, int64_t n,
                                            int64_t k, double alpha, double beta, double* C,
                                            int64_t rsC, int64_t csC, const double* A,
                                            int64_t rsA, int64_t csA, const double* B,
                                            int64_t rsB, int64_t csB, const char* ws,
                                            hipStream_t st) {
  const void* rc = ws + L.off_rc;
  const int* ea = (const int*)(ws + L.off_ea);
  const int* eb = (const int*)(ws + L.off_eb);
  dim3 grid((unsigned)ceil_div(m, 256 * kCrtRows), (unsigned)(n < 65535 ? n : 65535)), block(256);
#define MOMMS_RC(N_) hipLaunchKernelGGL((crt_reconstruct<N_, true>), grid, block, 0, st, m, n, k, L.beta, alpha, beta, C, rsC, csC, rc, L.mpad, L.npad, ea, eb, A, rsA, csA, B, rsB, csB)
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

// Argument checks shared by the entry points (argument positions given).
static int crt_i8_check(const char* fn, int64_t m, int64_t n, int64_t k, int moduli,
                        int moduli_arg, const void* workspace, size_t workspace_bytes,
                        int ws_arg) {
  if (m < 0) return bad_arg(fn, 1, "m < 0");
  if (n < 0) return bad_arg(fn, 2, "n < 0");
  if (k < 0) return bad_arg(fn, 3, "k < 0");
  if (k > kI8MaxK) return bad_arg(fn, 3, "k > 131072 (the INT32 sums would overflow)");
  if (moduli < 1 || moduli > kCrtMaxModuli) return bad_arg(fn, moduli_arg, "moduli must be 1..16");
  if (crt_i8_beta(k, moduli) < 1) return bad_arg(fn, moduli_arg, "too few moduli for this k");
  if (m == 0 || n == 0) return 0;
  if (workspace == nullptr) return bad_arg(fn, ws_arg, "workspace is null");
  if (workspace_bytes < crt_i8_layout(m, n, k, moduli).bytes)
    return bad_arg(fn, ws_arg + 1, "workspace too small");
  return 0;
}

}  // namespace momms

// ---------------------------------------------------------------------------
// C ABI

using namespace momms;

extern "C" int momms_hip_crt_i8_beta(int64_t k, int moduli) {
  return (k < 0 || moduli < 1 || moduli > kCrtMaxModuli) ? -1 : crt_i8_beta(k, moduli);
}

extern "C" int momms_hip_crt_i8_modulus(int index) {
  return (index >= 0 && index < kCrtMaxModuli) ? kCrtI8Moduli[index] : 0;
}

extern "C" size_t momms_hip_crt_i8_workspace_size(int64_t m, int64_t n, int64_t k, int moduli) {
  if (m < 0 || n < 0 || k < 0 || moduli < 1 || moduli > kCrtMaxModuli) return 0;
  if (crt_i8_beta(k, moduli) < 1) return 0;
  return crt_i8_layout(m, n, k, moduli).bytes;
}

extern "C" int momms_hip_dgemm_crt_i8(int64_t m, int64_t n, int64_t k, const double* alpha,
                                      const double* A, int64_t rsA, int64_t csA,
                                      const double* B, int64_t rsB, int64_t csB,
                                      const double* beta, double* C, int64_t rsC, int64_t csC,
                                      int moduli, void* workspace, size_t workspace_bytes,
                                      void* stream) {
  static const char* fn = "momms_hip_dgemm_crt_i8";
  int st;
  if ((st = crt_i8_check(fn, m, n, k, moduli, 15, workspace, workspace_bytes, 16))) return st;
  if (alpha == nullptr) return bad_arg(fn, 4, "alpha is null");
  if ((st = check_operand(fn, 5, A, m, k, rsA, csA))) return st;
  if ((st = check_operand(fn, 8, B, k, n, rsB, csB))) return st;
  if (beta == nullptr) return bad_arg(fn, 11, "beta is null");
  if ((st = check_operand(fn, 12, C, m, n, rsC, csC))) return st;
  if (m == 0 || n == 0) return 0;
  const CrtI8Layout L = crt_i8_layout(m, n, k, moduli);
  char* ws = (char*)workspace;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = crt_i8_prepare(L, m, n, k, A, rsA, csA, B, rsB, csB, ws, s);
  if (e == hipSuccess) e = crt_i8_multiply_launch(L, ws, s);
  if (e == hipSuccess)
    e = crt_i8_reconstruct_launch(L, m, n, k, *alpha, *beta, C, rsC, csC, A, rsA, csA, B, rsB,
                                  csB, ws, s);
  return hip_status(fn, e);
}

extern "C" int momms_hip_crt_i8_prepare(int64_t m, int64_t n, int64_t k, const double* A,
                                        int64_t rsA, int64_t csA, const double* B, int64_t rsB,
                                        int64_t csB, int moduli, void* workspace,
                                        size_t workspace_bytes, void* stream) {
  static const char* fn = "momms_hip_crt_i8_prepare";
  int st;
  if ((st = crt_i8_check(fn, m, n, k, moduli, 10, workspace, workspace_bytes, 11))) return st;
  if ((st = check_operand(fn, 4, A, m, k, rsA, csA))) return st;
  if ((st = check_operand(fn, 7, B, k, n, rsB, csB))) return st;
  if (m == 0 || n == 0) return 0;
  return hip_status(fn, crt_i8_prepare(crt_i8_layout(m, n, k, moduli), m, n, k, A, rsA, csA, B,
                                       rsB, csB, (char*)workspace, (hipStream_t)stream));
}

extern "C" int momms_hip_crt_i8_products(int64_t m, int64_t n, int64_t k, int moduli,
                                         void* workspace, size_t workspace_bytes, void* stream) {
  static const char* fn = "momms_hip_crt_i8_products";
  int st;
  if ((st = crt_i8_check(fn, m, n, k, moduli, 4, workspace, workspace_bytes, 5))) return st;
  if (m == 0 || n == 0) return 0;
  return hip_status(fn, crt_i8_multiply_launch(crt_i8_layout(m, n, k, moduli),
                                               (const char*)workspace, (hipStream_t)stream));
}

extern "C" int momms_hip_crt_i8_reconstruct(int64_t m, int64_t n, int64_t k, const double* alpha,
                                            const double* A, int64_t rsA, int64_t csA,
                                            const double* B, int64_t rsB, int64_t csB,
                                            const double* beta, double* C, int64_t rsC,
                                            int64_t csC, int moduli, const void* workspace,
                                            size_t workspace_bytes, void* stream) {
  static const char* fn = "momms_hip_crt_i8_reconstruct";
  int st;
  if ((st = crt_i8_check(fn, m, n, k, moduli, 15, workspace, workspace_bytes, 16))) return st;
  if (alpha == nullptr) return bad_arg(fn, 4, "alpha is null");
  if ((st = check_operand(fn, 5, A, m, k, rsA, csA))) return st;
  if ((st = check_operand(fn, 8, B, k, n, rsB, csB))) return st;
  if (beta == nullptr) return bad_arg(fn, 11, "beta is null");
  if ((st = check_operand(fn, 12, C, m, n, rsC, csC))) return st;
  if (m == 0 || n == 0) return 0;
  return hip_status(fn, crt_i8_reconstruct_launch(crt_i8_layout(m, n, k, moduli), m, n, k,
                                                  *alpha, *beta, C, rsC, csC, A, rsA, csA, B,
                                                  rsB, csB, (const char*)workspace,
                                                  (hipStream_t)stream));
}
