// Reduced-precision GEMMs with FP32 accumulation and FP32 output:
//   momms_hip_sgemm       FP32 in  -> v_mfma_f32_32x32x2_f32
//   momms_hip_gemm_bf16   BF16 in  -> v_mfma_f32_32x32x16_bf16
// C := beta*C + alpha*A*B, C float. The FP32 form is the counterpart of the
// reference's KNM sgemm leaf (KnmKernel -> sgemm_knm_asm_16x24,
// src/kern/knm_kernel.rs:8-81, ukernels/sgemm_knm_asm_24x16.c:30-146).
//
// Large BF16 problems take the packed-panel form below (launch_bf16_panels).
// Otherwise one template serves both: a slab of BK k-values is staged in LDS as
// [ks][kh][row][V] where V is the elements per lane of one MFMA operand
// (1 for f32, 8 for bf16) and kh = lane/32 picks the operand's k half, so
// every lane reads its fragment with one contiguous load. Workgroup 256
// threads (2 x 2 wavefronts), tile 128 x 128, each wavefront 64 x 64 as
// 2 x 2 blocks of 32 x 32. Operands are swapped at the MFMA (B slab as the
// "A" operand) so stores to column-major C are coalesced, as in dgemm.hip.
#include "dgemm_tile.hpp"
#include "mfma.hpp"
#include "panel_gemm.hpp"
#include "scratch.hpp"
#include "status.hpp"

namespace momms {

template <class TIn>
struct MfmaTraits;
template <>
struct MfmaTraits<float> {
  static constexpr int V = 1;      // elements per lane per operand
  static constexpr int KSTEP = 2;  // k per MFMA
  typedef float frag;
  __device__ static v16f32 mma(frag a, frag b, v16f32 c) { return mfma_f32_32x32x2(a, b, c); }
};
template <>
struct MfmaTraits<__bf16> {
  static constexpr int V = 8;
  static constexpr int KSTEP = 16;
  typedef v8bf16 frag;
  __device__ static v16f32 mma(frag a, frag b, v16f32 c) { return mfma_bf16_32x32x16(a, b, c); }
};

template <class TIn, int BK, int LA, int LB>
__global__ void __launch_bounds__(256)
gemm_f32acc_kernel(int64_t m, int64_t n, int64_t k, float alpha, const TIn* __restrict__ A,
                   int64_t rsA, int64_t csA, const TIn* __restrict__ B, int64_t rsB, int64_t csB,
                   float beta, float* __restrict__ C, int64_t rsC, int64_t csC) {
  typedef MfmaTraits<TIn> Tr;
  constexpr int BM = 128, BN = 128, NT = 256;
  constexpr int V = Tr::V;
  constexpr int KS = BK / Tr::KSTEP;  // MFMA steps per slab
  constexpr int A_PER = BM * BK / NT, B_PER = BN * BK / NT;
  static_assert(BK % Tr::KSTEP == 0, "slab");
  __shared__ __attribute__((aligned(16))) TIn As[2][BM * BK];
  __shared__ __attribute__((aligned(16))) TIn Bs[2][BN * BK];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int64_t tiles_m = ceil_div(m, BM), tiles_n = ceil_div(n, BN);
  const int64_t t = xcd_remap(blockIdx.x, (int64_t)gridDim.x);
  int64_t tm, tn;
  grouped_tile(t, tiles_m, tiles_n, 8, tm, tn);
  const int64_t row0 = tm * BM, col0 = tn * BN;
  const bool m_full = row0 + BM <= m, n_full = col0 + BN <= n;
  const int64_t ktiles = ceil_div(k, BK);

  // LDS offset of slab element (r, kk): [ks][kh][r][e]
  auto lds_off = [](int r, int kk, int rows) {
    const int ks = kk / (2 * V), kh = (kk % (2 * V)) / V, e = kk % V;
    return ((ks * 2 + kh) * rows + r) * V + e;
  };

  TIn ra[A_PER], rb[B_PER];
  auto load_slab = [&](int64_t kt) {
    const int64_t k0 = kt * BK;
    const bool k_full = k0 + BK <= k;
#pragma unroll
    for (int p = 0; p < A_PER; ++p) {
      const int e = tid + NT * p;
      const int i = (LA == kRow) ? e / BK : e % BM;
      const int kk = (LA == kRow) ? e % BK : e / BM;
      const int64_t gi = row0 + i, gk = k0 + kk;
      TIn v = TIn(0.0f);
      if ((m_full && k_full) || (gi < m && gk < k)) v = A[gi * rsA + gk * csA];
      ra[p] = v;
    }
#pragma unroll
    for (int p = 0; p < B_PER; ++p) {
      const int e = tid + NT * p;
      const int j = (LB == kRow) ? e % BN : e / BK;
      const int kk = (LB == kRow) ? e / BN : e % BK;
      const int64_t gj = col0 + j, gk = k0 + kk;
      TIn v = TIn(0.0f);
      if ((n_full && k_full) || (gj < n && gk < k)) v = B[gk * rsB + gj * csB];
      rb[p] = v;
    }
  };
  auto store_slab = [&](int buf) {
#pragma unroll
    for (int p = 0; p < A_PER; ++p) {
      const int e = tid + NT * p;
      const int i = (LA == kRow) ? e / BK : e % BM;
      const int kk = (LA == kRow) ? e % BK : e / BM;
      As[buf][lds_off(i, kk, BM)] = ra[p];
    }
#pragma unroll
    for (int p = 0; p < B_PER; ++p) {
      const int e = tid + NT * p;
      const int j = (LB == kRow) ? e % BN : e / BK;
      const int kk = (LB == kRow) ? e / BN : e % BK;
      Bs[buf][lds_off(j, kk, BN)] = rb[p];
    }
  };

  v16f32 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.0f;

  const int kh = lane >> 5, lr = lane & 31;
  if (ktiles > 0) {
    load_slab(0);
    store_slab(0);
  }
  __syncthreads();
  for (int64_t kt = 0; kt < ktiles; ++kt) {
    const int cur = (int)(kt & 1);
    if (kt + 1 < ktiles) load_slab(kt + 1);
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      typename Tr::frag af[2], bf[2];
#pragma unroll
      for (int a = 0; a < 2; ++a)
        af[a] = *reinterpret_cast<const typename Tr::frag*>(
            &As[cur][((ks * 2 + kh) * BM + wm * 64 + a * 32 + lr) * V]);
#pragma unroll
      for (int b = 0; b < 2; ++b)
        bf[b] = *reinterpret_cast<const typename Tr::frag*>(
            &Bs[cur][((ks * 2 + kh) * BN + wn * 64 + b * 32 + lr) * V]);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = Tr::mma(bf[b], af[a], acc[a][b]);
    }
    if (kt + 1 < ktiles) store_slab(cur ^ 1);
    __syncthreads();
  }

  // lane holds row i = l%32 of each block, columns 8(r/4) + 4(l/32) + r%4
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int64_t gi = row0 + wm * 64 + a * 32 + lr;
    if (gi >= m) continue;
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t gj = col0 + wn * 64 + b * 32 + 8 * (r >> 2) + 4 * kh + (r & 3);
        if (gj < n) {
          float* cp = C + gi * rsC + gj * csC;
          const float s = alpha * acc[a][b][r];
          *cp = (beta == 0.0f) ? s : __builtin_fmaf(beta, *cp, s);
        }
      }
  }
}

// FP32 in, column-major A, B and C with 16-byte aligned columns: a 128 x 128
// tile (4 wavefronts of 64 x 64 as 2 x 2 blocks of v_mfma_f32_32x32x2_f32)
// over 32-deep slabs moved HBM -> LDS by buffer_load ... lds (1 KiB per
// wavefront load, no register staging), two slabs in flight in turn; the
// two stages are distinct LDS objects and the loop is unrolled by two, so
// the reads of one stage never wait for the loads into the other.
//   A slab [32][128]: one wavefront load = two k-rows (lanes 0-31 the first,
//     32-63 the second), four rows of A per lane.
//   B slab: 16 blocks of 8 columns x 32 k; column c of a block keeps its
//     k-quad q at 16-byte slot c*8 + (q ^ c).
// k order inside a slab: MFMA step s (0..15) gives the lanes of half h the
// k-value 8(s/4) + 4h + s%4. So a lane's B values for four steps are one
// 16-byte read (quad 2(s/4) + h of its column; the 8 columns of a group of
// 8 lanes sit in distinct banks), and the A values of a half-wavefront are
// 32 consecutive floats of one k-row.
// Out-of-range elements: the buffer resources cover exactly the matrix
// extents (reads past them return 0); rows >= m of A and columns >= n of B
// may read neighbouring values, which reach only outputs that are not stored;
// the k >= K entries of the last slab are zeroed in LDS after they land.
__device__ __forceinline__ void sgemm_issue(__amdgpu_buffer_rsrc_t ra, int64_t lda,
                                            __amdgpu_buffer_rsrc_t rb, int64_t ldb, int64_t row0,
                                            int64_t col0, int64_t k0, float* As, float* Bs) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int bc = lane >> 3, bq = (lane & 7) ^ (lane >> 3);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int pr = wave * 4 + q;  // k-rows 2pr, 2pr+1
    const int64_t e = (row0 + 4 * (lane & 31)) + (k0 + 2 * pr + (lane >> 5)) * lda;
    dma16(ra, &As[pr * 256], (uint32_t)(e * 4));
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int b = wave * 4 + q;  // columns col0 + 8b ..
    const int64_t e = (col0 + 8 * b + bc) * ldb + k0 + 4 * bq;
    dma16(rb, &Bs[b * 256], (uint32_t)(e * 4));
  }
}

// the k >= kv entries of a landed slab set to zero (the last, partial slab)
__device__ __forceinline__ void sgemm_clear_tail(float* As, float* Bs, int kv) {
  for (int e = threadIdx.x; e < (32 - kv) * 128; e += 256) As[kv * 128 + e] = 0.0f;
  for (int e = threadIdx.x; e < 16 * 256; e += 256) {
    const int slot = (e & 255) >> 2, c = slot >> 3, q = (slot & 7) ^ c;
    if (4 * q + (e & 3) >= kv) Bs[e] = 0.0f;
  }
}

// k-groups [G0, G1) (of four, 8 k-values each) of a landed 32-deep slab
template <int G0, int G1>
__device__ __forceinline__ void sgemm_slab(v16f32 (&acc)[2][2], const float* As, const float* Bs) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1, h = lane >> 5, lr = lane & 31;
#pragma unroll
  for (int g = G0; g < G1; ++g) {
    v4f32 bq[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int j = wn * 64 + b * 32 + lr, c = j & 7, q = 2 * g + h;
      bq[b] = *reinterpret_cast<const v4f32*>(&Bs[(j >> 3) * 256 + (c * 8 + (q ^ c)) * 4]);
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int kk = 8 * g + 4 * h + s;
      float af[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) af[a] = As[kk * 128 + wm * 64 + a * 32 + lr];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = mfma_f32_32x32x2(bq[b][s], af[a], acc[a][b]);
    }
  }
}

__global__ void __launch_bounds__(256, 2)
sgemm_dma_kernel(int64_t m, int64_t n, int64_t k, float alpha, const float* __restrict__ A,
                 int64_t lda, uint32_t a_bytes, const float* __restrict__ B, int64_t ldb,
                 uint32_t b_bytes, float beta, float* __restrict__ C, int64_t ldc) {
  __shared__ __attribute__((aligned(16))) float As0[32 * 128];
  __shared__ __attribute__((aligned(16))) float Bs0[32 * 128];
  __shared__ __attribute__((aligned(16))) float As1[32 * 128];
  __shared__ __attribute__((aligned(16))) float Bs1[32 * 128];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1, h = lane >> 5, lr = lane & 31;
  const int64_t tiles_m = ceil_div(m, 128), tiles_n = ceil_div(n, 128);
  const int64_t t = xcd_remap(blockIdx.x, (int64_t)gridDim.x);
  int64_t tm, tn;
  grouped_tile(t, tiles_m, tiles_n, 8, tm, tn);
  const int64_t row0 = tm * 128, col0 = tn * 128;
  const __amdgpu_buffer_rsrc_t ra = make_rsrc(A, a_bytes), rb = make_rsrc(B, b_bytes);
  const int64_t ktiles = ceil_div(k, 32);

  v16f32 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.0f;

  // slab kt: wait for it, then (everyone done with the other stage) multiply
  // its first half, start slab kt + 1 into the other stage (the matrix cores
  // hold work while the loads are issued) and multiply the second half
#define MOMMS_SG_STAGE(KT, Ac, Bc, An, Bn)                             \
  {                                                                    \
    wait_vmcnt<0>();                                                   \
    lds_barrier();                                                     \
    if ((KT) * 32 + 32 > k) {                                          \
      sgemm_clear_tail(Ac, Bc, (int)(k - (KT) * 32));                  \
      lds_barrier();                                                   \
    }                                                                  \
    sgemm_slab<0, 2>(acc, Ac, Bc);                                     \
    if ((KT) + 1 < ktiles)                                             \
      sgemm_issue(ra, lda, rb, ldb, row0, col0, ((KT) + 1) * 32, An, Bn); \
    sgemm_slab<2, 4>(acc, Ac, Bc);                                     \
  }
  if (ktiles > 0) sgemm_issue(ra, lda, rb, ldb, row0, col0, 0, As0, Bs0);
  for (int64_t kt = 0; kt < ktiles; kt += 2) {
    MOMMS_SG_STAGE(kt, As0, Bs0, As1, Bs1)
    if (kt + 1 < ktiles) MOMMS_SG_STAGE(kt + 1, As1, Bs1, As0, Bs0)
  }
#undef MOMMS_SG_STAGE

  // every C value of a block row read before the first store; lane holds row
  // l%32 of each block, columns 8(r/4) + 4(l/32) + r%4
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int64_t gi = row0 + wm * 64 + a * 32 + lr;
    float cv[32];
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t gj = col0 + wn * 64 + b * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
        const bool ok = gi < m && gj < n;
        cv[b * 16 + r] = (beta != 0.0f) ? C[ok ? gi + gj * ldc : 0] : 0.0f;
      }
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t gj = col0 + wn * 64 + b * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
        if (gi < m && gj < n) {
          const float sv = alpha * acc[a][b][r];
          C[gi + gj * ldc] = (beta == 0.0f) ? sv : __builtin_fmaf(beta, cv[b * 16 + r], sv);
        }
      }
  }
}

// The DMA path's conditions: column-major A, B, C, columns 16-byte aligned
// (leading dimensions multiples of 4, 16-byte aligned bases) and operand
// extents a 32-bit buffer offset can address.
static bool sgemm_dma_ok(int64_t m, int64_t n, int64_t k, const float* A, int64_t rsA,
                         int64_t csA, const float* B, int64_t rsB, int64_t csB, int64_t rsC,
                         uint32_t& a_bytes, uint32_t& b_bytes) {
  if (rsA != 1 || rsB != 1 || rsC != 1 || k == 0) return false;
  if ((csA & 3) || (csB & 3)) return false;
  if ((reinterpret_cast<uintptr_t>(A) & 15) || (reinterpret_cast<uintptr_t>(B) & 15)) return false;
  const uint64_t ea = (uint64_t)((k - 1) * csA + m) * 4;
  const uint64_t eb = (uint64_t)((n - 1) * csB + k) * 4;
  const uint64_t lim = (1ull << 32) - (1ull << 20);
  if (ea > lim || eb > lim) return false;
  a_bytes = (uint32_t)ea;
  b_bytes = (uint32_t)eb;
  return true;
}

// BF16 in, FP32 out, large problems: both operands packed into 256-wide
// panels (pack_bf16_panels, into library scratch), then the DMA-ring panel
// GEMM of panel_gemm.hpp with a 256 x 256 tile and the alpha / beta epilogue.
constexpr int kBfFM = 2, kBfFN = 4, kBfWM = 4, kBfWN = 2, kBfSPS = 2;
typedef PanelGemm<kBfFM, kBfFN, kBfWM, kBfWN, kBfSPS> BfPanel;

__global__ void __launch_bounds__(64 * kBfWM * kBfWN, 1)
bf16_panel_gemm(int64_t m, int64_t n, int64_t tiles_m, int64_t tiles_n, int64_t nslabs,
                const uint16_t* __restrict__ PA, const uint16_t* __restrict__ PB, float alpha,
                float beta, float* __restrict__ C, int64_t rsC, int64_t csC) {
  typedef BfPanel G;
  constexpr int FM = kBfFM, FN = kBfFN, WM = kBfWM;
  __shared__ __attribute__((aligned(16))) uint16_t S0[G::SLOT];
  __shared__ __attribute__((aligned(16))) uint16_t S1[G::SLOT];
  __shared__ __attribute__((aligned(16))) uint16_t S2[G::SLOT];
  __shared__ __attribute__((aligned(16))) uint16_t S3[G::SLOT];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave % WM, wn = wave / WM, kh = lane >> 5;
  const int64_t t = xcd_remap(blockIdx.x, (int64_t)gridDim.x);
  int64_t tm, tn;
  grouped_tile(t, tiles_m, tiles_n, 8, tm, tn);
  v16f32 acc[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.0f;
  panel_mainloop<FM, FN, kBfWM, kBfWN, kBfSPS, false>(acc,
                                                      PA + tm * panel_stride(nslabs) * G::ABLK,
                                                      PB + tn * panel_stride(nslabs) * G::BBLK,
                                                      nslabs, 0.0f,
                                                      0.0f, S0, S1, S2, S3);
#pragma unroll
  for (int a = 0; a < FM; ++a) {
    const int64_t gi = tm * G::BM + wm * 32 * FM + a * 32 + (lane & 31);
    if (gi >= m) continue;
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t gj = tn * G::BN + wn * 32 * FN + b * 32 + 8 * (r >> 2) + 4 * kh + (r & 3);
        if (gj < n) {
          float* cp = C + gi * rsC + gj * csC;
          const float s = alpha * acc[a][b][r];
          *cp = (beta == 0.0f) ? s : __builtin_fmaf(beta, *cp, s);
        }
      }
  }
}

static hipError_t launch_bf16_panels(int64_t m, int64_t n, int64_t k, float alpha,
                                     const uint16_t* A, int64_t rsA, int64_t csA,
                                     const uint16_t* B, int64_t rsB, int64_t csB, float beta,
                                     float* C, int64_t rsC, int64_t csC, hipStream_t st) {
  const int64_t mpad = ceil_div(m, BfPanel::BM) * BfPanel::BM;
  const int64_t npad = ceil_div(n, BfPanel::BN) * BfPanel::BN;
  const int64_t kpad = ceil_div(k, BfPanel::KALIGN) * BfPanel::KALIGN;
  uint16_t* pa = nullptr;
  const int64_t kstride = panel_stride(kpad / kSlab) * kSlab;  // k-values per panel, with the gap
  hipError_t e = stream_scratch(st, (size_t)(mpad + npad) * kstride * 2, (void**)&pa);
  if (e != hipSuccess) return e;
  uint16_t* pb = pa + mpad * kstride;
  dim3 block(256);
  hipLaunchKernelGGL((pack_bf16_panels<BfPanel::BM>),
                     dim3((unsigned)ceil_div(mpad, 64), (unsigned)ceil_div(kpad, 64)), block, 0,
                     st, A, m, k, rsA, csA, mpad, kpad, pa);
  hipLaunchKernelGGL((pack_bf16_panels<BfPanel::BN>),
                     dim3((unsigned)ceil_div(npad, 64), (unsigned)ceil_div(kpad, 64)), block, 0,
                     st, B, n, k, csB, rsB, npad, kpad, pb);
  const int64_t tm = mpad / BfPanel::BM, tn = npad / BfPanel::BN;
  hipLaunchKernelGGL(bf16_panel_gemm, dim3((unsigned)(tm * tn)), dim3(64 * kBfWM * kBfWN), 0, st,
                     m, n, tm, tn, kpad / kSlab, pa, pb, alpha, beta, C, rsC, csC);
  return hipGetLastError();
}

static int layout_of(int64_t rs, int64_t cs, int64_t h, int64_t w) {
  if (rs == 1 || w == 1) return kCol;
  if (cs == 1 || h == 1) return kRow;
  return kGen;
}

template <class TIn, int BK>
static hipError_t launch_f32acc(int64_t m, int64_t n, int64_t k, float alpha, const TIn* A,
                                int64_t rsA, int64_t csA, const TIn* B, int64_t rsB, int64_t csB,
                                float beta, float* C, int64_t rsC, int64_t csC, hipStream_t st) {
  const int la = layout_of(rsA, csA, m, k) == kRow ? kRow : kCol;
  const int lb = layout_of(rsB, csB, k, n) == kRow ? kRow : kCol;
  dim3 grid((unsigned)(ceil_div(m, 128) * ceil_div(n, 128))), block(256);
#define MOMMS_F(LA_, LB_) hipLaunchKernelGGL((gemm_f32acc_kernel<TIn, BK, LA_, LB_>), grid, block, 0, st, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C, rsC, csC)
  if (la == kCol && lb == kCol) MOMMS_F(kCol, kCol);
  else if (la == kCol) MOMMS_F(kCol, kRow);
  else if (lb == kCol) MOMMS_F(kRow, kCol);
  else MOMMS_F(kRow, kRow);
#undef MOMMS_F
  return hipGetLastError();
}

}  // namespace momms

using namespace momms;

template <class T>
static int f32acc_entry(const char* fn, int64_t m, int64_t n, int64_t k, const float* alpha,
                        const T* A, int64_t rsA, int64_t csA, const T* B, int64_t rsB,
                        int64_t csB, const float* beta, float* C, int64_t rsC, int64_t csC,
                        void* stream) {
  if (m < 0) return bad_arg(fn, 1, "m < 0");
  if (n < 0) return bad_arg(fn, 2, "n < 0");
  if (k < 0) return bad_arg(fn, 3, "k < 0");
  if (alpha == nullptr) return bad_arg(fn, 4, "alpha is null");
  int st;
  if ((st = check_operand(fn, 5, A, m, k, rsA, csA))) return st;
  if ((st = check_operand(fn, 8, B, k, n, rsB, csB))) return st;
  if (beta == nullptr) return bad_arg(fn, 11, "beta is null");
  if ((st = check_operand(fn, 12, C, m, n, rsC, csC))) return st;
  if (m == 0 || n == 0) return 0;
  constexpr int BK = std::is_same<T, float>::value ? 16 : 32;
  // the packed-panel form once the 256 x 256 tiles cover the device and k is
  // deep enough to repay the packing pass
  if constexpr (std::is_same<T, float>::value) {
    uint32_t ab = 0, bb = 0;
    if (ceil_div(m, 128) * ceil_div(n, 128) >= 256 &&
        sgemm_dma_ok(m, n, k, A, rsA, csA, B, rsB, csB, rsC, ab, bb)) {
      hipLaunchKernelGGL(sgemm_dma_kernel, dim3((unsigned)(ceil_div(m, 128) * ceil_div(n, 128))),
                         dim3(256), 0, (hipStream_t)stream, m, n, k, *alpha, A, csA, ab, B, csB,
                         bb, *beta, C, csC);
      return hip_status(fn, hipGetLastError());
    }
  }
  if (!std::is_same<T, float>::value && ceil_div(m, 256) * ceil_div(n, 256) >= 128 && k >= 256)
    return hip_status(fn, launch_bf16_panels(m, n, k, *alpha, reinterpret_cast<const uint16_t*>(A),
                                             rsA, csA, reinterpret_cast<const uint16_t*>(B), rsB,
                                             csB, *beta, C, rsC, csC, (hipStream_t)stream));
  return hip_status(fn, launch_f32acc<T, BK>(m, n, k, *alpha, A, rsA, csA, B, rsB, csB, *beta, C,
                                             rsC, csC, (hipStream_t)stream));
}

extern "C" int momms_hip_sgemm(int64_t m, int64_t n, int64_t k, const float* alpha,
                               const float* A, int64_t rsA, int64_t csA, const float* B,
                               int64_t rsB, int64_t csB, const float* beta, float* C,
                               int64_t rsC, int64_t csC, void* stream) {
  return f32acc_entry<float>("momms_hip_sgemm", m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta,
                             C, rsC, csC, stream);
}

extern "C" int momms_hip_gemm_bf16(int64_t m, int64_t n, int64_t k, const float* alpha,
                                   const uint16_t* A, int64_t rsA, int64_t csA,
                                   const uint16_t* B, int64_t rsB, int64_t csB,
                                   const float* beta, float* C, int64_t rsC, int64_t csC,
                                   void* stream) {
  return f32acc_entry<__bf16>("momms_hip_gemm_bf16", m, n, k, alpha,
                              reinterpret_cast<const __bf16*>(A), rsA, csA,
                              reinterpret_cast<const __bf16*>(B), rsB, csB, beta, C, rsC, csC,
                              stream);
}
