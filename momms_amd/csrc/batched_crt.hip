// Batched FP64-accurate GEMMs in the residue form (Ozaki scheme II, the
// arithmetic of crt_bf16.hip): one persistent launch for a whole table of
// uneven problems, every problem's result bit-identical to
// momms_hip_dgemm_crt_bf16 with the same N moduli.
//
// No reference counterpart: the batching and precision additions together.
// The single-GEMM form streams its operands through HBM three times (residue
// pack, per-modulus products, reconstruction), which suits large problems;
// for thousands of small ones those passes would move several times the
// bytes of the operands. Here a 64 x 64 tile does everything on chip:
//   - it reads its 64 x 16 slab of A and 16 x 64 slab of B in FP64 (the
//     exponents of the rows of A and columns of B come from a pre-pass,
//     batched_exponents), turns every element into its integer image
//     rint(x 2^(beta - E)) and that into its balanced residue modulo each
//     p_l, and stores the residues as BF16 in LDS;
//   - each wavefront (32 x 32 of the tile) runs one 32x32x16 BF16 MFMA per
//     modulus and slab, the N accumulators held in registers; any 256
//     consecutive k-values add at most 256 * 255^2 to an accumulator that
//     starts below 0.51 p, so every partial sum is an exact FP32 integer, and
//     each is reduced modulo p_l every 16 slabs and at the end;
//   - the epilogue turns the N residues of every entry into C' (the Chinese
//     remainder sum of crt_recon.hpp), converts, scales and updates C.
// Rows of A or columns of B holding Inf or NaN take residue 0 and their
// entries are recomputed as alpha * (fma chain over k), as in crt_bf16.hip.
#include <hip/hip_runtime.h>
#include <math.h>

#include <type_traits>

#include "batch_tiles.hpp"
#include "batched.hpp"
#include "crt_common.hpp"
#include "crt_recon.hpp"
#include "dgemm_tile.hpp"
#include "mfma.hpp"
#include "split_tile.hpp"

namespace momms {

// f(std::integral_constant<int, i>) for i = 0 .. N-1, unrolled at compile time.
template <int I, int N, typename F>
__device__ __forceinline__ void static_for_from(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for_from<I + 1, N>(f);
  }
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_from<0, N>(f);
}

// LDS of a tile: a ring of three FP64 slab pairs filled by buffer -> LDS
// DMA (A: [16 k][64 rows], B: [64 columns][16 k], doubles, as in memory),
// and the BF16 residues of one slab pair: for modulus l, operand o (0: rows
// of A, 1: columns of B), k-half h and vector v (row or column of the tile),
// the 8 k-values 8h .. 8h+7 at r[l][o][h][v], which is the lane layout of a
// 32x32x16 BF16 MFMA operand (lane: vector l % 32, k-half l / 32), read as
// one 16-byte load per lane.
constexpr int kCrtRing = 3;
template <int N>
struct CrtBatchSmem {
  union {
    struct {
      double a[kCrtRing][16][64];
      double b[kCrtRing][64][16];
      __attribute__((aligned(16))) uint16_t r[N][2][2][64][8];
    };
    // after the k loop: the balanced residues of entry (row i, column j) of
    // the tile, moduli 0..15 (int16), at d[j][i]
    __attribute__((aligned(16))) int16_t d[64][64][16];
  };
};

// buffer -> LDS DMA of 4 bytes per lane (lane l's dword lands at lds + 4 l)
// at byte offset voff of the buffer rsrc (past its records: zero, never a
// fault), issued so that the compiler does not track it (as
// lds_dma16_untracked in panel_gemm.hpp, whose notes on m0 hold here too):
// the tile waits for it with wait_vmcnt before any wavefront reads the
// destination.
typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4u32 rsrc_words(const void* base, uint32_t bytes) {
  const uint64_t a = (uint64_t)(uintptr_t)base;
  v4u32 r;
  r.x = __builtin_amdgcn_readfirstlane((uint32_t)a);
  r.y = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32) & 0xFFFFu);  // stride 0
  r.z = __builtin_amdgcn_readfirstlane(bytes);
  r.w = 0x00020000u;  // the word make_rsrc uses
  return r;
}

#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void lds_dma4_untracked(v4u32 rsrc, uint32_t voff, const void* lds) {
  const uint32_t m0v = __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)lds);
  asm volatile("s_mov_b32 m0, %2\n\ts_nop 0\n\tbuffer_load_dword %0, %1, 0 offen lds"
               :
               : "v"(voff), "s"(rsrc), "s"(m0v)
               : "memory", "m0");
}
#pragma clang diagnostic pop

__device__ __forceinline__ uint32_t clamp_bytes(int64_t elems) {
  const int64_t b = elems > 0 ? elems * 8 : 0;
  return (uint32_t)(b < 0x7FFFFFFF ? b : 0x7FFFFFFF);
}

// Residues without fp64 division: the integer image q (|q| < 2^53) is
// taken as a sign and five unsigned 11-bit digits, |q| = sum_i d_i 2^(11 i),
// and its residue modulo p as x = sum_i (+-d_i) c_i with c_i = 2^(11 i) mod p
// balanced (|c_i| <= (p-1)/2 <= 255): every partial sum is an integer below
// 2047 (1 + 4 * 255) < 2^21, exact in fp32. Then x - p rint(x / p) is the
// balanced residue exactly: fp32(x * fp32(1/p)) is within |x/p| 2^-23 (1 +
// 2^-24) < 6e-4 of x / p, nearer than the 1/(2p) >= 9.7e-4 by which x / p
// (p odd) stays clear of a half-integer, so the rounding picks the nearest
// integer and |r| <= (p-1)/2 <= 255, an integer bf16 holds exactly (the high
// half of its float).
typedef _Float16 h2 __attribute__((ext_vector_type(2)));

struct CrtDigitTab {
  float c[kCrtMaxModuli][5];
};

constexpr CrtDigitTab make_crt_digit_tab(const int* mods) {
  CrtDigitTab t{};
  for (int l = 0; l < kCrtMaxModuli; ++l) {
    const int p = mods[l];
    int64_t w = 1;
    for (int i = 0; i < 5; ++i) {
      int r = (int)(w % p);
      if (r > (p - 1) / 2) r -= p;
      t.c[l][i] = (float)r;
      w = (w << 11) % p;
    }
  }
  return t;
}
constexpr CrtDigitTab kCrtDigits = make_crt_digit_tab(kCrtModuli);

// The signed 11-bit digits of q: d0 as a float, (d1, d2) and (d3, d4) as
// fp16 pairs (integers up to 2048 are exact in fp16).
struct CrtDigits {
  float d0;
  h2 d12, d34;
};

__device__ __forceinline__ CrtDigits crt_digits(double q) {
  const double a = __builtin_fabs(q);
  const uint32_t hi = (uint32_t)(a * 0x1p-32);              // floor(|q| / 2^32) < 2^21
  const uint32_t lo = (uint32_t)__builtin_fma(-(double)hi, 0x1p32, a);
  const float sg = q < 0.0 ? -1.0f : 1.0f;
  CrtDigits d;
  d.d0 = sg * (float)(lo & 2047u);
  d.d12 = h2{(_Float16)(sg * (float)((lo >> 11) & 2047u)),
             (_Float16)(sg * (float)(((lo >> 22) | (hi << 10)) & 2047u))};
  d.d34 = h2{(_Float16)(sg * (float)((hi >> 1) & 2047u)), (_Float16)(sg * (float)(hi >> 12))};
  return d;
}

// Float bits of the balanced residue modulo the l-th modulus of the number
// whose digits are d (its BF16 bits are the high half). The two fp16 dot
// products see exact integer products below 2^22 and partial sums below
// 2^21, so they are exact whatever their internal order.
template <int L>
__device__ __forceinline__ uint32_t crt_residue_bits(const CrtDigits& d) {
  constexpr const CrtDigitTab& D = kCrtDigits;
  constexpr float pf = (float)kCrtModuli[L], ipf = 1.0f / (float)kCrtModuli[L];
  constexpr h2 c12 = h2{(_Float16)D.c[L][1], (_Float16)D.c[L][2]};
  constexpr h2 c34 = h2{(_Float16)D.c[L][3], (_Float16)D.c[L][4]};
  const float x = __builtin_amdgcn_fdot2(d.d34, c34, __builtin_amdgcn_fdot2(d.d12, c12, d.d0, false), false);
  const float r = __builtin_fmaf(-__builtin_rintf(x * ipf), pf, x);
  return __float_as_uint(r);
}

// The moduli tables as compile-time constants (instruction literals).
constexpr CrtTables kCrtBatchTab = make_crt_tables(kCrtModuli);

template <int N>
__device__ __forceinline__ void crt_batch_tile(CrtBatchSmem<N>& sm, const momms_gemm_desc& p,
                                               const SplitProb& sp, int64_t row0, int64_t col0,
                                               double alpha, const double* __restrict__ A,
                                               const double* __restrict__ B, double beta,
                                               double* __restrict__ C,
                                               const char* __restrict__ ws) {
  constexpr const CrtTables& T = kCrtBatchTab;
  // eight wavefronts: wave w owns the 32 x 32 block (w % 2, (w / 2) % 2) of
  // the tile for moduli [0, N0) (w < 4) or [N0, N) (w >= 4), so a wave holds
  // at most N0 accumulator blocks and two waves share each SIMD
  constexpr int N0 = (N + 1) / 2;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = (wave >> 1) & 1, mh = wave >> 2;
  const int64_t m = p.m, n = p.n, k = p.k, lda = p.lda, ldb = p.ldb, ldc = p.ldc;
  const int bits = sp.beta;
  const int* EA = reinterpret_cast<const int*>(ws + sp.ea);
  const int* EB = reinterpret_cast<const int*>(ws + sp.eb);

  // conversion roles: threads 0..255 convert A[row0 + t%64][k0 + 4(t/64) + e],
  // threads 256..511 convert B[k0 + 4(t%4) + e][col0 + (t-256)/4], e < 4
  const bool isb = tid >= 256;
  const int t2 = tid & 255;
  const int vv = isb ? (t2 >> 2) : (t2 & 63), vq = isb ? (t2 & 3) : (t2 >> 6);
  const int64_t gv = (isb ? col0 : row0) + vv;
  // the exponent tables hold mpad and npad (multiples of 64) entries
  const int ev = isb ? EB[gv] : EA[gv];
  const bool vfin = gv < (isb ? n : m) && ev != kNonFinite;
  const int64_t ktiles = ceil_div(k, 16);

  // DMA roles: wavefront w moves A columns 2w, 2w+1 (two 256-byte halves
  // each) and B columns 8w .. 8w+7 (two per 256 bytes) of every slab: four
  // A and four B wavefront loads per slab. The buffer records cover the
  // problem's block from the slab's corner, so no read leaves it; what lands
  // for a row past m, a column past n or a k past k (the two slabs issued
  // past k read nothing and land zeros) is never converted (stage() masks it).
  const uint32_t la8 = (uint32_t)lda * 8u, lb8 = (uint32_t)ldb * 8u;
  const uint32_t va = (uint32_t)(lane >> 1) * 8u + 4u * (lane & 1);
  const uint32_t vb = (uint32_t)(lane >> 5) * lb8 + (uint32_t)((lane & 31) >> 1) * 8u + 4u * (lane & 1);
  const int64_t a_end = (k - 1) * lda + m, b_end = (n - 1) * ldb + k;
  auto issue = [&](int64_t kt, int st) {
    const int64_t k0 = kt * 16;
    const int64_t aoff = row0 + k0 * lda, boff = k0 + col0 * ldb;
    const v4u32 ra = rsrc_words(A + p.a_off + aoff, clamp_bytes(a_end - aoff));
    const v4u32 rb = rsrc_words(B + p.b_off + boff, clamp_bytes(b_end - boff));
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int kk = 2 * wave + (q >> 1), half = q & 1;
      lds_dma4_untracked(ra, va + (uint32_t)(32 * half) * 8u + (uint32_t)kk * la8, &sm.a[st][kk][32 * half]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = 8 * wave + 2 * q;
      lds_dma4_untracked(rb, vb + (uint32_t)c * lb8, &sm.b[st][c][0]);
    }
  };

  auto stage = [&](int st, int64_t kt) {
    const int kleft = (int)(k - kt * 16 < 16 ? k - kt * 16 : 16);
    double x[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int kk = 4 * vq + e;
      const double v = isb ? sm.b[st][vv][kk] : sm.a[st][kk][vv];
      x[e] = (vfin && kk < kleft) ? v : 0.0;
    }
    CrtDigits dg[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) dg[e] = crt_digits(__builtin_rint(ldexp(x[e], bits - ev)));
    static_for<N>([&](auto L) {
      constexpr int l = decltype(L)::value;
      uint32_t rr[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) rr[e] = crt_residue_bits<l>(dg[e]);
      // high halves of two floats into one word (v_perm_b32)
      uint2 w;
      w.x = __builtin_amdgcn_perm(rr[1], rr[0], 0x07060302u);
      w.y = __builtin_amdgcn_perm(rr[3], rr[2], 0x07060302u);
      *reinterpret_cast<uint2*>(&sm.r[l][isb ? 1 : 0][vq >> 1][vv][(vq & 1) * 4]) = w;
      __builtin_amdgcn_sched_barrier(0);  // one modulus at a time (registers)
    });
  };

  v16f32 acc[N0];
#pragma unroll
  for (int l = 0; l < N0; ++l)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[l][r] = 0.0f;

  // accumulator block l of this wave holds modulus l (mh = 0) or N0 + l
  // (mh = 1; for odd N the last block of that half is a spare that repeats
  // modulus N0 - 1 and is never read out)
  auto reduce = [&]() {
#pragma unroll
    for (int l = 0; l < N0; ++l) {
      const float pf = mh ? T.pf[N0 + l < N ? N0 + l : N0 - 1] : T.pf[l];
      const float ipf = mh ? T.invpf[N0 + l < N ? N0 + l : N0 - 1] : T.invpf[l];
#pragma unroll
      for (int r = 0; r < 16; ++r)
        acc[l][r] = __builtin_fmaf(-__builtin_rintf(acc[l][r] * ipf), pf, acc[l][r]);
      __builtin_amdgcn_sched_barrier(0);
    }
  };

  const int vr = lane & 31, kh = lane >> 5;
  auto multiply = [&]() {
#pragma unroll
    for (int l = 0; l < N0; ++l) {
      // the second half has N - N0 moduli; its spare block (N odd) multiplies
      // modulus N0 - 1's residues again and is never read
      const int g = mh ? (N0 + l < N ? N0 + l : N0 - 1) : l;
      const v8bf16 af = *reinterpret_cast<const v8bf16*>(&sm.r[g][0][kh][wm * 32 + vr][0]);
      const v8bf16 bf = *reinterpret_cast<const v8bf16*>(&sm.r[g][1][kh][wn * 32 + vr][0]);
      acc[l] = mfma_bf16_32x32x16(bf, af, acc[l]);
    }
  };

  // slab kt: its DMAs were issued two slabs earlier into ring slot kt % 3;
  // each wave has the 8 loads of slab kt+1 still in flight when it waits for
  // slab kt. Slot (kt+2) % 3 was last read by stage(kt-1), and the residues
  // by multiply(kt-1), both before this slab's first barrier.
  if (ktiles > 0) {
    issue(0, 0);
    issue(1, 1);
    int st = 0;
    for (int64_t kt = 0; kt < ktiles; ++kt) {
      wait_vmcnt<8>();
      lds_barrier();
      issue(kt + 2, st == 0 ? 2 : st - 1);
      stage(st, kt);
      lds_barrier();
      multiply();
      if ((kt & 15) == 15) reduce();
      st = st == 2 ? 0 : st + 1;
    }
    wait_vmcnt<0>();  // the two slabs issued past k (zeros) have landed
  }
  reduce();
  lds_barrier();  // every wave is done with the ring and the residues

  // balanced residues of the wave's 32 x 32 entries and moduli to LDS:
  // register r of lane holds entry (row wm*32 + l%32, column wn*32 + 8(r/4)
  // + 4(l/32) + r%4); two moduli per 4-byte store
  {
    const int il = wm * 32 + vr;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int jl = wn * 32 + 8 * (r >> 2) + 4 * kh + (r & 3);
      int32_t v[N0 + 1];
#pragma unroll
      for (int l = 0; l < N0; ++l) {
        const float pf = mh ? T.pf[N0 + l < N ? N0 + l : N0 - 1] : T.pf[l];
        const float hf = 0.5f * (pf - 1.0f);
        float x = acc[l][r];
        x = x > hf ? x - pf : x;
        x = x < -hf ? x + pf : x;
        v[l] = (int32_t)x;
      }
      v[N0] = 0;
      const int base = mh ? N0 : 0, cnt = mh ? N - N0 : N0;
#pragma unroll
      for (int l = 0; l < N0; l += 2) {
        if (l < cnt) {
          const uint32_t w = (uint32_t)(v[l] & 0xFFFF) | (l + 1 < cnt ? ((uint32_t)v[l + 1] << 16) : 0u);
          if (l + 1 < cnt)
            *reinterpret_cast<uint32_t*>(&sm.d[jl][il][base + l]) = w;
          else
            sm.d[jl][il][base + l] = (int16_t)(w & 0xFFFF);
        }
      }
    }
  }
  lds_barrier();

  // entry e = tid + 512u of the tile (row e % 64, column e / 64): a
  // wavefront updates 64 consecutive rows of one column at a time
  constexpr CrtRecon K = crt_recon_constants(N, kCrtModuli);
  const int il = tid & 63;
  const int64_t i = row0 + il;
  if (i < m) {
    const int ei = EA[i];
    double* Crow = C + p.c_off + i;
    int64_t j = col0 + (tid >> 6);
    // the next entry's exponent and C value are loaded while this one is
    // reconstructed
    int ejn = j < n ? EB[j] : 0;
    double cn = (beta != 0.0 && j < n) ? Crow[j * ldc] : 0.0;
#pragma unroll 1
    for (int u = 0; u < 8 && j < n; ++u, j += 8) {
      const int ej = ejn;
      const double c0 = cn;
      if (j + 8 < n) {
        ejn = EB[j + 8];
        if (beta != 0.0) cn = Crow[(j + 8) * ldc];
      }
      double v;
      if (ei == kNonFinite || ej == kNonFinite) {
        double s = 0.0;
        for (int64_t q = 0; q < k; ++q)
          s = __builtin_fma(A[p.a_off + i + q * lda], B[p.b_off + q + j * ldb], s);
        v = alpha * s;
      } else {
        const int jl = (int)(j - col0);
        int16_t rr[16];
        *reinterpret_cast<uint4*>(&rr[0]) = *reinterpret_cast<const uint4*>(&sm.d[jl][il][0]);
        *reinterpret_cast<uint4*>(&rr[8]) = *reinterpret_cast<const uint4*>(&sm.d[jl][il][8]);
        CrtSums a;
        crt_sums_init(a);
#pragma unroll
        for (int l = 0; l < N; ++l) crt_sums_add(a, (int)rr[l], l, K);
        v = alpha * ldexp(crt_i128_to_f64(crt_sums_value(a, K)), ei + ej - 2 * bits);
      }
      Crow[j * ldc] = beta == 0.0 ? v : __builtin_fma(beta, c0, v);
    }
  }
  lds_barrier();  // LDS free for the next tile of the persistent loop
}

// One workgroup of eight wavefronts per CU (two per SIMD): up to 8
// accumulator blocks of 16 registers a wave, 128 KiB of LDS.
template <int N>
__global__ void __launch_bounds__(512, 1)
batched_crt_kernel(const momms_gemm_desc* __restrict__ d, const SplitProb* __restrict__ sp,
                   const BatchItem* __restrict__ items, QueueBounds qb, unsigned* counters,
                   double alpha, const double* __restrict__ A, const double* __restrict__ B,
                   double beta, double* __restrict__ C, const char* __restrict__ ws) {
  __shared__ CrtBatchSmem<N> sm;
  run_queues(qb, counters, [&](int it) {
    const BatchItem w = items[uniform(it)];
    const int prob = uniform(w.prob);
    const momms_gemm_desc p = d[prob];
    const SplitProb s = sp[prob];
    crt_batch_tile<N>(sm, p, s, (int64_t)uniform(w.tm) * 64, (int64_t)uniform(w.tn) * 64, alpha,
                      A, B, beta, C, ws);
  });
}

#define MOMMS_CRT_BATCH_SWITCH(N_, CALL) \
  switch (N_) {                          \
    case 1: CALL(1); break;              \
    case 2: CALL(2); break;              \
    case 3: CALL(3); break;              \
    case 4: CALL(4); break;              \
    case 5: CALL(5); break;              \
    case 6: CALL(6); break;              \
    case 7: CALL(7); break;              \
    case 8: CALL(8); break;              \
    case 9: CALL(9); break;              \
    case 10: CALL(10); break;            \
    case 11: CALL(11); break;            \
    case 12: CALL(12); break;            \
    case 13: CALL(13); break;            \
    case 14: CALL(14); break;            \
    case 15: CALL(15); break;            \
    default: CALL(16); break;            \
  }

const void* batched_crt_kernel_ptr(int N) {
  const void* f = nullptr;
#define MOMMS_P(N_) f = reinterpret_cast<const void*>(&batched_crt_kernel<N_>)
  MOMMS_CRT_BATCH_SWITCH(N, MOMMS_P)
#undef MOMMS_P
  return f;
}

void launch_batched_crt(int N, int grid, hipStream_t st, const momms_gemm_desc* d,
                        const SplitProb* sp, const BatchItem* items, QueueBounds qb,
                        unsigned* counters, double alpha, const double* A, const double* B,
                        double beta, double* C, const char* ws) {
#define MOMMS_L(N_)                                                                          \
  hipLaunchKernelGGL((batched_crt_kernel<N_>), dim3((unsigned)grid), dim3(512), 0, st, d, sp, \
                     items, qb, counters, alpha, A, B, beta, C, ws)
  MOMMS_CRT_BATCH_SWITCH(N, MOMMS_L)
#undef MOMMS_L
}

}  // namespace momms
