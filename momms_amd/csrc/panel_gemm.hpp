// BF16 (or INT8) matrix-core GEMM over packed panels, shared by the residue
// split GEMMs (crt_bf16.hip, one BF16 GEMM per modulus, FP32 sums reduced
// modulo p; crt_i8.hip, one INT8 GEMM per modulus, exact INT32 sums) and the
// BF16 -> FP32 GEMM (gemm_f32acc.hip, alpha/beta epilogue). An INT8 panel has
// the byte layout of a BF16 one, each 16-byte group holding 16 int8 k-values
// instead of 8 bf16 ones: a slab is 32 k-values deep.
//
// Panel layout (written by the pack kernels): for a panel of PANEL rows of A
// (columns of B) and 16-deep k-slab ks, one contiguous block [kh][PANEL][8]
// of bf16 (kh = which 8-wide half of the slab); the slabs of a panel are
// contiguous, and panel pv starts at slab pv * (nslabs + 1) (panel_stride):
// one unused slab between panels. Without it every panel is a power of two
// long at power-of-two k, so the slabs that the concurrently running tiles
// read at the same k all share their low address bits and, depending on the
// physical pages, compete for the same L2 sets (measured: 2.7x the HBM
// fetches and 6 % of the residue GEMMs' time).
//
// Workgroup: WM x WN wavefronts, each FM x FN blocks of 32 x 32 (tile BM x BN,
// BM = 32*FM*WM, BN = 32*FN*WN). A stage is SPS consecutive slabs of A and
// B; stages move HBM -> LDS by global_load_lds_dwordx4 (1 KiB per wavefront
// load, no register staging) through a ring of four LDS slots (distinct
// __shared__ objects, A then B in each), three stages in flight while one is
// multiplied: at stage KS's barrier every wavefront has finished reading slot
// KS % 4, and the DMA of stage KS + 4 is issued into it.
// The DMA is issued from inline assembly, so the compiler does not see an LDS
// store in flight and adds no waits of its own before the fragment reads (it
// would otherwise wait for every DMA before any read once the ring has more
// slots than its alias tracking follows); the loop waits explicitly: before
// stage KS + 1 is read, vmcnt <= 2 stages (KS + 2, KS + 3 may be in flight),
// then a barrier. Past the last stage the ring is kept full with reloads of
// stage 0 into slots nobody reads again, so every wait has the same count;
// the loop ends with vmcnt(0), so no DMA outlives the workgroup's LDS.
// Fragments are software-pipelined: a wavefront holds two register sets and
// reads the fragments of slab s+1 between the MFMAs of slab s (one read after
// each MFMA), so LDS latency hides behind the matrix cores. The stage barrier
// sits before the last slab's MFMAs; after it the first fragments of the
// next stage are read, and the DMA of stage KS + 4 is issued one load at a
// time between the last slab's final MFMAs, so the matrix cores hold queued
// work while the loads go out (issued all at once after the barrier by every
// wavefront, they left the matrix cores idle: 3-5 % of the residue GEMMs'
// time). The slab that carries a block's modular reduction places its vector
// instructions between the MFMAs explicitly (panel_slab_order_valu). The MFMA
// order per accumulator is slab order, as in any k loop over the same panels.
// The MFMA takes the B block as its "A" operand, so lane l holds row
// i = l%32 of a block and columns 8(r/4) + 4(l/32) + r%4: column-major
// stores along i are coalesced.
#pragma once
#include <type_traits>

#include "mfma.hpp"
#include "split_tile.hpp"

namespace momms {

// slabs from one panel's start to the next (one unused slab between panels)
__host__ __device__ constexpr int64_t panel_stride(int64_t nslabs) { return nslabs + 1; }

template <int FM, int FN, int WM, int WN, int SPS>
struct PanelGemm {
  static constexpr int NW = WM * WN, BM = 32 * FM * WM, BN = 32 * FN * WN;
  static constexpr int ABLK = BM * kSlab, BBLK = BN * kSlab;   // bf16 per slab block
  static constexpr int LA = ABLK / 512 * SPS, LB = BBLK / 512 * SPS;  // 1 KiB loads per stage
  static constexpr int LOADS = LA + LB;
  static_assert(LOADS % NW == 0, "every wavefront issues the same number of loads");
  static constexpr int LW = LOADS / NW;
  static constexpr int KALIGN = kSlab * SPS;  // k padding of the panels
  static constexpr int SLOT = (ABLK + BBLK) * SPS;  // bf16 per LDS ring slot: A, then B
  static constexpr int BOFF = ABLK * SPS;           // B's offset in a slot
};

// global -> LDS DMA of 16 bytes per lane (a wavefront writes 1 KiB at lds,
// lane l at lds + 16 l), issued so that the compiler does not track it: the
// caller waits for it (vmcnt) before any wavefront reads the destination.
// No VGPR is written by the instruction. m0 is declared clobbered; the
// compiler reserves m0 and warns that it may hold a value of its own across
// the statement: the kernels built on this loop keep nothing in m0 (their
// only m0 writes are these statements, each setting it right before its use).
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void lds_dma16_untracked(const void* src, const void* lds) {
  const uint32_t m0v = __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)lds);
  asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off"
               :
               : "v"(src), "s"(m0v)
               : "memory", "m0");
}
#pragma clang diagnostic pop

// The matrix-core operation on 16-byte fragments of a slab: BF16 (a 16-deep
// slab, FP32 sums) or INT8 (the same bytes hold a 32-deep slab, exact INT32
// sums). Both put lane l's fragment at row l%32 and k-half l/32 of a 32-row
// block, and both return a 32 x 32 block in the same lane layout.
template <bool I8>
struct PanelOps {
  typedef v8bf16 frag;
  typedef v16f32 acc;
  static __device__ __forceinline__ acc mma(frag a, frag b, acc c) {
    return mfma_bf16_32x32x16(a, b, c);
  }
};
template <>
struct PanelOps<true> {
  typedef v4i32 frag;
  typedef v16i32 acc;
  static __device__ __forceinline__ acc mma(frag a, frag b, acc c) {
    return mfma_i8_32x32x32(a, b, c);
  }
};

template <int FM, int FN>
__device__ __forceinline__ void panel_reduce_mod(v16f32 (&acc)[FM][FN], float p, float invp) {
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const float x = acc[a][b][r];
        acc[a][b][r] = __builtin_fmaf(-__builtin_rintf(x * invp), p, x);
      }
}

// One slab's MFMAs with the next slab's fragment reads placed between them:
// MFMA, then (read, MFMA) pairs, then the remaining MFMAs. The reads name the
// other register set, so only the waits at the region's start (for reads of
// the previous region) remain.
template <int FM, int FN>
__device__ __forceinline__ void panel_slab_order() {
  __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
#pragma unroll
  for (int q = 0; q < FM + FN; ++q) {
    __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // DS read
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
  }
  if constexpr (FM * FN > 1 + FM + FN)
    __builtin_amdgcn_sched_group_barrier(0x008, FM * FN - 1 - FM - FN, 0);
}

// The same with up to VPER vector ALU instructions after each MFMA (the slab
// that carries a block's modular reduction: its 48 instructions spread over
// the shadows of the other blocks' MFMAs instead of bunching where the
// compiler puts them; the reduced block's MFMA comes last in the slab).
template <int FM, int FN, int VPER>
__device__ __forceinline__ void panel_slab_order_valu() {
#pragma unroll
  for (int q = 0; q < FM + FN; ++q) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);     // MFMA
    __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);     // DS read
    __builtin_amdgcn_sched_group_barrier(0x002, VPER, 0);  // VALU
  }
#pragma unroll
  for (int q = FM + FN; q < FM * FN; ++q) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);     // MFMA
    __builtin_amdgcn_sched_group_barrier(0x002, VPER, 0);  // VALU
  }
}

// The k loop: acc += A_panel * B_panel over nslabs slabs (a multiple of SPS).
// With MOD, each accumulator x is kept an exact FP32 integer by replacing it
// with x - p rint(x / p) (exact by fma for |x| < 2^24) once every 256
// k-values. The reductions are staggered: the loop is unrolled by one chunk
// (CHUNK_STAGES stages = 256 k), and at stage j of a chunk the blocks q with
// q * CHUNK_STAGES / (FM FN) == j are reduced before that stage's first slab,
// whose MFMAs then start with the next block, so the VALU work overlaps the
// other blocks' MFMAs instead of stalling both wavefronts of a SIMD at the
// chunk end. Between two reductions of a block lie exactly 256 k-values
// (|x| < 0.51 p + 256 * 255^2 < 2^24), and the caller reduces once more after
// the loop, so the residues equal those of a reduction at every chunk end.
template <int FM, int FN, int WM, int WN, int SPS, bool MOD, bool I8 = false>
__device__ __forceinline__ void panel_mainloop(typename PanelOps<I8>::acc (&acc)[FM][FN],
                                               const uint16_t* __restrict__ Abase,
                                               const uint16_t* __restrict__ Bbase,
                                               int64_t nslabs, float p, float invp, uint16_t* S0,
                                               uint16_t* S1, uint16_t* S2, uint16_t* S3) {
  typedef PanelGemm<FM, FN, WM, WN, SPS> G;
  constexpr int CHUNK_STAGES = 256 / kSlab / SPS;
  constexpr int NQ = FM * FN;
  static_assert((256 / kSlab) % SPS == 0 && CHUNK_STAGES % 4 == 0 && CHUNK_STAGES <= 16,
                "a chunk is whole stages and whole turns of the four-slot ring");
  static_assert(!MOD || NQ >= CHUNK_STAGES, "every stage of a chunk reduces some block");
  static_assert(!(MOD && I8), "INT8 sums are exact without reductions");
  static_assert(G::LW <= NQ, "a stage's loads go between the last slab's MFMAs");
  typedef typename PanelOps<I8>::frag Frag;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int64_t nst = nslabs / SPS;

  // load i (of LW) of this wavefront's part of stage kst, into ring slot S
  auto issue_one = [&](int64_t kst, uint16_t* S, int i) {
    if (kst >= nst) kst = 0;  // keeps the ring full: stage 0 again, into a slot no one reads
    const int c = wave + G::NW * i;
    if (c < G::LA)
      lds_dma16_untracked(Abase + kst * SPS * G::ABLK + c * 512 + lane * 8, &S[c * 512]);
    else
      lds_dma16_untracked(Bbase + kst * SPS * G::BBLK + (c - G::LA) * 512 + lane * 8,
                          &S[G::BOFF + (c - G::LA) * 512]);
  };
  auto issue = [&](int64_t kst, uint16_t* S) {
#pragma unroll
    for (int i = 0; i < G::LW; ++i) issue_one(kst, S, i);
  };

  const int kh = lane >> 5;
  const int a_off = kh * G::BM * 8 + (wm * 32 * FM + (lane & 31)) * 8;
  const int b_off = G::BOFF + kh * G::BN * 8 + (wn * 32 * FN + (lane & 31)) * 8;
  Frag af[FM], bf[FN], an[FM], bn[FN];
  auto rd = [&](const uint16_t* S, int ss, Frag (&xa)[FM], Frag (&xb)[FN]) {
#pragma unroll
    for (int a = 0; a < FM; ++a)
      xa[a] = *reinterpret_cast<const Frag*>(&S[ss * G::ABLK + a_off + a * 32 * 8]);
#pragma unroll
    for (int b = 0; b < FN; ++b)
      xb[b] = *reinterpret_cast<const Frag*>(&S[ss * G::BBLK + b_off + b * 32 * 8]);
  };
  // one slab's MFMAs, block q = a FN + b, in the order first, first + 1, ... (mod NQ)
  auto mm = [&](int first) {
#pragma unroll
    for (int t = 0; t < NQ; ++t) {
      const int q = (first + t) % NQ, a = q / FN, b = q % FN;
      acc[a][b] = PanelOps<I8>::mma(bf[b], af[a], acc[a][b]);
    }
  };
  // stage KS (J: its index in the chunk) from slot Sc; slot Sn holds stage KS + 1
  auto stage = [&](int64_t KS, uint16_t* Sc, uint16_t* Sn, auto jc) {
    constexpr int J = decltype(jc)::value;
    int first = 0;
    if constexpr (MOD) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        if (q * CHUNK_STAGES / NQ != J) continue;
        first = (q + 1) % NQ;
        const int a = q / FN, b = q % FN;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const float x = acc[a][b][r];
          acc[a][b][r] = __builtin_fmaf(-__builtin_rintf(x * invp), p, x);
        }
      }
    }
#pragma unroll
    for (int ss = 0; ss < SPS; ++ss) {
      const int f = ss == 0 ? first : 0;
      if (ss + 1 < SPS) {
        rd(Sc, ss + 1, an, bn);
        mm(f);
        if (MOD && ss == 0)
          panel_slab_order_valu<FM, FN, 7>();
        else
          panel_slab_order<FM, FN>();
      } else {
        __builtin_amdgcn_sched_barrier(0);
        split_wait_vmcnt<2 * G::LW>();  // this wavefront's part of stage KS + 1 landed
        split_barrier();                // everyone's has; everyone is done reading slot KS
        rd(Sn, 0, an, bn);              // past the last stage: unused
        __builtin_amdgcn_sched_barrier(0);
        // the slab's MFMAs with the DMA of stage KS + 4 (into slot KS, free
        // since the barrier) after the last LW of them, one load each: the
        // matrix cores have queued work while the loads are issued, instead
        // of every wavefront issuing its loads at once after the barrier
#pragma unroll
        for (int t = 0; t < NQ; ++t) {
          const int q = (f + t) % NQ, a = q / FN, b = q % FN;
          acc[a][b] = PanelOps<I8>::mma(bf[b], af[a], acc[a][b]);
          if (t >= NQ - G::LW) {
            __builtin_amdgcn_sched_barrier(0);
            issue_one(KS + 4, Sc, t - (NQ - G::LW));
            __builtin_amdgcn_sched_barrier(0);
          }
        }
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int a = 0; a < FM; ++a) af[a] = an[a];
#pragma unroll
      for (int b = 0; b < FN; ++b) bf[b] = bn[b];
    }
  };

  if (nst <= 0) return;
  // prologue: stages 0..3 in flight, wait for stage 0, its first fragments
  issue(0, S0);
  issue(1, S1);
  issue(2, S2);
  issue(3, S3);
  split_wait_vmcnt<3 * G::LW>();
  split_barrier();
  rd(S0, 0, af, bf);
  (void)p;
  (void)invp;
  for (int64_t ks = 0; ks < nst; ks += CHUNK_STAGES) {
#define MOMMS_PANEL_STAGE(J, Sc, Sn)                      \
  if constexpr (J < CHUNK_STAGES) {                       \
    if (ks + J < nst) stage(ks + J, Sc, Sn, std::integral_constant<int, J>()); \
  }
    MOMMS_PANEL_STAGE(0, S0, S1)
    MOMMS_PANEL_STAGE(1, S1, S2)
    MOMMS_PANEL_STAGE(2, S2, S3)
    MOMMS_PANEL_STAGE(3, S3, S0)
    MOMMS_PANEL_STAGE(4, S0, S1)
    MOMMS_PANEL_STAGE(5, S1, S2)
    MOMMS_PANEL_STAGE(6, S2, S3)
    MOMMS_PANEL_STAGE(7, S3, S0)
    MOMMS_PANEL_STAGE(8, S0, S1)
    MOMMS_PANEL_STAGE(9, S1, S2)
    MOMMS_PANEL_STAGE(10, S2, S3)
    MOMMS_PANEL_STAGE(11, S3, S0)
    MOMMS_PANEL_STAGE(12, S0, S1)
    MOMMS_PANEL_STAGE(13, S1, S2)
    MOMMS_PANEL_STAGE(14, S2, S3)
    MOMMS_PANEL_STAGE(15, S3, S0)
#undef MOMMS_PANEL_STAGE
  }
  split_wait_vmcnt<0>();  // the ring's last DMAs (reloads) have landed
}

// Packs count vectors of length len (vector v, element p at X[v*sv + p*sp],
// bf16 bit patterns) into the panel layout (vpad / PANEL panels of
// panel_stride(lpad / 16) slabs), padding with zeros to vpad vectors and
// lpad k-values (the unused slab between panels is not written): one 64 x 64 block of (v, p) per 256 threads,
// through LDS so the global reads (along whichever index is unit-stride) and
// the 16-byte panel writes are both coalesced.
template <int PANEL>
__global__ void __launch_bounds__(256)
pack_bf16_panels(const uint16_t* __restrict__ X, int64_t count, int64_t len, int64_t sv,
                 int64_t sp, int64_t vpad, int64_t lpad, uint16_t* __restrict__ Q) {
  __shared__ uint16_t tile[64][72];
  const int tid = threadIdx.x;
  const int64_t v0 = (int64_t)blockIdx.x * 64, p0 = (int64_t)blockIdx.y * 64;
  const bool v_unit = (sv == 1);
  if (count > 0 && len > 0) {  // (the same for the whole grid: X has an element 0)
    // all 16 loads first (X[0] for elements outside X, masked after), then the
    // LDS stores, so the loads of a thread are in flight together
    uint16_t x[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int e = tid + 256 * r;
      const int a = e & 63, b = e >> 6;  // a runs fastest
      const int64_t gv = v0 + (v_unit ? a : b), gp = p0 + (v_unit ? b : a);
      x[r] = X[(gv < count && gp < len) ? gv * sv + gp * sp : 0];
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int e = tid + 256 * r;
      const int a = e & 63, b = e >> 6;
      const int vv = v_unit ? a : b, pp = v_unit ? b : a;
      tile[vv][pp] = (v0 + vv < count && p0 + pp < len) ? x[r] : (uint16_t)0;
    }
  } else {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int e = tid + 256 * r;
      const int a = e & 63, b = e >> 6;
      tile[v_unit ? a : b][v_unit ? b : a] = (uint16_t)0;
    }
  }
  __syncthreads();
  const int64_t nslabs = lpad / kSlab;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int c = tid + 256 * r;
    const int vv = c & 63;
    const int pc = c >> 6;  // chunk of 8 along p
    const int64_t gv = v0 + vv;
    const int64_t gp = p0 + pc * 8;
    if (gv >= vpad || gp >= lpad) continue;
    uint4 w;
    w.x = (uint32_t)tile[vv][pc * 8 + 0] | ((uint32_t)tile[vv][pc * 8 + 1] << 16);
    w.y = (uint32_t)tile[vv][pc * 8 + 2] | ((uint32_t)tile[vv][pc * 8 + 3] << 16);
    w.z = (uint32_t)tile[vv][pc * 8 + 4] | ((uint32_t)tile[vv][pc * 8 + 5] << 16);
    w.w = (uint32_t)tile[vv][pc * 8 + 6] | ((uint32_t)tile[vv][pc * 8 + 7] << 16);
    const int64_t pv = gv / PANEL, ks = gp / kSlab;
    const int kh = (int)((gp % kSlab) / 8);
    *reinterpret_cast<uint4*>(Q + (pv * panel_stride(nslabs) + ks) * (PANEL * kSlab) +
                              (int64_t)kh * PANEL * 8 + (gv % PANEL) * 8) = w;
  }
}

}  // namespace momms
