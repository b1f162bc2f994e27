// Device building blocks of the split-BF16 FP64-accurate GEMM (the algorithm
// is described in split_bf16.hip): packed-panel slicing of one 64 x 64 block
// and the fused multi-slice multiply of one tile. Used by the one-problem
// kernels (split_bf16.hip) and the persistent batched kernels (batched.hip).
#pragma once
#include <math.h>

#include "c_update.hpp"
#include "mfma.hpp"

namespace momms {

constexpr int kSlab = 16;     // k per slab (= one v_mfma_f32_32x32x16_bf16 step)
constexpr int kMaxSlices = 8;

// Exponent flag of a vector holding Inf or NaN (its values pass unsliced).
constexpr int kNonFinite = 0x7fffffff;

// ---------------------------------------------------------------------------
// 2. slicing into packed panels.
//
// Vector v (a row of A or a column of B), element p lives at
// X[v*sv + p*sp]. Output for slice t: panel pv = v / PANEL, slab ks = p / 16,
//   Q[((pv*nslabs + ks)*S + t)*(PANEL*16) + kh*PANEL*8 + (v%PANEL)*8 + p%8]
// so the S slice blocks of one slab are one contiguous S*PANEL*32-byte chunk
// (and nothing else in a tile's stream sits at a power-of-two stride).
// with kh = (p % 16) / 8. Padding (v >= count or p >= len) is written as 0.
// A 64 x 64 tile of (v, p) goes through LDS so that both the global read
// (along whichever index is unit-stride) and the packed write are coalesced.

template <int S, int PANEL>
__device__ __forceinline__ void slice_block(const double* __restrict__ X, int64_t count,
                                            int64_t len, int64_t sv, int64_t sp,
                                            const int* __restrict__ ex, int64_t vpad,
                                            int64_t lpad, uint16_t* __restrict__ Q, int64_t v0,
                                            int64_t p0) {
  __shared__ double tile[64][65];
  const int tid = threadIdx.x;
  const bool v_unit = (sv == 1);
    // load: 64 x 64 doubles, 16 per thread
  if (count > 0 && len > 0) {  // (the same for the whole grid: X has an element 0)
    // all 16 loads first (X[0] for elements outside X, masked after), then the
    // LDS stores, so the loads of a thread are in flight together
    double x[16];
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
      tile[vv][pp] = (v0 + vv < count && p0 + pp < len) ? x[r] : 0.0;
    }
  } else {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int e = tid + 256 * r;
      const int a = e & 63, b = e >> 6;
      tile[v_unit ? a : b][v_unit ? b : a] = 0.0;
    }
  }
  __syncthreads();
  const int64_t nslabs = lpad / 16;
  // write: each thread produces 8 consecutive p (one 16-byte chunk) of one v
  // per step; 64 v x 8 chunks = 512 chunks, 2 per thread.
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int c = tid + 256 * r;
    const int vv = c & 63;
    const int pc = c >> 6;  // chunk of 8 along p
    const int64_t gv = v0 + vv;
    const int64_t gp = p0 + pc * 8;
    if (gv >= vpad || gp >= lpad) continue;
    const int E = (gv < count) ? ex[gv] : 0;
    uint16_t q[S][8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const double x = tile[vv][pc * 8 + e];
      if (E == kNonFinite) {
        const uint32_t fb = __float_as_uint((float)x);
        q[0][e] = (uint16_t)(fb >> 16);
#pragma unroll
        for (int t = 1; t < S; ++t) q[t][e] = 0;
      } else {
        double rr = ldexp(x, -E);
#pragma unroll
        for (int t = 1; t <= S; ++t) {
          const double qt = rint(ldexp(rr, 8 * t));
          rr -= ldexp(qt, -8 * t);
          // qt is an integer with |qt| <= 256: exact in fp32 and bf16
          q[t - 1][e] = (uint16_t)(__float_as_uint((float)qt) >> 16);
        }
      }
    }
    const int64_t pv = gv / PANEL, ks = gp / 16;
    const int kh = (int)((gp % 16) / 8);
    const int64_t within = (int64_t)kh * PANEL * 8 + (gv % PANEL) * 8;
#pragma unroll
    for (int t = 0; t < S; ++t) {
      const int64_t base = ((pv * nslabs + ks) * S + t) * (PANEL * 16) + within;
      uint4 w;
      w.x = (uint32_t)q[t][0] | ((uint32_t)q[t][1] << 16);
      w.y = (uint32_t)q[t][2] | ((uint32_t)q[t][3] << 16);
      w.z = (uint32_t)q[t][4] | ((uint32_t)q[t][5] << 16);
      w.w = (uint32_t)q[t][6] | ((uint32_t)q[t][7] << 16);
      *reinterpret_cast<uint4*>(Q + base) = w;
    }
  }
}

// ---------------------------------------------------------------------------
// 3. fused multi-slice multiply.
//
// Workgroup: 8 wavefronts (4 x 2), tile 128 x 64 of C; each wavefront one
// 32 x 32 block. Per 16-deep slab every wavefront loads S A-fragments and S
// B-fragments (16 bytes per lane each) and issues the S(S+1)/2 products
// t+u <= S+1 into S fp32 accumulators, one per level d = t+u. The MFMA takes
// the B slice as its "A" operand, so lane l holds column j = 8(r/4)+4(l/32)+r%4
// and row i = l%32 of the block: column-major stores are 256-byte coalesced.

// One (32*WM) x (32*WN) tile (tm, tn) of one problem, from slices packed in
// panels of BM = 32*WM rows of A and BN = 32*WN columns of B. Slab blocks move
// HBM -> LDS by global_load_lds_dwordx4 (1 KiB per wavefront load, no
// register staging) through a ring of STAGES slabs, STAGES-1 in flight.
// Ends after a workgroup barrier (LDS reusable at once); lanes past the edge
// return early from the epilogue.
template <int S, int WM, int WN, int STAGES_>
struct SplitRing {
  static constexpr int BM = 32 * WM, BN = 32 * WN;
  static constexpr int ABLK = BM * kSlab, BBLK = BN * kSlab;  // bf16 per slice-slab block
  static constexpr int STAGES = STAGES_;
  static_assert(STAGES >= 2 && STAGES <= 4, "ring depth");
};

__device__ __forceinline__ void lds_dma16(const void* src, void* lds) {
  __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1)))*)src,
                                   (void __attribute__((address_space(3)))*)lds, 16, 0, 0);
}

// s_waitcnt on the vector-memory counter only (gfx9 encoding).
template <int N>
__device__ __forceinline__ void split_wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

// Workgroup barrier without a memory fence (a fence would drain the ring).
__device__ __forceinline__ void split_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

template <int S, int WM, int WN, int STAGES_ = 3>
__device__ __forceinline__ void split_tile(int64_t m, int64_t n, int64_t kpad, double alpha,
                                           double beta, double* __restrict__ C, int64_t rsC,
                                           int64_t csC, const uint16_t* __restrict__ Aq,
                                           const uint16_t* __restrict__ Bq,
                                           const int* __restrict__ EA,
                                           const int* __restrict__ EB, int64_t tiles_m,
                                           int64_t tiles_n, int slabs_per_chunk, int64_t tm,
                                           int64_t tn) {
  typedef SplitRing<S, WM, WN, STAGES_> R;
  constexpr int NW = WM * WN;
  constexpr int BM = R::BM, BN = R::BN, ABLK = R::ABLK, BBLK = R::BBLK, STAGES = R::STAGES;
  constexpr int LA = ABLK / 512, LB = BBLK / 512;  // 1 KiB wavefront loads per slice block
  constexpr int LOADS = S * (LA + LB);              // per slab, whole workgroup
  constexpr int LO = LOADS / NW;                     // per wave: LO or LO+1
  __shared__ __attribute__((aligned(16))) uint16_t As[STAGES][S * ABLK];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[STAGES][S * BBLK];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave % WM;
  const int wn = wave / WM;
  const int my_loads = LO + (wave < LOADS % NW ? 1 : 0);

  const int64_t nslabs = kpad / kSlab;
  // slab ks of this tile: S contiguous slice blocks at Abase + ks*S*ABLK
  const uint16_t* Abase = Aq + tm * nslabs * S * ABLK;
  const uint16_t* Bbase = Bq + tn * nslabs * S * BBLK;
  (void)tiles_m;
  (void)tiles_n;

  auto issue = [&](int64_t ks, int st) {
    for (int c = wave; c < LOADS; c += NW) {
      if (c < S * LA) {
        const int t = c / LA, q = c % LA;
        lds_dma16(Abase + (ks * S + t) * ABLK + q * 512 + lane * 8,
                  &As[st][t * ABLK + q * 512]);
      } else {
        const int cb = c - S * LA, t = cb / LB, q = cb % LB;
        lds_dma16(Bbase + (ks * S + t) * BBLK + q * 512 + lane * 8,
                  &Bs[st][t * BBLK + q * 512]);
      }
    }
  };
  v16f32 lev[S];  // level accumulators; the first slab of a chunk starts them from zero
  double acc[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0;

  // fragment offsets inside a slice-slab block: [kh][row][8]
  const int kh = lane >> 5;
  const int a_off = kh * BM * 8 + (wm * 32 + (lane & 31)) * 8;
  const int b_off = kh * BN * 8 + (wn * 32 + (lane & 31)) * 8;

  auto fold = [&]() {
#pragma unroll
    for (int d = S - 1; d >= 0; --d) {  // level d+2, smallest first
      const double sc = ldexp(1.0, -8 * (d + 2));
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        acc[r] = __builtin_fma((double)lev[d][r], sc, acc[r]);
      }
    }
  };

  for (int st = 0; st < STAGES - 1; ++st)
    if (st < nslabs) issue(st, st);
  int in_chunk = 0;
  for (int64_t ks = 0; ks < nslabs; ++ks) {
    // slab ks has landed once at most (slabs issued after it) x my_loads remain
    const int64_t later = nslabs - 1 - ks < STAGES - 2 ? nslabs - 1 - ks : STAGES - 2;
    if (later >= 2) {
      if (my_loads == LO) split_wait_vmcnt<2 * LO>();
      else split_wait_vmcnt<2 * LO + 2>();
    } else if (later == 1) {
      if (my_loads == LO) split_wait_vmcnt<LO>();
      else split_wait_vmcnt<LO + 1>();
    } else {
      split_wait_vmcnt<0>();
    }
    split_barrier();  // slab ks visible to all; everyone is done with slab ks-1
    if (ks + STAGES - 1 < nslabs) issue(ks + STAGES - 1, (int)((ks + STAGES - 1) % STAGES));
    const int cur = (int)(ks % STAGES);
    v8bf16 af[S], bf[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      af[s] = *reinterpret_cast<const v8bf16*>(&As[cur][s * ABLK + a_off]);
      bf[s] = *reinterpret_cast<const v8bf16*>(&Bs[cur][s * BBLK + b_off]);
    }
    // A slice ta+1 times B slice d-ta+1 adds to level d+2; the inner loop
    // walks the levels so consecutive MFMAs use different accumulators. On
    // the first slab of a chunk the A-slice-1 products (one per level) take a
    // zero accumulator operand instead of clearing the levels after a fold.
    if (in_chunk == 0) {
#pragma unroll
      for (int d = 0; d < S; ++d) lev[d] = mfma_bf16_32x32x16(bf[d], af[0], v16f32{});
    } else {
#pragma unroll
      for (int d = 0; d < S; ++d) lev[d] = mfma_bf16_32x32x16(bf[d], af[0], lev[d]);
    }
#pragma unroll
    for (int ta = 1; ta < S; ++ta)
#pragma unroll
      for (int d = ta; d < S; ++d) lev[d] = mfma_bf16_32x32x16(bf[d - ta], af[ta], lev[d]);
    if (++in_chunk == slabs_per_chunk) {
      fold();
      in_chunk = 0;
    }
  }
  split_barrier();  // LDS free for the next tile of a persistent caller
  if (in_chunk) fold();

  // epilogue
  const int64_t gi = tm * BM + wm * 32 + (lane & 31);
  if (gi >= m) return;
  const int ei = EA[gi];
  // EB is padded to whole tiles, so every column of the tile indexes it
  int eb[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) eb[r] = EB[tn * BN + wn * 32 + 8 * (r >> 2) + 4 * kh + (r & 3)];
  int64_t idx[16];
  bool ok[16];
  double v[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int64_t gj = tn * BN + wn * 32 + 8 * (r >> 2) + 4 * kh + (r & 3);
    v[r] = alpha * acc[r];
    if (ei != kNonFinite && eb[r] != kNonFinite) v[r] = ldexp(v[r], ei + eb[r]);
    idx[r] = gi * rsC + gj * csC;
    ok[r] = gj < n;
  }
  update_c(C, idx, ok, v, beta);
}

}  // namespace momms
