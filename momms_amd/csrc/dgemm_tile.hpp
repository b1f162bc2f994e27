// The FP64 matrix-core tile: one BM x BN tile of C := beta*C + alpha*A*B,
// used by the one-problem kernel (dgemm.hip) and the persistent batched
// kernel (batched.hip). See dgemm.hip for how it maps the Goto tree.
#pragma once
#include "c_update.hpp"
#include "mfma.hpp"

namespace momms {

// One BM x BN tile of C at (row0, col0): the whole k loop and the epilogue.
// Shared by the one-problem kernel below and the persistent batched kernel.
// Ends after a workgroup barrier, so the LDS slabs may be reused at once.
template <int BM, int BN, int BK, int WAVES_M, int WAVES_N, int LA, int LB>
__device__ __forceinline__ void dgemm_tile(int64_t m, int64_t n, int64_t k, double alpha,
                                           const double* __restrict__ A, int64_t rsA,
                                           int64_t csA, const double* __restrict__ B,
                                           int64_t rsB, int64_t csB, double beta,
                                           double* __restrict__ C, int64_t rsC, int64_t csC,
                                           int64_t row0, int64_t col0) {
  constexpr int NT = 64 * WAVES_M * WAVES_N;
  constexpr int WTM = BM / WAVES_M;  // wave tile rows
  constexpr int WTN = BN / WAVES_N;  // wave tile cols
  constexpr int FM = WTM / 16;       // 16x16 blocks per wave, M direction
  constexpr int FN = WTN / 16;
  constexpr int PAD = 16;            // row pitch = 128 B mod 256 B: conflict-free reads
  constexpr int LDA = BM + PAD;
  constexpr int LDB = BN + PAD;
  constexpr int A_PER = BM * BK / NT;
  constexpr int B_PER = BK * BN / NT;
  static_assert(BM * BK % NT == 0 && BK * BN % NT == 0, "slab must split evenly");
  static_assert(WTM % 16 == 0 && WTN % 16 == 0 && BK % 4 == 0, "tile shape");

  __shared__ double As[2][BK][LDA];
  __shared__ double Bs[2][BK][LDB];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave % WAVES_M;
  const int wn = wave / WAVES_M;

  const int64_t ktiles = ceil_div(k, BK);
  const bool m_full = row0 + BM <= m;
  const bool n_full = col0 + BN <= n;

  double ra[A_PER];
  double rb[B_PER];

  // Global -> registers for slab kt. Element (i, kk) of the A slab and
  // (kk, j) of the B slab; lanes run along the unit-stride dimension.
  auto load_slab = [&](int64_t kt) {
    const int64_t k0 = kt * BK;
    const bool k_full = k0 + BK <= k;
#pragma unroll
    for (int p = 0; p < A_PER; ++p) {
      const int e = tid + NT * p;
      const int i = (LA == kRow) ? e / BK : e % BM;
      const int kk = (LA == kRow) ? e % BK : e / BM;
      const int64_t gi = row0 + i, gk = k0 + kk;
      const int64_t idx = (LA == kCol) ? gi + gk * csA : (LA == kRow) ? gi * rsA + gk
                                                                     : gi * rsA + gk * csA;
      if (m_full && k_full) {
        ra[p] = A[idx];
      } else {
        // out-of-range elements load A[0] (always valid) and are replaced by 0
        const bool ok = gi < m && gk < k;
        const double v = A[ok ? idx : 0];
        ra[p] = ok ? v : 0.0;
      }
    }
#pragma unroll
    for (int p = 0; p < B_PER; ++p) {
      const int e = tid + NT * p;
      const int j = (LB == kRow) ? e % BN : e / BK;
      const int kk = (LB == kRow) ? e / BN : e % BK;
      const int64_t gj = col0 + j, gk = k0 + kk;
      const int64_t idx = (LB == kCol) ? gk + gj * csB : (LB == kRow) ? gk * rsB + gj
                                                                     : gk * rsB + gj * csB;
      if (n_full && k_full) {
        rb[p] = B[idx];
      } else {
        const bool ok = gj < n && gk < k;
        const double v = B[ok ? idx : 0];
        rb[p] = ok ? v : 0.0;
      }
    }
  };
  auto store_slab = [&](int buf) {
#pragma unroll
    for (int p = 0; p < A_PER; ++p) {
      const int e = tid + NT * p;
      const int i = (LA == kRow) ? e / BK : e % BM;
      const int kk = (LA == kRow) ? e % BK : e / BM;
      As[buf][kk][i] = ra[p];
    }
#pragma unroll
    for (int p = 0; p < B_PER; ++p) {
      const int e = tid + NT * p;
      const int j = (LB == kRow) ? e % BN : e / BK;
      const int kk = (LB == kRow) ? e / BN : e % BK;
      Bs[buf][kk][j] = rb[p];
    }
  };

  v4f64 acc[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b) acc[a][b] = v4f64{0.0, 0.0, 0.0, 0.0};

  const int lr = lane & 15;   // row (A) / column (B) inside a 16-block
  const int lk = lane >> 4;   // k inside a 4-deep step
  const int arow = wm * WTM + lr;
  const int bcol = wn * WTN + lr;

  if (ktiles > 0) {
    load_slab(0);
    store_slab(0);
  }
  __syncthreads();

  for (int64_t kt = 0; kt < ktiles; ++kt) {
    const int cur = (int)(kt & 1);
    if (kt + 1 < ktiles) load_slab(kt + 1);
#pragma unroll
    for (int k4 = 0; k4 < BK / 4; ++k4) {
      double af[FM], bf[FN];
#pragma unroll
      for (int a = 0; a < FM; ++a) af[a] = As[cur][k4 * 4 + lk][arow + a * 16];
#pragma unroll
      for (int b = 0; b < FN; ++b) bf[b] = Bs[cur][k4 * 4 + lk][bcol + b * 16];
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) acc[a][b] = mfma_f64_16x16x4(bf[b], af[a], acc[a][b]);
    }
    if (kt + 1 < ktiles) store_slab(cur ^ 1);
    __syncthreads();
  }

  // Epilogue. In the C^T block layout lane l holds C[i][j] with i = l%16 and
  // j = l/16 + 4r (r < 4) of each 16x16 block.
#pragma unroll
  for (int a = 0; a < FM; ++a) {
    const int64_t gi = row0 + wm * WTM + a * 16 + lr;
    int64_t idx[FN * 4];
    bool ok[FN * 4];
    double v[FN * 4];
#pragma unroll
    for (int b = 0; b < FN; ++b) {
      const int64_t gj0 = col0 + wn * WTN + b * 16 + lk;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gj = gj0 + 4 * r;
        idx[b * 4 + r] = gi * rsC + gj * csC;
        ok[b * 4 + r] = gi < m && gj < n;
        v[b * 4 + r] = alpha * acc[a][b][r];
      }
    }
    update_c(C, idx, ok, v, beta);
  }
}

// 64 x 64-style tile of C := beta*C + alpha*A*B for column-major A, B and C
// (the batched kernel's case), with D slabs of global loads in flight: D
// register sets of the staged slab, the k loop unrolled by D, each set
// reloaded D slabs ahead right after it went to LDS. Each thread keeps base
// pointers to the row of A and the columns of B it stages. Loads are issued
// unconditionally: an element outside the problem (row >= m, column >= n or
// k-index >= k) is read from the problem's element (0, 0) instead, which
// exists whenever k > 0, and is masked when stored to LDS. So the compiler's
// wait before a set is stored covers that set only. For short tiles of
// small problems the round trip per 16 k, not the MFMA, is the cost; D = 3
// hides most of it. Each element sums its k-values in order, four per MFMA
// step.
template <int BM, int BN, int D>
__device__ __forceinline__ void dgemm_tile_deep(int64_t m, int64_t n, int64_t k, double alpha,
                                                const double* __restrict__ A, int64_t lda,
                                                const double* __restrict__ B, int64_t ldb,
                                                double beta, double* __restrict__ C, int64_t ldc,
                                                int64_t row0, int64_t col0) {
  constexpr int BK = 16, NT = 256, WTM = BM / 2, WTN = BN / 2, FM = WTM / 16, FN = WTN / 16;
  // A slab k-rows of BM doubles (lanes store and read along rows); B slab
  // n-rows of BK doubles, k contiguous as in memory: a row stride of BK + 2
  // doubles puts the 16 rows a fragment read touches (and the 16 k-values a
  // store writes) on distinct bank pairs
  constexpr int PAD = 16, LDA = BM + PAD, LDBK = BK + 2;
  constexpr int A_PER = BM * BK / NT, B_PER = BK * BN / NT;
  constexpr int AK = NT / BM, BJ = NT / BK;
  static_assert(NT % BM == 0 && A_PER * AK == BK && B_PER * BJ == BN, "slab split");
  __shared__ double As[2][BK][LDA];
  __shared__ double Bs[2][BN][LDBK];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % 2, wn = wave / 2;
  const int ai = tid % BM, ak = tid / BM;
  const int bk = tid % BK, bj = tid / BK;
  const bool arow_ok = row0 + ai < m;
  const double* pa = A + (arow_ok ? row0 + ai + (int64_t)ak * lda : 0);
  const double* pb[B_PER];
  bool bcol_ok[B_PER];
#pragma unroll
  for (int q = 0; q < B_PER; ++q) {
    const int64_t j = col0 + bj + BJ * q;
    bcol_ok[q] = j < n;
    pb[q] = B + (bcol_ok[q] ? bk + j * ldb : 0);
  }
  const int64_t ktiles = ceil_div(k, BK);
  double ra[D][A_PER], rb[D][B_PER];
  // loads of slab kt (any kt >= 0; pa and pb[q] are dereferenced only at
  // elements inside the problem, the others read A[0] / B[0])
  auto load_slab = [&](int64_t kt, double (&xa)[A_PER], double (&xb)[B_PER]) {
    const int64_t k0 = kt * BK;
#pragma unroll
    for (int q = 0; q < A_PER; ++q) {
      const bool ok = arow_ok && k0 + ak + AK * q < k;
      xa[q] = *(ok ? pa + (k0 * lda + (int64_t)AK * q * lda) : A);
    }
    const bool kok = k0 + bk < k;
#pragma unroll
    for (int q = 0; q < B_PER; ++q) xb[q] = *((bcol_ok[q] && kok) ? pb[q] + k0 : B);
  };
  auto store_slab = [&](int64_t kt, int buf, const double (&xa)[A_PER], const double (&xb)[B_PER]) {
    const int64_t k0 = kt * BK;
#pragma unroll
    for (int q = 0; q < A_PER; ++q)
      As[buf][ak + AK * q][ai] = (arow_ok && k0 + ak + AK * q < k) ? xa[q] : 0.0;
    const bool kok = k0 + bk < k;
#pragma unroll
    for (int q = 0; q < B_PER; ++q) Bs[buf][bj + BJ * q][bk] = (bcol_ok[q] && kok) ? xb[q] : 0.0;
  };

  v4f64 acc[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b) acc[a][b] = v4f64{0.0, 0.0, 0.0, 0.0};
  const int lr = lane & 15, lk = lane >> 4;
  const int arow = wm * WTM + lr, bcol = wn * WTN + lr;

  // one slab: to LDS, barrier, (reload the set D slabs ahead), multiply
  auto step = [&](int64_t kt, double (&xa)[A_PER], double (&xb)[B_PER], bool reload) {
    const int buf = (int)(kt & 1);
    store_slab(kt, buf, xa, xb);
    __syncthreads();  // slab kt visible; everyone is done with slab kt-1
    if (reload) load_slab(kt + D, xa, xb);
    // keep the multiplies below the reload, so the wait before the next set
    // is stored leaves this set's loads in flight
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k4 = 0; k4 < BK / 4; ++k4) {
      double af[FM], bf[FN];
#pragma unroll
      for (int a = 0; a < FM; ++a) af[a] = As[buf][k4 * 4 + lk][arow + a * 16];
#pragma unroll
      for (int b = 0; b < FN; ++b) bf[b] = Bs[buf][bcol + b * 16][k4 * 4 + lk];
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) acc[a][b] = mfma_f64_16x16x4(bf[b], af[a], acc[a][b]);
    }
  };
  if (ktiles > 0) {  // (k = 0: no loads; A and B may hold no element)
#pragma unroll
    for (int d = 0; d < D; ++d) load_slab(d, ra[d], rb[d]);
    // whole groups of D slabs with no branch inside, so the compiler's
    // accounting of loads in flight stays exact across the loop
    int64_t kt = 0;
#pragma unroll 1
    for (; kt + D <= ktiles; kt += D) {
#pragma unroll
      for (int d = 0; d < D; ++d) step(kt + d, ra[d], rb[d], true);
    }
#pragma unroll
    for (int d = 0; d + 1 < D; ++d)
      if (kt + d < ktiles) step(kt + d, ra[d], rb[d], false);
  }
  __syncthreads();  // LDS free for the next tile of a persistent caller
  int64_t idx[FM * FN * 4];
  bool ok[FM * FN * 4];
  double v[FM * FN * 4];
#pragma unroll
  for (int a = 0; a < FM; ++a) {
    const int64_t gi = row0 + wm * WTM + a * 16 + lr;
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int t = (a * FN + b) * 4 + r;
        const int64_t gj = col0 + wn * WTN + b * 16 + lk + 4 * r;
        idx[t] = gi + gj * ldc;
        ok[t] = gi < m && gj < n;
        v[t] = alpha * acc[a][b][r];
      }
  }
  update_c(C, idx, ok, v, beta);
}

}  // namespace momms

namespace momms {

// ---------------------------------------------------------------------------
// Column-major fast path: slabs move HBM -> LDS by buffer_load ... lds (no
// register staging) through a ring of STAGES slabs, so STAGES-1 slabs are in
// flight while one is multiplied. Each wavefront load moves 1 KiB: lane l
// writes LDS bytes [16l, 16l+16) of the destination chosen by the wave.
//   A slab [BK][LDA]: one load = one k-column of the tile (BM = 128 rows).
//   B slab, n-major: 8 columns x 16 k per 1 KiB block; column c of a block
//   keeps its k-pair kp at 16-byte slot c*8 + (kp ^ c), so the MFMA reads
//   (16 columns at one k) fall in distinct banks.
// Out-of-range elements: the buffer resource covers exactly the matrix
// extent (reads past it return zero, never fault); rows >= m of A and
// columns >= n of B may read neighbouring values, which only reach outputs
// that are not stored; the k >= K entries of the last slab are zeroed in LDS
// after they land.
// Per element, the k-values of a slab are summed in the order 0, 2, 4, 6, 1,
// 3, 5, 7, 8, 10, ... (MFMA steps 2u + t cover 8u + 2lk + t over the four
// lane groups lk); slabs in order.

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, bytes, 0x00020000);
}

__device__ __forceinline__ void dma16(__amdgpu_buffer_rsrc_t r, void* lds, uint32_t voff) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)lds, 16,
                                           voff, 0, 0, 0);
}

// s_waitcnt with only the vector-memory counter constrained (gfx9 encoding:
// vmcnt in bits 3:0 and 15:14, expcnt 6:4 and lgkmcnt 11:8 left at maximum).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

// Workgroup barrier without a memory fence: the caller has already waited for
// what the other wavefronts must see (vmcnt for LDS DMA, lgkmcnt for LDS
// stores). A fence here would make the compiler drain every outstanding
// buffer->LDS load, which is exactly what the ring is meant to keep in flight.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

template <int STAGES>
struct DgemmDmaSmem {
  static constexpr int BM = 128, BN = 128, BK = 16, LDA = BM + 16;
  double A[STAGES][BK][LDA];
  __attribute__((aligned(16))) double B[STAGES][BN * BK];
};

template <int STAGES>
__device__ __forceinline__ void dgemm_tile_dma(DgemmDmaSmem<STAGES>& sm, int64_t m, int64_t n,
                                               int64_t k, double alpha,
                                               __amdgpu_buffer_rsrc_t ra, int64_t lda,
                                               __amdgpu_buffer_rsrc_t rb, int64_t ldb,
                                               double beta, double* __restrict__ C,
                                               int64_t ldc, int64_t row0, int64_t col0) {
  constexpr int BM = 128, BN = 128, BK = 16, WTM = 64, WTN = 64, FM = 4, FN = 4;
  constexpr int LOADS = (BK + BN / 8) / 4;  // wavefront loads per slab per wave (8)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int64_t ktiles = ceil_div(k, BK);

  // this wave's loads of slab kt into stage st: A rows kk = wave*4 .. +3,
  // B blocks b = wave*4 .. +3
  const int bc = lane >> 3, bkp = (lane & 7) ^ (lane >> 3);  // column in block, k-pair fetched
  auto issue = [&](int64_t kt, int st) {
    const int64_t k0 = kt * BK;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int kk = wave * 4 + q;
      const int64_t e = (row0 + 2 * lane) + (k0 + kk) * lda;  // element offset in A
      dma16(ra, &sm.A[st][kk][0], (uint32_t)(e * 8));
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int b = wave * 4 + q;
      const int64_t e = (col0 + 8 * b + bc) * ldb + k0 + 2 * bkp;
      dma16(rb, &sm.B[st][b * 128], (uint32_t)(e * 8));
    }
  };

  v4f64 acc[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b) acc[a][b] = v4f64{0.0, 0.0, 0.0, 0.0};

  const int lr = lane & 15, lk = lane >> 4;
  for (int s = 0; s < STAGES - 1; ++s)
    if (s < ktiles) issue(s, s);

  for (int64_t kt = 0; kt < ktiles; ++kt) {
    const int64_t left = ktiles - 1 - kt;  // slabs issued after this one
    if (left >= STAGES - 2) wait_vmcnt<(STAGES - 2) * LOADS>();
    else if (left == 1) wait_vmcnt<LOADS>();
    else wait_vmcnt<0>();
    lds_barrier();  // slab kt visible; every wave is done with slab kt-1
    const int cur = (int)(kt % STAGES);
    if (kt * BK + BK > k) {
      // last, partial slab: clear the k >= K entries of both slabs (their
      // buffer offsets may have wrapped into the matrix), then sync again
      const int kv = (int)(k - kt * BK);
      for (int e = tid; e < BN * BK; e += 256) {
        const int slot = e >> 1, c = (slot >> 3) & 7, kp = (slot & 7) ^ c;
        if (2 * kp + (e & 1) >= kv) sm.B[cur][e] = 0.0;
      }
      for (int e = tid; e < (BK - kv) * BM; e += 256) sm.A[cur][kv + e / BM][e % BM] = 0.0;
      lds_barrier();
    }
    // k order inside the slab: MFMA step 2u + t gives lane group lk the
    // k-value 8u + 2lk + t, so a lane's B values for two steps are its
    // column's k-pair 4u + lk, one 16-byte read (the 8 columns of a group of
    // 8 lanes sit in distinct banks). The loads of slab kt + STAGES - 1 (into
    // the slot every wave finished with before the barrier) go out after the
    // first half of the MFMAs, so the matrix cores hold work meanwhile.
#pragma unroll
    for (int u = 0; u < BK / 8; ++u) {
      if (u == 1 && kt + STAGES - 1 < ktiles)
        issue(kt + STAGES - 1, (int)((kt + STAGES - 1) % STAGES));
      const int kp = 4 * u + lk;
      v2f64 bp[FN];
#pragma unroll
      for (int b = 0; b < FN; ++b) {
        const int j = wn * WTN + b * 16 + lr, c = j & 7;
        bp[b] = *reinterpret_cast<const v2f64*>(&sm.B[cur][(j >> 3) * 128 + (c * 8 + (kp ^ c)) * 2]);
      }
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int kk = 2 * kp + t;
        double af[FM];
#pragma unroll
        for (int a = 0; a < FM; ++a) af[a] = sm.A[cur][kk][wm * WTM + a * 16 + lr];
#pragma unroll
        for (int a = 0; a < FM; ++a)
#pragma unroll
          for (int b = 0; b < FN; ++b) acc[a][b] = mfma_f64_16x16x4(bp[b][t], af[a], acc[a][b]);
      }
    }
  }
  lds_barrier();  // LDS free for the next tile of a persistent caller

#pragma unroll
  for (int a = 0; a < FM; ++a) {
    const int64_t gi = row0 + wm * WTM + a * 16 + lr;
    int64_t idx[FN * 4];
    bool ok[FN * 4];
    double v[FN * 4];
#pragma unroll
    for (int b = 0; b < FN; ++b) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gj = col0 + wn * WTN + b * 16 + lk + 4 * r;
        idx[b * 4 + r] = gi + gj * ldc;
        ok[b * 4 + r] = gi < m && gj < n;
        v[b * 4 + r] = alpha * acc[a][b][r];
      }
    }
    update_c(C, idx, ok, v, beta);
  }
}

}  // namespace momms
