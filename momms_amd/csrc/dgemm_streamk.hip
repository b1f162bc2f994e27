// Persistent stream-K form of the FP64 matrix-core GEMM, for shapes whose
// output tiles alone cannot keep every CU busy (fewer tiles than workgroup
// slots, or a count that leaves the last wave of tiles half empty).
//
// This is the device form of the reference's thread team: SpawnThreads
// starts one persistent worker per core (src/composables/spawn.rs:100-135),
// ParallelN/ParallelM hand each worker an even, contiguous share of the loop
// range (src/composables/parallel_range.rs:98-163), and ThreadComm's barrier
// and broadcast combine their results (src/thread_comm.rs:50-107). Here the
// grid holds at most as many workgroups as the device keeps resident; the
// range shared out is the flattened (tile, 16-deep k-slab) iteration space,
// and workgroup g takes iterations [floor(g I / G), floor((g+1) I / G)). A
// tile whose k-slabs all fall to one workgroup is written to C at once. A
// tile split between workgroups leaves one part per workgroup in its own
// workspace slot; a second, short kernel (the kernel boundary is the
// barrier) has the workgroup that holds the tile's first k-slabs add the
// parts in the fixed order of their k-ranges, apply alpha and beta and write
// C. The sum therefore does not depend on timing, and no workgroup waits for
// another. (Combining inside the first kernel, through device-scope fences
// and an arrival counter per tile, measured slower by two orders of
// magnitude at 512^3: the fences serialize across the device.)
//
// Tiles (4 wavefronts of (BM/2) x (BN/2), v_mfma_f64_16x16x4_f64): 64 x 64
// and 128 x 64, the slabs moved HBM -> LDS by buffer loads to LDS
// through a ring of STAGES slabs (STAGES - 1 in flight), as the 128 x 128
// data-parallel kernel does. Column-major A, B, C only (the caller checks).
#include "dgemm.hpp"
#include "dgemm_streamk.hpp"
#include "dgemm_tile.hpp"
#include "mfma.hpp"

#include <mutex>

namespace momms {

template <int BM, int BN, int STAGES>
struct SkSmem {
  static constexpr int BK = 16;
  // A slab: k-rows of BM doubles. A 128-row k-row is one wavefront load
  // (1 KiB); two 64-row k-rows are one load, so they sit back to back.
  static constexpr int LDA = (BM == 128) ? BM + 16 : BM;
  double A[STAGES][BK][LDA];
  __attribute__((aligned(16))) double B[STAGES][BN * BK];
};

// acc += the product of slabs [kt0, kt1) of the tile at (row0, col0).
template <int BM, int BN, int STAGES>
__device__ __forceinline__ void sk_tile_range(SkSmem<BM, BN, STAGES>& sm, int64_t k,
                                              __amdgpu_buffer_rsrc_t ra, int64_t lda,
                                              __amdgpu_buffer_rsrc_t rb, int64_t ldb,
                                              int64_t row0, int64_t col0, int64_t kt0,
                                              int64_t kt1, v4f64 (&acc)[BM / 32][BN / 32]) {
  constexpr int BK = 16, WTM = BM / 2, WTN = BN / 2, FM = WTM / 16, FN = WTN / 16;
  constexpr int AL = (BM == 128) ? 4 : 2;  // A loads per wave per slab
  constexpr int BL = BN / 32;              // B loads (8-column blocks) per wave per slab
  constexpr int LOADS = AL + BL;
  static_assert(BM == 64 || BM == 128, "BM");
  static_assert(BN == 64 || BN == 128, "BN");
  static_assert(STAGES >= 2 && STAGES <= 6, "STAGES");
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int64_t nk = kt1 - kt0;

  const int bc = lane >> 3, bkp = (lane & 7) ^ (lane >> 3);  // column in block, k-pair fetched
  auto issue = [&](int64_t kt, int st) {
    const int64_t k0 = kt * BK;
    if constexpr (BM == 128) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int kk = wave * 4 + q;
        const int64_t e = (row0 + 2 * lane) + (k0 + kk) * lda;
        dma16(ra, &sm.A[st][kk][0], (uint32_t)(e * 8));
      }
    } else {
      // lanes 0-31: k-row kk, lanes 32-63: k-row kk + 1, 64 rows each
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int kk = wave * 4 + 2 * q;
        const int64_t e = (row0 + 2 * (lane & 31)) + (k0 + kk + (lane >> 5)) * lda;
        dma16(ra, &sm.A[st][kk][0], (uint32_t)(e * 8));
      }
    }
#pragma unroll
    for (int q = 0; q < BL; ++q) {
      const int b = wave * BL + q;
      const int64_t e = (col0 + 8 * b + bc) * ldb + k0 + 2 * bkp;
      dma16(rb, &sm.B[st][b * 128], (uint32_t)(e * 8));
    }
  };

  const int lr = lane & 15, lk = lane >> 4;
  // earlier stores (the previous tile's C or part) leave the count exact
  wait_vmcnt<0>();
#pragma unroll
  for (int s = 0; s < STAGES - 1; ++s)
    if (s < nk) issue(kt0 + s, s);

  for (int64_t j = 0; j < nk; ++j) {
    const int64_t kt = kt0 + j;
    const int64_t left = nk - 1 - j;  // slabs of this range issued after this one
    if constexpr (STAGES >= 5) {
      if (left >= STAGES - 2) wait_vmcnt<(STAGES - 2) * LOADS>();
      else if (left == 3) wait_vmcnt<3 * LOADS>();
      else if (left == 2) wait_vmcnt<2 * LOADS>();
      else if (left == 1) wait_vmcnt<LOADS>();
      else wait_vmcnt<0>();
    } else if constexpr (STAGES == 4) {
      if (left >= 2) wait_vmcnt<2 * LOADS>();
      else if (left == 1) wait_vmcnt<LOADS>();
      else wait_vmcnt<0>();
    } else if constexpr (STAGES == 3) {
      if (left >= 1) wait_vmcnt<LOADS>();
      else wait_vmcnt<0>();
    } else {
      wait_vmcnt<0>();
    }
    lds_barrier();  // slab visible; every wave is done with the previous one
    const int cur = (int)(j % STAGES);
    if (kt * BK + BK > k) {
      // the matrix's last, partial slab: clear its k >= K entries (their
      // buffer offsets may have wrapped into the matrix), then sync again
      const int kv = (int)(k - kt * BK);
      for (int e = tid; e < BN * BK; e += 256) {
        const int slot = e >> 1, c = (slot >> 3) & 7, kp = (slot & 7) ^ c;
        if (2 * kp + (e & 1) >= kv) sm.B[cur][e] = 0.0;
      }
      for (int e = tid; e < (BK - kv) * BM; e += 256) sm.A[cur][kv + e / BM][e % BM] = 0.0;
      lds_barrier();
    }
    // MFMA step 2u + t gives lane group lk the k-value 8u + 2lk + t; a lane's
    // B values for two steps are its column's k-pair 4u + lk, one 16-byte read.
    // The loads of slab kt + STAGES - 1 (into the slot every wave finished
    // with before the barrier) go out after the first half of the MFMAs, so
    // the matrix cores hold work while they are issued.
#pragma unroll
    for (int u = 0; u < BK / 8; ++u) {
      if (u == 1 && j + STAGES - 1 < nk) issue(kt + STAGES - 1, (int)((j + STAGES - 1) % STAGES));
      const int kp = 4 * u + lk;
      v2f64 bp[FN];
#pragma unroll
      for (int b = 0; b < FN; ++b) {
        const int jj = wn * WTN + b * 16 + lr, c = jj & 7;
        bp[b] = *reinterpret_cast<const v2f64*>(
            &sm.B[cur][(jj >> 3) * 128 + (c * 8 + (kp ^ c)) * 2]);
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
  lds_barrier();  // LDS free for the next range
}

// C := beta*C + alpha*v for the tile's outputs this lane holds.
template <int BM, int BN>
__device__ __forceinline__ void sk_store_c(int64_t m, int64_t n, double alpha, double beta,
                                           double* __restrict__ C, int64_t ldc, int64_t row0,
                                           int64_t col0, const v4f64 (&v)[BM / 32][BN / 32]) {
  constexpr int WTM = BM / 2, WTN = BN / 2, FM = WTM / 16, FN = WTN / 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1, lr = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int a = 0; a < FM; ++a) {
    const int64_t gi = row0 + wm * WTM + a * 16 + lr;
    int64_t idx[FN * 4];
    bool ok[FN * 4];
    double x[FN * 4];
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gj = col0 + wn * WTN + b * 16 + lk + 4 * r;
        idx[b * 4 + r] = gi + gj * ldc;
        ok[b * 4 + r] = gi < m && gj < n;
        x[b * 4 + r] = alpha * v[a][b][r];
      }
    update_c(C, idx, ok, x, beta);
  }
}

// First workgroup whose range holds iteration x: floor(((x + 1) G - 1) / I).
__device__ __forceinline__ int64_t sk_owner(int64_t x, int64_t G, int64_t I) {
  return ((x + 1) * G - 1) / I;
}

// Workspace slot of workgroup h's part of tile t: 2h for the first tile h
// touches, 2h + 1 for its last.
template <int BM, int BN>
__device__ __forceinline__ double* sk_slot(double* parts, int64_t h, int64_t t, int64_t G,
                                           int64_t I, int64_t S) {
  const int64_t h_first = (h * I / G) / S;
  return parts + (2 * h + (t == h_first ? 0 : 1)) * (int64_t)(BM * BN);
}

template <int BM, int BN>
__device__ __forceinline__ int64_t sk_frag_off(int wave, int a, int b, int lane) {
  return (((wave * (BM / 32) + a) * (BN / 32) + b) * 64 + lane) * 4;
}

template <int BM, int BN, int STAGES, int MINB>
__global__ void __launch_bounds__(256, MINB)
dgemm_streamk_kernel(int64_t m, int64_t n, int64_t k, double alpha, const double* __restrict__ A,
                     int64_t lda, uint32_t a_bytes, const double* __restrict__ B, int64_t ldb,
                     uint32_t b_bytes, double beta, double* __restrict__ C, int64_t ldc,
                     int64_t group_m, double* __restrict__ parts) {
  constexpr int FM = BM / 32, FN = BN / 32;
  __shared__ SkSmem<BM, BN, STAGES> sm;
  const int64_t G = gridDim.x;
  const int64_t g = xcd_remap(blockIdx.x, G);  // consecutive ranges on one XCD
  const int64_t tiles_m = ceil_div(m, BM), tiles_n = ceil_div(n, BN);
  const int64_t S = ceil_div(k, 16);
  const int64_t I = tiles_m * tiles_n * S;
  const int64_t it0 = g * I / G, it1 = (g + 1) * I / G;
  const __amdgpu_buffer_rsrc_t ra = make_rsrc(A, a_bytes), rb = make_rsrc(B, b_bytes);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;

  for (int64_t it = it0; it < it1;) {
    const int64_t t = it / S;
    const int64_t ks = it - t * S;
    const int64_t ke = (S - ks) < (it1 - it) ? S : ks + (it1 - it);
    int64_t tm, tn;
    grouped_tile(t, tiles_m, tiles_n, group_m, tm, tn);
    const int64_t row0 = tm * BM, col0 = tn * BN;
    v4f64 acc[FM][FN];
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int b = 0; b < FN; ++b) acc[a][b] = v4f64{0.0, 0.0, 0.0, 0.0};
    sk_tile_range<BM, BN, STAGES>(sm, k, ra, lda, rb, ldb, row0, col0, ks, ke, acc);
    if (ks == 0 && ke == S) {
      sk_store_c<BM, BN>(m, n, alpha, beta, C, ldc, row0, col0, acc);
    } else {
      double* mine = sk_slot<BM, BN>(parts, g, t, G, I, S);
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b)
          *reinterpret_cast<v4f64*>(mine + sk_frag_off<BM, BN>(wave, a, b, lane)) = acc[a][b];
    }
    it += ke - ks;
  }
}

// The split tiles: workgroup g handles the split tile whose first k-slab
// lies in its range of the main kernel (at most one: a split tile reaches
// past the end of that range), adding the parts of workgroups g, g+1, ...
// in that order.
template <int BM, int BN>
__global__ void __launch_bounds__(256)
dgemm_streamk_fixup(int64_t m, int64_t n, int64_t k, double alpha, double beta,
                    double* __restrict__ C, int64_t ldc, int64_t group_m, int64_t G,
                    const double* __restrict__ parts) {
  constexpr int FM = BM / 32, FN = BN / 32;
  const int64_t g = blockIdx.x;
  const int64_t tiles_m = ceil_div(m, BM), tiles_n = ceil_div(n, BN);
  const int64_t S = ceil_div(k, 16);
  const int64_t I = tiles_m * tiles_n * S;
  const int64_t it0 = g * I / G, it1 = (g + 1) * I / G;
  const int64_t t = (it1 - 1) / S;           // the last tile g touches
  if (t * S < it0 || (t + 1) * S <= it1) return;  // starts before g, or not split
  const int64_t gl = sk_owner((t + 1) * S - 1, G, I);
  int64_t tm, tn;
  grouped_tile(t, tiles_m, tiles_n, group_m, tm, tn);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double* const ps = const_cast<double*>(parts);
  // the parts in chunks of four, every fragment of a chunk loaded at once and
  // then added in k order: one round trip per four parts
  v4f64 acc[FM][FN];
  for (int64_t h0 = g; h0 <= gl; h0 += 4) {
    v4f64 x[4][FM][FN];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      // a part past the last re-reads the last (never added)
      const int64_t h = h0 + q <= gl ? h0 + q : gl;
      const double* p = sk_slot<BM, BN>(ps, h, t, G, I, S);
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b)
          x[q][a][b] = *reinterpret_cast<const v4f64*>(p + sk_frag_off<BM, BN>(wave, a, b, lane));
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (h0 + q > gl) break;
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b)
          acc[a][b] = (h0 + q == g) ? x[q][a][b] : acc[a][b] + x[q][a][b];
    }
  }
  sk_store_c<BM, BN>(m, n, alpha, beta, C, ldc, tm * BM, tn * BN, acc);
}

namespace {

struct SkKernel {
  const void* fn;
  const void* fixup;
  int bm, bn;
};

template <int BM, int BN, int STAGES, int MINB>
SkKernel sk_kernel() {
  return {reinterpret_cast<const void*>(&dgemm_streamk_kernel<BM, BN, STAGES, MINB>),
          reinterpret_cast<const void*>(&dgemm_streamk_fixup<BM, BN>), BM, BN};
}

// cfg 0: 64 x 64, three-slab ring, three workgroups per CU; 1: 128 x 64,
// three-slab ring, two per CU; 2: 64 x 64, four-slab ring, two per CU;
// 3: 64 x 64, six-slab ring, one per CU. (64 x 128 and 128 x 128 forms need
// more registers than two workgroups per CU leave and spill: their
// epilogues hold twice the outputs per row block.)
SkKernel sk_pick(int cfg) {
  switch (cfg) {
    case 0: return sk_kernel<64, 64, 3, 3>();
    case 1: return sk_kernel<128, 64, 3, 2>();
    case 2: return sk_kernel<64, 64, 4, 2>();
    default: return sk_kernel<64, 64, 6, 1>();
  }
}

std::mutex g_mu;
int g_slots[64][kStreamkConfigs];  // resident workgroups per device and tile form

int resident_workgroups(int cfg, hipError_t& err) {
  int dev = 0;
  err = hipGetDevice(&dev);
  if (err != hipSuccess) return 0;
  if (dev < 0 || dev >= 64) {
    err = hipErrorInvalidDevice;
    return 0;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_slots[dev][cfg] == 0) {
    int cus = 0, per = 0;
    if ((err = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) !=
        hipSuccess)
      return 0;
    if ((err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, sk_pick(cfg).fn, 256, 0)) !=
        hipSuccess)
      return 0;
    if (per < 1) per = 1;
    g_slots[dev][cfg] = cus * per;
  }
  return g_slots[dev][cfg];
}

}  // namespace

int64_t streamk_grid(const DgemmArgs& g, int cfg, int min_slabs, hipError_t& err) {
  const SkKernel kk = sk_pick(cfg);
  const int64_t iters = ceil_div(g.m, kk.bm) * ceil_div(g.n, kk.bn) * ceil_div(g.k, 16);
  int64_t G = resident_workgroups(cfg, err);
  if (err != hipSuccess) return 0;
  // at least min_slabs k-slabs per workgroup: fewer workgroups on small
  // problems, so that tiles split into fewer parts
  const int64_t cap = ceil_div(iters, min_slabs > 0 ? min_slabs : 1);
  if (cap < G) G = cap;
  if (G >= 8) G = G / 8 * 8;  // whole rounds over the 8 XCDs (xcd_remap is a bijection)
  return iters < G ? iters : (G < 1 ? 1 : G);
}

size_t streamk_workspace_bytes(int64_t grid, int cfg) {
  const SkKernel kk = sk_pick(cfg);
  return (size_t)grid * 2 * kk.bm * kk.bn * sizeof(double);
}

hipError_t streamk_launch(const DgemmArgs& g, int cfg, int64_t grid, uint32_t a_bytes,
                          uint32_t b_bytes, double* parts) {
  const SkKernel kk = sk_pick(cfg);
  int64_t group_m = 8;
  void* args[] = {(void*)&g.m,   (void*)&g.n,     (void*)&g.k,     (void*)&g.alpha,
                  (void*)&g.A,   (void*)&g.csA,   (void*)&a_bytes, (void*)&g.B,
                  (void*)&g.csB, (void*)&b_bytes, (void*)&g.beta,  (void*)&g.C,
                  (void*)&g.csC, (void*)&group_m, (void*)&parts};
  hipError_t e = hipLaunchKernel(kk.fn, dim3((unsigned)grid), dim3(256), args, 0, g.stream);
  if (e != hipSuccess) return e;
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // a split tile exists unless every workgroup's range is whole tiles
  const int64_t S = ceil_div(g.k, 16);
  const int64_t I = ceil_div(g.m, kk.bm) * ceil_div(g.n, kk.bn) * S;
  if (I % grid == 0 && (I / grid) % S == 0) return hipSuccess;
  const double* cparts = parts;
  void* fargs[] = {(void*)&g.m,    (void*)&g.n,   (void*)&g.k,     (void*)&g.alpha,
                   (void*)&g.beta, (void*)&g.C,   (void*)&g.csC,   (void*)&group_m,
                   (void*)&grid,   (void*)&cparts};
  e = hipLaunchKernel(kk.fixup, dim3((unsigned)grid), dim3(256), fargs, 0, g.stream);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

}  // namespace momms
