// FP64-accurate GEMM from BF16 slices (Ozaki scheme I) on BF16 matrix cores.
//
// No counterpart in the reference; it is the precision addition the engine
// offers beside the plain FP64 path. Exact arithmetic (restated on the CPU
// for bit-exact comparison):
//   row i of A:  E_i with max_p |a_ip| < 2^E_i;  r = a * 2^-E_i;
//                q_t = rint(r * 2^(8t)), r -= q_t * 2^(-8t)     (t = 1..s)
//   column j of B likewise (exponent F_j).  |q_1| <= 256, |q_t| <= 128:
//   every slice is an integer that bf16 holds exactly.
//   C_ij = beta*C0_ij + alpha * 2^(E_i+F_j) * sum_{d=2}^{s+1} 2^(-8d) S_d,
//   S_d = sum_{t+u=d} sum_p qa_t[i,p] qb_u[p,j], the level-d slice products.
// The k range is cut into chunks of kch(s) values (a multiple of 16). Within
// a chunk every S_d is an integer below 2^24 in magnitude, so the fp32
// accumulation of the matrix cores is exact in any order; after each chunk
// the levels are folded into an fp64 accumulator, d = s+1 down to 2, by
// acc = fma(S_d, 2^(-8d), acc).
//
// Data movement (the Goto tree on the device):
//   slicing writes A and B slices straight into packed panels, the device
//   form of PackA<RowPanelMatrix> / PackB<ColumnPanelMatrix>
//   (src/composables/pack.rs:51-133): for slice t, row-tile tm and 16-deep
//   k-slab ks, one contiguous block [kh][PANEL][8] of bf16, kh = k half.
//   The multiply kernel copies one block per slice per slab HBM -> LDS with
//   16-byte loads (double-buffered), and each wavefront reads 16 bytes per
//   lane conflict-free straight into the MFMA operand layout.
#include <math.h>

#include "mfma.hpp"
#include "split_bf16.hpp"
#include "status.hpp"

namespace momms {

constexpr int kPanelA = 128;  // rows of A per packed panel (= multiply tile BM)
constexpr int kPanelB = 64;   // columns of B per packed panel (= multiply tile BN)
constexpr int kSlab = 16;     // k per slab (= one v_mfma_f32_32x32x16_bf16 step)
constexpr int kMaxSlices = 8;

int split_chunk(int s) {
  int kch = 16 * (1024 / (16 * (s + 2)));
  return kch > 256 ? 256 : kch;
}

static inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

SplitLayout split_layout(int64_t m, int64_t n, int64_t k, int s) {
  SplitLayout L;
  L.s = s;
  L.mpad = ceil_div(m, kPanelA) * kPanelA;
  L.npad = ceil_div(n, kPanelB) * kPanelB;
  L.kpad = ceil_div(k > 0 ? k : 1, kSlab) * kSlab;
  size_t off = 0;
  L.off_aq = off;
  off = align256(off + (size_t)s * L.mpad * L.kpad * 2);
  L.off_bq = off;
  off = align256(off + (size_t)s * L.npad * L.kpad * 2);
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
// 1. max |x| per vector, as the bit pattern of a non-negative double (which
//    orders like the value; NaN patterns sort above +Inf).

__global__ void zero_u64(uint64_t* p, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
}

// One thread per vector (coalesced when vectors are adjacent, stride_vec == 1),
// a grid.y split of the length, and an atomic max.
__global__ void absmax_strided(const double* __restrict__ X, int64_t count, int64_t len,
                               int64_t stride_vec, int64_t stride_elem, int64_t per_block,
                               uint64_t* mx) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= count) return;
  const int64_t p0 = (int64_t)blockIdx.y * per_block;
  const int64_t p1 = p0 + per_block < len ? p0 + per_block : len;
  uint64_t best = 0;
  for (int64_t p = p0; p < p1; ++p) {
    const uint64_t b = (uint64_t)__double_as_longlong(X[v * stride_vec + p * stride_elem]) & 0x7fffffffffffffffull;
    best = b > best ? b : best;
  }
  if (p1 > p0) atomicMax((unsigned long long*)&mx[v], (unsigned long long)best);
}

// Elements of a vector are adjacent (stride_elem == 1): one wavefront per vector.
__global__ void absmax_contig(const double* __restrict__ X, int64_t count, int64_t len,
                              int64_t stride_vec, uint64_t* mx) {
  const int64_t v = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (v >= count) return;
  const double* x = X + v * stride_vec;
  uint64_t best = 0;
  for (int64_t p = lane; p < len; p += 64) {
    const uint64_t b = (uint64_t)__double_as_longlong(x[p]) & 0x7fffffffffffffffull;
    best = b > best ? b : best;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t o = __shfl_xor(best, off);
    best = o > best ? o : best;
  }
  if (lane == 0) mx[v] = best;
}

// Exponent E with max < 2^E (frexp convention); 0 for an all-zero vector.
// A vector holding Inf or NaN gets the flag value kNonFinite.
constexpr int kNonFinite = 0x7fffffff;
__global__ void exponents(const uint64_t* __restrict__ mx, int64_t count, int64_t padded,
                          int* __restrict__ e) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= padded) return;
  if (v >= count) {
    e[v] = 0;
    return;
  }
  const uint64_t b = mx[v];
  if (b >= 0x7ff0000000000000ull) {
    e[v] = kNonFinite;
  } else if (b == 0) {
    e[v] = 0;
  } else {
    int E;
    frexp(__longlong_as_double((long long)b), &E);
    e[v] = E;
  }
}

// ---------------------------------------------------------------------------
// 2. slicing into packed panels.
//
// Vector v (a row of A or a column of B), element p lives at
// X[v*sv + p*sp]. Output for slice t: panel pv = v / PANEL, slab ks = p / 16,
//   Q[((t*npanels + pv)*nslabs + ks)*(PANEL*16) + kh*PANEL*8 + (v%PANEL)*8 + p%8]
// with kh = (p % 16) / 8. Padding (v >= count or p >= len) is written as 0.
// A 64 x 64 tile of (v, p) goes through LDS so that both the global read
// (along whichever index is unit-stride) and the packed write are coalesced.

template <int S, int PANEL>
__global__ void __launch_bounds__(256)
slice_pack(const double* __restrict__ X, int64_t count, int64_t len, int64_t sv, int64_t sp,
           const int* __restrict__ ex, int64_t vpad, int64_t lpad, uint16_t* __restrict__ Q) {
  __shared__ double tile[64][65];
  const int tid = threadIdx.x;
  const int64_t v0 = (int64_t)blockIdx.x * 64;
  const int64_t p0 = (int64_t)blockIdx.y * 64;
  const bool v_unit = (sv == 1);
  // load: 64 x 64 doubles, 16 per thread
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int e = tid + 256 * r;
    const int a = e & 63, b = e >> 6;  // a runs fastest
    const int vv = v_unit ? a : b;
    const int pp = v_unit ? b : a;
    const int64_t gv = v0 + vv, gp = p0 + pp;
    double x = 0.0;
    if (gv < count && gp < len) x = X[gv * sv + gp * sp];
    tile[vv][pp] = x;
  }
  __syncthreads();
  const int64_t nslabs = lpad / 16;
  const int64_t npanels = vpad / PANEL;
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
      const int64_t base = (((int64_t)t * npanels + pv) * nslabs + ks) * (PANEL * 16) + within;
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

template <int S>
__global__ void __launch_bounds__(512)
split_multiply(int64_t m, int64_t n, int64_t kpad, double alpha, double beta,
               double* __restrict__ C, int64_t rsC, int64_t csC,
               const uint16_t* __restrict__ Aq, const uint16_t* __restrict__ Bq,
               const int* __restrict__ EA, const int* __restrict__ EB, int64_t mpad, int64_t npad,
               int slabs_per_chunk) {
  constexpr int BM = kPanelA, BN = kPanelB;
  constexpr int ABLK = BM * kSlab;  // bf16 per A slice-slab block (4 KiB)
  constexpr int BBLK = BN * kSlab;  // (2 KiB)
  constexpr int CH_A = S * ABLK / 8;  // 16-byte chunks per slab
  constexpr int CH_B = S * BBLK / 8;
  constexpr int CH = CH_A + CH_B;
  constexpr int PER = (CH + 511) / 512;
  __shared__ __attribute__((aligned(16))) uint16_t As[2][S * ABLK];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][S * BBLK];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave & 3;   // 4 waves along M
  const int wn = wave >> 2;  // 2 along N

  const int64_t tiles_m = mpad / BM;
  const int64_t tiles_n = npad / BN;
  const int64_t t = xcd_remap(blockIdx.x, (int64_t)gridDim.x);
  int64_t tm, tn;
  grouped_tile(t, tiles_m, tiles_n, 8, tm, tn);

  const int64_t nslabs = kpad / kSlab;
  // slab block of slice t: A panel tm, slab ks
  const uint16_t* Abase = Aq + tm * nslabs * ABLK;
  const uint16_t* Bbase = Bq + tn * nslabs * BBLK;
  const int64_t a_slice_stride = tiles_m * nslabs * ABLK;
  const int64_t b_slice_stride = tiles_n * nslabs * BBLK;

  uint4 stage[PER];
  auto load_slab = [&](int64_t ks) {
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      const int c = tid + 512 * r;
      if (c < CH_A) {
        const int sl = c / (ABLK / 8), w = c % (ABLK / 8);
        stage[r] = *reinterpret_cast<const uint4*>(Abase + sl * a_slice_stride + ks * ABLK + w * 8);
      } else if (c < CH) {
        const int cb = c - CH_A;
        const int sl = cb / (BBLK / 8), w = cb % (BBLK / 8);
        stage[r] = *reinterpret_cast<const uint4*>(Bbase + sl * b_slice_stride + ks * BBLK + w * 8);
      }
    }
  };
  auto store_slab = [&](int buf) {
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      const int c = tid + 512 * r;
      if (c < CH_A) *reinterpret_cast<uint4*>(&As[buf][c * 8]) = stage[r];
      else if (c < CH) *reinterpret_cast<uint4*>(&Bs[buf][(c - CH_A) * 8]) = stage[r];
    }
  };

  v16f32 lev[S];
#pragma unroll
  for (int d = 0; d < S; ++d)
#pragma unroll
    for (int r = 0; r < 16; ++r) lev[d][r] = 0.0f;
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
        lev[d][r] = 0.0f;
      }
    }
  };

  if (nslabs > 0) {
    load_slab(0);
    store_slab(0);
  }
  __syncthreads();
  int in_chunk = 0;
  for (int64_t ks = 0; ks < nslabs; ++ks) {
    const int cur = (int)(ks & 1);
    if (ks + 1 < nslabs) load_slab(ks + 1);
    v8bf16 af[S], bf[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      af[s] = *reinterpret_cast<const v8bf16*>(&As[cur][s * ABLK + a_off]);
      bf[s] = *reinterpret_cast<const v8bf16*>(&Bs[cur][s * BBLK + b_off]);
    }
#pragma unroll
    for (int d = 0; d < S; ++d)       // level d+2
#pragma unroll
      for (int ta = 0; ta <= d; ++ta)  // A slice ta+1, B slice d-ta+1
        lev[d] = mfma_bf16_32x32x16(bf[d - ta], af[ta], lev[d]);
    if (ks + 1 < nslabs) store_slab(cur ^ 1);
    if (++in_chunk == slabs_per_chunk) {
      fold();
      in_chunk = 0;
    }
    __syncthreads();
  }
  if (in_chunk) fold();

  // epilogue
  const int64_t gi = tm * BM + wm * 32 + (lane & 31);
  if (gi >= m) return;
  const int ei = EA[gi];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int64_t gj = tn * BN + wn * 32 + 8 * (r >> 2) + 4 * kh + (r & 3);
    if (gj < n) {
      double v = alpha * acc[r];
      if (ei != kNonFinite && EB[gj] != kNonFinite) v = ldexp(v, ei + EB[gj]);
      double* cp = C + gi * rsC + gj * csC;
      *cp = (beta == 0.0) ? v : __builtin_fma(beta, *cp, v);
    }
  }
}

// ---------------------------------------------------------------------------
// host side

// Vector v, element p at X[v*sv + p*se]. The strided form needs vectors one
// element apart (sv == 1); any other layout with se != 1 is gathered by it too,
// only less coalesced.
static hipError_t absmax(const double* X, int64_t count, int64_t len, int64_t sv, int64_t se,
                         uint64_t* mx, hipStream_t st) {
  if (count == 0) return hipSuccess;
  if (se == 1 && sv != 1) {
    hipLaunchKernelGGL(absmax_contig, dim3((unsigned)ceil_div(count, 4)), dim3(256), 0, st, X,
                       count, len, sv, mx);
  } else {
    hipLaunchKernelGGL(zero_u64, dim3((unsigned)ceil_div(count, 256)), dim3(256), 0, st, mx,
                       count);
    // enough blocks to fill the device: split the length in pieces of >= 64
    const int64_t vblocks = ceil_div(count, 256);
    int64_t ysplit = ceil_div(2048, vblocks);
    if (ysplit > ceil_div(len, 64)) ysplit = ceil_div(len, 64);
    if (ysplit < 1) ysplit = 1;
    const int64_t per = ceil_div(len, ysplit);
    hipLaunchKernelGGL(absmax_strided, dim3((unsigned)vblocks, (unsigned)ysplit), dim3(256), 0,
                       st, X, count, len, sv, se, per, mx);
  }
  return hipGetLastError();
}

template <int PANEL>
static hipError_t slice_launch(int s, const double* X, int64_t count, int64_t len, int64_t sv,
                               int64_t sp, const int* ex, int64_t vpad, int64_t lpad,
                               uint16_t* Q, hipStream_t st) {
  dim3 grid((unsigned)ceil_div(vpad, 64), (unsigned)ceil_div(lpad, 64)), block(256);
#define MOMMS_SL(S_) hipLaunchKernelGGL((slice_pack<S_, PANEL>), grid, block, 0, st, X, count, len, sv, sp, ex, vpad, lpad, Q)
  switch (s) {
    case 1: MOMMS_SL(1); break;
    case 2: MOMMS_SL(2); break;
    case 3: MOMMS_SL(3); break;
    case 4: MOMMS_SL(4); break;
    case 5: MOMMS_SL(5); break;
    case 6: MOMMS_SL(6); break;
    case 7: MOMMS_SL(7); break;
    default: MOMMS_SL(8); break;
  }
#undef MOMMS_SL
  return hipGetLastError();
}

hipError_t split_prepare(const SplitLayout& L, int64_t m, int64_t n, int64_t k, const double* A,
                         int64_t rsA, int64_t csA, const double* B, int64_t rsB, int64_t csB,
                         char* ws, hipStream_t st) {
  uint16_t* aq = (uint16_t*)(ws + L.off_aq);
  uint16_t* bq = (uint16_t*)(ws + L.off_bq);
  int* ea = (int*)(ws + L.off_ea);
  int* eb = (int*)(ws + L.off_eb);
  uint64_t* ma = (uint64_t*)(ws + L.off_maxa);
  uint64_t* mb = (uint64_t*)(ws + L.off_maxb);
  hipError_t e;
  // rows of A: vector stride rsA, element stride csA
  if ((e = absmax(A, m, k, rsA, csA, ma, st)) != hipSuccess) return e;
  // columns of B: vector stride csB, element stride rsB
  if ((e = absmax(B, n, k, csB, rsB, mb, st)) != hipSuccess) return e;
  hipLaunchKernelGGL(exponents, dim3((unsigned)ceil_div(L.mpad, 256)), dim3(256), 0, st, ma, m,
                     L.mpad, ea);
  hipLaunchKernelGGL(exponents, dim3((unsigned)ceil_div(L.npad, 256)), dim3(256), 0, st, mb, n,
                     L.npad, eb);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = slice_launch<kPanelA>(L.s, A, m, k, rsA, csA, ea, L.mpad, L.kpad, aq, st)) != hipSuccess)
    return e;
  return slice_launch<kPanelB>(L.s, B, n, k, csB, rsB, eb, L.npad, L.kpad, bq, st);
}

hipError_t split_multiply_launch(const SplitLayout& L, int64_t m, int64_t n, int64_t k,
                                 double alpha, double beta, double* C, int64_t rsC, int64_t csC,
                                 const char* ws, hipStream_t st) {
  const uint16_t* aq = (const uint16_t*)(ws + L.off_aq);
  const uint16_t* bq = (const uint16_t*)(ws + L.off_bq);
  const int* ea = (const int*)(ws + L.off_ea);
  const int* eb = (const int*)(ws + L.off_eb);
  const int64_t tiles = (L.mpad / kPanelA) * (L.npad / kPanelB);
  const int spc = split_chunk(L.s) / kSlab;
  const int64_t kp = (k > 0) ? L.kpad : 0;
  dim3 grid((unsigned)tiles), block(512);
#define MOMMS_MU(S_) hipLaunchKernelGGL((split_multiply<S_>), grid, block, 0, st, m, n, kp, alpha, beta, C, rsC, csC, aq, bq, ea, eb, L.mpad, L.npad, spc)
  switch (L.s) {
    case 1: MOMMS_MU(1); break;
    case 2: MOMMS_MU(2); break;
    case 3: MOMMS_MU(3); break;
    case 4: MOMMS_MU(4); break;
    case 5: MOMMS_MU(5); break;
    case 6: MOMMS_MU(6); break;
    case 7: MOMMS_MU(7); break;
    default: MOMMS_MU(8); break;
  }
#undef MOMMS_MU
  return hipGetLastError();
}

}  // namespace momms

// ---------------------------------------------------------------------------
// C ABI

using namespace momms;

extern "C" size_t momms_hip_split_bf16_workspace_size(int64_t m, int64_t n, int64_t k,
                                                      int slices) {
  if (m < 0 || n < 0 || k < 0 || slices < 1 || slices > kMaxSlices) return 0;
  return split_layout(m, n, k, slices).bytes;
}

extern "C" int momms_hip_split_bf16_prepare(int64_t m, int64_t n, int64_t k, const double* A,
                                            int64_t rsA, int64_t csA, const double* B,
                                            int64_t rsB, int64_t csB, int slices,
                                            void* workspace, size_t workspace_bytes,
                                            void* stream) {
  static const char* fn = "momms_hip_split_bf16_prepare";
  // argument numbering of this entry point: m n k A rsA csA B rsB csB slices ws bytes stream
  if (m < 0) return bad_arg(fn, 1, "m < 0");
  if (n < 0) return bad_arg(fn, 2, "n < 0");
  if (k < 0) return bad_arg(fn, 3, "k < 0");
  int st;
  if ((st = check_operand(fn, 4, A, m, k, rsA, csA))) return st;
  if ((st = check_operand(fn, 7, B, k, n, rsB, csB))) return st;
  if (slices < 1 || slices > kMaxSlices) return bad_arg(fn, 10, "slices must be 1..8");
  const SplitLayout L = split_layout(m, n, k, slices);
  if (workspace == nullptr) return bad_arg(fn, 11, "workspace is null");
  if (workspace_bytes < L.bytes) return bad_arg(fn, 12, "workspace too small");
  if (m == 0 && n == 0) return 0;
  return hip_status(fn, split_prepare(L, m, n, k, A, rsA, csA, B, rsB, csB, (char*)workspace,
                                      (hipStream_t)stream));
}

extern "C" int momms_hip_split_bf16_multiply(int64_t m, int64_t n, int64_t k,
                                             const double* alpha, const double* beta, double* C,
                                             int64_t rsC, int64_t csC, int slices,
                                             const void* workspace, size_t workspace_bytes,
                                             void* stream) {
  static const char* fn = "momms_hip_split_bf16_multiply";
  if (m < 0) return bad_arg(fn, 1, "m < 0");
  if (n < 0) return bad_arg(fn, 2, "n < 0");
  if (k < 0) return bad_arg(fn, 3, "k < 0");
  if (alpha == nullptr) return bad_arg(fn, 4, "alpha is null");
  if (beta == nullptr) return bad_arg(fn, 5, "beta is null");
  int st;
  if ((st = check_operand(fn, 6, C, m, n, rsC, csC))) return st;
  if (slices < 1 || slices > kMaxSlices) return bad_arg(fn, 9, "slices must be 1..8");
  const SplitLayout L = split_layout(m, n, k, slices);
  if (workspace == nullptr) return bad_arg(fn, 10, "workspace is null");
  if (workspace_bytes < L.bytes) return bad_arg(fn, 11, "workspace too small");
  if (m == 0 || n == 0) return 0;
  return hip_status(fn, split_multiply_launch(L, m, n, k, *alpha, *beta, C, rsC, csC,
                                              (const char*)workspace, (hipStream_t)stream));
}

extern "C" int momms_hip_dgemm_split_bf16(int64_t m, int64_t n, int64_t k, const double* alpha,
                                          const double* A, int64_t rsA, int64_t csA,
                                          const double* B, int64_t rsB, int64_t csB,
                                          const double* beta, double* C, int64_t rsC,
                                          int64_t csC, int slices, void* workspace,
                                          size_t workspace_bytes, void* stream) {
  static const char* fn = "momms_hip_dgemm_split_bf16";
  if (m < 0) return bad_arg(fn, 1, "m < 0");
  if (n < 0) return bad_arg(fn, 2, "n < 0");
  if (k < 0) return bad_arg(fn, 3, "k < 0");
  if (alpha == nullptr) return bad_arg(fn, 4, "alpha is null");
  int st;
  if ((st = check_operand(fn, 5, A, m, k, rsA, csA))) return st;
  if ((st = check_operand(fn, 8, B, k, n, rsB, csB))) return st;
  if (beta == nullptr) return bad_arg(fn, 11, "beta is null");
  if ((st = check_operand(fn, 12, C, m, n, rsC, csC))) return st;
  if (slices < 1 || slices > kMaxSlices) return bad_arg(fn, 15, "slices must be 1..8");
  const SplitLayout L = split_layout(m, n, k, slices);
  if (workspace == nullptr) return bad_arg(fn, 16, "workspace is null");
  if (workspace_bytes < L.bytes) return bad_arg(fn, 17, "workspace too small");
  if (m == 0 || n == 0) return 0;
  hipError_t e = split_prepare(L, m, n, k, A, rsA, csA, B, rsB, csB, (char*)workspace,
                               (hipStream_t)stream);
  if (e != hipSuccess) return hip_status(fn, e);
  return hip_status(fn, split_multiply_launch(L, m, n, k, *alpha, *beta, C, rsC, csC,
                                              (const char*)workspace, (hipStream_t)stream));
}
