// Kernels shared by the two residue forms of the FP64-accurate GEMM
// (crt_bf16.hip: BF16 residues of moduli <= 511; crt_i8.hip: INT8 residues of
// moduli <= 255): the moduli constants, the residue pack and the Chinese
// remainder reconstruction, each a template on the form.
#pragma once
#include <math.h>

#include "c_update.hpp"
#include "crt_recon.hpp"
#include "mfma.hpp"
#include "panel_gemm.hpp"
#include "split_tile.hpp"

namespace momms {

struct CrtTables {
  float pf[kCrtMaxModuli], invpf[kCrtMaxModuli];
  double pd[kCrtMaxModuli], invpd[kCrtMaxModuli];
};

constexpr CrtTables make_crt_tables(const int* mods) {
  CrtTables T{};
  for (int l = 0; l < kCrtMaxModuli; ++l) {
    T.pf[l] = (float)mods[l];
    T.invpf[l] = 1.0f / (float)mods[l];
    T.pd[l] = (double)mods[l];
    T.invpd[l] = 1.0 / (double)mods[l];
  }
  return T;
}

static __constant__ CrtTables kCrtBf16Tab = make_crt_tables(kCrtModuli);
static __constant__ CrtTables kCrtI8Tab = make_crt_tables(kCrtI8Moduli);

template <bool I8>
__device__ __forceinline__ const CrtTables& crt_tables() {
  return I8 ? kCrtI8Tab : kCrtBf16Tab;
}

// Bits of the exponent-aligned integers for k and the first N moduli of a
// table: beta = min(53, floor((bitlen(P) - 2 - kb) / 2), floor((125 - kb) / 2))
// with kb = ceil(log2 k), so |C'| <= k 2^(2 beta) < P/2 and <= 2^125; -1 if
// N is out of range, below 1 if N moduli cannot hold a k-term product.
int crt_beta_for(int64_t k, int N, const int* mods);

// ---------------------------------------------------------------------------
// 2. residues in packed panels.
//
// Vector v (a row of A or a column of B), element p at X[v*sv + p*sp]. With
// G residues per 16-byte group (BF16: G = 8, INT8: G = 16) and slabs of
// D = 2G k-values, the residue modulo p_l goes to byte group
//   ((l*(vpad/PANEL) + pv)*S + ks)*(2*PANEL) + kh*PANEL + v%PANEL, element p%G
// with pv = v / PANEL, ks = p / D, kh = (p % D) / G and S = panel_stride(lpad/D)
// slabs per panel (one unused slab between panels); padding is 0. A 64 x 64
// block of (v, p) goes through LDS so the global read (along whichever index
// is unit-stride) and the 16-byte packed writes are both coalesced.

template <int PANEL, bool I8>
__global__ void __launch_bounds__(256)
crt_pack(const double* __restrict__ X, int64_t count, int64_t len, int64_t sv, int64_t sp,
         const int* __restrict__ ex, int64_t vpad, int64_t lpad, int nmod, int beta,
         uint16_t* __restrict__ R) {
  constexpr int G = I8 ? 16 : 8, D = 2 * G;  // residues per 16 bytes, k per slab
  __shared__ double tile[64][65];
  const int tid = threadIdx.x;
  const int64_t v0 = (int64_t)blockIdx.x * 64, p0 = (int64_t)blockIdx.y * 64;
  const bool v_unit = (sv == 1);
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
  const CrtTables& T = crt_tables<I8>();
  const int64_t nslabs = lpad / D;
  const int64_t mod_stride = vpad * panel_stride(nslabs) * 16;  // uint16 per modulus
#pragma unroll
  for (int r = 0; r < 4096 / (256 * G); ++r) {
    const int c = tid + 256 * r;
    const int vv = c & 63;
    const int pc = c >> 6;  // group of G along p
    const int64_t gv = v0 + vv;
    const int64_t gp = p0 + pc * G;
    if (gv >= vpad || gp >= lpad) continue;
    const int E = (gv < count) ? ex[gv] : 0;
    const bool finite = (E != kNonFinite);
    double q[G];
#pragma unroll
    for (int e = 0; e < G; ++e) q[e] = finite ? rint(ldexp(tile[vv][pc * G + e], beta - E)) : 0.0;
    const int64_t pv = gv / PANEL, ks = gp / D;
    const int kh = (int)((gp % D) / G);
    uint16_t* dst = R + ((pv * panel_stride(nslabs) + ks) * (2 * PANEL) + (int64_t)kh * PANEL +
                         (gv % PANEL)) * 8;
    for (int l = 0; l < nmod; ++l) {
      const double p = T.pd[l], invp = T.invpd[l];
      const float pf = T.pf[l], invpf = T.invpf[l];
      uint32_t bits[G];
#pragma unroll
      for (int e = 0; e < G; ++e) {
        // x = q - p*rint(q/p) is an integer below 0.51 p in magnitude (exact
        // by fma); the same step once more in fp32 on x (|x| < 2^9, and x/p
        // is at least 1/(2p) away from +-1/2, so rint(x/p) is exact) gives the
        // balanced residue, |rr| <= (p-1)/2
        const float x = (float)__builtin_fma(-rint(q[e] * invp), p, q[e]);
        const float rr = __builtin_fmaf(-__builtin_rintf(x * invpf), pf, x);
        // BF16: |rr| <= 255 is exact in bf16 (the high half of the float);
        // INT8: |rr| <= 127, its two's complement byte
        bits[e] = I8 ? ((uint32_t)(int)rr & 0xFFu) : (__float_as_uint((float)rr) >> 16);
      }
      uint4 w;
      if constexpr (I8) {
        w.x = bits[0] | (bits[1] << 8) | (bits[2] << 16) | (bits[3] << 24);
        w.y = bits[4] | (bits[5] << 8) | (bits[6] << 16) | (bits[7] << 24);
        w.z = bits[8] | (bits[9] << 8) | (bits[10] << 16) | (bits[11] << 24);
        w.w = bits[12] | (bits[13] << 8) | (bits[14] << 16) | (bits[15] << 24);
      } else {
        w.x = bits[0] | (bits[1] << 16);
        w.y = bits[2] | (bits[3] << 16);
        w.z = bits[4] | (bits[5] << 16);
        w.w = bits[6] | (bits[7] << 16);
      }
      *reinterpret_cast<uint4*>(dst + (int64_t)l * mod_stride) = w;
    }
  }
}

// ---------------------------------------------------------------------------
// 4. reconstruction.

// One thread: kCrtRows consecutive rows i0.. of one column j (the residue
// matrices are padded to whole tiles, so the residue loads of kCrtRows int16
// stay in bounds); the constants of each modulus are loaded once for all of
// them. Two rows per thread keep the sums and the loads in flight at 77
// registers, six wavefronts per SIMD, so the arithmetic of some overlaps the
// memory traffic of others (eight rows: 150 registers, three wavefronts per
// SIMD, 0.88 ms at 8192^2 and 13 moduli; four rows 0.81 ms; two rows 0.79 ms).
constexpr int kCrtRows = 2;

// The residue matrices hold int16 (BF16 form) or int8 (INT8 form) balanced
// residues, column-major with leading dimension mpad.
template <int N, bool I8>
__global__ void __launch_bounds__(256)
crt_reconstruct(int64_t m, int64_t n, int64_t k, int beta_bits, double alpha, double beta,
                double* __restrict__ C, int64_t rsC, int64_t csC, const void* __restrict__ RCv,
                int64_t mpad, int64_t npad, const int* __restrict__ EA,
                const int* __restrict__ EB, const double* __restrict__ A, int64_t rsA,
                int64_t csA, const double* __restrict__ B, int64_t rsB, int64_t csB) {
  constexpr CrtRecon K = crt_recon_constants(N, I8 ? kCrtI8Moduli : kCrtModuli);
  // a thread reconstructs rows i0 .. i0+kCrtRows-1 of column j;
  // the results go through LDS so that the C update of the block runs along
  // columns, 64 consecutive rows per wavefront access
  __shared__ double vt[256 * (kCrtRows + 1)];
  const int tid = threadIdx.x;
  const int64_t rb = (int64_t)blockIdx.x * (256 * kCrtRows);
  const int64_t i0 = rb + (int64_t)tid * kCrtRows;
  const bool active = i0 < m;  // (rows up to i0 + kCrtRows - 1 < mpad then)
  for (int64_t j = blockIdx.y; j < n; j += gridDim.y) {
    double v[kCrtRows];
#pragma unroll
    for (int e = 0; e < kCrtRows; ++e) v[e] = 0.0;
    if (active) {
      CrtSums a[kCrtRows];
#pragma unroll
      for (int e = 0; e < kCrtRows; ++e) crt_sums_init(a[e]);
#pragma unroll
      for (int l = 0; l < N; ++l) {
        int rr[kCrtRows];
        if constexpr (I8) {
          const int8_t* src = (const int8_t*)RCv + (int64_t)l * mpad * npad + j * mpad + i0;
          const uint16_t h = *reinterpret_cast<const uint16_t*>(src);  // two int8 residues
#pragma unroll
          for (int e = 0; e < kCrtRows; ++e) rr[e] = (int8_t)(h >> (8 * e));
        } else {
          const int16_t* src = (const int16_t*)RCv + (int64_t)l * mpad * npad + j * mpad + i0;
          const uint32_t h = *reinterpret_cast<const uint32_t*>(src);  // two int16 residues
#pragma unroll
          for (int e = 0; e < kCrtRows; ++e) rr[e] = (int16_t)(h >> (16 * e));
        }
#pragma unroll
        for (int e = 0; e < kCrtRows; ++e) crt_sums_add(a[e], rr[e], l, K);
      }
      const int ej = EB[j];
#pragma unroll
      for (int e = 0; e < kCrtRows; ++e) {
        const int64_t i = i0 + e;
        const int ei = EA[i];  // EA holds mpad entries
        if (i < m && (ei == kNonFinite || ej == kNonFinite)) {
          double s = 0.0;
          for (int64_t p = 0; p < k; ++p) s = __builtin_fma(A[i * rsA + p * csA], B[p * rsB + j * csB], s);
          v[e] = alpha * s;
        } else {
          v[e] = alpha * ldexp(crt_i128_to_f64(crt_sums_value(a[e], K)), ei + ej - 2 * beta_bits);
        }
      }
    }
    // row rb + q sits at vt[(q / kCrtRows) * (kCrtRows + 1) + q % kCrtRows] (the
    // pad keeps the stores of a wavefront on distinct banks)
#pragma unroll
    for (int e = 0; e < kCrtRows; ++e) vt[tid * (kCrtRows + 1) + e] = v[e];
    __syncthreads();
    int64_t idx[kCrtRows];
    bool ok[kCrtRows];
    double w[kCrtRows];
#pragma unroll
    for (int e = 0; e < kCrtRows; ++e) {
      const int q = tid + 256 * e;
      const int64_t i = rb + q;
      ok[e] = i < m;
      idx[e] = i * rsC + j * csC;
      w[e] = vt[(q / kCrtRows) * (kCrtRows + 1) + q % kCrtRows];
    }
    __syncthreads();  // vt free for the next column
    update_c(C, idx, ok, w, beta);
  }
}

}  // namespace momms
