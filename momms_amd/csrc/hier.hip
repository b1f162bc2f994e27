// Device-resident hierarchical (multi-level blocked) matrices: the layout of
// momms' Hierarch (src/matrix/hierarch.rs:18-184), packing a strided matrix
// into it (Packer<Matrix, Hierarch>, src/composables/pack.rs:153-309) and
// adding one back into a strided matrix (Unpacker<Matrix, Hierarch>,
// src/composables/unpack.rs:61-98, :174-213).
//
// The offset of element (y, x) is a y part plus an x part. Both kernels
// therefore work from two small tables, Oy[y] over the padded height and
// Ox[x] over the padded width (one pass of the level walk per row and per
// column instead of per element), kept in the library's per-stream scratch.
// A place outside [0, capacity) (a hand-made layout; never one that
// momms_hier_layout_init returns) is neither written nor read: it unpacks as 0.
// A workgroup moves 256 consecutive indices along the strided matrix's unit
// stride times G consecutive indices across it through LDS, so that both the
// strided side and the layout side are accessed in long contiguous runs.
#include <hip/hip_runtime.h>

#include "mfma.hpp"
#include "momms_hip.h"
#include "scratch.hpp"
#include "status.hpp"

namespace momms {

// offset tables: t[i] for i < h_pad is the y part of row i, t[h_pad + j] the x
// part of column j (hierarch.rs:153-184)
__global__ void __launch_bounds__(256)
hier_tables(momms_hier_layout L, int64_t* __restrict__ t) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.h_pad + L.w_pad) return;
  const bool is_y = i < L.h_pad;
  int64_t v = is_y ? i : i - L.h_pad;
  const int64_t n = is_y ? L.ny : L.nx;
  const int64_t* blk = is_y ? L.y_blksz : L.x_blksz;
  const int64_t* str = is_y ? L.y_stride : L.x_stride;
  int64_t off = 0;
  for (int64_t l = 0; l < n; ++l) {
    const int64_t b = v / blk[l];
    off += b * str[l];
    v -= b * blk[l];
  }
  t[i] = off + v * (is_y ? L.lrs : L.lcs);
}

// A workgroup moves a tile of 256 lane indices u (rows when LX is false,
// columns when it is true) by G minor indices v through LDS:
//  - on the strided side, lane u holds u and steps through the G minor
//    indices, so every access runs along the matrix's unit stride;
//  - on the layout side, element e = q*256 + tid of the tile is taken in
//    layout order: runs of R consecutive lane indices (the leaf's contiguous
//    run along u, a power of two dividing 256; 1 when the leaf runs along v)
//    by the G minor indices, so a wave touches 64 consecutive places.
// Placement is from the offset tables in either order; the order only decides
// which places one wave instruction touches.
template <typename T, int G>
struct HierTile {
  static constexpr int kStride = 256 + 4;  // conflict-free reads in layout order
  T v[G * kStride];
  int64_t ou[256];
  int64_t ov[G];
};

template <int G>
__device__ __forceinline__ void layout_order(int q, int log2r, int& uu, int& c) {
  const uint32_t e = (uint32_t)q * 256 + threadIdx.x;
  const uint32_t i = (e >> log2r) / G;  // group of R lane indices x G
  const uint32_t r = e - i * (G << log2r);
  uu = (int)((i << log2r) + (r & ((1u << log2r) - 1)));
  c = (int)(r >> log2r);
}

// h == 0 or w == 0 (grid-uniform): nothing is read and every place is padding.
// PADW: padding places are written (zeros); otherwise only the matrix's own
// places are.
template <typename T, int G, bool LX, bool PADW>
__global__ void __launch_bounds__(256)
hier_pack_kernel(int64_t h, int64_t w, int64_t h_pad, int64_t w_pad, uint64_t cap, uint32_t nbu,
                 int log2r, const int64_t* __restrict__ oy, const int64_t* __restrict__ ox,
                 const T* __restrict__ A, int64_t rs, int64_t cs, T* __restrict__ P) {
  __shared__ HierTile<T, G> S;
  const uint32_t bu = blockIdx.x % nbu, bv = blockIdx.x / nbu;
  const int64_t u0 = (int64_t)bu * 256, v0 = (int64_t)bv * G;
  const int64_t u = u0 + threadIdx.x;
  const int64_t U = LX ? w_pad : h_pad, V = LX ? h_pad : w_pad;
  const int64_t hu = LX ? w : h, hv = LX ? h : w;  // logical extents
  const int64_t su = LX ? cs : rs, sv = LX ? rs : cs;
  const int64_t* ou = LX ? ox : oy;
  const int64_t* ov = LX ? oy : ox;
  T v[G];
  if (h > 0 && w > 0) {
    // all loads first (clamped into the matrix; padding is replaced by 0)
    const int64_t uc = u < hu ? u : hu - 1;
#pragma unroll
    for (int c = 0; c < G; ++c) {
      const int64_t j = v0 + c < hv ? v0 + c : hv - 1;
      v[c] = A[uc * su + j * sv];
    }
#pragma unroll
    for (int c = 0; c < G; ++c)
      if (u >= hu || v0 + c >= hv) v[c] = T(0);
  } else {
#pragma unroll
    for (int c = 0; c < G; ++c) v[c] = T(0);
  }
  S.ou[threadIdx.x] = ou[u < U ? u : U - 1];
  if (threadIdx.x < G) S.ov[threadIdx.x] = ov[v0 + threadIdx.x < V ? v0 + threadIdx.x : V - 1];
#pragma unroll
  for (int c = 0; c < G; ++c) S.v[c * HierTile<T, G>::kStride + threadIdx.x] = v[c];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < G; ++q) {
    int uu, c;
    layout_order<G>(q, log2r, uu, c);
    const int64_t o = S.ou[uu] + S.ov[c];
    const bool place = PADW ? (u0 + uu < U && v0 + c < V) : (u0 + uu < hu && v0 + c < hv);
    if (place && (uint64_t)o < cap) P[o] = S.v[c * HierTile<T, G>::kStride + uu];
  }
}

// ADD: C += P at the matrix's places; otherwise C := P there (C not read).
template <typename T, int G, bool LX, bool ADD>
__global__ void __launch_bounds__(256)
hier_unpack_add_kernel(int64_t h, int64_t w, uint64_t cap, uint32_t nbu, int log2r,
                       const int64_t* __restrict__ oy, const int64_t* __restrict__ ox,
                       const T* __restrict__ P, T* __restrict__ C, int64_t rs,
                       int64_t cs) {
  __shared__ HierTile<T, G> S;
  const uint32_t bu = blockIdx.x % nbu, bv = blockIdx.x / nbu;
  const int64_t u0 = (int64_t)bu * 256, v0 = (int64_t)bv * G;
  const int64_t u = u0 + threadIdx.x;
  const int64_t hu = LX ? w : h, hv = LX ? h : w;
  const int64_t su = LX ? cs : rs, sv = LX ? rs : cs;
  const int64_t* ou = LX ? ox : oy;
  const int64_t* ov = LX ? oy : ox;
  // C first (its loads overlap the layout side), clamped into the matrix
  const int64_t uc = u < hu ? u : hu - 1;
  T c0[G];
#pragma unroll
  for (int c = 0; c < G; ++c) {
    const int64_t j = v0 + c < hv ? v0 + c : hv - 1;
    c0[c] = ADD ? C[uc * su + j * sv] : T(0);
  }
  S.ou[threadIdx.x] = ou[uc];
  if (threadIdx.x < G) S.ov[threadIdx.x] = ov[v0 + threadIdx.x < hv ? v0 + threadIdx.x : hv - 1];
  __syncthreads();
  T p[G];
#pragma unroll
  for (int q = 0; q < G; ++q) {
    int uu, c;
    layout_order<G>(q, log2r, uu, c);
    const int64_t o = S.ou[uu] + S.ov[c];
    const bool in = (uint64_t)o < cap;
    p[q] = P[in ? o : 0];
    p[q] = in ? p[q] : T(0);
  }
#pragma unroll
  for (int q = 0; q < G; ++q) {
    int uu, c;
    layout_order<G>(q, log2r, uu, c);
    S.v[c * HierTile<T, G>::kStride + uu] = p[q];
  }
  __syncthreads();
  if (u >= hu) return;
#pragma unroll
  for (int c = 0; c < G; ++c)
    if (v0 + c < hv)
      C[u * su + (v0 + c) * sv] =
          ADD ? S.v[c * HierTile<T, G>::kStride + threadIdx.x] + c0[c]
              : S.v[c * HierTile<T, G>::kStride + threadIdx.x];
}

// Lanes along the source's unit stride (rows of a column-major matrix, the
// default; columns of a row-major one). G tiles the leaf's contiguous run
// along the minor index when the leaf runs that way (its width X for leaves
// stored by rows and lanes down columns, its height Y in the transposed
// case), else 16; R is the power-of-two part (at most 256) of the leaf's run
// along the lane index, 1 when the leaf runs along the minor index.
struct HierShape {
  bool lx;
  int g;
  int log2r;
};

static HierShape hier_shape(const momms_hier_layout& L, int64_t rs, int64_t cs) {
  HierShape sh;
  sh.lx = cs == 1 && rs != 1;
  const int64_t Y = L.ny ? L.y_blksz[L.ny - 1] : L.h_pad, X = L.nx ? L.x_blksz[L.nx - 1] : L.w_pad;
  const bool run_u = sh.lx ? L.lcs == 1 : L.lrs == 1;  // leaf contiguous along the lanes
  const int64_t Ru = run_u ? (sh.lx ? X : Y) : 1, Rv = run_u ? 1 : (sh.lx ? Y : X);
  sh.g = Rv % 16 == 0 || Rv == 1 ? 16 : Rv % 12 == 0 ? 12 : 8;
  sh.log2r = 0;
  while (sh.log2r < 8 && Ru % ((int64_t)2 << sh.log2r) == 0) ++sh.log2r;
  return sh;
}

template <typename T, int G, bool LX, bool PADW>
static void launch_pack(const momms_hier_layout& L, int log2r, const int64_t* t, const T* A,
                        int64_t rs, int64_t cs, T* P, hipStream_t s) {
  const int64_t U = LX ? L.w_pad : L.h_pad, V = LX ? L.h_pad : L.w_pad;
  const uint32_t nbu = (uint32_t)ceil_div(U, 256);
  const uint32_t blocks = nbu * (uint32_t)ceil_div(V, G);
  hipLaunchKernelGGL((hier_pack_kernel<T, G, LX, PADW>), dim3(blocks), dim3(256), 0, s, L.h, L.w,
                     L.h_pad, L.w_pad, (uint64_t)L.capacity, nbu, log2r, t, t + L.h_pad, A, rs, cs,
                     P);
}

template <typename T, int G, bool LX, bool ADD>
static void launch_unpack(const momms_hier_layout& L, int log2r, const int64_t* t,
                          const T* P, T* C, int64_t rs, int64_t cs, hipStream_t s) {
  const int64_t U = LX ? L.w : L.h, V = LX ? L.h : L.w;
  const uint32_t nbu = (uint32_t)ceil_div(U, 256);
  const uint32_t blocks = nbu * (uint32_t)ceil_div(V, G);
  hipLaunchKernelGGL((hier_unpack_add_kernel<T, G, LX, ADD>), dim3(blocks), dim3(256), 0, s, L.h,
                     L.w, (uint64_t)L.capacity, nbu, log2r, t, t + L.h_pad, P, C, rs, cs);
}

static hipError_t hier_launch_tables(const momms_hier_layout& L, hipStream_t st, int64_t** t) {
  hipError_t e = stream_scratch(st, (size_t)(L.h_pad + L.w_pad) * sizeof(int64_t), (void**)t);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(hier_tables, dim3((unsigned)ceil_div(L.h_pad + L.w_pad, 256)), dim3(256), 0,
                     st, L, *t);
  return hipGetLastError();
}

static int check_layout(const char* fn, int arg, const momms_hier_layout* L) {
  if (!L) return bad_arg(fn, arg, "layout is null");
  if (L->h < 0 || L->w < 0 || L->h_pad < L->h || L->w_pad < L->w || L->ny < 0 || L->nx < 0 ||
      L->ny > MOMMS_HIER_MAX_LEVELS || L->nx > MOMMS_HIER_MAX_LEVELS ||
      L->h_pad > ((int64_t)1 << 31) || L->w_pad > ((int64_t)1 << 31) ||
      L->capacity != L->h_pad * L->w_pad)
    return bad_arg(fn, arg, "layout is inconsistent");
  // one launch of at most 2^32 - 1 workgroups of 256 lanes x 8 indices
  if (ceil_div(L->h_pad, 256) * ceil_div(L->w_pad, 8) >= ((int64_t)1 << 32))
    return bad_arg(fn, arg, "layout too large for one launch");
  for (int64_t l = 0; l < L->ny; ++l)
    if (L->y_blksz[l] <= 0) return bad_arg(fn, arg, "layout has a non-positive block size");
  for (int64_t l = 0; l < L->nx; ++l)
    if (L->x_blksz[l] <= 0) return bad_arg(fn, arg, "layout has a non-positive block size");
  return 0;
}

// strided matrix -> layout (all padding places written as 0 when padw)
template <typename T, bool PADW>
static hipError_t hier_pack_impl(const momms_hier_layout& L, const T* A, int64_t rs,
                                 int64_t cs, T* P, hipStream_t s) {
  if (L.capacity == 0 || (!PADW && (L.h == 0 || L.w == 0))) return hipSuccess;
  int64_t* t = nullptr;
  hipError_t e = hier_launch_tables(L, s, &t);
  if (e != hipSuccess) return e;
  const HierShape sh = hier_shape(L, rs, cs);
  if (sh.lx) {
    if (sh.g == 16) launch_pack<T, 16, true, PADW>(L, sh.log2r, t, A, rs, cs, P, s);
    else if (sh.g == 12) launch_pack<T, 12, true, PADW>(L, sh.log2r, t, A, rs, cs, P, s);
    else launch_pack<T, 8, true, PADW>(L, sh.log2r, t, A, rs, cs, P, s);
  } else {
    if (sh.g == 16) launch_pack<T, 16, false, PADW>(L, sh.log2r, t, A, rs, cs, P, s);
    else if (sh.g == 12) launch_pack<T, 12, false, PADW>(L, sh.log2r, t, A, rs, cs, P, s);
    else launch_pack<T, 8, false, PADW>(L, sh.log2r, t, A, rs, cs, P, s);
  }
  return hipGetLastError();
}

// layout -> strided matrix: C += P (ADD) or C := P at the matrix's places
template <typename T, bool ADD>
static hipError_t hier_unpack_impl(const momms_hier_layout& L, const T* P, T* C,
                                   int64_t rs, int64_t cs, hipStream_t s) {
  if (L.h == 0 || L.w == 0) return hipSuccess;
  int64_t* t = nullptr;
  hipError_t e = hier_launch_tables(L, s, &t);
  if (e != hipSuccess) return e;
  const HierShape sh = hier_shape(L, rs, cs);
  if (sh.lx) {
    if (sh.g == 16) launch_unpack<T, 16, true, ADD>(L, sh.log2r, t, P, C, rs, cs, s);
    else if (sh.g == 12) launch_unpack<T, 12, true, ADD>(L, sh.log2r, t, P, C, rs, cs, s);
    else launch_unpack<T, 8, true, ADD>(L, sh.log2r, t, P, C, rs, cs, s);
  } else {
    if (sh.g == 16) launch_unpack<T, 16, false, ADD>(L, sh.log2r, t, P, C, rs, cs, s);
    else if (sh.g == 12) launch_unpack<T, 12, false, ADD>(L, sh.log2r, t, P, C, rs, cs, s);
    else launch_unpack<T, 8, false, ADD>(L, sh.log2r, t, P, C, rs, cs, s);
  }
  return hipGetLastError();
}

}  // namespace momms

using namespace momms;

extern "C" int momms_hier_layout_init(momms_hier_layout* L, int64_t h, int64_t w,
                                      const momms_algo_step* steps, int64_t nsteps, int32_t ydim,
                                      int32_t xdim, int64_t lh, int64_t lw, int64_t lrs,
                                      int64_t lcs) {
  if (!L || h < 0 || w < 0 || nsteps < 0 || (nsteps > 0 && !steps) || lrs < 0 || lcs < 0)
    return -1;
  // outermost step of a dimension: its block is the top-level one
  // (get_top_level_dim_size, hierarch.rs:50-58)
  auto top = [&](int32_t dim, int64_t dflt) {
    for (int64_t i = nsteps - 1; i >= 0; --i)
      if (steps[i].kind == dim) return steps[i].bsz;
    return dflt;
  };
  const int64_t yt = top(ydim, lh), xt = top(xdim, lw);
  if (yt <= 0 || xt <= 0) return -1;
  const int64_t nby = h == 0 ? 1 : (h - 1) / yt + 1, nbx = w == 0 ? 1 : (w - 1) / xt + 1;
  momms_hier_layout R{};
  R.h = h;
  R.w = w;
  R.h_pad = nby * yt;
  R.w_pad = nbx * xt;
  R.capacity = R.h_pad * R.w_pad;
  R.lrs = lrs;
  R.lcs = lcs;
  // parse_input_hierarchy (hierarch.rs:59-89): outermost step first; an x
  // level strides over the current block height, a y level over its width
  int64_t ht = R.h_pad, wt = R.w_pad;
  for (int64_t i = nsteps - 1; i >= 0; --i) {
    if (steps[i].bsz <= 0) return -1;
    if (steps[i].kind == xdim) {
      if (R.nx == MOMMS_HIER_MAX_LEVELS) return -2;
      R.x_blksz[R.nx] = steps[i].bsz;
      R.x_stride[R.nx] = ht * steps[i].bsz;
      ++R.nx;
      wt = steps[i].bsz;
    }
    if (steps[i].kind == ydim) {
      if (R.ny == MOMMS_HIER_MAX_LEVELS) return -2;
      R.y_blksz[R.ny] = steps[i].bsz;
      R.y_stride[R.ny] = wt * steps[i].bsz;
      ++R.ny;
      ht = steps[i].bsz;
    }
  }
  // every place holds one element only when each block splits evenly into
  // the next level's and the leaf (innermost y block x innermost x block) is
  // stored densely, by columns or by rows; the reference builds such layouts
  // only and would overlap elements otherwise
  for (int64_t l = 1; l < R.ny; ++l)
    if (R.y_blksz[l - 1] % R.y_blksz[l]) return -3;
  for (int64_t l = 1; l < R.nx; ++l)
    if (R.x_blksz[l - 1] % R.x_blksz[l]) return -3;
  const int64_t Y = R.ny ? R.y_blksz[R.ny - 1] : R.h_pad, X = R.nx ? R.x_blksz[R.nx - 1] : R.w_pad;
  if (!((lrs == 1 && lcs == Y) || (lcs == 1 && lrs == X))) return -3;
  *L = R;
  return 0;
}

extern "C" int momms_hip_hier_pack(const momms_hier_layout* L, const double* A, int64_t rsA,
                                   int64_t csA, double* P, void* stream) {
  static const char* fn = "momms_hip_hier_pack";
  int st;
  if ((st = check_layout(fn, 1, L))) return st;
  if ((st = check_operand(fn, 2, A, L->h, L->w, rsA, csA))) return st;
  if (!P && L->capacity) return bad_arg(fn, 5, "P is null");
  return hip_status(fn, hier_pack_impl<double, true>(*L, A, rsA, csA, P, (hipStream_t)stream));
}

extern "C" int momms_hip_hier_unpack_add(const momms_hier_layout* L, const double* P, double* C,
                                         int64_t rsC, int64_t csC, void* stream) {
  static const char* fn = "momms_hip_hier_unpack_add";
  int st;
  if ((st = check_layout(fn, 1, L))) return st;
  if (!P && L->capacity) return bad_arg(fn, 2, "P is null");
  if ((st = check_operand(fn, 3, C, L->h, L->w, rsC, csC))) return st;
  return hip_status(fn, hier_unpack_impl<double, true>(*L, P, C, rsC, csC, (hipStream_t)stream));
}

extern "C" int momms_hip_hier_pack_f32(const momms_hier_layout* L, const float* A, int64_t rsA,
                                       int64_t csA, float* P, void* stream) {
  static const char* fn = "momms_hip_hier_pack_f32";
  int st;
  if ((st = check_layout(fn, 1, L))) return st;
  if ((st = check_operand(fn, 2, A, L->h, L->w, rsA, csA))) return st;
  if (!P && L->capacity) return bad_arg(fn, 5, "P is null");
  return hip_status(fn, hier_pack_impl<float, true>(*L, A, rsA, csA, P, (hipStream_t)stream));
}

extern "C" int momms_hip_hier_unpack_add_f32(const momms_hier_layout* L, const float* P,
                                             float* C, int64_t rsC, int64_t csC, void* stream) {
  static const char* fn = "momms_hip_hier_unpack_add_f32";
  int st;
  if ((st = check_layout(fn, 1, L))) return st;
  if (!P && L->capacity) return bad_arg(fn, 2, "P is null");
  if ((st = check_operand(fn, 3, C, L->h, L->w, rsC, csC))) return st;
  return hip_status(fn, hier_unpack_impl<float, true>(*L, P, C, rsC, csC, (hipStream_t)stream));
}

extern "C" int momms_hip_hier_unpack_f32(const momms_hier_layout* L, const float* P, float* C,
                                         int64_t rsC, int64_t csC, void* stream) {
  static const char* fn = "momms_hip_hier_unpack_f32";
  int st;
  if ((st = check_layout(fn, 1, L))) return st;
  if (!P && L->capacity) return bad_arg(fn, 2, "P is null");
  if ((st = check_operand(fn, 3, C, L->h, L->w, rsC, csC))) return st;
  return hip_status(fn, hier_unpack_impl<float, false>(*L, P, C, rsC, csC, (hipStream_t)stream));
}

extern "C" int momms_hip_dgemm_hier(const momms_hier_layout* LA, const double* PA,
                                    const momms_hier_layout* LB, const double* PB,
                                    const double* alpha, const double* beta,
                                    const momms_hier_layout* LC, double* PC, void* stream) {
  static const char* fn = "momms_hip_dgemm_hier";
  int st;
  if ((st = check_layout(fn, 1, LA))) return st;
  if ((st = check_layout(fn, 3, LB))) return st;
  if ((st = check_layout(fn, 7, LC))) return st;
  if (!alpha) return bad_arg(fn, 5, "alpha is null");
  if (!beta) return bad_arg(fn, 6, "beta is null");
  const int64_t m = LC->h, n = LC->w, k = LA->w;
  if (LA->h != m) return bad_arg(fn, 1, "A has a different height from C");
  if (LB->h != k) return bad_arg(fn, 3, "B's height differs from A's width");
  if (LB->w != n) return bad_arg(fn, 3, "B has a different width from C");
  if (m == 0 || n == 0) return 0;
  if (!PA && m * k) return bad_arg(fn, 2, "PA is null");
  if (!PB && k * n) return bad_arg(fn, 4, "PB is null");
  if (!PC) return bad_arg(fn, 8, "PC is null");
  hipStream_t s = (hipStream_t)stream;
  // the operands column-major in scratch slot 1 (the GEMM's own scratch use,
  // and the offset tables, are slot 0)
  double* W = nullptr;
  hipError_t e = stream_scratch(s, (size_t)(m * k + k * n + m * n) * sizeof(double), (void**)&W, 1);
  if (e != hipSuccess) return hip_status(fn, e);
  double *A = W, *B = W + m * k, *C = W + m * k + k * n;
  if ((e = hier_unpack_impl<double, false>(*LA, PA, A, 1, m, s)) != hipSuccess) return hip_status(fn, e);
  if ((e = hier_unpack_impl<double, false>(*LB, PB, B, 1, k > 0 ? k : 1, s)) != hipSuccess)
    return hip_status(fn, e);
  if (*beta != 0.0 && (e = hier_unpack_impl<double, false>(*LC, PC, C, 1, m, s)) != hipSuccess)
    return hip_status(fn, e);
  if ((st = momms_hip_dgemm(m, n, k, alpha, A, 1, m, B, 1, k > 0 ? k : 1, beta, C, 1, m, stream)))
    return st;
  return hip_status(fn, hier_pack_impl<double, false>(*LC, C, 1, m, PC, s));
}
