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
// A workgroup takes 256 consecutive rows of 8 consecutive columns, so the
// strided side is read or written along its unit stride when that is the
// row stride (the column-major case) and the layout side stays within the
// leaves those rows share.
#include <hip/hip_runtime.h>

#include "mfma.hpp"
#include "momms_hip.h"
#include "scratch.hpp"
#include "status.hpp"

namespace momms {

constexpr int kHierCols = 8;  // columns per workgroup

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

// h == 0 or w == 0 (grid-uniform): nothing is read and every place is padding
__global__ void __launch_bounds__(256)
hier_pack_kernel(int64_t h, int64_t w, int64_t h_pad, int64_t w_pad, uint64_t cap,
                 const int64_t* __restrict__ oy, const int64_t* __restrict__ ox,
                 const double* __restrict__ A, int64_t rs, int64_t cs, double* __restrict__ P) {
  const int64_t y = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= h_pad) return;
  const int64_t x0 = (int64_t)blockIdx.y * kHierCols;
  const int64_t py = oy[y];
  double v[kHierCols];
  if (h > 0 && w > 0) {
    // all loads first (A[0] stands in for padding and is replaced by 0),
    // then the stores
    const int64_t yc = y < h ? y : h - 1;
#pragma unroll
    for (int c = 0; c < kHierCols; ++c) {
      const int64_t x = x0 + c < w ? x0 + c : w - 1;
      v[c] = A[yc * rs + x * cs];
    }
#pragma unroll
    for (int c = 0; c < kHierCols; ++c)
      if (y >= h || x0 + c >= w) v[c] = 0.0;
  } else {
#pragma unroll
    for (int c = 0; c < kHierCols; ++c) v[c] = 0.0;
  }
#pragma unroll
  for (int c = 0; c < kHierCols; ++c) {
    if (x0 + c >= w_pad) break;
    const int64_t o = py + ox[x0 + c];
    if ((uint64_t)o < cap) P[o] = v[c];
  }
}

__global__ void __launch_bounds__(256)
hier_unpack_add_kernel(int64_t h, int64_t w, uint64_t cap, const int64_t* __restrict__ oy,
                       const int64_t* __restrict__ ox, const double* __restrict__ P,
                       double* __restrict__ C, int64_t rs, int64_t cs) {
  const int64_t y = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= h) return;
  const int64_t x0 = (int64_t)blockIdx.y * kHierCols;
  const int64_t py = oy[y];
  double p[kHierCols], c0[kHierCols];
#pragma unroll
  for (int c = 0; c < kHierCols; ++c) {
    const int64_t x = x0 + c < w ? x0 + c : w - 1;  // past the edge: re-read column w-1
    const int64_t o = py + ox[x];
    const bool in = (uint64_t)o < cap;
    p[c] = P[in ? o : 0];
    p[c] = in ? p[c] : 0.0;
    c0[c] = C[y * rs + x * cs];
  }
#pragma unroll
  for (int c = 0; c < kHierCols; ++c)
    if (x0 + c < w) C[y * rs + (x0 + c) * cs] = p[c] + c0[c];
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
  if (L->w_pad > 65535 * (int64_t)kHierCols)
    return bad_arg(fn, arg, "layout is wider than 524280 padded columns");
  for (int64_t l = 0; l < L->ny; ++l)
    if (L->y_blksz[l] <= 0) return bad_arg(fn, arg, "layout has a non-positive block size");
  for (int64_t l = 0; l < L->nx; ++l)
    if (L->x_blksz[l] <= 0) return bad_arg(fn, arg, "layout has a non-positive block size");
  return 0;
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
  if (L->capacity == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  int64_t* t = nullptr;
  hipError_t e = hier_launch_tables(*L, s, &t);
  if (e != hipSuccess) return hip_status(fn, e);
  dim3 grid((unsigned)ceil_div(L->h_pad, 256), (unsigned)ceil_div(L->w_pad, kHierCols));
  hipLaunchKernelGGL(hier_pack_kernel, grid, dim3(256), 0, s, L->h, L->w, L->h_pad, L->w_pad,
                     (uint64_t)L->capacity, t,
                     t + L->h_pad, A, rsA, csA, P);
  return hip_status(fn, hipGetLastError());
}

extern "C" int momms_hip_hier_unpack_add(const momms_hier_layout* L, const double* P, double* C,
                                         int64_t rsC, int64_t csC, void* stream) {
  static const char* fn = "momms_hip_hier_unpack_add";
  int st;
  if ((st = check_layout(fn, 1, L))) return st;
  if (!P && L->capacity) return bad_arg(fn, 2, "P is null");
  if ((st = check_operand(fn, 3, C, L->h, L->w, rsC, csC))) return st;
  if (L->h == 0 || L->w == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  int64_t* t = nullptr;
  hipError_t e = hier_launch_tables(*L, s, &t);
  if (e != hipSuccess) return hip_status(fn, e);
  dim3 grid((unsigned)ceil_div(L->h, 256), (unsigned)ceil_div(L->w, kHierCols));
  hipLaunchKernelGGL(hier_unpack_add_kernel, grid, dim3(256), 0, s, L->h, L->w,
                     (uint64_t)L->capacity, t, t + L->h_pad,
                     P, C, rsC, csC);
  return hip_status(fn, hipGetLastError());
}
