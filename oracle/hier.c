/*
 * ORACLE — test infrastructure only (see oracle.h).
 *
 * The hierarchical matrix layout of momms (`Hierarch`, src/matrix/hierarch.rs),
 * restated:
 *   oracle_hier_layout  Hierarch::new (hierarch.rs:90-140): padded size from
 *                       the top-level block of each dimension
 *                       (get_top_level_dim_size, :50-58), then the level
 *                       strides (parse_input_hierarchy, :59-89)
 *   oracle_hier_offset  get_offset_y / get_offset_x (hierarch.rs:153-184);
 *                       oracle_hier_offsets: the same over the padded domain
 *   oracle_hier_pack    Packer<Matrix, Hierarch>::pack (pack.rs:153-309):
 *                       every element of the matrix to its place; padding is
 *                       zero (the reference zeroes the padding of the leaves it
 *                       visits, pack.rs:186-195, and leaves whole padding leaves
 *                       unwritten)
 *   oracle_hier_unpack_add  Unpacker<Matrix, Hierarch>::add (unpack.rs:61-98,
 *                       :174-213): C[y][x] += P[offset(y, x)] over the matrix
 * Pinned by tests/test_oracle.py: closed forms of the layouts the reference's
 * experiments build (Goto A/B/C panels, l3.rs:56-58, :84-86), the offset map
 * being a bijection onto [0, capacity) over the padded domain, and round trips.
 *
 * A hierarchy step is (kind, bsz) with kind 0 = M, 1 = N, 2 = K, listed
 * innermost first as a control tree's hierarchy_description() returns them.
 */
#include <stdint.h>
#include <string.h>

#include "oracle.h"

/* bsz of the outermost step of dimension `dim` (hierarch.rs:50-58), or 0 */
static int64_t top_level(const int32_t* kind, const int64_t* bsz, int64_t n, int32_t dim) {
  for (int64_t i = n - 1; i >= 0; --i)
    if (kind[i] == dim) return bsz[i];
  return 0;
}

int oracle_hier_layout(int64_t h, int64_t w, const int32_t* kind, const int64_t* bsz,
                       int64_t nsteps, int32_t ydim, int32_t xdim, int64_t lh, int64_t lw,
                       int64_t lrs, int64_t lcs, int64_t* out) {
  /* out: [h, w, h_pad, w_pad, ny, nx, capacity, lrs, lcs,
           y_blksz[ORACLE_HIER_LEVELS], y_stride[..], x_blksz[..], x_stride[..]] */
  int64_t yt = top_level(kind, bsz, nsteps, ydim), xt = top_level(kind, bsz, nsteps, xdim);
  if (yt == 0) yt = lh;
  if (xt == 0) xt = lw;
  if (yt <= 0 || xt <= 0) return -1;
  const int64_t nby = h == 0 ? 1 : (h - 1) / yt + 1, nbx = w == 0 ? 1 : (w - 1) / xt + 1;
  const int64_t hp = nby * yt, wp = nbx * xt;
  int64_t* yb = out + 9;
  int64_t* ys = yb + ORACLE_HIER_LEVELS;
  int64_t* xb = ys + ORACLE_HIER_LEVELS;
  int64_t* xs = xb + ORACLE_HIER_LEVELS;
  int64_t ny = 0, nx = 0, ht = hp, wt = wp;
  for (int64_t i = nsteps - 1; i >= 0; --i) {
    if (kind[i] == xdim) {
      if (nx == ORACLE_HIER_LEVELS) return -2;
      xb[nx] = bsz[i];
      xs[nx] = ht * bsz[i];
      ++nx;
      wt = bsz[i];
    }
    if (kind[i] == ydim) {
      if (ny == ORACLE_HIER_LEVELS) return -2;
      yb[ny] = bsz[i];
      ys[ny] = wt * bsz[i];
      ++ny;
      ht = bsz[i];
    }
  }
  out[0] = h;
  out[1] = w;
  out[2] = hp;
  out[3] = wp;
  out[4] = ny;
  out[5] = nx;
  out[6] = hp * wp;
  out[7] = lrs;
  out[8] = lcs;
  return 0;
}

int64_t oracle_hier_offset(const int64_t* L, int64_t y, int64_t x) {
  const int64_t ny = L[4], nx = L[5];
  const int64_t* yb = L + 9;
  const int64_t* ys = yb + ORACLE_HIER_LEVELS;
  const int64_t* xb = ys + ORACLE_HIER_LEVELS;
  const int64_t* xs = xb + ORACLE_HIER_LEVELS;
  int64_t off = 0;
  for (int64_t i = 0; i < ny; ++i) {
    const int64_t b = y / yb[i];
    off += b * ys[i];
    y -= b * yb[i];
  }
  for (int64_t i = 0; i < nx; ++i) {
    const int64_t b = x / xb[i];
    off += b * xs[i];
    x -= b * xb[i];
  }
  return off + y * L[7] + x * L[8];
}

void oracle_hier_offsets(const int64_t* L, int64_t* out) {
  for (int64_t y = 0; y < L[2]; ++y)
    for (int64_t x = 0; x < L[3]; ++x) out[y * L[3] + x] = oracle_hier_offset(L, y, x);
}

void oracle_hier_pack(const int64_t* L, const double* A, int64_t rs, int64_t cs, double* P) {
  const int64_t h = L[0], w = L[1], hp = L[2], wp = L[3];
  memset(P, 0, (size_t)L[6] * sizeof(double));
  for (int64_t x = 0; x < wp; ++x)
    for (int64_t y = 0; y < hp; ++y)
      if (y < h && x < w) P[oracle_hier_offset(L, y, x)] = A[y * rs + x * cs];
}

void oracle_hier_unpack_add(const int64_t* L, const double* P, double* C, int64_t rs,
                            int64_t cs) {
  for (int64_t x = 0; x < L[1]; ++x)
    for (int64_t y = 0; y < L[0]; ++y) C[y * rs + x * cs] = P[oracle_hier_offset(L, y, x)] + C[y * rs + x * cs];
}
