"""Parity of the FP64 matrix-core GEMM (momms_hip_dgemm) with the oracle.

Bar: componentwise |C_gpu - C_ref| <= 2*gamma_{k+1}*(|alpha||A||B| + |beta||C0|)
against the restated Goto path (both are k-term sums with different rounding
orders, each within gamma_{k+1} of the exact value), and within gamma_{k+1}
of the double-double oracle.
"""
import numpy as np
import pytest
import torch

import pyoracle as O
from conftest import SEED, gamma, uniform_colmajor
from momms_amd import capi

pytestmark = pytest.mark.gpu


def to_dev(x: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(x).cuda()


def layout_view(rng, h, w, layout, device="cuda"):
    """An h x w float64 tensor in the given storage order, U[-1,1)."""
    x = torch.from_numpy(rng.uniform(-1, 1, (h, w)))
    if layout == "col":
        t = torch.empty((w, h), dtype=torch.float64, device=device).T
    elif layout == "row":
        t = torch.empty((h, w), dtype=torch.float64, device=device)
    else:  # general strides, both > 1, with an offset
        big = torch.empty((2 * h + 1, 3 * w + 2), dtype=torch.float64, device=device)
        t = big[1::2, 2::3][:h, :w]
    t.copy_(x)
    return t


def host(t: torch.Tensor) -> np.ndarray:
    return np.asfortranarray(t.detach().cpu().numpy())


def check_against_oracle(A, B, C0, C, alpha, beta, k):
    a, b, c0, c = host(A), host(B), host(C0), host(C)
    ref, mag = O.dd_gemm(a, b, c0, alpha, beta)
    err = np.abs(c - ref)
    tol = gamma(k + 1) * mag
    bad = err > tol
    assert not bad.any(), f"{bad.sum()} entries off; max err/mag {np.max(err / np.maximum(mag, 1e-300))}"
    goto = c0.copy(order="F")
    O.goto_dgemm(a, b, goto, alpha, beta, nthreads=4)
    assert (np.abs(c - goto) <= 2 * tol).all()


def test_config1_1024_cube():
    """m=n=k=1024, column-major U[-1,1), C += A*B."""
    rng = np.random.default_rng(SEED)
    m = n = k = 1024
    A = to_dev(uniform_colmajor(rng, m, k))
    B = to_dev(uniform_colmajor(rng, k, n))
    C0 = to_dev(uniform_colmajor(rng, m, n))
    C = C0.clone()
    capi.dgemm(A, B, C, 1.0, 1.0)
    torch.cuda.synchronize()
    check_against_oracle(A, B, C0, C, 1.0, 1.0, k)


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (5, 7, 3), (129, 65, 17), (200, 300, 129),
                                   (64, 64, 64), (257, 129, 1000), (33, 500, 2)])
@pytest.mark.parametrize("la,lb,lc", [("col", "col", "col"), ("row", "row", "row"),
                                      ("gen", "col", "row"), ("col", "gen", "gen")])
def test_ragged_shapes_and_layouts(m, n, k, la, lb, lc):
    rng = np.random.default_rng(SEED + m * 7 + n * 3 + k)
    A = layout_view(rng, m, k, la)
    B = layout_view(rng, k, n, lb)
    C = layout_view(rng, m, n, lc)
    C0 = C.clone()
    capi.dgemm(A, B, C, 1.0, 1.0)
    torch.cuda.synchronize()
    check_against_oracle(A, B, C0, C, 1.0, 1.0, k)


@pytest.mark.parametrize("tile", list(range(12)))
def test_every_tile_config(tile):
    rng = np.random.default_rng(SEED + tile)
    m, n, k = 300, 250, 99
    A, B, C = (to_dev(uniform_colmajor(rng, *s)) for s in ((m, k), (k, n), (m, n)))
    C0 = C.clone()
    capi.dgemm(A, B, C, -1.5, 0.25, tile_cfg=tile)
    torch.cuda.synchronize()
    check_against_oracle(A, B, C0, C, -1.5, 0.25, k)


def test_beta_zero_does_not_read_c():
    rng = np.random.default_rng(SEED)
    m, n, k = 150, 140, 60
    A, B = to_dev(uniform_colmajor(rng, m, k)), to_dev(uniform_colmajor(rng, k, n))
    C = torch.full((n, m), float("nan"), dtype=torch.float64, device="cuda").T
    capi.dgemm(A, B, C, 2.0, 0.0)
    torch.cuda.synchronize()
    assert torch.isfinite(C).all()
    check_against_oracle(A, B, torch.zeros_like(C), C, 2.0, 0.0, k)


def test_k_zero_scales_c_by_beta():
    rng = np.random.default_rng(SEED)
    A = torch.empty((10, 0), dtype=torch.float64, device="cuda")
    B = torch.empty((0, 12), dtype=torch.float64, device="cuda")
    C0 = to_dev(uniform_colmajor(rng, 10, 12))
    C = C0.clone()
    capi.dgemm(A, B, C, 1.0, 0.5)
    torch.cuda.synchronize()
    assert torch.equal(C, C0 * 0.5)


def test_nan_and_inf_propagate():
    m = n = k = 40
    rng = np.random.default_rng(SEED)
    A, B, C = (to_dev(uniform_colmajor(rng, *s)) for s in ((m, k), (k, n), (m, n)))
    A[3, 5] = float("nan")
    B[7, 9] = float("inf")
    capi.dgemm(A, B, C, 1.0, 1.0)
    torch.cuda.synchronize()
    c = C.cpu()
    assert torch.isnan(c[3, :]).all()
    assert not torch.isfinite(c[:, 9]).all()
    mask = torch.ones_like(c, dtype=torch.bool)
    mask[3, :] = False
    mask[:, 9] = False
    assert torch.isfinite(c[mask]).all()


def test_submatrix_view_leaves_rest_untouched():
    """A Matrix<T> view (src/matrix/general_stride.rs:244-249): only the view
    of C changes."""
    rng = np.random.default_rng(SEED)
    big = to_dev(uniform_colmajor(rng, 300, 300))
    before = big.clone()
    A = to_dev(uniform_colmajor(rng, 100, 50))
    B = to_dev(uniform_colmajor(rng, 50, 120))
    Cv = big[37:137, 11:131]
    C0 = Cv.clone()
    capi.dgemm(A, B, Cv, 1.0, 1.0)
    torch.cuda.synchronize()
    check_against_oracle(A, B, C0, Cv, 1.0, 1.0, 50)
    mask = torch.ones_like(big, dtype=torch.bool)
    mask[37:137, 11:131] = False
    assert torch.equal(big[mask], before[mask])


@pytest.mark.parametrize("m,n,k", [(512, 512, 512), (1024, 1024, 1024), (2048, 2048, 2048),
                                   (4096, 4096, 4096), (8192, 8192, 8192), (8192, 8192, 256),
                                   (8192, 64, 8192), (2048, 2048, 256), (1000, 300, 5000)])
def test_large_sampled_against_dd(m, n, k):
    """Full-size shapes (every shape of the plain FP64 sweep: squares 512 to
    8192, the rank-256 update and the 64-column skinny product): sampled
    entries against the double-double oracle, and the randomized residual of
    src/util.rs:55-85 (a size-independent check)."""
    rng = np.random.default_rng(SEED + m + n + k)
    a, b, c0 = uniform_colmajor(rng, m, k), uniform_colmajor(rng, k, n), uniform_colmajor(rng, m, n)
    A, B, C = to_dev(a), to_dev(b), to_dev(c0)
    capi.dgemm(A, B, C, 1.0, 1.0)
    torch.cuda.synchronize()
    c = host(C)
    rows = rng.integers(0, m, 4096)
    cols = rng.integers(0, n, 4096)
    ref, mag = O.dd_entries(a, b, c0, rows, cols)
    assert (np.abs(c[rows, cols] - ref) <= gamma(k + 1) * mag).all()
    w = rng.uniform(-1, 1, n)
    res = O.residual(a, b, c, c0, 1.0, 1.0, w)
    scale = np.linalg.norm(np.abs(a) @ (np.abs(b) @ np.abs(w)))
    assert np.sqrt(res) <= 4 * gamma(k + n + 2) * scale


@pytest.mark.parametrize("m,n,k", [(1024, 1024, 1024), (300, 250, 4099), (8192, 64, 2048),
                                   (512, 512, 512), (1024, 1024, 512), (256, 256, 2048)])
def test_split_k_is_deterministic_and_accurate(m, n, k):
    """The split-k form (k parts reduced in fixed order): repeat calls give the
    same bits, and the result is within gamma_{k+1} of double-double."""
    rng = np.random.default_rng(SEED + m + k)
    a, b, c0 = (uniform_colmajor(rng, *s) for s in ((m, k), (k, n), (m, n)))
    A, B = to_dev(a), to_dev(b)
    C1, C2 = to_dev(c0), to_dev(c0)
    capi.dgemm(A, B, C1, 0.5, -1.0, tile_cfg=5)
    capi.dgemm(A, B, C2, 0.5, -1.0, tile_cfg=5)
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)
    rows = rng.integers(0, m, 2048)
    cols = rng.integers(0, n, 2048)
    ref, mag = O.dd_entries(a, b, c0, rows, cols, 0.5, -1.0)
    c = host(C1)
    assert (np.abs(c[rows, cols] - ref) <= gamma(k + 1) * mag).all()


@pytest.mark.parametrize("tile", [6, 7, 8, 9])
@pytest.mark.parametrize("m,n,k", [(1024, 1024, 1024), (512, 512, 512), (8192, 64, 8192),
                                   (300, 250, 4099), (129, 65, 1000), (2000, 1500, 777),
                                   (64, 64, 16), (70, 33, 5)])
def test_stream_k_is_deterministic_and_accurate(tile, m, n, k):
    """The persistent stream-K form (tiles split between workgroups, their
    parts added in k order by whichever workgroup arrives last): repeat calls
    give the same bits, and every sampled entry is within gamma_{k+1} of
    double-double. A third call on a C holding NaN with beta = 0 shows that
    C is not read and that every entry is written."""
    rng = np.random.default_rng(SEED + m + 3 * n + 7 * k + tile)
    a, b, c0 = (uniform_colmajor(rng, *s) for s in ((m, k), (k, n), (m, n)))
    A, B = to_dev(a), to_dev(b)
    C1, C2 = to_dev(c0), to_dev(c0)
    capi.dgemm(A, B, C1, 0.5, -1.0, tile_cfg=tile)
    capi.dgemm(A, B, C2, 0.5, -1.0, tile_cfg=tile)
    C3 = torch.full((n, m), float("nan"), dtype=torch.float64, device="cuda").T
    capi.dgemm(A, B, C3, 0.5, 0.0, tile_cfg=tile)
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)
    assert torch.isfinite(C3).all()
    rows = rng.integers(0, m, 2048)
    cols = rng.integers(0, n, 2048)
    ref, mag = O.dd_entries(a, b, c0, rows, cols, 0.5, -1.0)
    c = host(C1)
    assert (np.abs(c[rows, cols] - ref) <= gamma(k + 1) * mag).all()
    ref0, mag0 = O.dd_entries(a, b, c0, rows, cols, 0.5, 0.0)
    assert (np.abs(host(C3)[rows, cols] - ref0) <= gamma(k + 1) * mag0).all()


@pytest.mark.parametrize("tree", sorted(capi.REFERENCE_TREES))
@pytest.mark.parametrize("m,n,k", [(1000, 3100, 600), (1300, 1250, 400), (257, 130, 77)])
def test_tree_order_gives_the_bits_of_the_128_tile(tree, m, n, k):
    """momms_hip_dgemm_tree with the reference's variant trees: the tile order
    changes which panels share an L2, never a tile's arithmetic, so every
    tree gives the bits of tile configuration 0; and within gamma_{k+1} of
    double-double at sampled entries."""
    rng = np.random.default_rng(SEED + m + n + k + len(tree))
    a, b, c0 = (uniform_colmajor(rng, *s) for s in ((m, k), (k, n), (m, n)))
    A, B = to_dev(a), to_dev(b)
    C1, C2 = to_dev(c0), to_dev(c0)
    capi.dgemm_tree(A, B, C1, capi.REFERENCE_TREES[tree], 1.5, -0.5)
    capi.dgemm(A, B, C2, 1.5, -0.5, tile_cfg=0)
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)
    rows, cols = rng.integers(0, m, 2048), rng.integers(0, n, 2048)
    ref, mag = O.dd_entries(a, b, c0, rows, cols, 1.5, -0.5)
    assert (np.abs(host(C1)[rows, cols] - ref) <= gamma(k + 1) * mag).all()


def test_tree_order_on_row_major_operands_and_an_empty_tree():
    rng = np.random.default_rng(SEED + 99)
    m, n, k = 300, 200, 150
    A, B, C = layout_view(rng, m, k, "row"), layout_view(rng, k, n, "col"), layout_view(rng, m, n, "row")
    C0 = C.clone()
    C2 = C.clone()
    capi.dgemm_tree(A, B, C, capi.REFERENCE_TREES["l3c"], 1.0, 1.0)
    capi.dgemm(A, B, C2, 1.0, 1.0, tile_cfg=0)
    torch.cuda.synchronize()
    assert torch.equal(C, C2)
    check_against_oracle(A, B, C0, C, 1.0, 1.0, k)
    A2, B2 = to_dev(uniform_colmajor(rng, m, k)), to_dev(uniform_colmajor(rng, k, n))
    C3, C4 = to_dev(uniform_colmajor(rng, m, n)), None
    C4 = C3.clone()
    capi.dgemm_tree(A2, B2, C3, [], 1.0, 1.0)
    capi.dgemm(A2, B2, C4, 1.0, 1.0, tile_cfg=0)
    torch.cuda.synchronize()
    assert torch.equal(C3, C4)


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (32, 32, 32), (70, 33, 5), (96, 64, 128),
                                   (129, 65, 300), (300, 200, 260), (17, 900, 41)])
@pytest.mark.parametrize("layout", ["col", "row", "gen"])
def test_small_gemm_leaf(m, n, k, layout):
    """momms_hip_dgemm_small (the libxsmm-leaf counterpart): column-major and
    row-major operands take the single-launch wavefront tiles, general strides
    the ordinary path; within gamma_{k+1} of double-double either way, and
    beta = 0 never reads C."""
    rng = np.random.default_rng(SEED + m + 5 * n + 11 * k + len(layout))
    A = layout_view(rng, m, k, layout)
    B = layout_view(rng, k, n, layout)
    C = layout_view(rng, m, n, layout)
    C0 = C.clone()
    capi.dgemm_small(A, B, C, -0.75, 1.25)
    torch.cuda.synchronize()
    check_against_oracle(A, B, C0, C, -0.75, 1.25, k)
    Cz = layout_view(rng, m, n, layout)
    Cz.fill_(float("nan"))
    capi.dgemm_small(A, B, Cz, 1.0, 0.0)
    torch.cuda.synchronize()
    assert torch.isfinite(Cz).all()
    check_against_oracle(A, B, torch.zeros_like(Cz), Cz, 1.0, 0.0, k)


@pytest.mark.parametrize("tile", [10, 11])
@pytest.mark.parametrize("m,n,k", [(512, 512, 512), (1024, 1024, 1024), (333, 77, 1000),
                                   (64, 64, 7), (100, 1, 300), (1, 130, 64), (40, 50, 0)])
def test_k_split_in_workgroup_is_deterministic_and_accurate(tile, m, n, k):
    """The form with k split over the four wavefronts of a workgroup (partial
    tiles added in LDS in wavefront order): ragged edges, k not a multiple of
    the 8-deep chunks or of 4, one row or one column, k = 0. Repeat calls give
    the same bits; sampled entries are within gamma_{k+1} of double-double;
    beta = 0 never reads C."""
    rng = np.random.default_rng(SEED + m + 3 * n + 7 * k + tile)
    a, b, c0 = (uniform_colmajor(rng, *s) for s in ((m, k), (k, n), (m, n)))
    A = to_dev(a) if k else torch.empty((m, 0), dtype=torch.float64, device="cuda")
    B = to_dev(b) if k else torch.empty((0, n), dtype=torch.float64, device="cuda")
    C1, C2 = to_dev(c0), to_dev(c0)
    capi.dgemm(A, B, C1, -0.75, 1.25, tile_cfg=tile)
    capi.dgemm(A, B, C2, -0.75, 1.25, tile_cfg=tile)
    C3 = torch.full((n, m), float("nan"), dtype=torch.float64, device="cuda").T
    capi.dgemm(A, B, C3, -0.75, 0.0, tile_cfg=tile)
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)
    assert torch.isfinite(C3).all()
    rows = rng.integers(0, m, 2048)
    cols = rng.integers(0, n, 2048)
    ref, mag = O.dd_entries(a, b, c0, rows, cols, -0.75, 1.25)
    assert (np.abs(host(C1)[rows, cols] - ref) <= gamma(k + 1) * mag).all()
    ref0, mag0 = O.dd_entries(a, b, c0, rows, cols, -0.75, 0.0)
    assert (np.abs(host(C3)[rows, cols] - ref0) <= gamma(k + 1) * mag0).all()

