"""Parity of the device-resident hierarchical layout (momms_hip_hier_pack,
momms_hip_hier_unpack_add) with the oracle: bit-exact (pure data movement),
on the reference's Goto / L3b layouts and random nested hierarchies, ragged
and empty sizes, three storage orders; at 8192 x 8192 through the closed form
of the Goto A buffer and a pack -> unpack round trip."""
import random

import numpy as np
import pytest
import torch

import pyoracle as po
from momms_amd import capi
from test_gpu_dgemm import host, layout_view
from test_hier import FORMS, GOTO, dense_leaf, goto_a, random_nested_steps

pytestmark = pytest.mark.gpu
M, K = po.M, po.K


def oracle_layout(L: capi.HierLayout) -> np.ndarray:
    return np.array(L.as_list(), dtype=np.int64)


def check_pack_unpack(L, rng, layout):
    A = layout_view(rng, L.h, L.w, layout)
    P = torch.full((max(L.capacity, 1),), float("nan"), dtype=torch.float64, device="cuda")
    capi.hier_pack(L, A, P)
    C = layout_view(rng, L.h, L.w, layout)
    C0 = host(C)
    capi.hier_unpack_add(L, P, C)
    torch.cuda.synchronize()
    OL = oracle_layout(L)
    want_p = po.hier_pack(OL, host(A))
    np.testing.assert_array_equal(P.cpu().numpy()[:L.capacity], want_p)
    want_c = po.hier_unpack_add(OL, want_p, C0.copy())
    np.testing.assert_array_equal(host(C), want_c)


@pytest.mark.parametrize("name", sorted(FORMS))
@pytest.mark.parametrize("h,w", [(1, 1), (121, 193), (250, 130), (0, 9), (7, 0)])
@pytest.mark.parametrize("layout", ["col", "row", "strided"])
def test_pack_unpack_match_oracle(name, h, w, layout):
    _, steps, yd, xd, lh, lw, lrs, lcs = FORMS[name]
    L = capi.hier_layout(h, w, steps, yd, xd, lh, lw, lrs, lcs)
    print(f"hier {name} {h}x{w} {layout}", flush=True)
    check_pack_unpack(L, np.random.default_rng(h * 1000 + w), layout)


@pytest.mark.parametrize("seed", range(10))
def test_random_nested_layouts(seed):
    rng = random.Random(100 + seed)
    steps, _ = random_nested_steps(rng)
    yd, xd = rng.sample([0, 1, 2], 2)
    h, w = rng.randint(0, 600), rng.randint(0, 600)
    hp = po.hier_layout(h, w, steps, yd, xd, 3, 5, 1, 1)
    lrs, lcs = dense_leaf(rng, steps, yd, xd, int(hp[2]), int(hp[3]))
    L = capi.hier_layout(h, w, steps, yd, xd, 3, 5, lrs, lcs)
    print(f"hier random seed {seed}: {h}x{w} steps {steps}", flush=True)
    check_pack_unpack(L, np.random.default_rng(seed), ["col", "row", "strided"][seed % 3])


def test_goto_a_8192_closed_form_and_round_trip():
    n = 8192
    L = capi.hier_layout(n, n, GOTO, M, K, 4, 192, 1, 4)
    print(f"hier goto_a {n}x{n}: capacity {L.capacity}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(7)
    A = torch.empty((n, n), dtype=torch.float64, device="cuda").T  # column-major
    A.uniform_(-1, 1, generator=g)
    P = torch.full((L.capacity,), float("nan"), dtype=torch.float64, device="cuda")
    capi.hier_pack(L, A, P)
    # every element at its closed-form place (GotoA, computed on the GPU)
    y = torch.arange(L.h_pad, device="cuda").view(-1, 1)
    x = torch.arange(L.w_pad, device="cuda").view(1, -1)
    off = (x // 192) * L.h_pad * 192 + (y // 120) * 192 * 120 + ((y % 120) // 4) * 192 * 4 \
        + (y % 4) + (x % 192) * 4
    assert torch.equal(P[off[:n, :n]], A)
    pad = torch.ones((L.h_pad, L.w_pad), dtype=torch.bool, device="cuda")
    pad[:n, :n] = False
    assert int(pad.sum()) == L.capacity - n * n
    assert bool((P[off[pad]] == 0).all())
    # pack -> unpack into zeros gives A back; unpacking twice adds it twice
    C = torch.zeros((n, n), dtype=torch.float64, device="cuda").T
    capi.hier_unpack_add(L, P, C)
    assert torch.equal(C, A)
    capi.hier_unpack_add(L, P, C)
    assert torch.equal(C, 2 * A)


def test_pack_small_matches_closed_form_on_a_row_major_source():
    L = capi.hier_layout(300, 500, GOTO, M, K, 4, 192, 1, 4)
    A = torch.randn((300, 500), dtype=torch.float64, device="cuda")
    P = torch.empty((L.capacity,), dtype=torch.float64, device="cuda")
    capi.hier_pack(L, A, P)
    hp, wp, off = goto_a(300, 500)
    Pn = P.cpu().numpy()
    np.testing.assert_array_equal(Pn[off[:300, :500]], A.cpu().numpy())


def colmajor(rng, h, w):
    t = torch.empty((w, h), dtype=torch.float64, device="cuda").T
    t.copy_(torch.from_numpy(rng.uniform(-1, 1, (h, w))))
    return t


GOTO_ABC = [(M, K, 4, 192, 1, 4), (K, po.N, 192, 12, 12, 1), (M, po.N, 4, 12, 12, 1)]


@pytest.mark.parametrize("m,n,k", [(250, 130, 77), (1000, 700, 300), (64, 3001, 193), (5, 1, 1),
                                   (120, 12, 0)])
@pytest.mark.parametrize("alpha,beta", [(1.0, 1.0), (-0.5, 2.0), (1.0, 0.0)])
def test_dgemm_hier_equals_dgemm_on_the_same_matrices(m, n, k, alpha, beta):
    """The GEMM on Goto A/B/C hierarchical buffers (l3.rs:56-58) gives the bits
    of momms_hip_dgemm on the column-major matrices, and leaves C's padding
    places as they were."""
    print(f"dgemm_hier {m}x{n}x{k} alpha={alpha} beta={beta}", flush=True)
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    dims = [(m, k), (k, n), (m, n)]
    Ls = [capi.hier_layout(h, w, GOTO, *spec) for (h, w), spec in zip(dims, GOTO_ABC)]
    A, B, C = (colmajor(rng, h, w) for h, w in dims)
    Ps = [torch.zeros(max(L.capacity, 1), dtype=torch.float64, device="cuda") for L in Ls]
    for L, X, P in zip(Ls, (A, B, C), Ps):
        capi.hier_pack(L, X, P)
    # C's padding places hold NaN; with beta == 0 its own places do too
    offC = torch.from_numpy(po.hier_offsets(oracle_layout(Ls[2]))).cuda()
    pad = torch.ones(offC.shape, dtype=torch.bool, device="cuda")
    pad[:m, :n] = False
    Ps[2][offC[pad]] = float("nan")
    if beta == 0.0:
        Ps[2][offC[:m, :n].reshape(-1)] = float("nan")
    capi.dgemm_hier(Ls[0], Ps[0], Ls[1], Ps[1], Ls[2], Ps[2], alpha, beta)
    Cref = C.clone()
    capi.dgemm(A, B, Cref, alpha, beta)
    got = Ps[2][offC[:m, :n].reshape(-1)].view(m, n)
    torch.cuda.synchronize()
    assert torch.equal(got, Cref)
    assert bool(torch.isnan(Ps[2][offC[pad]]).all())


def test_dgemm_hier_goto_8192_matches_dgemm():
    n = 8192
    print(f"dgemm_hier goto {n}^3", flush=True)
    g = torch.Generator(device="cuda").manual_seed(11)
    X = [torch.empty((n, n), dtype=torch.float64, device="cuda").T.uniform_(-1, 1, generator=g)
         for _ in range(3)]
    Ls = [capi.hier_layout(n, n, GOTO, *spec) for spec in GOTO_ABC]
    Ps = [torch.empty(L.capacity, dtype=torch.float64, device="cuda") for L in Ls]
    for L, x, P in zip(Ls, X, Ps):
        capi.hier_pack(L, x, P)
    capi.dgemm_hier(Ls[0], Ps[0], Ls[1], Ps[1], Ls[2], Ps[2], 1.0, 1.0)
    capi.dgemm(X[0], X[1], X[2], 1.0, 1.0)
    back = torch.zeros((n, n), dtype=torch.float64, device="cuda").T
    capi.hier_unpack_add(Ls[2], Ps[2], back)
    assert torch.equal(back, X[2])


@pytest.mark.parametrize("name", sorted(FORMS))
@pytest.mark.parametrize("layout", ["col", "row", "strided"])
def test_f32_pack_unpack_match_oracle(name, layout):
    """The float32 forms (Hierarch<f32>, the KNM tree's operands): the same
    places as the oracle's layout, values moved bit for bit; the copy form of
    the unpack gives the source back."""
    _, steps, yd, xd, lh, lw, lrs, lcs = FORMS[name]
    h, w = 203, 157
    L = capi.hier_layout(h, w, steps, yd, xd, lh, lw, lrs, lcs)
    rng = np.random.default_rng(h + w + len(layout))
    A = layout_view(rng, h, w, layout).float()
    if layout == "row":
        A = A.contiguous()
    elif layout == "col":
        A = A.T.contiguous().T
    P = torch.full((L.capacity,), float("nan"), dtype=torch.float32, device="cuda")
    capi.hier_pack(L, A, P)
    C = torch.rand((w, h), dtype=torch.float32, device="cuda").T
    C0 = C.clone()
    capi.hier_unpack_add(L, P, C)
    R = torch.empty((w, h), dtype=torch.float32, device="cuda").T
    capi.hier_unpack_f32(L, P, R)
    torch.cuda.synchronize()
    OL = oracle_layout(L)
    want_p = po.hier_pack(OL, host(A).astype(np.float64)).astype(np.float32)
    np.testing.assert_array_equal(P.cpu().numpy(), want_p)
    assert torch.equal(R, A)
    assert torch.equal(C, C0 + A)
