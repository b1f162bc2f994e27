"""The hierarchical (multi-level blocked) layout on the CPU: the oracle's
restatement of Hierarch (src/matrix/hierarch.rs:50-184) pinned by closed forms
of the packed formats the reference's experiments build, by the offset map
being a bijection onto [0, capacity), and by pack -> unpack round trips; the
library's host-side layout builder against the oracle; argument checks of the
device entry points that return before any GPU work."""
import ctypes
import random

import numpy as np
import pytest

from momms_amd import capi
import pyoracle as po

M, N, K = po.M, po.N, po.K

# Control trees of src/exper/l3.rs:80-108, as hierarchy_description lists them
# (innermost first): KernelNM<NR=12, MR=4> gives [M4, N12] (kernel_nm.rs:112-117),
# each Part* appends its block (part.rs:37-41, :74-78, :115-119).
GOTO = [(M, 4), (N, 12), (M, 120), (K, 192), (N, 3000)]
L3B = [(M, 4), (N, 12), (K, 192), (M, 120), (K, 768), (N, 768)]


def rup(v, b):
    return max(1, -(-v // b)) * b


def grid(hp, wp):
    return np.meshgrid(np.arange(hp), np.arange(wp), indexing="ij")


# Closed forms of the BLIS/Goto packed buffers (written from the packing
# scheme itself, not from the level walk): A is k-panels of KC columns, each
# holding MC x KC blocks of MR-row micro-panels stored column by column; B is
# NC-column panels of KC x NC blocks of NR-column micro-panels stored row by
# row; C is MR x NR micro-tiles (row-major) ordered N-panel, M-block,
# NR-panel, MR-panel.
def goto_a(m, k):
    hp, wp = rup(m, 120), rup(k, 192)
    y, x = grid(hp, wp)
    off = (x // 192) * hp * 192 + (y // 120) * 192 * 120 + ((y % 120) // 4) * 192 * 4 \
        + (y % 4) + (x % 192) * 4
    return hp, wp, off


def goto_b(k, n):
    hp, wp = rup(k, 192), rup(n, 3000)
    y, x = grid(hp, wp)
    off = (x // 3000) * hp * 3000 + (y // 192) * 3000 * 192 + ((x % 3000) // 12) * 192 * 12 \
        + (y % 192) * 12 + (x % 12)
    return hp, wp, off


def goto_c(m, n):
    hp, wp = rup(m, 120), rup(n, 3000)
    y, x = grid(hp, wp)
    off = (x // 3000) * hp * 3000 + (y // 120) * 3000 * 120 + ((x % 3000) // 12) * 120 * 12 \
        + ((y % 120) // 4) * 12 * 4 + (y % 4) * 12 + (x % 12)
    return hp, wp, off


def l3b_a(m, k):
    hp, wp = rup(m, 120), rup(k, 768)
    y, x = grid(hp, wp)
    off = (x // 768) * hp * 768 + (y // 120) * 768 * 120 + ((x % 768) // 192) * 120 * 192 \
        + ((y % 120) // 4) * 192 * 4 + (y % 4) + (x % 192) * 4
    return hp, wp, off


# (closed form, steps, ydim, xdim, LH, LW, LRS, LCS) of GotoA/B/C and L3bA
# (l3.rs:84-86, :99-101)
FORMS = {
    "goto_a": (goto_a, GOTO, M, K, 4, 192, 1, 4),
    "goto_b": (goto_b, GOTO, K, N, 192, 12, 12, 1),
    "goto_c": (goto_c, GOTO, M, N, 4, 12, 12, 1),
    "l3b_a": (l3b_a, L3B, M, K, 4, 192, 1, 4),
}


@pytest.mark.parametrize("name", sorted(FORMS))
@pytest.mark.parametrize("h,w", [(1, 1), (50, 50), (121, 193), (240, 384), (0, 7)])
def test_oracle_offsets_match_packed_formats(name, h, w):
    form, steps, yd, xd, lh, lw, lrs, lcs = FORMS[name]
    L = po.hier_layout(h, w, steps, yd, xd, lh, lw, lrs, lcs)
    hp, wp, want = form(h, w)
    assert (int(L[2]), int(L[3]), int(L[6])) == (hp, wp, hp * wp)
    np.testing.assert_array_equal(po.hier_offsets(L), want)


def test_oracle_offsets_catch_a_swapped_leaf_stride():
    # the same GotoA layout with the leaf strides exchanged is a different map
    L = po.hier_layout(121, 193, GOTO, M, K, 4, 192, 4, 1)
    assert not np.array_equal(po.hier_offsets(L), goto_a(121, 193)[2])


def random_nested_steps(rng):
    """A random hierarchy whose blocks of each kind nest (outer multiple of inner)."""
    steps = []
    inner = {M: rng.choice([2, 3, 4]), N: rng.choice([2, 3, 5]), K: rng.choice([2, 4])}
    cur = dict(inner)
    for d in (M, N, K):
        steps.append((d, cur[d]))
    for _ in range(rng.randint(1, 4)):
        d = rng.choice([M, N, K])
        cur[d] *= rng.choice([2, 3])
        steps.append((d, cur[d]))
    return steps, inner


def dense_leaf(rng, steps, yd, xd, hp, wp):
    ys = [b for k, b in steps if k == yd]
    xs = [b for k, b in steps if k == xd]
    Y = ys[0] if ys else hp
    X = xs[0] if xs else wp
    return (1, Y) if rng.random() < 0.5 else (X, 1)


@pytest.mark.parametrize("seed", range(12))
def test_oracle_offsets_are_a_bijection(seed):
    rng = random.Random(seed)
    steps, _ = random_nested_steps(rng)
    yd, xd = rng.sample([M, N, K], 2)
    h, w = rng.randint(0, 40), rng.randint(0, 40)
    probe = po.hier_layout(h, w, steps, yd, xd, 3, 5, 1, 1)
    lrs, lcs = dense_leaf(rng, steps, yd, xd, int(probe[2]), int(probe[3]))
    L = po.hier_layout(h, w, steps, yd, xd, 3, 5, lrs, lcs)
    off = po.hier_offsets(L)
    assert off.size == int(L[6])
    np.testing.assert_array_equal(np.sort(off.ravel()), np.arange(off.size))


@pytest.mark.parametrize("seed", range(4))
def test_oracle_pack_unpack_round_trip(seed):
    rng = np.random.default_rng(seed)
    form, steps, yd, xd, lh, lw, lrs, lcs = FORMS[sorted(FORMS)[seed]]
    h, w = int(rng.integers(1, 150)), int(rng.integers(1, 250))
    A = np.asfortranarray(rng.standard_normal((h, w)))
    L = po.hier_layout(h, w, steps, yd, xd, lh, lw, lrs, lcs)
    P = po.hier_pack(L, A)
    hp, wp, off = form(h, w)
    # every element sits at its closed-form place, padding places hold zero
    inside = np.zeros((hp, wp), bool)
    inside[:h, :w] = True
    np.testing.assert_array_equal(P[off[:h, :w]], A)
    assert not P[off[~inside]].any()
    C = np.asfortranarray(rng.standard_normal((h, w)))
    C0 = C.copy()
    po.hier_unpack_add(L, P, C)
    np.testing.assert_array_equal(C, C0 + A)


def test_library_layout_matches_oracle():
    rng = random.Random(5)
    cases = [(h, w, *FORMS[n][1:]) for n in sorted(FORMS) for h, w in [(0, 0), (1, 1), (8192, 8192),
                                                                     (121, 4001)]]
    for seed in range(40):
        steps, _ = random_nested_steps(rng)
        yd, xd = rng.sample([M, N, K], 2)
        h, w = rng.randint(0, 300), rng.randint(0, 300)
        probe = po.hier_layout(h, w, steps, yd, xd, 3, 5, 1, 1)
        lrs, lcs = dense_leaf(rng, steps, yd, xd, int(probe[2]), int(probe[3]))
        cases.append((h, w, steps, yd, xd, 3, 5, lrs, lcs))
    for c in cases:
        assert capi.hier_layout(*c).as_list() == [int(v) for v in po.hier_layout(*c)], c


def test_library_layout_rejects():
    with pytest.raises(ValueError, match="-1"):
        capi.hier_layout(-1, 4, GOTO, M, K, 4, 192, 1, 4)
    with pytest.raises(ValueError, match="-1"):
        capi.hier_layout(4, 4, [(M, 0)], M, K, 4, 192, 1, 4)
    with pytest.raises(ValueError, match="-2"):
        capi.hier_layout(4, 4, [(M, 2 ** (i + 1)) for i in range(9)], M, K, 4, 192, 1, 2)
    with pytest.raises(ValueError, match="-3"):  # 120 does not split into blocks of 7
        capi.hier_layout(4, 4, [(M, 7), (M, 120)], M, K, 4, 192, 1, 7)
    with pytest.raises(ValueError, match="-3"):  # leaf of 4 rows with column stride 6
        capi.hier_layout(4, 4, GOTO, M, K, 4, 192, 1, 6)


def test_device_entry_points_validate_before_gpu_work():
    l = capi.lib()
    L = capi.hier_layout(10, 10, GOTO, M, K, 4, 192, 1, 4)
    A = ctypes.c_void_p(16)
    assert l.momms_hip_hier_pack(None, A, 1, 10, A, None) == -1
    assert l.momms_hip_hier_pack(ctypes.byref(L), None, 1, 10, A, None) == -2
    assert l.momms_hip_hier_pack(ctypes.byref(L), A, 0, 10, A, None) == -3
    assert l.momms_hip_hier_pack(ctypes.byref(L), A, 1, 10, None, None) == -5
    assert l.momms_hip_hier_unpack_add(ctypes.byref(L), None, A, 1, 10, None) == -2
    assert l.momms_hip_hier_unpack_add(ctypes.byref(L), A, A, 1, 0, None) == -5
    bad = capi.HierLayout.from_buffer_copy(L)
    bad.capacity += 1
    assert l.momms_hip_hier_pack(ctypes.byref(bad), A, 1, 10, A, None) == -1
    assert "inconsistent" in capi.last_error()
    # 2^23 x 4096 workgroups: more than one launch holds
    huge = capi.hier_layout(2 ** 31, 32768, [], M, N, 1, 1, 1, 2 ** 31)
    assert l.momms_hip_hier_pack(ctypes.byref(huge), A, 1, 2 ** 31, A, None) == -1
    assert "too large" in capi.last_error()


def test_dgemm_hier_validates_before_gpu_work():
    l = capi.lib()
    LA = capi.hier_layout(10, 7, GOTO, M, K, 4, 192, 1, 4)
    LB = capi.hier_layout(7, 5, GOTO, K, N, 192, 12, 12, 1)
    LC = capi.hier_layout(10, 5, GOTO, M, N, 4, 12, 12, 1)
    bad = capi.hier_layout(9, 5, GOTO, M, N, 4, 12, 12, 1)
    p = ctypes.c_void_p(16)
    one = ctypes.c_double(1.0)
    r = ctypes.byref
    f = l.momms_hip_dgemm_hier
    assert f(None, p, r(LB), p, r(one), r(one), r(LC), p, None) == -1
    assert f(r(LA), p, r(LB), p, None, r(one), r(LC), p, None) == -5
    assert f(r(LA), p, r(LB), p, r(one), None, r(LC), p, None) == -6
    assert f(r(LA), p, r(LB), p, r(one), r(one), r(bad), p, None) == -1
    assert "height" in capi.last_error()
    assert f(r(LA), p, r(LA), p, r(one), r(one), r(LC), p, None) == -3
    assert f(r(LA), None, r(LB), p, r(one), r(one), r(LC), p, None) == -2
    assert f(r(LA), p, r(LB), p, r(one), r(one), r(LC), None, None) == -8
