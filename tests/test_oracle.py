"""Pins the CPU oracle before anything is compared with it.

Each oracle function is checked against something that does not share its
code: exact-rational known answers (tests/golden), closed forms, invariants,
and brute force. A planted-mistake test shows the checks have teeth.
"""
import math

import numpy as np
import pytest

import pyoracle as O
from conftest import SEED, gamma, unhex, uniform_colmajor


# --- micro-kernel (ukernels/bli_gemm_haswell_asm_d6x8.c:666-1221) -----------

def test_ukernel_matches_exact_fma_chain(golden):
    for case in golden["ukernel"]:
        k = case["k"]
        a6 = unhex(case["a"], 6, k)     # a[i][p]
        b8 = unhex(case["b"], k, 8)     # b[p][j]
        # packed panels: MR (=6) contiguous per k for A, NR (=8) per k for B
        apack = np.ascontiguousarray(a6.T).ravel() if k else np.zeros(8)
        bpack = np.ascontiguousarray(b8).ravel() if k else np.zeros(8)
        c = unhex(case["c"], 6, 8)
        out = unhex(case["out"], 6, 8)
        cc = np.asfortranarray(c.copy())
        O.ukernel_6x8(k, float.fromhex(case["alpha"]), apack, bpack, float.fromhex(case["beta"]),
                      cc, 1, 6)
        assert np.array_equal(cc.view(np.int64), out.view(np.int64)), f"k={k}"


def test_ukernel_beta_zero_never_reads_c():
    rng = np.random.default_rng(SEED)
    k = 7
    a = rng.uniform(-1, 1, 6 * k)
    b = rng.uniform(-1, 1, 8 * k)
    c = np.full((6, 8), np.nan, order="F")
    O.ukernel_6x8(k, 1.0, a, b, 0.0, c, 1, 6)
    assert np.isfinite(c).all()


# --- Goto tree ---------------------------------------------------------------

def test_goto_matches_golden_rounding_order(golden):
    for case in golden["goto"]:
        m, n, k = case["m"], case["n"], case["k"]
        A, B, C = unhex(case["A"], m, k), unhex(case["B"], k, n), unhex(case["C"], m, n)
        out = unhex(case["out"], m, n)
        al, be = float.fromhex(case["alpha"]), float.fromhex(case["beta"])
        for nt in (1, 3):
            c1 = C.copy(order="F")
            O.goto_dgemm(A, B, c1, al, be, nthreads=nt)
            assert np.array_equal(c1.view(np.int64), out.view(np.int64)), (m, n, k, nt)
        c2 = C.copy(order="F")
        O.goto_scalar(A, B, c2, al, be)
        assert np.array_equal(c2.view(np.int64), out.view(np.int64)), (m, n, k)


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (121, 250, 193), (250, 3001, 20), (6, 8, 384),
                                   (7, 13, 0)])
@pytest.mark.parametrize("layout", ["col", "row", "gen"])
def test_goto_equals_scalar_restatement(m, n, k, layout):
    rng = np.random.default_rng(SEED + m + n + k)

    def mk(h, w):
        x = rng.uniform(-1, 1, (h, w))
        if layout == "col":
            return np.asfortranarray(x)
        if layout == "row":
            return np.ascontiguousarray(x)
        big = np.zeros((2 * h + 1, 3 * w + 2))
        v = big[1::2, 2::3][:h, :w]
        v[...] = x
        return v

    A, B, C = mk(m, k), mk(k, n), mk(m, n)
    c1 = C.copy()
    c1 = mk(m, n)
    c1[...] = C
    O.goto_dgemm(A, B, c1, 1.0, 1.0, nthreads=4)
    c2 = np.asfortranarray(C.copy())
    O.goto_scalar(np.asfortranarray(A), np.asfortranarray(B), c2, 1.0, 1.0)
    assert np.array_equal(np.asarray(c1).view(np.int64), c2.view(np.int64))


def test_goto_within_gamma_of_dd_oracle():
    rng = np.random.default_rng(SEED)
    m, n, k = 130, 70, 400
    A, B, C = (uniform_colmajor(rng, *s) for s in ((m, k), (k, n), (m, n)))
    ref, mag = O.dd_gemm(A, B, C)
    c = C.copy(order="F")
    O.goto_dgemm(A, B, c, 1.0, 1.0, nthreads=2)
    err = np.abs(c - ref)
    assert (err <= gamma(k + 1) * mag + 1e-300).all()
    # and it is not trivially exact: the rounding order shows up
    assert err.max() > 0


# --- TripleLoop --------------------------------------------------------------

def test_triple_loop_golden(golden):
    case = golden["triple"][0]
    m, n, k = case["m"], case["n"], case["k"]
    A, B, C = unhex(case["A"], m, k), unhex(case["B"], k, n), unhex(case["C"], m, n)
    O.triple_loop(A, B, C)
    out = unhex(case["out"], m, n)
    assert np.array_equal(C.view(np.int64), out.view(np.int64))


# --- double-double oracle ----------------------------------------------------

def test_dd_within_faithful_bound_of_exact(golden):
    u = 2.0 ** -53
    for case in golden["exact"]:
        m, n, k = case["m"], case["n"], case["k"]
        A, B, C = unhex(case["A"], m, k), unhex(case["B"], k, n), unhex(case["C"], m, n)
        al, be = float.fromhex(case["alpha"]), float.fromhex(case["beta"])
        exact = unhex(case["out"], m, n)
        got, mag = O.dd_gemm(A, B, C, al, be)
        # Dot2: |res - x'y| <= u|x'y| + gamma_n^2 |x|'|y|
        bound = 2 * u * np.abs(exact) + gamma(k + 2) ** 2 * mag
        assert (np.abs(got - exact) <= bound).all()


def test_dd_exact_on_integer_inputs():
    # Products and sums of small integers are exact in doubles: closed form.
    rng = np.random.default_rng(SEED)
    A = np.asfortranarray(rng.integers(-1000, 1000, (9, 50)).astype(np.float64))
    B = np.asfortranarray(rng.integers(-1000, 1000, (50, 5)).astype(np.float64))
    C = np.asfortranarray(rng.integers(-1000, 1000, (9, 5)).astype(np.float64))
    got, mag = O.dd_gemm(A, B, C, 1.0, 1.0)
    exact = (A.astype(np.int64) @ B.astype(np.int64) + C.astype(np.int64)).astype(np.float64)
    assert np.array_equal(got, exact)
    assert np.array_equal(mag, (np.abs(A) @ np.abs(B) + np.abs(C)))


def test_dd_recovers_cancellation_plain_sum_loses():
    # x'y = 1 exactly, but the naive sum is 0: (1e20 - 1e20 + 1).
    A = np.asfortranarray(np.array([[1e20, 1.0, -1e20]]))
    B = np.asfortranarray(np.array([[1.0], [1.0], [1.0]]))
    C = np.zeros((1, 1), order="F")
    got, _ = O.dd_gemm(A, B, C, 1.0, 0.0)
    assert got[0, 0] == 1.0


def test_dd_entries_match_full():
    rng = np.random.default_rng(SEED)
    A, B, C = (uniform_colmajor(rng, *s) for s in ((40, 30), (30, 20), (40, 20)))
    full, magf = O.dd_gemm(A, B, C, 1.5, -0.5)
    rows = rng.integers(0, 40, 100)
    cols = rng.integers(0, 20, 100)
    e, mag = O.dd_entries(A, B, C, rows, cols, 1.5, -0.5)
    assert np.array_equal(e, full[rows, cols])
    assert np.array_equal(mag, magf[rows, cols])


# --- residual check (src/util.rs:55-85) -------------------------------------

def test_residual_zero_for_exact_product_and_catches_planted_error():
    rng = np.random.default_rng(SEED)
    A = np.asfortranarray(rng.integers(-8, 8, (20, 15)).astype(np.float64))
    B = np.asfortranarray(rng.integers(-8, 8, (15, 10)).astype(np.float64))
    C0 = np.asfortranarray(rng.integers(-8, 8, (20, 10)).astype(np.float64))
    C = np.asfortranarray(A @ B + C0)
    w = rng.integers(-4, 4, 10).astype(np.float64)
    assert O.residual(A, B, C, C0, 1.0, 1.0, w) == 0.0
    C[3, 4] += 2.0  # plant a mistake: residual = (2 * w_4)^2
    assert O.residual(A, B, C, C0, 1.0, 1.0, w) == (2.0 * w[4]) ** 2


# --- split-BF16 restatement --------------------------------------------------

def test_ozaki_golden(golden):
    for case in golden["ozaki"]:
        m, n, k, s = case["m"], case["n"], case["k"], case["s"]
        A, B, C = unhex(case["A"], m, k), unhex(case["B"], k, n), unhex(case["C"], m, n)
        al, be = float.fromhex(case["alpha"]), float.fromhex(case["beta"])
        O.ozaki_gemm(A, B, C, s, al, be)
        out = unhex(case["out"], m, n)
        assert np.array_equal(C.view(np.int64), out.view(np.int64)), (m, n, k, s)


@pytest.mark.parametrize("s", [1, 3, 7])
@pytest.mark.parametrize("e", [0, 16])
def test_ozaki_slices_reconstruct_exactly(s, e):
    rng = np.random.default_rng(SEED + s + e)
    A = np.asfortranarray(rng.uniform(-1, 1, (17, 33)) * np.exp2(rng.integers(-e, e + 1, (17, 33))))
    q, E = O.ozaki_split_rows(A, s)
    assert (q == np.rint(q)).all()
    assert np.abs(q[:, 0, :]).max() <= 256 and (s == 1 or np.abs(q[:, 1:, :]).max() <= 128)
    scale = np.exp2(-8.0 * np.arange(1, s + 1))
    recon = np.einsum("itp,t->ip", q, scale)      # partial sums are exact multiples
    scaled = A * np.exp2(-E.astype(np.float64))[:, None]
    resid = scaled - recon
    assert (np.abs(resid) <= 2.0 ** (-8 * s - 1)).all()
    # exactness: the residual is representable, so recon + resid == scaled bitwise
    assert np.array_equal(recon + resid, scaled)
    # every row max lies in [1/2, 1) after scaling
    mx = np.abs(scaled).max(axis=1)
    assert ((mx >= 0.5) & (mx < 1.0)).all()


@pytest.mark.parametrize("s", [1, 2, 3, 4, 5, 6, 7, 8])
def test_ozaki_chunk_keeps_fp32_sums_exact(s):
    kch = O.ozaki_chunk(s)
    assert kch % 16 == 0 and kch <= 256
    worst = 2 ** 16 if s == 1 else (s + 2) * 2 ** 14
    assert kch * max(worst, 2 ** 16 if s >= 1 else 0) <= 2 ** 24 or kch * worst <= 2 ** 24
    assert kch * worst <= 2 ** 24
    assert kch * 2 ** 16 <= 2 ** 24


@pytest.mark.parametrize("s,bound", [(3, 2.0 ** -20), (5, 2.0 ** -36), (7, 2.0 ** -50)])
def test_ozaki_error_shrinks_with_slices(s, bound):
    rng = np.random.default_rng(SEED + s)
    m, n, k = 24, 20, 300
    A, B, C = (uniform_colmajor(rng, *sh) for sh in ((m, k), (k, n), (m, n)))
    ref, mag = O.dd_gemm(A, B, C)
    c = C.copy(order="F")
    O.ozaki_gemm(A, B, c, s)
    err = np.abs(c - ref) / mag
    assert err.max() <= bound
    if s < 7:
        assert err.max() > bound / 2 ** 24   # and truncation is visible


def test_planted_oracle_mistake_is_caught(golden):
    """A dropped term in a restated product must fail the golden comparison."""
    case = golden["goto"][0]
    m, n, k = case["m"], case["n"], case["k"]
    A, B, C = unhex(case["A"], m, k), unhex(case["B"], k, n), unhex(case["C"], m, n)
    A2 = A.copy(order="F")
    A2[:, -1] = 0.0                     # drop the last rank-1 term
    O.goto_scalar(A2, B, C, float.fromhex(case["alpha"]), float.fromhex(case["beta"]))
    out = unhex(case["out"], m, n)
    assert not np.array_equal(C.view(np.int64), out.view(np.int64))


# --- Chinese-remainder split-BF16 restatement ---------------------------------

def _balanced(x, p):
    r = x % p
    return r - p if r > (p - 1) // 2 else r


def test_crt_golden(golden):
    for case in golden["crt"]:
        m, n, k, N = case["m"], case["n"], case["k"], case["N"]
        A, B, C = unhex(case["A"], m, k), unhex(case["B"], k, n), unhex(case["C"], m, n)
        al, be = float.fromhex(case["alpha"]), float.fromhex(case["beta"])
        O.crt_gemm(A, B, C, N, al, be)
        out = unhex(case["out"], m, n)
        assert np.array_equal(C.view(np.int64), out.view(np.int64)), (m, n, k, N)


def test_crt_moduli_are_the_largest_coprime_odd_set():
    mods = O.crt_moduli(O.CRT_MAX_MODULI)
    want, p = [], 511
    while len(want) < O.CRT_MAX_MODULI:       # greedy from 511 down
        if p % 2 and all(math.gcd(p, q) == 1 for q in want):
            want.append(p)
        p -= 1
    assert mods == want
    # balanced residues |r| <= 255 are integers bf16 holds exactly, and a
    # 256-long chunk of their products (+ a reduced start <= 0.51 p) stays < 2^24
    assert max(mods) <= 511 and 256 * 255 ** 2 + 0.51 * 511 < 2 ** 24


@pytest.mark.parametrize("k", [0, 1, 2, 3, 300, 4096, 4097, 8192, 1 << 20, 1 << 24])
def test_crt_beta_fits_the_product_of_moduli(k):
    for N in range(1, O.CRT_MAX_MODULI + 1):
        b = O.crt_beta(k, N)
        P = math.prod(O.crt_moduli(N))
        kb = (k - 1).bit_length() if k > 1 else 0
        if b < 1:   # even beta = 1 would not fit: too few moduli for this k
            assert 2 * 2 ** kb * 2 ** 2 > 2 ** (P.bit_length() - 2)
            continue
        # |C'| <= k 2^(2b) must lie strictly inside the balanced range and in 2^125
        assert 2 * max(k, 1) * 2 ** (2 * b) < P
        assert max(k, 1) * 2 ** (2 * b) <= 2 ** 125
        # and b is the largest such (or capped at 53)
        if b < 53:
            assert 2 * 2 ** kb * 2 ** (2 * (b + 1)) >= 2 ** (P.bit_length() - 1) or \
                2 ** kb * 2 ** (2 * (b + 1)) > 2 ** 125


@pytest.mark.parametrize("e", [0, 16, 300])
def test_crt_integers_are_exponent_aligned_roundings(e):
    from fractions import Fraction as Fr
    rng = np.random.default_rng(SEED + e)
    A = np.asfortranarray(rng.uniform(-1, 1, (9, 40)) * np.exp2(rng.integers(-e, e + 1, (9, 40))))
    A[3, :] = 0.0
    b = 53
    ia, E = O.crt_integers_rows(A, b)
    assert E[3] == 0 and (ia[3] == 0).all()
    for i in range(9):
        if i == 3:
            continue
        mx = np.abs(A[i]).max()
        assert 2.0 ** (E[i] - 1) <= mx < 2.0 ** E[i]
        assert np.abs(ia[i]).max() <= 2 ** b and np.abs(ia[i]).max() >= 2 ** (b - 1)
        for p in range(40):
            assert ia[i, p] == round(Fr(float(A[i, p])) * Fr(2) ** (b - int(E[i])))


def test_crt_residues_recover_the_integer():
    """The device's reconstruction: balanced residues -> balanced mixed-radix
    digits (Garner) -> value. Covers (-P/2, P/2) at N = 2 (every 7th integer)
    and its ends and random members at N = 14..16."""
    mods = O.crt_moduli(2)
    P = mods[0] * mods[1]

    def recover(x, mods):
        r = [_balanced(x, p) for p in mods]
        v = []
        for l, p in enumerate(mods):
            t = r[l]
            for j in range(l):
                t = _balanced((t - v[j]) * pow(mods[j], -1, p), p)
            v.append(t)
        y = 0
        for l in range(len(mods) - 1, -1, -1):
            y = y * mods[l] + v[l]
        return y

    for x in range(-(P // 2), P // 2 + 1, 7):
        assert recover(x, mods) == x
    import random
    gen = random.Random(SEED)
    for N in (14, 15, 16):
        mods = O.crt_moduli(N)
        half = math.prod(mods) // 2
        for x in [half, -half] + [gen.randrange(-half, half + 1) for _ in range(300)]:
            assert recover(x, mods) == x


@pytest.mark.parametrize("N,bound", [(8, 2.0 ** -18), (11, 2.0 ** -33), (14, 2.0 ** -51)])
def test_crt_error_against_double_double(N, bound):
    rng = np.random.default_rng(SEED + N)
    m, n, k = 24, 20, 300
    A, B, C = (uniform_colmajor(rng, *sh) for sh in ((m, k), (k, n), (m, n)))
    ref, mag = O.dd_gemm(A, B, C)
    c = C.copy(order="F")
    O.crt_gemm(A, B, c, N)
    err = np.abs(c - ref) / mag
    assert err.max() <= bound
    if N < 14:
        assert err.max() > bound / 2 ** 20   # the truncation at beta bits is visible


def test_crt_planted_mistake_is_caught(golden):
    """One modulus too few changes beta and must fail the golden comparison."""
    case = golden["crt"][0]
    m, n, k, N = case["m"], case["n"], case["k"], case["N"]
    A, B, C = unhex(case["A"], m, k), unhex(case["B"], k, n), unhex(case["C"], m, n)
    O.crt_gemm(A, B, C, N - 2, float.fromhex(case["alpha"]), float.fromhex(case["beta"]))
    out = unhex(case["out"], m, n)
    assert not np.array_equal(C.view(np.int64), out.view(np.int64))


# --- the INT8 residue form -----------------------------------------------------

def test_crt_i8_moduli_are_the_largest_coprime_odd_set_below_256():
    mods = O.crt_i8_moduli(O.CRT_MAX_MODULI)
    want, p = [], 255
    while len(want) < O.CRT_MAX_MODULI:       # greedy from 255 down
        if p % 2 and all(math.gcd(p, q) == 1 for q in want):
            want.append(p)
        p -= 1
    assert mods == want
    # balanced residues |r| <= 127 are int8 values, and a k-term sum of their
    # products stays an exact INT32 for every k the form accepts
    assert max(mods) <= 255 and 131072 * 127 ** 2 < 2 ** 31
    assert O.crt_i8_beta(131073, 16) == -1 and O.crt_i8_beta(131072, 16) > 0


@pytest.mark.parametrize("k", [0, 1, 2, 3, 300, 4096, 4097, 8192, 1 << 17])
def test_crt_i8_beta_fits_the_product_of_moduli(k):
    for N in range(1, O.CRT_MAX_MODULI + 1):
        b = O.crt_i8_beta(k, N)
        P = math.prod(O.crt_i8_moduli(N))
        kb = (k - 1).bit_length() if k > 1 else 0
        if b < 1:
            assert 2 * 2 ** kb * 2 ** 2 > 2 ** (P.bit_length() - 2)
            continue
        assert 2 * max(k, 1) * 2 ** (2 * b) < P
        assert max(k, 1) * 2 ** (2 * b) <= 2 ** 125
        if b < 53:
            assert 2 * 2 ** kb * 2 ** (2 * (b + 1)) >= 2 ** (P.bit_length() - 1) or \
                2 ** kb * 2 ** (2 * (b + 1)) > 2 ** 125
    # 15 INT8 moduli reach the integer width of 13 BF16 moduli at k = 8192
    assert O.crt_i8_beta(8192, 15) == O.crt_beta(8192, 13) == 51


def test_crt_i8_golden_through_equal_beta(golden):
    """Either residue form's result is fixed by beta alone: for every golden
    case (exact Python integers, BF16 moduli count N) and every INT8 moduli
    count with the same beta, the INT8 restatement gives the golden bits."""
    hits = 0
    for case in golden["crt"]:
        m, n, k, N = case["m"], case["n"], case["k"], case["N"]
        b = O.crt_beta(k, N)
        for N8 in range(1, O.CRT_MAX_MODULI + 1):
            if O.crt_i8_beta(k, N8) != b:
                continue
            A, B, C = unhex(case["A"], m, k), unhex(case["B"], k, n), unhex(case["C"], m, n)
            O.crt_i8_gemm(A, B, C, N8, float.fromhex(case["alpha"]), float.fromhex(case["beta"]))
            out = unhex(case["out"], m, n)
            assert np.array_equal(C.view(np.int64), out.view(np.int64)), (m, n, k, N, N8)
            hits += 1
    assert hits >= 3


@pytest.mark.parametrize("N,bound", [(10, 2.0 ** -18), (13, 2.0 ** -30), (16, 2.0 ** -51)])
def test_crt_i8_error_against_double_double(N, bound):
    rng = np.random.default_rng(SEED + 100 + N)
    m, n, k = 24, 20, 300
    A, B, C = (uniform_colmajor(rng, *sh) for sh in ((m, k), (k, n), (m, n)))
    ref, mag = O.dd_gemm(A, B, C)
    c = C.copy(order="F")
    O.crt_i8_gemm(A, B, c, N)
    err = np.abs(c - ref) / mag
    assert err.max() <= bound
    if N < 16:
        assert err.max() > bound / 2 ** 24   # the truncation at beta bits is visible
