"""Parity of the split-BF16 FP64-accurate GEMM and of the GPU accuracy checker.

Bar for the split GEMM: bit-exact against the oracle's restatement
(oracle/ozaki.c), which tests/test_oracle.py pins to exact-rational golden
vectors and to the double-double oracle. At full sizes, where the CPU
restatement is too slow, the error against double-double values at sampled
entries is checked against the truncation bound of the slice count.
"""
import numpy as np
import pytest
import torch

import pyoracle as O
from conftest import SEED, gamma
from momms_amd import capi

pytestmark = pytest.mark.gpu


def make(rng, h, w, e=0, layout="col"):
    x = rng.uniform(-1, 1, (h, w)) * np.exp2(rng.integers(-e, e + 1, (h, w)))
    if layout == "col":
        t = torch.empty((w, h), dtype=torch.float64, device="cuda").T
    elif layout == "row":
        t = torch.empty((h, w), dtype=torch.float64, device="cuda")
    else:
        big = torch.empty((2 * h + 1, 3 * w + 2), dtype=torch.float64, device="cuda")
        t = big[1::2, 2::3][:h, :w]
    t.copy_(torch.from_numpy(x))
    return t


def host(t):
    return np.asfortranarray(t.cpu().numpy())


def same_bits(a: np.ndarray, b: np.ndarray) -> bool:
    return np.array_equal(np.ascontiguousarray(a).view(np.int64), np.ascontiguousarray(b).view(np.int64))


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (5, 7, 3), (129, 65, 17), (200, 300, 129),
                                   (257, 129, 1000), (64, 64, 0)])
@pytest.mark.parametrize("s", [1, 3, 7])
def test_bitwise_against_restatement(m, n, k, s):
    rng = np.random.default_rng(SEED + m + n + k + s)
    A, B, C = make(rng, m, k, 8), make(rng, k, n, 8), make(rng, m, n, 8)
    want = host(C)
    O.ozaki_gemm(host(A), host(B), want, s, 1.0, 1.0)
    capi.dgemm_split_bf16(A, B, C, s, 1.0, 1.0)
    torch.cuda.synchronize()
    assert same_bits(host(C), want)


@pytest.mark.parametrize("s", [2, 4, 5, 6, 8])
@pytest.mark.parametrize("layouts", [("row", "row", "row"), ("gen", "col", "gen")])
def test_bitwise_slices_layouts_scalars(s, layouts):
    rng = np.random.default_rng(SEED + s)
    m, n, k = 150, 90, 300
    A, B, C = make(rng, m, k, 16, layouts[0]), make(rng, k, n, 16, layouts[1]), make(rng, m, n, 16, layouts[2])
    want = host(C)
    O.ozaki_gemm(host(A), host(B), want, s, -0.75, 2.0)
    capi.dgemm_split_bf16(A, B, C, s, -0.75, 2.0)
    torch.cuda.synchronize()
    assert same_bits(host(C), want)


def test_staged_equals_one_shot_and_workspace_reuse():
    rng = np.random.default_rng(SEED)
    m, n, k, s = 300, 200, 500, 5
    A, B, C0 = make(rng, m, k), make(rng, k, n), make(rng, m, n)
    ws = capi.SplitWorkspace(m, n, k, s)
    C1 = C0.clone()
    capi.dgemm_split_bf16(A, B, C1, s, 1.0, 1.0, workspace=ws)
    C2 = C0.clone()
    ws.prepare(A, B)
    ws.multiply(C2, 1.0, 1.0)
    C3 = C0.clone()
    ws.multiply(C3, 1.0, 1.0)   # slices stay valid in the workspace
    torch.cuda.synchronize()
    assert torch.equal(C1, C2) and torch.equal(C1, C3)


def test_beta_zero_does_not_read_c():
    rng = np.random.default_rng(SEED)
    m, n, k = 130, 70, 40
    A, B = make(rng, m, k), make(rng, k, n)
    C = torch.full((n, m), float("nan"), dtype=torch.float64, device="cuda").T
    capi.dgemm_split_bf16(A, B, C, 4, 1.0, 0.0)
    torch.cuda.synchronize()
    want = np.zeros((m, n), order="F")
    O.ozaki_gemm(host(A), host(B), want, 4, 1.0, 0.0)
    assert same_bits(host(C), want)


def test_nonfinite_inputs_propagate():
    rng = np.random.default_rng(SEED)
    m = n = k = 50
    A, B, C = make(rng, m, k), make(rng, k, n), make(rng, m, n)
    A[3, 5] = float("nan")
    B[7, 9] = float("inf")
    capi.dgemm_split_bf16(A, B, C, 3, 1.0, 1.0)
    torch.cuda.synchronize()
    c = C.cpu()
    assert torch.isnan(c[3, :]).all()
    assert not torch.isfinite(c[:, 9]).any()
    mask = torch.ones_like(c, dtype=torch.bool)
    mask[3, :] = False
    mask[:, 9] = False
    assert torch.isfinite(c[mask]).all()


# Truncation bound of the split product, scaled by |A||B|: the dropped levels
# and residuals are below about 4*s*2^(-8s) of the row/column maxima, and for
# entries of mixed exponent (spread e) the row maximum can exceed a typical
# |a| by 2^e on each side.
def split_bound(s, e):
    return max(8.0 * s * 2.0 ** (-8 * s) * 2.0 ** (2 * e), 4 * 2.0 ** -53)


@pytest.mark.parametrize("size", [1024, 4096, 8192])
@pytest.mark.parametrize("e", [0, 8, 16])
def test_config3_sampled_error_per_slice_count(size, e):
    """Config 3 inputs (U[-1,1)*2^E, E uniform in [-e, e]) at full sizes."""
    rng = np.random.default_rng(SEED + size + e)
    m = n = k = size
    a = np.asfortranarray(rng.uniform(-1, 1, (m, k)) * np.exp2(rng.integers(-e, e + 1, (m, k))))
    b = np.asfortranarray(rng.uniform(-1, 1, (k, n)) * np.exp2(rng.integers(-e, e + 1, (k, n))))
    c0 = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    A, B, C0 = (torch.from_numpy(x).cuda() for x in (a, b, c0))
    rows = rng.integers(0, m, 2048)
    cols = rng.integers(0, n, 2048)
    ref, mag = O.dd_entries(a, b, c0, rows, cols)
    ri, ci = torch.from_numpy(rows).cuda(), torch.from_numpy(cols).cuda()
    del a, b, c0
    ws = None
    prev = None
    for s in range(3, 8):
        ws = capi.SplitWorkspace(m, n, k, s)
        C = C0.clone()
        capi.dgemm_split_bf16(A, B, C, s, 1.0, 1.0, workspace=ws)
        torch.cuda.synchronize()
        got = C[ri, ci].cpu().numpy()
        err = float(np.max(np.abs(got - ref) / mag))
        assert err <= split_bound(s, e), (s, e, err)
        if prev is not None and e == 0 and s <= 6:
            assert err < prev     # each extra slice helps until fp64 rounding dominates
        prev = err
        del ws


def test_checker_matches_cpu_double_double():
    rng = np.random.default_rng(SEED)
    m, n, k = 300, 257, 999
    a, b, c0 = (np.asfortranarray(rng.uniform(-1, 1, s)) for s in ((m, k), (k, n), (m, n)))
    A, B, C0 = (torch.from_numpy(x).cuda() for x in (a, b, c0))
    C = C0.clone()
    capi.dgemm(A, B, C, 1.25, -0.5)
    samples = 5000
    ref_dev = torch.empty(samples, dtype=torch.float64, device="cuda")
    res = capi.sampled_error(A, B, C0, C, 1.25, -0.5, samples=samples, seed=77, ref_out=ref_dev)
    rows, cols = capi.sample_positions(77, samples, m, n)
    want, mag = O.dd_entries(a, b, c0, rows, cols, 1.25, -0.5)
    got = ref_dev.cpu().numpy()
    # both are double-double accurate (error <= u|x'y| + O(k^2 u^2)|x|'|y|):
    # they agree to about one ulp of the value, plus a term at the u^2 level
    diff = np.abs(got - want)
    bound = 4 * 2.0 ** -53 * np.abs(want) + 64 * gamma(k) ** 2 * mag
    worst = int(np.argmax(diff / bound))
    assert (diff <= bound).all(), (worst, got[worst], want[worst], mag[worst], diff[worst] / mag[worst])
    c = C.cpu().numpy()[rows, cols]
    scaled = np.max(np.abs(c - want) / mag)
    assert res["max_scaled"] == pytest.approx(scaled, rel=1e-6, abs=1e-30)
    assert 0 < res["max_scaled"] <= gamma(k + 1)


def test_checker_flags_a_planted_error():
    rng = np.random.default_rng(SEED)
    m, n, k = 64, 64, 64
    A, B, C0 = (torch.from_numpy(np.asfortranarray(rng.uniform(-1, 1, s))).cuda()
                for s in ((m, k), (k, n), (m, n)))
    C = C0.clone()
    capi.dgemm(A, B, C, 1.0, 1.0)
    C += 1e-6    # every entry off by 1e-6
    res = capi.sampled_error(A, B, C0, C, 1.0, 1.0, samples=100, seed=1)
    assert res["max_scaled"] > 1e-8
