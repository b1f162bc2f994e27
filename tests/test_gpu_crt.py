"""Parity of the Chinese-remainder split-BF16 FP64-accurate GEMM.

Bar: bit-exact against the oracle's restatement (oracle/crt.c computes the
exponent-aligned integer product directly in 128-bit integers, so equality
checks the device's residue GEMMs and reconstruction), which
tests/test_oracle.py pins to exact-integer golden vectors and to the
double-double oracle. At full sizes the error against double-double values at
sampled entries is checked against the bound of the integer width beta.
"""
import numpy as np
import pytest
import torch

import pyoracle as O
from conftest import SEED
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


def run_both(A, B, C, N, alpha, beta):
    want = host(C)
    O.crt_gemm(host(A), host(B), want, N, alpha, beta)
    capi.dgemm_crt_bf16(A, B, C, N, alpha, beta)
    torch.cuda.synchronize()
    return host(C), want


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (5, 7, 3), (129, 65, 17), (200, 300, 129),
                                   (257, 129, 1000), (64, 64, 0), (300, 40, 4100)])
@pytest.mark.parametrize("N", [14, 9, 16])
def test_bitwise_against_restatement(m, n, k, N):
    rng = np.random.default_rng(SEED + m + n + k + N)
    A, B, C = make(rng, m, k, 8), make(rng, k, n, 8), make(rng, m, n, 8)
    got, want = run_both(A, B, C, N, 1.0, 1.0)
    assert same_bits(got, want)


@pytest.mark.parametrize("sign", [1.0, -1.0])
@pytest.mark.parametrize("m,n,k", [(260, 70, 2100), (64, 300, 4096)])
def test_same_sign_accumulation_at_the_fp32_limit(sign, m, n, k):
    """Constant A and B whose integer images are (p-1)/2 = 255 modulo the
    largest modulus p = 511: every residue product of that GEMM is the same
    +-255^2, so each FP32 sum climbs by 256 * 255^2 = 16,646,400 per 256
    k-values, 0.8 % below 2^24. The sums stay exact only if every accumulator
    is reduced modulo p at least once per 256 k-values (259 products already
    exceed 2^24 with an odd value), so equality with the exact restatement pins
    the reduction schedule of the residue GEMM."""
    N = 13
    p1 = O.crt_moduli(N)[0]
    assert p1 == 511
    beta = O.crt_beta(k, N)
    base = 1 << (beta - 1)
    ap = base + ((((p1 - 1) // 2) - base) % p1)  # smallest A' >= 2^(beta-1), A' = 255 mod 511
    assert ap < 2 * base and ap % p1 == (p1 - 1) // 2
    a = float(ap) * 2.0 ** (1 - beta)  # in [1, 2): row exponent 1, image exactly ap
    rng = np.random.default_rng(SEED + k)
    A = torch.full((k, m), a, dtype=torch.float64, device="cuda").T
    B = torch.full((n, k), sign * a, dtype=torch.float64, device="cuda").T
    C = make(rng, m, n)
    got, want = run_both(A, B, C, N, 1.0, 1.0)
    assert same_bits(got, want)


@pytest.mark.parametrize("layouts", [("row", "row", "row"), ("gen", "col", "gen"),
                                     ("col", "gen", "row")])
@pytest.mark.parametrize("e", [0, 16])
def test_bitwise_layouts_scalars_spread(layouts, e):
    rng = np.random.default_rng(SEED + e + len(layouts[0]))
    m, n, k = 150, 90, 300
    A, B, C = make(rng, m, k, e, layouts[0]), make(rng, k, n, e, layouts[1]), make(rng, m, n, e, layouts[2])
    got, want = run_both(A, B, C, 14, -0.75, 2.0)
    assert same_bits(got, want)


def test_zero_rows_and_columns_and_beta_zero_does_not_read_c():
    rng = np.random.default_rng(SEED)
    m, n, k = 130, 70, 40
    A, B = make(rng, m, k), make(rng, k, n)
    A[5, :] = 0.0
    B[:, 9] = 0.0
    C = torch.full((n, m), float("nan"), dtype=torch.float64, device="cuda").T
    capi.dgemm_crt_bf16(A, B, C, 14, 1.0, 0.0)
    torch.cuda.synchronize()
    want = np.zeros((m, n), order="F")
    O.crt_gemm(host(A), host(B), want, 14, 1.0, 0.0)
    assert same_bits(host(C), want)
    assert (host(C)[5, :] == 0).all() and (host(C)[:, 9] == 0).all()


def test_nonfinite_rows_and_columns_take_the_fp64_path():
    rng = np.random.default_rng(SEED)
    m = n = k = 50
    A, B, C = make(rng, m, k), make(rng, k, n), make(rng, m, n)
    A[3, 5] = float("nan")
    B[7, 9] = float("inf")
    got, want = run_both(A, B, C, 14, 1.0, 1.0)
    assert np.isnan(got[3, :]).all()
    assert not np.isfinite(got[:, 9]).any()
    mask = np.ones((m, n), dtype=bool)
    mask[3, :] = False
    mask[:, 9] = False
    assert np.isfinite(got[mask]).all()
    assert same_bits(got[mask], want[mask])
    assert np.array_equal(np.isnan(got), np.isnan(want))


def test_staged_equals_one_shot():
    rng = np.random.default_rng(SEED + 1)
    m, n, k = 260, 130, 700
    A, B, C0 = make(rng, m, k, 4), make(rng, k, n, 4), make(rng, m, n)
    ws = capi.CrtWorkspace(m, n, k, 13)
    C1, C2 = C0.clone(), C0.clone()
    capi.dgemm_crt_bf16(A, B, C1, 13, 0.5, -1.0, workspace=ws)
    ws.prepare(A, B)
    ws.products()
    ws.reconstruct(A, B, C2, 0.5, -1.0)
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)


def test_workspace_reuse_is_deterministic():
    rng = np.random.default_rng(SEED)
    m, n, k = 300, 200, 500
    A, B, C0 = make(rng, m, k), make(rng, k, n), make(rng, m, n)
    ws = capi.CrtWorkspace(m, n, k, 14)
    C1, C2 = C0.clone(), C0.clone()
    capi.dgemm_crt_bf16(A, B, C1, 14, workspace=ws)
    capi.dgemm_crt_bf16(A, B, C2, 14, workspace=ws)
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)


@pytest.mark.parametrize("size", [1024, 4096, 8192])
@pytest.mark.parametrize("e", [0, 8, 16])
def test_config3_sampled_error_per_moduli(size, e):
    """Config 3 inputs (U[-1,1)*2^E, E uniform in [-e, e]) at full sizes: the
    error of each modulus count stays within the bound of its integer width
    beta (truncation at 2^-beta of the row / column maxima, which exceed a
    typical entry by up to 2^e each), and at 14 moduli reaches fp64 level."""
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
    for N in (10, 12, 13, 14, 16):
        beta = capi.crt_beta(k, N)
        C = C0.clone()
        capi.dgemm_crt_bf16(A, B, C, N, 1.0, 1.0)
        torch.cuda.synchronize()
        got = C[ri, ci].cpu().numpy()
        err = float(np.max(np.abs(got - ref) / mag))
        bound = 4.0 * 2.0 ** (-beta) * 2.0 ** (2 * e) + 4 * 2.0 ** -53
        assert err <= bound, (N, beta, e, err)
