"""Parity of the FP64-accurate GEMM from INT8 residues (momms_hip_dgemm_crt_i8).

Bar: bit-exact against the oracle's INT8 restatement (oracle/crt.c computes
the exponent-aligned integer product directly in 128-bit integers with the
integer width of the INT8 moduli, so equality checks the device's INT8
residue GEMMs, their final modular reduction and the reconstruction), which
tests/test_oracle.py pins to the exact-integer golden vectors through equal
beta and to the double-double oracle. At full sizes the error against
double-double values at sampled entries is checked against the bound of the
integer width beta.
"""
import numpy as np
import pytest
import torch

import pyoracle as O
from conftest import SEED
from momms_amd import capi
from test_gpu_crt import host, make, same_bits

pytestmark = pytest.mark.gpu


def run_both(A, B, C, N, alpha, beta):
    want = host(C)
    O.crt_i8_gemm(host(A), host(B), want, N, alpha, beta)
    capi.dgemm_crt_i8(A, B, C, N, alpha, beta)
    torch.cuda.synchronize()
    return host(C), want


def test_moduli_and_beta_match_the_oracle():
    assert capi.crt_moduli(16, "i8") == O.crt_i8_moduli(16)
    for k in (1, 300, 8192, 131072):
        for N in range(1, 17):
            assert capi.crt_beta(k, N, "i8") == O.crt_i8_beta(k, N), (k, N)


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (5, 7, 3), (129, 65, 17), (200, 300, 129),
                                   (257, 129, 1000), (64, 64, 0), (300, 40, 4100), (70, 90, 65)])
@pytest.mark.parametrize("N", [15, 10, 16])
def test_bitwise_against_restatement(m, n, k, N):
    rng = np.random.default_rng(SEED + 7 * m + n + k + N)
    A, B, C = make(rng, m, k, 8), make(rng, k, n, 8), make(rng, m, n, 8)
    got, want = run_both(A, B, C, N, 1.0, 1.0)
    assert same_bits(got, want)


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_same_sign_accumulation_at_the_int32_limit(sign):
    """Constant A and B whose integer images are (p-1)/2 = 127 modulo the
    largest modulus p = 255, at the largest k the form accepts: every residue
    product of that GEMM is the same +-127^2, so its INT32 sum reaches
    131072 * 127^2 = 2,114,060,288, 1.6 % below 2^31, and its reduction modulo
    p starts from a float estimate of the quotient that is off by one. Equality
    with the exact restatement pins both."""
    m, n, k, N = 64, 64, 131072, 16
    p1 = O.crt_i8_moduli(N)[0]
    assert p1 == 255
    beta = O.crt_i8_beta(k, N)
    base = 1 << (beta - 1)
    ap = base + ((((p1 - 1) // 2) - base) % p1)  # smallest A' >= 2^(beta-1), A' = 127 mod 255
    assert ap < 2 * base and ap % p1 == (p1 - 1) // 2
    a = float(ap) * 2.0 ** (1 - beta)  # in [1, 2): row exponent 1, image exactly ap
    rng = np.random.default_rng(SEED + 3)
    A = torch.full((k, m), a, dtype=torch.float64, device="cuda").T
    B = torch.full((n, k), sign * a, dtype=torch.float64, device="cuda").T
    C = make(rng, m, n)
    got, want = run_both(A, B, C, N, 1.0, 1.0)
    assert same_bits(got, want)


@pytest.mark.parametrize("layouts", [("row", "row", "row"), ("gen", "col", "gen"),
                                     ("col", "gen", "row")])
@pytest.mark.parametrize("e", [0, 16])
def test_bitwise_layouts_scalars_spread(layouts, e):
    rng = np.random.default_rng(SEED + 5 * e + len(layouts[0]))
    m, n, k = 150, 90, 300
    A, B, C = make(rng, m, k, e, layouts[0]), make(rng, k, n, e, layouts[1]), make(rng, m, n, e, layouts[2])
    got, want = run_both(A, B, C, 16, -0.75, 2.0)
    assert same_bits(got, want)


def test_zero_rows_and_nonfinite_rows():
    rng = np.random.default_rng(SEED + 11)
    m, n, k = 130, 70, 40
    A, B = make(rng, m, k), make(rng, k, n)
    A[5, :] = 0.0
    B[:, 9] = 0.0
    A[7, 3] = float("nan")
    B[2, 11] = float("inf")
    C = torch.full((n, m), float("nan"), dtype=torch.float64, device="cuda").T
    capi.dgemm_crt_i8(A, B, C, 16, 1.0, 0.0)
    torch.cuda.synchronize()
    want = np.zeros((m, n), order="F")
    O.crt_i8_gemm(host(A), host(B), want, 16, 1.0, 0.0)
    got = host(C)
    assert np.array_equal(np.isnan(got), np.isnan(want))
    fin = np.isfinite(want)
    assert same_bits(got[fin], want[fin])
    assert (got[5, :11] == 0).all()


def test_staged_equals_one_shot():
    rng = np.random.default_rng(SEED + 12)
    m, n, k = 260, 130, 700
    A, B, C0 = make(rng, m, k, 4), make(rng, k, n, 4), make(rng, m, n)
    ws = capi.CrtWorkspace(m, n, k, 15, form="i8")
    C1, C2 = C0.clone(), C0.clone()
    capi.dgemm_crt_i8(A, B, C1, 15, 0.5, -1.0, workspace=ws)
    ws.prepare(A, B)
    ws.products()
    ws.reconstruct(A, B, C2, 0.5, -1.0)
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)


@pytest.mark.parametrize("size", [1024, 4096, 8192])
@pytest.mark.parametrize("e", [0, 8, 16])
def test_config3_sampled_error_per_moduli(size, e):
    """Config 3 inputs (U[-1,1)*2^E, E uniform in [-e, e]) at full sizes: the
    error of each INT8 moduli count stays within the bound of its integer
    width beta, and equals the BF16 form's bit for bit where the widths agree
    (15 INT8 and 13 BF16 moduli at k = 8192)."""
    rng = np.random.default_rng(SEED + size + e + 1)
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
    for N in (12, 14, 15, 16):
        beta = capi.crt_beta(k, N, "i8")
        C = C0.clone()
        capi.dgemm_crt_i8(A, B, C, N, 1.0, 1.0)
        torch.cuda.synchronize()
        got = C[ri, ci].cpu().numpy()
        err = float(np.max(np.abs(got - ref) / mag))
        bound = 4.0 * 2.0 ** (-beta) * 2.0 ** (2 * e) + 4 * 2.0 ** -53
        assert err <= bound, (N, beta, e, err)
        for Nb in range(1, 17):
            if capi.crt_beta(k, Nb) == beta:
                Cb = C0.clone()
                capi.dgemm_crt_bf16(A, B, Cb, Nb, 1.0, 1.0)
                assert torch.equal(C, Cb), (N, Nb)
                break
