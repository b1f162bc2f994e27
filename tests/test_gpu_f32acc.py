"""FP32 and BF16-in/FP32-accumulate GEMMs against float64 references.

Bar: |C - C*| <= gamma_{k+2}(u32) * (|alpha||A||B| + |beta||C0|) componentwise,
u32 = 2^-24, with C* the exact product of the given (already rounded) inputs
(computed in float64, whose own error is negligible at this scale). The max
relative error against the FP64 oracle is what config 4 reports.
"""
import numpy as np
import pytest
import torch

from conftest import SEED
from momms_amd import capi

pytestmark = pytest.mark.gpu


def gamma32(n):
    u = 2.0 ** -24
    return n * u / (1 - n * u)


def mk(rng, h, w, dtype, layout="col"):
    x = torch.from_numpy(rng.uniform(-1, 1, (h, w)))
    t = torch.empty((w, h), dtype=dtype, device="cuda").T if layout == "col" else \
        torch.empty((h, w), dtype=dtype, device="cuda")
    t.copy_(x)
    return t


def check(A, B, C0, C, alpha, beta, k):
    a = A.double().cpu().numpy()
    b = B.double().cpu().numpy()
    c0 = C0.double().cpu().numpy()
    exact = alpha * (a @ b) + beta * c0
    mag = abs(alpha) * (np.abs(a) @ np.abs(b)) + abs(beta) * np.abs(c0)
    err = np.abs(C.double().cpu().numpy() - exact)
    assert (err <= gamma32(k + 2) * mag + 1e-30).all(), float(np.max(err / mag))


@pytest.mark.parametrize("fn,dtype", [(capi.sgemm, torch.float32), (capi.gemm_bf16, torch.bfloat16)])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (129, 65, 17), (300, 200, 515), (256, 256, 256)])
@pytest.mark.parametrize("layout", ["col", "row"])
def test_against_float64(fn, dtype, m, n, k, layout):
    rng = np.random.default_rng(SEED + m + n + k)
    A, B = mk(rng, m, k, dtype, layout), mk(rng, k, n, dtype, layout)
    C0 = mk(rng, m, n, torch.float32, layout)
    C = C0.clone()
    fn(A, B, C, 1.5, -0.5)
    torch.cuda.synchronize()
    check(A, B, C0, C, 1.5, -0.5, k)


@pytest.mark.parametrize("fn,dtype", [(capi.sgemm, torch.float32), (capi.gemm_bf16, torch.bfloat16)])
def test_beta_zero_does_not_read_c(fn, dtype):
    rng = np.random.default_rng(SEED)
    A, B = mk(rng, 70, 40, dtype), mk(rng, 40, 90, dtype)
    C = torch.full((90, 70), float("nan"), dtype=torch.float32, device="cuda").T
    fn(A, B, C, 1.0, 0.0)
    torch.cuda.synchronize()
    assert torch.isfinite(C).all()
    check(A, B, torch.zeros_like(C), C, 1.0, 0.0, 40)


@pytest.mark.parametrize("fn,dtype", [(capi.sgemm, torch.float32), (capi.gemm_bf16, torch.bfloat16)])
def test_config4_shape_sampled(fn, dtype):
    """8192^3 with U[-1,1) inputs: sampled entries within the fp32 bound."""
    rng = np.random.default_rng(SEED)
    n = 8192
    g = torch.Generator(device="cuda").manual_seed(SEED & 0xFFFFFFFF)
    A = (torch.rand((n, n), generator=g, device="cuda", dtype=torch.float64) * 2 - 1).to(dtype).T
    B = (torch.rand((n, n), generator=g, device="cuda", dtype=torch.float64) * 2 - 1).to(dtype).T
    C = torch.zeros((n, n), dtype=torch.float32, device="cuda").T
    fn(A, B, C, 1.0, 0.0)
    torch.cuda.synchronize()
    rows = torch.from_numpy(rng.integers(0, n, 512)).cuda()
    cols = torch.from_numpy(rng.integers(0, n, 512)).cuda()
    a = A.double()[rows, :]
    b = B.double()[:, cols]
    exact = (a * b.T).sum(dim=1)
    mag = (a.abs() * b.abs().T).sum(dim=1)
    err = (C.double()[rows, cols] - exact).abs()
    assert bool((err <= gamma32(n + 2) * mag).all())


@pytest.mark.parametrize("m,n,k,layout", [(3000, 2900, 1000, "col"), (4100, 2060, 300, "row"),
                                          (2048, 4096, 777, "col")])
def test_bf16_packed_panel_form(m, n, k, layout):
    """Shapes large enough for the packed-panel BF16 form (256 x 256 tiles
    covering the device, k >= 256), ragged in every dimension."""
    rng = np.random.default_rng(SEED + m + k)
    A, B = mk(rng, m, k, torch.bfloat16, layout), mk(rng, k, n, torch.bfloat16, layout)
    C0 = mk(rng, m, n, torch.float32, layout)
    C = C0.clone()
    capi.gemm_bf16(A, B, C, -0.75, 2.0)
    torch.cuda.synchronize()
    check(A, B, C0, C, -0.75, 2.0, k)


@pytest.mark.parametrize("m,n,k,ld_pad", [(2100, 2050, 1000, 0), (2048, 2176, 31, 0),
                                          (2052, 2048, 4097, 0), (2100, 2049, 513, 12),
                                          (2049, 2048, 300, 0)])
def test_sgemm_dma_form(m, n, k, ld_pad):
    """Column-major FP32 shapes with at least 256 tiles of 128 x 128: the DMA
    tile (leading dimensions multiples of 4) with ragged m, n, k (partial last
    slab), padded leading dimensions, and the fallback when a leading
    dimension is not a multiple of 4 (m = 2049)."""
    print(f"sgemm dma {m}x{n}x{k} ld+{ld_pad}", flush=True)
    rng = np.random.default_rng(SEED + m + n + k)

    def colmajor(h, w):
        big = torch.empty((w, h + ld_pad), dtype=torch.float32, device="cuda")
        t = big.T[:h, :]
        t.copy_(torch.from_numpy(rng.uniform(-1, 1, (h, w))))
        return t

    A, B, C0 = colmajor(m, k), colmajor(k, n), colmajor(m, n)
    C = C0.clone() if ld_pad == 0 else colmajor(m, n)
    C.copy_(C0)
    capi.sgemm(A, B, C, 1.25, -0.5)
    torch.cuda.synchronize()
    check(A, B, C0, C, 1.25, -0.5, k)
