"""Exact properties of the FP64-accurate forms at the benchmark's full size,
m = n = k = 8192, where the CPU restatement would take hours.

The residue forms compute the exponent-aligned integer product exactly and
round once, so their result is a function of the integer images alone:
  - the BF16 form at 13 moduli and the INT8 form at 15 moduli reach the same
    integer width beta at k = 8192 and must agree bit for bit (two different
    kernels: BF16 matrix cores with FP32 sums reduced modulo p, INT8 matrix
    cores with exact INT32 sums);
  - permuting k (columns of A with the rows of B) leaves every result bit
    unchanged: the row and column exponents are maxima over k, the integer
    sums are exact in any order;
  - (A B)^T = B^T A^T bit for bit: the exponents of the rows of B^T are those
    of the columns of B;
  - scaling A by 2^5 and B by 2^-3 scales the result by exactly 4.
The slice form folds its levels in FP64, so it keeps the transpose and
power-of-two properties but not the k-permutation one. Each property was
checked on the oracle (oracle/crt.c, oracle/ozaki.c) before being asserted
here; a dropped term, a wrong exponent or an inexact partial sum breaks it.
The plain FP64 kernels (data-parallel, stream-K, k-split) scale exactly by
powers of two too, and repeat their bits from run to run.
"""
import pytest
import torch

from conftest import SEED
from momms_amd import capi

pytestmark = pytest.mark.gpu

N = 8192


def colmajor(x):
    """The same matrix with unit row stride."""
    return x.T.contiguous().T


def rand_colmajor(rows, cols, gen, e):
    x = torch.rand((cols, rows), generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
    if e:
        x = x * torch.pow(2.0, torch.randint(-e, e + 1, (cols, rows), generator=gen,
                                             device="cuda").double())
    return x.T


def bits(x):
    return x.contiguous().view(torch.int64)


def same_bits(x, y):
    return torch.equal(bits(x), bits(y))


def run(form, A, B, count):
    C = torch.empty((B.shape[1], A.shape[0]), dtype=torch.float64, device="cuda").T
    if form == "bf16":
        capi.dgemm_crt_bf16(A, B, C, count, 1.0, 0.0)
    elif form == "i8":
        capi.dgemm_crt_i8(A, B, C, count, 1.0, 0.0)
    else:
        capi.dgemm_split_bf16(A, B, C, count, 1.0, 0.0)
    torch.cuda.synchronize()
    return C


@pytest.fixture(scope="module", params=[0, 8])
def inputs(request):
    gen = torch.Generator(device="cuda").manual_seed((SEED + request.param) & 0xFFFFFFFF)
    A = rand_colmajor(N, N, gen, request.param)
    B = rand_colmajor(N, N, gen, request.param)
    yield A, B, gen
    del A, B
    torch.cuda.empty_cache()


def test_bf16_and_int8_residue_forms_agree_bitwise(inputs):
    A, B, _ = inputs
    assert capi.crt_beta(N, 13) == capi.crt_beta(N, 15, "i8") == 51
    c16 = run("bf16", A, B, 13)
    c8 = run("i8", A, B, 15)
    assert torch.isfinite(c16).all()
    assert same_bits(c16, c8)
    # a different width gives different bits: the comparison is not vacuous
    assert not same_bits(c16, run("bf16", A, B, 14))


@pytest.mark.parametrize("form,count", [("bf16", 13), ("i8", 15)])
def test_residues_do_not_depend_on_the_order_of_k(inputs, form, count):
    A, B, gen = inputs
    perm = torch.randperm(N, generator=gen, device="cuda")
    Ap, Bp = colmajor(A[:, perm]), colmajor(B[perm, :])
    assert same_bits(run(form, A, B, count), run(form, Ap, Bp, count))


@pytest.mark.parametrize("form,count", [("bf16", 13), ("i8", 15), ("slices", 7)])
def test_transpose_symmetry(inputs, form, count):
    A, B, _ = inputs
    c = run(form, A, B, count)
    ct = run(form, colmajor(B.T), colmajor(A.T), count)
    assert same_bits(c, ct.T)


@pytest.mark.parametrize("form,count", [("bf16", 13), ("i8", 15), ("slices", 7)])
def test_power_of_two_scaling(inputs, form, count):
    A, B, _ = inputs
    c = run(form, A, B, count)
    cs = run(form, colmajor(A * 32.0), colmajor(B * 0.125), count)
    assert same_bits(c * 4.0, cs)


@pytest.mark.parametrize("m,n,k", [(8192, 8192, 8192), (1024, 1024, 1024), (8192, 64, 8192),
                                   (512, 512, 512), (8192, 8192, 256)])
def test_plain_fp64_power_of_two_scaling_and_repeatability(m, n, k):
    """Every product and sum of the FP64 kernels (the data-parallel tile, the
    stream-K and in-workgroup k-split forms with their fixed-order sums of
    parts) scales exactly by a power of two, so the result must too; a second
    run gives the same bits."""
    gen = torch.Generator(device="cuda").manual_seed((SEED + m + n + k) & 0xFFFFFFFF)
    A, B = rand_colmajor(m, k, gen, 8), rand_colmajor(k, n, gen, 8)

    def plain(X, Y):
        C = torch.empty((n, m), dtype=torch.float64, device="cuda").T
        capi.dgemm(X, Y, C, 1.0, 0.0)
        torch.cuda.synchronize()
        return C

    c = plain(A, B)
    assert same_bits(c, plain(A, B))
    assert same_bits(c * 4.0, plain(colmajor(A * 32.0), colmajor(B * 0.125)))
