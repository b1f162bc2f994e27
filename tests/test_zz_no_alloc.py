"""momms_hip_dgemm_noalloc (the allocation-free FP64 GEMM) against
momms_hip_dgemm, through the C ABI: the same bits on two calls in a row, the
second smaller, both on shapes that take the split-k form (whose partial
products are the fixed device array's contents), then a shape that needs no
workspace; and the same first call again after the smaller one overwrote
the array. The first result is also within gamma_{k+1} of double-double on
sampled entries, the bar of the FP64 GEMM."""
import ctypes

import numpy as np
import pytest
import torch

import pyoracle as O
from conftest import SEED, gamma, uniform_colmajor
from momms_amd import capi

pytestmark = pytest.mark.gpu


def call(name, A, B, C, alpha, beta):
    m, k = A.shape
    n = B.shape[1]
    al, be = ctypes.c_double(alpha), ctypes.c_double(beta)
    st = getattr(capi.lib(), name)(m, n, k, ctypes.byref(al), A.data_ptr(), A.stride(0),
                                   A.stride(1), B.data_ptr(), B.stride(0), B.stride(1),
                                   ctypes.byref(be), C.data_ptr(), C.stride(0), C.stride(1),
                                   capi.stream_handle())
    assert st == 0, (name, st, capi.last_error())


def case(rng, m, n, k):
    a, b, c0 = (uniform_colmajor(rng, *s) for s in ((m, k), (k, n), (m, n)))
    return a, b, c0, *(torch.from_numpy(x).cuda() for x in (a, b, c0))


def test_noalloc_matches_dgemm_on_two_calls_then_a_smaller_one():
    rng = np.random.default_rng(SEED + 77)
    shapes = [(1024, 1024, 1024), (512, 512, 512), (4096, 4096, 64), (1024, 1024, 1024)]
    inputs = {}
    for m, n, k in shapes:
        if (m, n, k) not in inputs:
            inputs[(m, n, k)] = case(rng, m, n, k)
        a, b, c0, A, B, C0 = inputs[(m, n, k)]
        Cref, Cnoa = C0.clone(), C0.clone()
        call("momms_hip_dgemm", A, B, Cref, 0.75, -1.25)
        call("momms_hip_dgemm_noalloc", A, B, Cnoa, 0.75, -1.25)
        torch.cuda.synchronize()
        assert torch.equal(Cref, Cnoa), (m, n, k)
        if (m, n, k) == shapes[0]:
            rows, cols = rng.integers(0, m, 2048), rng.integers(0, n, 2048)
            ref, mag = O.dd_entries(a, b, c0, rows, cols, 0.75, -1.25)
            got = Cnoa.cpu().numpy()[rows, cols]
            assert (np.abs(got - ref) <= gamma(k + 1) * mag).all()


def test_noalloc_argument_errors_before_gpu_work():
    al = ctypes.c_double(1.0)
    st = capi.lib().momms_hip_dgemm_noalloc(-1, 1, 1, ctypes.byref(al), 8, 1, 1, 8, 1, 1,
                                            ctypes.byref(al), 8, 1, 1, None)
    assert st == -1 and "momms_hip_dgemm_noalloc" in capi.last_error()
    st = capi.lib().momms_hip_dgemm_noalloc(4, 4, 4, ctypes.byref(al), 8, 1, 4, 8, 1, 4,
                                            None, 8, 1, 4, None)
    assert st == -11
