"""CPU checks of the C ABI: the library loads, exports what include/momms_hip.h
declares, and rejects invalid arguments before touching the GPU."""
import ctypes

import pytest

from momms_amd import capi


def test_library_exports_every_declared_symbol():
    lib = capi.lib()
    names = capi.declared_symbols()
    assert len(names) >= 3
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    # every declared function has argtypes in the ctypes layer
    assert set(capi.declared_symbols()) <= set(capi._SIGS)


D = ctypes.c_double


def _dgemm(m, n, k, rsA=1, csA=1, rsB=1, csB=1, rsC=1, csC=1, A=8, B=8, C=8, alpha=True,
           beta=True):
    al, be = D(1.0), D(1.0)
    return capi.lib().momms_hip_dgemm(m, n, k, ctypes.byref(al) if alpha else None, A, rsA, csA,
                                      B, rsB, csB, ctypes.byref(be) if beta else None, C, rsC,
                                      csC, None)


@pytest.mark.parametrize("kw,code", [
    (dict(m=-1, n=1, k=1), -1),
    (dict(m=1, n=-1, k=1), -2),
    (dict(m=1, n=1, k=-1), -3),
    (dict(m=1, n=1, k=1, alpha=False), -4),
    (dict(m=1, n=1, k=1, A=None), -5),
    (dict(m=1, n=1, k=1, rsA=0), -6),
    (dict(m=1, n=1, k=1, csA=0), -7),
    (dict(m=1, n=1, k=1, B=None), -8),
    (dict(m=1, n=1, k=1, csB=-3), -10),
    (dict(m=1, n=1, k=1, beta=False), -11),
    (dict(m=1, n=1, k=1, C=None), -12),
    (dict(m=1, n=1, k=1, rsC=0), -13),
])
def test_dgemm_argument_errors(kw, code):
    st = _dgemm(**kw)
    assert st == code
    assert "momms_hip_dgemm" in capi.last_error()


def test_dgemm_empty_is_a_no_op():
    # m == 0 or n == 0 returns before any GPU work, with null matrices allowed
    assert _dgemm(0, 5, 3, A=None, C=None) == 0
    assert _dgemm(5, 0, 3, B=None, C=None) == 0


def test_dgemm_tiled_rejects_unknown_tile():
    al = D(1.0)
    st = capi.lib().momms_hip_dgemm_tiled(1, 1, 1, ctypes.byref(al), 8, 1, 1, 8, 1, 1,
                                          ctypes.byref(al), 8, 1, 1, 99, None)
    assert st == -15
