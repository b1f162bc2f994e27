"""ORACLE — test infrastructure only.

ctypes wrapper of oracle/liboracle.so for tests/, the smoke check and
bench.py's CPU-baseline leg. Arrays are numpy float64; a matrix argument is
(array, row_stride, col_stride) in elements, like the reference's Matrix<T>.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

_DIR = Path(__file__).resolve().parent
_PATH = _DIR / "liboracle.so"
_lib = None

_i64 = ctypes.c_int64
_vp = ctypes.c_void_p


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not _PATH.exists():
            import subprocess

            subprocess.run(["make", "-s", "-C", str(_DIR), "liboracle.so"], check=True)
        _lib = ctypes.CDLL(str(_PATH))
        _lib.oracle_residual.restype = ctypes.c_double
        _lib.oracle_ozaki_chunk.restype = ctypes.c_int
        _lib.oracle_hier_offset.restype = ctypes.c_int64
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(_vp)


def _d(x: float):
    return ctypes.byref(ctypes.c_double(x))


def strides(a: np.ndarray) -> tuple[int, int]:
    """Element strides (row, col) of a 2-D float64 numpy view."""
    s = a.strides
    return s[0] // a.itemsize, s[1] // a.itemsize


def colmajor(h: int, w: int) -> np.ndarray:
    return np.zeros((w, h), dtype=np.float64).T


def goto_dgemm(A, B, C, alpha=1.0, beta=1.0, nthreads=1, cpus=None):
    """Restated Goto tree, in place on C (any strides); thread t pinned to
    core cpus[t] when a list of nthreads cores is given."""
    m, k = A.shape
    n = B.shape[1]
    ra, ca = strides(A)
    rb, cb = strides(B)
    rc, cc = strides(C)
    pins = None
    if cpus is not None:
        if len(cpus) < nthreads:
            raise ValueError("need one core per thread")
        pins = (ctypes.c_int * nthreads)(*cpus[:nthreads])
    st = lib().oracle_goto_dgemm_pinned(_i64(m), _i64(n), _i64(k), _d(alpha), _p(A), _i64(ra),
                                        _i64(ca), _p(B), _i64(rb), _i64(cb), _d(beta), _p(C),
                                        _i64(rc), _i64(cc), ctypes.c_int(nthreads), pins)
    if st != 0:
        raise MemoryError("oracle_goto_dgemm could not allocate its pack buffers")
    return C


def goto_scalar(A, B, C, alpha=1.0, beta=1.0):
    m, k = A.shape
    n = B.shape[1]
    ra, ca = strides(A)
    rb, cb = strides(B)
    rc, cc = strides(C)
    lib().oracle_goto_scalar(_i64(m), _i64(n), _i64(k), _d(alpha), _p(A), _i64(ra), _i64(ca),
                             _p(B), _i64(rb), _i64(cb), _d(beta), _p(C), _i64(rc), _i64(cc))
    return C


def ukernel_6x8(k, alpha, a, b, beta, c, rs_c, cs_c):
    lib().oracle_ukernel_6x8(_i64(k), _d(alpha), _p(a), _p(b), _d(beta), _p(c), _i64(rs_c),
                             _i64(cs_c))
    return c


def triple_loop(A, B, C):
    m, k = A.shape
    n = B.shape[1]
    ra, ca = strides(A)
    rb, cb = strides(B)
    rc, cc = strides(C)
    lib().oracle_triple_loop(_i64(m), _i64(n), _i64(k), _p(A), _i64(ra), _i64(ca), _p(B),
                             _i64(rb), _i64(cb), _p(C), _i64(rc), _i64(cc))
    return C


def dd_gemm(A, B, C0, alpha=1.0, beta=1.0):
    """Double-double C* = beta*C0 + alpha*A*B and magnitudes |alpha||A||B| + |beta||C0|,
    both column-major m x n."""
    m, k = A.shape
    n = B.shape[1]
    ra, ca = strides(A)
    rb, cb = strides(B)
    rc, cc = strides(C0)
    out = colmajor(m, n)
    mag = colmajor(m, n)
    lib().oracle_dd_gemm(_i64(m), _i64(n), _i64(k), _d(alpha), _p(A), _i64(ra), _i64(ca), _p(B),
                         _i64(rb), _i64(cb), _d(beta), _p(C0), _i64(rc), _i64(cc), _p(out),
                         _p(mag))
    return out, mag


def dd_entries(A, B, C0, rows, cols, alpha=1.0, beta=1.0):
    k = A.shape[1]
    ra, ca = strides(A)
    rb, cb = strides(B)
    rc, cc = strides(C0)
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    out = np.zeros(len(rows))
    mag = np.zeros(len(rows))
    lib().oracle_dd_entries(_i64(k), _d(alpha), _p(A), _i64(ra), _i64(ca), _p(B), _i64(rb),
                            _i64(cb), _d(beta), _p(C0), _i64(rc), _i64(cc), _i64(len(rows)),
                            _p(rows), _p(cols), _p(out), _p(mag))
    return out, mag


def residual(A, B, C, C0, alpha, beta, w):
    m, k = A.shape
    n = B.shape[1]
    ra, ca = strides(A)
    rb, cb = strides(B)
    rc, cc = strides(C)
    w = np.ascontiguousarray(w, dtype=np.float64)
    return lib().oracle_residual(_i64(m), _i64(n), _i64(k), _p(A), _i64(ra), _i64(ca), _p(B),
                                 _i64(rb), _i64(cb), _p(C), _p(C0) if C0 is not None else None,
                                 _i64(rc), _i64(cc), ctypes.c_double(alpha),
                                 ctypes.c_double(beta), _p(w))


def ozaki_chunk(s: int) -> int:
    return lib().oracle_ozaki_chunk(ctypes.c_int(s))


def ozaki_split_rows(A, s):
    """Slices of every row of A: q[i, t, p] (m x s x k) and exponents E[i]."""
    m, k = A.shape
    ra, ca = strides(A)
    q = np.zeros((m, s, k))
    e = np.zeros(m, dtype=np.int32)
    lib().oracle_ozaki_split(_i64(m), _i64(k), _p(A), _i64(ra), _i64(ca), ctypes.c_int(s), _p(q),
                             _p(e))
    return q, e


def ozaki_split_cols(B, s):
    """Slices of every column of B: q[j, t, p] (n x s x k) and exponents F[j]."""
    k, n = B.shape
    rb, cb = strides(B)
    q = np.zeros((n, s, k))
    e = np.zeros(n, dtype=np.int32)
    lib().oracle_ozaki_split(_i64(n), _i64(k), _p(B), _i64(cb), _i64(rb), ctypes.c_int(s), _p(q),
                             _p(e))
    return q, e


def ozaki_gemm(A, B, C, s, alpha=1.0, beta=1.0):
    m, k = A.shape
    n = B.shape[1]
    ra, ca = strides(A)
    rb, cb = strides(B)
    rc, cc = strides(C)
    lib().oracle_ozaki_gemm(_i64(m), _i64(n), _i64(k), _d(alpha), _p(A), _i64(ra), _i64(ca),
                            _p(B), _i64(rb), _i64(cb), _d(beta), _p(C), _i64(rc), _i64(cc),
                            ctypes.c_int(s))
    return C


CRT_MAX_MODULI = 16


def crt_moduli(N: int) -> list[int]:
    return [lib().oracle_crt_modulus(ctypes.c_int(l)) for l in range(N)]


def crt_beta(k: int, N: int) -> int:
    return lib().oracle_crt_beta(_i64(k), ctypes.c_int(N))


def crt_integers_rows(A, beta):
    """Integer images of the rows of A: (m x k int64, exponents int32)."""
    m, k = A.shape
    ra, ca = strides(A)
    out = np.zeros((m, max(k, 1)), dtype=np.int64)
    e = np.zeros(m, dtype=np.int32)
    lib().oracle_crt_integers(_i64(m), _i64(k), _p(A), _i64(ra), _i64(ca), ctypes.c_int(beta),
                              _p(out), _p(e))
    return out[:, :k], e


def crt_gemm(A, B, C, N, alpha=1.0, beta=1.0):
    """Chinese-remainder split restatement with N moduli, in place on C."""
    m, k = A.shape
    n = B.shape[1]
    ra, ca = strides(A)
    rb, cb = strides(B)
    rc, cc = strides(C)
    st = lib().oracle_crt_gemm(_i64(m), _i64(n), _i64(k), _d(alpha), _p(A), _i64(ra), _i64(ca),
                               _p(B), _i64(rb), _i64(cb), _d(beta), _p(C), _i64(rc), _i64(cc),
                               ctypes.c_int(N))
    if st != 0:
        raise ValueError(f"oracle_crt_gemm: status {st} (too few moduli for k={k}?)")
    return C


def crt_i8_moduli(N: int) -> list[int]:
    return [lib().oracle_crt_i8_modulus(ctypes.c_int(l)) for l in range(N)]


def crt_i8_beta(k: int, N: int) -> int:
    return lib().oracle_crt_i8_beta(_i64(k), ctypes.c_int(N))


def crt_i8_gemm(A, B, C, N, alpha=1.0, beta=1.0):
    """The INT8 residue form with N moduli (<= 255), in place on C."""
    m, k = A.shape
    n = B.shape[1]
    ra, ca = strides(A)
    rb, cb = strides(B)
    rc, cc = strides(C)
    st = lib().oracle_crt_i8_gemm(_i64(m), _i64(n), _i64(k), _d(alpha), _p(A), _i64(ra),
                                  _i64(ca), _p(B), _i64(rb), _i64(cb), _d(beta), _p(C),
                                  _i64(rc), _i64(cc), ctypes.c_int(N))
    if st != 0:
        raise ValueError(f"oracle_crt_i8_gemm: status {st} (too few moduli for k={k}?)")
    return C


HIER_LEVELS = 8
M, N, K = 0, 1, 2  # hierarchy step kinds


def hier_layout(h, w, steps, ydim, xdim, lh, lw, lrs, lcs) -> np.ndarray:
    """Hierarch::new restated: steps = [(kind, bsz), ...] innermost first.
    Returns the int64 layout vector (see oracle.h)."""
    kind = np.array([s[0] for s in steps] or [0], dtype=np.int32)
    bsz = np.array([s[1] for s in steps] or [1], dtype=np.int64)
    out = np.zeros(9 + 4 * HIER_LEVELS, dtype=np.int64)
    st = lib().oracle_hier_layout(_i64(h), _i64(w), _p(kind), _p(bsz), _i64(len(steps)),
                                  ctypes.c_int32(ydim), ctypes.c_int32(xdim), _i64(lh), _i64(lw),
                                  _i64(lrs), _i64(lcs), _p(out))
    if st != 0:
        raise ValueError(f"oracle_hier_layout: status {st}")
    return out


def hier_offset(L, y, x) -> int:
    return int(lib().oracle_hier_offset(_p(L), _i64(y), _i64(x)))


def hier_offsets(L) -> np.ndarray:
    """offset(y, x) over the padded domain, [h_pad, w_pad]."""
    out = np.empty((int(L[2]), int(L[3])), dtype=np.int64)
    lib().oracle_hier_offsets(_p(L), _p(out))
    return out


def hier_pack(L, A) -> np.ndarray:
    ra, ca = strides(A)
    P = np.empty(int(L[6]), dtype=np.float64)
    lib().oracle_hier_pack(_p(L), _p(A), _i64(ra), _i64(ca), _p(P))
    return P


def hier_unpack_add(L, P, C):
    rc, cc = strides(C)
    lib().oracle_hier_unpack_add(_p(L), _p(P), _p(C), _i64(rc), _i64(cc))
    return C
