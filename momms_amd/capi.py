"""ctypes binding of libmomms_hip.so (the C ABI declared in include/momms_hip.h).

The functions here take torch tensors that already live on the GPU and pass
their data pointers and element strides straight through; torch provides the
device memory and the stream, the arithmetic is the HIP library's. There is
no fallback: if the library is not built, loading it raises.

Shapes follow the reference's Matrix<T> (src/matrix/general_stride.rs:7-53):
an operand is an h x w view with a row stride and a column stride, so a
column-major matrix is `torch.empty((w, h)).T` and a row-major one is
`torch.empty((h, w))`.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libmomms_hip.so"

_c_i64 = ctypes.c_int64
_c_int = ctypes.c_int
_c_vp = ctypes.c_void_p
_c_dp = ctypes.POINTER(ctypes.c_double)
_c_fp = ctypes.POINTER(ctypes.c_float)

_lib = None


class MommsError(RuntimeError):
    """A momms_hip_* entry point returned a non-zero status."""

    def __init__(self, fn: str, status: int, text: str):
        super().__init__(f"{fn} returned {status}: {text}")
        self.status = status
        self.fn = fn


# name -> argtypes; all return int unless listed in _RESTYPES.
_SIGS: dict[str, list] = {
    "momms_hip_dgemm": [_c_i64, _c_i64, _c_i64, _c_dp, _c_vp, _c_i64, _c_i64, _c_vp, _c_i64,
                        _c_i64, _c_dp, _c_vp, _c_i64, _c_i64, _c_vp],
    "momms_hip_dgemm_tiled": [_c_i64, _c_i64, _c_i64, _c_dp, _c_vp, _c_i64, _c_i64, _c_vp,
                              _c_i64, _c_i64, _c_dp, _c_vp, _c_i64, _c_i64, _c_int, _c_vp],
    "momms_last_error_string": [],
    "momms_hip_split_bf16_workspace_size": [_c_i64, _c_i64, _c_i64, _c_int],
    "momms_hip_dgemm_split_bf16": [_c_i64, _c_i64, _c_i64, _c_dp, _c_vp, _c_i64, _c_i64, _c_vp,
                                   _c_i64, _c_i64, _c_dp, _c_vp, _c_i64, _c_i64, _c_int, _c_vp,
                                   ctypes.c_size_t, _c_vp],
    "momms_hip_split_bf16_prepare": [_c_i64, _c_i64, _c_i64, _c_vp, _c_i64, _c_i64, _c_vp,
                                     _c_i64, _c_i64, _c_int, _c_vp, ctypes.c_size_t, _c_vp],
    "momms_hip_split_bf16_multiply": [_c_i64, _c_i64, _c_i64, _c_dp, _c_dp, _c_vp, _c_i64,
                                      _c_i64, _c_int, _c_vp, ctypes.c_size_t, _c_vp],
    "momms_hip_gemm_check_sampled": [_c_i64, _c_i64, _c_i64, _c_dp, _c_vp, _c_i64, _c_i64,
                                     _c_vp, _c_i64, _c_i64, _c_dp, _c_vp, _c_vp, _c_i64, _c_i64,
                                     _c_i64, ctypes.c_uint64, _c_vp, _c_dp, _c_dp, _c_vp],
    "momms_sample_index": [ctypes.c_uint64, ctypes.c_uint64, _c_int, ctypes.c_uint64],
}
_RESTYPES = {"momms_last_error_string": ctypes.c_char_p,
             "momms_hip_split_bf16_workspace_size": ctypes.c_size_t,
             "momms_sample_index": ctypes.c_uint64}


def lib() -> ctypes.CDLL:
    """Loads libmomms_hip.so from the package directory (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python -m momms_amd._build`")
    l = ctypes.CDLL(str(LIB_PATH))
    for name, args in _SIGS.items():
        f = getattr(l, name, None)
        if f is None:
            continue
        f.argtypes = args
        f.restype = _RESTYPES.get(name, ctypes.c_int)
    _lib = l
    return l


def last_error() -> str:
    return lib().momms_last_error_string().decode()


def check(fn: str, status: int) -> None:
    if status != 0:
        raise MommsError(fn, status, last_error())


def stream_handle(stream=None) -> int | None:
    """hipStream_t of a torch stream (the current one by default)."""
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def _mat(t, h: int, w: int, name: str):
    if t.dim() != 2 or tuple(t.shape) != (h, w):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected ({h}, {w})")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a GPU tensor")
    rs, cs = t.stride()
    return t.data_ptr(), max(rs, 1), max(cs, 1)


def dgemm(A, B, C, alpha: float = 1.0, beta: float = 1.0, stream=None, tile_cfg: int | None = None):
    """C := beta*C + alpha*A*B on FP64 matrix cores (in place on C)."""
    import torch

    for t, nm in ((A, "A"), (B, "B"), (C, "C")):
        if t.dtype != torch.float64:
            raise TypeError(f"{nm} must be float64")
    m, k = A.shape
    n = B.shape[1]
    a, rsa, csa = _mat(A, m, k, "A")
    b, rsb, csb = _mat(B, k, n, "B")
    c, rsc, csc = _mat(C, m, n, "C")
    al = ctypes.c_double(alpha)
    be = ctypes.c_double(beta)
    s = stream_handle(stream)
    if tile_cfg is None:
        st = lib().momms_hip_dgemm(m, n, k, ctypes.byref(al), a, rsa, csa, b, rsb, csb,
                                   ctypes.byref(be), c, rsc, csc, s)
        check("momms_hip_dgemm", st)
    else:
        st = lib().momms_hip_dgemm_tiled(m, n, k, ctypes.byref(al), a, rsa, csa, b, rsb, csb,
                                         ctypes.byref(be), c, rsc, csc, int(tile_cfg), s)
        check("momms_hip_dgemm_tiled", st)
    return C


def has(name: str) -> bool:
    """True when the built library exports `name`."""
    return hasattr(lib(), name)


def _f64(*ts):
    import torch

    for t, nm in ts:
        if t.dtype != torch.float64:
            raise TypeError(f"{nm} must be float64")


class SplitWorkspace:
    """Device workspace of the split-BF16 GEMM for one (m, n, k, slices)."""

    def __init__(self, m: int, n: int, k: int, slices: int, device="cuda"):
        import torch

        self.m, self.n, self.k, self.slices = m, n, k, slices
        self.nbytes = int(lib().momms_hip_split_bf16_workspace_size(m, n, k, slices))
        if self.nbytes == 0:
            raise ValueError(f"invalid split problem m={m} n={n} k={k} slices={slices}")
        self.buf = torch.empty(self.nbytes, dtype=torch.uint8, device=device)

    @property
    def ptr(self) -> int:
        return self.buf.data_ptr()

    def prepare(self, A, B, stream=None):
        _f64((A, "A"), (B, "B"))
        a, rsa, csa = _mat(A, self.m, self.k, "A")
        b, rsb, csb = _mat(B, self.k, self.n, "B")
        st = lib().momms_hip_split_bf16_prepare(self.m, self.n, self.k, a, rsa, csa, b, rsb, csb,
                                               self.slices, self.ptr, self.nbytes,
                                               stream_handle(stream))
        check("momms_hip_split_bf16_prepare", st)

    def multiply(self, C, alpha=1.0, beta=1.0, stream=None):
        _f64((C, "C"))
        c, rsc, csc = _mat(C, self.m, self.n, "C")
        al, be = ctypes.c_double(alpha), ctypes.c_double(beta)
        st = lib().momms_hip_split_bf16_multiply(self.m, self.n, self.k, ctypes.byref(al),
                                                ctypes.byref(be), c, rsc, csc, self.slices,
                                                self.ptr, self.nbytes, stream_handle(stream))
        check("momms_hip_split_bf16_multiply", st)
        return C


def dgemm_split_bf16(A, B, C, slices: int, alpha: float = 1.0, beta: float = 1.0,
                     workspace: SplitWorkspace | None = None, stream=None):
    """FP64-accurate C := beta*C + alpha*A*B from BF16 slices (in place on C)."""
    _f64((A, "A"), (B, "B"), (C, "C"))
    m, k = A.shape
    n = B.shape[1]
    ws = workspace or SplitWorkspace(m, n, k, slices, device=A.device)
    if (ws.m, ws.n, ws.k, ws.slices) != (m, n, k, slices):
        raise ValueError("workspace was made for another problem")
    a, rsa, csa = _mat(A, m, k, "A")
    b, rsb, csb = _mat(B, k, n, "B")
    c, rsc, csc = _mat(C, m, n, "C")
    al, be = ctypes.c_double(alpha), ctypes.c_double(beta)
    st = lib().momms_hip_dgemm_split_bf16(m, n, k, ctypes.byref(al), a, rsa, csa, b, rsb, csb,
                                          ctypes.byref(be), c, rsc, csc, slices, ws.ptr,
                                          ws.nbytes, stream_handle(stream))
    check("momms_hip_dgemm_split_bf16", st)
    return C


def sampled_error(A, B, C0, C, alpha=1.0, beta=1.0, samples=65536, seed=0, ref_out=None,
                  stream=None) -> dict:
    """Max scaled and relative error of C against double-double values of
    beta*C0 + alpha*A*B at `samples` hashed positions (computed on the GPU)."""
    _f64((A, "A"), (B, "B"), (C, "C"))
    m, k = A.shape
    n = B.shape[1]
    a, rsa, csa = _mat(A, m, k, "A")
    b, rsb, csb = _mat(B, k, n, "B")
    c, rsc, csc = _mat(C, m, n, "C")
    if C0 is not None:
        c0, rs0, cs0 = _mat(C0, m, n, "C0")
        if (rs0, cs0) != (rsc, csc):
            raise ValueError("C0 and C must have the same strides")
    else:
        c0 = None
    al, be = ctypes.c_double(alpha), ctypes.c_double(beta)
    ms, mr = ctypes.c_double(0.0), ctypes.c_double(0.0)
    ref = ref_out.data_ptr() if ref_out is not None else None
    st = lib().momms_hip_gemm_check_sampled(m, n, k, ctypes.byref(al), a, rsa, csa, b, rsb, csb,
                                            ctypes.byref(be), c0, c, rsc, csc, samples, seed, ref,
                                            ctypes.byref(ms), ctypes.byref(mr),
                                            stream_handle(stream))
    check("momms_hip_gemm_check_sampled", st)
    return {"max_scaled": ms.value, "max_rel": mr.value, "samples": samples}


def sample_positions(seed: int, samples: int, m: int, n: int):
    """(rows, cols) that momms_hip_gemm_check_sampled visits."""
    f = lib().momms_sample_index
    rows = [f(seed, e, 0, m) for e in range(samples)]
    cols = [f(seed, e, 1, n) for e in range(samples)]
    return rows, cols


def declared_symbols(header: os.PathLike | None = None) -> list[str]:
    """Every function name declared in include/momms_hip.h."""
    import re

    h = Path(header) if header else _PKG.parent / "include" / "momms_hip.h"
    text = h.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(momms_\w+)\s*\(", text)))
