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
    "momms_hip_dgemm_tree": [_c_i64, _c_i64, _c_i64, _c_dp, _c_vp, _c_i64, _c_i64, _c_vp,
                             _c_i64, _c_i64, _c_dp, _c_vp, _c_i64, _c_i64, _c_vp, _c_i64, _c_vp],
    "momms_hip_dgemm_small": [_c_i64, _c_i64, _c_i64, _c_dp, _c_vp, _c_i64, _c_i64, _c_vp,
                              _c_i64, _c_i64, _c_dp, _c_vp, _c_i64, _c_i64, _c_vp],
    "momms_hip_hier_pack_f32": [_c_vp, _c_vp, _c_i64, _c_i64, _c_vp, _c_vp],
    "momms_hip_hier_unpack_add_f32": [_c_vp, _c_vp, _c_vp, _c_i64, _c_i64, _c_vp],
    "momms_hip_hier_unpack_f32": [_c_vp, _c_vp, _c_vp, _c_i64, _c_i64, _c_vp],
    "momms_hip_dgemm_noalloc": [_c_i64, _c_i64, _c_i64, _c_dp, _c_vp, _c_i64, _c_i64, _c_vp,
                                _c_i64, _c_i64, _c_dp, _c_vp, _c_i64, _c_i64, _c_vp],
    "momms_last_error_string": [],
    "momms_hip_batch_class": [_c_vp, _c_i64, _c_vp],
    "momms_hip_tuning_export": [_c_vp],
    "momms_hip_tuning_import": [_c_vp, _c_int],
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
    "momms_hip_sgemm": [_c_i64, _c_i64, _c_i64, _c_fp, _c_vp, _c_i64, _c_i64, _c_vp, _c_i64,
                        _c_i64, _c_fp, _c_vp, _c_i64, _c_i64, _c_vp],
    "momms_hip_batch_plan_create": [_c_vp, _c_i64, _c_int, _c_int, _c_int, _c_vp, _c_vp,
                                    ctypes.POINTER(_c_vp), _c_vp],
    "momms_hip_batch_run": [_c_vp, _c_dp, _c_vp, _c_vp, _c_dp, _c_vp, _c_vp],
    "momms_hip_batch_capture": [_c_vp, _c_dp, _c_vp, _c_vp, _c_dp, _c_vp],
    "momms_hip_batch_launch": [_c_vp, _c_vp],
    "momms_hip_batch_plan_describe": [_c_vp],
    "momms_hip_batch_plan_destroy": [_c_vp],
    "momms_hip_dgemm_batched_loop": [_c_vp, _c_i64, _c_dp, _c_vp, _c_vp, _c_dp, _c_vp, _c_vp],
    "momms_hip_gemm_bf16": [_c_i64, _c_i64, _c_i64, _c_fp, _c_vp, _c_i64, _c_i64, _c_vp, _c_i64,
                            _c_i64, _c_fp, _c_vp, _c_i64, _c_i64, _c_vp],
    "momms_hip_crt_bf16_beta": [_c_i64, _c_int],
    "momms_hip_crt_bf16_modulus": [_c_int],
    "momms_hip_crt_bf16_workspace_size": [_c_i64, _c_i64, _c_i64, _c_int],
    "momms_hip_dgemm_crt_bf16": [_c_i64, _c_i64, _c_i64, _c_dp, _c_vp, _c_i64, _c_i64, _c_vp,
                                 _c_i64, _c_i64, _c_dp, _c_vp, _c_i64, _c_i64, _c_int, _c_vp,
                                 ctypes.c_size_t, _c_vp],
    "momms_hip_crt_bf16_prepare": [_c_i64, _c_i64, _c_i64, _c_vp, _c_i64, _c_i64, _c_vp, _c_i64,
                                   _c_i64, _c_int, _c_vp, ctypes.c_size_t, _c_vp],
    "momms_hip_crt_bf16_products": [_c_i64, _c_i64, _c_i64, _c_int, _c_vp, ctypes.c_size_t,
                                    _c_vp],
    "momms_hip_crt_i8_beta": [_c_i64, _c_int],
    "momms_hip_crt_i8_modulus": [_c_int],
    "momms_hip_crt_i8_workspace_size": [_c_i64, _c_i64, _c_i64, _c_int],
    "momms_hip_dgemm_crt_i8": [_c_i64, _c_i64, _c_i64, _c_dp, _c_vp, _c_i64, _c_i64, _c_vp,
                               _c_i64, _c_i64, _c_dp, _c_vp, _c_i64, _c_i64, _c_int, _c_vp,
                               ctypes.c_size_t, _c_vp],
    "momms_hip_crt_i8_prepare": [_c_i64, _c_i64, _c_i64, _c_vp, _c_i64, _c_i64, _c_vp, _c_i64,
                                 _c_i64, _c_int, _c_vp, ctypes.c_size_t, _c_vp],
    "momms_hip_crt_i8_products": [_c_i64, _c_i64, _c_i64, _c_int, _c_vp, ctypes.c_size_t,
                                  _c_vp],
    "momms_hip_crt_i8_reconstruct": [_c_i64, _c_i64, _c_i64, _c_dp, _c_vp, _c_i64, _c_i64,
                                     _c_vp, _c_i64, _c_i64, _c_dp, _c_vp, _c_i64, _c_i64,
                                     _c_int, _c_vp, ctypes.c_size_t, _c_vp],
    "momms_hier_layout_init": [_c_vp, _c_i64, _c_i64, _c_vp, _c_i64, ctypes.c_int32,
                               ctypes.c_int32, _c_i64, _c_i64, _c_i64, _c_i64],
    "momms_hip_hier_pack": [_c_vp, _c_vp, _c_i64, _c_i64, _c_vp, _c_vp],
    "momms_hip_hier_unpack_add": [_c_vp, _c_vp, _c_vp, _c_i64, _c_i64, _c_vp],
    "momms_hip_dgemm_hier": [_c_vp, _c_vp, _c_vp, _c_vp, _c_dp, _c_dp, _c_vp, _c_vp, _c_vp],
    "momms_hip_crt_bf16_reconstruct": [_c_i64, _c_i64, _c_i64, _c_dp, _c_vp, _c_i64, _c_i64,
                                       _c_vp, _c_i64, _c_i64, _c_dp, _c_vp, _c_i64, _c_i64,
                                       _c_int, _c_vp, ctypes.c_size_t, _c_vp],
}
_RESTYPES = {"momms_last_error_string": ctypes.c_char_p,
             "momms_hip_split_bf16_workspace_size": ctypes.c_size_t,
             "momms_hip_crt_bf16_workspace_size": ctypes.c_size_t,
             "momms_hip_crt_i8_workspace_size": ctypes.c_size_t,
             "momms_sample_index": ctypes.c_uint64,
             "momms_hip_batch_plan_describe": ctypes.c_char_p}


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


def _mat(t, h: int, w: int, name: str, output: bool = False):
    """(pointer, row stride, column stride) of a 2-D GPU tensor. A stride is
    replaced by 1 only along a dimension of extent <= 1, where it never
    scales an index. A stride of 0 (a broadcast view, from expand) along a
    longer dimension is rejected, and so is an output whose elements
    overlap: the kernels would read past the storage or race on one place."""
    if t.dim() != 2 or tuple(t.shape) != (h, w):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected ({h}, {w})")
    rs, cs = t.stride()
    if h <= 1:
        rs = 1
    if w <= 1:
        cs = 1
    if rs < 1 or cs < 1:
        raise ValueError(f"{name} has stride {tuple(t.stride())}: a broadcast (stride 0) or "
                         "negative stride along a dimension longer than 1")
    if output and h > 1 and w > 1:
        # the shorter stride's run must end before the longer stride's step
        (s_small, ext_small), s_big = sorted(((rs, h), (cs, w)))[0], max(rs, cs)
        if s_small * ext_small > s_big:
            raise ValueError(f"{name} has overlapping elements (shape {(h, w)}, strides "
                             f"{tuple(t.stride())})")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a GPU tensor")
    return t.data_ptr(), rs, cs


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
    c, rsc, csc = _mat(C, m, n, "C", output=True)
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


# Hierarchy descriptions (innermost step first, (kind, block size)) of the
# reference's control trees with MR = 4, NR = 12: Goto (src/exper/l3_shapes.rs:
# 58-64), resident A, B, C (l3_shapes.rs:69-100) and L4C (src/exper/l4.rs:
# 71-78); kinds 0 = M, 1 = N, 2 = K.
REFERENCE_TREES = {
    "goto": [(0, 4), (1, 12), (0, 120), (2, 192), (1, 3000)],
    "l3a": [(1, 12), (0, 4), (2, 192), (1, 120), (2, 768), (0, 768)],
    "l3b": [(0, 4), (1, 12), (2, 192), (0, 120), (2, 768), (1, 768)],
    "l3c": [(0, 4), (1, 12), (0, 156), (2, 156), (0, 624), (1, 624)],
    "l4c": [(0, 4), (1, 12), (0, 120), (2, 192), (1, 3600), (0, 3600)],
}


def dgemm_small(A, B, C, alpha: float = 1.0, beta: float = 1.0, stream=None):
    """C := beta*C + alpha*A*B by the single-launch small-GEMM leaf."""
    _f64((A, "A"), (B, "B"), (C, "C"))
    m, k = A.shape
    n = B.shape[1]
    a, rsa, csa = _mat(A, m, k, "A")
    b, rsb, csb = _mat(B, k, n, "B")
    c, rsc, csc = _mat(C, m, n, "C", output=True)
    al, be = ctypes.c_double(alpha), ctypes.c_double(beta)
    check("momms_hip_dgemm_small",
          lib().momms_hip_dgemm_small(m, n, k, ctypes.byref(al), a, rsa, csa, b, rsb, csb,
                                      ctypes.byref(be), c, rsc, csc, stream_handle(stream)))
    return C


def dgemm_tree(A, B, C, steps, alpha: float = 1.0, beta: float = 1.0, stream=None):
    """C := beta*C + alpha*A*B with the tiles in the order of a control tree
    whose hierarchy description (innermost first) is `steps`, a list of
    (STEP_M | STEP_N | STEP_K, block size) pairs."""
    _f64((A, "A"), (B, "B"), (C, "C"))
    m, k = A.shape
    n = B.shape[1]
    a, rsa, csa = _mat(A, m, k, "A")
    b, rsb, csb = _mat(B, k, n, "B")
    c, rsc, csc = _mat(C, m, n, "C", output=True)
    arr = (AlgoStep * max(1, len(steps)))()
    for i, (kind, bsz) in enumerate(steps):
        arr[i].kind, arr[i].bsz = int(kind), int(bsz)
    al, be = ctypes.c_double(alpha), ctypes.c_double(beta)
    st = lib().momms_hip_dgemm_tree(m, n, k, ctypes.byref(al), a, rsa, csa, b, rsb, csb,
                                    ctypes.byref(be), c, rsc, csc, ctypes.addressof(arr),
                                    len(steps), stream_handle(stream))
    check("momms_hip_dgemm_tree", st)
    return C


def _gemm_f32acc(fn: str, A, B, C, alpha, beta, in_dtype, stream):
    import torch

    for t, nm in ((A, "A"), (B, "B")):
        if t.dtype != in_dtype:
            raise TypeError(f"{nm} must be {in_dtype}")
    if C.dtype != torch.float32:
        raise TypeError("C must be float32")
    m, k = A.shape
    n = B.shape[1]
    a, rsa, csa = _mat(A, m, k, "A")
    b, rsb, csb = _mat(B, k, n, "B")
    c, rsc, csc = _mat(C, m, n, "C", output=True)
    al, be = ctypes.c_float(alpha), ctypes.c_float(beta)
    st = getattr(lib(), fn)(m, n, k, ctypes.byref(al), a, rsa, csa, b, rsb, csb, ctypes.byref(be),
                            c, rsc, csc, stream_handle(stream))
    check(fn, st)
    return C


def sgemm(A, B, C, alpha: float = 1.0, beta: float = 1.0, stream=None):
    """C := beta*C + alpha*A*B, float32 operands, FP32 matrix cores."""
    import torch

    return _gemm_f32acc("momms_hip_sgemm", A, B, C, alpha, beta, torch.float32, stream)


def gemm_bf16(A, B, C, alpha: float = 1.0, beta: float = 1.0, stream=None):
    """C := beta*C + alpha*A*B, bfloat16 A and B, float32 accumulation and C."""
    import torch

    return _gemm_f32acc("momms_hip_gemm_bf16", A, B, C, alpha, beta, torch.bfloat16, stream)


class GemmDesc(ctypes.Structure):
    """momms_gemm_desc: one column-major problem at element offsets into the arenas."""
    _fields_ = [(f, ctypes.c_int64) for f in
                ("m", "n", "k", "a_off", "b_off", "c_off", "lda", "ldb", "ldc")]


def desc_array(rows):
    """ctypes array of GemmDesc from (m, n, k, a_off, b_off, c_off, lda, ldb, ldc) tuples."""
    arr = (GemmDesc * len(rows))()
    for i, r in enumerate(rows):
        arr[i] = GemmDesc(*[int(x) for x in r])
    return arr


class BatchPlan:
    """A plan for one batch table (momms_hip_batch_plan_*)."""

    def __init__(self, rows, mode: int = 0, slices: int = 0, autotune: bool = False, A=None,
                 B=None, stream=None):
        self.descs = desc_array(rows)
        self.count = len(rows)
        h = ctypes.c_void_p()
        st = lib().momms_hip_batch_plan_create(
            ctypes.cast(self.descs, _c_vp), self.count, mode, slices, int(bool(autotune)),
            A.data_ptr() if A is not None else None, B.data_ptr() if B is not None else None,
            ctypes.byref(h), stream_handle(stream))
        check("momms_hip_batch_plan_create", st)
        self.handle = h

    def describe(self) -> dict:
        import json

        return json.loads(lib().momms_hip_batch_plan_describe(self.handle).decode())

    def run(self, A, B, C, alpha=1.0, beta=1.0, stream=None):
        al, be = ctypes.c_double(alpha), ctypes.c_double(beta)
        st = lib().momms_hip_batch_run(self.handle, ctypes.byref(al), A.data_ptr(), B.data_ptr(),
                                       ctypes.byref(be), C.data_ptr(), stream_handle(stream))
        check("momms_hip_batch_run", st)

    def capture(self, A, B, C, alpha=1.0, beta=1.0):
        al, be = ctypes.c_double(alpha), ctypes.c_double(beta)
        st = lib().momms_hip_batch_capture(self.handle, ctypes.byref(al), A.data_ptr(),
                                           B.data_ptr(), ctypes.byref(be), C.data_ptr())
        check("momms_hip_batch_capture", st)

    def launch(self, stream=None):
        check("momms_hip_batch_launch", lib().momms_hip_batch_launch(self.handle, stream_handle(stream)))

    def close(self):
        if getattr(self, "handle", None):
            lib().momms_hip_batch_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dgemm_batched_loop(rows_or_descs, A, B, C, alpha=1.0, beta=1.0, stream=None):
    d = rows_or_descs if isinstance(rows_or_descs, ctypes.Array) else desc_array(rows_or_descs)
    al, be = ctypes.c_double(alpha), ctypes.c_double(beta)
    st = lib().momms_hip_dgemm_batched_loop(ctypes.cast(d, _c_vp), len(d), ctypes.byref(al),
                                            A.data_ptr(), B.data_ptr(), ctypes.byref(be),
                                            C.data_ptr(), stream_handle(stream))
    check("momms_hip_dgemm_batched_loop", st)


TUNING_VERSION = 1
TUNING_MAX = 256
TUNE_KINDS = {"batch_fp64": 1, "batch_split": 2, "batch_crt": 3}


class _TuningEntry(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("key", ctypes.c_int32), ("value", ctypes.c_int32 * 4)]


class _Tuning(ctypes.Structure):
    """momms_tuning (include/momms_hip.h)."""
    _fields_ = [("version", ctypes.c_uint32), ("count", ctypes.c_uint32),
                ("entry", _TuningEntry * TUNING_MAX)]


def batch_class(rows_or_descs) -> int:
    """The class key the batched planner files a batch table's tuned choices under."""
    d = rows_or_descs if isinstance(rows_or_descs, ctypes.Array) else desc_array(rows_or_descs)
    key = ctypes.c_int32()
    check("momms_hip_batch_class",
          lib().momms_hip_batch_class(ctypes.cast(d, _c_vp), len(d), ctypes.byref(key)))
    return key.value


def tuning_export() -> dict:
    """The process's tuning table: {"version": 1, "entries": [{"kind", "key", "value"}]},
    kinds by name (TUNE_KINDS)."""
    t = _Tuning()
    check("momms_hip_tuning_export", lib().momms_hip_tuning_export(ctypes.byref(t)))
    names = {v: k for k, v in TUNE_KINDS.items()}
    return {"version": t.version,
            "entries": [{"kind": names[t.entry[i].kind], "key": t.entry[i].key,
                         "value": list(t.entry[i].value)} for i in range(t.count)]}


def tuning_import(table: dict, replace: bool = False) -> None:
    """Loads a table of the form tuning_export returns (the library checks every entry)."""
    entries = table.get("entries", [])
    if len(entries) > TUNING_MAX:
        raise ValueError(f"tuning table has {len(entries)} entries, at most {TUNING_MAX}")
    t = _Tuning()
    t.version = int(table.get("version", -1)) & 0xFFFFFFFF
    t.count = len(entries)
    for i, e in enumerate(entries):
        kind = e["kind"]
        t.entry[i].kind = TUNE_KINDS[kind] if isinstance(kind, str) else int(kind)
        t.entry[i].key = int(e["key"])
        vals = list(e["value"]) + [0] * (4 - len(e["value"]))
        if len(vals) != 4:
            raise ValueError(f"entry {i} has {len(e['value'])} values, at most 4")
        for j in range(4):
            t.entry[i].value[j] = int(vals[j])
    check("momms_hip_tuning_import", lib().momms_hip_tuning_import(ctypes.byref(t), int(bool(replace))))


def save_tuning(path) -> None:
    """Writes the process's tuning table to `path` as JSON."""
    import json

    Path(path).write_text(json.dumps(tuning_export(), indent=1) + "\n")


def load_tuning(path, replace: bool = False) -> None:
    """Imports a tuning table saved by save_tuning."""
    import json

    tuning_import(json.loads(Path(path).read_text()), replace)


def uneven_batch(count: int, lo: int = 16, hi: int = 512, seed: int = 0x6D6F6D6D73):
    """The uneven batch workload: m, n, k each floor(exp(U[ln lo, ln(hi+1)))) clamped to
    [lo, hi], packed back to back column-major. Returns (rows, a_len, b_len, c_len)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    dims = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi + 1), (count, 3)))).astype(np.int64)
    dims = np.clip(dims, lo, hi)
    rows = []
    ao = bo = co = 0
    for m, n, k in dims:
        rows.append((m, n, k, ao, bo, co, m, k, m))
        ao += m * k
        bo += k * n
        co += m * n
    return rows, ao, bo, co


def has(name: str) -> bool:
    """True when the built library exports `name`."""
    return hasattr(lib(), name)


def _f64(*ts):
    import torch

    for t, nm in ts:
        if t.dtype != torch.float64:
            raise TypeError(f"{nm} must be float64")


def _owned(ws, stream):
    """A workspace made for one call and dropped when it returns: its memory
    goes back to torch's caching allocator, which may hand it out again at
    once on the stream it was allocated on (the current one). When the
    kernels run on another stream, the allocator is told to wait for that
    stream's work before reusing the block."""
    if stream is not None:
        ws.buf.record_stream(stream)
    return ws


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
        c, rsc, csc = _mat(C, self.m, self.n, "C", output=True)
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
    ws = workspace or _owned(SplitWorkspace(m, n, k, slices, device=A.device), stream)
    if (ws.m, ws.n, ws.k, ws.slices) != (m, n, k, slices):
        raise ValueError("workspace was made for another problem")
    a, rsa, csa = _mat(A, m, k, "A")
    b, rsb, csb = _mat(B, k, n, "B")
    c, rsc, csc = _mat(C, m, n, "C", output=True)
    al, be = ctypes.c_double(alpha), ctypes.c_double(beta)
    st = lib().momms_hip_dgemm_split_bf16(m, n, k, ctypes.byref(al), a, rsa, csa, b, rsb, csb,
                                          ctypes.byref(be), c, rsc, csc, slices, ws.ptr,
                                          ws.nbytes, stream_handle(stream))
    check("momms_hip_dgemm_split_bf16", st)
    return C


class CrtWorkspace:
    """Device workspace of a residue split GEMM for one (m, n, k, moduli):
    form "bf16" (BF16 residues, moduli <= 511) or "i8" (INT8 residues,
    moduli <= 255)."""

    def __init__(self, m: int, n: int, k: int, moduli: int, device="cuda", form: str = "bf16"):
        import torch

        if form not in ("bf16", "i8"):
            raise ValueError(f"unknown residue form {form!r}")
        self.m, self.n, self.k, self.moduli, self.form = m, n, k, moduli, form
        self.nbytes = int(getattr(lib(), f"momms_hip_crt_{form}_workspace_size")(m, n, k, moduli))
        if self.nbytes == 0:
            raise ValueError(f"invalid residue problem m={m} n={n} k={k} moduli={moduli}")
        self.buf = torch.empty(self.nbytes, dtype=torch.uint8, device=device)

    @property
    def ptr(self) -> int:
        return self.buf.data_ptr()

    def _fn(self, stage: str):
        name = f"momms_hip_crt_{self.form}_{stage}"
        return name, getattr(lib(), name)

    def prepare(self, A, B, stream=None):
        _f64((A, "A"), (B, "B"))
        a, rsa, csa = _mat(A, self.m, self.k, "A")
        b, rsb, csb = _mat(B, self.k, self.n, "B")
        name, f = self._fn("prepare")
        check(name, f(self.m, self.n, self.k, a, rsa, csa, b, rsb, csb, self.moduli, self.ptr,
                      self.nbytes, stream_handle(stream)))

    def products(self, stream=None):
        name, f = self._fn("products")
        check(name, f(self.m, self.n, self.k, self.moduli, self.ptr, self.nbytes,
                      stream_handle(stream)))

    def reconstruct(self, A, B, C, alpha=1.0, beta=1.0, stream=None):
        _f64((A, "A"), (B, "B"), (C, "C"))
        a, rsa, csa = _mat(A, self.m, self.k, "A")
        b, rsb, csb = _mat(B, self.k, self.n, "B")
        c, rsc, csc = _mat(C, self.m, self.n, "C", output=True)
        al, be = ctypes.c_double(alpha), ctypes.c_double(beta)
        name, f = self._fn("reconstruct")
        check(name, f(self.m, self.n, self.k, ctypes.byref(al), a, rsa, csa, b, rsb, csb,
                      ctypes.byref(be), c, rsc, csc, self.moduli, self.ptr, self.nbytes,
                      stream_handle(stream)))
        return C


def crt_beta(k: int, moduli: int, form: str = "bf16") -> int:
    """Bits of the exponent-aligned integers for k and `moduli` (-1: too few)."""
    return int(getattr(lib(), f"momms_hip_crt_{form}_beta")(k, moduli))


def crt_moduli(count: int, form: str = "bf16") -> list[int]:
    f = getattr(lib(), f"momms_hip_crt_{form}_modulus")
    return [int(f(i)) for i in range(count)]


def _dgemm_crt(form, A, B, C, moduli, alpha, beta, workspace, stream):
    _f64((A, "A"), (B, "B"), (C, "C"))
    m, k = A.shape
    n = B.shape[1]
    ws = workspace or _owned(CrtWorkspace(m, n, k, moduli, device=A.device, form=form), stream)
    if (ws.m, ws.n, ws.k, ws.moduli, ws.form) != (m, n, k, moduli, form):
        raise ValueError("workspace was made for another problem")
    a, rsa, csa = _mat(A, m, k, "A")
    b, rsb, csb = _mat(B, k, n, "B")
    c, rsc, csc = _mat(C, m, n, "C", output=True)
    al, be = ctypes.c_double(alpha), ctypes.c_double(beta)
    name = f"momms_hip_dgemm_crt_{form}"
    st = getattr(lib(), name)(m, n, k, ctypes.byref(al), a, rsa, csa, b, rsb, csb,
                              ctypes.byref(be), c, rsc, csc, moduli, ws.ptr, ws.nbytes,
                              stream_handle(stream))
    check(name, st)
    return C


def dgemm_crt_bf16(A, B, C, moduli: int, alpha: float = 1.0, beta: float = 1.0,
                   workspace: CrtWorkspace | None = None, stream=None):
    """FP64-accurate C := beta*C + alpha*A*B from BF16 residues (in place on C)."""
    return _dgemm_crt("bf16", A, B, C, moduli, alpha, beta, workspace, stream)


def dgemm_crt_i8(A, B, C, moduli: int, alpha: float = 1.0, beta: float = 1.0,
                 workspace: CrtWorkspace | None = None, stream=None):
    """FP64-accurate C := beta*C + alpha*A*B from INT8 residues (in place on C)."""
    return _dgemm_crt("i8", A, B, C, moduli, alpha, beta, workspace, stream)


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


HIER_MAX_LEVELS = 8
STEP_M, STEP_N, STEP_K = 0, 1, 2  # AlgorithmStep kinds


class AlgoStep(ctypes.Structure):
    """momms_algo_step: AlgorithmStep::{M, N, K}{bsz} (src/composables/mod.rs:25-29)."""

    _fields_ = [("kind", ctypes.c_int32), ("pad", ctypes.c_int32), ("bsz", _c_i64)]


class HierLayout(ctypes.Structure):
    """momms_hier_layout: the shape of a Hierarch (src/matrix/hierarch.rs:18-184)."""

    _fields_ = [("h", _c_i64), ("w", _c_i64), ("h_pad", _c_i64), ("w_pad", _c_i64),
                ("ny", _c_i64), ("nx", _c_i64), ("capacity", _c_i64), ("lrs", _c_i64),
                ("lcs", _c_i64), ("y_blksz", _c_i64 * HIER_MAX_LEVELS),
                ("y_stride", _c_i64 * HIER_MAX_LEVELS), ("x_blksz", _c_i64 * HIER_MAX_LEVELS),
                ("x_stride", _c_i64 * HIER_MAX_LEVELS)]

    def as_list(self) -> list[int]:
        """[h, w, h_pad, w_pad, ny, nx, capacity, lrs, lcs, y_blksz..., y_stride...,
        x_blksz..., x_stride...] (levels beyond ny / nx are 0)."""
        head = [self.h, self.w, self.h_pad, self.w_pad, self.ny, self.nx, self.capacity,
                self.lrs, self.lcs]
        return head + list(self.y_blksz) + list(self.y_stride) + list(self.x_blksz) + \
            list(self.x_stride)


def hier_layout(h: int, w: int, steps, ydim: int, xdim: int, lh: int, lw: int, lrs: int,
                lcs: int) -> HierLayout:
    """Hierarch::new (src/matrix/hierarch.rs:90-140) for an h x w matrix: rows
    follow the steps of kind `ydim`, columns those of kind `xdim`; `steps` is a
    control tree's hierarchy description, [(kind, bsz), ...] innermost first;
    lh x lw leaves with strides lrs, lcs. Host only."""
    arr = (AlgoStep * max(len(steps), 1))(*[AlgoStep(int(k), 0, int(b)) for k, b in steps])
    L = HierLayout()
    st = lib().momms_hier_layout_init(ctypes.byref(L), h, w, arr, len(steps), ydim, xdim, lh, lw,
                                      lrs, lcs)
    if st != 0:
        why = {-2: "more than 8 levels in one dimension",
               -3: "blocks do not nest or the leaf is not dense"}.get(st, "bad argument")
        raise ValueError(f"momms_hier_layout_init returned {st}: {why}")
    return L


def _hier_dtype(*ts):
    """'' for float64 tensors, '_f32' for float32 ones (all alike)."""
    import torch

    dts = {t.dtype for t, _ in ts}
    if dts == {torch.float64}:
        return ""
    if dts == {torch.float32}:
        return "_f32"
    raise TypeError(f"{', '.join(n for _, n in ts)} must be all float64 or all float32")


def _hier_buf(L, P):
    if P.numel() < L.capacity or not P.is_cuda or (P.numel() > 1 and not P.is_contiguous()):
        raise ValueError(f"P must be a contiguous GPU tensor of >= {L.capacity} elements")


def hier_pack(L: HierLayout, A, P, stream=None):
    """Packer<Matrix, Hierarch>::pack (src/composables/pack.rs:153-309): P (a
    float64 or float32 GPU tensor of L.capacity elements, as A) := A in
    layout L, padding 0."""
    sfx = _hier_dtype((A, "A"), (P, "P"))
    a, ra, ca = _mat(A, L.h, L.w, "A")
    _hier_buf(L, P)
    name = "momms_hip_hier_pack" + sfx
    check(name, getattr(lib(), name)(ctypes.byref(L), a, ra, ca, P.data_ptr(),
                                     stream_handle(stream)))
    return P


def hier_unpack_add(L: HierLayout, P, C, stream=None):
    """Unpacker<Matrix, Hierarch>::add (src/composables/unpack.rs:174-213):
    C[y][x] += P[offset(y, x)] over the h x w matrix C (in place)."""
    sfx = _hier_dtype((P, "P"), (C, "C"))
    c, rc, cc = _mat(C, L.h, L.w, "C", output=True)
    _hier_buf(L, P)
    name = "momms_hip_hier_unpack_add" + sfx
    check(name, getattr(lib(), name)(ctypes.byref(L), P.data_ptr(), c, rc, cc,
                                     stream_handle(stream)))
    return C


def hier_unpack_f32(L: HierLayout, P, C, stream=None):
    """C[y][x] := P[offset(y, x)] (float32, C not read)."""
    _hier_dtype((P, "P"), (C, "C"))
    c, rc, cc = _mat(C, L.h, L.w, "C", output=True)
    _hier_buf(L, P)
    check("momms_hip_hier_unpack_f32",
          lib().momms_hip_hier_unpack_f32(ctypes.byref(L), P.data_ptr(), c, rc, cc,
                                          stream_handle(stream)))
    return C


def dgemm_hier(LA: HierLayout, PA, LB: HierLayout, PB, LC: HierLayout, PC, alpha: float = 1.0,
               beta: float = 1.0, stream=None):
    """C := beta*C + alpha*A*B on hierarchical buffers (float64 GPU tensors of at
    least LA/LB/LC.capacity elements), in place on PC."""
    _f64((PA, "PA"), (PB, "PB"), (PC, "PC"))
    for P, L, nm in ((PA, LA, "PA"), (PB, LB, "PB"), (PC, LC, "PC")):
        if P.numel() < L.capacity or not P.is_cuda or (P.numel() > 1 and not P.is_contiguous()):
            raise ValueError(f"{nm} must be a contiguous GPU tensor of >= {L.capacity} elements")
    a, b = ctypes.c_double(alpha), ctypes.c_double(beta)
    check("momms_hip_dgemm_hier",
          lib().momms_hip_dgemm_hier(ctypes.byref(LA), PA.data_ptr(), ctypes.byref(LB),
                                     PB.data_ptr(), ctypes.byref(a), ctypes.byref(b),
                                     ctypes.byref(LC), PC.data_ptr(), stream_handle(stream)))
    return PC


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
