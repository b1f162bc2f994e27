"""momms_amd — MI355X (gfx950) GEMM engine for the momms GemmNode API.

Layers:
  include/momms_hip.h, momms_amd/csrc/   HIP kernels behind a C ABI
  momms_amd/host/include/momms/          C++ GemmNode layer (Matrix<T> views,
                                         PartM/N/K, TripleLoop, HipGemm nodes)
  momms_amd/capi.py                      ctypes binding used by tests and bench
"""
from .capi import MommsError, dgemm, lib, last_error  # noqa: F401

__all__ = ["MommsError", "dgemm", "lib", "last_error"]
