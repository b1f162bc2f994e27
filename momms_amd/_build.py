"""Build recipe for the native parts of momms_amd.

Produces, in-tree:
  momms_amd/libmomms_hip.so   the HIP kernel library behind include/momms_hip.h
  tests/native/bin/*          test programs (GemmNode host layer, native checks)
  oracle/liboracle.so         the CPU oracle (test infrastructure; built by
                              its own Makefile)

Every HIP source is compiled for gfx950 only, with packed-FP32 instruction
selection disabled (-packed-fp32-ops) and no floating-point contraction.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libmomms_hip.so"
HOST = PKG / "host"
ORACLE = ROOT / "oracle"

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # arithmetic exactly as written: the error-free transformations (TwoSum,
    # TwoProduct) and the bit-exact split reconstruction depend on it
    "-ffp-contract=off",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def _objects(jobs: int) -> list[Path]:
    """Compiles every csrc/*.hip and csrc/*.cpp to an object file, in parallel."""
    from concurrent.futures import ThreadPoolExecutor

    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))

    def one(src: Path) -> Path:
        obj = objdir / (src.name + ".o")
        if src.suffix == ".cpp":
            flags = [f for f in HIP_FLAGS if not f.startswith("--offload-arch")]
            _run([HIPCC, *flags, "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include",
                  "-c", str(src), "-o", str(obj)])
        else:
            _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
        return obj

    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        return list(ex.map(one, srcs))


def build_hip_library(jobs: int = 8) -> Path:
    objs = _objects(jobs)
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    shutil.rmtree(PKG / "_build", ignore_errors=True)
    return LIB


def _host_flags() -> list[str]:
    return ["-O2", "-std=c++17", "-Wall", "-pthread", "-mavx2", "-mfma",
            f"-I{HOST / 'include'}", f"-I{ROOT / 'include'}",
            "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include"]


def _host_libs() -> list[str]:
    return [f"-L{PKG}", "-lmomms_hip", f"-Wl,-rpath,{PKG}",
            f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib"]


def build_host_programs() -> list[Path]:
    """Test programs of the GemmNode host layer (header-only, links the ABI):
    tests/native/host/*.cpp -> tests/native/bin/, plus a ThreadSanitizer
    build of the CPU self-test (host code only, compiled by g++)."""
    progs = []
    src_dir = ROOT / "tests" / "native" / "host"
    if not src_dir.is_dir():
        return progs
    out_dir = ROOT / "tests" / "native" / "bin"
    out_dir.mkdir(exist_ok=True)
    for src in sorted(src_dir.glob("*.cpp")):
        out = out_dir / src.stem
        _run(["g++", *_host_flags(), str(src), "-o", str(out), *_host_libs()])
        progs.append(out)
    selftest = src_dir / "host_selftest.cpp"
    if selftest.exists():
        out = out_dir / "host_selftest_tsan"
        _run(["g++", *_host_flags(), "-g", "-fsanitize=thread", str(selftest), "-o", str(out),
              *_host_libs()])
        progs.append(out)
    return progs


def build_test_programs() -> list[Path]:
    """Programs the tests run: tests/native/*.hip (GPU) and *.cpp (host builds of
    device arithmetic) -> tests/native/bin/."""
    src_dir = ROOT / "tests" / "native"
    out_dir = src_dir / "bin"
    progs = []
    if not src_dir.is_dir():
        return progs
    out_dir.mkdir(exist_ok=True)
    for src in sorted(src_dir.glob("*.hip")):
        out = out_dir / src.stem
        _run([HIPCC, *HIP_FLAGS, str(src), "-o", str(out)])
        progs.append(out)
    # host builds of device-side arithmetic (CPU tests)
    for src in sorted(src_dir.glob("*.cpp")):
        out = out_dir / src.stem
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wno-unknown-pragmas", "-ffp-contract=off",
              f"-I{CSRC}", str(src),
              "-o", str(out)])
        progs.append(out)
    return progs


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ORACLE), "liboracle.so"])
    return ORACLE / "liboracle.so"


def build_all(jobs: int = 8) -> None:
    build_hip_library(jobs)
    build_host_programs()
    build_test_programs()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print(f"built {LIB}")
