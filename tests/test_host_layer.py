"""The C++ GemmNode layer (momms_amd/host): CPU self-test of views, partitioning
and thread-team nodes, and (GPU) the HipGemm leaves inside reference-shaped
trees, each against TripleLoop."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "native" / "bin"


def _run(name, timeout):
    exe = BIN / name
    assert exe.exists(), f"{exe} is not built"
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=timeout)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    return p.stdout


def test_host_selftest():
    out = _run("host_selftest", 120)
    assert "PASS (0 failures)" in out


def test_host_selftest_under_thread_sanitizer():
    """The thread-team nodes (ThreadComm barrier and broadcast with
    acquire/release, SpawnThreads, ParallelM/N, Barrier) run under
    ThreadSanitizer without a reported race."""
    exe = BIN / "host_selftest_tsan"
    assert exe.exists(), f"{exe} is not built"
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600,
                       env={**__import__("os").environ, "TSAN_OPTIONS": "halt_on_error=1"})
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-4000:]
    assert "PASS (0 failures)" in p.stdout
    assert "WARNING: ThreadSanitizer" not in p.stderr


@pytest.mark.gpu
def test_hip_nodes_in_gemmnode_trees():
    out = _run("l3_gpu", 300)
    assert out.count("PASS") == 24 and "FAIL" not in out
