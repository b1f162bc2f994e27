"""The C++ GemmNode layer (momms_amd/host): CPU self-test of views, partitioning
and thread-team nodes, and (GPU) the HipGemm leaves inside reference-shaped
trees, each against TripleLoop."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent.parent / "momms_amd" / "host" / "bin"


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


@pytest.mark.gpu
def test_hip_nodes_in_gemmnode_trees():
    out = _run("l3_gpu", 300)
    assert out.count("PASS") == 9 and "FAIL" not in out
