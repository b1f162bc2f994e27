"""The MFMA lane layouts the kernels are written against (momms_amd/csrc/mfma.hpp),
checked on the device against a CPU matmul for every instruction used."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
BIN = Path(__file__).resolve().parent / "native" / "bin" / "mfma_layout"


def test_mfma_layouts():
    assert BIN.exists(), "build() first"
    p = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=60)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.count("PASS") == 5
