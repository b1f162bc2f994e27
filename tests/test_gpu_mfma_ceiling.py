"""The matrix-core ceiling program (tests/native/mfma_ceiling.hip) that bench.py
runs beside the residue GEMMs: it runs, prints one JSON line, and its flop
accounting never puts a rate above the dense peak of the form it times."""
import json
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
BIN = Path(__file__).resolve().parent / "native" / "bin" / "mfma_ceiling"
PEAK_BF16 = 2516.6  # TFLOP/s, dense BF16 MFMA at 2.4 GHz
PEAK_INT8 = 5033.2  # TOP/s, dense INT8 MFMA at 2.4 GHz


def test_mfma_ceiling_rates_within_the_dense_peaks():
    assert BIN.exists(), "build() first"
    # a short run: 20000 iterations, a few milliseconds a launch
    p = subprocess.run([str(BIN), "20000"], capture_output=True, text=True, timeout=60)
    print(p.stdout, p.stderr)
    assert p.returncode == 0, p.stdout + p.stderr
    r = json.loads(p.stdout.strip().splitlines()[-1])
    assert r["iterations"] == 20000 and r["cus"] > 0
    for kind in ("random", "zero"):
        assert 0 < r["bf16_32x32x16"][f"{kind}_tflops"] <= PEAK_BF16 * 1.02
        assert 0 < r["i8_32x32x32"][f"{kind}_tops"] <= PEAK_INT8 * 1.02


def test_mfma_ceiling_rejects_a_bad_count():
    assert BIN.exists(), "build() first"
    p = subprocess.run([str(BIN), "0"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 2 and "positive" in p.stderr
