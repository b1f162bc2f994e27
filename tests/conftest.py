import json
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

SEED = 0x6D6F6D6D73  # "momms"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on cuda:0)")


def _ensure_built():
    """Build what is missing (never rebuilds what exists)."""
    lib = ROOT / "momms_amd" / "libmomms_hip.so"
    orc = ROOT / "oracle" / "liboracle.so"
    if not orc.exists() or not lib.exists():
        from momms_amd import _build

        if not lib.exists():
            _build.build_hip_library()
            _build.build_host_programs()
        if not orc.exists():
            _build.build_oracle()


_ensure_built()


@pytest.fixture(scope="session")
def golden():
    return json.loads((ROOT / "tests" / "golden" / "golden.json").read_text())


def unhex(lst, h, w):
    """Column-major h x w array from a list of float.hex strings."""
    a = np.array([float.fromhex(x) for x in lst], dtype=np.float64)
    return np.asfortranarray(a.reshape((w, h)).T) if h * w else np.zeros((h, w), order="F")


def gamma(n, u=2.0 ** -53):
    return n * u / (1 - n * u)


def uniform_colmajor(rng, h, w, lo=-1.0, hi=1.0):
    """h x w column-major U[lo, hi) (the configs' input distribution)."""
    return np.asfortranarray(rng.uniform(lo, hi, (h, w)))
