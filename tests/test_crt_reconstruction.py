"""CPU check of the residue split GEMMs' Chinese remainder step: the host build
of momms_amd/csrc/crt_recon.hpp (the code the reconstruction kernels run),
for the BF16 form's moduli (<= 511) and the INT8 form's (<= 255), recovers
every integer it is given from its balanced residues, at the ends of
the balanced range (-P/2, P/2) (where the quotient estimate can be one off
and the fix applies), at the 2^125 bound of the capped moduli counts, and at
random points; and its double is the correctly rounded one (Python's
float(int) rounds to nearest, ties to even)."""
import math
import random
import subprocess
from pathlib import Path

import pytest

import pyoracle as O

PROG = Path(__file__).resolve().parent / "native" / "bin" / "crt_recon_check"


def balanced(x, p):
    r = x % p
    return r - p if r > (p - 1) // 2 else r


@pytest.fixture(scope="module", params=["bf16", "i8"])
def cases(request):
    if not PROG.exists():
        pytest.fail(f"{PROG} is not built (run the build)")
    gen = random.Random(20261203)
    table = request.param
    mods = (O.crt_moduli if table == "bf16" else O.crt_i8_moduli)(O.CRT_MAX_MODULI)
    rows = []
    for N in range(1, O.CRT_MAX_MODULI + 1):
        P = math.prod(mods[:N])
        lim = min(P // 2, 2 ** 125)
        w = min(10 ** 6, lim)
        vals = [0, 1, -1, lim, -lim, lim - 1, -(lim - 1)]
        vals += [gen.randrange(-lim, lim + 1) for _ in range(400)]
        vals += [gen.randrange(-lim, -lim + w + 1) for _ in range(50)]
        vals += [gen.randrange(lim - w, lim + 1) for _ in range(50)]
        # values with ties and near-ties in the conversion to double
        for e in (60, 100, 120):
            if 2 ** e + 2 ** (e - 53) <= lim:
                vals += [2 ** e + 2 ** (e - 53), 2 ** e + 3 * 2 ** (e - 53), -(2 ** e + 2 ** (e - 53)),
                         2 ** e + 2 ** (e - 53) + 1]
        rows += [(N, c) for c in vals]
    return table, mods, rows


def test_reconstruction_recovers_every_integer(cases):
    table, mods, rows = cases
    text = "\n".join(f"{N} " + " ".join(str(balanced(c, p)) for p in mods[:N]) for N, c in rows)
    out = subprocess.run([str(PROG), table], input=text + "\n", capture_output=True, text=True,
                         check=True).stdout.split("\n")
    assert len([l for l in out if l]) == len(rows)
    for (N, c), line in zip(rows, out):
        hexval, dec = line.split()
        assert int(dec) == c, (N, c)
        got = float.fromhex(hexval)
        assert got == float(c) and math.copysign(1.0, got) == math.copysign(1.0, float(c)), (N, c)
