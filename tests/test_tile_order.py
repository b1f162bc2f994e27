"""CPU check of the control-tree tile order (momms_amd/csrc/tile_order.hpp)
that momms_hip_dgemm_tree runs: the host build of the same code visits every
tile of the grid exactly once, and in the order of a direct restatement of
the tree's loops (nested PartM / PartN loops over whole 128-row / 128-column
tiles, src/composables/part.rs:7-79, written here as recursion, not as the
closed-form index arithmetic the kernel uses)."""
import random
import subprocess
from pathlib import Path

import pytest

PROG = Path(__file__).resolve().parent / "native" / "bin" / "tile_order_check"
M, N, K = 0, 1, 2

# hierarchy descriptions (innermost first) of the reference's trees, MR = 4,
# NR = 12 (src/exper/l3_shapes.rs:58-100, src/exper/l4.rs:71-78)
TREES = {
    "goto": [(M, 4), (N, 12), (M, 120), (K, 192), (N, 3000)],
    "l3a": [(N, 12), (M, 4), (K, 192), (N, 120), (K, 768), (M, 768)],
    "l3b": [(M, 4), (N, 12), (K, 192), (M, 120), (K, 768), (N, 768)],
    "l3c": [(M, 4), (N, 12), (M, 156), (K, 156), (M, 624), (N, 624)],
    "l4c": [(M, 4), (N, 12), (M, 120), (K, 192), (N, 3600), (M, 3600)],
}


def loops(steps, TM, TN, tile=128):
    """The tiles in the order the tree's M / N loops visit them."""
    levels = [(kind, max(1, bsz // tile)) for kind, bsz in reversed(steps) if kind != K]
    out = []

    def walk(i, m0, m1, n0, n1):
        if i == len(levels):
            for tn in range(n0, n1):        # what is left: m fastest
                for tm in range(m0, m1):
                    out.append((tm, tn))
            return
        kind, s = levels[i]
        if kind == M:
            for b in range(m0, m1, s):
                walk(i + 1, b, min(b + s, m1), n0, n1)
        else:
            for b in range(n0, n1, s):
                walk(i + 1, m0, m1, b, min(b + s, n1))

    walk(0, 0, TM, 0, TN)
    return out


def run(cases):
    text = "\n".join(f"{TM} {TN} {len(st)} " + " ".join(f"{k} {b}" for k, b in st)
                     for TM, TN, st in cases)
    res = subprocess.run([str(PROG)], input=text + "\n", capture_output=True, text=True,
                         check=True).stdout.strip().split("\n")
    return [None if r == "ERR" else [tuple(map(int, p.split(","))) for p in r.split()]
            for r in res]


@pytest.fixture(scope="module")
def prog():
    if not PROG.exists():
        pytest.fail(f"{PROG} is not built (run the build)")
    return PROG


def test_reference_trees_on_ragged_grids(prog):
    cases = [(TM, TN, st) for st in TREES.values() for TM, TN in ((64, 64), (7, 30), (33, 5), (1, 1))]
    for (TM, TN, st), got in zip(cases, run(cases)):
        want = loops(st, TM, TN)
        assert got == want, (TM, TN, st)
        assert sorted(got) == [(i, j) for i in range(TM) for j in range(TN)]


def test_random_trees(prog):
    gen = random.Random(20261204)
    cases = []
    for _ in range(300):
        steps = [(gen.choice([M, N, K]), gen.choice([4, 100, 128, 300, 1000, 5000]))
                 for _ in range(gen.randint(0, 7))]
        cases.append((gen.randint(1, 40), gen.randint(1, 40), steps))
    for (TM, TN, st), got in zip(cases, run(cases)):
        assert got == loops(st, TM, TN), (TM, TN, st)


def test_refused_steps(prog):
    nine = [(M, 128)] * 9
    got = run([(3, 3, nine), (3, 3, [(M, 0)]), (3, 3, [(5, 128)])])
    assert got == [None, None, None]
