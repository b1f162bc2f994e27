"""The bench line's roofline.traffic: the counter arithmetic of
measure_traffic on recorded counter tables, and the source digest that keeps
a stored reading from outliving the kernel it was taken on. No GPU: the
rocprofv3 child runs are replaced by writers of counter_collection.csv files
of the layout rocprofv3 --output-format csv produces."""
import csv
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


def _fake_rocprof(tables):
    """A subprocess.run stand-in that writes, for `rocprofv3 --pmc X -d D ...`,
    D/host/1/p_counter_collection.csv with the rows tables[X] (kernel name,
    dispatch id, value), split in two partial rows per dispatch."""

    def run(cmd, **kw):
        counter = cmd[cmd.index("--pmc") + 1]
        d = Path(cmd[cmd.index("-d") + 1]) / "host" / "1"
        d.mkdir(parents=True)
        with open(d / "p_counter_collection.csv", "w", newline="") as f:
            w = csv.writer(f)
            w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
            for name, disp, value in tables[counter]:
                w.writerow([disp, name, counter, value / 4])
                w.writerow([disp, name, counter, 3 * value / 4])
                w.writerow([disp, name, "OTHER", 1e9])
        return subprocess.CompletedProcess(cmd, 0, b"", None)

    return run


def test_measure_traffic_scales_by_the_copy(monkeypatch):
    copy, crt = "__amd_rocclr_copyBuffer", "void momms::crt_multiply<13>(...)"
    tables = {
        # a small host copy first, then five 1 GiB device copies (FETCH reads
        # half of the bytes, WRITE all of them, as measured on MI355X)
        "FETCH_SIZE": [(copy, 1, 4.0)] + [(copy, 2 + i, 524288.0) for i in range(5)]
                      + [(crt, 10, 1.0e7), (crt, 11, 1.1e7), (crt, 12, 1.2e7)],
        "WRITE_SIZE": [(copy, 1, 4.0)] + [(copy, 2 + i, 1048576.0) for i in range(5)]
                      + [(crt, 10, 2.0e6), (crt, 11, 2.5e6), (crt, 12, 2.2e6)],
    }
    monkeypatch.setattr(subprocess, "run", _fake_rocprof(tables))
    t = bench.measure_traffic(8192, 13)
    assert t["fetch_factor"] == 2.0 and t["write_factor"] == 1.0
    assert t["fetch_kib"] == 1.1e7 and t["write_kib"] == 2.2e6  # medians over the dispatches
    assert t["traffic_bytes"] == int((1.1e7 * 2 + 2.2e6) * 1024)


def test_measure_traffic_reports_a_failed_run(monkeypatch):
    def run(cmd, **kw):
        return subprocess.CompletedProcess(cmd, 137, b"line\nKilled\n", None)

    monkeypatch.setattr(subprocess, "run", run)
    with pytest.raises(RuntimeError, match="exited 137"):
        bench.measure_traffic(8192, 13)


def test_measure_traffic_needs_the_kernel(monkeypatch):
    copy = "__amd_rocclr_copyBuffer"
    tables = {c: [(copy, 1, 524288.0)] for c in ("FETCH_SIZE", "WRITE_SIZE")}
    monkeypatch.setattr(subprocess, "run", _fake_rocprof(tables))
    with pytest.raises(RuntimeError, match="crt_multiply"):
        bench.measure_traffic(8192, 13)


def test_sources_digest_follows_included_headers(tmp_path, monkeypatch):
    csrc, inc = tmp_path / "momms_amd" / "csrc", tmp_path / "include"
    csrc.mkdir(parents=True)
    inc.mkdir()
    (csrc / "k.hip").write_text('#include <math.h>\n#include "a.hpp"\n#include "api.h"\n')
    (csrc / "a.hpp").write_text('#include "b.hpp"\n')
    (csrc / "b.hpp").write_text("// leaf\n")
    (inc / "api.h").write_text("// interface\n")
    (csrc / "other.hip").write_text("// another kernel\n")
    monkeypatch.setattr(bench, "ROOT", tmp_path)
    base = bench.sources_digest("k.hip")
    assert base == bench.sources_digest("k.hip")
    # a header reached only through another header, and one under include/
    for f in (csrc / "b.hpp", inc / "api.h"):
        before = bench.sources_digest("k.hip")
        f.write_text(f.read_text() + "// changed\n")
        assert bench.sources_digest("k.hip") != before
    # a file the kernel does not include leaves it alone
    before = bench.sources_digest("k.hip")
    (csrc / "other.hip").write_text("// changed\n")
    assert bench.sources_digest("k.hip") == before


def test_sources_digest_of_the_headline_kernel_reads_its_headers():
    # every local header crt_bf16.hip names exists where the digest looks
    assert len(bench.sources_digest("crt_bf16.hip")) == 64
