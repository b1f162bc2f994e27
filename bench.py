#!/usr/bin/env python3
"""Benchmark: FP64-accurate GEMM at m=n=k=8192 on one MI355X.

One step = one C := C + A*B at m=n=k=8192 (column-major f64, entries U[-1,1)
from a fixed seed, inputs resident in HBM) through a split-BF16 path, with the
plain FP64-MFMA GEMM timed beside it on the same inputs. Two split forms run,
each at the smallest count whose sampled errors (scaled and relative, against
double-double values) stay within 2x plain FP64's: BF16 slices (Ozaki scheme
I, s slices, s(s+1)/2 slice products) and BF16 residues (Ozaki scheme II, N
moduli, N residue GEMMs). Prints ONE JSON line:

  value        FP64-equivalent TFLOP/s (2mnk / time of one whole split GEMM,
               conversion + products + reconstruction) of the faster form
  roofline     that form's dominant kernel (crt_multiply / split_multiply)
               timed alone: its BF16 matrix-core flops per launch against the
               dense BF16 MFMA peak
  crt, split   both forms: count, rate, fraction of the FP64 peak, errors
  plain        the FP64-MFMA GEMM: TFLOP/s, fraction of the FP64 peak, errors
  configs      the other configurations' lines (sweep, slices/moduli x
               exponent spread, FP32/BF16, batched, hierarchical layout
               pack / unpack-add) unless --quick
  mfma_ceiling the same GPU's BF16 / INT8 matrix-core rate with operands in
               registers (random: the power-limited ceiling; zero: full
               clock), and the residue GEMMs' fractions of it
  vendor_reference  the vendor library's BF16 and INT8 GEMMs at the same
               size on the same GPU, and the residue GEMMs' rates against them
  cpu_baseline the restated Goto + 6x8 CPU path on this host's cores

roofline.traffic (HBM bytes per launch of crt_multiply) is measured in the
run: two rocprofv3 --pmc runs (FETCH_SIZE, then WRITE_SIZE) of this script's
--only traffic probe in a child process, each counter scaled by its reading
for a 1 GiB device copy in the same probe. --no-traffic skips them and uses
the reading stored in profiles/roofline_traffic.json only when it was taken
on the current kernel sources (sha256 over crt_bf16.hip and its headers).

--only {plain,split,multiply,crt,batched,batched_plain,batched_split,batched_crt,copy,hier,fp32,
        bf16,vendor,traffic}
    runs one kernel family in a loop
(for profilers) and prints nothing else but its timing line; vendor times the
vendor library's plain BF16 and INT8 GEMMs (torch.matmul, torch._int_mm) on
the same GPU as reference points for the residue GEMMs, and its FP64 GEMM
(torch.addmm) beside this library's over the plain sweep's shapes.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "FP64-accurate TFLOP/s @8192^3: split-BF16 MFMA vs plain FP64 MFMA, % peak"
PEAK_FP64 = 78.6      # TFLOP/s, MI355X dense FP64 matrix
PEAK_FP32 = 157.3     # TFLOP/s, MI355X dense FP32 matrix
PEAK_BF16 = 2516.6    # TFLOP/s, MI355X dense BF16 matrix
PEAK_INT8 = 5033.2    # TOP/s, MI355X dense INT8 matrix
PEAK_HBM = 8000.0     # GB/s
SEED = 0x6D6F6D6D73
# sampled entries of every accuracy statement; the slice-count selection and the
# reported errors use the same set, so the reported errors meet the selection rule
CHECK_SAMPLES = 65536


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


LAST_TIMING: dict = {}


def time_loop(fn, steps: int, warmup: int, group: int = 0) -> float:
    """Median milliseconds per call, from HIP events on torch's current stream
    (the stream every call launches on). After `warmup` calls and a device
    synchronization, the timed calls are queued behind a short spin kernel,
    so the GPU runs them back to back whatever the host's launch rate, and
    an event pair brackets each group of `group` calls (0: one call per pair
    when a call takes over 0.5 ms, else enough calls for about 0.5 ms, which
    keeps the events' own cost, a few microseconds, out of the figure). The
    median is over `steps` groups; LAST_TIMING also holds the mean over the
    whole bracket and the group size."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    if group <= 0:
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        one = e0.elapsed_time(e1)
        group = 1 if one >= 0.5 else min(50, max(1, int(0.5 / max(one, 1e-3))))
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    # about 40 us of host time per call to enqueue (spin counted in ~2.5 GHz cycles)
    torch.cuda._sleep(int(steps * group * 40e-6 * 2.5e9))
    ev[0].record()
    for i in range(steps):
        for _ in range(group):
            fn()
        ev[i + 1].record()
    torch.cuda.synchronize()
    per = sorted(ev[i].elapsed_time(ev[i + 1]) / group for i in range(steps))
    LAST_TIMING.clear()
    LAST_TIMING.update({"median_ms": per[len(per) // 2],
                        "mean_ms": ev[0].elapsed_time(ev[steps]) / (steps * group),
                        "group": group, "groups": steps})
    return per[len(per) // 2]


def rand_colmajor(n_rows, n_cols, gen, scale_e=0):
    """Column-major f64 U[-1,1) (times 2^E, E uniform in [-e, e]) generated on the GPU."""
    import torch

    x = torch.rand((n_cols, n_rows), generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
    if scale_e:
        e = torch.randint(-scale_e, scale_e + 1, (n_cols, n_rows), generator=gen, device="cuda")
        x = x * torch.pow(2.0, e.double())
    return x.T


def host_cpus() -> dict:
    """This host's CPUs as the process may use them: the model, the logical
    CPUs in its affinity mask, one logical CPU per physical core (SMT
    siblings share a (package, core) id in /sys/devices/system/cpu/*/topology),
    the cgroup CPU share (cpu.max quota / period, if one is set), and the
    maximum clock."""
    aff = sorted(os.sched_getaffinity(0))
    first = {}
    for c in aff:
        try:
            top = Path(f"/sys/devices/system/cpu/cpu{c}/topology")
            key = ((top / "physical_package_id").read_text().strip(),
                   (top / "core_id").read_text().strip())
        except OSError:
            key = ("?", str(c))
        first.setdefault(key, c)
    share = None
    try:
        quota, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()
        if quota != "max":
            share = max(1, int(quota) // int(period))
    except (OSError, ValueError):
        pass
    model, mhz, avx512 = "unknown", None, False
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name") and model == "unknown":
                model = line.split(":", 1)[1].strip()
            elif line.startswith("cpu MHz") and mhz is None:
                mhz = float(line.split(":", 1)[1])
            elif line.startswith("flags") and " avx512f" in line:
                avx512 = True
    except OSError:
        pass
    try:
        mhz = int(Path("/sys/devices/system/cpu/cpu0/cpufreq/cpuinfo_max_freq").read_text()) / 1000
    except (OSError, ValueError):
        pass
    return {"model": model, "logical_cpus": len(aff), "physical_cores": len(first),
            "cpu_share": share, "cores": sorted(first.values()), "max_mhz": mhz,
            "avx512": avx512}


def _best_of(run, c0, reps: int, flush) -> float:
    """Best wall-clock seconds of `reps` runs on fresh copies of c0, each after
    touching a 256 MiB buffer (l3_packing.rs:24-46: fresh inputs, flush, time
    run only, best of the repetitions)."""
    best = float("inf")
    for _ in range(reps):
        c = c0.copy(order="F")
        flush += 1.0
        t0 = time.perf_counter()
        run(c)
        best = min(best, time.perf_counter() - t0)
    return best


def cpu_baseline(a, b, c0, a1, b1, c1, threads: int, info: dict) -> dict:
    """The restated reference path (Goto tree, Nc/Kc/Mc = 3000/192/120, 6x8
    FMA micro-kernel, one pinned thread per physical core) on the same inputs
    as the GPU legs: the 8192^3 workload (A with the first n columns of B and
    C, all of them by default), and the m=n=k=1024 configuration; all threads
    and one thread; best of 5 (1024^3 at one thread: best of 3)."""
    import numpy as np

    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as O

    cpus = info["cores"][:threads]
    flush = np.zeros(256 * 1024 * 1024 // 8)
    m, k = a.shape
    n = b.shape[1]

    def goto(A, B, nt, pins):
        return lambda c: O.goto_dgemm(A, B, c, 1.0, 1.0, nthreads=nt, cpus=pins)

    t_all = _best_of(goto(a, b, threads, cpus), c0, 5, flush)
    t_c1 = _best_of(goto(a1, b1, threads, cpus), c1, 5, flush)
    t_c1_one = _best_of(goto(a1, b1, 1, cpus[:1]), c1, 3, flush)
    tf = 2.0 * m * n * k / t_all / 1e12
    gf_core = tf * 1e3 / threads
    c1_flops = 2.0 * a1.shape[0] * b1.shape[1] * a1.shape[1]
    # 2 FMA pipes x 4 doubles x 2 flops per cycle at the clock the host reports
    # (its boost clock can be higher: the one-thread rate is the better yardstick)
    peak_core = 16.0 * info["max_mhz"] / 1e3 if info["max_mhz"] else None
    return {"value": round(tf, 4), "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "sample": f"plain FP64 Goto+6x8 GEMM (the reference path restated in C) on "
                      + ("the whole workload's A, B and C" if n == m else
                         f"the workload's A and its first {n} of {m} columns of B and C")
                      + f" (m={m} n={n} k={k}), {threads} threads pinned one per physical "
                      f"core, best of 5 with fresh C and a 256 MiB flush, {t_all:.2f} s a run",
            "cpu_model": info["model"], "physical_cores": info["physical_cores"],
            "logical_cpus": info["logical_cpus"], "cgroup_cpu_share": info["cpu_share"],
            "gflops_per_core": round(gf_core, 2),
            "reported_mhz": info["max_mhz"],
            "avx2_fma_peak_per_core_gflops_at_reported_mhz": round(peak_core, 1) if peak_core else None,
            "frac_of_core_peak_at_reported_mhz": round(gf_core / peak_core, 3) if peak_core else None,
            "frac_of_one_thread_rate": round(gf_core / (c1_flops / t_c1_one / 1e9), 3),
            "config1_1024cube": {"all_threads_gflops": round(c1_flops / t_c1 / 1e9, 2),
                                 "one_thread_gflops": round(c1_flops / t_c1_one / 1e9, 2),
                                 "seconds_all": round(t_c1, 4),
                                 "seconds_one": round(t_c1_one, 4)}}


def sweep_plain(capi, torch, gen, steps):
    out = []
    shapes = [(s, s, s) for s in (512, 1024, 2048, 4096, 8192)] + [(8192, 8192, 256), (8192, 64, 8192)]
    for m, n, k in shapes:
        A, B, C = rand_colmajor(m, k, gen), rand_colmajor(k, n, gen), rand_colmajor(m, n, gen)
        ms = time_loop(lambda: capi.dgemm(A, B, C, 1.0, 1.0), steps, 3)
        tf = 2.0 * m * n * k / ms / 1e9
        out.append({"m": m, "n": n, "k": k, "ms": round(ms, 4), "tflops": round(tf, 3),
                    "frac_fp64_peak": round(tf / PEAK_FP64, 4),
                    "timing": f"median of {steps} event-timed groups of {LAST_TIMING['group']} "
                              "queued calls"})
        del A, B, C
    return out


def variant_trees(capi, torch, gen, steps, sizes=(4096, 8192)):
    """The reference's variant control trees (Goto, resident A / B / C, L4C)
    as tile orders of the FP64 GEMM (momms_hip_dgemm_tree), beside the
    built-in grouped raster of the same 128 x 128 tile (configuration 0):
    same bits, different L2 sharing."""
    out = []
    for n in sizes:
        A, B, C = rand_colmajor(n, n, gen), rand_colmajor(n, n, gen), rand_colmajor(n, n, gen)
        row = {"m": n, "n": n, "k": n}
        ms = time_loop(lambda: capi.dgemm(A, B, C, 1.0, 1.0, tile_cfg=0), steps, 3)
        row["grouped_raster"] = round(2.0 * n ** 3 / ms / 1e9, 3)
        for name, tree in capi.REFERENCE_TREES.items():
            ms = time_loop(lambda: capi.dgemm_tree(A, B, C, tree, 1.0, 1.0), steps, 3)
            row[name] = round(2.0 * n ** 3 / ms / 1e9, 3)
        row["unit"] = "TFLOP/s"
        out.append(row)
        del A, B, C
    return out


def split_grid(capi, torch, gen, steps, sizes=(4096, 8192), spreads=(0, 8, 16), slices=range(3, 8)):
    out = []
    for n in sizes:
        for e in spreads:
            A, B, C0 = rand_colmajor(n, n, gen, e), rand_colmajor(n, n, gen, e), rand_colmajor(n, n, gen)
            C = C0.clone()
            Cp = C0.clone()
            capi.dgemm(A, B, Cp, 1.0, 1.0)
            errp = capi.sampled_error(A, B, C0, Cp, samples=8192, seed=SEED)
            for s in slices:
                ws = capi.SplitWorkspace(n, n, n, s)
                ms = time_loop(lambda: capi.dgemm_split_bf16(A, B, C, s, 1.0, 1.0, workspace=ws), steps, 1)
                C.copy_(C0)
                capi.dgemm_split_bf16(A, B, C, s, 1.0, 1.0, workspace=ws)
                err = capi.sampled_error(A, B, C0, C, samples=8192, seed=SEED)
                tf = 2.0 * n ** 3 / ms / 1e9
                out.append({"n": n, "e": e, "slices": s, "ms": round(ms, 4), "tflops_fp64_equiv": round(tf, 3),
                            "frac_fp64_peak": round(tf / PEAK_FP64, 4),
                            "max_scaled_err": err["max_scaled"], "max_rel_err": err["max_rel"],
                            "plain_max_scaled_err": errp["max_scaled"]})
                del ws
            del A, B, C0, C, Cp
    return out


def crt_grid(capi, torch, gen, steps, sizes=(4096, 8192), spreads=(0, 8, 16),
             moduli=(10, 12, 13, 14, 16), form="bf16"):
    """Config 3 inputs through a residue form: error and rate per moduli count."""
    out = []
    for n in sizes:
        for e in spreads:
            A, B, C0 = rand_colmajor(n, n, gen, e), rand_colmajor(n, n, gen, e), rand_colmajor(n, n, gen)
            C = C0.clone()
            for nm in moduli:
                ws = capi.CrtWorkspace(n, n, n, nm, form=form)
                fn = capi.dgemm_crt_bf16 if form == "bf16" else capi.dgemm_crt_i8
                ms = time_loop(lambda: fn(A, B, C, nm, 1.0, 1.0, workspace=ws), steps, 1)
                C.copy_(C0)
                fn(A, B, C, nm, 1.0, 1.0, workspace=ws)
                err = capi.sampled_error(A, B, C0, C, samples=8192, seed=SEED)
                tf = 2.0 * n ** 3 / ms / 1e9
                out.append({"n": n, "e": e, "form": form, "moduli": nm,
                            "beta_bits": capi.crt_beta(n, nm, form),
                            "ms": round(ms, 4), "tflops_fp64_equiv": round(tf, 3),
                            "frac_fp64_peak": round(tf / PEAK_FP64, 4),
                            "max_scaled_err": err["max_scaled"], "max_rel_err": err["max_rel"]})
                del ws
            del A, B, C0, C
    return out


def reduced_precision(capi, torch, gen, steps, n=8192):
    out = []
    A64, B64 = rand_colmajor(n, n, gen), rand_colmajor(n, n, gen)
    for name, dt, fn, peak in (("fp32", torch.float32, capi.sgemm, PEAK_FP32),
                               ("bf16_fp32acc", torch.bfloat16, capi.gemm_bf16, PEAK_BF16)):
        A = A64.to(dt)
        B = B64.to(dt)
        C = torch.zeros((n, n), device="cuda", dtype=torch.float32).T
        ms = time_loop(lambda: fn(A, B, C, 1.0, 0.0), steps, 2)
        fn(A, B, C, 1.0, 0.0)
        # error against the FP64 oracle of the FP64 inputs (sampled, on the GPU)
        err = capi.sampled_error(A64, B64, None, C.double(), 1.0, 0.0, samples=8192, seed=SEED)
        tf = 2.0 * n ** 3 / ms / 1e9
        out.append({"kind": name, "n": n, "ms": round(ms, 4), "tflops": round(tf, 3),
                    "frac_peak": round(tf / peak, 4), "max_scaled_err": err["max_scaled"],
                    "max_rel_err": err["max_rel"]})
        del A, B, C
    return out


def batched(capi, torch, gen, steps):
    rows, al, bl, cl = capi.uneven_batch(4096)
    flops = sum(2.0 * r[0] * r[1] * r[2] for r in rows)
    A = torch.rand(al, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
    B = torch.rand(bl, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
    C = torch.rand(cl, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
    res = {"problems": len(rows), "gflop": round(flops / 1e9, 3)}
    plan = capi.BatchPlan(rows, mode=0, autotune=True, A=A, B=B)
    plan.capture(A, B, C, 1.0, 1.0)
    # a launch takes about a millisecond: at least 20 timed and 10 warm-up
    # launches each, the eager ones between two graph measurements (the
    # faster graph measurement is reported, so clock ramp-up does not count
    # against whichever runs first)
    bsteps, bwarm = max(steps, 20), 10
    ms_g1 = time_loop(plan.launch, bsteps, bwarm)
    ms_e = time_loop(lambda: plan.run(A, B, C, 1.0, 1.0), bsteps, bwarm)
    ms_g2 = time_loop(plan.launch, bsteps, bwarm)
    ms = min(ms_g1, ms_g2)
    res["plain_graph"] = {"ms": round(ms, 4), "tflops": round(flops / ms / 1e9, 3),
                          "plan": plan.describe()}
    res["plain_eager"] = {"ms": round(ms_e, 4), "tflops": round(flops / ms_e / 1e9, 3)}
    plan.close()
    descs = capi.desc_array(rows)
    ms = time_loop(lambda: capi.dgemm_batched_loop(descs, A, B, C, 1.0, 1.0), 2, 1)
    res["loop_of_launches"] = {"ms": round(ms, 4), "tflops": round(flops / ms / 1e9, 3)}
    splan = capi.BatchPlan(rows, mode=1, slices=7)
    splan.capture(A, B, C, 1.0, 1.0)
    ms = time_loop(splan.launch, bsteps, bwarm)
    res["split7_graph"] = {"ms": round(ms, 4), "tflops_fp64_equiv": round(flops / ms / 1e9, 3),
                           "plan": splan.describe()}
    splan.close()
    # the residue form, its moduli count chosen by the autotuner (the smallest
    # whose sampled error is within 2x plain FP64's on this batch)
    cplan = capi.BatchPlan(rows, mode=2, slices=0, autotune=True, A=A, B=B)
    cplan.capture(A, B, C, 1.0, 1.0)
    ms = time_loop(cplan.launch, max(steps, 10), 3)
    res["crt_graph"] = {"ms": round(ms, 4), "tflops_fp64_equiv": round(flops / ms / 1e9, 3),
                        "plan": cplan.describe()}
    cplan.close()
    return res


def hier(capi, torch, gen, steps, n=8192):
    """Device-resident hierarchical matrices at n x n: pack a column-major
    matrix into the Goto A (MR=4, MC=120, KC=192) and Goto C (MR=4, NR=12,
    MC=120, NC=3000) buffers of the reference's l3 experiment and add the C
    buffer back (algorithmic bytes = reads + writes of each element once,
    padding places written as zeros); then the GEMM C += A*B with A, B, C in
    the Goto A/B/C buffers."""
    M, N, K = capi.STEP_M, capi.STEP_N, capi.STEP_K
    goto = [(M, 4), (N, 12), (M, 120), (K, 192), (N, 3000)]
    out = {}
    X = rand_colmajor(n, n, gen)
    for name, (yd, xd, lrs, lcs) in {"goto_a": (M, K, 1, 4), "goto_c": (M, N, 12, 1)}.items():
        L = capi.hier_layout(n, n, goto, yd, xd, 4, 12, lrs, lcs)
        P = torch.empty(L.capacity, dtype=torch.float64, device="cuda")
        ms = time_loop(lambda: capi.hier_pack(L, X, P), steps, 3)
        b = 8.0 * (n * n + L.capacity)
        out[f"pack_{name}"] = {"ms": round(ms, 4), "bytes": b, "GBps": round(b / ms / 1e6, 1),
                               "frac_hbm": round(b / ms / 1e6 / PEAK_HBM, 4),
                               "capacity": L.capacity}
        if name == "goto_c":
            ms = time_loop(lambda: capi.hier_unpack_add(L, P, X), steps, 3)
            b = 8.0 * 3 * n * n
            out[f"unpack_add_{name}"] = {"ms": round(ms, 4), "bytes": b,
                                         "GBps": round(b / ms / 1e6, 1),
                                         "frac_hbm": round(b / ms / 1e6 / PEAK_HBM, 4)}
        del P
    # the GEMM on Goto A/B/C hierarchical buffers (C += A*B), against the plain
    # GEMM's 2mnk
    specs = [(M, K, 1, 4), (K, N, 12, 1), (M, N, 12, 1)]
    Ls = [capi.hier_layout(n, n, goto, yd, xd, 4, 12, lrs, lcs) for yd, xd, lrs, lcs in specs]
    Ps = [torch.empty(L.capacity, dtype=torch.float64, device="cuda") for L in Ls]
    for L, P in zip(Ls, Ps):
        capi.hier_pack(L, X, P)
    ms = time_loop(lambda: capi.dgemm_hier(Ls[0], Ps[0], Ls[1], Ps[1], Ls[2], Ps[2], 1.0, 1.0),
                   max(2, steps // 3), 1)
    out["dgemm_hier_goto"] = {"ms": round(ms, 4), "tflops": round(2.0 * n ** 3 / ms / 1e9, 3),
                              "frac_fp64_peak": round(2.0 * n ** 3 / ms / 1e9 / PEAK_FP64, 4)}
    del Ps
    return out


def sources_digest(name: str) -> str:
    """sha256 over a kernel's source file under momms_amd/csrc and every local
    header it includes, transitively: the files whose change can change the
    kernel, so a stored counter reading is used only for the sources it was
    taken on."""
    import hashlib
    import re

    dirs = [ROOT / "momms_amd" / "csrc", ROOT / "include"]
    seen, todo = {}, [ROOT / "momms_amd" / "csrc" / name]
    while todo:
        p = todo.pop()
        rel = str(p.relative_to(ROOT))
        if rel in seen:
            continue
        seen[rel] = p.read_bytes()
        for inc in re.findall(rb'#include\s+"([^"]+)"', seen[rel]):
            for d in [p.parent] + dirs:
                q = d / inc.decode()
                if q.exists():
                    todo.append(q.resolve())
                    break
    h = hashlib.sha256()
    for rel in sorted(seen):
        h.update(rel.encode() + b"\0" + seen[rel])
    return h.hexdigest()


COPY_BYTES = 1 << 30  # the calibration copy of the traffic probe


def traffic_probe(capi, torch, gen, n: int, moduli: int) -> dict:
    """The launches whose HBM counters give the headline kernel's traffic:
    five copies of a 1 GiB device buffer (a known 1 GiB read + 1 GiB write
    each, which scales the counters), then three launches of the residue
    products (crt_multiply) at m=n=k=n with `moduli` moduli."""
    src = torch.empty(COPY_BYTES, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    A, B = rand_colmajor(n, n, gen), rand_colmajor(n, n, gen)
    ws = capi.CrtWorkspace(n, n, n, moduli)
    ws.prepare(A, B)
    for _ in range(3):
        ws.products()
    torch.cuda.synchronize()
    return {"only": "traffic", "copies": 5, "products": 3, "moduli": moduli, "size": n}


def _pmc_dispatches(counter: str, args: list, out: Path, limit: int) -> dict:
    """KiB per dispatch of one rocprofv3 --pmc counter over this script run
    with `args` in a child process, grouped by kernel name."""
    import csv
    import subprocess

    d = out / counter
    cmd = ["timeout", "-s", "KILL", str(limit), "rocprofv3", "--pmc", counter, "-d", str(d),
           "-o", "p", "--output-format", "csv", "--", sys.executable, str(ROOT / "bench.py"),
           *args]
    log("bench: traffic pass: " + " ".join(cmd[4:]))
    r = subprocess.run(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"),
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=limit + 30)
    if r.returncode != 0:
        tail = r.stdout.decode(errors="replace").strip().splitlines()[-3:]
        raise RuntimeError(f"rocprofv3 --pmc {counter} exited {r.returncode}: {tail}")
    files = sorted(d.rglob("*counter_collection.csv"))
    if not files:
        raise RuntimeError(f"rocprofv3 --pmc {counter} wrote no counter_collection.csv")
    per: dict = {}
    with open(files[0], newline="") as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] != counter:
                continue
            key = (row["Kernel_Name"], int(row["Dispatch_Id"]))
            per[key] = per.get(key, 0.0) + float(row["Counter_Value"])
    by_kernel: dict = {}
    for (name, _), v in sorted(per.items(), key=lambda kv: kv[0][1]):
        by_kernel.setdefault(name, []).append(v)
    return by_kernel


def _median_of(by_kernel: dict, part: str) -> float:
    vals = sorted(v for name, vs in by_kernel.items() if part in name for v in vs)
    if not vals:
        raise RuntimeError(f"no dispatch of a kernel named *{part}*")
    return vals[len(vals) // 2]


def measure_traffic(n: int, moduli: int, limit: int = 180) -> dict:
    """HBM bytes per launch of crt_multiply at m=n=k=n, `moduli` moduli,
    measured now: FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each from its
    own rocprofv3 --pmc run of the traffic probe in a child process, each
    scaled by what the same counter reads for the probe's 1 GiB device copy
    (the medians over the dispatches of each kernel)."""
    import shutil
    import tempfile

    out = Path(tempfile.mkdtemp(prefix="momms_traffic_", dir="/tmp"))
    try:
        probe = ["--only", "traffic", "--size", str(n), "--moduli", str(moduli), "--no-cpu"]
        kib = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            by_kernel = _pmc_dispatches(counter, probe, out, limit)
            kib[counter] = (_median_of(by_kernel, "copyBuffer"), _median_of(by_kernel, "crt_multiply"))
    finally:
        shutil.rmtree(out, ignore_errors=True)
    ff = COPY_BYTES / 1024 / kib["FETCH_SIZE"][0]
    wf = COPY_BYTES / 1024 / kib["WRITE_SIZE"][0]
    fetch, write = kib["FETCH_SIZE"][1], kib["WRITE_SIZE"][1]
    return {"fetch_kib": fetch, "write_kib": write, "fetch_factor": round(ff, 4),
            "write_factor": round(wf, 4), "traffic_bytes": int(round((fetch * ff + write * wf) * 1024))}


def vendor_lowp_gemms(torch, gen, n: int, steps: int, warmup: int) -> tuple:
    """Milliseconds per call of the vendor library's plain BF16 GEMM
    (torch.matmul, hipBLASLt) and INT8 GEMM (torch._int_mm) at n^3 on this
    GPU: the same-box reference points for the residue GEMMs."""
    Ah = (torch.rand((n, n), generator=gen, device="cuda") * 2 - 1).to(torch.bfloat16)
    Bh = (torch.rand((n, n), generator=gen, device="cuda") * 2 - 1).to(torch.bfloat16)
    ms_bf = time_loop(lambda: torch.matmul(Ah, Bh), steps, warmup)
    del Ah, Bh
    Ai = torch.randint(-128, 128, (n, n), generator=gen, device="cuda", dtype=torch.int8)
    Bi = torch.randint(-128, 128, (n, n), generator=gen, device="cuda", dtype=torch.int8)
    ms_i8 = time_loop(lambda: torch._int_mm(Ai, Bi), steps, warmup)
    del Ai, Bi
    return ms_bf, ms_i8


def mfma_ceiling(limit: int = 60) -> dict:
    """The matrix cores' sustained rate on this GPU with the operands in
    registers (tests/native/mfma_ceiling.hip, run as a child process): the
    BF16 and INT8 MFMA forms of the residue GEMMs with random operands, which
    bound any GEMM kernel at the clock the chip holds under full matrix-core
    load, and with zero operands, which run at the full clock."""
    import subprocess

    exe = ROOT / "tests" / "native" / "bin" / "mfma_ceiling"
    if not exe.exists():
        raise RuntimeError(f"{exe} is not built")
    r = subprocess.run(["timeout", "-k", "5", str(limit), str(exe)], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, timeout=limit + 15)
    text = r.stdout.decode(errors="replace").strip()
    if r.returncode != 0 or not text.startswith("{"):
        raise RuntimeError(f"mfma_ceiling exited {r.returncode}: {text[-300:]}")
    return json.loads(text.splitlines()[-1])


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=8192)
    ap.add_argument("--slices", type=int, default=0,
                    help="0 = smallest count within 2x plain FP64's error (FP64-accurate)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline only, no other configs")
    ap.add_argument("--tile", type=int, default=None, help="FP64 tile configuration (--only plain)")
    ap.add_argument("--moduli", type=int, default=14, help="moduli of the residue form (--only crt)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter runs that measure roofline.traffic")
    ap.add_argument("--traffic-out", default=None,
                    help="also store the measured traffic, keyed by workload and kernel sources, "
                         "in this JSON file (the format of profiles/roofline_traffic.json)")
    ap.add_argument("--only", choices=["plain", "plain_sweep", "split", "multiply", "crt", "crt_i8",
                                       "batched",
                                       "batched_plain", "batched_split", "batched_crt", "copy", "hier",
                                       "fp32", "bf16", "vendor", "traffic"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported. This engine "
              "runs on a single MI355X; multi-GPU sharding is design-only.", file=sys.stderr)
        return 2

    import torch

    from momms_amd import capi

    n = args.size
    s = args.slices if args.slices else 7
    gen = torch.Generator(device="cuda").manual_seed(SEED & 0xFFFFFFFF)
    flops = 2.0 * n ** 3
    if args.only == "traffic":
        print(json.dumps(traffic_probe(capi, torch, gen, n, args.moduli)))
        return 0
    log(f"bench: inputs {n}^3 on {torch.cuda.get_device_name(0)}")
    A = rand_colmajor(n, n, gen)
    B = rand_colmajor(n, n, gen)
    C0 = rand_colmajor(n, n, gen)
    C = C0.clone()

    if args.only == "copy":
        # a plain copy of known size, for scaling HBM counters
        src = torch.empty(1 << 30, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        ms = time_loop(lambda: dst.copy_(src), args.steps, args.warmup)
        print(json.dumps({"only": "copy", "bytes": 1 << 30, "ms": ms,
                          "GBps": 2 * (1 << 30) / ms / 1e6}))
        return 0
    if args.only == "plain":
        ms = time_loop(lambda: capi.dgemm(A, B, C, 1.0, 1.0, tile_cfg=args.tile), args.steps,
                       args.warmup)
        print(json.dumps({"only": "plain", "tile": args.tile, "ms": ms, "tflops": flops / ms / 1e9}))
        return 0
    if args.only == "plain_sweep":
        # the plain FP64 sweep shapes at the built-in tile choice
        del A, B, C0, C
        print(json.dumps({"only": "plain_sweep",
                          "sweep": sweep_plain(capi, torch, gen, args.steps)}))
        return 0
    if args.only in ("split", "multiply"):
        ws = capi.SplitWorkspace(n, n, n, s)
        if args.only == "split":
            ms = time_loop(lambda: capi.dgemm_split_bf16(A, B, C, s, 1.0, 1.0, workspace=ws),
                           args.steps, args.warmup)
        else:
            ws.prepare(A, B)
            ms = time_loop(lambda: ws.multiply(C, 1.0, 1.0), args.steps, args.warmup)
        print(json.dumps({"only": args.only, "slices": s, "ms": ms, "tflops": flops / ms / 1e9}))
        return 0
    if args.only in ("crt", "crt_i8"):
        form = "bf16" if args.only == "crt" else "i8"
        fn = capi.dgemm_crt_bf16 if form == "bf16" else capi.dgemm_crt_i8
        ws = capi.CrtWorkspace(n, n, n, args.moduli, form=form)
        ms = time_loop(lambda: fn(A, B, C, args.moduli, 1.0, 1.0, workspace=ws),
                       args.steps, args.warmup)
        ws.prepare(A, B)
        pms = time_loop(ws.products, args.steps, args.warmup)
        print(json.dumps({"only": args.only, "moduli": args.moduli, "ms": ms,
                          "tflops": flops / ms / 1e9, "products_ms": pms}))
        return 0
    if args.only == "fp32":
        A32, B32 = A.float(), B.float()
        C32 = torch.zeros((n, n), device="cuda", dtype=torch.float32).T
        del A, B, C0, C
        ms = time_loop(lambda: capi.sgemm(A32, B32, C32, 1.0, 1.0), args.steps, args.warmup)
        print(json.dumps({"only": "fp32", "ms": ms, "tflops": flops / ms / 1e9}))
        return 0
    if args.only == "bf16":
        # the BF16 -> FP32 GEMM alone (bf16_panel_gemm), C := A*B
        Ah, Bh = A.to(torch.bfloat16), B.to(torch.bfloat16)
        C32 = torch.zeros((n, n), device="cuda", dtype=torch.float32).T
        del A, B, C0, C
        ms = time_loop(lambda: capi.gemm_bf16(Ah, Bh, C32, 1.0, 0.0), args.steps, args.warmup)
        print(json.dumps({"only": "bf16", "ms": ms, "tflops": flops / ms / 1e9,
                          "frac_peak": flops / ms / 1e9 / PEAK_BF16}))
        return 0
    if args.only == "vendor":
        # reference points on the same GPU: the vendor library's plain BF16 GEMM
        # (torch.matmul, hipBLASLt) and INT8 GEMM (torch._int_mm), n^3
        del A, B, C0, C
        ms_bf, ms_i8 = vendor_lowp_gemms(torch, gen, n, args.steps, args.warmup)
        # the vendor library's FP64 GEMM (torch.addmm on column-major operands,
        # C := C + A*B) over the plain sweep's shapes, beside this library's
        fp64 = []
        for m, nn, k in [(s, s, s) for s in (512, 1024, 2048, 8192)] + [(8192, 8192, 256),
                                                                         (8192, 64, 8192)]:
            A, B, C = rand_colmajor(m, k, gen), rand_colmajor(k, nn, gen), rand_colmajor(m, nn, gen)
            f = 2.0 * m * nn * k
            ms_v = time_loop(lambda: torch.addmm(C, A, B), max(args.steps, 20), args.warmup)
            ms_o = time_loop(lambda: capi.dgemm(A, B, C, 1.0, 1.0), max(args.steps, 20), args.warmup)
            fp64.append({"m": m, "n": nn, "k": k,
                         "vendor_ms": round(ms_v, 4),
                         "vendor_frac_fp64_peak": round(f / ms_v / 1e9 / PEAK_FP64, 4),
                         "momms_ms": round(ms_o, 4),
                         "momms_frac_fp64_peak": round(f / ms_o / 1e9 / PEAK_FP64, 4)})
            del A, B, C
        print(json.dumps({"only": "vendor", "fp64_addmm": fp64,
                          "bf16_matmul": {"ms": ms_bf, "tflops": flops / ms_bf / 1e9,
                                          "frac_peak": flops / ms_bf / 1e9 / PEAK_BF16},
                          "int8_int_mm": {"ms": ms_i8, "tops": flops / ms_i8 / 1e9,
                                          "frac_peak": flops / ms_i8 / 1e9 / PEAK_INT8}}))
        return 0
    if args.only == "hier":
        del A, B, C0, C
        print(json.dumps({"only": "hier", **hier(capi, torch, gen, args.steps, n)}))
        return 0
    if args.only == "batched":
        del A, B, C0, C
        print(json.dumps({"only": "batched", **batched(capi, torch, gen, args.steps)}))
        return 0
    if args.only == "batched_split":
        # the uneven batch through the slice plan at --slices slices, eager launches
        del A, B, C0, C
        rows, al, bl, cl = capi.uneven_batch(4096)
        flops = sum(2.0 * r[0] * r[1] * r[2] for r in rows)
        Ab = torch.rand(al, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
        Bb = torch.rand(bl, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
        Cb = torch.rand(cl, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
        plan = capi.BatchPlan(rows, mode=1, slices=s)
        ms = time_loop(lambda: plan.run(Ab, Bb, Cb, 1.0, 1.0), args.steps, args.warmup)
        print(json.dumps({"only": "batched_split", "ms": ms, "tflops_fp64_equiv": flops / ms / 1e9,
                          "plan": plan.describe()}))
        return 0
    if args.only == "batched_crt":
        # the uneven batch through the residue plan at --moduli moduli, eager launches
        del A, B, C0, C
        rows, al, bl, cl = capi.uneven_batch(4096)
        flops = sum(2.0 * r[0] * r[1] * r[2] for r in rows)
        Ab = torch.rand(al, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
        Bb = torch.rand(bl, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
        Cb = torch.rand(cl, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
        plan = capi.BatchPlan(rows, mode=2, slices=args.moduli)
        ms = time_loop(lambda: plan.run(Ab, Bb, Cb, 1.0, 1.0), args.steps, args.warmup)
        print(json.dumps({"only": "batched_crt", "ms": ms, "tflops_fp64_equiv": flops / ms / 1e9,
                          "plan": plan.describe()}))
        return 0
    if args.only == "batched_plain":
        # the uneven batch through the tuned plain plan only, eager launches
        del A, B, C0, C
        rows, al, bl, cl = capi.uneven_batch(4096)
        flops = sum(2.0 * r[0] * r[1] * r[2] for r in rows)
        Ab = torch.rand(al, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
        Bb = torch.rand(bl, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
        Cb = torch.rand(cl, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
        plan = capi.BatchPlan(rows, mode=0, autotune=True, A=Ab, B=Bb)
        ms = time_loop(lambda: plan.run(Ab, Bb, Cb, 1.0, 1.0), args.steps, args.warmup)
        print(json.dumps({"only": "batched_plain", "ms": ms, "tflops": flops / ms / 1e9,
                          "plan": plan.describe()}))
        return 0

    log("bench: plain FP64 MFMA")
    plain_ms = time_loop(lambda: capi.dgemm(A, B, C, 1.0, 1.0), args.steps, args.warmup)
    plain_tf = flops / plain_ms / 1e9

    # FP64-accurate: the smallest slice count (moduli count) whose sampled
    # errors, scaled and relative, are within 2x the plain FP64 GEMM's on these
    # inputs, on the same sample set as the reported errors
    Cp = C0.clone()
    capi.dgemm(A, B, Cp, 1.0, 1.0)
    err_p = capi.sampled_error(A, B, C0, Cp, 1.0, 1.0, samples=CHECK_SAMPLES, seed=SEED)
    del Cp

    def accurate(e):
        return e["max_scaled"] <= 2 * err_p["max_scaled"] and e["max_rel"] <= 2 * err_p["max_rel"]

    if args.slices == 0:
        s = 8
        for cand in range(4, 9):
            wsc = capi.SplitWorkspace(n, n, n, cand)
            Cc = C0.clone()
            capi.dgemm_split_bf16(A, B, Cc, cand, 1.0, 1.0, workspace=wsc)
            e = capi.sampled_error(A, B, C0, Cc, 1.0, 1.0, samples=CHECK_SAMPLES, seed=SEED)
            del wsc, Cc
            log(f"bench: slices={cand}: scaled {e['max_scaled']:.3e} rel {e['max_rel']:.3e} "
                f"(plain {err_p['max_scaled']:.3e} / {err_p['max_rel']:.3e})")
            if accurate(e):
                s = cand
                break
    nm = 16
    for cand in range(10, 17):
        wsc = capi.CrtWorkspace(n, n, n, cand)
        Cc = C0.clone()
        capi.dgemm_crt_bf16(A, B, Cc, cand, 1.0, 1.0, workspace=wsc)
        e = capi.sampled_error(A, B, C0, Cc, 1.0, 1.0, samples=CHECK_SAMPLES, seed=SEED)
        del wsc, Cc
        log(f"bench: moduli={cand}: scaled {e['max_scaled']:.3e} rel {e['max_rel']:.3e}")
        if accurate(e):
            nm = cand
            break

    log(f"bench: split-BF16 slices, s={s}")
    ws = capi.SplitWorkspace(n, n, n, s)
    Cs = C0.clone()
    split_ms = time_loop(lambda: capi.dgemm_split_bf16(A, B, Cs, s, 1.0, 1.0, workspace=ws),
                         args.steps, args.warmup)
    ws.prepare(A, B)
    mult_ms = time_loop(lambda: ws.multiply(Cs, 1.0, 1.0), args.steps, args.warmup)
    Cs.copy_(C0)
    capi.dgemm_split_bf16(A, B, Cs, s, 1.0, 1.0, workspace=ws)
    err_s = capi.sampled_error(A, B, C0, Cs, 1.0, 1.0, samples=CHECK_SAMPLES, seed=SEED)
    del ws, Cs
    split_tf = flops / split_ms / 1e9
    P = s * (s + 1) // 2

    log(f"bench: split-BF16 residues, N={nm}")
    wr = capi.CrtWorkspace(n, n, n, nm)
    Cr = C0.clone()
    crt_ms = time_loop(lambda: capi.dgemm_crt_bf16(A, B, Cr, nm, 1.0, 1.0, workspace=wr),
                       args.steps, args.warmup)
    wr.prepare(A, B)
    prod_ms = time_loop(wr.products, args.steps, args.warmup)
    Cr.copy_(C0)
    capi.dgemm_crt_bf16(A, B, Cr, nm, 1.0, 1.0, workspace=wr)
    err_r = capi.sampled_error(A, B, C0, Cr, 1.0, 1.0, samples=CHECK_SAMPLES, seed=SEED)
    del wr, Cr

    # the INT8 residue form (not the headline: the metric names BF16), at the
    # smallest moduli count that meets the same accuracy rule
    n8 = 16
    for cand in range(13, 17):
        wsc = capi.CrtWorkspace(n, n, n, cand, form="i8")
        Cc = C0.clone()
        capi.dgemm_crt_i8(A, B, Cc, cand, 1.0, 1.0, workspace=wsc)
        e = capi.sampled_error(A, B, C0, Cc, 1.0, 1.0, samples=CHECK_SAMPLES, seed=SEED)
        del wsc, Cc
        log(f"bench: int8 moduli={cand}: scaled {e['max_scaled']:.3e} rel {e['max_rel']:.3e}")
        if accurate(e):
            n8 = cand
            break
    log(f"bench: INT8 residues, N={n8}")
    w8 = capi.CrtWorkspace(n, n, n, n8, form="i8")
    C8 = C0.clone()
    i8_ms = time_loop(lambda: capi.dgemm_crt_i8(A, B, C8, n8, 1.0, 1.0, workspace=w8),
                      args.steps, args.warmup)
    w8.prepare(A, B)
    i8_prod_ms = time_loop(w8.products, args.steps, args.warmup)
    C8.copy_(C0)
    capi.dgemm_crt_i8(A, B, C8, n8, 1.0, 1.0, workspace=w8)
    err_8 = capi.sampled_error(A, B, C0, C8, 1.0, 1.0, samples=CHECK_SAMPLES, seed=SEED)
    del w8, C8, C
    crt_tf = flops / crt_ms / 1e9
    i8_tf = flops / i8_ms / 1e9

    use_crt = crt_tf >= split_tf
    if use_crt:
        value, step_ms = crt_tf, crt_ms
        workload = f"FP64-accurate GEMM C += A*B, m=n=k={n}, split-BF16 residues, {nm} moduli"
        dtype = "f64 in/out; bf16 residues x bf16 residues -> fp32 (exact, mod p), 128-bit CRT"
        # HBM bytes per launch: measured now by two rocprofv3 --pmc runs of a
        # child process (FETCH_SIZE, WRITE_SIZE, scaled on a 1 GiB copy); with
        # --no-traffic or when a run fails, the stored reading of
        # profiles/roofline_traffic.json if it was taken on these kernel
        # sources, else null
        digest = sources_digest("crt_bf16.hip")
        key = f"crt_multiply moduli={nm} m=n=k={n}"
        traffic, source = None, "not measured"
        if not args.no_traffic:
            try:
                t = measure_traffic(n, nm)
                traffic, source = t["traffic_bytes"], {"measured": "in this run", **t}
                if args.traffic_out:
                    out = Path(args.traffic_out)
                    table = json.loads(out.read_text()) if out.exists() else {}
                    table[key] = {**t, "sources_sha256": digest,
                                  "command": f"python bench.py --size {n} (moduli {nm} chosen)"}
                    out.write_text(json.dumps(table, indent=1) + "\n")
            except Exception as exc:  # the profiler is optional for the bench line
                log(f"bench: traffic measurement failed: {exc}")
                source = f"measurement failed: {exc}"
        if traffic is None:
            tpath = ROOT / "profiles" / "roofline_traffic.json"
            stored = json.loads(tpath.read_text()).get(key, {}) if tpath.exists() else {}
            if stored.get("sources_sha256") == digest:
                traffic = stored["traffic_bytes"]
                source = {"measured": "stored reading for these kernel sources",
                          **{k: v for k, v in stored.items() if k != "sources_sha256"}}
        roof = {"bound": "mfma", "achieved": round(nm * flops / prod_ms / 1e9, 2),
                "peak": PEAK_BF16, "unit": "TFLOP/s",
                "frac": round(nm * flops / prod_ms / 1e9 / PEAK_BF16, 4),
                "traffic": traffic, "traffic_source": source,
                # BF16 residues of A and B in, one 16-bit residue of C out, per modulus
                "algorithmic_bytes_per_launch": 2 * nm * 3 * n * n,
                "kernel": "crt_multiply (all moduli, one launch)", "ms_per_launch": round(prod_ms, 4),
                "algorithmic_flops_per_launch": nm * flops}
    else:
        value, step_ms = split_tf, split_ms
        workload = f"FP64-accurate GEMM C += A*B, m=n=k={n}, split-BF16 slices, {s} slices"
        dtype = "f64 in/out; bf16 slices x bf16 slices -> fp32 (exact), f64 fold"
        roof = {"bound": "mfma", "achieved": round(P * flops / mult_ms / 1e9, 2),
                "peak": PEAK_BF16, "unit": "TFLOP/s",
                "frac": round(P * flops / mult_ms / 1e9 / PEAK_BF16, 4), "traffic": None,
                "kernel": f"split_multiply<{s}>", "ms_per_launch": round(mult_ms, 4),
                "algorithmic_flops_per_launch": P * flops}

    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic: column-major f64 U[-1,1) from a fixed seed, resident in HBM",
        "config": {"workload": workload, "m": n, "n": n, "k": n,
                   "form": "residues" if use_crt else "slices",
                   "moduli" if use_crt else "slices": nm if use_crt else s},
        "roofline": roof,
        "speedup_vs_plain": round(value / plain_tf, 4),
        "crt": {"moduli": nm, "beta_bits": capi.crt_beta(n, nm), "gemms": nm,
                "tflops_fp64_equiv": round(crt_tf, 3), "ms": round(crt_ms, 4),
                "products_ms": round(prod_ms, 4),
                "frac_fp64_peak": round(crt_tf / PEAK_FP64, 4),
                "speedup_vs_plain": round(crt_tf / plain_tf, 4),
                "max_scaled_err": err_r["max_scaled"], "max_rel_err": err_r["max_rel"]},
        "crt_i8": {"moduli": n8, "beta_bits": capi.crt_beta(n, n8, "i8"), "gemms": n8,
                   "tflops_fp64_equiv": round(i8_tf, 3), "ms": round(i8_ms, 4),
                   "products_ms": round(i8_prod_ms, 4),
                   "products_tops": round(n8 * flops / i8_prod_ms / 1e9, 2),
                   "products_frac_int8_peak": round(n8 * flops / i8_prod_ms / 1e9 / PEAK_INT8, 4),
                   "frac_fp64_peak": round(i8_tf / PEAK_FP64, 4),
                   "speedup_vs_plain": round(i8_tf / plain_tf, 4),
                   "max_scaled_err": err_8["max_scaled"], "max_rel_err": err_8["max_rel"],
                   "note": "INT8 residues (v_mfma_i32_32x32x32_i8); reported beside the "
                           "headline, which the metric fixes to the split-BF16 forms"},
        "split": {"slices": s, "products": P, "tflops_fp64_equiv": round(split_tf, 3),
                  "ms": round(split_ms, 4), "multiply_ms": round(mult_ms, 4),
                  "frac_fp64_peak": round(split_tf / PEAK_FP64, 4),
                  "speedup_vs_plain": round(split_tf / plain_tf, 4),
                  "max_scaled_err": err_s["max_scaled"], "max_rel_err": err_s["max_rel"]},
        "plain": {"tflops": round(plain_tf, 3), "ms": round(plain_ms, 4),
                  "frac_fp64_peak": round(plain_tf / PEAK_FP64, 4),
                  "max_scaled_err": err_p["max_scaled"], "max_rel_err": err_p["max_rel"]},
        "accuracy_rule": f"max scaled and max relative error over {CHECK_SAMPLES} sampled "
                         "entries (double-double reference) within 2x plain FP64's",
    }
    # the same GPU's matrix-core rate with operands in registers, measured now:
    # the residue GEMMs against what the chip sustains at its power limit
    log("bench: matrix-core ceiling with operands in registers (BF16, INT8)")
    try:
        ceil = mfma_ceiling()
        bf_r, i8_r = ceil["bf16_32x32x16"]["random_tflops"], ceil["i8_32x32x32"]["random_tops"]
        ceil["crt_multiply_frac_of_bf16_random"] = round(nm * flops / prod_ms / 1e9 / bf_r, 4)
        ceil["crt_i8_products_frac_of_i8_random"] = round(n8 * flops / i8_prod_ms / 1e9 / i8_r, 4)
        ceil["note"] = ("MFMA chains with operands in registers on every CU (no memory traffic): "
                        "random operands give the rate the chip sustains at its power limit, "
                        "zero operands the full-clock rate")
        result["mfma_ceiling"] = ceil
    except Exception as exc:  # a reported reference point, not part of the metric
        log(f"bench: matrix-core ceiling not measured: {exc}")
        result["mfma_ceiling"] = {"error": str(exc)}
    # the vendor library's plain BF16 and INT8 GEMMs at the same size on the
    # same GPU: what its own kernels reach under the same power limit
    log("bench: vendor BF16 / INT8 GEMMs at the same size")
    v_bf, v_i8 = vendor_lowp_gemms(torch, gen, n, args.steps, args.warmup)
    result["vendor_reference"] = {
        "bf16_matmul": {"ms": round(v_bf, 4), "tflops": round(flops / v_bf / 1e9, 2),
                        "frac_peak": round(flops / v_bf / 1e9 / PEAK_BF16, 4)},
        "int8_int_mm": {"ms": round(v_i8, 4), "tops": round(flops / v_i8 / 1e9, 2),
                        "frac_peak": round(flops / v_i8 / 1e9 / PEAK_INT8, 4)},
        "crt_multiply_rate_vs_bf16_matmul": round(nm * flops / prod_ms / (flops / v_bf), 4),
        "crt_i8_products_rate_vs_int_mm": round(n8 * flops / i8_prod_ms / (flops / v_i8), 4),
        "note": "torch.matmul on bf16 and torch._int_mm on int8, m=n=k as the workload; "
                "rates are per GEMM, the residue kernels' per modulus"}
    cpu_inputs, c1_gpu_tf = None, None
    if not args.no_cpu:
        # host copies for the CPU leg: the workload's A, B and C0 whole (A, B,
        # C0 are transposed views of row-major tensors, so .T is contiguous;
        # ncols below n would time a column sample of the same product)
        ncols = n
        cpu_inputs = [A.T.cpu().numpy().T, B.T[:ncols].cpu().numpy().T,
                      C0.T[:ncols].cpu().numpy().T]
        # configuration 1 (m=n=k=1024): drawn on the GPU, timed there too
        g1 = torch.Generator(device="cuda").manual_seed(SEED & 0xFFFFFFFF)
        A1, B1, C1 = rand_colmajor(1024, 1024, g1), rand_colmajor(1024, 1024, g1), \
            rand_colmajor(1024, 1024, g1)
        cpu_inputs += [A1.T.cpu().numpy().T, B1.T.cpu().numpy().T, C1.T.cpu().numpy().T]
        C1w = C1.clone()
        c1_ms = time_loop(lambda: capi.dgemm(A1, B1, C1w, 1.0, 1.0), 20, 3)
        c1_gpu_tf = 2.0 * 1024 ** 3 / c1_ms / 1e9
        del A1, B1, C1, C1w
    del A, B, C0
    torch.cuda.empty_cache()

    if not args.quick:
        cfg = {}
        log("bench: config 2 sweep (plain FP64)")
        cfg["plain_sweep"] = sweep_plain(capi, torch, gen, 20)
        log("bench: variant control trees as tile orders")
        cfg["variant_trees"] = variant_trees(capi, torch, gen, 20)
        log("bench: config 3 grid (split-BF16: size x spread x slices)")
        cfg["split_grid"] = split_grid(capi, torch, gen, 10)
        log("bench: config 3 grid (split-BF16 residues: size x spread x moduli)")
        cfg["crt_grid"] = crt_grid(capi, torch, gen, 10)
        log("bench: config 3 grid (INT8 residues: size x spread x moduli)")
        cfg["crt_i8_grid"] = crt_grid(capi, torch, gen, 10, moduli=(12, 14, 15, 16), form="i8")
        log("bench: config 4 (FP32, BF16->FP32)")
        cfg["reduced_precision"] = reduced_precision(capi, torch, gen, 20)
        log("bench: config 5 (4096 uneven GEMMs)")
        cfg["batched"] = batched(capi, torch, gen, 20)
        log("bench: hierarchical layout pack / unpack-add at 8192^2")
        cfg["hier_layout"] = hier(capi, torch, gen, 20)
        result["configs"] = cfg
        torch.cuda.empty_cache()

    if not args.no_cpu:
        info = host_cpus()
        # one thread per physical core the process may use, within its cgroup share
        threads = info["physical_cores"]
        if info["cpu_share"]:
            threads = min(threads, info["cpu_share"])
        threads = args.cpu_threads or max(1, threads)
        log(f"bench: CPU baseline on {threads} threads ({info['model']}, "
            f"{info['physical_cores']} physical cores, share {info['cpu_share']})")
        cb = cpu_baseline(*cpu_inputs, threads, info)
        cb["config1_1024cube"]["gpu_tflops"] = round(c1_gpu_tf, 3)
        result["cpu_baseline"] = cb

    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
