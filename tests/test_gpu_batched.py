"""Batched uneven GEMMs in one persistent launch (and one hipGraph launch).

Bars: plain FP64 batch within gamma_{k+1} of the double-double oracle, and so
within 2 gamma_{k+1} of the loop of per-problem momms_hip_dgemm launches
(the two take different summation orders: one-wavefront 32 x 32 tiles,
128 x 128 tiles, split k); the same batch twice gives the same bits;
split-BF16 batch and residue batch == the oracle's restatements bit for bit;
graph replays == eager runs.
"""
import numpy as np
import pytest
import torch

import pyoracle as O
from conftest import SEED, gamma
from momms_amd import capi

pytestmark = pytest.mark.gpu


def arenas(rows, a_len, b_len, c_len, seed=SEED):
    g = torch.Generator(device="cuda").manual_seed(seed & 0xFFFFFFFF)
    A = torch.rand(a_len, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    B = torch.rand(b_len, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    C = torch.rand(c_len, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    return A, B, C


def view(arena, off, h, w, ld):
    return arena[off: off + ld * w].view(w, ld)[:, :h].T if h and w else np.zeros((h, w))


def np_view(arena, off, h, w, ld):
    if h == 0 or w == 0:
        return np.zeros((h, w), order="F")
    return np.asfortranarray(arena[off: off + ld * w].reshape(w, ld)[:, :h].T)


def edge_rows():
    """A table with ragged, degenerate and padded-leading-dimension problems."""
    shapes = [(1, 1, 1), (17, 33, 5), (64, 64, 64), (65, 127, 200), (300, 7, 513), (0, 5, 5),
              (5, 0, 5), (9, 9, 0), (129, 260, 31), (512, 512, 16), (16, 512, 512)]
    rows, ao, bo, co = [], 0, 0, 0
    for i, (m, n, k) in enumerate(shapes):
        lda, ldb, ldc = m + (i % 3), k + 1, m + 2
        rows.append((m, n, k, ao, bo, co, max(lda, 1), max(ldb, 1), max(ldc, 1)))
        ao += max(lda, 1) * k
        bo += max(ldb, 1) * n
        co += max(ldc, 1) * n
    return rows, ao, bo, co


@pytest.mark.parametrize("table", ["edge", "uneven"])
@pytest.mark.parametrize("autotune", [False, True])
def test_plain_batch_within_gamma_of_oracle(table, autotune):
    rows, al, bl, cl = edge_rows() if table == "edge" else capi.uneven_batch(300, 16, 300)
    A, B, C0 = arenas(rows, al, bl, cl)
    plan = capi.BatchPlan(rows, mode=0, autotune=autotune, A=A, B=B)
    info = plan.describe()
    assert info["mode"] == "fp64"
    C1 = C0.clone()
    plan.run(A, B, C1, 1.0, 1.0)
    C3 = C0.clone()
    plan.run(A, B, C3, 1.0, 1.0)
    C2 = C0.clone()
    capi.dgemm_batched_loop(rows, A, B, C2, 1.0, 1.0)
    torch.cuda.synchronize()
    assert torch.equal(C1, C3)                     # deterministic
    a, b, c0, c1, c2 = (x.cpu().numpy() for x in (A, B, C0, C1, C2))
    for (m, n, k, ao, bo, co, lda, ldb, ldc) in rows[:80]:
        if m == 0 or n == 0:
            continue
        Am, Bm, Cm0 = np_view(a, ao, m, k, lda), np_view(b, bo, k, n, ldb), np_view(c0, co, m, n, ldc)
        ref, mag = O.dd_gemm(Am, Bm, Cm0)
        for cc in (c1, c2):
            assert (np.abs(np_view(cc, co, m, n, ldc) - ref) <= gamma(k + 1) * mag).all(), (m, n, k)
    # padding between problems (ldc > m) is untouched
    untouched = np.ones(cl, dtype=bool)
    for (m, n, k, ao, bo, co, lda, ldb, ldc) in rows:
        for j in range(n):
            untouched[co + j * ldc: co + j * ldc + m] = False
    assert np.array_equal(c1[untouched], c0[untouched])
    plan.close()


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_batch_reads_only_its_problems(mode):
    """Arenas whose padding (lda > m, ldb > k, gaps) is NaN and whose last
    problem ends at the last element: shapes where tiles reach past k (k < 16,
    k not a multiple of 16, one-wavefront tiles at k = 512), past m and n, and
    k = 0. A masked-out element that leaked into a sum would make it NaN."""
    shapes = [(100, 120, 5), (30, 40, 37), (200, 96, 3), (32, 32, 512), (17, 300, 0), (96, 100, 1),
              (5, 7, 9), (32, 32, 512)]
    rows, ao, bo, co = [], 0, 0, 0
    for i, (m, n, k) in enumerate(shapes):
        last = i == len(shapes) - 1
        lda, ldb = m + (0 if last else 3), max(k + (0 if last else 2), 1)
        rows.append((m, n, k, ao + 1, bo + 1, co, lda, ldb, m))  # one NaN before each
        ao += 1 + lda * k
        bo += 1 + ldb * n
        co += m * n
    al, bl = ao, bo
    g = torch.Generator(device="cuda").manual_seed(SEED + 31 + mode)
    A = torch.full((al,), float("nan"), device="cuda", dtype=torch.float64)
    B = torch.full((bl,), float("nan"), device="cuda", dtype=torch.float64)
    C0 = torch.rand(co, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    for (m, n, k, a0, b0, c0_, lda, ldb, ldc) in rows:
        if k:
            view(A, a0, m, k, lda).copy_(torch.rand(m, k, generator=g, device="cuda", dtype=torch.float64) - 0.5)
        if k and n:
            view(B, b0, k, n, ldb).copy_(torch.rand(k, n, generator=g, device="cuda", dtype=torch.float64) - 0.5)
    assert rows[-1][3] + rows[-1][6] * rows[-1][2] == al and rows[-1][4] + rows[-1][7] * rows[-1][1] == bl
    plan = capi.BatchPlan(rows, mode=mode, slices={0: 0, 1: 7, 2: 13}[mode], autotune=mode == 0,
                          A=A, B=B)
    C = C0.clone()
    plan.run(A, B, C, 1.0, 1.0)
    torch.cuda.synchronize()
    a, b, c0, c = (x.cpu().numpy() for x in (A, B, C0, C))
    assert np.isfinite(c).all()
    for (m, n, k, a0, b0, c0_, lda, ldb, ldc) in rows:
        Am, Bm, Cm0 = np_view(a, a0, m, k, lda), np_view(b, b0, k, n, ldb), np_view(c0, c0_, m, n, ldc)
        got = np_view(c, c0_, m, n, ldc)
        if mode == 0:
            ref, mag = O.dd_gemm(Am, Bm, Cm0)
            assert (np.abs(got - ref) <= gamma(k + 1) * mag).all(), (m, n, k)
        else:
            want = Cm0.copy(order="F")
            if mode == 1:
                O.ozaki_gemm(Am, Bm, want, 7, 1.0, 1.0)
            else:
                O.crt_gemm(Am, Bm, want, 13, 1.0, 1.0)
            assert np.array_equal(got.view(np.int64), want.view(np.int64)), (m, n, k)
    plan.close()


def test_batch_wide_leading_dimensions():
    """Problems with lda, ldb or ldc above 2^21 (the 64-bit-addressed tiles)
    in one batch with ordinary ones; padding is NaN."""
    W = (1 << 21) + 5
    specs = [(70, 50, 3, W, 3, 70), (40, 3, 20, 40, W, 40), (33, 2, 17, 33, 17, W),
             (64, 64, 64, 64, 64, 64), (20, 30, 40, 21, 41, 22), (9, 9, 0, 9, 1, 9)]
    rows, ao, bo, co = [], 0, 0, 0
    for (m, n, k, lda, ldb, ldc) in specs:
        rows.append((m, n, k, ao, bo, co, lda, ldb, ldc))
        ao += lda * k
        bo += ldb * n
        co += ldc * n
    g = torch.Generator(device="cuda").manual_seed(SEED + 77)
    A = torch.full((ao,), float("nan"), device="cuda", dtype=torch.float64)
    B = torch.full((bo,), float("nan"), device="cuda", dtype=torch.float64)
    C0 = torch.full((co,), float("nan"), device="cuda", dtype=torch.float64)
    for (m, n, k, a0, b0, c0_, lda, ldb, ldc) in rows:
        if k:
            view(A, a0, m, k, lda).copy_(torch.rand(m, k, generator=g, device="cuda", dtype=torch.float64) - 0.5)
            view(B, b0, k, n, ldb).copy_(torch.rand(k, n, generator=g, device="cuda", dtype=torch.float64) - 0.5)
        view(C0, c0_, m, n, ldc).copy_(torch.rand(m, n, generator=g, device="cuda", dtype=torch.float64) - 0.5)
    plan = capi.BatchPlan(rows, mode=0, autotune=False)
    C = C0.clone()
    plan.run(A, B, C, 1.0, 1.0)
    torch.cuda.synchronize()
    for (m, n, k, a0, b0, c0_, lda, ldb, ldc) in rows:
        Am = view(A, a0, m, k, lda).cpu().numpy() if k else np.zeros((m, 0))
        Bm = view(B, b0, k, n, ldb).cpu().numpy() if k else np.zeros((0, n))
        Cm0 = view(C0, c0_, m, n, ldc).cpu().numpy()
        got = view(C, c0_, m, n, ldc).cpu().numpy()
        ref, mag = O.dd_gemm(np.asfortranarray(Am), np.asfortranarray(Bm), np.asfortranarray(Cm0))
        assert (np.abs(got - ref) <= gamma(k + 1) * mag).all(), (m, n, k, lda, ldb, ldc)
    # nothing outside the problems was written
    written = torch.zeros(co, dtype=torch.bool, device="cuda")
    for (m, n, k, a0, b0, c0_, lda, ldb, ldc) in rows:
        view(written, c0_, m, n, ldc).fill_(True)
    assert torch.isnan(C[~written]).all()
    plan.close()


@pytest.mark.parametrize("s", [3, 7])
def test_split_batch_bitwise_against_restatement(s):
    rows, al, bl, cl = edge_rows()
    A, B, C0 = arenas(rows, al, bl, cl, seed=SEED + s)
    plan = capi.BatchPlan(rows, mode=1, slices=s)
    assert plan.describe()["slices"] == s
    C = C0.clone()
    plan.run(A, B, C, -0.5, 1.5)
    torch.cuda.synchronize()
    a, b, c0, c = (x.cpu().numpy() for x in (A, B, C0, C))
    for (m, n, k, ao, bo, co, lda, ldb, ldc) in rows:
        if m == 0 or n == 0:
            continue
        want = np_view(c0, co, m, n, ldc)
        O.ozaki_gemm(np_view(a, ao, m, k, lda), np_view(b, bo, k, n, ldb), want, s, -0.5, 1.5)
        got = np_view(c, co, m, n, ldc)
        assert np.array_equal(got.view(np.int64), want.view(np.int64)), (m, n, k)
    plan.close()


@pytest.mark.parametrize("moduli", [9, 13, 16])
def test_crt_batch_bitwise_against_restatement(moduli):
    """The residue batch equals the CPU restatement of the residue form
    (oracle/crt.c) problem by problem, bit for bit: ragged and degenerate
    shapes, padded leading dimensions, k = 0, k > 256 (the mid-loop modular
    reduction), alpha and beta other than 1, and rows of A / columns of B
    holding Inf and NaN (their entries take the fma-chain path)."""
    rows, al, bl, cl = edge_rows()
    A, B, C0 = arenas(rows, al, bl, cl, seed=SEED + moduli)
    # Inf in row 3 of problem 3's A, NaN in column 5 of problem 4's B
    m3, n3, k3, ao3, _, _, lda3, _, _ = rows[3]
    A[ao3 + 3 + 7 * lda3] = float("inf")
    m4, n4, k4, _, bo4, _, _, ldb4, _ = rows[4]
    B[bo4 + 5 * ldb4 + 11] = float("nan")
    plan = capi.BatchPlan(rows, mode=2, slices=moduli)
    info = plan.describe()
    assert info["mode"] == "crt_bf16" and info["moduli"] == moduli
    C = C0.clone()
    plan.run(A, B, C, -0.5, 1.5)
    torch.cuda.synchronize()
    a, b, c0, c = (x.cpu().numpy() for x in (A, B, C0, C))
    for (m, n, k, ao, bo, co, lda, ldb, ldc) in rows:
        if m == 0 or n == 0:
            continue
        want = np_view(c0, co, m, n, ldc)
        O.crt_gemm(np_view(a, ao, m, k, lda), np_view(b, bo, k, n, ldb), want, moduli, -0.5, 1.5)
        got = np_view(c, co, m, n, ldc)
        assert np.array_equal(got.view(np.int64), want.view(np.int64)), (m, n, k)
    # the fma-chain entries are there: non-finite in problem 3's row 3
    assert not np.isfinite(np_view(c, rows[3][5], m3, n3, rows[3][8])[3]).all()
    plan.close()


def test_crt_batch_beta_zero_never_reads_c():
    rows, al, bl, cl = edge_rows()
    A, B, _ = arenas(rows, al, bl, cl, seed=SEED + 77)
    C = torch.full((cl,), float("nan"), device="cuda", dtype=torch.float64)
    plan = capi.BatchPlan(rows, mode=2, slices=13)
    plan.run(A, B, C, 1.0, 0.0)
    torch.cuda.synchronize()
    a, b, c = (x.cpu().numpy() for x in (A, B, C))
    for (m, n, k, ao, bo, co, lda, ldb, ldc) in rows:
        if m == 0 or n == 0:
            continue
        want = np.zeros((m, n), order="F")
        O.crt_gemm(np_view(a, ao, m, k, lda), np_view(b, bo, k, n, ldb), want, 13, 1.0, 0.0)
        assert np.array_equal(np_view(c, co, m, n, ldc).view(np.int64), want.view(np.int64)), (m, n, k)
    plan.close()


def test_crt_moduli_autotune_meets_accuracy_and_is_remembered():
    """The autotuner picks the smallest moduli count whose sampled error is
    within 2x plain FP64's; the choice lands in the tuning table under the
    batch's class, and a later plan without autotuning takes it from there."""
    rows, al, bl, cl = capi.uneven_batch(64, 16, 256, seed=SEED + 19)
    A, B, C0 = arenas(rows, al, bl, cl)
    saved = capi.tuning_export()
    try:
        capi.tuning_import({"version": 1, "entries": []}, replace=True)
        plan = capi.BatchPlan(rows, mode=2, slices=0, autotune=True, A=A, B=B)
        N = plan.describe()["moduli"]
        assert 8 <= N <= 16
        entry = {(e["kind"], e["key"]): e["value"] for e in capi.tuning_export()["entries"]}
        assert entry[("batch_crt", capi.batch_class(rows))] == [N, 0, 0, 0]
        C = C0.clone()
        plan.run(A, B, C, 1.0, 0.0)
        torch.cuda.synchronize()
        a, b, c = (x.cpu().numpy() for x in (A, B, C))
        worst = 0.0
        for (m, n, k, ao, bo, co, lda, ldb, ldc) in rows[:20]:
            ref, mag = O.dd_gemm(np_view(a, ao, m, k, lda), np_view(b, bo, k, n, ldb),
                                 np.zeros((m, n), order="F"), 1.0, 0.0)
            worst = max(worst, float(np.max(np.abs(np_view(c, co, m, n, ldc) - ref) / mag)))
        assert worst <= 4 * gamma(256)
        plan.close()
        again = capi.BatchPlan(rows, mode=2, slices=0)
        assert again.describe()["moduli"] == N
        again.close()
    finally:
        capi.tuning_import(saved, replace=True)


def test_tuned_plain_thresholds_are_used_from_an_imported_table():
    rows, al, bl, cl = capi.uneven_batch(128, 16, 300, seed=SEED + 23)
    A, B, C0 = arenas(rows, al, bl, cl)
    key = capi.batch_class(rows)
    saved = capi.tuning_export()
    try:
        outs = []
        for T, T2 in [(16, 1 << 30), (16, 16)]:
            capi.tuning_import({"version": 1, "entries": [
                {"kind": "batch_fp64", "key": key, "value": [T, T2, 0, 0]}]}, replace=True)
            plan = capi.BatchPlan(rows, mode=0)
            info = plan.describe()
            assert (info["threshold"], info["threshold_128"]) == (T, T2)
            assert info["tiles_32x32_wavefront"] == 0
            assert (info["tiles_128x128"] == 0) == (T2 == 1 << 30)
            C = C0.clone()
            plan.run(A, B, C, 1.0, 1.0)
            outs.append(C)
            plan.close()
        torch.cuda.synchronize()
        a, b, c0 = (x.cpu().numpy() for x in (A, B, C0))
        for C in outs:
            c = C.cpu().numpy()
            for (m, n, k, ao, bo, co, lda, ldb, ldc) in rows[:40]:
                ref, mag = O.dd_gemm(np_view(a, ao, m, k, lda), np_view(b, bo, k, n, ldb),
                                     np_view(c0, co, m, n, ldc))
                assert (np.abs(np_view(c, co, m, n, ldc) - ref) <= gamma(k + 1) * mag).all()
    finally:
        capi.tuning_import(saved, replace=True)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_graph_replays_match_eager(mode):
    rows, al, bl, cl = capi.uneven_batch(200, 16, 256, seed=SEED + mode)
    A, B, C0 = arenas(rows, al, bl, cl)
    plan = capi.BatchPlan(rows, mode=mode, slices=12 if mode == 2 else 5)
    Ce = C0.clone()
    plan.run(A, B, Ce, 1.0, 1.0)
    Cg = C0.clone()
    plan.capture(A, B, Cg, 1.0, 1.0)
    for rep in range(3):
        Cg.copy_(C0)                      # eager work between replays
        torch.cuda.synchronize()
        plan.launch()
        torch.cuda.synchronize()
        assert torch.equal(Cg, Ce), f"replay {rep}"
    plan.close()


def test_split_slice_autotune_meets_accuracy():
    rows, al, bl, cl = capi.uneven_batch(64, 16, 256, seed=SEED + 9)
    A, B, C0 = arenas(rows, al, bl, cl)
    plan = capi.BatchPlan(rows, mode=1, slices=0, autotune=True, A=A, B=B)
    s = plan.describe()["slices"]
    assert 4 <= s <= 8
    C = C0.clone()
    plan.run(A, B, C, 1.0, 0.0)
    torch.cuda.synchronize()
    a, b, c = (x.cpu().numpy() for x in (A, B, C))
    worst = 0.0
    for (m, n, k, ao, bo, co, lda, ldb, ldc) in rows[:20]:
        ref, mag = O.dd_gemm(np_view(a, ao, m, k, lda), np_view(b, bo, k, n, ldb),
                             np.zeros((m, n), order="F"), 1.0, 0.0)
        worst = max(worst, float(np.max(np.abs(np_view(c, co, m, n, ldc) - ref) / mag)))
    assert worst <= 4 * gamma(512)
    plan.close()


def test_config5_full_batch_against_loop():
    """4096 problems, m, n, k log-uniform in [16, 512]: the captured batch and
    the loop of per-problem launches agree within 2 gamma_{k+1} (|A||B| + |C0|)
    in every entry (each is within gamma_{k+1} of the exact value)."""
    rows, al, bl, cl = capi.uneven_batch(4096)
    A, B, C0 = arenas(rows, al, bl, cl)
    plan = capi.BatchPlan(rows, mode=0, autotune=True, A=A, B=B)
    C1 = C0.clone()
    plan.capture(A, B, C1, 1.0, 1.0)
    plan.launch()
    C2 = C0.clone()
    capi.dgemm_batched_loop(rows, A, B, C2, 1.0, 1.0)
    torch.cuda.synchronize()
    worst = 0.0
    for (m, n, k, ao, bo, co, lda, ldb, ldc) in rows:
        Am = A[ao: ao + lda * k].view(k, lda)[:, :m].T
        Bm = B[bo: bo + ldb * n].view(n, ldb)[:, :k].T
        Cm0 = C0[co: co + ldc * n].view(n, ldc)[:, :m].T
        mag = Am.abs().contiguous() @ Bm.abs().contiguous() + Cm0.abs()
        d = (C1[co: co + ldc * n].view(n, ldc)[:, :m].T - C2[co: co + ldc * n].view(n, ldc)[:, :m].T).abs()
        r = float((d / (2 * gamma(k + 1) * mag)).max())
        worst = max(worst, r)
    assert worst <= 1.0, worst
    plan.close()


def test_config5_split_graph_bitwise_on_sampled_problems():
    """The 4096-problem uneven batch through the split-BF16 plan at 7 slices,
    captured and launched as one graph (the form the batched benchmark times):
    64 problems drawn with a fixed seed equal the oracle's restatement bit for
    bit, and the whole output is finite."""
    rows, al, bl, cl = capi.uneven_batch(4096)
    A, B, C0 = arenas(rows, al, bl, cl, seed=SEED + 5)
    plan = capi.BatchPlan(rows, mode=1, slices=7)
    assert plan.describe()["slices"] == 7
    C = C0.clone()
    plan.capture(A, B, C, 1.0, 1.0)
    plan.launch()
    torch.cuda.synchronize()
    assert torch.isfinite(C).all()
    pick = np.random.default_rng(SEED + 4096).choice(len(rows), 64, replace=False)
    for i in sorted(pick):
        m, n, k, ao, bo, co, lda, ldb, ldc = rows[i]
        a = view(A, ao, m, k, lda).cpu().numpy()
        b = view(B, bo, k, n, ldb).cpu().numpy()
        want = np.asfortranarray(view(C0, co, m, n, ldc).cpu().numpy())
        O.ozaki_gemm(np.asfortranarray(a), np.asfortranarray(b), want, 7, 1.0, 1.0)
        got = np.asfortranarray(view(C, co, m, n, ldc).cpu().numpy())
        assert np.array_equal(got.view(np.int64), want.view(np.int64)), (i, m, n, k)
    plan.close()


def test_plan_rejects_bad_tables():
    with pytest.raises(capi.MommsError):
        capi.BatchPlan([(4, 4, 4, 0, 0, 0, 2, 4, 4)])   # lda < m
    with pytest.raises(capi.MommsError):
        capi.BatchPlan([(4, 4, 4, 0, 0, 0, 4, 4, 4)], mode=3)
    # one modulus cannot hold the products of a k = 4096 problem
    assert O.crt_beta(4096, 1) < 1
    with pytest.raises(capi.MommsError, match="too few moduli"):
        capi.BatchPlan([(4, 4, 4096, 0, 0, 0, 4, 4096, 4)], mode=2, slices=1)


def test_config5_crt_graph_bitwise_on_sampled_problems():
    """The 4096-problem uneven batch through the residue plan at 13 moduli,
    captured and launched as one graph: 64 problems drawn with a fixed seed
    equal the oracle's restatement bit for bit, and the whole output is
    finite."""
    rows, al, bl, cl = capi.uneven_batch(4096)
    A, B, C0 = arenas(rows, al, bl, cl, seed=SEED + 13)
    plan = capi.BatchPlan(rows, mode=2, slices=13)
    assert plan.describe()["moduli"] == 13
    C = C0.clone()
    plan.capture(A, B, C, 1.0, 1.0)
    plan.launch()
    torch.cuda.synchronize()
    assert torch.isfinite(C).all()
    pick = np.random.default_rng(SEED + 4097).choice(len(rows), 64, replace=False)
    for i in sorted(pick):
        m, n, k, ao, bo, co, lda, ldb, ldc = rows[i]
        a = view(A, ao, m, k, lda).cpu().numpy()
        b = view(B, bo, k, n, ldb).cpu().numpy()
        want = np.asfortranarray(view(C0, co, m, n, ldc).cpu().numpy())
        O.crt_gemm(np.asfortranarray(a), np.asfortranarray(b), want, 13, 1.0, 1.0)
        got = np.asfortranarray(view(C, co, m, n, ldc).cpu().numpy())
        assert np.array_equal(got.view(np.int64), want.view(np.int64)), (i, m, n, k)
    plan.close()
