"""CPU checks of the C ABI: the library loads, exports what include/momms_hip.h
declares, and rejects invalid arguments before touching the GPU."""
import ctypes

import pytest

from momms_amd import capi


def test_library_exports_every_declared_symbol():
    lib = capi.lib()
    names = capi.declared_symbols()
    assert len(names) >= 3
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    # every declared function has argtypes in the ctypes layer
    assert set(capi.declared_symbols()) <= set(capi._SIGS)


D = ctypes.c_double


def _dgemm(m, n, k, rsA=1, csA=1, rsB=1, csB=1, rsC=1, csC=1, A=8, B=8, C=8, alpha=True,
           beta=True):
    al, be = D(1.0), D(1.0)
    return capi.lib().momms_hip_dgemm(m, n, k, ctypes.byref(al) if alpha else None, A, rsA, csA,
                                      B, rsB, csB, ctypes.byref(be) if beta else None, C, rsC,
                                      csC, None)


@pytest.mark.parametrize("kw,code", [
    (dict(m=-1, n=1, k=1), -1),
    (dict(m=1, n=-1, k=1), -2),
    (dict(m=1, n=1, k=-1), -3),
    (dict(m=1, n=1, k=1, alpha=False), -4),
    (dict(m=1, n=1, k=1, A=None), -5),
    (dict(m=1, n=1, k=1, rsA=0), -6),
    (dict(m=1, n=1, k=1, csA=0), -7),
    (dict(m=1, n=1, k=1, B=None), -8),
    (dict(m=1, n=1, k=1, csB=-3), -10),
    (dict(m=1, n=1, k=1, beta=False), -11),
    (dict(m=1, n=1, k=1, C=None), -12),
    (dict(m=1, n=1, k=1, rsC=0), -13),
])
def test_dgemm_argument_errors(kw, code):
    st = _dgemm(**kw)
    assert st == code
    assert "momms_hip_dgemm" in capi.last_error()


def test_dgemm_empty_is_a_no_op():
    # m == 0 or n == 0 returns before any GPU work, with null matrices allowed
    assert _dgemm(0, 5, 3, A=None, C=None) == 0
    assert _dgemm(5, 0, 3, B=None, C=None) == 0


def test_dgemm_tiled_rejects_unknown_tile():
    al = D(1.0)
    st = capi.lib().momms_hip_dgemm_tiled(1, 1, 1, ctypes.byref(al), 8, 1, 1, 8, 1, 1,
                                          ctypes.byref(al), 8, 1, 1, 99, None)
    assert st == -15


def test_batch_plan_rejects_bad_tables_before_gpu_work():
    h = ctypes.c_void_p()
    bad = capi.desc_array([(4, 4, 4, 0, 0, 0, 2, 4, 4)])  # lda < m
    st = capi.lib().momms_hip_batch_plan_create(ctypes.cast(bad, ctypes.c_void_p), 1, 0, 0, 0,
                                                None, None, ctypes.byref(h), None)
    assert st == -1 and "descs[0]" in capi.last_error()
    good = capi.desc_array([(4, 4, 4, 0, 0, 0, 4, 4, 4)])
    st = capi.lib().momms_hip_batch_plan_create(ctypes.cast(good, ctypes.c_void_p), 1, 3, 0, 0,
                                                None, None, ctypes.byref(h), None)
    assert st == -3
    st = capi.lib().momms_hip_batch_plan_create(ctypes.cast(good, ctypes.c_void_p), 1, 1, 9, 0,
                                                None, None, ctypes.byref(h), None)
    assert st == -4
    st = capi.lib().momms_hip_batch_plan_create(ctypes.cast(good, ctypes.c_void_p), 1, 2, 17, 0,
                                                None, None, ctypes.byref(h), None)
    assert st == -4 and "moduli" in capi.last_error()
    wide = capi.desc_array([(4, 4, 4, 0, 0, 0, (1 << 21) + 1, 4, 4)])
    st = capi.lib().momms_hip_batch_plan_create(ctypes.cast(wide, ctypes.c_void_p), 1, 2, 13, 0,
                                                None, None, ctypes.byref(h), None)
    assert st == -1 and "2^21" in capi.last_error()
    st = capi.lib().momms_hip_batch_plan_create(ctypes.cast(good, ctypes.c_void_p), 1, 0, 0, 1,
                                                None, None, ctypes.byref(h), None)
    assert st == -6


def test_split_entry_points_validate_workspace():
    al = D(1.0)
    size = capi.lib().momms_hip_split_bf16_workspace_size(100, 50, 30, 7)
    assert size > 7 * 128 * 32 * 2
    assert capi.lib().momms_hip_split_bf16_workspace_size(100, 50, 30, 9) == 0
    st = capi.lib().momms_hip_dgemm_split_bf16(100, 50, 30, ctypes.byref(al), 8, 1, 100, 8, 1, 30,
                                               ctypes.byref(al), 8, 1, 100, 9, 8, size, None)
    assert st == -15
    st = capi.lib().momms_hip_dgemm_split_bf16(100, 50, 30, ctypes.byref(al), 8, 1, 100, 8, 1, 30,
                                               ctypes.byref(al), 8, 1, 100, 7, 8, size - 1, None)
    assert st == -17
    st = capi.lib().momms_hip_split_bf16_multiply(100, 50, 30, ctypes.byref(al), ctypes.byref(al),
                                                  8, 1, 100, 7, None, size, None)
    assert st == -10


def test_sample_index_is_splitmix64():
    # SplitMix64 (Steele, Lea, Flood 2014) reference values for seed 0
    def sm(x):
        x = (x + 0x9E3779B97F4A7C15) & (2**64 - 1)
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
        return x ^ (x >> 31)
    assert sm(0) == 0xE220A8397B1DCDAF     # published first output of SplitMix64(0)
    for e in range(5):
        assert capi.lib().momms_sample_index(7, e, 0, 1000) == sm(7 + 2 * e) % 1000
        assert capi.lib().momms_sample_index(7, e, 1, 999) == sm(7 + 2 * e + 1) % 999


def test_crt_constants_match_the_oracle():
    import pyoracle as O
    assert capi.crt_moduli(16) == O.crt_moduli(16)
    assert capi.lib().momms_hip_crt_bf16_modulus(16) == 0
    for k in (0, 1, 2, 17, 300, 4096, 4097, 8192, 65536, 1 << 20, 1 << 26):
        for N in range(1, 17):
            assert capi.crt_beta(k, N) == O.crt_beta(k, N), (k, N)
    assert capi.crt_beta(10, 0) == -1 and capi.crt_beta(10, 17) == -1


def test_crt_entry_point_validates_arguments():
    al = D(1.0)
    size = capi.lib().momms_hip_crt_bf16_workspace_size(100, 50, 30, 14)
    assert size >= 14 * 256 * 32 * 2 * 2
    assert capi.lib().momms_hip_crt_bf16_workspace_size(100, 50, 30, 17) == 0
    assert capi.lib().momms_hip_crt_bf16_workspace_size(100, 50, 8192, 1) == 0  # too few moduli

    def call(moduli=14, ws=8, nbytes=size, k=30):
        return capi.lib().momms_hip_dgemm_crt_bf16(100, 50, k, ctypes.byref(al), 8, 1, 100, 8, 1,
                                                   k, ctypes.byref(al), 8, 1, 100, moduli, ws,
                                                   nbytes, None)
    assert call(moduli=0) == -15 and "moduli" in capi.last_error()
    assert call(moduli=17) == -15
    assert call(moduli=1, k=8192) == -15 and "too few" in capi.last_error()
    assert call(ws=None) == -16
    assert call(nbytes=size - 1) == -17


def test_crt_staged_entry_points_validate_arguments():
    al = D(1.0)
    size = capi.lib().momms_hip_crt_bf16_workspace_size(100, 50, 30, 14)
    L = capi.lib()
    assert L.momms_hip_crt_bf16_prepare(100, 50, 30, 8, 1, 100, 8, 1, 30, 0, 8, size, None) == -10
    assert L.momms_hip_crt_bf16_prepare(100, 50, 30, 8, 1, 100, 8, 1, 30, 14, None, size,
                                        None) == -11
    assert L.momms_hip_crt_bf16_prepare(100, 50, 30, None, 1, 100, 8, 1, 30, 14, 8, size,
                                        None) == -4
    assert L.momms_hip_crt_bf16_products(100, 50, 30, 17, 8, size, None) == -4
    assert L.momms_hip_crt_bf16_products(100, 50, 30, 14, 8, size - 1, None) == -6
    assert L.momms_hip_crt_bf16_reconstruct(100, 50, 30, ctypes.byref(al), 8, 1, 100, 8, 1, 30,
                                            None, 8, 1, 100, 14, 8, size, None) == -11
    assert L.momms_hip_crt_bf16_reconstruct(100, 50, 30, ctypes.byref(al), 8, 1, 100, 8, 1, 30,
                                            ctypes.byref(al), 8, 1, 100, 14, 8, 10, None) == -17
    assert "momms_hip_crt_bf16_reconstruct" in capi.last_error()


def test_wrapper_rejects_broadcast_and_overlapping_views():
    """A stride of 0 along a dimension longer than 1 (torch.expand) would make
    the kernels read past the tensor's storage, and an output whose elements
    overlap would have two places written at once: both are refused before
    any GPU work. A stride along a dimension of extent 1 is free."""
    import torch

    row = torch.zeros(4, dtype=torch.float64)
    with pytest.raises(ValueError, match="broadcast"):
        capi._mat(row.expand(3, 4), 3, 4, "A")
    with pytest.raises(ValueError, match="broadcast"):
        capi._mat(row[:3].unsqueeze(1).expand(3, 4), 3, 4, "B")
    base = torch.zeros(16, dtype=torch.float64)
    with pytest.raises(ValueError, match="overlapping"):
        capi._mat(base.as_strided((4, 4), (1, 2)), 4, 4, "C", output=True)
    # the same view as an input (read only) passes the stride check
    with pytest.raises(ValueError, match="GPU tensor"):
        capi._mat(base.as_strided((4, 4), (1, 2)), 4, 4, "A")
    with pytest.raises(ValueError, match="GPU tensor"):
        capi._mat(torch.zeros(1, 5, dtype=torch.float64).expand(1, 5), 1, 5, "C", output=True)


def test_crt_i8_rejects_k_beyond_exact_int32_sums():
    al = D(1.0)
    lib = capi.lib()
    assert lib.momms_hip_crt_i8_workspace_size(8, 8, 131073, 16) == 0
    assert lib.momms_hip_crt_i8_workspace_size(8, 8, 131072, 16) > 0
    st = lib.momms_hip_dgemm_crt_i8(8, 8, 131073, ctypes.byref(al), 8, 1, 8, 8, 1, 131073,
                                    ctypes.byref(al), 8, 1, 8, 16, 8, 1 << 30, None)
    assert st == -3 and "131072" in capi.last_error()
    assert lib.momms_hip_crt_i8_beta(8192, 15) == 51
    assert [lib.momms_hip_crt_i8_modulus(i) for i in (0, 15, 16)] == [255, 191, 0]


def test_dgemm_tree_rejects_bad_steps_before_gpu_work():
    al = D(1.0)
    lib = capi.lib()
    steps = (capi.AlgoStep * 9)()
    for s in steps:
        s.kind, s.bsz = capi.STEP_M, 128
    args = (4, 4, 4, ctypes.byref(al), 8, 1, 4, 8, 1, 4, ctypes.byref(al), 8, 1, 4)
    assert lib.momms_hip_dgemm_tree(*args, ctypes.addressof(steps), 9, None) == -15
    steps[0].bsz = 0
    assert lib.momms_hip_dgemm_tree(*args, ctypes.addressof(steps), 1, None) == -15
    steps[0].kind, steps[0].bsz = 7, 128
    assert lib.momms_hip_dgemm_tree(*args, ctypes.addressof(steps), 1, None) == -15
    assert lib.momms_hip_dgemm_tree(*args, None, 2, None) == -16
    assert lib.momms_hip_dgemm_tree(*args, None, -1, None) == -16
    assert "momms_hip_dgemm_tree" in capi.last_error()


def test_batch_class_follows_its_formula():
    import math

    rows, _, _, _ = capi.uneven_batch(300)
    rows = rows + [(0, 5, 7, 0, 0, 0, 1, 7, 1)]  # an empty problem does not count
    live = [r for r in rows if r[0] and r[1]]
    mean_k = sum(r[2] for r in live) / len(live)
    want = 16 * min(15, int(math.log2(1 + len(live)))) + min(8, int(math.log2(1 + mean_k)))
    assert capi.batch_class(rows) == want
    assert capi.batch_class([]) == 0


def test_tuning_table_round_trip_and_refusals(tmp_path):
    saved = capi.tuning_export()
    try:
        table = {"version": capi.TUNING_VERSION, "entries": [
            {"kind": "batch_fp64", "key": 200, "value": [48, 256, 0, 0]},
            {"kind": "batch_split", "key": 200, "value": [6, 0, 0, 0]},
            {"kind": "batch_crt", "key": 17, "value": [13, 0, 0, 0]}]}
        capi.tuning_import(table, replace=True)
        got = capi.tuning_export()
        assert got["version"] == capi.TUNING_VERSION
        assert sorted((e["kind"], e["key"], tuple(e["value"])) for e in got["entries"]) == sorted(
            (e["kind"], e["key"], tuple(e["value"])) for e in table["entries"])
        # merge: same (kind, key) replaced, others kept
        capi.tuning_import({"version": 1, "entries": [
            {"kind": "batch_fp64", "key": 200, "value": [32, 1 << 30, 0, 0]}]})
        got = {(e["kind"], e["key"]): e["value"] for e in capi.tuning_export()["entries"]}
        assert got[("batch_fp64", 200)] == [32, 1 << 30, 0, 0] and len(got) == 3
        # through a file
        path = tmp_path / "tuning.json"
        capi.save_tuning(path)
        capi.tuning_import({"version": 1, "entries": []}, replace=True)
        assert capi.tuning_export()["entries"] == []
        capi.load_tuning(path)
        assert len(capi.tuning_export()["entries"]) == 3
        before = capi.tuning_export()
        bad = [
            {"version": 2, "entries": []},
            {"version": 1, "entries": [{"kind": 9, "key": 0, "value": [1, 0, 0, 0]}]},
            {"version": 1, "entries": [{"kind": "batch_fp64", "key": 256, "value": [48, 256, 0, 0]}]},
            {"version": 1, "entries": [{"kind": "batch_fp64", "key": 3, "value": [300, 256, 0, 0]}]},
            {"version": 1, "entries": [{"kind": "batch_fp64", "key": 3, "value": [8, 256, 0, 0]}]},
            {"version": 1, "entries": [{"kind": "batch_split", "key": 3, "value": [9, 0, 0, 0]}]},
            {"version": 1, "entries": [{"kind": "batch_crt", "key": 3, "value": [17, 0, 0, 0]}]},
            {"version": 1, "entries": [{"kind": "batch_crt", "key": 3, "value": [13, 1, 0, 0]}]},
            # one bad entry refuses the whole table
            {"version": 1, "entries": [{"kind": "batch_split", "key": 4, "value": [5, 0, 0, 0]},
                                       {"kind": "batch_split", "key": 5, "value": [0, 0, 0, 0]}]},
        ]
        for t in bad:
            with pytest.raises(capi.MommsError):
                capi.tuning_import(t, replace=True)
            assert capi.tuning_export() == before
        with pytest.raises(ValueError):
            capi.tuning_import({"version": 1, "entries": [
                {"kind": "batch_split", "key": i % 256, "value": [5]} for i in range(257)]})
    finally:
        capi.tuning_import(saved, replace=True)
