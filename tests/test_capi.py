"""C-ABI checks that need no GPU: the library loads, exports every symbol
include/bulkx.h declares, and rejects bad arguments before touching a device."""
import ctypes

import pytest

from bulk_amd import _capi as C


def test_library_exports_every_declared_symbol():
    lib = C.lib()
    declared = C.header_symbols()
    assert len(declared) >= 20
    missing = [s for s in declared if not hasattr(lib, s)]
    assert missing == []
    # the ctypes table binds every declared function with a signature
    assert sorted(set(declared) - set(C.SIGNATURES)) == []


def test_status_strings():
    lib = C.lib()
    for code, name in C.STATUS_NAMES.items():
        assert lib.bulkx_status_string(code).decode() == name
    assert lib.bulkx_status_string(999).decode() == "BULKX_ERR_UNKNOWN"
    assert lib.bulkx_version().decode().count(".") == 2


def test_dtype_sizes():
    lib = C.lib()
    assert [lib.bulkx_dtype_size(d) for d in range(6)] == [4, 4, 8, 8, 4, 8]
    assert lib.bulkx_dtype_size(6) == 0


@pytest.mark.parametrize("call", [
    lambda l: l.bulkx_inclusive_scan(99, 0, None, None, 10, None, None, 0, None),
    lambda l: l.bulkx_inclusive_scan(0, 7, None, None, 10, None, None, 0, None),
    lambda l: l.bulkx_inclusive_scan(0, 0, None, None, -1, None, None, 0, None),
    lambda l: l.bulkx_inclusive_scan(0, 0, None, None, 10, None, None, 0, None),
    lambda l: l.bulkx_exclusive_scan(0, 0, None, None, 10, None, None, 0, None),
    lambda l: l.bulkx_reduce(0, 0, None, 10, None, None, None, 0, None),
    lambda l: l.bulkx_merge(0, 5, None, 1, None, 1, None, None, 0, None),
    lambda l: l.bulkx_stable_sort_by_key(0, 0, 0, None, None, -3, None, 0, None),
    lambda l: l.bulkx_reduce_by_key(0, 4, 0, 0, None, None, 10, None, None, None, None, 0, None),
    lambda l: l.bulkx_reduce_intervals(0, 0, None, 10, 2, 0, None, None),
])
def test_invalid_arguments_rejected_before_launch(call):
    status = call(C.lib())
    assert status in (1, 2, 3)


def test_server_submit_host_rejects_bad_arguments():
    """Host-staged submission validates before touching a server or device."""
    lib = C.lib()
    rid, nb = ctypes.c_int64(), ctypes.c_int64()
    job = C.Job(0, 0, -1, 0, 16, 0, None, None, None, None, None, None, None)
    assert lib.bulkx_server_submit_host(None, ctypes.byref(job), -1.0, ctypes.byref(rid),
                                        ctypes.byref(nb)) == 1
    assert lib.bulkx_server_submit_host(ctypes.c_void_p(1), None, -1.0, ctypes.byref(rid),
                                        ctypes.byref(nb)) == 1


def test_zero_length_calls_are_noops():
    lib = C.lib()
    assert lib.bulkx_inclusive_scan(0, 0, None, None, 0, None, None, 0, None) == 0
    assert lib.bulkx_stable_sort_by_key(0, 0, 0, None, None, 0, None, 0, None) == 0
    assert lib.bulkx_merge(0, 0, None, 0, None, 0, None, None, 0, None) == 0


def test_workspace_size_query_is_host_only():
    lib = C.lib()
    nb = ctypes.c_size_t()
    ng, gs, gr = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
    # one look-back record per tile plus a 256-byte header; records are one
    # 64-bit word per 32 bits of the accumulator (int32: 1, float64 as
    # double-double: 4)
    # double-double: 4).  The workspace is sized for the smallest tile of any
    # compiled shape, so it serves whichever shape the tuning table picks.
    nvar = ctypes.c_int()
    assert lib.bulkx_tuning_variants(C.ALGO_INCLUSIVE_SCAN, ctypes.byref(nvar)) == 0
    for dt, n, words in [(C.INT32, 1 << 24, 1), (C.FLOAT64, 1 << 20, 4), (C.FLOAT32, 12345679, 2)]:
        assert lib.bulkx_workspace_size(C.ALGO_INCLUSIVE_SCAN, dt, -1, n, 0, ctypes.byref(nb)) == 0
        tiles = []
        for v in range(nvar.value):
            for sc in range(5):
                assert lib.bulkx_set_tuning(C.ALGO_INCLUSIVE_SCAN, dt, sc, v) == 0
            assert lib.bulkx_choose_sizes(C.ALGO_INCLUSIVE_SCAN, dt, n, ctypes.byref(ng),
                                          ctypes.byref(gs), ctypes.byref(gr)) == 0
            tile = gs.value * gr.value
            assert ng.value == -(-n // tile)
            tiles.append(ng.value)
        assert nb.value == 256 + 8 * words * max(tiles)
    assert lib.bulkx_load_tuning(None, None) == 0
    # a single tile needs no scratch
    assert lib.bulkx_workspace_size(C.ALGO_INCLUSIVE_SCAN, C.INT32, -1, 1000, 0, ctypes.byref(nb)) == 0
    assert nb.value == 0
    assert lib.bulkx_workspace_size(C.ALGO_INCLUSIVE_SCAN, 9, -1, 10, 0, ctypes.byref(nb)) == 2


def test_device_info_without_gpu_reports_no_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    lib = C.lib()
    cus = ctypes.c_int()
    st = lib.bulkx_device_info(ctypes.byref(cus), None, None, None, None)
    assert st in (5, 6)


def test_tuning_table_round_trip(tmp_path):
    """A table written in the autotuner's format changes choose_sizes; bad
    entries are ignored; bulkx_set_tuning validates its arguments."""
    lib = C.lib()
    ng, gs, gr = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
    shapes = {}
    nvar = ctypes.c_int()
    assert lib.bulkx_tuning_variants(C.ALGO_REDUCE_BY_KEY, ctypes.byref(nvar)) == 0
    assert nvar.value >= 6
    for v in range(nvar.value):
        path = tmp_path / f"t{v}.json"
        path.write_text('{"device": "x", "entries": ['
                        f'{{"algo": {C.ALGO_REDUCE_BY_KEY}, "dtype": {C.FLOAT32}, "size_class": 4, "variant": {v}}},'
                        f'{{"algo": {C.ALGO_REDUCE_BY_KEY}, "dtype": {C.FLOAT32}, "size_class": 4, "variant": 99}},'
                        '{"algo": 42, "dtype": 0, "size_class": 0, "variant": 0}]}')
        entries = ctypes.c_int()
        assert lib.bulkx_load_tuning(str(path).encode(), ctypes.byref(entries)) == 0
        assert entries.value == 1
        assert lib.bulkx_choose_sizes(C.ALGO_REDUCE_BY_KEY, C.FLOAT32, 1 << 26, ctypes.byref(ng),
                                      ctypes.byref(gs), ctypes.byref(gr)) == 0
        shapes[v] = (gs.value, gr.value)
        assert ng.value == (1 << 26) // (gs.value * gr.value)
    assert {v: shapes[v] for v in range(6)} == {0: (512, 16), 1: (512, 8), 2: (256, 16),
                                                 3: (1024, 16), 4: (1024, 8), 5: (256, 32)}
    assert lib.bulkx_set_tuning(C.ALGO_SCAN if hasattr(C, "ALGO_SCAN") else 0, 0, 5, 0) == 1  # size classes 0..4
    assert lib.bulkx_set_tuning(C.ALGO_REDUCE_BY_KEY, C.FLOAT32, 0, nvar.value) == 1
    assert lib.bulkx_load_tuning(str(tmp_path / "missing.json").encode(), None) == 1
    assert lib.bulkx_load_tuning(None, None) == 0
