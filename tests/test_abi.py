"""CPU checks of the C-ABI boundary: the HIP library loads without a GPU and
exports every entry point declared in include/trident_hip.h (no compute
calls), and the Python binding declares each one."""
import ctypes
import os
import re

from conftest import ROOT
from trident_amd import _lib

HEADER = os.path.join(ROOT, "include", "trident_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tr_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("tr_spgemm", "tr_sp_mma", "tr_spadd", "tr_symbolic", "tr_bitmask_gen", "tr_bitmask_and",
                 "tr_compress", "tr_csx_free", "tr_last_error", "tr_ctx_create", "tr_ctx_destroy"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_binding_covers_header():
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(declared_functions()) <= bound


def test_struct_layout_matches_header():
    # tr_csx_t: 4 int32, int64, 4 pointers, size_t
    assert ctypes.sizeof(_lib.tr_csx_t) == 4 * 4 + 8 + 4 * 8 + 8
    assert _lib.tr_csx_t.nnz.offset == 16 and _lib.tr_csx_t.ptr_vec.offset == 24


def test_version_string_without_gpu():
    lib = _lib.load()
    assert b"gfx950" in lib.tr_version()


def test_null_context_is_rejected_without_gpu():
    lib = _lib.load()
    assert lib.tr_spgemm(None, None, None, None) == _lib.TR_ERR_INVALID
    assert lib.tr_last_error(None) == b"null context"


def test_tuning_defaults_and_validation_without_gpu():
    from trident_amd import _lib
    lib = _lib.load()
    t = _lib.tr_tuning_t()
    assert lib.tr_tuning_default(ctypes.byref(t)) == _lib.TR_OK
    assert (t.light_max_flops, t.window_entries, t.item_windows, t.heavy_table, t.dense_width) == \
        (1024, 1152, 16, 2048, 2048)
    bad = [dict(light_max_flops=2048), dict(light_max_flops=32), dict(heavy_table=3000),
           dict(dense_width=-1), dict(dense_width=2049),
           dict(window_entries=1410), dict(window_entries=32), dict(item_windows=0), dict(item_windows=65)]
    for b in bad:
        u = _lib.tr_tuning_t()
        lib.tr_tuning_default(ctypes.byref(u))
        for k, v in b.items():
            setattr(u, k, v)
        assert lib.tr_set_default_tuning(ctypes.byref(u)) == _lib.TR_ERR_INVALID, b
    u = _lib.tr_tuning_t()
    lib.tr_tuning_default(ctypes.byref(u))
    u.heavy_table, u.window_entries, u.dense_width = 4096, 1920, 4096
    assert lib.tr_set_default_tuning(ctypes.byref(u)) == _lib.TR_OK
    u.dense_width = 4097
    assert lib.tr_set_default_tuning(ctypes.byref(u)) == _lib.TR_ERR_INVALID
    u.dense_width = 0
    u.heavy_table, u.window_entries = 2048, 1409  # 3/4 of the table - 127
    assert lib.tr_set_default_tuning(ctypes.byref(u)) == _lib.TR_OK
    assert lib.tr_set_default_tuning(None) == _lib.TR_OK
