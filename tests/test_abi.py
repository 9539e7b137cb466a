"""The C-ABI library builds for gfx950, loads, and exports every entry point
include/cgocn.h declares (no device calls: this runs without a GPU)."""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

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


STRUCTS = {"cgo_grid_desc": engine.GridDesc, "cgo_member_params": engine.MemberParams,
           "cgo_bgc_params": engine.BgcParams, "cgo_diag": engine.Diag}
# cgo_abi_info reports the first four; the layout check covers every struct
LAYOUTS = dict(STRUCTS, cgo_momentum_desc=engine.MomentumDesc, cgo_surflux_desc=engine.SurfluxDesc)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_struct_layouts_match_header(tmp_path):
    """Every field of the ctypes mirrors sits at the offset, and every struct
    has the size, that the C compiler gives include/cgocn.h."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "cgocn.h"', "int main(void){"]
    for cname, py in LAYOUTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            lines.append(f'printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for ln in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        a, b, v = ln.split()
        got[(a, b)] = int(v)
    for cname, py in LAYOUTS.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f in py._fields_:
            assert got[(cname, f[0])] == getattr(py, f[0]).offset, (cname, f[0])
    # the sizes the Fortran bind(C) mirrors are also checked against (test_fortran_host)
    assert [got[(n, "size")] for n in LAYOUTS] == [208, 24, 256, 48, 80, 208]


class _FakeLib:
    """Stands in for a libcgocn.so whose cgo_abi_info reports `sizes` and `stamp`."""

    def __init__(self, sizes, stamp):
        self.sizes, self.stamp = sizes, stamp

    def cgo_abi_info(self, sizes, stamp, n):
        for i, v in enumerate(self.sizes):
            sizes[i] = v
        C.memmove(stamp, self.stamp.encode() + b"\0", len(self.stamp) + 1)
        return 0


def test_mismatched_library_is_refused_at_load():
    from cgenie_amd.build import abi_stamp

    good = [C.sizeof(s) for s in STRUCTS.values()]
    engine.check_abi(_FakeLib(good, abi_stamp()))
    with pytest.raises(engine.CgoError, match="struct sizes"):
        engine.check_abi(_FakeLib([good[0] + 8] + good[1:], abi_stamp()))
    with pytest.raises(engine.CgoError, match="another include/cgocn.h"):
        engine.check_abi(_FakeLib(good, "0123456789abcdef"))


def test_built_library_matches_this_header():
    lib = engine.load()   # check_abi runs inside load(); repeat it explicitly
    engine.check_abi(lib)
