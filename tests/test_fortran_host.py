"""The Fortran host side (cgenie_amd/host): iso_c_binding interface, the
drop-in module with the reference's subroutine names and the genie-order
driver (genie.f90:306-313, 360-381).

CPU: the host code builds with amdflang and its bind(C) types have the sizes
the C compiler gives the structs of include/cgocn.h.
GPU: the driver, stepping in the genie.f90 order through tstepo /
step_biogem / biogem_tracercoupling, reproduces the Python path through the
same library bit for bit and the oracle to 1e-12."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from cgenie_amd import engine as E
import oracle_ffi as O
from cases import make_case, rel_err, wet

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "cgenie_amd", "host")
DRIVER = os.path.join(HOST, "_build", "cgenie_hip_driver")
LAYOUT = os.path.join(HOST, "_build", "cgo_api_layout")

needs_flang = pytest.mark.skipif(shutil.which("amdflang") is None and not os.path.exists(LAYOUT),
                                 reason="amdflang not available and the host code is not built")


def _build_host():
    if shutil.which("amdflang") is not None and os.path.exists(os.path.join(ROOT, "cgenie_amd", "libcgocn.so")):
        subprocess.run(["make", "-C", HOST], check=True, capture_output=True)


@needs_flang
def test_bind_c_types_match_c_structs(tmp_path):
    _build_host()
    out = subprocess.run([LAYOUT], check=True, capture_output=True, text=True).stdout.split()
    fsz = {out[i]: int(out[i + 1]) for i in range(0, len(out), 2)}
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include "cgocn.h"\nint main(void){printf("%zu %zu %zu %zu\\n",'
                   'sizeof(cgo_grid_desc),sizeof(cgo_member_params),sizeof(cgo_bgc_params),sizeof(cgo_diag));}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    csz = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert [fsz["cgo_grid_desc"], fsz["cgo_member_params"], fsz["cgo_bgc_params"], fsz["cgo_diag"]] == csz


def write_driver_input(path, c, nsteps, kbiogem=1, precision=E.FP64, bgc=True, mirror=False, dtyr=1.0 / 96):
    g = c.g
    hdr = np.array([g.maxi, g.maxj, g.maxk, c.tset.L, nsteps, kbiogem, 1, precision, 1 if bgc else 0, 0,
                    1 if mirror else 0], dtype="<i4")
    sc = np.array([g.dphi, g.rdphi, 5000.0, c.diff[0], c.diff[1], *list(g.ec)[:5]], dtype="<f8")
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(sc.tobytes())
        f.write(np.ascontiguousarray(c.k1, dtype="<i4").tobytes())
        for n in ("rc", "rc2", "cv", "cv2", "rds", "dsv", "rdsv", "ds", "sv"):
            f.write(np.ascontiguousarray(getattr(g, n), dtype="<f8")[: g.maxj + 2].tobytes())
        for n in ("dt", "rdz", "rdza", "dza", "ssmax", "dz"):
            f.write(np.ascontiguousarray(getattr(g, n), dtype="<f8")[: g.maxk + 2].tobytes())
        f.write(np.ascontiguousarray(c.u, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(c.ts, dtype="<f8").tobytes())
        if bgc:
            t = c.tset
            slots = [q + 1 if q >= 0 else 0 for q in (t.DIC, t.ALK, t.PO4, t.O2, t.DOM_C, t.DOM_P)]
            f.write(np.array(slots, dtype="<i4").tobytes())
            f.write(np.array([dtyr, 1e-12], dtype="<f8").tobytes())
            f.write(np.ascontiguousarray(c.light, dtype="<f8").tobytes())
            f.write(np.ascontiguousarray(c.mld, dtype="<f8").tobytes())


@pytest.mark.gpu
@pytest.mark.parametrize("mirror", [False, True])
def test_driver_genie_order_matches_engine_and_oracle(tmp_path, mirror):
    assert os.path.exists(DRIVER), "build the host driver first (__graft_entry__.build())"
    c = make_case(kind="config2", seed=81)
    nsteps = 6
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_driver_input(fin, c, nsteps, mirror=mirror)
    print("driver case config2 seed 81, mirror", mirror, flush=True)
    r = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    raw = np.fromfile(fout, dtype="<f8")
    got = raw[: c.ts.size].reshape(c.ts.shape)
    inv = raw[c.ts.size:]
    # the same library driven from Python
    eng = E.Engine(0)
    m = eng.add_member(eng.add_grid(c.g, c.k1), c.tset.L, E.FP64, c.diff)
    eng.set_velocity(m, c.u)
    eng.upload(m, c.ts)
    eng.bgc_init(m, E.default_bgc_params(1.0 / 96, c.tset, carb_tol=1e-12, carb_maxit=1000), c.light, c.mld)
    eng.step([m], nsteps)
    ref = eng.download(m, c.ts.shape)
    inv_ref = eng.inventory(m)
    eng.close()
    assert np.array_equal(got, ref)
    assert np.array_equal(inv, inv_ref)
    gold = O.GoldsteinOracle(c.og, c.k1, c.ts, c.u)
    bio = O.BiogemOracle(gold, O.bgc_params(1.0 / 96, c.tset), c.light, c.mld)
    bio.step(nsteps)
    err = rel_err(got, gold.ts, wet(c))
    print("driver vs oracle max rel err", err.max(), flush=True)
    assert np.all(err <= 1e-12)
