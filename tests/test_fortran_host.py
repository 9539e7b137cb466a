"""The Fortran host side (cgenie_amd/host): iso_c_binding interface, the
drop-in module with the reference's subroutine names and the genie-order
driver (genie.f90:306-313, 360-381).

CPU: the host code builds with amdflang and its bind(C) types have the sizes
the C compiler gives the structs of include/cgocn.h.
GPU: the driver, stepping in the genie.f90 order through tstepo /
step_biogem / biogem_tracercoupling, reproduces the Python path through the
same library bit for bit and the oracle to 1e-12."""
import ctypes as C
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
    src.write_text('#include <stdio.h>\n#include "cgocn.h"\nint main(void){printf("%zu %zu %zu %zu %zu %zu\\n",'
                   'sizeof(cgo_grid_desc),sizeof(cgo_member_params),sizeof(cgo_bgc_params),sizeof(cgo_diag),'
                   'sizeof(cgo_momentum_desc),sizeof(cgo_surflux_desc));}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    csz = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert [fsz["cgo_grid_desc"], fsz["cgo_member_params"], fsz["cgo_bgc_params"], fsz["cgo_diag"],
            fsz["cgo_momentum_desc"], fsz["cgo_surflux_desc"]] == csz


REF_SRC = "/root/reference/src"


@needs_flang
@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="the reference tree is not present (checked where it is)")
def test_reference_wrapper_calls_compile_against_module(tmp_path):
    """The reference's goldstein_wrapper, biogem_wrapper and
    biogem_tracercoupling_wrapper (genie_loop_wrappers.f90:122-151, 310-322),
    taken from the reference file with only their `USE goldstein` /
    `USE biogem` line swapped for `USE cgenie_hip`, compile against the module
    with the reference's genie_global (compiled where it lies, with the
    reference's -fdefault-real-8): the CALL lines of step_goldstein,
    step_biogem and biogem_tracercoupling are unchanged."""
    import re

    _build_host()
    flags = ["-cpp", "-fdefault-real-8", "-O0", "-module-dir", str(tmp_path), "-I", str(tmp_path)]
    for f in ("common/gem_cmn.f90", "wrappers/genie_control.f90", "wrappers/genie_global.f90"):
        r = subprocess.run(["amdflang", *flags, "-c", os.path.join(REF_SRC, f), "-o",
                            str(tmp_path / (os.path.basename(f) + ".o"))], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
    src = open(os.path.join(REF_SRC, "wrappers", "genie_loop_wrappers.f90")).read()
    parts = []
    for name, use in (("goldstein_wrapper", "USE goldstein"), ("biogem_wrapper", "USE biogem"),
                      ("biogem_tracercoupling_wrapper", "USE biogem")):
        m = re.search(rf"^[ \t]*SUBROUTINE {name}[ \t]*$.*?^[ \t]*END SUBROUTINE {name}[ \t]*$", src, re.M | re.S)
        assert m, name
        body = m.group(0)
        assert len(re.findall(rf"^[ \t]*{use}[ \t]*$", body, re.M)) == 1, name
        assert "CALL " in body
        parts.append(re.sub(rf"^([ \t]*){use}[ \t]*$", r"\1USE cgenie_hip", body, flags=re.M))
    w = tmp_path / "wrappers_on_gpu.f90"
    # IMPLICIT NONE (EXTERNAL): every CALL must resolve to a procedure with an
    # explicit interface (the module's), not to an implicit external one
    w.write_text("MODULE wrappers_on_gpu\n  USE genie_global\n  IMPLICIT NONE (TYPE, EXTERNAL)\nCONTAINS\n" +
                 "\n".join(parts) +
                 "\nEND MODULE wrappers_on_gpu\n")
    r = subprocess.run(["amdflang", *flags, "-I", os.path.join(HOST, "_build"), "-c", str(w), "-o",
                        str(tmp_path / "wrappers_on_gpu.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


def write_driver_input(path, c, nsteps, kbiogem=1, precision=E.FP64, bgc=True, mirror=False, dtyr=1.0 / 96,
                       mom=None, sfx=None):
    """mom: None, or (world, stress (maxj, maxi, 4), adrag, rel, scf) for the velocity solve;
    sfx: None, or (SurfluxSetup, (atemp, ashum, sich, sica, tice), stress) for the surface fluxes."""
    g = c.g
    hdr = np.array([g.maxi, g.maxj, g.maxk, c.tset.L, nsteps, kbiogem, 1, precision, 1 if bgc else 0, 0,
                    1 if mirror else 0, 1 if mom is not None else 0, 1 if sfx is not None else 0], dtype="<i4")
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
        if mom is not None:
            world, st, adrag, rel, scf = mom
            for n in ("s", "c", "rcv", "rds2"):
                f.write(np.ascontiguousarray(getattr(g, n), dtype="<f8")[: g.maxj + 2].tobytes())
            f.write(np.ascontiguousarray(world.labels, dtype="<i4").tobytes())
            f.write(np.array([world.isles], dtype="<i4").tobytes())
            if world.isles:
                f.write(np.ascontiguousarray(world.npi, dtype="<i4").tobytes())
                f.write(np.ascontiguousarray(world.paths, dtype="<i4").tobytes())
            f.write(np.array([adrag, rel, scf], dtype="<f8").tobytes())
            # stress(maxi, maxj, 4) column-major = (4, maxj, maxi) C order
            f.write(np.ascontiguousarray(np.moveaxis(st, 2, 0), dtype="<f8").tobytes())
        if sfx is not None:
            su, atm, st = sfx
            f.write(np.array([su.nyear], dtype="<i4").tobytes())
            f.write(np.ascontiguousarray(su.solfor, dtype="<f8").tobytes())       # solfor(maxj, nyear)
            for a in (su.albcl, su.ca, su.pmeadj):
                f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
            for a in (su.iroff, su.jroff):
                f.write(np.ascontiguousarray(a, dtype="<i4").tobytes())
            f.write(np.array([su.igrid, su.polar_avg, su.radfor_atchem], dtype="<i4").tobytes())
            f.write(np.array([getattr(su, n) for n in SFX_SCALARS], dtype="<f8").tobytes())
            for a in atm:
                f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
            if mom is None:
                f.write(np.ascontiguousarray(np.moveaxis(st, 2, 0), dtype="<f8").tobytes())


# initialise_surflux_hip's scl(17)
SFX_SCALARS = ("scf", "rmax", "delf2x", "rdtdim", "rsictscsf", "olr_adj", "olr_adj0", "t_eqm", "albsic_min",
               "albsic_max", "sich_max", "rate_co2", "rate_ch4", "rate_n2o", "co2", "ch4", "n2o")


CONV_YR_S = 3600.0 * (24.0 * 365.25)              # gem_cmn.f90:511-513
IO_OF = {"T": 1, "S": 2, "DIC": 3, "PO4": 8, "O2": 10, "ALK": 12, "DOM_C": 15, "DOM_P": 20}  # gem_cmn.f90:72-90
IS_OF = (3, 8, 14, 33, 34)                        # POC, POP, CaCO3, POC_frac2, CaCO3_frac2 (gem_cmn.f90:190-255)


@pytest.mark.gpu
@pytest.mark.parametrize("mirror", [False, True])
def test_driver_genie_order_matches_engine_and_oracle(tmp_path, mirror):
    """The driver calls step_goldstein, step_biogem and biogem_tracercoupling
    with the reference's argument lists: the state equals the Python path
    through the same library bit for bit and the oracle to 1e-12; with
    mirroring on, go_rho, the energy and water diagnostics, the bottom-water
    composition dum_sfcocn1 and the rain dum_sfxsed1 of the last step match
    the oracle's (biogem.f90:1719-1754, goldstein.f90:458-471)."""
    assert os.path.exists(DRIVER), "build the host driver first (__graft_entry__.build())"
    c = make_case(kind="config2", seed=81)
    g, t = c.g, c.tset
    nsteps = 6
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_driver_input(fin, c, nsteps, mirror=mirror)
    print("driver case config2 seed 81, mirror", mirror, flush=True)
    r = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    raw = np.fromfile(fout, dtype="<f8")
    n_ts, n_c, n_col = c.ts.size, g.maxi * g.maxj * g.maxk, g.maxi * g.maxj
    o = 0
    got = raw[o:o + n_ts].reshape(c.ts.shape); o += n_ts
    inv = raw[o:o + t.L]; o += t.L
    rho = raw[o:o + n_c].reshape(g.maxk, g.maxj, g.maxi); o += n_c
    energy, water = raw[o], raw[o + 1]; o += 2
    sfcocn = raw[o:o + 38 * n_col].reshape(g.maxj, g.maxi, 38); o += 38 * n_col
    sfxsed = raw[o:o + 34 * n_col].reshape(g.maxj, g.maxi, 34); o += 34 * n_col
    o += 3 * n_c + (g.maxi + 1) * (g.maxj + 1) + 2 * n_col    # go_u, go_psi, go_tau
    assert o == raw.size
    # the same library driven from Python
    eng = E.Engine(0)
    m = eng.add_member(eng.add_grid(c.g, c.k1), t.L, E.FP64, c.diff)
    eng.set_velocity(m, c.u)
    eng.upload(m, c.ts)
    eng.bgc_init(m, E.default_bgc_params(1.0 / 96, t, carb_tol=1e-12, carb_maxit=1000), c.light, c.mld)
    eng.step([m], nsteps)
    ref = eng.download(m, c.ts.shape)
    inv_ref = eng.inventory(m)
    rho_ref = eng.rho(m, g)
    eng.close()
    assert np.array_equal(got, ref)
    assert np.array_equal(inv, inv_ref)
    # the oracle, the last step split to see step_biogem's outputs before the coupling
    gold = O.GoldsteinOracle(c.og, c.k1, c.ts, c.u)
    bio = O.BiogemOracle(gold, O.bgc_params(1.0 / 96, t), c.light, c.mld)
    bio.step(nsteps - 1)
    gold.step(1)
    ts_transported = gold.ts.copy()        # what the last step_goldstein handed back
    L = O.lib()
    L.og_biogem_step(C.byref(bio.st), C.byref(gold.g), C.byref(bio.p), gold.k1, bio.vdocn)
    ocn_b, vd_b = bio.ocn(), bio.vdocn.copy()
    settle = bio.arr("settle", O.NSED).reshape(g.maxk, g.maxj, g.maxi, O.NSED).copy()
    L.og_tracercoupling(C.byref(bio.st), C.byref(gold.g), gold.k1, gold.ts, bio.vdocn)
    w = wet(c)
    err = rel_err(got, gold.ts, w)
    print("driver vs oracle max rel err", err.max(), flush=True)
    assert np.all(err <= 1e-12)
    if not mirror:
        assert not sfcocn.any() and not sfxsed.any()
        return
    assert np.array_equal(rho[w], rho_ref[w])
    # energy and water relative to the start, summed as goldstein.f90:458-471
    vsc = g.dphi * 6.37e6 * 6.37e6

    def tot(ts):
        e = wv = 0.0
        for k in range(1, g.maxk + 1):
            for j in range(1, g.maxj + 1):
                e = e + ts[k - 1, j - 1, :, 0].sum() * g.dz[k] * g.ds[j]
                wv = wv - ts[k - 1, j - 1, :, 1].sum() * g.dz[k] * g.ds[j]
        return e * vsc * 5000.0 * 1000.0 * 3981.1, 1000.0 * wv * vsc * 5000.0 / 34.9

    e0, w0 = tot(c.ts)
    e1, w1 = tot(ts_transported)
    print("energy, water diagnostics", energy, water, "oracle", e1 - e0, w1 - w0, flush=True)
    assert abs(energy - (e1 - e0)) <= 1e-12 * abs(e0)
    wscale = 1000.0 * vsc * 5000.0 / 34.9 * sum(np.abs(ts_transported[k - 1, j - 1, :, 1]).sum() * g.dz[k] * g.ds[j]
                                             for k in range(1, g.maxk + 1) for j in range(1, g.maxj + 1))
    assert abs(water - (w1 - w0)) <= 1e-12 * wscale
    # bottom-water composition and rain of the last step_biogem
    k1 = c.k1[1:-1, 1:-1]
    dts = (1.0 / 96) * CONV_YR_S
    dtyr = dts / CONV_YR_S
    worst = 0.0
    for j, i in np.argwhere(k1 <= g.maxk):
        kb = k1[j, i] - 1
        for name, io in IO_OF.items():
            slot = {"T": 0, "S": 1}.get(name, getattr(t, name, -1))
            if slot < 0:
                continue
            want = ocn_b[kb, j, i, slot] + vd_b[kb, j, i, slot]
            worst = max(worst, abs(sfcocn[j, i, io - 1] - want) / max(abs(want), 1e-300))
        S = sfcocn[j, i, 1]
        assert sfcocn[j, i, 34] == max(0.01028 * S / 35.0, 0.999999e-19)    # Ca from salinity
        A = 2.0 * 3.141592653589793 * (6.37e6 * 6.37e6) * (1.0 / g.maxi) * (g.sv[j + 1] - g.sv[j])
        for q, isx in enumerate(IS_OF):
            sv_ = settle[kb, j, i, q]
            want = sv_ * (1.0 / dts) * dtyr if q >= 3 else (1.0 / A) * sv_ * (1.0 / dts)
            worst = max(worst, abs(sfxsed[j, i, isx - 1] - want) / max(abs(want), 1e-300))
    print("interface arrays vs oracle max rel err", worst, flush=True)
    assert worst <= 1e-12


@pytest.mark.gpu
def test_driver_velocity_solve_matches_engine(tmp_path):
    """step_goldstein with the velocity solve (initialise_goldstein_momentum_hip
    and the wind stresses in step_goldstein's own arguments) on the worbe2
    world, mirroring on: the state, go_u and go_psi / 1592.5 equal the
    Python path through the same library bit for bit (which equals the
    oracle's coupled step, tests/test_gpu_velocity.py), and go_tau is the
    stresses' tau (goldstein.f90:102-107, 444-450)."""
    from types import SimpleNamespace

    from cgenie_amd import grid as G
    import test_gpu_velocity as V

    assert os.path.exists(DRIVER), "build the host driver first (__graft_entry__.build())"
    world = G.read_world(V.WORLD, V.MAXI, V.MAXJ)
    g = G.make_grid(V.MAXI, V.MAXJ, V.MAXK)
    L, nsteps = 4, 5
    ts = V.state(world.k1, L)
    st = V.wind()
    case = SimpleNamespace(g=g, k1=world.k1, u=np.zeros((g.maxk, g.maxj + 1, g.maxi + 1, 3)), ts=ts,
                           tset=SimpleNamespace(L=L), diff=V.DIFF)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_driver_input(fin, case, nsteps, bgc=False, mirror=True, mom=(world, st, 2.5, 0.9, 2.0))
    print("driver with the velocity solve, worbe2, steps", nsteps, flush=True)
    r = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    raw = np.fromfile(fout, dtype="<f8")
    n_c, n_col = g.maxi * g.maxj * g.maxk, g.maxi * g.maxj
    o = 0
    got = raw[o:o + ts.size].reshape(ts.shape); o += ts.size
    o += L + n_c + 2 + 38 * n_col + 34 * n_col        # inventory, rho, diagnostics, BIOGEM outputs
    go_u = raw[o:o + 3 * n_c].reshape(g.maxk, g.maxj, g.maxi, 3); o += 3 * n_c
    go_psi = raw[o:o + (g.maxi + 1) * (g.maxj + 1)].reshape(g.maxj + 1, g.maxi + 1); o += (g.maxi + 1) * (g.maxj + 1)
    go_tau = raw[o:o + 2 * n_col].reshape(g.maxj, g.maxi, 2); o += 2 * n_col
    assert o == raw.size
    eng = E.Engine(0)
    gid = eng.add_grid(g, world.k1)
    eng.set_momentum(gid, g, world)
    m = eng.add_member(gid, L, E.FP64, V.DIFF, diso=True)
    eng.upload(m, ts)
    eng.set_wind(m, st)
    eng.step([m], nsteps)
    ref, u_ref, psi_ref = eng.download(m, ts.shape), eng.velocity(m, g), eng.psi(m, g)
    eng.close()
    assert np.abs(u_ref).max() > 0
    assert np.array_equal(got, ref)
    assert np.array_equal(go_u, u_ref)
    assert np.array_equal(go_psi, 1592.5 * psi_ref)
    dzz = g.dz[g.maxk] * g.dza[g.maxk - 1] / 2
    den = 1.0e3 * 5000.0 * 0.05 * (2 * 7.2921e-5)
    assert np.array_equal(go_tau[..., 0], 2.0 * st[..., 0] / den / dzz * dzz)
    assert np.array_equal(go_tau[..., 1], 2.0 * st[..., 3] / den / dzz * dzz)


@pytest.mark.gpu
@pytest.mark.parametrize("mirror", [False, True])
def test_driver_surface_fluxes_match_engine(tmp_path, mirror):
    """surflux_hip before every step_goldstein, with the velocity solve, on the
    worbe2 world with its own winds and runoff routing and sea ice: the state
    and the last call's outputs, ashum and tice equal the Python path through
    the same library bit for bit (which matches the oracle,
    tests/test_gpu_surflux.py)."""
    from types import SimpleNamespace

    from cgenie_amd import grid as G
    from cgenie_amd import surface as SF
    import test_gpu_surflux as S

    assert os.path.exists(DRIVER), "build the host driver first (__graft_entry__.build())"
    world = G.read_world(os.path.join(S.GOLDEN, "worbe2"), S.MAXI, S.MAXJ)
    g = G.make_grid(S.MAXI, S.MAXJ, S.MAXK)
    L, nsteps = 4, 5
    ts = S.state(world.k1, L)
    st = S.stress()
    su = SF.surflux_setup(g, world.k1)
    atm = SF.atmosphere(g, world.k1, 13, ice=True)
    case = SimpleNamespace(g=g, k1=world.k1, u=np.zeros((g.maxk, g.maxj + 1, g.maxi + 1, 3)), ts=ts,
                           tset=SimpleNamespace(L=L), diff=S.DIFF)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_driver_input(fin, case, nsteps, bgc=False, mirror=mirror, mom=(world, st, 2.5, 0.9, 2.0),
                       sfx=(su, atm, st))
    print("driver with surface fluxes and the velocity solve, worbe2, mirror", mirror, flush=True)
    r = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    raw = np.fromfile(fout, dtype="<f8")
    n_c, n_col = g.maxi * g.maxj * g.maxk, g.maxi * g.maxj
    o = 0
    got = raw[o:o + ts.size].reshape(ts.shape); o += ts.size
    o += L + n_c + 2 + 38 * n_col + 34 * n_col + 3 * n_c + (g.maxi + 1) * (g.maxj + 1) + 2 * n_col
    out = raw[o:o + 28 * n_col].reshape(28, g.maxj, g.maxi); o += 28 * n_col
    ashum = raw[o:o + n_col].reshape(g.maxj, g.maxi); o += n_col
    tice = raw[o:o + n_col].reshape(g.maxj, g.maxi); o += n_col
    assert o == raw.size
    eng = E.Engine(0)
    gid = eng.add_grid(g, world.k1)
    eng.set_surflux(gid, su)
    eng.set_momentum(gid, g, world)
    m = eng.add_member(gid, L, E.FP64, S.DIFF, diso=True)
    eng.upload(m, ts)
    eng.set_wind(m, st)
    eng.set_surface(m, *atm)
    eng.step([m], nsteps)
    ref = eng.download(m, ts.shape)
    dev = eng.surface_outputs(m, g)
    eng.close()
    assert np.array_equal(got, ref)
    for q, n in enumerate(E.SF_NAMES):
        assert np.array_equal(out[q], dev[n]), n
    assert np.array_equal(ashum, dev["ashum"]) and np.array_equal(tice, dev["tice"])
    assert dev["istep"] == nsteps
