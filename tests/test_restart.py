"""Saved state: exact engine checkpoints (GPU) and restart files in the
reference's netCDF formats (CPU: layout, coordinates and round trips;
goldstein_data.f90:10-354, biogem_data_netCDF.f90:24-142,
biogem_data.f90:438-568, biogem.f90:2280-2360)."""
import math

import numpy as np
import pytest

from cgenie_amd import engine as E
from cgenie_amd import grid as G
from cgenie_amd import restart as R
from cases import make_case, wet


def test_goldstein_coordinates_closed_form():
    """Cell-centre longitudes 360/maxi apart from -260 + 180/maxi, latitudes
    asin of the equal-sin centres (symmetric about the equator), depths
    increasing downwards and summing with the layer thicknesses."""
    g = G.make_grid(36, 36, 16)
    lon, lat, depth = R.goldstein_coordinates(g)
    assert np.allclose(np.diff(lon), 10.0, rtol=0, atol=1e-12) and abs(lon[0] - (-255.0)) < 1e-12
    assert np.allclose(lat, -lat[::-1], rtol=0, atol=1e-12)
    assert np.allclose(np.sin(np.radians(lat)), g.s[1:37], rtol=0, atol=1e-14)
    assert np.all(np.diff(depth) < 0)          # level 1 is the bottom
    # each level's depth lies between its top (zw(k)) and bottom (zw(k-1)) face
    for k in range(1, 17):
        assert -G.DSC * g.zw[k] < depth[k - 1] < -G.DSC * g.zw[k - 1]


def test_goldstein_restart_round_trip_and_layout(tmp_path):
    from scipy.io import netcdf_file

    c = make_case(18, 18, 8, kind="ts", L=3, seed=4)
    g = c.g
    p = tmp_path / "goldstein_restart.nc"
    R.write_goldstein_restart(str(p), g, c.k1, c.ts, c.u, iyear=12, imonth=3, iday=7, ioffset=5)
    with netcdf_file(str(p), "r", mmap=False) as f:
        assert f.version_byte == 1                                   # NetCDF classic
        assert f.dimensions == {"nrecs": 1, "longitude": 18, "latitude": 18, "depth": 8}
        assert f.variables["temp"].dimensions == ("depth", "latitude", "longitude")
        assert f.variables["temp"].typecode() == "d" and f.variables["longitude"].typecode() == "f"
        assert f.variables["iyear"].typecode() == "i"
    r = R.read_goldstein_restart(str(p))
    colwet = c.k1[1:-1, 1:-1] <= g.maxk
    assert np.array_equal(r["temp"], np.where(colwet[None], c.ts[..., 0], 0.0))
    assert np.array_equal(r["salinity"], np.where(colwet[None], c.ts[..., 1], 0.0))
    assert np.array_equal(r["uvel"], c.u[:, 1:, 1:, 0]) and np.array_equal(r["vvel"], c.u[:, 1:, 1:, 1])
    assert (r["iyear"], r["imonth"], r["iday"], r["ioffset"]) == (12, 3, 7, 5)


def test_biogem_restart_round_trip_and_layout(tmp_path):
    from scipy.io import netcdf_file

    c = make_case(18, 18, 8, kind="config2", seed=6)
    g, t = c.g, c.tset
    rng = np.random.default_rng(1)
    ocn = np.zeros_like(c.ts)
    ocn[..., 0] = c.ts[..., 0] + 273.15
    ocn[..., 1] = c.ts[..., 1] + 34.9
    for q in (t.DIC, t.ALK, t.PO4, t.O2, t.DOM_P):
        ocn[..., q] = c.ts[..., q]
    part = rng.uniform(0, 1e-6, (g.maxk, g.maxj, g.maxi, 5))
    w = wet(c)
    p = tmp_path / "_restart.nc"
    R.write_biogem_restart(str(p), g, c.k1, ocn, part, t, year=40.0)
    with netcdf_file(str(p), "r", mmap=False) as f:
        names = set(f.variables)
        assert {"ocn_temp", "ocn_sal", "ocn_DIC", "ocn_ALK", "ocn_PO4", "ocn_O2", "ocn_DOM_P", "bio_part_POC",
                "bio_part_POP", "bio_part_CaCO3", "bio_part_POC_frac2", "bio_part_CaCO3_frac2"} <= names
        assert "ocn_DOM_C" not in names                               # not carried in this set
        v = f.variables["ocn_DIC"]
        assert v.dimensions == ("zt", "lat", "lon") and v.typecode() == "f"
        raw = np.array(v[:])
        # surface first; land is the fill value
        assert np.array_equal(raw[0][w[-1]], ocn[-1, ..., t.DIC][w[-1]].astype(np.float32))
        assert np.all(raw[::-1][~w] > 9e36)
        zt = np.array(f.variables["zt"][:])
        assert np.all(np.diff(zt) > 0) and abs(zt[0] - 0.5 * G.DSC * g.dz[g.maxk]) < 1e-9
    o2, p2 = R.read_biogem_restart(str(p), g, t)
    for q in (0, 1, t.DIC, t.ALK, t.PO4, t.O2, t.DOM_P):
        assert np.array_equal(o2[..., q][w], ocn[..., q][w].astype(np.float32).astype(np.float64))
        assert not o2[..., q][~w].any()
    assert np.array_equal(p2[w], part[w].astype(np.float32).astype(np.float64))


def test_biogem_restart_state_returns_the_rain():
    """biogem.f90:2298-2315: the rain at k1 goes back into the bottom cell
    through conv_sed_ocn, DIC += rM (POC + CaCO3), ALK += rM (2 CaCO3 - 16 POP),
    PO4 += rM POP, O2 += rM (cO2POC POC - 2 POP); other cells unchanged."""
    c = make_case(18, 18, 8, kind="config2", seed=8)
    g, t = c.g, c.tset
    ocn = np.zeros_like(c.ts)
    settle = np.zeros((g.maxj, g.maxi, 5))
    j, i = np.argwhere(c.k1[1:-1, 1:-1] <= g.maxk)[3]
    settle[j, i] = [3e-3, 3e-3 / 106, 1e-3, 0.05, 0.5]
    rM = np.full((g.maxk, g.maxj, g.maxi), 2e-15)
    out = R.biogem_restart_state(ocn, settle, rM, c.k1, t)
    kb = c.k1[j + 1, i + 1] - 1
    d = out - ocn
    POC, POP, Ca = settle[j, i, :3] * 2e-15
    assert math.isclose(d[kb, j, i, t.DIC], POC + Ca, rel_tol=1e-14)
    assert math.isclose(d[kb, j, i, t.ALK], 2 * Ca - 16 * POP, rel_tol=1e-12)
    assert math.isclose(d[kb, j, i, t.PO4], POP, rel_tol=1e-14)
    assert math.isclose(d[kb, j, i, t.O2], (-138 / 106 + 2 / 106) * POC - 2 * POP, rel_tol=1e-12)
    d[kb, j, i] = 0.0
    assert not d.any()


def _bgc_member(eng, c, prec):
    m = eng.add_member(eng.add_grid(c.g, c.k1), c.tset.L, prec, c.diff)
    eng.set_velocity(m, c.u)
    eng.upload(m, c.ts)
    eng.bgc_init(m, E.default_bgc_params(1.0 / 96, c.tset, carb_tol=1e-12, carb_maxit=1000), c.light, c.mld)
    return m


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [E.FP64, E.FP32C])
def test_checkpoint_restore_continues_bitwise(tmp_path, prec):
    """A member checkpointed after 7 steps and restored into a fresh member
    (a new engine) continues to the bits of the uninterrupted run."""
    c = make_case(kind="config2", seed=91)
    eng = E.Engine(0)
    try:
        a = _bgc_member(eng, c, prec)
        eng.step([a], 7)
        path = tmp_path / "member.ckpt"
        nbytes = R.save_checkpoint(eng, a, str(path))
        eng.step([a], 9)
        want = eng.download(a, c.ts.shape)
        want_ocn = eng.bgc_download(a, 0, c.ts.shape)
        want_carb = eng.bgc_download(a, 2, (c.g.maxj, c.g.maxi, 10))
        steps = eng.diag(a).steps
    finally:
        eng.close()
    print("checkpoint bytes", nbytes, flush=True)
    eng = E.Engine(0)
    try:
        b = eng.add_member(eng.add_grid(c.g, c.k1), c.tset.L, prec, c.diff)   # no BIOGEM yet: the restore adds it
        R.load_checkpoint(eng, b, str(path))
        eng.step([b], 9)
        assert eng.diag(b).steps == steps
        assert np.array_equal(eng.download(b, c.ts.shape), want)
        assert np.array_equal(eng.bgc_download(b, 0, c.ts.shape), want_ocn)
        assert np.array_equal(eng.bgc_download(b, 2, (c.g.maxj, c.g.maxi, 10)), want_carb)
        # refused: another grid, another tracer count
        c2 = make_case(18, 18, 8, kind="config2", seed=92)
        x = eng.add_member(eng.add_grid(c2.g, c2.k1), c2.tset.L, prec, c2.diff)
        with pytest.raises(E.CgoError, match="rc=-1"):
            R.load_checkpoint(eng, x, str(path))
        y = eng.add_member(eng.add_grid(c.g, c.k1), c.tset.L - 1, prec, c.diff)
        with pytest.raises(E.CgoError, match="rc=-1"):
            R.load_checkpoint(eng, y, str(path))
    finally:
        eng.close()


@pytest.mark.gpu
def test_reference_restart_files_from_the_engine(tmp_path):
    """GOLDSTEIN and BIOGEM restart files written from a member's device
    state hold its T and S exactly and its BIOGEM tracers to float precision."""
    c = make_case(18, 18, 8, kind="config2", seed=93)
    eng = E.Engine(0)
    try:
        m = _bgc_member(eng, c, E.FP64)
        eng.step([m], 5)
        ts = eng.download(m, c.ts.shape)
        ocn = eng.bgc_download(m, 0, c.ts.shape)
        part = eng.bgc_download(m, 1, (c.g.maxk, c.g.maxj, c.g.maxi, 5))
    finally:
        eng.close()
    R.write_goldstein_restart(str(tmp_path / "g.nc"), c.g, c.k1, ts, c.u)
    R.write_biogem_restart(str(tmp_path / "b.nc"), c.g, c.k1, ocn, part, c.tset)
    r = R.read_goldstein_restart(str(tmp_path / "g.nc"))
    w = wet(c)
    assert np.array_equal(r["temp"][w], ts[..., 0][w]) and np.array_equal(r["salinity"][w], ts[..., 1][w])
    o2, p2 = R.read_biogem_restart(str(tmp_path / "b.nc"), c.g, c.tset)
    for q in (0, 1, c.tset.DIC, c.tset.PO4):
        assert np.all(np.abs(o2[..., q][w] - ocn[..., q][w]) <= 6e-8 * np.abs(ocn[..., q][w]))
