"""Air-sea gas exchange and the ATCHEM atmosphere on the device, through the
C ABI (needs an MI355X): CO2 and O2 exchange at the surface of every wet
column (biogem_box.f90:27-299, biogem.f90:1044-1046, 1605-1614) and the
well-mixed atmosphere it feeds (atchem.f90:63-158, 306-320), against the
oracle.

Bar: as the other BIOGEM parity tests, FP64 within 1e-12 per cell, relative
(the device's exp, log and pow differ from glibc's by an ulp at most); the
atmosphere within 1e-12 relative; the FP32 form within 1e-6 of FP64."""
import sys

import numpy as np
import pytest

from cgenie_amd import engine as E
from cgenie_amd import grid as G
import oracle_ffi as O
from cases import make_case, rel_err, wet

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = E.Engine(0)
    yield e
    e.close()


def _wind(c, seed=7):
    return np.random.default_rng(seed).uniform(2.0, 14.0, (c.g.maxj, c.g.maxi))


def _member(eng, c, prec=E.FP64, pco2=250e-6):
    m = eng.add_member(eng.add_grid(c.g, c.k1), c.tset.L, prec, c.diff)
    eng.set_velocity(m, c.u)
    eng.upload(m, c.ts)
    p = E.default_bgc_params(1.0 / 96, c.tset, carb_tol=1e-12, carb_maxit=1000, airsea=1, atm_pCO2=pco2)
    eng.bgc_init(m, p, c.light, c.mld)
    eng.set_windspeed(m, _wind(c))
    return m


@pytest.mark.parametrize("kind,pco2", [("config2", 250e-6), ("domc", 400e-6)])
def test_airsea_fp64_matches_oracle(eng, kind, pco2):
    """GOLDSTEIN, BIOGEM with air-sea exchange, the tracer coupling and
    ATCHEM, 12 steps in the genie.f90 order, against the oracle: state,
    BIOGEM tracers, the last flux and the atmosphere."""
    say("airsea case", kind, pco2)
    c = make_case(kind=kind, seed=41)
    m = _member(eng, c, pco2=pco2)
    gold = O.GoldsteinOracle(c.og, c.k1, c.ts, c.u)
    bio = O.BiogemOracle(gold, O.bgc_params(1.0 / 96, c.tset, airsea=1, atm_pCO2=pco2), c.light, c.mld)
    bio.set_windspeed(_wind(c))
    n = 12
    eng.step([m], n)
    bio.step(n)
    w = wet(c)
    err = rel_err(eng.download(m, c.ts.shape), gold.ts, w)
    say("  ts max rel err", err.max())
    assert np.all(err <= 1e-12)
    ocn, oref = eng.bgc_download(m, 0, c.ts.shape), bio.ocn()
    sl = [0, 1] + [q for q in (c.tset.DIC, c.tset.ALK, c.tset.PO4, c.tset.O2, c.tset.DOM_C, c.tset.DOM_P) if q >= 0]
    eo = rel_err(ocn[..., sl], oref[..., sl], w)
    say("  ocn max rel err", eo)
    assert np.all(eo <= 1e-12)
    f, fref = eng.airsea_flux(m, c.g.maxi, c.g.maxj), bio.fatm()
    ws = w[-1]
    assert not f[~ws].any() and np.abs(fref[ws]).min() > 0
    ef = np.abs(f - fref).max(axis=(0, 1)) / np.abs(fref).max(axis=(0, 1))
    say("  flux err (CO2, O2)", ef)
    assert np.all(ef <= 1e-10)
    a, aref = eng.atmosphere(m, c.g.maxi, c.g.maxj), bio.atm()
    ea = np.abs(a - aref).max(axis=(0, 1)) / np.abs(aref).max(axis=(0, 1))
    say("  atm", a[0, 0], "oracle", aref[0, 0], "err", ea)
    assert np.all(ea <= 1e-12)
    assert np.all(a == a[:1, :1]) and a[0, 0, 0] != pco2   # one partial pressure, moved by the exchange
    eng.remove_member(m)


def test_airsea_off_leaves_the_atmosphere(eng):
    """With the exchange off the flux stays zero and the atmosphere at its
    initial partial pressures, whatever the wind."""
    c = make_case(18, 18, 8, kind="config2", seed=42)
    m = eng.add_member(eng.add_grid(c.g, c.k1), c.tset.L, E.FP64, c.diff)
    eng.set_velocity(m, c.u)
    eng.upload(m, c.ts)
    eng.bgc_init(m, E.default_bgc_params(1.0 / 96, c.tset, carb_tol=1e-12, carb_maxit=1000), c.light, c.mld)
    eng.set_windspeed(m, _wind(c))
    eng.step([m], 4)
    assert not eng.airsea_flux(m, c.g.maxi, c.g.maxj).any()
    a = eng.atmosphere(m, c.g.maxi, c.g.maxj)
    assert np.all(a[..., 0] == 278e-6) and np.all(a[..., 1] == 0.2095)
    with pytest.raises(E.CgoError, match="rc="):
        eng.set_windspeed(m, -_wind(c))
    eng.remove_member(m)


def test_airsea_fp32_form_tracks_fp64(eng):
    """The FP32 form with air-sea exchange stays within 1e-6 of FP64 in the
    state and in the atmosphere over 24 steps."""
    c = make_case(kind="config2", seed=43)
    a = _member(eng, c, E.FP64)
    b = _member(eng, c, E.FP32C)
    eng.step([a, b], 24)
    w = wet(c)
    xb, xa = eng.download(b, c.ts.shape), eng.download(a, c.ts.shape)
    xb[..., 1] += G.SALN0  # absolute salinity, not the anomaly GOLDSTEIN stores
    xa[..., 1] += G.SALN0
    err = rel_err(xb, xa, w)
    say("fp32c vs fp64 max rel err", err.max())
    assert np.all(err <= 1e-6)
    ea = eng.atmosphere(a, c.g.maxi, c.g.maxj)
    eb = eng.atmosphere(b, c.g.maxi, c.g.maxj)
    assert np.all(np.abs(eb - ea) <= 1e-6 * np.abs(ea))
    eng.remove_member(a)
    eng.remove_member(b)


def test_airsea_graph_replay_equals_eager(eng):
    """Steps captured as a graph and replayed three times, with eager steps
    and downloads between, give the eager run's bits, atmosphere included."""
    c = make_case(18, 18, 8, kind="config2", seed=44)
    a = _member(eng, c)
    b = _member(eng, c)
    gid = eng.capture([b], 4)
    for r in range(3):
        eng.step([a], 4)
        eng.launch(gid)
        eng.sync()
        say("replay", r)
        assert np.array_equal(eng.download(a, c.ts.shape), eng.download(b, c.ts.shape))
        assert np.array_equal(eng.atmosphere(a, c.g.maxi, c.g.maxj), eng.atmosphere(b, c.g.maxi, c.g.maxj))
        assert np.array_equal(eng.airsea_flux(a, c.g.maxi, c.g.maxj), eng.airsea_flux(b, c.g.maxi, c.g.maxj))
    eng.drop_graph(gid)
    eng.remove_member(a)
    eng.remove_member(b)


def test_airsea_checkpoint_restore_continues_bitwise(eng):
    """A checkpoint holds the atmosphere and the wind: the restored member
    continues to the bits of the uninterrupted one."""
    c = make_case(18, 18, 8, kind="config2", seed=45)
    a = _member(eng, c)
    eng.step([a], 5)
    blob = eng.checkpoint(a)
    b = eng.add_member(eng.add_grid(c.g, c.k1), c.tset.L, E.FP64, c.diff)
    eng.restore(b, blob)
    eng.step([a], 6)
    eng.step([b], 6)
    assert np.array_equal(eng.download(a, c.ts.shape), eng.download(b, c.ts.shape))
    assert np.array_equal(eng.atmosphere(a, c.g.maxi, c.g.maxj), eng.atmosphere(b, c.g.maxi, c.g.maxj))
    eng.remove_member(a)
    eng.remove_member(b)
