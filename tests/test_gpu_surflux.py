"""EMBM's surface fluxes on the GPU (cgo_grid_set_surflux, cgo_set_surface,
cgo_surflux, cgo_step) against the oracle's og_surflux (tests/test_surflux.py
pins that one), through the C ABI, on the reference's worbe2 world
(36 x 36 x 8) with its own wind stress and runoff routing.  Needs an MI355X.

Bar: the kernel is the oracle's arithmetic in the oracle's order, except
that exp, log and pow are the device's, which may round an ulp away from
glibc's.  So one call's outputs match the oracle's to 1e-12 relative to
each field's scale (the sea-ice temperature, a Newton iterate converged to
1e-10, to 1e-9 absolute), and a coupled run -- surface fluxes, velocity
solve and tracer step, 24 steps -- matches the oracle's to 1e-10 relative
per cell; the counts of bit-identical values are printed.  The FP32 form
follows FP64 to 1e-6 of each tracer's scale."""
import os
import sys

import numpy as np
import pytest

from cgenie_amd import engine as E
from cgenie_amd import grid as G
from cgenie_amd import surface as SF
import oracle_ffi as O

pytestmark = pytest.mark.gpu

MAXI, MAXJ, MAXK = 36, 36, 8
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
DIFF = G.diffusivities(2000.0, 2.7e-5)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = E.Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def world():
    return G.read_world(os.path.join(GOLDEN, "worbe2"), MAXI, MAXJ)


def stress():
    return np.load(os.path.join(GOLDEN, "worbe2_wind.npz"))["stress"]


def state(k1, L=4, seed=3):
    """T from 28 degC at the equator to -1.9 at the poles at the top,
    stratified below; S anomalies; L - 2 dyes; zero on land."""
    rng = np.random.default_rng(seed)
    g = G.make_grid(MAXI, MAXJ, MAXK)
    lat = np.arcsin(np.asarray(g.s[1:MAXJ + 1]))
    top = -1.9 + 29.9 * np.cos(lat) ** 2
    ts = np.zeros((MAXK, MAXJ, MAXI, L))
    frac = (np.arange(MAXK) + 1.0) / MAXK
    ts[..., 0] = 1.0 + (top[None, :, None] - 1.0) * frac[:, None, None] ** 2 + \
        0.05 * rng.standard_normal((MAXK, MAXJ, MAXI))
    ts[..., 1] = 0.3 * rng.standard_normal((MAXK, MAXJ, MAXI))
    for l in range(2, L):
        ts[..., l] = rng.uniform(0.5, 1.5, (MAXK, MAXJ, MAXI))
    wet = np.arange(1, MAXK + 1)[:, None, None] >= k1[None, 1:-1, 1:-1]
    ts[~wet] = 0.0
    return ts


def wet3(k1):
    return np.arange(1, MAXK + 1)[:, None, None] >= k1[None, 1:-1, 1:-1]


def new_member(eng, world, prec=E.FP64, L=4, su_kw=None, ice=True, mom=True, seed=3):
    g = G.make_grid(MAXI, MAXJ, MAXK)
    gid = eng.add_grid(g, world.k1)
    su = SF.surflux_setup(g, world.k1, **(su_kw or {}))
    eng.set_surflux(gid, su)
    if mom:
        eng.set_momentum(gid, g, world)
    m = eng.add_member(gid, L, prec, DIFF, diso=True)
    ts = state(world.k1, L, seed)
    eng.upload(m, ts)
    eng.set_wind(m, stress())
    atm = SF.atmosphere(g, world.k1, seed + 10, ice=ice)
    eng.set_surface(m, *atm)
    return g, gid, m, ts, su, atm


def surf_oracle(world, su, atm):
    so = O.SurfluxOracle(O.grid(MAXI, MAXJ, MAXK), world.k1, su, *atm)
    so.set_wind(stress())
    return so


def compare_outputs(dev, so, label, rel=1e-12):
    worst = 0.0
    same = total = 0
    for q, n in enumerate(O.SF_NAMES):
        a, b = dev[n], so.out[q]
        scale = max(np.abs(b).max(), 1e-300)
        e = np.abs(a - b).max() / scale
        worst = max(worst, e)
        same += np.count_nonzero(a == b)
        total += a.size
        assert e <= rel, (label, n, e)
    say(f"  {label}: outputs max rel err {worst:.3g}, bit-identical {same} of {total}")
    assert np.abs(dev["tice"] - so.tice).max() <= 1e-9
    assert np.abs(dev["ashum"] - so.ashum).max() <= rel * np.abs(so.ashum).max()


@pytest.mark.parametrize("ice,kw", [(False, {}), (True, {}), (True, {"olr_adj": 0.5, "olr_adj0": 1.0}),
                                    (True, {"igrid": 1, "rate_co2": 0.001})])
def test_one_call_matches_oracle(eng, world, ice, kw):
    say("surflux one call, ice", ice, kw)
    g, gid, m, ts, su, atm = new_member(eng, world, su_kw=kw, ice=ice, mom=False)
    so = surf_oracle(world, su, atm)
    for it in range(2):
        eng.surflux([m])
        so.step(ts[MAXK - 1, ..., 0], ts[MAXK - 1, ..., 1])
        dev = eng.surface_outputs(m, g)
        compare_outputs(dev, so, f"call {it + 1}")
        assert dev["istep"] == it + 1 and dev["co2"] == so.S.co2
    assert (so.sica > 0).any() == ice
    assert np.abs(dev["runoff_ocn"]).max() > 0 or not (so.out[O.SF["pptn"]] > 0).any()
    eng.remove_member(m)


class CoupledOracle:
    """og_surflux, step_goldstein's surface slot, og_mom_step and
    og_goldstein_step per ocean step, in genie.f90's order; the fluxes read
    the top level the last tracer step left."""

    def __init__(self, world, ts, su, atm):
        self.og = O.grid(MAXI, MAXJ, MAXK)
        self.so = surf_oracle(world, su, atm)
        self.mo = O.MomOracle(self.og, world.k1, world.labels, world.isles, world.npi, world.paths)
        self.mo.set_wind(stress())
        self.u = np.zeros((MAXK, MAXJ + 1, MAXI + 1, 3))
        self.gold = O.GoldsteinOracle(self.og, world.k1, ts, self.u, sflux=np.zeros((MAXJ, MAXI, ts.shape[-1])))
        self.gold.rho[...] = O.rho_from_ts(self.og, world.k1, ts)
        self.L = ts.shape[-1]

    def step(self, n=1):
        for _ in range(n):
            top = self.gold.ts[MAXK - 1]
            self.so.step(top[..., 0].copy(), top[..., 1].copy())
            self.gold.sflux[...] = self.so.slot(self.L)
            self.mo.step(self.gold.rho, self.u)
            self.gold.step(1)


def rel_cells(a, b, mask):
    aw, bw = a[mask], b[mask]
    scale = np.maximum(np.abs(bw).max(axis=0), 1e-300)
    return float((np.abs(aw - bw) / np.maximum(np.abs(bw), 1e-3 * scale)).max())


def test_coupled_fp64_matches_oracle(eng, world):
    say("surflux + velocity solve + tracer step, 24 steps, FP64")
    g, gid, m, ts, su, atm = new_member(eng, world)
    orc = CoupledOracle(world, ts, su, atm)
    eng.step([m], 24)
    orc.step(24)
    got = eng.download(m, ts.shape)
    w = wet3(world.k1)
    err = rel_cells(got, orc.gold.ts, w)
    say("  state max rel err", err, "bit-identical cells", np.count_nonzero(got[w] == orc.gold.ts[w]), "of",
        got[w].size)
    assert err <= 1e-10
    dev = eng.surface_outputs(m, g)
    compare_outputs(dev, orc.so, "step 24", rel=1e-10)
    # the fluxes move the surface: the polar top layer cools, the tropics warm
    d = got[MAXK - 1, ..., 0] - ts[MAXK - 1, ..., 0]
    assert np.abs(d[w[MAXK - 1]]).max() > 1e-3
    eng.remove_member(m)


def test_fp32c_tracks_fp64(eng, world):
    say("FP32C member with surface fluxes follows FP64, 24 steps")
    g, _, m64, ts, su, atm = new_member(eng, world, E.FP64)
    _, _, m32, _, _, _ = new_member(eng, world, E.FP32C)
    eng.step([m64, m32], 24)
    a, b = eng.download(m32, ts.shape), eng.download(m64, ts.shape)
    # each tracer's error relative to its largest value: the FP32 form keeps
    # ~40 significant bits of every value, so anomalies near zero (S) carry
    # the absolute error of their tracer's scale
    w = wet3(world.k1)
    err = max(float(np.abs(a[..., l][w] - b[..., l][w]).max() / np.abs(b[..., l][w]).max()) for l in range(ts.shape[-1]))
    o32, o64 = eng.surface_outputs(m32, g), eng.surface_outputs(m64, g)
    eo = max(np.abs(o32[n] - o64[n]).max() / max(np.abs(o64[n]).max(), 1e-300) for n in E.SF_NAMES)
    say("  state rel err", err, "outputs rel err", eo)
    assert err <= 1e-6 and eo <= 1e-5
    eng.remove_member(m64)
    eng.remove_member(m32)


def test_graph_replay_equals_eager(eng, world):
    say("surface fluxes in a replayed graph = eager steps")
    g, _, a, ts, _, _ = new_member(eng, world)
    _, _, b, _, _, _ = new_member(eng, world)
    eng.step([a], 8)
    gr = eng.capture([b], 2)
    for _ in range(4):
        eng.launch(gr, 1)
        eng.sync()
    eng.drop_graph(gr)
    assert np.array_equal(eng.download(a, ts.shape), eng.download(b, ts.shape))
    oa, ob = eng.surface_outputs(a, g), eng.surface_outputs(b, g)
    for n in list(E.SF_NAMES) + ["ashum", "tice"]:
        assert np.array_equal(oa[n], ob[n]), n
    assert oa["istep"] == ob["istep"] == 8
    eng.remove_member(a)
    eng.remove_member(b)


def test_prognostic_graph_replay_equals_eager(eng, world):
    """The whole prognostic step in a replayed graph: surface fluxes with the
    members' own ATCHEM CO2, the velocity solve (FP32 members: one dense
    product for the grid), transport, BIOGEM with air-sea exchange on the
    second stream (forked after the solve's first column kernel) and the
    coupling.  Two members captured for 2 steps and replayed 4 times, two
    more stepped eagerly 8 times: state, atmosphere and surface outputs
    equal bit for bit."""
    say("prognostic step (surflux + velocity + BIOGEM/ATCHEM) replayed = eager")
    g = G.make_grid(MAXI, MAXJ, MAXK)
    tset = G.tracer_set("config2")
    ts = G.make_tracers(g, world.k1, tset, 61)
    light, mld = G.make_light(g, 62), G.make_mld(g, 63, 0.0)
    su = SF.surflux_setup(g, world.k1, radfor_atchem=1)
    atm = SF.atmosphere(g, world.k1, 64, ice=True)
    gid = eng.add_grid(g, world.k1)
    eng.set_momentum(gid, g, world)
    eng.set_surflux(gid, su)
    ms = []
    for q in range(4):
        m = eng.add_member(gid, tset.L, E.FP32C, DIFF, diso=True)
        eng.upload(m, ts)
        eng.bgc_init(m, E.default_bgc_params(1.0 / 96, tset, carb_tol=1e-12, carb_maxit=1000, airsea=1,
                                             atm_pCO2=400e-6), light, mld)
        eng.set_windspeed(m, np.full((MAXJ, MAXI), 8.0))
        eng.set_wind(m, stress())
        eng.set_surface(m, *atm)
        ms.append(m)
    gr = eng.capture(ms[:2], 2)
    for _ in range(4):
        eng.launch(gr, 1)
        eng.sync()
    eng.drop_graph(gr)
    eng.step(ms[2:], 8)
    for a, b in ((ms[0], ms[2]), (ms[1], ms[3])):
        assert np.array_equal(eng.download(a, ts.shape), eng.download(b, ts.shape))
        assert np.array_equal(eng.velocity(a, g), eng.velocity(b, g))
        assert np.array_equal(eng.atmosphere(a, MAXI, MAXJ), eng.atmosphere(b, MAXI, MAXJ))
        oa, ob = eng.surface_outputs(a, g), eng.surface_outputs(b, g)
        for n in list(E.SF_NAMES) + ["ashum", "tice", "co2"]:
            assert np.array_equal(oa[n], ob[n]), n
    for m in ms:
        eng.remove_member(m)


def test_checkpoint_continues_bitwise(eng, world):
    say("checkpoint of a member with surface fluxes continues bit for bit")
    g, gid, a, ts, su, atm = new_member(eng, world)
    eng.step([a], 3)
    blob = eng.checkpoint(a)
    eng.step([a], 3)
    b = eng.add_member(gid, ts.shape[-1], E.FP64, DIFF, diso=True)
    eng.restore(b, blob)
    eng.step([b], 3)
    assert np.array_equal(eng.download(a, ts.shape), eng.download(b, ts.shape))
    oa, ob = eng.surface_outputs(a, g), eng.surface_outputs(b, g)
    for n in list(E.SF_NAMES) + ["ashum", "tice", "istep", "co2"]:
        assert np.array_equal(oa[n], ob[n]), n
    eng.remove_member(a)
    eng.remove_member(b)


def test_biogem_atchem_co2_forcing_matches_oracle(eng, world):
    """Surface fluxes with the CO2 of the member's own ATCHEM atmosphere
    (atchem_radfor), BIOGEM with air-sea exchange, 12 steps in genie.f90's
    order against the oracle: the fluxes read the top level as tstepo left
    it, before the tracer coupling rewrites T and S."""
    say("surflux + BIOGEM + air-sea + ATCHEM CO2, 12 steps")
    g = G.make_grid(MAXI, MAXJ, MAXK)
    og = O.grid(MAXI, MAXJ, MAXK)
    tset = G.tracer_set("config2")
    ts = G.make_tracers(g, world.k1, tset, 45)
    u = G.make_velocity(g, world.k1, 46)
    light, mld = G.make_light(g, 47), G.make_mld(g, 48, 0.0)
    su = SF.surflux_setup(g, world.k1, radfor_atchem=1)
    atm = SF.atmosphere(g, world.k1, 49, ice=True)
    wspeed = np.random.default_rng(50).uniform(2.0, 14.0, (MAXJ, MAXI))
    gid = eng.add_grid(g, world.k1)
    eng.set_surflux(gid, su)
    m = eng.add_member(gid, tset.L, E.FP64, DIFF)
    eng.set_velocity(m, u)
    eng.upload(m, ts)
    pco2 = 560e-6
    eng.bgc_init(m, E.default_bgc_params(1.0 / 96, tset, carb_tol=1e-12, carb_maxit=1000, airsea=1, atm_pCO2=pco2),
                 light, mld)
    eng.set_windspeed(m, wspeed)
    eng.set_wind(m, stress())
    eng.set_surface(m, *atm)
    gold = O.GoldsteinOracle(og, world.k1, ts, u, sflux=np.zeros((MAXJ, MAXI, tset.L)))
    bio = O.BiogemOracle(gold, O.bgc_params(1.0 / 96, tset, airsea=1, atm_pCO2=pco2), light, mld)
    bio.set_windspeed(wspeed)
    so = surf_oracle(world, su, atm)
    top = ts[MAXK - 1, ..., :2].copy()
    n = 12
    eng.step([m], n)
    for _ in range(n):
        so.step(top[..., 0], top[..., 1], bio.atm()[..., 0])
        gold.sflux[...] = so.slot(tset.L)
        gold.step(1)
        top = gold.ts[MAXK - 1, ..., :2].copy()
        bio.step(1, goldstein=False)
    w = wet3(world.k1)
    err = rel_cells(eng.download(m, ts.shape), gold.ts, w)
    say("  state max rel err", err)
    assert err <= 1e-10
    dev = eng.surface_outputs(m, g)
    compare_outputs(dev, so, "step 12", rel=1e-10)
    a = eng.atmosphere(m, MAXI, MAXJ)
    assert abs(a[0, 0, 0] / bio.atm()[0, 0, 0] - 1) <= 1e-10
    # the forcing of the member's own CO2 shows in the outgoing longwave:
    # a member at the reference CO2 has fx0a lower by delf2x ln(pCO2 / 278e-6)
    assert a[0, 0, 0] > 400e-6
    eng.remove_member(m)


def test_refusals(eng, world):
    say("surface-flux refusals")
    g, gid, m, ts, su, atm = new_member(eng, world, mom=False)
    with pytest.raises(E.CgoError, match="computes its surface slot"):
        eng.set_sflux(m, np.zeros((MAXJ, MAXI, ts.shape[-1])))
    eng.set_surface(m, None)
    eng.set_sflux(m, np.zeros((MAXJ, MAXI, ts.shape[-1])))
    with pytest.raises(E.CgoError, match="no surface fluxes"):
        eng.surflux([m])
    with pytest.raises(E.CgoError, match="already has"):
        eng.set_surflux(gid, su)
    g2 = G.make_grid(MAXI, MAXJ, MAXK)
    gid2 = eng.add_grid(g2, world.k1)
    eng.set_surflux(gid2, SF.surflux_setup(g2, world.k1, radfor_atchem=1))
    m2 = eng.add_member(gid2, 4, E.FP64, DIFF)
    eng.upload(m2, ts)
    eng.set_surface(m2, *atm)
    with pytest.raises(E.CgoError, match="atmosphere it does not have"):
        eng.step([m2], 1)
    bad = SF.surflux_setup(g2, world.k1)
    land = np.argwhere(world.k1[1:-1, 1:-1] > MAXK)[0]
    bad.iroff = bad.iroff.copy()
    bad.jroff = bad.jroff.copy()
    bad.iroff[land[0], land[1]], bad.jroff[land[0], land[1]] = land[1] + 1, land[0] + 1
    gid3 = eng.add_grid(g2, world.k1)
    with pytest.raises(E.CgoError, match="not an ocean cell"):
        eng.set_surflux(gid3, bad)
    for x in (m, m2):
        eng.remove_member(x)


def test_surface_slot_with_sea_ice_fluxes(eng, world):
    """cgo_surface_slot: the sea-ice model's heat and freshwater, handed over
    after surflux (as genie.f90 orders them), join the slot the call formed;
    one tstepo with it against the oracle's slot and tracer step."""
    say("surface slot with the sea-ice model's fluxes")
    g, gid, m, ts, su, atm = new_member(eng, world, mom=False)
    u = G.make_velocity(g, world.k1, 46)
    eng.set_velocity(m, u)
    rng = np.random.default_rng(9)
    ice = atm[3] > 0
    fx0sic = np.where(ice, rng.normal(-30.0, 10.0, ice.shape), 0.0)
    fwsic = np.where(ice, rng.normal(0.0, 2e-5, ice.shape), 0.0)
    eng.surflux([m])
    eng.surface_slot(m, fx0sic, fwsic)
    eng.tstepo([m], 1)
    so = surf_oracle(world, su, atm)
    so.step(ts[MAXK - 1, ..., 0], ts[MAXK - 1, ..., 1])
    gold = O.GoldsteinOracle(O.grid(MAXI, MAXJ, MAXK), world.k1, ts, u, sflux=so.slot(ts.shape[-1], fx0sic, fwsic))
    gold.step(1)
    err = rel_cells(eng.download(m, ts.shape), gold.ts, wet3(world.k1))
    say("  state max rel err", err)
    assert err <= 1e-12
    # without the sea-ice fluxes the state differs where there is ice
    gold0 = O.GoldsteinOracle(O.grid(MAXI, MAXJ, MAXK), world.k1, ts, u, sflux=so.slot(ts.shape[-1]))
    gold0.step(1)
    assert np.abs(gold.ts[MAXK - 1, ..., 0] - gold0.ts[MAXK - 1, ..., 0])[ice].max() > 1e-6
    eng.remove_member(m)
