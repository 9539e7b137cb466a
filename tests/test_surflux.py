"""EMBM's surface fluxes (surflux, embm.f90:2548-3738, without ENTS) in the
oracle, and the set-up tables of cgenie_amd.surface, pinned by closed forms
and budgets that no single planted mistake keeps:

- insolation (radfor): the equinox form S0 cos(lat)/pi and the global annual
  mean S0 / (4 sqrt(1 - e^2)); the albedo's latitude form;
- the wind speed usurf: the bulk formula |tau| = rho_air cd u^2 inverted,
  and the zonal mean of the polar rows;
- the three energy closures: the ocean's four components sum to fx0neto,
  the atmosphere's four to fx0a, and (fx0a + fx0o) moves only by the latent
  heat of the extra evaporation when the ocean warms (the sensible and
  longwave exchange cancel between air and sea);
- water: the rain equals the humidity the clamp removes, runoff from every
  land cell arrives at the cell readroff routes it to, and the ocean's runoff
  totals the land's rain;
- sea ice: at the converged ice temperature the surface flux balances the
  conduction through the ice (the Newton root), the growth cap;
- freezing point against Millero (1978), saturation humidity against Bolton
  (1980), 4 W m-2 of forcing per CO2 doubling, the seasonal cycle's period;
- the ocean's surface slot, through one tracer step of a resting uniform
  ocean: the top layer warms by F dt / (rho cp dz) and freshens by
  S0 P dt / dz.

The reference's EMBM module uses its netCDF module and cannot be compiled
here, so no compiled-reference comparison backs these pins.
"""
import math
import os

import numpy as np
import pytest

import oracle_ffi as O
from cgenie_amd import grid as G
from cgenie_amd import surface as SF

MAXI = MAXJ = 36
MAXK = 8
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

RHOAIR, CD, RHO0, HLV, CPSC = 1.25, 0.0013, 1.0E3, 2.501E6, 3981.1
DSC, USC, RSC = 5.0E3, 0.05, 6.37E6
SALN0 = 34.9
CONSIC = 2.166


def world():
    return G.read_world(os.path.join(GOLDEN, "worbe2"), MAXI, MAXJ)


def wind():
    return np.load(os.path.join(GOLDEN, "worbe2_wind.npz"))["stress"]


def make(ice=True, seed=5, **kw):
    g = G.make_grid(MAXI, MAXJ, MAXK)
    w = world()
    su = SF.surflux_setup(g, w.k1, **kw)
    atemp, ashum, sich, sica, tice = SF.atmosphere(g, w.k1, seed, ice=ice)
    so = O.SurfluxOracle(O.grid(MAXI, MAXJ, MAXK), w.k1, su, atemp, ashum, sich, sica, tice)
    so.set_wind(wind())
    return g, w, su, so


def ocean(g, w, so, seed=6):
    """A top ocean level consistent with the air: 1-3 K warmer, at the freezing point under ice."""
    rng = np.random.default_rng(seed)
    ot = np.maximum(so.atemp + rng.uniform(1.0, 3.0, so.atemp.shape), -1.5)
    ot = np.where(so.sica > 0, -1.95, ot)
    osal = rng.normal(0.2, 0.3, ot.shape)
    return ot, osal


def wet_of(w):
    return w.k1[1:MAXJ + 1, 1:MAXI + 1] <= MAXK


def out_of(so, name):
    return so.out[O.SF[name]]


# ---------------------------------------------------------------- set-up tables
def test_radfor_equinox_closed_form():
    g = G.make_grid(MAXI, MAXJ, MAXK)
    sol = SF.radfor(g, osce=0.0, oscsob=0.0)
    want = 1368.0 * np.array(g.c[1:MAXJ + 1]) / math.pi
    assert np.allclose(sol, want[None, :], rtol=1e-13, atol=0)


def test_radfor_global_annual_mean_and_seasons():
    g = G.make_grid(MAXI, MAXJ, MAXK)
    sol = SF.radfor(g)
    ds = np.array(g.ds[1:MAXJ + 1])
    gm = (sol.mean(axis=0) * ds).sum() / ds.sum()
    assert abs(gm / (1368.0 / (4 * math.sqrt(1 - 0.0167 ** 2))) - 1) < 1e-3
    # polar night and polar day: the south pole row dark at some step, and
    # the summer pole receiving more than the equator
    assert sol[:, 0].min() == 0.0 and sol[:, -1].max() > sol[:, MAXJ // 2].max()
    flat = SF.radfor(g, dosc=False)
    assert np.allclose(flat, sol.mean(axis=0)[None, :], rtol=1e-12)


def test_albedo_latitude_form():
    g = G.make_grid(MAXI, MAXJ, MAXK)
    alb = SF.climatological_albedo(g)
    s = np.array(g.s[1:MAXJ + 1])
    assert np.allclose(alb, (0.20 + 0.36 * s * s)[:, None], rtol=0, atol=1e-14)


def test_runoff_routing_worbe2():
    w = world()
    iro, jro = SF.runoff_routing(w.k1, MAXI, MAXJ, MAXK)
    wet = wet_of(w)
    J, I = np.meshgrid(np.arange(1, MAXJ + 1), np.arange(1, MAXI + 1), indexing="ij")
    assert np.all(iro[wet] == I[wet]) and np.all(jro[wet] == J[wet])
    assert np.all(w.k1[jro, iro] <= MAXK)
    # every land cell's first move follows its own compass code
    for j in range(1, MAXJ + 1):
        for i in range(1, MAXI + 1):
            code = w.k1[j, i]
            if code <= MAXK:
                continue
            di, dj = {91: (1, 0), 92: (0, -1), 93: (-1, 0), 94: (0, 1)}[int(code)]
            ni, nj = (i + di - 1) % MAXI + 1, j + dj
            assert (iro[j - 1, i - 1], jro[j - 1, i - 1]) == (iro[nj - 1, ni - 1], jro[nj - 1, ni - 1])


def test_runoff_routing_rejects_bad_masks():
    w = world()
    k1 = w.k1.copy()
    land = np.argwhere(k1[1:MAXJ + 1, 1:MAXI + 1] > MAXK)[0] + 1
    k1[land[0], land[1]] = 90
    with pytest.raises(ValueError, match="no direction"):
        SF.runoff_routing(k1, MAXI, MAXJ, MAXK)
    k1 = np.full_like(w.k1, 94)
    k1[1, 1] = 3
    with pytest.raises(ValueError, match="leaves the grid"):
        SF.runoff_routing(k1, MAXI, MAXJ, MAXK)


# ---------------------------------------------------------------- wind speed
def test_usurf_inverts_the_bulk_stress_formula():
    g, w, su, so = make(ice=False)
    tx, ty = 0.07, -0.04
    st = np.zeros((MAXJ, MAXI, 4))
    st[..., 0] = st[..., 2] = tx
    st[..., 1] = st[..., 3] = ty
    so.set_wind(st)
    want = math.sqrt(math.hypot(tx, ty) / (RHOAIR * CD))
    row1 = math.sqrt(math.hypot(tx, ty / 2) / (RHOAIR * CD))
    assert np.allclose(so.usurf[1:], want, rtol=1e-13)
    assert np.allclose(so.usurf[0], row1, rtol=1e-13)
    # the wind scale cancels
    su.scf = 3.0
    so2 = O.SurfluxOracle(O.grid(MAXI, MAXJ, MAXK), w.k1, su, so.atemp, so.ashum)
    so2.set_wind(st)
    assert np.allclose(so2.usurf, so.usurf, rtol=1e-14)


def test_usurf_polar_rows_are_zonal_means():
    g, w, su, so = make(ice=False)
    su.polar_avg = 2
    raw = O.SurfluxOracle(O.grid(MAXI, MAXJ, MAXK), w.k1, su, so.atemp, so.ashum)
    raw.set_wind(wind())
    for j in (0, 1, MAXJ - 2, MAXJ - 1):
        assert np.allclose(so.usurf[j], raw.usurf[j].mean(), rtol=1e-13)
        assert raw.usurf[j].std() > 0
    assert np.array_equal(so.usurf[2:MAXJ - 2], raw.usurf[2:MAXJ - 2])


# ---------------------------------------------------------------- budgets
def test_ocean_and_atmosphere_heat_closures():
    g, w, su, so = make(ice=True)
    ot, osal = ocean(g, w, so)
    so.step(ot, osal)
    wet = wet_of(w)
    ocn = sum(out_of(so, n) for n in ("fxlho", "fxsho", "fxswo", "fxlwo"))
    scale = np.abs(out_of(so, "fx0neto")).max()
    assert np.abs(ocn - out_of(so, "fx0neto"))[wet].max() <= 1e-12 * scale
    atm = sum(out_of(so, n) for n in ("fxlha", "fxsha", "fxswa", "fxlwa"))
    assert np.abs(atm - out_of(so, "fx0a")).max() <= 1e-12 * np.abs(out_of(so, "fx0a")).max()
    # signs: the warmer ocean loses sensible heat and longwave; shortwave enters
    open_ = wet & (so.sica == 0)
    assert np.all(out_of(so, "fxsho")[open_] < 0) and np.all(out_of(so, "fxlwo")[open_] < 0)
    assert np.all(out_of(so, "fxswo")[open_] >= 0)
    assert (so.sica > 0).sum() > 10


def test_air_sea_exchange_cancels_between_air_and_sea():
    """Warming the ocean changes fx0a + fx0o only through the latent heat of
    the extra evaporation: the sensible and longwave exchange leave one side
    and enter the other."""
    outs, evs = [], []
    for dT in (0.0, 1.5):
        g, w, su, so = make(ice=False)
        ot, osal = ocean(g, w, so)
        so.step(ot + dT, osal)
        outs.append(out_of(so, "fx0a") + out_of(so, "fx0o"))
        evs.append(out_of(so, "evap").copy())
        assert np.abs(out_of(so, "fxsen")[wet_of(w)]).min() > 0
    wet = wet_of(w)
    d = outs[1] - outs[0]
    want = -RHO0 * HLV * (evs[1] - evs[0])
    assert np.abs(want[wet]).max() > 1.0
    assert np.abs(d - want)[wet].max() <= 1e-9 * np.abs(outs[0][wet]).max()


def test_rain_is_the_humidity_the_clamp_removes():
    g, w, su, so = make(ice=False)
    so.ashum *= 1.3                     # supersaturate most cells
    before = so.ashum.copy()
    ot, osal = ocean(g, w, so)
    so.step(ot, osal)
    rain = out_of(so, "pptn")
    dq = before - so.ashum
    assert (dq > 0).sum() > 100 and np.all(rain[dq <= 0] == 0)
    assert np.allclose(rain, dq * (RHOAIR / RHO0) * 1800. * su.rdtdim, rtol=1e-13, atol=0)
    assert np.all(out_of(so, "pptn_atm") == rain * 1000.0)


def test_saturation_humidity_bolton():
    g, w, su, so = make(ice=False, rmax=1.0)
    rng = np.random.default_rng(3)
    so.atemp[...] = rng.uniform(-20.0, 30.0, so.atemp.shape)
    so.ashum[...] = 0.05                # far above saturation everywhere
    ot, osal = ocean(g, w, so)
    so.step(ot, osal)
    T = so.atemp
    es = 6.112 * np.exp(17.67 * T / (T + 243.5))          # hPa, Bolton (1980) eq. 10
    q = 0.622 * es / (1013.25 - 0.378 * es)
    assert np.all(np.abs(so.ashum / q - 1) < 0.03)
    assert abs(so.ashum[T == T.flat[0]][0] - q.flat[0]) / q.flat[0] < 0.03


def test_freezing_point_millero():
    g, w, su, so = make(ice=False)
    ot, _ = ocean(g, w, so)
    osal = np.random.default_rng(1).uniform(-4.0, 3.0, ot.shape)
    so.step(ot, osal)
    wet = wet_of(w)
    S = SALN0 + osal
    tf = -0.0575 * S + 1.710523e-3 * S ** 1.5 - 2.154996e-4 * S * S   # Millero (1978), p = 0
    assert np.abs(out_of(so, "tsfreez") - tf)[wet].max() < 0.03


def test_runoff_arrives_where_routed_and_totals_the_land_rain():
    g, w, su, so = make(ice=False)
    wet = wet_of(w)
    ot, osal = ocean(g, w, so)
    qs = 3.80E-3 * np.exp(17.67 * so.atemp / (so.atemp + 243.5))
    land = np.argwhere(~wet)
    jl, il = land[len(land) // 3]
    so.ashum[...] = 0.5 * qs
    so.ashum[jl, il] = 1.2 * qs[jl, il]
    so.step(ot, osal)
    ro = out_of(so, "runoff_ocn")
    dest = (su.jroff[jl, il] - 1, su.iroff[jl, il] - 1)
    assert np.count_nonzero(ro) == 1 and ro[dest] > 0
    assert ro[dest] == out_of(so, "pptn")[jl, il] * 1000.0
    assert out_of(so, "runoff_land")[jl, il] == ro[dest]
    for igrid in (0, 1):
        g, w, su, so = make(ice=False, igrid=igrid)
        if igrid:
            # igrid 1 and 2 are grids of unequal rows: weight by a row width
            # that varies, so that the area weighting shows
            for j in range(1, MAXJ + 1):
                so.g.ds[j] = 0.02 + 0.05 * math.cos(math.asin(g.s[j]))
                so.g.rds[j] = 1.0 / so.g.ds[j]
        so.ashum *= 1.3
        so.step(ot, osal)
        ro, rain = out_of(so, "runoff_ocn"), out_of(so, "pptn")
        ds = np.array(so.g.ds[1:MAXJ + 1])[:, None] * np.ones((1, MAXI))
        wgt = ds if igrid else np.ones_like(ds)
        tot_in, tot_out = (rain * 1000.0 * wgt)[~wet].sum(), (ro * wgt)[wet].sum()
        assert tot_in > 0 and abs(tot_out / tot_in - 1) < 1e-12
        # rain over the ocean stays in pptn_ocn, plus pmeadj (zero here)
        assert np.allclose(out_of(so, "pptn_ocn")[wet], rain[wet] * 1000.0, rtol=1e-15)
        assert np.all(out_of(so, "pptn_ocn")[~wet] == 0) and np.all(out_of(so, "runoff_ocn")[~wet] == 0)


def test_freshwater_adjustment_and_evaporation_signs():
    g = G.make_grid(MAXI, MAXJ, MAXK)
    w = world()
    pm = np.random.default_rng(2).normal(0.0, 1e-8, (MAXJ, MAXI))
    su = SF.surflux_setup(g, w.k1, pmeadj=pm)
    atemp, ashum, *_ = SF.atmosphere(g, w.k1, 5)
    so = O.SurfluxOracle(O.grid(MAXI, MAXJ, MAXK), w.k1, su, atemp, ashum)
    so.set_wind(wind())
    ot, osal = ocean(g, w, so)
    so.step(ot, osal)
    wet = wet_of(w)
    assert np.allclose(out_of(so, "pptn_ocn")[wet], (out_of(so, "pptn") + pm)[wet] * 1000.0, rtol=1e-12, atol=1e-20)
    assert np.all(out_of(so, "evap_ocn") == -out_of(so, "evap_atm"))
    assert np.all(out_of(so, "evap_ocn")[wet] <= 0)
    # saturated air over the ocean takes up nothing
    so2 = O.SurfluxOracle(O.grid(MAXI, MAXJ, MAXK), w.k1, su, atemp, np.full_like(ashum, 0.2), None, None, None)
    so2.su.rmax = 100.0
    so2.P.rmax = 100.0
    so2.set_wind(wind())
    so2.step(ot, osal)
    assert np.all(out_of(so2, "evap") == 0)


def test_bulk_transfer_coefficient_is_observed_magnitude():
    """The sensible heat flux over open water is rho_air cp_air C_H u (To - Ta)
    with C_H of order 1e-3 (observed air-sea transfer coefficients), larger
    in unstable air and stronger wind."""
    g, w, su, so = make(ice=False)
    ot, osal = ocean(g, w, so)
    so.step(ot, osal)
    wet = wet_of(w)
    dT = ot - so.atemp
    m = wet & (so.usurf > 3) & (so.usurf < 12) & (np.abs(dT) > 0.5)
    ch = out_of(so, "fxsen")[m] / (RHOAIR * 1004.0 * so.usurf[m] * dT[m])
    assert m.sum() > 50 and ch.min() > 0.8e-3 and ch.max() < 1.8e-3
    # stronger wind, larger coefficient (same stability class)
    u = so.usurf[m]
    assert np.corrcoef(u, ch)[0, 1] > 0.3


def test_sea_ice_albedo_falls_with_air_temperature():
    g, w, su, so = make(ice=True)
    ot, osal = ocean(g, w, so)
    so.step(ot, osal)
    ice = so.sica > 0
    a, T = out_of(so, "albice")[ice], so.atemp[ice]
    order = np.argsort(T)
    assert np.all(np.diff(a[order]) <= 0) and a.max() <= 0.7 and a.min() >= 0.2
    assert a[order][0] > a[order][-1]


# ---------------------------------------------------------------- sea ice
def test_sea_ice_temperature_is_the_surface_energy_balance_root():
    g, w, su, so = make(ice=True)
    ot, osal = ocean(g, w, so)
    so.step(ot, osal)
    ice = so.sica > 0
    tf = out_of(so, "tsfreez")
    bal = so.sich * out_of(so, "fx0sic") + CONSIC * (tf - so.tice)
    cond = np.abs(CONSIC * (tf - so.tice))
    free = ice & (so.tice < 0)
    assert free.sum() > 10
    assert np.all(np.abs(bal[free]) <= 1e-6 * np.maximum(cond[free], 1.0))
    assert np.all(so.iters[ice] >= 1) and np.all(so.iters[ice] <= 21)
    # where the iteration ends above freezing the ice surface melts: tice = 0
    # and the balance holds a surplus
    melt = ice & (so.tice == 0)
    assert np.all(bal[melt] > 0)
    assert np.all(so.tice[~ice] == 0)


def test_sea_ice_growth_cap():
    g, w, su, so = make(ice=True)
    ot, osal = ocean(g, w, so)
    so.step(ot, osal)
    free = O.SurfluxOracle(O.grid(MAXI, MAXJ, MAXK), w.k1, su, so.atemp, so.ashum, so.sich * 0, so.sica * 0)
    free.usurf[...] = so.usurf
    free.step(ot, osal)
    dho = out_of(free, "dthsic")          # open-water growth; the ice does not change it
    s = so.sica
    dhsic = np.where(s > 0, (out_of(so, "dthsic") - (1 - s) * dho) / np.where(s > 0, s, 1), 0)
    grow = (s > 0) & (dhsic > 1e-12)
    assert grow.sum() > 0
    su2 = SF.surflux_setup(g, w.k1, sich_max=0.1)
    atemp, ashum, sich, sica, tice = SF.atmosphere(g, w.k1, 5, ice=True)
    cap = O.SurfluxOracle(O.grid(MAXI, MAXJ, MAXK), w.k1, su2, atemp, ashum, sich, sica, tice)
    cap.usurf[...] = so.usurf
    cap.step(ot, osal)
    got = out_of(cap, "dthsic")[grow]
    assert np.allclose(got, ((1 - s) * dho)[grow], rtol=1e-9, atol=1e-20)
    assert np.all(out_of(cap, "dthsic")[(s > 0) & ~grow] == out_of(so, "dthsic")[(s > 0) & ~grow])


# ---------------------------------------------------------------- forcing
def test_co2_doubling_forcing():
    res = []
    for p in (278.0E-6, 2 * 278.0E-6):
        g, w, su, so = make(ice=True, radfor_atchem=1)
        ot, osal = ocean(g, w, so)
        so.step(ot, osal, np.full((MAXJ, MAXI), p))
        res.append(out_of(so, "fx0a").copy())
    d = res[1] - res[0]
    assert np.allclose(d, 5.77 * math.log(2.0), rtol=1e-9)
    assert abs(5.77 * math.log(2.0) - 4.0) < 0.01
    # a pCO2 below a thousandth of the reference reads as the reference
    g, w, su, so = make(ice=True, radfor_atchem=1)
    so.step(ot, osal, np.full((MAXJ, MAXI), 1e-9))
    assert np.array_equal(out_of(so, "fx0a"), res[0])


def test_compound_co2_rise():
    rate = 0.01
    g, w, su, so = make(ice=False)
    su.rate_co2 = rate
    so.P.rate_co2 = rate
    base = make(ice=False)[3]
    ot, osal = ocean(g, w, so)
    for n in range(1, 4):
        so.step(ot, osal)
        base.step(ot, osal)
        assert so.S.co2 == pytest.approx(278.0E-6 * (1 + rate) ** n, rel=1e-14)
    d = out_of(so, "fx0a") - out_of(base, "fx0a")
    # the lower outgoing longwave shows in fx0a; the rain was spent on the first call
    assert np.allclose(d, 5.77 * math.log((1 + rate) ** 3), rtol=1e-8)


def test_seasonal_cycle_period():
    g, w, su, so = make(ice=False)
    ot, osal = ocean(g, w, so)
    sw = []
    for n in range(su.nyear + 1):
        so.step(ot, osal)
        if n in (0, 1, su.nyear):
            sw.append(out_of(so, "fxsw").copy())
    assert np.array_equal(sw[0], sw[2]) and not np.array_equal(sw[0], sw[1])
    assert np.allclose(sw[0], su.solfor[0][:, None] * (1 - su.albcl), rtol=1e-15)


# ---------------------------------------------------------------- the ocean's slot
def test_surface_slot_heats_and_freshens_the_top_layer():
    """One tracer step of a resting, uniform ocean with a uniform surface
    flux: the top layer's temperature rises by F dt / (rho0 cp dz) and its
    salinity falls by S0 P dt / dz, the layers below are untouched."""
    g, w, su, so = make(ice=False)
    og = so.g
    F, P = 150.0, 3e-8            # W m-2 into the ocean; m s-1 of fresh water
    so.out[...] = 0.0
    so.out[O.SF["fxswo"]] = F
    so.out[O.SF["pptn_ocn"]] = P * 1000.0
    sflux = so.slot(2)
    ts = np.zeros((MAXK, MAXJ, MAXI, 2))
    ts[..., 0], ts[..., 1] = 10.0, 0.0
    gold = O.GoldsteinOracle(og, w.k1, ts, np.zeros((MAXK, MAXJ + 1, MAXI + 1, 3)), sflux=sflux, diso=False)
    gold.step(1)
    wet = wet_of(w)
    dt_s = og.dt[MAXK] * og.tsc
    dz_m = og.dz[MAXK] * DSC
    dT = gold.ts[MAXK - 1, ..., 0] - 10.0
    dS = gold.ts[MAXK - 1, ..., 1]
    assert np.allclose(dT[wet], F * dt_s / (RHO0 * CPSC * dz_m), rtol=1e-9)
    assert np.allclose(dS[wet], -SALN0 * P * dt_s / dz_m, rtol=1e-9)
    below = gold.ts[:MAXK - 1, ..., 0][np.broadcast_to(wet, (MAXK - 1,) + wet.shape)]
    assert np.abs(below - 10.0).max() <= 1e-12     # rounding of the uniform field's zero fluxes
