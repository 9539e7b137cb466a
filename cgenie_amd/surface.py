"""Set-up of EMBM's surface fluxes (surflux) for a grid: the tables and
constants initialise_embm leaves for the run, in the EMBM configuration
without ENTS (embm.f90:198-1530, defaults of embm-defaults.nml).

The tables are plain arrays in the reference's shapes, so a cGENIE host can
pass its own EMBM arrays instead:

    solfor  (nyear, maxj)   solfor(j, m) at [m-1, j-1]    radfor, embm.f90:2383-2522
    albcl   (maxj, maxi)    albcl(i, j)                    embm.f90:911-920
    ca      (maxj, maxi)    ca(i, j)                       embm.f90:924-929
    pmeadj  (maxj, maxi)    pmeadj(i, j)                   embm.f90:1292-1345
    iroff, jroff (maxj, maxi) int32, 1-based               readroff, embm.f90:3779-3836
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .grid import Grid, _powi

# embm_lib.f90:36-42
USC, RSC, DSC, FSC = 0.05, 6.37E6, 5.0E3, 2 * 7.2921E-5
RH0SC = 1.0E3
TSC = RSC / USC
RHO0 = 1.0E3
CO20, CH40, N2O0 = 278.0E-6, 700.0E-9, 275.0E-9   # embm_lib.f90:59

# readroff's compass codes for a dry cell's runoff direction (embm.f90:3791)
IROE, IROS, IROW, IRON = 91, 92, 93, 94


@dataclass
class SurfluxSetup:
    solfor: np.ndarray
    albcl: np.ndarray
    ca: np.ndarray
    pmeadj: np.ndarray
    iroff: np.ndarray
    jroff: np.ndarray
    igrid: int = 0
    polar_avg: int = 0          # par_wind_polar_avg
    radfor_atchem: int = 0      # atchem_radfor = 'y'
    scf: float = 2.0
    rmax: float = 0.85
    delf2x: float = 5.77
    rdtdim: float = 0.0
    rsictscsf: float = 0.0
    olr_adj: float = 0.0
    olr_adj0: float = 0.0
    t_eqm: float = 12.371
    albsic_min: float = 0.2
    albsic_max: float = 0.7
    sich_max: float = 9999.9
    rate_co2: float = 0.0
    rate_ch4: float = 0.0
    rate_n2o: float = 0.0
    co2: float = CO20           # the greenhouse gases' initial values, radfor_scl_* x the references
    ch4: float = CH40
    n2o: float = N2O0
    extra: dict = field(default_factory=dict)

    @property
    def nyear(self) -> int:
        return self.solfor.shape[0]


def radfor(g: Grid, solconst: float = 1368.0, daysperyear: float = 365.25, dosc: bool = True,
           osce: float = 0.0167, oscsob: float = 0.397789, oscgam: float = 1.352631,
           osctau0: float = -0.5) -> np.ndarray:
    """Daily-mean top-of-atmosphere insolation solfor(j, istep) for each of
    the nyear ocean steps of a year at a fixed orbit (radfor, embm.f90:2383-2522
    with orbit_radfor = 'n'); dosc = .false. replaces the seasonal cycle by its
    annual mean.  Returns (nyear, maxj)."""
    nyear, maxj = g.nyear, g.maxj
    pi = 4 * math.atan(1.0)
    rpi = 1.0 / pi
    tv = osce * osce
    osce1 = osce * (2.0 - 0.25 * tv)
    osce2 = 1.25 * tv
    osce3 = osce * tv * 13. / 12.
    q = (1.0 + 0.5 * tv) / (1.0 - tv)
    osce4 = q * q
    oscryr = 2.0 * pi / float(nyear)
    osctau1 = osctau0 + 0.5
    out = np.zeros((nyear, maxj))
    for istep in range(1, nyear + 1):
        osct = (float((istep - 1) % nyear + 1) - (nyear * osctau1 / daysperyear)) * oscryr
        for j in range(1, maxj + 1):
            oscv = osct + osce1 * math.sin(osct) + osce2 * math.sin(2.0 * osct) + osce3 * math.sin(3.0 * osct)
            a = 1.0 + osce * math.cos(oscv)
            oscsolf = osce4 * (a * a)
            oscsind = oscsob * math.sin(oscv - oscgam)
            oscss = oscsind * g.s[j]
            osccc = math.sqrt(1.0 - oscsind * oscsind) * g.c[j]
            osctt = min(1.0, max(-1.0, oscss / osccc))
            oscday = math.acos(-osctt)
            out[istep - 1, j - 1] = solconst * oscsolf * rpi * (oscss * oscday + osccc * math.sin(oscday))
    if not dosc:
        for j in range(maxj):
            acc = 0.0
            for istep in range(nyear):
                acc = acc + out[istep, j]
            out[:, j] = acc / nyear
    return out


def climatological_albedo(g: Grid, offs: float = 0.20, amp: float = 0.36, skew: float = 0.0, skewp: int = 0,
                          mod2: float = 0.0, mod4: float = 0.0, mod6: float = 0.0) -> np.ndarray:
    """albcl(i, j), a function of latitude only (embm.f90:911-920).  (maxj, maxi)."""
    out = np.zeros((g.maxj, g.maxi))
    for j in range(1, g.maxj + 1):
        scl = _powi((skew - g.s[j]) / 2.0, skewp)
        tv = math.asin(g.s[j])
        out[j - 1, :] = offs + amp * 0.5 * (1.0 - math.cos(2.0 * tv) + scl * mod2 * math.cos(2.0 * tv) +
                                            scl * mod4 * math.cos(4.0 * tv) + scl * mod6 * math.cos(6.0 * tv))
    return out


def runoff_routing(k1: np.ndarray, maxi: int, maxj: int, maxk: int) -> tuple[np.ndarray, np.ndarray]:
    """readroff (embm.f90:3779-3836): from every cell follow the dry cells'
    compass codes (91 east, 92 south, 93 west, 94 north; periodic in i) to
    the first wet cell.  k1 is (maxj+2, maxi+2) with the reference halo.
    Returns iroff, jroff (maxj, maxi), 1-based; raises ValueError where a
    path leaves the grid, meets a dry cell without a code or does not end."""
    iroff = np.zeros((maxj, maxi), dtype=np.int32)
    jroff = np.zeros((maxj, maxi), dtype=np.int32)
    for j in range(1, maxj + 1):
        for i in range(1, maxi + 1):
            ii, jj, loop = i, j, 0
            while k1[jj, ii] > maxk:
                code = k1[jj, ii]
                if code == IROE:
                    ii += 1
                elif code == IROS:
                    jj -= 1
                elif code == IROW:
                    ii -= 1
                elif code == IRON:
                    jj += 1
                else:
                    raise ValueError(f"runoff from ({i},{j}) meets dry cell ({ii},{jj}) with k1 = {code}, no direction")
                if ii == maxi + 1:
                    ii = 1
                elif ii == 0:
                    ii = maxi
                if not 1 <= jj <= maxj:
                    raise ValueError(f"runoff from ({i},{j}) leaves the grid at j = {jj}")
                loop += 1
                if loop > 100000:
                    raise ValueError(f"runoff from ({i},{j}) does not reach the ocean")
            iroff[j - 1, i - 1], jroff[j - 1, i - 1] = ii, jj
    return iroff, jroff


def surflux_setup(g: Grid, k1: np.ndarray, pmeadj: np.ndarray | None = None, solconst: float = 1368.0,
                  daysperyear: float = 365.25, dosc: bool = True, ndta: int = 5,
                  radfor_scl_co2: float = 1.0, radfor_pc_co2_rise: float = 0.0,
                  radfor_scl_ch4: float = 1.0, radfor_pc_ch4_rise: float = 0.0,
                  radfor_scl_n2o: float = 1.0, radfor_pc_n2o_rise: float = 0.0,
                  **params) -> SurfluxSetup:
    """The tables and constants of initialise_embm for a grid g with the
    mask k1 (maxj+2, maxi+2): insolation, albedo, the shortwave absorption ca
    (0.3 over the ocean, 1 over land), the runoff routing, the time-step
    constants rdtdim (embm.f90:742-770) and rsictscsf (1133), the greenhouse
    gases' initial values and compound rates (948-963).  pmeadj (maxj, maxi)
    is the freshwater adjustment field (zero if None); params override the
    namelist scalars (rmax, delf2x, scf, olr_adj, ...)."""
    mi, mj, mk = g.maxi, g.maxj, g.maxk
    wet = k1[1:mj + 1, 1:mi + 1] <= mk
    iroff, jroff = runoff_routing(k1, mi, mj, mk)
    tv = 86400.0 * g.yearlen / (g.nyear * TSC)
    dtatm = tv / ndta
    ryear = 1.0 / (g.yearlen * 86400)
    cpo_ice = 4044.0
    su = SurfluxSetup(
        solfor=radfor(g, solconst, daysperyear, dosc),
        albcl=climatological_albedo(g),
        ca=np.where(wet, 0.3, 1.0),
        pmeadj=np.zeros((mj, mi)) if pmeadj is None else np.ascontiguousarray(pmeadj, dtype=np.float64),
        iroff=iroff, jroff=jroff,
        rdtdim=1.0 / (TSC * tv),
        rsictscsf=DSC * g.dz[mk] * RHO0 * cpo_ice / (17.5 * 86400.0),
        rate_co2=radfor_pc_co2_rise * 0.01 * TSC * dtatm * ndta * ryear,
        rate_ch4=radfor_pc_ch4_rise * 0.01 * TSC * dtatm * ndta * ryear,
        rate_n2o=radfor_pc_n2o_rise * 0.01 * TSC * dtatm * ndta * ryear,
        co2=radfor_scl_co2 * CO20, ch4=radfor_scl_ch4 * CH40, n2o=radfor_scl_n2o * N2O0)
    for k, v in params.items():
        if not hasattr(su, k):
            raise TypeError(f"surflux_setup: unknown parameter {k}")
        setattr(su, k, v)
    return su


def atmosphere(g: Grid, k1: np.ndarray, seed: int, relh: float = 0.8, ice: bool = False):
    """A synthetic atmosphere and sea-ice state for the surface fluxes:
    air temperature (deg C) 27 cos^2(lat) - 12 plus noise, specific humidity
    at relh of saturation (const1 exp(const4 T/(T + const5))) plus noise,
    and, with ice, sea ice over the cold wet cells (height 0.5-2.5 m, cover
    0.2-0.9, surface temperature -5..-1 deg C).  Returns atemp, ashum, sich,
    sica, tice as (maxj, maxi) arrays."""
    rng = np.random.default_rng(seed)
    mi, mj, mk = g.maxi, g.maxj, g.maxk
    lat = np.arcsin(np.asarray(g.s[1:mj + 1]))
    atemp = (27.0 * np.cos(lat) ** 2 - 12.0)[:, None] + rng.normal(0.0, 1.5, (mj, mi))
    qs = 3.80E-3 * np.exp(17.67 * atemp / (atemp + 243.5))
    ashum = relh * qs * (1.0 + rng.normal(0.0, 0.1, (mj, mi)))
    sich = np.zeros((mj, mi))
    sica = np.zeros((mj, mi))
    tice = np.zeros((mj, mi))
    if ice:
        wet = k1[1:mj + 1, 1:mi + 1] <= mk
        cold = wet & (atemp < -2.0)
        sich[cold] = rng.uniform(0.5, 2.5, cold.sum())
        sica[cold] = rng.uniform(0.2, 0.9, cold.sum())
        tice[cold] = rng.uniform(-5.0, -1.0, cold.sum())
    return atemp, ashum, sich, sica, tice
