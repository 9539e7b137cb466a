"""GOLDSTEIN grid, metrics and the seeded synthetic inputs of the tracer step.

The metric formulas restate goldstein.f90:906-1075 (equal-sin(lat) grid,
exponential ``ez0 = 0.1`` vertical grid, EOS coefficients), 1406-1407
(diffusivity scaling) and 2057-2070 (``ssmax``), with the scales of
goldstein_lib.f90:48-53.  They are evaluated with :mod:`math` (glibc libm),
operation by operation in the reference order, so the arrays are the same
doubles the reference computes.

Arrays handed to the engine use the reference's own Fortran layouts, held in
numpy as C-order arrays with the axes reversed:

=========  ==========================  ===========================
name       Fortran (reference)         numpy shape
=========  ==========================  ===========================
go_ts      (maxl, maxi, maxj, maxk)    (maxk, maxj, maxi, maxl)
k1         (0:maxi+1, 0:maxj+1)        (maxj+2, maxi+2) int32
u          (3, 0:maxi, 0:maxj, maxk)   (maxk, maxj+1, maxi+1, 3)
sflux      (maxl, maxi, maxj)          (maxj, maxi, maxl)
=========  ==========================  ===========================
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# goldstein_lib.f90:48-53
USC, RSC, DSC, FSC = 0.05, 6.37e6, 5.0e3, 2 * 7.2921e-5
GSC, RH0SC = 9.81, 1.0e3
RHOSC = RH0SC * FSC * USC * RSC / GSC / DSC
TSC = RSC / USC
SALN0 = 34.9  # goldstein-defaults.nml:60
ZEROC = 273.15  # gem_cmn.f90:690
LAND_K1 = 94  # any k1 > maxk marks a land column


def _powi(x: float, m: int) -> float:
    """REAL**INTEGER as gfortran evaluates it (libgcc __powidf2)."""
    n = abs(m)
    y = x if n % 2 else 1.0
    n >>= 1
    while n:
        x = x * x
        if n % 2:
            y = y * x
        n >>= 1
    return 1.0 / y if m < 0 else y


@dataclass
class Grid:
    """Metrics of one GOLDSTEIN grid; Fortran index == array index."""

    maxi: int
    maxj: int
    maxk: int
    nyear: int = 96
    yearlen: float = 365.25
    dphi: float = 0.0
    rdphi: float = 0.0
    ec: np.ndarray = field(default_factory=lambda: np.zeros(5))
    # per-j arrays, index 0..maxj+1
    s: np.ndarray = None
    sv: np.ndarray = None
    c: np.ndarray = None
    cv: np.ndarray = None
    ds: np.ndarray = None
    rds: np.ndarray = None
    rc: np.ndarray = None
    rc2: np.ndarray = None
    dsv: np.ndarray = None
    rdsv: np.ndarray = None
    rcv: np.ndarray = None
    cv2: np.ndarray = None
    rds2: np.ndarray = None
    # per-k arrays, index 0..maxk+1
    dz: np.ndarray = None
    dza: np.ndarray = None
    rdz: np.ndarray = None
    rdza: np.ndarray = None
    zw: np.ndarray = None
    zro: np.ndarray = None
    dt: np.ndarray = None
    ssmax: np.ndarray = None

    @property
    def dtyr(self) -> float:
        return 1.0 / self.nyear


def make_grid(maxi=36, maxj=36, maxk=16, nyear=96, yearlen=365.25,
              ssmaxsurf=10.0, ssmaxdeep=10.0) -> Grid:
    """goldstein.f90:906-1075 and 2057-2070 (igrid = 0, ez0 = 0.1, ieos = 0)."""
    J, K = maxj + 2, maxk + 2
    z = lambda n: np.zeros(n)  # noqa: E731
    g = Grid(maxi, maxj, maxk, nyear, yearlen)
    for name in ("s", "sv", "c", "cv", "ds", "rds", "rc", "rc2", "dsv", "rdsv", "rcv", "cv2", "rds2"):
        setattr(g, name, z(J))
    for name in ("dz", "dza", "rdz", "rdza", "zw", "zro", "dt", "ssmax"):
        setattr(g, name, z(K))
    pi = 4 * math.atan(1.0)
    th0, th1 = -pi / 2, pi / 2
    s0, s1 = math.sin(th0), math.sin(th1)
    g.dphi = 2 * pi / maxi
    g.rdphi = 1.0 / g.dphi
    g.sv[0] = s0
    g.cv[0] = math.cos(th0)
    dscon = (s1 - s0) / maxj
    for j in range(1, maxj + 1):
        g.sv[j] = s0 + j * dscon
        g.cv[j] = math.sqrt(1 - g.sv[j] * g.sv[j])
        g.s[j] = g.sv[j] - 0.5 * dscon
    for j in range(1, maxj + 1):
        g.ds[j] = g.sv[j] - g.sv[j - 1]
        g.rds[j] = 1.0 / g.ds[j]
        g.c[j] = math.sqrt(1 - g.s[j] * g.s[j])
        g.rc[j] = 1.0 / g.c[j]
        g.rc2[j] = g.rc[j] * g.rc[j] * g.rdphi
        if j < maxj:
            g.dsv[j] = g.s[j + 1] - g.s[j]
            g.rdsv[j] = 1.0 / g.dsv[j]
            g.rcv[j] = 1.0 / g.cv[j]
            g.cv2[j] = g.cv[j] * g.cv[j] * g.rdsv[j]
            if j > 1:
                g.rds2[j] = 2.0 / (g.dsv[j] + g.dsv[j - 1])
    tv = 86400.0 * yearlen / (nyear * TSC)
    g.dt[:] = tv
    ez0 = 0.1
    z1 = ez0 * ((1.0 + 1 / ez0) ** (1.0 / maxk) - 1.0)
    tv4 = ez0 * ((z1 / ez0 + 1) ** 0.5 - 1)
    tv2 = 0.0
    g.zro[maxk] = -tv4
    g.zw[maxk] = tv2
    for k in range(1, maxk + 1):
        tv3 = ez0 * (_powi(z1 / ez0 + 1, k) - 1)
        g.dz[maxk - k + 1] = tv3 - tv2
        tv2 = tv3
        tv5 = ez0 * ((z1 / ez0 + 1) ** (k + 0.5) - 1)
        if k < maxk:
            g.dza[maxk - k] = tv5 - tv4
        tv4 = tv5
    for k in range(maxk, 0, -1):
        if k > 1:
            g.zro[k - 1] = g.zro[k] - g.dza[k - 1]
        g.zw[k - 1] = g.zw[k] - g.dz[k]
    for k in range(1, maxk):
        g.rdz[k] = 1.0 / g.dz[k]
        g.rdza[k] = 1.0 / g.dza[k]
    g.rdz[maxk] = 1.0 / g.dz[maxk]
    g.dza[maxk] = 0.0
    g.ec = np.array([-0.0559 / RHOSC, 0.7968 / RHOSC, -0.0063 / RHOSC, 3.7315e-5 / RHOSC, 0.0])
    if -1.0e-7 < ssmaxsurf - ssmaxdeep < 1.0e-7:
        g.ssmax[1:maxk] = ssmaxdeep
    else:
        mid = 0.5 * (math.log(ssmaxsurf) + math.log(ssmaxdeep))
        dif = 0.5 * (math.log(ssmaxsurf) - math.log(ssmaxdeep))
        efold, dep0 = 200 / DSC, -300 / DSC
        for k in range(1, maxk):
            g.ssmax[k] = math.exp(mid + dif * math.tanh((g.zw[k] - dep0) / efold))
    return g


@dataclass
class World:
    """A GOLDSTEIN world's files: the bathymetry k1(0:maxi+1, 0:maxj+1) as
    (maxj+2, maxi+2), the landmass label of every psi point (maxi, 0:maxj) as
    (maxj+1, maxi), the island count and the island paths."""
    k1: np.ndarray
    labels: np.ndarray
    isles: int
    npi: np.ndarray
    paths: np.ndarray


def read_world(prefix: str, maxi: int, maxj: int) -> World:
    """Read <prefix>.k1, .psiles and .paths as initialise_goldstein does
    (goldstein.f90:1113-1120, 1500-1574): k1 rows from j = maxj+1 down to 0
    with the periodic columns copied from the interior, .psiles rows from
    j = maxj down to 0, isles = the largest label - 1, then the path lengths
    and the (lpisl, ipisl, jpisl) of each path point, island by island."""
    def numbers(path, dtype):
        with open(path) as f:
            return np.array(f.read().split(), dtype=dtype)

    v = numbers(prefix + ".k1", np.int64)
    if v.size < (maxj + 2) * (maxi + 2):
        raise ValueError(f"{prefix}.k1: {v.size} values, {(maxj + 2) * (maxi + 2)} needed")
    k1 = np.zeros((maxj + 2, maxi + 2), dtype=np.int32)
    for r in range(maxj + 2):
        k1[maxj + 1 - r] = v[r * (maxi + 2):(r + 1) * (maxi + 2)]
    k1[:, 0] = k1[:, maxi]
    k1[:, maxi + 1] = k1[:, 1]
    lab = numbers(prefix + ".psiles", np.float64)
    if lab.size < (maxj + 1) * maxi:
        raise ValueError(f"{prefix}.psiles: {lab.size} values, {(maxj + 1) * maxi} needed")
    labels = np.zeros((maxj + 1, maxi), dtype=np.int32)
    for r in range(maxj + 1):
        labels[maxj - r] = lab[r * maxi:(r + 1) * maxi].astype(np.int32)
    isles = max(int(labels.max()) - 1, 0)
    npi = np.zeros(0, dtype=np.int32)
    paths = np.zeros(0, dtype=np.int32)
    if isles > 0:
        tok = numbers(prefix + ".paths", np.int64)
        npi = tok[:isles].astype(np.int32)
        paths = tok[isles:isles + 3 * int(npi.sum())].astype(np.int32)
        if paths.size != 3 * int(npi.sum()):
            raise ValueError(f"{prefix}.paths: {paths.size // 3} path points, {int(npi.sum())} needed")
    return World(k1, labels, isles, npi, paths)


def diffusivities(kappa_h: float, kappa_v: float) -> tuple[float, float]:
    """Non-dimensional diff(1), diff(2) (goldstein.f90:1406-1407)."""
    return kappa_h / (RSC * USC), kappa_v * RSC / (USC * DSC * DSC)


# --------------------------------------------------------------------------
# seeded synthetic inputs
# --------------------------------------------------------------------------
def _smooth(a: np.ndarray, sigma: float) -> np.ndarray:
    """Gaussian smoothing, periodic in i (last axis), reflecting in j."""
    from scipy.ndimage import gaussian_filter

    return gaussian_filter(a, sigma=sigma, mode=("reflect", "wrap"))


def make_k1(g: Grid, seed: int, land_frac: float = 0.30) -> np.ndarray:
    """Land mask and column depth k1 (0:maxi+1, 0:maxj+1) as numpy (maxj+2, maxi+2).

    Land: smoothed N(0,1) noise (sigma = 3 cells) below its land_frac quantile.
    Ocean columns: k1 ~ U{1..maxk}, smoothed (sigma = 3) and rounded into
    [1, maxk].  Rows j = 0 and maxj+1 are land; i = 0 and maxi+1 are the
    periodic copies (goldstein.f90:1114-1123)."""
    rng = np.random.default_rng(seed)
    mi, mj, mk = g.maxi, g.maxj, g.maxk
    noise = _smooth(rng.standard_normal((mj, mi)), 3.0)
    land = noise < np.quantile(noise, land_frac)
    depth = _smooth(rng.integers(1, mk + 1, size=(mj, mi)).astype(np.float64), 3.0)
    kk = np.clip(np.rint(depth), 1, mk).astype(np.int32)
    kk[land] = LAND_K1
    k1 = np.full((mj + 2, mi + 2), LAND_K1, dtype=np.int32)
    k1[1:mj + 1, 1:mi + 1] = kk
    k1[:, 0] = k1[:, mi]
    k1[:, mi + 1] = k1[:, 1]
    return k1


def wet_mask(g: Grid, k1: np.ndarray) -> np.ndarray:
    """Boolean (maxk, maxj, maxi): True where k >= k1(i,j)."""
    k = np.arange(1, g.maxk + 1)[:, None, None]
    return k >= k1[None, 1:g.maxj + 1, 1:g.maxi + 1]


def make_velocity(g: Grid, k1: np.ndarray, seed: int, umax_ms: float = 0.1,
                  scale: float = 1.0) -> np.ndarray:
    """Divergence-free C-grid velocity u (3, 0:maxi, 0:maxj, maxk).

    A smooth random streamfunction psi(i, j, k) at psi points is zeroed where
    any adjacent tracer cell is land at level k; u and v follow ubarsolv's
    discrete form (goldstein.f90:3536-3548, without rh), w follows velc's
    continuity loop (3668-3678).  The field is scaled so max|u,v| equals
    umax_ms (m/s) times ``scale``."""
    rng = np.random.default_rng(seed)
    mi, mj, mk = g.maxi, g.maxj, g.maxk
    wet = wet_mask(g, k1)
    psi = np.zeros((mk, mj + 1, mi + 1))
    base = _smooth(rng.standard_normal((mj + 1, mi)), 3.0)
    for k in range(mk):
        lev = base + 0.5 * _smooth(rng.standard_normal((mj + 1, mi)), 3.0)
        p = np.zeros((mj + 1, mi + 1))
        p[:, 1:] = lev
        p[:, 0] = p[:, mi]
        # psi(i,j) touches tracer cells (i,j),(i+1,j),(i,j+1),(i+1,j+1)
        wk = np.zeros((mj + 2, mi + 2), dtype=bool)
        wk[1:mj + 1, 1:mi + 1] = wet[k]
        wk[:, 0] = wk[:, mi]
        wk[:, mi + 1] = wk[:, 1]
        ok = wk[0:mj + 1, 0:mi + 1] & wk[0:mj + 1, 1:mi + 2] & wk[1:mj + 2, 0:mi + 1] & wk[1:mj + 2, 1:mi + 2]
        p[~ok] = 0.0
        psi[k] = p

    def uv(ps):
        u = np.zeros((mk, mj + 1, mi + 1, 3))
        for j in range(1, mj + 1):
            u[:, j, 1:, 0] = -g.c[j] * (ps[:, j, 1:] - ps[:, j - 1, 1:]) * g.rds[j]
        for j in range(1, mj):
            u[:, j, 1:, 1] = (ps[:, j, 1:] - ps[:, j, :-1]) * g.rcv[j] * g.rdphi
        u[:, :, 0, 0] = u[:, :, mi, 0]
        return u

    u = uv(psi)
    m = max(np.abs(u[..., 0]).max(), np.abs(u[..., 1]).max())
    target = umax_ms * scale / USC
    psi *= target / m if m > 0 else 0.0
    u = uv(psi)
    # w from continuity, velc 3668-3678
    for j in range(1, mj + 1):
        for i in range(1, mi + 1):
            tv = 0.0
            for k in range(int(k1[j, i]), mk):
                tv1 = (u[k - 1, j, i, 0] - u[k - 1, j, i - 1, 0]) * g.rdphi * g.rc[j]
                tv2 = (u[k - 1, j, i, 1] * g.cv[j] - u[k - 1, j - 1, i, 1] * g.cv[j - 1]) * g.rds[j]
                u[k - 1, j, i, 2] = tv - g.dz[k] * (tv1 + tv2)
                tv = u[k - 1, j, i, 2]
    return u


def make_random_velocity(g: Grid, seed: int, amp: float = 2.0) -> np.ndarray:
    """A non-solenoidal random field with a vertical component, for the
    stress tests of flux-form conservation (any u conserves the tracer
    inventory in the flux form)."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(-amp, amp, size=(g.maxk, g.maxj + 1, g.maxi + 1, 3))
    u[..., 2] *= 1e-3
    u[:, :, 0, 0] = u[:, :, g.maxi, 0]
    return u


@dataclass
class TracerSet:
    """Compact tracer slots.  Slot 0 = T, slot 1 = S (as in GOLDSTEIN ts)."""

    names: list
    DIC: int = -1
    ALK: int = -1
    PO4: int = -1
    O2: int = -1
    DOM_C: int = -1
    DOM_P: int = -1

    @property
    def L(self) -> int:
        return len(self.names)


def tracer_set(kind: str = "config2", L: int | None = None) -> TracerSet:
    """'ts'     : T, S (+ L-2 dyes if L given)
    'config2': T, S, DIC, ALK, PO4, O2, DOM_P and 17 dyes (24 tracers)
    'domc'   : T, S, DIC, ALK, PO4, O2, DOM_C, DOM_P and dyes to L (default 24)"""
    if kind == "ts":
        n = 2 if L is None else L
        return TracerSet(["T", "S"] + [f"dye{q}" for q in range(n - 2)])
    if kind == "config2":
        names = ["T", "S", "DIC", "ALK", "PO4", "O2", "DOM_P"]
        n = 24 if L is None else L
        names += [f"dye{q}" for q in range(n - len(names))]
        return TracerSet(names, DIC=2, ALK=3, PO4=4, O2=5, DOM_P=6)
    if kind == "domc":
        names = ["T", "S", "DIC", "ALK", "PO4", "O2", "DOM_C", "DOM_P"]
        n = 24 if L is None else L
        names += [f"dye{q}" for q in range(n - len(names))]
        return TracerSet(names, DIC=2, ALK=3, PO4=4, O2=5, DOM_C=6, DOM_P=7)
    raise ValueError(kind)


def make_tracers(g: Grid, k1: np.ndarray, ts: TracerSet, seed: int) -> np.ndarray:
    """Initial go_ts (maxk, maxj, maxi, L); land cells are zero.

    T = 2 + 18 exp(-z/1000 m) + N(0, 0.1) degC; S stored as S - 34.9 with
    S = 34.9 + N(0, 0.1); DIC 2000+-100, ALK 2300+-50, PO4 U[0,3],
    O2 U[0,300] umol/kg; DOM U[0, 0.1] umol/kg (DOM_C 106x); dyes U[0,1]."""
    rng = np.random.default_rng(seed)
    mi, mj, mk = g.maxi, g.maxj, g.maxk
    shp = (mk, mj, mi)
    depth = -np.array([g.zro[k] for k in range(1, mk + 1)]) * DSC
    out = np.zeros(shp + (ts.L,))
    out[..., 0] = 2.0 + 18.0 * np.exp(-depth / 1000.0)[:, None, None] + rng.normal(0, 0.1, shp)
    out[..., 1] = rng.normal(0, 0.1, shp)
    um = 1e-6
    for q, name in enumerate(ts.names[2:], start=2):
        if name == "DIC":
            out[..., q] = rng.uniform(1900, 2100, shp) * um
        elif name == "ALK":
            out[..., q] = rng.uniform(2250, 2350, shp) * um
        elif name == "PO4":
            out[..., q] = rng.uniform(0, 3, shp) * um
        elif name == "O2":
            out[..., q] = rng.uniform(0, 300, shp) * um
        elif name == "DOM_P":
            out[..., q] = rng.uniform(0, 0.1, shp) * um
        elif name == "DOM_C":
            out[..., q] = rng.uniform(0, 10.6, shp) * um
        else:
            out[..., q] = rng.uniform(0, 1, shp)
    out[~wet_mask(g, k1)] = 0.0
    return out


def make_light(g: Grid, seed: int) -> np.ndarray:
    """Seeded light field solfor/solar_constant in [0, 1], (maxj, maxi)."""
    rng = np.random.default_rng(seed)
    lat = np.array([math.asin(g.s[j]) for j in range(1, g.maxj + 1)])
    base = np.cos(lat)[:, None] * np.ones((g.maxj, g.maxi))
    return np.clip(base * (0.75 + 0.25 * _smooth(rng.standard_normal((g.maxj, g.maxi)), 2.0)), 0.0, 1.0)


def make_mld(g: Grid, seed: int, max_m: float = 0.0) -> np.ndarray:
    """Mixed-layer depth (m) per column; zero (surface layer only) by default."""
    if max_m <= 0:
        return np.zeros((g.maxj, g.maxi))
    rng = np.random.default_rng(seed)
    return rng.uniform(0, max_m, (g.maxj, g.maxi))
