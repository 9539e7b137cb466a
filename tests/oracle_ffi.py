"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

Test infrastructure: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py only.  It is the checker, never the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "liboracle.so")
REF_LIB = os.path.join(ORACLE_DIR, "_ref", "libref_carbchem.so")
REF_AIRSEA_LIB = os.path.join(ORACLE_DIR, "_ref", "libref_airsea.so")
REF_GOLD_LIB = os.path.join(ORACLE_DIR, "_ref", "libref_gold.so")

OG_MAXJ, OG_MAXK = 256, 64
NSED, NCARB, NCC = 5, 10, 17

_d = C.c_double
_i = C.c_int
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class OgGrid(C.Structure):
    J, K = OG_MAXJ + 2, OG_MAXK + 2
    _fields_ = [("maxi", _i), ("maxj", _i), ("maxk", _i), ("dphi", _d), ("rdphi", _d)] + \
        [(n, _d * (OG_MAXJ + 2)) for n in ("s", "sv", "c", "cv", "ds", "rds", "rc", "rc2", "dsv", "rdsv",
                                           "rcv", "cv2", "rds2")] + \
        [(n, _d * (OG_MAXK + 2)) for n in ("dz", "dza", "rdz", "rdza", "zw", "zro", "dt", "ssmax")] + \
        [("ec", _d * 5), ("diff", _d * 2), ("tsc", _d), ("dtyr", _d)]


class OgBgcParams(C.Structure):
    _fields_ = [("l_DIC", _i), ("l_ALK", _i), ("l_PO4", _i), ("l_O2", _i), ("l_DOMC", _i), ("l_DOMP", _i),
                ("caco3", _i), ("dtyr", _d), ("k0_PO4", _d), ("c0_PO4", _d), ("DOMfrac", _d),
                ("DOMlifetime", _d), ("red_POP_POC", _d), ("red_POP_PO2", _d), ("red_PON_ALK", _d),
                ("red_POP_PON", _d), ("red_POC_CaCO3", _d), ("red_POC_CaCO3_pP", _d), ("POC_frac2", _d),
                ("POC_dfrac2", _d), ("POC_c0frac2", _d), ("POC_eL1", _d), ("POC_eL2", _d),
                ("CaCO3_frac2", _d), ("CaCO3_eL1", _d), ("CaCO3_eL2", _d), ("sinkrate", _d),
                ("remin_k_O2", _d), ("remin_c0_O2", _d), ("carb_tol", _d), ("carb_maxit", _i), ("dsc", _d),
                ("airsea", _i), ("gastransfer_a", _d), ("atm_pCO2", _d), ("atm_pO2", _d)]


class OgBgcState(C.Structure):
    _fields_ = [("maxi", _i), ("maxj", _i), ("maxk", _i), ("maxl", _i)] + \
        [(n, C.POINTER(_d)) for n in ("ocn", "part", "settle", "carb", "carbconst", "M", "rM", "V")] + \
        [(n, _d * (OG_MAXK + 2)) for n in ("Dbot", "Dmid", "dD")] + \
        [("light", C.POINTER(_d)), ("mld", C.POINTER(_d)), ("carb_fail", C.c_int32), ("carb_iter", C.c_int32)] + \
        [(n, C.POINTER(_d)) for n in ("wspeed", "atm", "fatm", "sumatm")]


class OgMom(C.Structure):
    _fields_ = [("maxi", _i), ("maxj", _i), ("maxk", _i), ("isles", _i), ("mpi", _i), ("rel", _d)] + \
        [(n, C.POINTER(_d)) for n in ("h", "rh")] + [(n, C.POINTER(C.c_int32)) for n in ("ku", "mk")] + \
        [("getj", C.POINTER(C.c_uint8))] + \
        [(n, C.POINTER(_d)) for n in ("drag", "rtv", "rtv3", "gap", "ratm", "gb", "gbold", "bp", "sbp", "ub", "psi",
                                      "ubisl", "psisl", "erisl", "psibc")] + \
        [(n, C.POINTER(C.c_int32)) for n in ("npi", "lpisl", "ipisl", "jpisl")] + \
        [(n, C.POINTER(_d)) for n in ("tau", "dztau", "dztav", "u1")]


class OgSfxPar(C.Structure):
    _fields_ = [(n, _i) for n in ("maxi", "maxj", "maxk", "nyear", "igrid", "polar_avg", "radfor_atchem")] + \
        [(n, _d) for n in ("scf", "rmax", "delf2x", "rdtdim", "rsictscsf", "olr_adj", "olr_adj0", "t_eqm",
                           "albsic_min", "albsic_max", "sich_max", "rate_co2", "rate_ch4", "rate_n2o")] + \
        [(n, C.POINTER(_d)) for n in ("solfor", "albcl", "ca", "pmeadj")] + \
        [(n, C.POINTER(C.c_int32)) for n in ("iroff", "jroff")]


class OgSfxState(C.Structure):
    _fields_ = [(n, C.POINTER(_d)) for n in ("atemp", "ashum", "sich", "sica", "tice")] + \
        [(n, _d) for n in ("co2", "ch4", "n2o")] + [("istep", C.c_int64)]


# surflux outputs, in the oracle's order (cgo_oracle.h)
SF_NAMES = ("fxlho", "fxsho", "fxswo", "fxlwo", "evap_ocn", "pptn_ocn", "runoff_ocn", "runoff_land", "fxlha",
            "fxsha", "fxswa", "fxlwa", "evap_atm", "pptn_atm", "dthsic", "dtareasic", "albedo", "albice", "fxsw",
            "fx0a", "fx0o", "fxsen", "fxlw", "evap", "pptn", "fx0sic", "fx0neto", "tsfreez")
SF = {n: q for q, n in enumerate(SF_NAMES)}


def build(force: bool = False) -> None:
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-C", ORACLE_DIR, "liboracle.so"], check=True, capture_output=True)


_lib = None


def _bind(L: C.CDLL) -> C.CDLL:
    L.og_grid_init.argtypes = [C.POINTER(OgGrid), _i, _i, _i, _i, _d, _d, _d, _d, _d]
    L.og_goldstein_step.argtypes = [C.POINTER(OgGrid), _i, _i32p, _dp, _dp, C.c_void_p, _dp, _dp,
                                    C.c_void_p, _i, C.POINTER(C.c_int32), C.POINTER(_d)]
    L.og_inventory.argtypes = [C.POINTER(OgGrid), _i, _i32p, _dp, _dp]
    L.og_calc_carbconst.argtypes = [_d, _d, _d, _dp]
    L.og_calc_carb.argtypes = [_d] * 10 + [_dp, _dp, C.c_void_p, _d, _i]
    L.og_calc_carb.restype = _i
    L.og_calc_carb_RF0.argtypes = [_d] * 9 + [_dp, _dp, _d, _i]
    L.og_calc_carb_RF0.restype = _i
    for f in ("og_calc_Ca", "og_calc_Btot", "og_calc_SO4tot", "og_calc_Ftot"):
        getattr(L, f).argtypes = [_d]
        getattr(L, f).restype = _d
    L.og_eos.argtypes = [_dp, _d, _d]
    L.og_eos.restype = _d
    L.og_bgc_default_params.argtypes = [C.POINTER(OgBgcParams), _d]
    L.og_bgc_init.argtypes = [C.POINTER(OgBgcState), C.POINTER(OgGrid), C.POINTER(OgBgcParams), _i, _i32p,
                              _dp, _dp, _dp]
    L.og_bgc_free.argtypes = [C.POINTER(OgBgcState)]
    L.og_biogem_step.argtypes = [C.POINTER(OgBgcState), C.POINTER(OgGrid), C.POINTER(OgBgcParams), _i32p, _dp]
    L.og_tracercoupling.argtypes = [C.POINTER(OgBgcState), C.POINTER(OgGrid), _i32p, _dp, _dp]
    L.og_calc_rho.argtypes = [_d, _d]
    L.og_calc_rho.restype = _d
    L.og_airsea_cell.argtypes = [_d] * 14 + [_dp]
    L.og_bgc_set_windspeed.argtypes = [C.POINTER(OgBgcState), _dp]
    L.og_atchem_step.argtypes = [C.POINTER(OgBgcState), C.POINTER(OgGrid), C.POINTER(OgBgcParams)]
    L.og_mom_init.argtypes = [C.POINTER(OgMom), C.POINTER(OgGrid), _i32p, _i32p, _i, _i32p, _i32p, _d, _d]
    L.og_mom_init.restype = _i
    L.og_mom_free.argtypes = [C.POINTER(OgMom)]
    L.og_mom_set_wind.argtypes = [C.POINTER(OgMom), C.POINTER(OgGrid), _dp, _d]
    L.og_mom_step.argtypes = [C.POINTER(OgMom), C.POINTER(OgGrid), _i32p, _dp, _dp]
    L.og_mom_island.argtypes = [C.POINTER(OgMom), C.POINTER(OgGrid), _i]
    L.og_mom_island.restype = _d
    L.og_mom_island_of.argtypes = [C.POINTER(OgMom), C.POINTER(OgGrid), _dp, _i, _i]
    L.og_mom_island_of.restype = _d
    L.og_rho_from_ts.argtypes = [C.POINTER(OgGrid), _i, _i32p, _dp, _dp]
    L.og_sfx_usurf.argtypes = [C.POINTER(OgSfxPar), C.POINTER(OgGrid), _dp, _dp, _dp, _dp, _dp]
    L.og_surflux.argtypes = [C.POINTER(OgSfxPar), C.POINTER(OgGrid), _i32p, C.POINTER(OgSfxState), _dp, _dp, _dp,
                             C.c_void_p, _dp, C.c_void_p]
    L.og_surface_slot.argtypes = [C.POINTER(OgSfxPar), _dp, C.c_void_p, C.c_void_p, _i, _dp]
    return L


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = _bind(C.CDLL(LIB))
    return _lib


class use_library:
    """Context manager: route every oracle call to another build of
    cgo_oracle.c (the mutation tests load copies with a planted mistake)."""

    def __init__(self, path: str):
        self.path = path

    def __enter__(self):
        global _lib
        self.saved = lib()
        _lib = _bind(C.CDLL(self.path))
        return _lib

    def __exit__(self, *exc):
        global _lib
        _lib = self.saved
        return False


def grid(maxi=36, maxj=36, maxk=16, nyear=96, yearlen=365.25, kappa_h=2000.0, kappa_v=2.7e-5,
         ssmaxsurf=10.0, ssmaxdeep=10.0) -> OgGrid:
    g = OgGrid()
    lib().og_grid_init(C.byref(g), maxi, maxj, maxk, nyear, yearlen, kappa_h, kappa_v, ssmaxsurf, ssmaxdeep)
    return g


def bgc_params(dtyr: float, ts=None, carb_tol=1e-12, carb_maxit=1000, **kw) -> OgBgcParams:
    p = OgBgcParams()
    lib().og_bgc_default_params(C.byref(p), dtyr)
    if ts is not None:
        p.l_DIC, p.l_ALK, p.l_PO4, p.l_O2 = ts.DIC, ts.ALK, ts.PO4, ts.O2
        p.l_DOMC, p.l_DOMP = ts.DOM_C, ts.DOM_P
        p.caco3 = 1 if (ts.DIC >= 0 and ts.ALK >= 0) else 0
    p.carb_tol, p.carb_maxit = carb_tol, carb_maxit
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class GoldsteinOracle:
    """Holds the arrays of one member and steps it with the C restatement."""

    def __init__(self, g: OgGrid, k1: np.ndarray, go_ts: np.ndarray, u: np.ndarray,
                 sflux: np.ndarray | None = None, diso: bool = True):
        self.g, self.k1 = g, np.ascontiguousarray(k1, dtype=np.int32)
        self.ts = np.ascontiguousarray(go_ts, dtype=np.float64).copy()
        self.u = np.ascontiguousarray(u, dtype=np.float64)
        self.L = self.ts.shape[-1]
        self.sflux = None if sflux is None else np.ascontiguousarray(sflux, dtype=np.float64)
        self.rho = np.zeros((g.maxk + 1, g.maxj + 2, g.maxi + 2))
        self.cost = np.zeros((g.maxj, g.maxi))
        self.diso = 1 if diso else 0
        self.limps = C.c_int32(0)
        self.dmax = _d(0.0)

    def step(self, n: int = 1):
        L = lib()
        sf = None if self.sflux is None else self.sflux.ctypes.data_as(C.c_void_p)
        for _ in range(n):
            L.og_goldstein_step(C.byref(self.g), self.L, self.k1, self.ts, self.u, sf, self.rho, self.cost,
                                None, self.diso, C.byref(self.limps), C.byref(self.dmax))

    def inventory(self) -> np.ndarray:
        inv = np.zeros(self.L)
        lib().og_inventory(C.byref(self.g), self.L, self.k1, self.ts, inv)
        return inv


class BiogemOracle:
    """BIOGEM state of one member plus the GOLDSTEIN oracle it couples to."""

    def __init__(self, gold: GoldsteinOracle, p: OgBgcParams, light: np.ndarray, mld: np.ndarray):
        self.gold, self.p = gold, p
        self.st = OgBgcState()
        self.light = np.ascontiguousarray(light, dtype=np.float64)
        self.mld = np.ascontiguousarray(mld, dtype=np.float64)
        lib().og_bgc_init(C.byref(self.st), C.byref(gold.g), C.byref(p), gold.L, gold.k1, gold.ts, self.light,
                          self.mld)
        g = gold.g
        self.ncell = g.maxi * g.maxj * g.maxk
        self.vdocn = np.zeros_like(gold.ts)

    def __del__(self):
        try:
            lib().og_bgc_free(C.byref(self.st))
        except Exception:
            pass

    def arr(self, name: str, per_cell: int = 1, cells: int | None = None) -> np.ndarray:
        n = (self.ncell if cells is None else cells) * per_cell
        return np.ctypeslib.as_array(getattr(self.st, name), shape=(n,))

    def set_windspeed(self, wspeed: np.ndarray):
        lib().og_bgc_set_windspeed(C.byref(self.st), np.ascontiguousarray(wspeed, dtype=np.float64))

    def step(self, n: int = 1, goldstein: bool = True):
        """genie.f90 order: step_goldstein, step_biogem, tracercoupling, then
        with air-sea exchange the ocean -> atmosphere coupling and ATCHEM."""
        L = lib()
        for _ in range(n):
            if goldstein:
                self.gold.step(1)
            L.og_biogem_step(C.byref(self.st), C.byref(self.gold.g), C.byref(self.p), self.gold.k1, self.vdocn)
            L.og_tracercoupling(C.byref(self.st), C.byref(self.gold.g), self.gold.k1, self.gold.ts, self.vdocn)
            if self.p.airsea:
                L.og_atchem_step(C.byref(self.st), C.byref(self.gold.g), C.byref(self.p))

    def atm(self) -> np.ndarray:
        """(maxj, maxi, 2) atmospheric pCO2 and pO2 (atm)."""
        g = self.gold.g
        return self.arr("atm", 2, g.maxi * g.maxj).reshape(g.maxj, g.maxi, 2).copy()

    def fatm(self) -> np.ndarray:
        """(maxj, maxi, 2) the last step's ocean -> atmosphere CO2 and O2 flux (mol yr-1)."""
        g = self.gold.g
        return self.arr("fatm", 2, g.maxi * g.maxj).reshape(g.maxj, g.maxi, 2).copy()

    def ocn(self) -> np.ndarray:
        g = self.gold.g
        return self.arr("ocn", self.gold.L).reshape(g.maxk, g.maxj, g.maxi, self.gold.L).copy()

    def part(self) -> np.ndarray:
        g = self.gold.g
        return self.arr("part", NSED).reshape(g.maxk, g.maxj, g.maxi, NSED).copy()

    def carb(self) -> np.ndarray:
        g = self.gold.g
        return self.arr("carb", NCARB, g.maxi * g.maxj).reshape(g.maxj, g.maxi, NCARB).copy()


_ref = None


def ref_carbchem():
    """The reference's gem_carbchem compiled in place (oracle/_ref), or None."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        R = C.CDLL(REF_LIB)
        R.ref_set_params.argtypes = [_d, _i]
        R.ref_calc_carbconst.argtypes = [_d, _d, _d, _dp]
        R.ref_salt_totals.argtypes = [_d, _dp]
        R.ref_calc_carb.argtypes = [_d] * 10 + [_dp, _dp, _dp]
        R.ref_calc_carb.restype = _i
        R.ref_calc_carb_rf0.argtypes = [_d] * 9 + [_dp, _dp]
        _ref = R
    return _ref


_ref_as = None


def ref_airsea():
    """The reference's BIOGEM air-sea routines compiled in place (oracle/_ref), or None."""
    global _ref_as
    if _ref_as is None and os.path.exists(REF_AIRSEA_LIB):
        R = C.CDLL(REF_AIRSEA_LIB)
        R.ref_calc_rho.argtypes = [_d, _d]
        R.ref_calc_rho.restype = _d
        R.ref_airsea_cell.argtypes = [_d] * 14 + [_dp]
        _ref_as = R
    return _ref_as


GOLDEN = os.path.join(ROOT, "tests", "golden")


def read_world(name: str = "worbe2", maxi: int = 36, maxj: int = 36):
    """The reference's island data of a world as goldstein.f90 reads it
    (1113-1120, 1500-1575): k1(0:maxi+1, 0:maxj+1) as (maxj+2, maxi+2) with
    the periodic columns enforced, the .psiles landmass labels as
    (maxj+1, maxi) (row j = 0 first), isles, and the .paths path lengths and
    (lpi, ipi, jpi) triples."""
    v = np.array(open(os.path.join(GOLDEN, name + ".k1")).read().split(), dtype=np.int32)
    k1 = v[:(maxj + 2) * (maxi + 2)].reshape(maxj + 2, maxi + 2)[::-1].copy()
    k1[:, 0] = k1[:, maxi]
    k1[:, maxi + 1] = k1[:, 1]
    lab = np.array(open(os.path.join(GOLDEN, name + ".psiles")).read().split(), dtype=float)
    labels = lab[:(maxj + 1) * maxi].reshape(maxj + 1, maxi)[::-1].astype(np.int32).copy()
    isles = int(labels.max()) - 1
    tok = open(os.path.join(GOLDEN, name + ".paths")).read().split()
    npi = np.array(tok[:isles], dtype=np.int32)
    paths = np.array(tok[isles:isles + 3 * int(npi.sum())], dtype=np.int32)
    return k1, labels, isles, npi, paths


class MomOracle:
    """The GOLDSTEIN velocity solve of one member (og_mom_*)."""

    def __init__(self, g: OgGrid, k1, labels, isles, npi, paths, adrag=2.5, rel=0.9):
        self.g = g
        self.k1 = np.ascontiguousarray(k1, dtype=np.int32)
        self.mo = OgMom()
        rc = lib().og_mom_init(C.byref(self.mo), C.byref(g), self.k1, np.ascontiguousarray(labels, dtype=np.int32),
                               isles, np.ascontiguousarray(npi, dtype=np.int32),
                               np.ascontiguousarray(paths, dtype=np.int32), adrag, rel)
        if rc != 0:
            raise ValueError("og_mom_init: bad island data")

    def __del__(self):
        try:
            lib().og_mom_free(C.byref(self.mo))
        except Exception:
            pass

    def set_wind(self, stress: np.ndarray, scf: float = 2.0):
        """stress (maxj, maxi, 4): x and y at u points, x and y at v points (N m-2)."""
        lib().og_mom_set_wind(C.byref(self.mo), C.byref(self.g), np.ascontiguousarray(stress, dtype=np.float64), scf)

    def step(self, rho: np.ndarray, u: np.ndarray):
        lib().og_mom_step(C.byref(self.mo), C.byref(self.g), self.k1, rho, u)

    def island(self, isl: int) -> float:
        return lib().og_mom_island(C.byref(self.mo), C.byref(self.g), isl)

    def island_of(self, ub: np.ndarray, isl: int, indj: int) -> float:
        """The island path integral of a given ub ((maxj+1, maxi+2, 2), as ub() gives it)."""
        return lib().og_mom_island_of(C.byref(self.mo), C.byref(self.g), np.ascontiguousarray(ub, dtype=np.float64),
                                      isl, indj)

    def _arr(self, name, shape, dtype=np.float64):
        return np.ctypeslib.as_array(getattr(self.mo, name), shape=shape).copy()

    def ub(self) -> np.ndarray:
        """ub(2, 0:maxi+1, 0:maxj) as (maxj+1, maxi+2, 2)."""
        g = self.g
        return self._arr("ub", ((g.maxj + 1), (g.maxi + 2), 2))

    def psi(self) -> np.ndarray:
        g = self.g
        return self._arr("psi", ((g.maxj + 1), (g.maxi + 1)))

    def field(self, name: str) -> np.ndarray:
        """(l, 0:maxi+1, 0:maxj+1) fields h, rh (l = 3) and drag (l = 2) as (maxj+2, maxi+2, l)."""
        g = self.g
        nl = 3 if name in ("h", "rh") else 2
        return self._arr(name, ((g.maxj + 2), (g.maxi + 2), nl))


_ref_gold = None


def ref_gold():
    """The reference's GOLDSTEIN island path integral compiled in place (oracle/_ref), or None."""
    global _ref_gold
    if _ref_gold is None and os.path.exists(REF_GOLD_LIB):
        R = C.CDLL(REF_GOLD_LIB)
        R.ref_gold_island.argtypes = [_i] * 5 + [C.c_void_p] * 4 + [C.c_void_p] * 6 + [_d] + \
            [C.c_void_p] * 8 + [_i, _i, _dp]
        _ref_gold = R
    return _ref_gold


def ref_island(m: "MomOracle", isl: int, indj: int = 1, ub: np.ndarray = None) -> float:
    """goldstein_lib.f90's island() on the oracle member's arrays, each cut to
    the bounds goldstein.f90 allocates it with; ub as MomOracle.ub() gives it
    (default: the member's own)."""
    R = ref_gold()
    g, mo = m.g, m.mo
    MI, MJ, MK = g.maxi, g.maxj, g.maxk
    F = lambda a, dt=np.float64: np.asfortranarray(a, dtype=dt)
    vec = lambda name, lo, hi: F(np.array(getattr(g, name)[lo:hi + 1]))
    nis = max(mo.isles, 1)
    npi = F(np.ctypeslib.as_array(mo.npi, shape=(nis,)), np.int32)
    # (mpi, isles) column-major = the C arrays' layout p + mpi * q
    paths = [F(np.ctypeslib.as_array(getattr(mo, n), shape=(nis, mo.mpi)).T, np.int32)
             for n in ("lpisl", "ipisl", "jpisl")]
    drag = m.field("drag").transpose(2, 1, 0)[:, 1:MI + 2, 1:MJ + 1]
    rh = m.field("rh").transpose(2, 1, 0)
    tau = np.ctypeslib.as_array(mo.tau, shape=(MJ, MI, 2)).transpose(2, 1, 0)
    bp = np.ctypeslib.as_array(mo.bp, shape=(MK + 1, MJ + 2, MI + 2)).transpose(2, 1, 0)[1:MI + 2, 1:MJ + 1, 1:MK + 1]
    sbp = np.ctypeslib.as_array(mo.sbp, shape=(MJ + 2, MI + 2)).T[1:MI + 2, 1:MJ + 1]
    ku = np.ctypeslib.as_array(mo.ku, shape=(MJ, MI, 2)).transpose(2, 1, 0)
    mk = np.ctypeslib.as_array(mo.mk, shape=(MJ + 2, MI + 2)).T[1:MI + 2, 1:MJ + 1]
    ubv = (m.ub() if ub is None else ub).transpose(2, 1, 0)
    keep = [F(drag), F(rh), F(tau), F(bp), F(sbp), F(ku, np.int32), F(mk, np.int32), F(ubv)]
    vecs = [vec("s", 0, MJ), vec("sv", 0, MJ), vec("c", 0, MJ), vec("rcv", 1, MJ - 1), vec("dsv", 1, MJ - 1),
            vec("dz", 1, MK)]
    out = np.zeros(1)
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    R.ref_gold_island(MI, MJ, MK, mo.isles, mo.mpi, ptr(npi), *[ptr(a) for a in paths], *[ptr(a) for a in vecs],
                      g.dphi, *[ptr(a) for a in keep], isl, indj, out)
    return float(out[0])


def rho_from_ts(g: OgGrid, k1, go_ts) -> np.ndarray:
    rho = np.zeros((g.maxk + 1, g.maxj + 2, g.maxi + 2))
    lib().og_rho_from_ts(C.byref(g), go_ts.shape[-1], np.ascontiguousarray(k1, dtype=np.int32),
                         np.ascontiguousarray(go_ts, dtype=np.float64), rho)
    return rho


class SurfluxOracle:
    """EMBM's surflux for one member (og_surflux): the set-up tables of
    cgenie_amd.surface.SurfluxSetup, the atmosphere and sea-ice state
    ((maxj, maxi) arrays, copied) and the wind speed usurf from the four
    stress fields."""

    def __init__(self, g: OgGrid, k1, su, atemp, ashum, sich=None, sica=None, tice=None):
        self.g, self.su = g, su
        self.k1 = np.ascontiguousarray(k1, dtype=np.int32)
        mj, mi = g.maxj, g.maxi
        z = np.zeros((mj, mi))
        f = lambda a: np.ascontiguousarray(z if a is None else a, dtype=np.float64).copy()  # noqa: E731
        self.atemp, self.ashum, self.sich, self.sica, self.tice = f(atemp), f(ashum), f(sich), f(sica), f(tice)
        self.tabs = [np.ascontiguousarray(a, dtype=np.float64) for a in (su.solfor, su.albcl, su.ca, su.pmeadj)]
        self.rout = [np.ascontiguousarray(a, dtype=np.int32) for a in (su.iroff, su.jroff)]
        P = OgSfxPar()
        P.maxi, P.maxj, P.maxk, P.nyear = mi, mj, g.maxk, su.solfor.shape[0]
        for n in ("igrid", "polar_avg", "radfor_atchem", "scf", "rmax", "delf2x", "rdtdim", "rsictscsf", "olr_adj",
                  "olr_adj0", "t_eqm", "albsic_min", "albsic_max", "sich_max", "rate_co2", "rate_ch4", "rate_n2o"):
            setattr(P, n, getattr(su, n))
        P.solfor, P.albcl, P.ca, P.pmeadj = [a.ctypes.data_as(C.POINTER(_d)) for a in self.tabs]
        P.iroff, P.jroff = [a.ctypes.data_as(C.POINTER(C.c_int32)) for a in self.rout]
        self.P = P
        S = OgSfxState()
        S.atemp, S.ashum, S.sich, S.sica, S.tice = [a.ctypes.data_as(C.POINTER(_d)) for a in
                                                    (self.atemp, self.ashum, self.sich, self.sica, self.tice)]
        S.co2, S.ch4, S.n2o, S.istep = su.co2, su.ch4, su.n2o, 0
        self.S = S
        self.usurf = np.zeros((mj, mi))
        self.out = np.zeros((len(SF_NAMES), mj, mi))
        self.iters = np.zeros((mj, mi), dtype=np.int32)

    def set_wind(self, stress: np.ndarray):
        """stress (maxj, maxi, 4) as MomOracle.set_wind takes it."""
        st = np.ascontiguousarray(stress, dtype=np.float64)
        parts = [np.ascontiguousarray(st[..., q]) for q in range(4)]
        lib().og_sfx_usurf(C.byref(self.P), C.byref(self.g), *parts, self.usurf)

    def step(self, otemp: np.ndarray, osaln: np.ndarray, pco2: np.ndarray | None = None) -> np.ndarray:
        """One call; otemp, osaln (maxj, maxi) the ocean's top level; returns the outputs (SF_N, maxj, maxi)."""
        pc = None if pco2 is None else np.ascontiguousarray(pco2, dtype=np.float64)
        self._pc = pc
        lib().og_surflux(C.byref(self.P), C.byref(self.g), self.k1, C.byref(self.S), self.usurf,
                         np.ascontiguousarray(otemp, dtype=np.float64), np.ascontiguousarray(osaln, dtype=np.float64),
                         None if pc is None else pc.ctypes.data_as(C.c_void_p), self.out,
                         self.iters.ctypes.data_as(C.c_void_p))
        return self.out

    def slot(self, maxl: int, fx0sic=None, fwsic=None) -> np.ndarray:
        """The surface slot sflux (maxj, maxi, maxl) from the last call's outputs."""
        sf = np.zeros((self.g.maxj, self.g.maxi, maxl))
        keep = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (fx0sic, fwsic)]
        lib().og_surface_slot(C.byref(self.P), self.out, *[None if a is None else a.ctypes.data_as(C.c_void_p)
                                                           for a in keep], maxl, sf)
        return sf
