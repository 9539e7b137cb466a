"""Python binding of libcgocn.so (the C ABI in include/cgocn.h).

A thin ctypes layer: arrays cross the boundary in the reference's Fortran
layouts (see cgenie_amd/grid.py), exactly as a Fortran caller would pass
them.  There is no fallback: if the HIP library is missing or no GPU is
visible the constructor raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import grid as G

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcgocn.so")

FP64, FP32C = 0, 1

_d, _i32, _i64 = C.c_double, C.c_int32, C.c_int64
_dp = C.POINTER(C.c_double)


class GridDesc(C.Structure):
    _fields_ = [("maxi", _i32), ("maxj", _i32), ("maxk", _i32), ("k1", C.POINTER(_i32)), ("dphi", _d), ("rdphi", _d)] + \
        [(n, _dp) for n in ("rc", "rc2", "cv", "cv2", "rds", "dsv", "rdsv", "ds", "sv",
                            "dt", "rdz", "rdza", "dza", "ssmax", "dz")] + \
        [("ec", _d * 5), ("dsc", _d)]


class MemberParams(C.Structure):
    _fields_ = [("diff", _d * 2), ("diso", _i32), ("carry", _i32)]


class BgcParams(C.Structure):
    _fields_ = [("slot_DIC", _i32), ("slot_ALK", _i32), ("slot_PO4", _i32), ("slot_O2", _i32),
                ("slot_DOM_C", _i32), ("slot_DOM_P", _i32), ("caco3", _i32)] + \
        [(n, _d) for n in ("dtyr", "k0_PO4", "c0_PO4", "DOMfrac", "DOMlifetime", "red_POP_POC", "red_POP_PO2",
                           "red_PON_ALK", "red_POP_PON", "red_POC_CaCO3", "red_POC_CaCO3_pP", "POC_frac2",
                           "POC_dfrac2", "POC_c0frac2", "POC_eL1", "POC_eL2", "CaCO3_frac2", "CaCO3_eL1",
                           "CaCO3_eL2", "sinkrate", "remin_k_O2", "remin_c0_O2", "carb_tol", "dsc")] + \
        [("carb_maxit", _i32), ("airsea", _i32), ("gastransfer_a", _d), ("atm_pCO2", _d), ("atm_pO2", _d)]


class MomentumDesc(C.Structure):
    _fields_ = [(n, _dp) for n in ("s", "c", "rcv", "rds2")] + \
        [("labels", C.POINTER(_i32)), ("isles", _i32), ("npi", C.POINTER(_i32)), ("paths", C.POINTER(_i32)),
         ("adrag", _d), ("rel", _d)]


class SurfluxDesc(C.Structure):
    _fields_ = [("solfor", _dp), ("nyear", _i32), ("albcl", _dp), ("ca", _dp), ("pmeadj", _dp),
                ("iroff", C.POINTER(_i32)), ("jroff", C.POINTER(_i32)), ("igrid", _i32), ("polar_avg", _i32),
                ("radfor_atchem", _i32)] + \
        [(n, _d) for n in ("scf", "rmax", "delf2x", "rdtdim", "rsictscsf", "olr_adj", "olr_adj0", "t_eqm",
                           "albsic_min", "albsic_max", "sich_max", "rate_co2", "rate_ch4", "rate_n2o", "co2", "ch4",
                           "n2o")]


# the surface fluxes' outputs (cgo_download_surflux), in the order of CGO_SF_*
SF_NAMES = ("fxlho", "fxsho", "fxswo", "fxlwo", "evap_ocn", "pptn_ocn", "runoff_ocn", "runoff_land", "fxlha",
            "fxsha", "fxswa", "fxlwa", "evap_atm", "pptn_atm", "dthsic", "dtareasic", "albedo", "albice", "fxsw",
            "fx0a", "fx0o", "fxsen", "fxlw", "evap", "pptn", "fx0sic", "fx0neto", "tsfreez")


class Diag(C.Structure):
    _fields_ = [("steps", _i64), ("limps", _i32), ("carb_fail", _i32), ("carb_iters", _i64), ("dmax", _d),
                ("cost", _d), ("smean", _d)]


class CgoError(RuntimeError):
    pass


_lib = None


def load() -> C.CDLL:
    """Load the HIP engine; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CgoError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950) first")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    ip = C.POINTER(_i32)
    sig = {
        "cgo_abi_info": ([C.POINTER(_i64), C.c_char_p, C.c_size_t], C.c_int),
        "cgo_create": ([_i32, C.POINTER(P)], C.c_int),
        "cgo_destroy": ([P], C.c_int),
        "cgo_last_error": ([P, C.c_char_p, C.c_size_t], C.c_int),
        "cgo_sync": ([P], C.c_int),
        "cgo_grid_create": ([P, C.POINTER(GridDesc), ip], C.c_int),
        "cgo_member_create": ([P, _i32, _i32, _i32, C.POINTER(MemberParams), ip], C.c_int),
        "cgo_member_destroy": ([P, _i32], C.c_int),
        "cgo_member_bytes": ([P, _i32, _i32, _i32, _i32], _i64),
        "cgo_set_velocity": ([P, _i32, _dp], C.c_int),
        "cgo_set_sflux": ([P, _i32, _dp], C.c_int),
        "cgo_upload_state": ([P, _i32, _dp], C.c_int),
        "cgo_download_state": ([P, _i32, _dp], C.c_int),
        "cgo_download_rho": ([P, _i32, _dp], C.c_int),
        "cgo_inventory": ([P, _i32, _dp], C.c_int),
        "cgo_inventory_noalloc": ([P, _i32, _dp], C.c_int),
        "cgo_get_diag": ([P, _i32, C.POINTER(Diag)], C.c_int),
        "cgo_set_cost": ([P, _i32, _dp], C.c_int),
        "cgo_checkpoint_bytes": ([P, _i32], _i64),
        "cgo_checkpoint": ([P, _i32, C.c_void_p, C.c_size_t], C.c_int),
        "cgo_restore": ([P, _i32, C.c_void_p, C.c_size_t], C.c_int),
        "cgo_download_cost": ([P, _i32, _dp], C.c_int),
        "cgo_bgc_default_params": ([C.POINTER(BgcParams), _d], None),
        "cgo_bgc_init": ([P, _i32, C.POINTER(BgcParams), _dp, _dp], C.c_int),
        "cgo_bgc_download": ([P, _i32, _i32, _dp], C.c_int),
        "cgo_set_windspeed": ([P, _i32, _dp], C.c_int),
        "cgo_step": ([P, ip, _i32, _i32], C.c_int),
        "cgo_tstepo": ([P, ip, _i32, _i32], C.c_int),
        "cgo_bio_step": ([P, ip, _i32], C.c_int),
        "cgo_tracercoupling": ([P, ip, _i32], C.c_int),
        "cgo_graph_capture": ([P, ip, _i32, _i32, ip], C.c_int),
        "cgo_graph_launch": ([P, _i32, _i32], C.c_int),
        "cgo_graph_destroy": ([P, _i32], C.c_int),
        "cgo_set_tile": ([P, _i32, _i32], C.c_int),
        "cgo_set_flux_cols": ([P, _i32, _i32], C.c_int),
        "cgo_set_flux_ring": ([P, _i32, _i32], C.c_int),
        "cgo_set_flux_iso": ([P, _i32, _i32], C.c_int),
        "cgo_time_kernel": ([P, _i32, ip, _i32, _i32, _dp], C.c_int),
        "cgo_grid_set_momentum": ([P, _i32, C.POINTER(MomentumDesc)], C.c_int),
        "cgo_set_wind": ([P, _i32, _dp, _dp, _dp, _dp, _d], C.c_int),
        "cgo_download_velocity": ([P, _i32, _dp], C.c_int),
        "cgo_download_psi": ([P, _i32, _dp], C.c_int),
        "cgo_grid_set_surflux": ([P, _i32, C.POINTER(SurfluxDesc)], C.c_int),
        "cgo_set_surface": ([P, _i32] + [_dp] * 7, C.c_int),
        "cgo_surflux": ([P, ip, _i32], C.c_int),
        "cgo_surface_slot": ([P, _i32, _dp, _dp], C.c_int),
        "cgo_download_surflux": ([P, _i32, _dp, _dp, _dp, _dp], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    check_abi(L)
    _lib = L
    return L


def check_abi(L: C.CDLL) -> None:
    """Refuse a library built from another version of include/cgocn.h: its
    struct sizes must equal the ctypes mirrors below and its build stamp the
    hash of the header in this tree (cgenie_amd/build.py writes both)."""
    from .build import abi_stamp

    sizes = (_i64 * 4)()
    stamp = C.create_string_buffer(64)
    if L.cgo_abi_info(sizes, stamp, 64) != 0:
        raise CgoError("cgo_abi_info failed")
    want = [C.sizeof(GridDesc), C.sizeof(MemberParams), C.sizeof(BgcParams), C.sizeof(Diag)]
    if list(sizes) != want:
        raise CgoError(f"{LIB_PATH}: struct sizes {list(sizes)} differ from this binding's {want}; rebuild it")
    if stamp.value.decode() != abi_stamp():
        raise CgoError(f"{LIB_PATH} was built from another include/cgocn.h (stamp {stamp.value.decode()}, "
                       f"header {abi_stamp()}); rebuild it with __graft_entry__.build()")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _ids(members) -> tuple:
    arr = (_i32 * len(members))(*members)
    return arr, len(members)


def default_bgc_params(dtyr: float, ts: G.TracerSet | None = None, **kw) -> BgcParams:
    p = BgcParams()
    load().cgo_bgc_default_params(C.byref(p), dtyr)
    if ts is not None:
        p.slot_DIC, p.slot_ALK, p.slot_PO4, p.slot_O2 = ts.DIC, ts.ALK, ts.PO4, ts.O2
        p.slot_DOM_C, p.slot_DOM_P = ts.DOM_C, ts.DOM_P
        p.caco3 = 1 if (ts.DIC >= 0 and ts.ALK >= 0) else 0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Engine:
    """One device context (stream, grids, members)."""

    def __init__(self, device: int = 0):
        self.L = load()
        self.h = C.c_void_p()
        rc = self.L.cgo_create(device, C.byref(self.h))
        if rc != 0:
            raise CgoError(f"cgo_create({device}) failed ({rc}): no usable HIP device")
        self._keep = []
        self.members = {}

    def close(self):
        if self.h:
            self.L.cgo_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int, what: str):
        if rc != 0:
            buf = C.create_string_buffer(512)
            self.L.cgo_last_error(self.h, buf, 512)
            raise CgoError(f"{what}: rc={rc}: {buf.value.decode()}")

    # -- grids and members
    def add_grid(self, g: G.Grid, k1: np.ndarray) -> int:
        k1 = np.ascontiguousarray(k1, dtype=np.int32)
        keep = [k1]
        d = GridDesc()
        d.maxi, d.maxj, d.maxk = g.maxi, g.maxj, g.maxk
        d.k1 = k1.ctypes.data_as(C.POINTER(_i32))
        d.dphi, d.rdphi = g.dphi, g.rdphi
        for n in ("rc", "rc2", "cv", "cv2", "rds", "dsv", "rdsv", "ds", "sv", "dt", "rdz", "rdza", "dza", "ssmax", "dz"):
            a = np.ascontiguousarray(getattr(g, n), dtype=np.float64)
            keep.append(a)
            setattr(d, n, _ptr(a))
        for q in range(5):
            d.ec[q] = float(g.ec[q])
        d.dsc = G.DSC
        gid = _i32()
        self._chk(self.L.cgo_grid_create(self.h, C.byref(d), C.byref(gid)), "grid_create")
        self._keep.append(keep)
        return gid.value

    def set_momentum(self, grid_id: int, g: G.Grid, world: "G.World", adrag: float = 2.5, rel: float = 0.9):
        """The grid's velocity solve (cgo_grid_set_momentum): world holds the
        .psiles labels and .paths island paths (grid.read_world); adrag in
        days and rel as in ini_gold_nml."""
        keep = [np.ascontiguousarray(getattr(g, n), dtype=np.float64) for n in ("s", "c", "rcv", "rds2")]
        labels = np.ascontiguousarray(world.labels, dtype=np.int32)
        npi = np.ascontiguousarray(world.npi if world.isles else np.zeros(1), dtype=np.int32)
        paths = np.ascontiguousarray(world.paths if world.isles else np.zeros(3), dtype=np.int32)
        keep += [labels, npi, paths]
        d = MomentumDesc()
        d.s, d.c, d.rcv, d.rds2 = (_ptr(a) for a in keep[:4])
        d.labels = labels.ctypes.data_as(C.POINTER(_i32))
        d.isles = int(world.isles)
        d.npi = npi.ctypes.data_as(C.POINTER(_i32))
        d.paths = paths.ctypes.data_as(C.POINTER(_i32))
        d.adrag, d.rel = float(adrag), float(rel)
        self._chk(self.L.cgo_grid_set_momentum(self.h, grid_id, C.byref(d)), "grid_set_momentum")

    def set_wind(self, m: int, stress: np.ndarray, scf: float = 2.0):
        """Wind stress (maxj, maxi, 4): x and y at u points, x and y at v
        points in N m-2 (step_goldstein's stressxu/yu/xv/yv); the first call
        switches the member to solved velocities."""
        st = np.asarray(stress, dtype=np.float64)
        parts = [np.ascontiguousarray(st[..., q]) for q in range(4)]
        self._chk(self.L.cgo_set_wind(self.h, m, *(_ptr(a) for a in parts), float(scf)), "set_wind")

    def set_surflux(self, grid_id: int, su) -> None:
        """The grid's surface fluxes (cgo_grid_set_surflux) from a
        cgenie_amd.surface.SurfluxSetup."""
        keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (su.solfor, su.albcl, su.ca, su.pmeadj)]
        rout = [np.ascontiguousarray(a, dtype=np.int32) for a in (su.iroff, su.jroff)]
        d = SurfluxDesc()
        d.solfor, d.albcl, d.ca, d.pmeadj = (_ptr(a) for a in keep)
        d.nyear = int(su.solfor.shape[0])
        d.iroff, d.jroff = (a.ctypes.data_as(C.POINTER(_i32)) for a in rout)
        d.igrid, d.polar_avg, d.radfor_atchem = int(su.igrid), int(su.polar_avg), int(su.radfor_atchem)
        for n in ("scf", "rmax", "delf2x", "rdtdim", "rsictscsf", "olr_adj", "olr_adj0", "t_eqm", "albsic_min",
                  "albsic_max", "sich_max", "rate_co2", "rate_ch4", "rate_n2o", "co2", "ch4", "n2o"):
            setattr(d, n, float(getattr(su, n)))
        self._chk(self.L.cgo_grid_set_surflux(self.h, grid_id, C.byref(d)), "grid_set_surflux")

    def set_surface(self, m: int, atemp, ashum=None, sich=None, sica=None, tice=None, fx0sic=None, fwsic=None):
        """Switch member m's surface fluxes on with this air and sea-ice
        state ((maxj, maxi) arrays; None = zero), or off with atemp None."""
        keep = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
                for a in (atemp, ashum, sich, sica, tice, fx0sic, fwsic)]
        self._chk(self.L.cgo_set_surface(self.h, m, *(None if a is None else _ptr(a) for a in keep)), "set_surface")

    def surflux(self, members) -> None:
        """One surflux call of the listed members."""
        ids = np.ascontiguousarray(list(members), dtype=np.int32)
        self._chk(self.L.cgo_surflux(self.h, ids.ctypes.data_as(C.POINTER(_i32)), len(ids)), "surflux")

    def surface_slot(self, m: int, fx0sic=None, fwsic=None) -> None:
        """Form member m's surface slot again with new sea-ice fluxes ((maxj, maxi); None = zero)."""
        keep = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (fx0sic, fwsic)]
        self._chk(self.L.cgo_surface_slot(self.h, m, *(None if a is None else _ptr(a) for a in keep)), "surface_slot")

    def surface_outputs(self, m: int, g: G.Grid) -> dict:
        """The last call's outputs by name ((maxj, maxi) each), with ashum,
        tice and the scalars co2, ch4, n2o, istep."""
        out = np.zeros((len(SF_NAMES), g.maxj, g.maxi))
        ashum = np.zeros((g.maxj, g.maxi))
        tice = np.zeros((g.maxj, g.maxi))
        gases = np.zeros(4)
        self._chk(self.L.cgo_download_surflux(self.h, m, _ptr(out), _ptr(ashum), _ptr(tice), _ptr(gases)),
                  "download_surflux")
        d = {n: out[q] for q, n in enumerate(SF_NAMES)}
        d.update(ashum=ashum, tice=tice, co2=gases[0], ch4=gases[1], n2o=gases[2], istep=int(gases[3]))
        return d

    def velocity(self, m: int, g: G.Grid) -> np.ndarray:
        """go_u(3, maxi, maxj, maxk) as (maxk, maxj, maxi, 3)."""
        out = np.zeros((g.maxk, g.maxj, g.maxi, 3))
        self._chk(self.L.cgo_download_velocity(self.h, m, _ptr(out)), "download_velocity")
        return out

    def psi(self, m: int, g: G.Grid) -> np.ndarray:
        """go_psi(0:maxi, 0:maxj) as (maxj+1, maxi+1)."""
        out = np.zeros((g.maxj + 1, g.maxi + 1))
        self._chk(self.L.cgo_download_psi(self.h, m, _ptr(out)), "download_psi")
        return out

    def add_member(self, grid_id: int, L: int, precision: int = FP64, diff=(None, None), diso: bool = True,
                   carry: bool = True) -> int:
        p = MemberParams()
        p.diff[0], p.diff[1] = diff
        p.diso = 1 if diso else 0
        p.carry = 1 if carry else 0
        mid = _i32()
        self._chk(self.L.cgo_member_create(self.h, grid_id, L, precision, C.byref(p), C.byref(mid)), "member_create")
        self.members[mid.value] = L
        return mid.value

    def remove_member(self, m: int):
        self._chk(self.L.cgo_member_destroy(self.h, m), "member_destroy")
        self.members.pop(m, None)

    def member_bytes(self, grid_id: int, L: int, precision: int, bgc: bool) -> int:
        return int(self.L.cgo_member_bytes(self.h, grid_id, L, precision, 1 if bgc else 0))

    # -- data
    def set_velocity(self, m: int, u: np.ndarray):
        u = np.ascontiguousarray(u, dtype=np.float64)
        self._chk(self.L.cgo_set_velocity(self.h, m, _ptr(u)), "set_velocity")

    def set_sflux(self, m: int, sflux: np.ndarray | None):
        if sflux is None:
            self._chk(self.L.cgo_set_sflux(self.h, m, None), "set_sflux")
        else:
            s = np.ascontiguousarray(sflux, dtype=np.float64)
            self._chk(self.L.cgo_set_sflux(self.h, m, _ptr(s)), "set_sflux")

    def upload(self, m: int, go_ts: np.ndarray):
        a = np.ascontiguousarray(go_ts, dtype=np.float64)
        self._chk(self.L.cgo_upload_state(self.h, m, _ptr(a)), "upload_state")

    def download(self, m: int, shape) -> np.ndarray:
        out = np.zeros(shape)
        self._chk(self.L.cgo_download_state(self.h, m, _ptr(out)), "download_state")
        return out

    def rho(self, m: int, g: G.Grid) -> np.ndarray:
        out = np.zeros((g.maxk, g.maxj, g.maxi))
        self._chk(self.L.cgo_download_rho(self.h, m, _ptr(out)), "download_rho")
        return out

    def inventory(self, m: int) -> np.ndarray:
        out = np.zeros(self.members[m])
        self._chk(self.L.cgo_inventory(self.h, m, _ptr(out)), "inventory")
        return out

    def inventory_noalloc(self, m: int) -> np.ndarray:
        """cgo_inventory through the library's static scratch (no device allocation)."""
        out = np.zeros(self.members[m])
        self._chk(self.L.cgo_inventory_noalloc(self.h, m, _ptr(out)), "inventory_noalloc")
        return out

    def diag(self, m: int) -> Diag:
        d = Diag()
        self._chk(self.L.cgo_get_diag(self.h, m, C.byref(d)), "get_diag")
        return d

    def checkpoint(self, m: int) -> bytes:
        """The member's exact resident state (cgo_checkpoint)."""
        n = int(self.L.cgo_checkpoint_bytes(self.h, m))
        if n < 0:
            self._chk(n, "checkpoint_bytes")
        buf = C.create_string_buffer(n)
        self._chk(self.L.cgo_checkpoint(self.h, m, buf, n), "checkpoint")
        return buf.raw

    def restore(self, m: int, blob: bytes):
        """Continue member m from a checkpoint (cgo_restore)."""
        self._chk(self.L.cgo_restore(self.h, m, blob, len(blob)), "restore")

    def bgc_init(self, m: int, p: BgcParams, light: np.ndarray, mld: np.ndarray):
        lt = np.ascontiguousarray(light, dtype=np.float64)
        ml = np.ascontiguousarray(mld, dtype=np.float64)
        self._chk(self.L.cgo_bgc_init(self.h, m, C.byref(p), _ptr(lt), _ptr(ml)), "bgc_init")

    def bgc_download(self, m: int, what: int, shape) -> np.ndarray:
        out = np.zeros(shape)
        self._chk(self.L.cgo_bgc_download(self.h, m, what, _ptr(out)), "bgc_download")
        return out

    def set_windspeed(self, m: int, wspeed: np.ndarray):
        """Wind speed (maxj, maxi) in m/s for the air-sea exchange
        (par_phys_windspeed, biogem.f90:2228-2232)."""
        w = np.ascontiguousarray(wspeed, dtype=np.float64)
        self._chk(self.L.cgo_set_windspeed(self.h, m, _ptr(w)), "set_windspeed")

    def atmosphere(self, m: int, maxi: int, maxj: int) -> np.ndarray:
        """(maxj, maxi, 2) atmospheric pCO2 and pO2 (atm)."""
        return self.bgc_download(m, 6, (maxj, maxi, 2))

    def airsea_flux(self, m: int, maxi: int, maxj: int) -> np.ndarray:
        """(maxj, maxi, 2) the last step's ocean -> atmosphere CO2 and O2 flux (mol yr-1)."""
        return self.bgc_download(m, 7, (maxj, maxi, 2))

    # -- stepping
    def step(self, members, nsteps: int = 1):
        arr, n = _ids(members)
        self._chk(self.L.cgo_step(self.h, arr, n, nsteps), "step")

    def tstepo(self, members, nsteps: int = 1):
        """tstepo (goldstein.f90:2280-2432): transport only."""
        arr, n = _ids(members)
        self._chk(self.L.cgo_tstepo(self.h, arr, n, nsteps), "tstepo")

    def bio_step(self, members):
        """step_biogem (biogem.f90:528-1877) for the carried species."""
        arr, n = _ids(members)
        self._chk(self.L.cgo_bio_step(self.h, arr, n), "bio_step")

    def tracercoupling(self, members):
        """biogem_tracercoupling (biogem.f90:1885-2077); follows bio_step."""
        arr, n = _ids(members)
        self._chk(self.L.cgo_tracercoupling(self.h, arr, n), "tracercoupling")

    def capture(self, members, nsteps: int) -> int:
        arr, n = _ids(members)
        gid = _i32()
        self._chk(self.L.cgo_graph_capture(self.h, arr, n, nsteps, C.byref(gid)), "graph_capture")
        return gid.value

    def launch(self, graph: int, nreps: int = 1):
        self._chk(self.L.cgo_graph_launch(self.h, graph, nreps), "graph_launch")

    def drop_graph(self, graph: int):
        self._chk(self.L.cgo_graph_destroy(self.h, graph), "graph_destroy")

    def sync(self):
        self._chk(self.L.cgo_sync(self.h), "sync")

    def set_tile(self, precision: int, tracer_block: int, cols: int | None = None, ring: int | None = None,
                 iso_prepass: int | None = None):
        self._chk(self.L.cgo_set_tile(self.h, precision, tracer_block), "set_tile")
        if cols is not None:
            self._chk(self.L.cgo_set_flux_cols(self.h, precision, cols), "set_flux_cols")
        if ring is not None:
            self._chk(self.L.cgo_set_flux_ring(self.h, precision, ring), "set_flux_ring")
        if iso_prepass is not None:
            self._chk(self.L.cgo_set_flux_iso(self.h, precision, iso_prepass), "set_flux_iso")

    def time_kernel(self, which: int, members, reps: int) -> float:
        arr, n = _ids(members)
        ms = _d()
        self._chk(self.L.cgo_time_kernel(self.h, which, arr, n, reps, C.byref(ms)), "time_kernel")
        return ms.value
