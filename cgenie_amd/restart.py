"""Restart files in the reference's formats, and exact engine checkpoints.

Two kinds of saved state:

* Exact checkpoints (``save_checkpoint`` / ``load_checkpoint``): the
  engine's own binary blob of one member (cgo_checkpoint / cgo_restore in
  include/cgocn.h: a header with the grid fingerprint, tracer count,
  precision, carry, parameters and BIOGEM flags, then every device array of
  the member).  Restoring one continues the run bit for bit; the serving
  layer uses them to suspend and resume runs.  File layout: the 8-byte magic
  ``CGOCKPT1`` opens the blob, which is written as it comes from the engine.

* Restart files in the reference's netCDF formats (NetCDF classic, the
  format ``NF90_CREATE(..., NF90_CLOBBER, ...)`` writes), so states move
  between this engine and a reference build:

  - GOLDSTEIN (``outm_netcdf``, goldstein_data.f90:153-354; read back by
    ``inm_netcdf``, 10-149): dimensions nrecs = 1, longitude, latitude,
    depth; float coordinates longitude/latitude/depth; int scalars ioffset,
    iyear, imonth, iday; double (depth, latitude, longitude) temp and
    salinity (GOLDSTEIN units: degrees C and the anomaly S - saln0; every
    level of a wet column is written, land columns as zero, as the
    reference's column mask does), uvel and vvel; double (latitude,
    longitude) evap, late and sens.
  - BIOGEM (``sub_data_netCDF_ncrstsave``, biogem_data_netCDF.f90:24-142,
    called from ``biogem_save_restart``, biogem.f90:2280-2360; read back by
    ``sub_data_load_rst``, biogem_data.f90:438-568): dimensions lon, lat,
    lon_edges, lat_edges, zt, zt_edges; double coordinates; one float
    variable per selected tracer, ``ocn_<name>`` (BIOGEM units: T in K,
    absolute S, concentrations in mol/kg) and ``bio_part_<name>``, with the
    depth axis surface first and land cells holding the netCDF double fill
    value.  The reference stores these in single precision, so a state that
    passes through them keeps its BIOGEM tracers to float precision only;
    before writing, the reference folds the rain that the closed-system
    sediment would return at the next step into the bottom cells
    (biogem.f90:2298-2315), as ``biogem_restart_state`` does here.

Host-side code (numpy + scipy.io.netcdf_file); the device state is
reached through the engine binding.
"""
from __future__ import annotations

import math

import numpy as np

from . import grid as G

NC_FILL_DOUBLE = 9.9692099683868690e36          # nf90_fill_double (netcdf.inc)
# tracer names of the reference's tracer tables (data/main/tracer_define.ocn, .sed)
OCN_NAMES = {"T": "temp", "S": "sal", "DIC": "DIC", "PO4": "PO4", "O2": "O2", "ALK": "ALK", "DOM_C": "DOM_C",
             "DOM_P": "DOM_P"}
SED_NAMES = ("POC", "POP", "CaCO3", "POC_frac2", "CaCO3_frac2")
CONV_M3_KG = 1027.649                            # gem_cmn.f90:510


# --------------------------------------------------------------- checkpoints
def save_checkpoint(eng, m: int, path: str) -> int:
    """Write member m's exact checkpoint to `path`; returns its size."""
    blob = eng.checkpoint(m)
    with open(path, "wb") as f:
        f.write(blob)
    return len(blob)


def load_checkpoint(eng, m: int, path: str) -> None:
    """Continue member m (same grid, tracer count, precision, carry) from `path`."""
    with open(path, "rb") as f:
        blob = f.read()
    if blob[:8] != b"CGOCKPT1":
        raise ValueError(f"{path}: not an engine checkpoint")
    eng.restore(m, blob)


# --------------------------------------------------------------- coordinates
def goldstein_coordinates(g: G.Grid):
    """Longitude, latitude and depth of the cell centres as GOLDSTEIN writes
    them (goldstein.f90:1842-1848, 1876-1884, 1919-1923; igrid = 0):
    degrees east from -260, degrees north, metres (positive down)."""
    deg_to_rad = math.pi / 180.0
    phi0 = -260.0 * deg_to_rad
    lon = np.array([360.0 * (i - 0.5) / g.maxi + phi0 / deg_to_rad for i in range(1, g.maxi + 1)])
    lat = np.array([math.asin(g.s[j]) * 180.0 / math.pi for j in range(1, g.maxj + 1)])
    depth = np.array([abs(G.DSC * g.zro[k]) for k in range(1, g.maxk + 1)])   # level k (bottom = 1)
    return lon, lat, depth


# --------------------------------------------------------------- GOLDSTEIN
def write_goldstein_restart(path: str, g: G.Grid, k1: np.ndarray, go_ts: np.ndarray, u: np.ndarray,
                            iyear: int = 0, imonth: int = 1, iday: int = 1, ioffset: int = 0,
                            evap=None, late=None, sens=None) -> None:
    """outm_netcdf: T and S of go_ts (numpy (maxk, maxj, maxi, maxl)) and
    the horizontal velocities of u ((maxk, maxj+1, maxi+1, 3))."""
    from scipy.io import netcdf_file

    lon, lat, depth = goldstein_coordinates(g)
    # landmask(i,j,:) = 1 for every level of a wet column (goldstein_data.f90:212-219)
    colwet = (k1[1:g.maxj + 1, 1:g.maxi + 1] <= g.maxk).astype(np.float64)
    zero2 = np.zeros((g.maxj, g.maxi))
    with netcdf_file(path, "w", version=1) as f:
        f.createDimension("nrecs", 1)
        f.createDimension("longitude", g.maxi)
        f.createDimension("latitude", g.maxj)
        f.createDimension("depth", g.maxk)
        for name, val, dim in (("longitude", lon, "longitude"), ("latitude", lat, "latitude"),
                               ("depth", depth, "depth")):
            v = f.createVariable(name, "f4", (dim,))
            v[:] = val.astype(np.float32)
        f.variables["longitude"].units = b"degrees_east"
        f.variables["longitude"].long_name = b"longitude"
        f.variables["latitude"].units = b"degrees_north"
        f.variables["latitude"].long_name = b"latitude"
        for name, val in (("ioffset", ioffset), ("iyear", iyear), ("imonth", imonth), ("iday", iday)):
            f.createVariable(name, "i4", ("nrecs",))[:] = np.array([val], dtype=np.int32)
        for name, val in (("temp", go_ts[..., 0] * colwet[None]), ("salinity", go_ts[..., 1] * colwet[None]),
                          ("uvel", u[:, 1:, 1:, 0]), ("vvel", u[:, 1:, 1:, 1])):
            f.createVariable(name, "f8", ("depth", "latitude", "longitude"))[:] = val
        for name, val in (("evap", evap), ("late", late), ("sens", sens)):
            f.createVariable(name, "f8", ("latitude", "longitude"))[:] = zero2 if val is None else val


def read_goldstein_restart(path: str) -> dict:
    """inm_netcdf: temp, salinity, uvel, vvel (depth, latitude, longitude),
    evap, late, sens and the date scalars."""
    from scipy.io import netcdf_file

    with netcdf_file(path, "r", mmap=False) as f:
        out = {k: np.array(f.variables[k][:], dtype=np.float64)
               for k in ("temp", "salinity", "uvel", "vvel", "evap", "late", "sens")}
        for k in ("ioffset", "iyear", "imonth", "iday"):
            out[k] = int(f.variables[k][0])
    return out


# --------------------------------------------------------------- BIOGEM
def biogem_restart_state(ocn: np.ndarray, settle: np.ndarray, rM: np.ndarray, k1: np.ndarray, tset: G.TracerSet,
                         caco3: bool = True, red_PON_ALK_x_POP_PON: float = -16.0) -> np.ndarray:
    """The ocean state biogem_save_restart writes with the closed-system
    sediment (biogem.f90:2298-2315): the rain at k1 (settle, numpy
    (maxj, maxi, 5): POC, POP, CaCO3 and the fractions) returned to the bottom
    cells through conv_sed_ocn (gem_util.f90:79-98) without the redox
    factor, ocn += rM conv settle.  ocn is numpy (maxk, maxj, maxi, maxl) in
    BIOGEM units; rM the cell 1/mass (maxk, maxj, maxi)."""
    out = ocn.copy()
    maxj, maxi = k1.shape[0] - 2, k1.shape[1] - 2
    maxk = ocn.shape[0]
    red_POP_PO2, red_POP_POC = -138.0, 106.0
    cO2POP = -4.0 / 2.0
    cO2POC = red_POP_PO2 / red_POP_POC - cO2POP / red_POP_POC
    conv = []   # (sediment slot, tracer slot, factor)
    if tset.DIC >= 0:
        conv.append((0, tset.DIC, 1.0))
    if tset.O2 >= 0:
        conv += [(0, tset.O2, cO2POC), (1, tset.O2, cO2POP)]
    if tset.PO4 >= 0:
        conv.append((1, tset.PO4, 1.0))
    if tset.ALK >= 0:
        conv.append((1, tset.ALK, red_PON_ALK_x_POP_PON))
    if caco3 and tset.DIC >= 0 and tset.ALK >= 0:
        conv += [(2, tset.DIC, 1.0), (2, tset.ALK, 2.0)]
    for j in range(maxj):
        for i in range(maxi):
            kb = k1[j + 1, i + 1]
            if kb > maxk:
                continue
            for s, l, f in conv:
                out[kb - 1, j, i, l] = out[kb - 1, j, i, l] + rM[kb - 1, j, i] * f * settle[j, i, s]
    return out


def write_biogem_restart(path: str, g: G.Grid, k1: np.ndarray, ocn: np.ndarray, part: np.ndarray,
                         tset: G.TracerSet, year: float = 0.0) -> None:
    """sub_data_netCDF_ncrstsave: ocn (maxk, maxj, maxi, maxl; BIOGEM units,
    carried species and T, S) and part (maxk, maxj, maxi, 5)."""
    from scipy.io import netcdf_file

    lon, lat, _ = goldstein_coordinates(g)
    # mid-level and bottom depths as sub_init_phys_ocn forms them (biogem_data.f90:1098-1140):
    # the half-thickness dza, with dz(maxk)/2 at the surface level
    dzal = g.dza.copy()
    dzal[g.maxk] = g.dz[g.maxk] / 2.0
    Dmid = np.array([sum(G.DSC * dzal[q] for q in range(k, g.maxk + 1)) if k >= 1 else 0.0
                     for k in range(0, g.maxk + 1)])
    Dbot = np.array([sum(G.DSC * g.dz[q] for q in range(k, g.maxk + 1)) if k >= 1 else 0.0
                     for k in range(0, g.maxk + 1)])
    wet = G.wet_mask(g, k1)                       # (maxk, maxj, maxi)
    with netcdf_file(path, "w", version=1) as f:
        f.title = f"BIOGEM restart @ year {int(year):8d}".encode()
        f.createDimension("lon", g.maxi)
        f.createDimension("lat", g.maxj)
        f.createDimension("lon_edges", g.maxi + 1)
        f.createDimension("lat_edges", g.maxj + 1)
        f.createDimension("zt", g.maxk)
        f.createDimension("zt_edges", g.maxk + 1)
        f.createVariable("lon", "f8", ("lon",))[:] = lon
        f.createVariable("lat", "f8", ("lat",))[:] = lat
        f.createVariable("lon_edges", "f8", ("lon_edges",))[:] = np.array(
            [360.0 * (i - 1.0) / g.maxi - 260.0 for i in range(1, g.maxi + 2)])
        f.createVariable("lat_edges", "f8", ("lat_edges",))[:] = np.array(
            [math.asin(g.sv[j]) * 180.0 / math.pi for j in range(0, g.maxj + 1)])
        f.createVariable("zt", "f8", ("zt",))[:] = Dmid[g.maxk:0:-1]
        f.createVariable("zt_edges", "f8", ("zt_edges",))[:] = np.concatenate([[0.0], Dbot[g.maxk:0:-1]])

        def put(name, field):
            v = f.createVariable(name, "f4", ("zt", "lat", "lon"))
            x = np.where(wet, field, NC_FILL_DOUBLE)[::-1]       # surface first
            v[:] = x.astype(np.float32)
            v.missing_value = NC_FILL_DOUBLE

        put("ocn_temp", ocn[..., 0])
        put("ocn_sal", ocn[..., 1])
        for key in ("DIC", "PO4", "O2", "ALK", "DOM_C", "DOM_P"):
            slot = getattr(tset, key)
            if slot >= 0:
                put("ocn_" + OCN_NAMES[key], ocn[..., slot])
        for q, name in enumerate(SED_NAMES):
            put("bio_part_" + name, part[..., q])


def read_biogem_restart(path: str, g: G.Grid, tset: G.TracerSet):
    """sub_data_load_rst: (ocn, part) in the layouts write_biogem_restart takes
    (land cells zero); variables the file lacks stay zero."""
    from scipy.io import netcdf_file

    ocn = np.zeros((g.maxk, g.maxj, g.maxi, tset.L))
    part = np.zeros((g.maxk, g.maxj, g.maxi, 5))
    with netcdf_file(path, "r", mmap=False) as f:
        def get(name):
            x = np.array(f.variables[name][:], dtype=np.float64)[::-1]
            return np.where(x >= 0.99 * NC_FILL_DOUBLE, 0.0, x)

        names = f.variables.keys()
        slots = {"T": 0, "S": 1}
        slots.update({k: getattr(tset, k) for k in ("DIC", "PO4", "O2", "ALK", "DOM_C", "DOM_P")})
        for key, slot in slots.items():
            if slot >= 0 and "ocn_" + OCN_NAMES[key] in names:
                ocn[..., slot] = get("ocn_" + OCN_NAMES[key])
        for q, name in enumerate(SED_NAMES):
            if "bio_part_" + name in names:
                part[..., q] = get("bio_part_" + name)
    return ocn, part
