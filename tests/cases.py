"""Seeded test cases shared by the oracle and the engine tests."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from cgenie_amd import grid as G
import oracle_ffi as O


@dataclass
class Case:
    g: G.Grid
    og: object
    k1: np.ndarray
    u: np.ndarray
    ts: np.ndarray
    tset: G.TracerSet
    diff: tuple
    sflux: np.ndarray | None = None
    light: np.ndarray | None = None
    mld: np.ndarray | None = None


def make_case(maxi=36, maxj=36, maxk=16, kind="ts", L=None, seed=1, kappa_h=2000.0, kappa_v=2.7e-5,
              random_u=False, sflux=False, mld_max=0.0, vscale=1.0) -> Case:
    g = G.make_grid(maxi, maxj, maxk)
    og = O.grid(maxi, maxj, maxk, kappa_h=kappa_h, kappa_v=kappa_v)
    k1 = G.make_k1(g, seed)
    u = G.make_random_velocity(g, seed + 1) if random_u else G.make_velocity(g, k1, seed + 1, scale=vscale)
    tset = G.tracer_set(kind, L)
    ts = G.make_tracers(g, k1, tset, seed + 2)
    sf = None
    if sflux:
        rng = np.random.default_rng(seed + 3)
        sf = rng.normal(0, 1e-3, (maxj, maxi, tset.L))
    return Case(g, og, k1, u, ts, tset, G.diffusivities(kappa_h, kappa_v), sf, G.make_light(g, seed + 4),
                G.make_mld(g, seed + 5, mld_max))


def wet(case: Case) -> np.ndarray:
    return G.wet_mask(case.g, case.k1)


def rel_err(a: np.ndarray, b: np.ndarray, mask: np.ndarray, floor: float = 1e-3) -> np.ndarray:
    """Per-tracer max over wet cells of |a-b| / max(|b|, floor * max|b_l|):
    relative per cell, with a floor for cells whose value is near zero
    (S is stored as the anomaly S - 34.9, dyes and nutrients reach 0)."""
    aw, bw = a[mask], b[mask]
    scale = np.maximum(np.abs(bw).max(axis=0), 1e-300)
    den = np.maximum(np.abs(bw), floor * scale[None, :])
    return (np.abs(aw - bw) / den).max(axis=0)
