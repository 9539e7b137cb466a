"""Writes tests/golden/drift_1000y.npz: the oracle's state after 1000 model
years of the precision-study configuration (36x36x16, 24 tracers: T, S, DIC,
ALK, PO4, O2, DOM_P and 17 dyes; BIOGEM every ocean step; kappa_h 2000 m2/s,
kappa_v 2.7e-5 m2/s; seed 2024; 96,000 ocean steps).

Stored: the wet-cell state (N_wet, 24) in float64 after year 1000, the
SHA-256 of the float64 wet-cell state and the per-tracer inventories at
years 1, 10, 100 and 1000.  The GPU test (tests/test_gpu_configs.py) runs the
same configuration in FP64 (bit-identical to these) and in the
FP32-compensated form (within 1e-6 per tracer).

    python tests/golden/make_drift_golden.py            # ~20 min on one core

Test infrastructure: the oracle is the checker here.
"""
from __future__ import annotations

import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tests")]

from cgenie_amd import grid as G  # noqa: E402

OUT = os.path.join(HERE, "drift_1000y.npz")
CHECKPOINTS = (1, 10, 100, 1000)
SEED = 2024


def inputs():
    g = G.make_grid(36, 36, 16)
    k1 = G.make_k1(g, SEED)
    tset = G.tracer_set("config2")
    u = G.make_velocity(g, k1, SEED + 1)
    ts = G.make_tracers(g, k1, tset, SEED + 2)
    light = G.make_light(g, SEED + 3)
    mld = G.make_mld(g, SEED + 4, 0.0)
    return g, k1, tset, u, ts, light, mld, G.diffusivities(2000.0, 2.7e-5)


def wet_state_hash(wet_state: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(wet_state, dtype="<f8").tobytes()).hexdigest()


def main():
    import oracle_ffi as O

    g, k1, tset, u, ts, light, mld, diff = inputs()
    og = O.grid(36, 36, 16, kappa_h=2000.0, kappa_v=2.7e-5)
    gold = O.GoldsteinOracle(og, k1, ts, u)
    bio = O.BiogemOracle(gold, O.bgc_params(1.0 / 96, tset, carb_tol=1e-12, carb_maxit=1000), light, mld)
    w = G.wet_mask(g, k1)
    hashes, invs, done, t0 = [], [], 0, time.time()
    for y in CHECKPOINTS:
        for _ in range(y - done):
            bio.step(96)
        done = y
        hashes.append(wet_state_hash(gold.ts[w]))
        invs.append(gold.inventory())
        print(f"oracle year {y} done ({time.time() - t0:.0f} s)", flush=True)
    np.savez_compressed(OUT, years=np.array(CHECKPOINTS), state=gold.ts[w], sha256=np.array(hashes),
                        inventory=np.array(invs), carb_fail=np.array([bio.st.carb_fail]))
    print("wrote", OUT)


if __name__ == "__main__":
    main()
