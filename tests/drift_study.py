"""Precision drift study (the 1000-model-year FP32 form vs FP64).

Config: 36x36x16, 24 tracers (T, S, DIC, ALK, PO4, O2, DOM_P, 17 dyes),
BIOGEM every step, kappa_h 2000 m2/s, kappa_v 2.7e-5 m2/s, seed 2024.

  python tests/drift_study.py oracle OUTDIR [--years N]
      the CPU oracle (FP64 restatement), states saved at the checkpoints
  python tests/drift_study.py gpu OUTDIR [--years N]
      FP64, FP32-compensated (carry) and plain FP32 members, and an FP64
      member started from a state perturbed by 1e-7 relative, stepped
      together on the GPU; states saved at the checkpoints
  python tests/drift_study.py compare DIR_ORACLE DIR_GPU
      per-tracer max relative error and inventory drift at each checkpoint

Relative error per cell uses |a-b| / max(|b|, 1e-3 max|b_l|) and, for S,
the absolute salinity (S + 34.9) since GOLDSTEIN stores the anomaly.
Test infrastructure: the oracle here is the checker.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tests")]

from cgenie_amd import grid as G  # noqa: E402

CHECKPOINTS = (1, 3, 10, 30, 100, 300, 1000)
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


def _save(outdir, tag, year, ts, extra=None):
    """Wet cells only (maxk, maxj, maxi, L)[wet] -> (N_wet, L)."""
    os.makedirs(outdir, exist_ok=True)
    g, k1 = G.make_grid(36, 36, 16), G.make_k1(G.make_grid(36, 36, 16), SEED)
    np.save(os.path.join(outdir, f"{tag}_y{year:04d}.npy"), ts[G.wet_mask(g, k1)])
    if extra:
        with open(os.path.join(outdir, f"{tag}_y{year:04d}.json"), "w") as f:
            json.dump(extra, f)


def run_oracle(outdir, years):
    import oracle_ffi as O

    g, k1, tset, u, ts, light, mld, diff = inputs()
    og = O.grid(36, 36, 16, kappa_h=2000.0, kappa_v=2.7e-5)
    gold = O.GoldsteinOracle(og, k1, ts, u)
    bio = O.BiogemOracle(gold, O.bgc_params(1.0 / 96, tset, carb_tol=1e-12, carb_maxit=1000), light, mld)
    done = 0
    t0 = time.time()
    for y in CHECKPOINTS:
        if y > years:
            break
        bio.step(96 * (y - done))
        done = y
        _save(outdir, "oracle", y, gold.ts, dict(inventory=list(gold.inventory()), wall_s=time.time() - t0))
        print(f"oracle year {y} done ({time.time() - t0:.0f} s)", flush=True)


def run_gpu(outdir, years):
    from cgenie_amd import engine as E

    g, k1, tset, u, ts, light, mld, diff = inputs()
    eng = E.Engine(0)
    gid = eng.add_grid(g, k1)
    p = E.default_bgc_params(1.0 / 96, tset, carb_tol=1e-12, carb_maxit=1000)
    ms = {}
    # fp64p: the FP64 path from a state perturbed by 1e-7 relative (seeded),
    # the model's own sensitivity to a perturbation of FP32 rounding size
    rng = np.random.default_rng(SEED + 9)
    ts_p = ts * (1.0 + 1e-7 * rng.standard_normal(ts.shape))
    for tag, prec, carry, t0s in (("fp64", E.FP64, True, ts), ("fp32c", E.FP32C, True, ts),
                                  ("fp32", E.FP32C, False, ts), ("fp64p", E.FP64, True, ts_p)):
        m = eng.add_member(gid, tset.L, prec, diff, diso=True, carry=carry)
        eng.set_velocity(m, u)
        eng.upload(m, t0s)
        eng.bgc_init(m, p, light, mld)
        ms[tag] = m
    graph = eng.capture(list(ms.values()), 96)
    done = 0
    t0 = time.time()
    for y in CHECKPOINTS:
        if y > years:
            break
        eng.launch(graph, y - done)
        eng.sync()
        done = y
        for tag, m in ms.items():
            d = eng.diag(m)
            _save(outdir, tag, y, eng.download(m, ts.shape),
                  dict(inventory=list(eng.inventory(m)), carb_fail=int(d.carb_fail), steps=int(d.steps)))
        print(f"gpu year {y} done ({time.time() - t0:.1f} s)", flush=True)
    eng.close()


def rel(a, b):
    """a, b: (N_wet, L) wet-cell states."""
    aw, bw = a.copy(), b.copy()
    aw[:, 1] += G.SALN0
    bw[:, 1] += G.SALN0
    scale = np.maximum(np.abs(bw).max(axis=0), 1e-300)
    den = np.maximum(np.abs(bw), 1e-3 * scale[None, :])
    return (np.abs(aw - bw) / den).max(axis=0)


def compare(dir_ref, dir_gpu, ref_tag="oracle"):
    g, k1, tset, u, ts, light, mld, diff = inputs()
    w = G.wet_mask(g, k1)
    out = {}
    for y in CHECKPOINTS:
        fr = os.path.join(dir_ref, f"{ref_tag}_y{y:04d}.npy")
        if not os.path.exists(fr):
            continue
        ref = np.load(fr)
        if ref.ndim == 4:
            ref = ref[w]
        iref = json.load(open(fr[:-4] + ".json"))["inventory"]
        row = {}
        for tag in ("fp64", "fp32c", "fp32", "fp64p"):
            f = os.path.join(dir_gpu, f"{tag}_y{y:04d}.npy")
            if not os.path.exists(f):
                continue
            a = np.load(f)
            inv = json.load(open(f[:-4] + ".json"))["inventory"]
            e = rel(a, ref)
            idr = [abs(x - y0) / max(abs(y0), 1e-300) for x, y0 in zip(inv, iref)]
            row[tag] = dict(max_rel=float(e.max()), per_tracer=[float(v) for v in e],
                            inventory_rel=[float(v) for v in idr])
        out[y] = row
        print(y, {k: f"{v['max_rel']:.3e}" for k, v in row.items()}, flush=True)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=("oracle", "gpu", "compare"))
    ap.add_argument("dirs", nargs="+")
    ap.add_argument("--years", type=int, default=1000)
    ap.add_argument("--ref-tag", default="oracle")
    a = ap.parse_args()
    if a.mode == "oracle":
        run_oracle(a.dirs[0], a.years)
    elif a.mode == "gpu":
        run_gpu(a.dirs[0], a.years)
    else:
        res = compare(a.dirs[0], a.dirs[1], a.ref_tag)
        if len(a.dirs) > 2:
            with open(a.dirs[2], "w") as f:
                json.dump(res, f, indent=1)
