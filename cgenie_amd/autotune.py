"""Tile-shape autotuner for the flux kernel.

For a batch of members (grid, tracer count, precision, member count) the
autotuner times every compiled tile shape of the flux kernel -- tracers per
block, wet columns per block, level planes resident in LDS, isoneutral
coefficients computed in every tracer block or once by a pre-pass -- with HIP
events on the engine's stream (cgo_time_kernel) and keeps the fastest.  The
flux kernel only reads the current state and writes the other ping-pong
buffer, so timing it repeatedly leaves the model state unchanged.

Results are cached per (grid, L, precision, member-count bucket) in
cgenie_amd/tuning.json so a serving process tunes each shape once.
"""
from __future__ import annotations

import json
import os

from . import engine as E

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "tuning.json")
TRACER_BLOCKS = (2, 4, 6, 8)
FLUX_COLS = (128, 256)
FLUX_RINGS = (2, 3)
ISO_PREPASS = (0, 1)


def _key(grid_dims, L, precision, n_members) -> str:
    bucket = 1
    while bucket < n_members:
        bucket *= 4
    return f"{grid_dims[0]}x{grid_dims[1]}x{grid_dims[2]}/L{L}/p{precision}/n{bucket}"


def load_cache() -> dict:
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(c: dict) -> None:
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as f:
        json.dump(c, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


def tune(eng: E.Engine, members, grid_dims, L: int, precision: int, reps: int = 10,
         use_cache: bool = True, write_cache: bool = True) -> dict:
    """Pick the flux kernel's tile shape for this batch.

    Returns {"tracer_block": tb, "cols": cols, "ring": planes, "iso_prepass": 0|1,
    "times_ms": {"tbxcolsxplanesxprepass": ms}} and applies the choice to the engine."""
    key = _key(grid_dims, L, precision, len(members))
    cache = load_cache() if use_cache else {}
    if key in cache:
        c = cache[key]
        eng.set_tile(precision, int(c["tracer_block"]), int(c.get("cols", 128)), int(c.get("ring", 3)),
                     int(c.get("iso_prepass", 0)))
        return c
    times = {}
    for pre in ISO_PREPASS:
        for ring in FLUX_RINGS:
            for cols in FLUX_COLS:
                for tb in TRACER_BLOCKS:
                    eng.set_tile(precision, tb, cols, ring, pre)
                    eng.time_kernel(0, members, 2)  # warm
                    times[(tb, cols, ring, pre)] = eng.time_kernel(0, members, reps)
    best = min(times, key=times.get)
    eng.set_tile(precision, *best)
    res = {"tracer_block": best[0], "cols": best[1], "ring": best[2], "iso_prepass": best[3],
           "times_ms": {f"{k[0]}x{k[1]}x{k[2]}x{k[3]}": v for k, v in times.items()}}
    if write_cache:
        cache = load_cache()
        cache[key] = res
        try:
            save_cache(cache)
        except OSError:
            pass
    return res
