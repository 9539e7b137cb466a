"""Benchmark of the MI355X ocean tracer step (GOLDSTEIN tstepo + BIOGEM +
tracer coupling) in member-years per second.

Workload (one line of JSON on stdout): an ensemble of 256 members on the
36x36x16 grid with 24 tracers (T, S, DIC, ALK, PO4, O2, DOM_P, 17 dyes),
seeded per-member parameters (kappa_h U[500,5000] m2/s, kappa_v
log-U[1e-5,1e-4] m2/s, velocity scale U[0.5,1.5], POC e-folding depth
U[300,900] m), BIOGEM every ocean step, FP32-compensated advection and
diffusion with FP64 convection, chemistry and sums.  One "step" is one ocean
step of the whole ensemble; a model year is 96 steps.  The run is captured
as a HIP graph of one model year and replayed.

Also reported: the same ensemble in FP64, the plain FP64 single-member path,
p50/p99 model-year latency, the roofline of the dominant kernel (flux) and
the CPU oracle timed on the host cores.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

STEPS_PER_YEAR = 96
HBM_PEAK_GBS = 8000.0


def ensemble_inputs(n_members: int, seed: int = 2024, kind: str = "config2"):
    from cgenie_amd import grid as G

    g = G.make_grid(36, 36, 16, nyear=STEPS_PER_YEAR)
    k1 = G.make_k1(g, seed)
    tset = G.tracer_set(kind)
    base_u = G.make_velocity(g, k1, seed + 1)
    ts0 = G.make_tracers(g, k1, tset, seed + 2)
    light = G.make_light(g, seed + 3)
    mld = G.make_mld(g, seed + 4, 0.0)
    rng = np.random.default_rng(seed + 5)
    members = []
    for q in range(n_members):
        kh = rng.uniform(500.0, 5000.0)
        kv = math.exp(rng.uniform(math.log(1e-5), math.log(1e-4)))
        vs = rng.uniform(0.5, 1.5)
        eL1 = rng.uniform(300.0, 900.0)
        members.append(dict(diff=G.diffusivities(kh, kv), u=base_u * vs, eL1=eL1))
    return g, k1, tset, ts0, light, mld, members


def algorithmic_bytes_per_member_step(n_wet: int, n_col: int, L: int, s_state: int, s_vel: int,
                                      s_rho: int = 8) -> int:
    """Bytes one ocean step of one member must move at least: read and write
    every tracer of every wet cell once (ping-pong, no copy), read u, v and w
    and write rho once per wet cell, and read the surface-flux slot of every
    tracer of every wet column."""
    return n_wet * (2 * L * s_state + 3 * s_vel + s_rho) + n_col * L * s_state


def build_ensemble(eng, E, n, precision, carry=True, seed=2024):
    g, k1, tset, ts0, light, mld, mem = ensemble_inputs(n, seed)
    gid = eng.add_grid(g, k1)
    ids = []
    for p in mem:
        m = eng.add_member(gid, tset.L, precision, p["diff"], diso=True, carry=carry)
        eng.set_velocity(m, p["u"])
        eng.upload(m, ts0)
        bp = E.default_bgc_params(1.0 / STEPS_PER_YEAR, tset, carb_tol=1e-12, carb_maxit=1000, POC_eL1=p["eL1"])
        eng.bgc_init(m, bp, light, mld)
        ids.append(m)
    wet = (np.arange(1, 17)[:, None, None]) >= k1[None, 1:37, 1:37]
    n_wet = int(wet.sum())
    n_col = int((k1[1:37, 1:37] <= 16).sum())
    return ids, tset, n_wet, n_col, g


def replay_steps(eng, year_graph, rest_graph, nsteps):
    """Launch exactly nsteps ocean steps: whole years on the one-year graph,
    the remainder on a graph of that many steps."""
    q, r = divmod(nsteps, STEPS_PER_YEAR)
    if q:
        eng.launch(year_graph, q)
    if r:
        eng.launch(rest_graph, 1)


def timed_years(eng, graph, years):
    """Replay the one-year graph `years` times, one year at a time; per-year
    wall seconds (each year ends with a synchronisation)."""
    out = []
    for _ in range(years):
        t0 = time.perf_counter()
        eng.launch(graph, 1)
        eng.sync()
        out.append(time.perf_counter() - t0)
    return out


def run_probe(args):
    """Counter runs (rocprofv3 --pmc): `args.steps` eager steps of the ensemble
    at the given tile, then a 1 GiB fill and a 1 GiB device-to-device copy
    whose FETCH_SIZE / WRITE_SIZE readings calibrate the byte counters."""
    import torch

    from cgenie_amd import engine as E

    prec = E.FP32C if args.probe == "fp32c" else E.FP64
    eng = E.Engine(0)
    ids, tset, n_wet, n_col, g = build_ensemble(eng, E, args.members, prec)
    eng.set_tile(prec, args.probe_tb, args.probe_cols, args.probe_ring, args.probe_iso)
    eng.step(ids, args.steps)
    eng.sync()
    print(f"# probe: {args.members} members x {args.steps} steps done", file=sys.stderr, flush=True)
    x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB written
    y = torch.empty_like(x)
    y.copy_(x)                                                    # 1 GiB read, 1 GiB written
    torch.cuda.synchronize()
    print("# probe: calibration fill and copy done", file=sys.stderr, flush=True)
    eng.close()


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "counters_fp32c_36x36x16x24.txt")


def measured_traffic(shape: str):
    """HBM bytes per flux launch at tile `shape` ("tb x cols x ring") from the
    committed counter summary (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs
    of `bench.py --probe fp32c`, scaled by the copy calibration), or None when
    the summary holds no figure for that tile."""
    try:
        for line in open(TRAFFIC_SUMMARY):
            if line.startswith(f"bytes_per_launch {shape} ="):
                return int(line.split("=")[1].split()[0])
    except OSError:
        return None
    return None


def profile_figure(path: str, key: str):
    """A figure `key = value` from a committed rocprofv3 summary, or None."""
    try:
        for line in open(path):
            if line.startswith(key + " ="):
                return float(line.split("=")[1].split()[0])
    except OSError:
        return None
    return None


BGC_FLOPS_SUMMARY = os.path.join(ROOT, "profiles", "fp64_counters_k_bgc_col_36x36x16x24.txt")


def run_gpu(args):
    from cgenie_amd import engine as E
    from cgenie_amd import autotune

    eng = E.Engine(0)
    res = {}
    n = args.members
    K, W = args.steps, args.warmup
    tuned = {}
    for prec, tag in ((E.FP32C, "fp32c"), (E.FP64, "fp64")):
        ids, tset, n_wet, n_col, g = build_ensemble(eng, E, n, prec)
        tuned[tag] = autotune.tune(eng, ids, (36, 36, 16), tset.L, prec, use_cache=False, write_cache=False)
        print(f"# autotune {tag}: {tuned[tag]}", file=sys.stderr, flush=True)
        gy = eng.capture(ids, STEPS_PER_YEAR)
        rest = {r: eng.capture(ids, r) for r in {K % STEPS_PER_YEAR, W % STEPS_PER_YEAR} if r}
        replay_steps(eng, gy, rest.get(W % STEPS_PER_YEAR), W)       # W untimed warm-up steps
        eng.sync()
        t0 = time.perf_counter()
        replay_steps(eng, gy, rest.get(K % STEPS_PER_YEAR), K)       # exactly K timed steps
        eng.sync()
        tot = time.perf_counter() - t0
        per_year = timed_years(eng, gy, args.latency_years)
        print(f"# {tag}: {K} steps in {tot:.3f} s; {args.latency_years} replayed years", file=sys.stderr, flush=True)
        flux_ms = eng.time_kernel(0, ids, 20)
        flux_step_ms = eng.time_kernel(3, ids, 20)  # beside the BIOGEM stream, as in the step
        co_ms = eng.time_kernel(1, ids, 20)
        bgc_cpl_ms = eng.time_kernel(2, ids, 5)
        d = eng.diag(ids[0])
        carb_fail = int(d.carb_fail)
        bgc_ms = eng.time_kernel(4, ids, 5)         # last: leaves the BIOGEM arrays for timing only
        s_state = 8 if prec == E.FP64 else 4
        s_vel = 8 if prec == E.FP64 else 4
        B = algorithmic_bytes_per_member_step(n_wet, n_col, tset.L, s_state, s_vel) * n
        Bc = algorithmic_bytes_per_member_step(n_wet, n_col, tset.L, s_state + (2 if prec == E.FP32C else 0), s_vel) * n
        res[tag] = dict(member_years_per_s=n * K / STEPS_PER_YEAR / tot, ms_per_step=1e3 * tot / K,
                        p50_year_s=float(np.percentile(per_year, 50)), p99_year_s=float(np.percentile(per_year, 99)),
                        latency_years=len(per_year), latency_member_years_per_s=n * len(per_year) / sum(per_year),
                        flux_ms=flux_ms, flux_step_ms=flux_step_ms, co_ms=co_ms, bgc_coupling_ms=bgc_cpl_ms,
                        bgc_ms=bgc_ms, flux_bytes=B, flux_bytes_with_carry=Bc,
                        flux_gbs=B / (flux_step_ms * 1e-3) / 1e9, flux_alone_gbs=B / (flux_ms * 1e-3) / 1e9,
                        carb_fail=carb_fail, n_wet=n_wet, n_col=n_col, L=tset.L,
                        tracer_block=tuned[tag]["tracer_block"], flux_cols=tuned[tag]["cols"],
                        flux_ring=tuned[tag]["ring"], flux_iso_prepass=tuned[tag]["iso_prepass"])
        eng.drop_graph(gy)
        for gr in rest.values():
            eng.drop_graph(gr)
        for m in ids:
            eng.remove_member(m)
    # plain FP64 single member, eager launches (the reference-shaped path:
    # a host calling the step once per ocean step)
    eng.set_tile(E.FP64, tuned["fp64"]["tracer_block"], tuned["fp64"]["cols"], tuned["fp64"]["ring"],
                 tuned["fp64"]["iso_prepass"])
    ids, tset, n_wet, n_col, g = build_ensemble(eng, E, 1, E.FP64)
    eng.step(ids, STEPS_PER_YEAR)
    eng.sync()
    ny = 3
    t0 = time.perf_counter()
    eng.step(ids, STEPS_PER_YEAR * ny)
    eng.sync()
    dt = time.perf_counter() - t0
    res["fp64_single"] = dict(member_years_per_s=ny / dt, ms_per_step=1e3 * dt / (ny * STEPS_PER_YEAR))
    for m in ids:
        eng.remove_member(m)
    eng.close()
    res["tuned_fp32c"] = tuned["fp32c"]
    return res


def run_prognostic(args, tuned):
    """The fully prognostic ocean: every member on the reference's worbe2
    world (36x36x8) with its own wind stress, EMBM's surface fluxes (sea
    ice, runoff, CO2 radiative forcing from the member's own ATCHEM
    atmosphere), the velocity solve, transport, BIOGEM with air-sea exchange
    and the tracer coupling; FP32C, replayed one-year graphs.  Each member
    has its own diffusivities, remineralisation length and air temperature."""
    from cgenie_amd import engine as E
    from cgenie_amd import grid as G
    from cgenie_amd import surface as SF

    here = os.path.dirname(os.path.abspath(__file__))
    world = G.read_world(os.path.join(here, "tests", "golden", "worbe2"), 36, 36)
    stress = np.load(os.path.join(here, "tests", "golden", "worbe2_wind.npz"))["stress"]
    g = G.make_grid(36, 36, 8, nyear=STEPS_PER_YEAR)
    tset = G.tracer_set("config2")
    ts0 = G.make_tracers(g, world.k1, tset, 11)
    light, mld = G.make_light(g, 12), G.make_mld(g, 13, 0.0)
    su = SF.surflux_setup(g, world.k1, radfor_atchem=1)
    atemp, ashum, sich, sica, tice = SF.atmosphere(g, world.k1, 14, ice=True)
    eng = E.Engine(0)
    eng.set_tile(E.FP32C, tuned["tracer_block"], tuned["cols"], tuned["ring"], tuned["iso_prepass"])
    gid = eng.add_grid(g, world.k1)
    eng.set_momentum(gid, g, world)
    eng.set_surflux(gid, su)
    rng = np.random.default_rng(15)
    ids = []
    for q in range(args.members):
        kh = rng.uniform(500.0, 5000.0)
        kv = math.exp(rng.uniform(math.log(1e-5), math.log(1e-4)))
        m = eng.add_member(gid, tset.L, E.FP32C, G.diffusivities(kh, kv), diso=True)
        eng.upload(m, ts0)
        bp = E.default_bgc_params(1.0 / STEPS_PER_YEAR, tset, carb_tol=1e-12, carb_maxit=1000,
                                  POC_eL1=rng.uniform(300.0, 900.0), airsea=1, atm_pCO2=278e-6)
        eng.bgc_init(m, bp, light, mld)
        eng.set_windspeed(m, np.full((36, 36), 8.0))
        eng.set_wind(m, stress)
        eng.set_surface(m, atemp + rng.normal(0.0, 1.0), ashum, sich, sica, tice)
        ids.append(m)
    gy = eng.capture(ids, STEPS_PER_YEAR)
    eng.launch(gy, 1)
    eng.sync()
    years = max(1, args.steps // STEPS_PER_YEAR)
    t0 = time.perf_counter()
    eng.launch(gy, years)
    eng.sync()
    tot = time.perf_counter() - t0
    out = dict(member_years_per_s=round(len(ids) * years / tot, 2),
               ms_per_step=round(1e3 * tot / (years * STEPS_PER_YEAR), 4), members=len(ids), years=years,
               grid="36x36x8 worbe2", tracers=tset.L,
               kernels_ms=dict(surflux=round(eng.time_kernel(6, ids, 20), 4), velocity=round(eng.time_kernel(5, ids, 5), 4),
                               flux=round(eng.time_kernel(0, ids, 10), 4), co=round(eng.time_kernel(1, ids, 10), 4)))
    print(f"# prognostic: {out}", file=sys.stderr, flush=True)
    eng.drop_graph(gy)
    eng.close()
    return out


def run_serving(seconds: float, rate: float = 4.0, seed: int = 5, uniform: bool = False,
                max_active: int | None = None, round_steps: int = 48):
    """The serving form: a seeded Poisson stream (mean `rate` requests/s)
    admitted against a 200 GB HBM budget and stepped together, served for
    `seconds` of wall time.  Mixed: runs on 18x18x8, 36x36x16 and 72x72x32
    with 2-40 tracers, 10-1000 model years, FP32C or FP64.  Uniform (the
    like-for-like stream): every run 36x36x16 with 24 tracers in FP32C.  At
    4 requests/s of runs averaging ~215 years the stream offers ~860
    member-years/s, below what the GPU serves; the saturated leg offers more
    (a faster stream, at most `max_active` runs on the GPU at once) and so
    measures the serving form's own throughput."""
    from cgenie_amd import engine as E
    from cgenie_amd import serving as S

    eng = E.Engine(0)
    srv = S.Server(eng, hbm_budget=200e9, round_steps=round_steps, max_active=max_active)
    kw = dict(grids=((36, 36, 16),), L_range=(24, 24), precisions=(E.FP32C,)) if uniform else {}
    for r in S.poisson_requests(int(rate * seconds * 2) + 8, rate, seed, **kw):
        srv.submit(r)
    st = srv.run(max_seconds=seconds, log=lambda m: print("# serve " + m, file=sys.stderr, flush=True))
    eng.close()
    return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in st.items()}


def _cpu_worker(args):
    """One serial member of the CPU oracle's -O3 build, nsteps ocean steps
    with BIOGEM every step; wall seconds."""
    q, nsteps = args
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ctypes as C

    import oracle_ffi as O
    from cgenie_amd import grid as G

    O._lib = O._bind(C.CDLL(os.path.join(ROOT, "oracle", "liboracle_o3.so")))
    g, k1, tset, ts0, light, mld, mem = ensemble_inputs(1, 2024 + q)
    og = O.grid(36, 36, 16, nyear=STEPS_PER_YEAR)
    og.diff[0], og.diff[1] = mem[0]["diff"]
    gold = O.GoldsteinOracle(og, k1, ts0, mem[0]["u"])
    bio = O.BiogemOracle(gold, O.bgc_params(1.0 / STEPS_PER_YEAR, tset, POC_eL1=mem[0]["eL1"]), light, mld)
    t0 = time.perf_counter()
    bio.step(nsteps)
    return time.perf_counter() - t0


def _compiler() -> str:
    import subprocess

    try:
        v = subprocess.run(["gcc", "-dumpfullversion"], capture_output=True, text=True, timeout=10).stdout.strip()
    except (OSError, subprocess.SubprocessError):
        v = "?"
    return f"gcc {v} -O3 -funroll-loops -ffp-contract=off (oracle/liboracle_o3.so)"


def run_cpu_baseline(cores: int, nsteps: int):
    """The oracle (a C restatement of the reference arithmetic, -O3), one
    serial member per core, each running nsteps ocean steps.  Runs before the
    GPU is touched."""
    import multiprocessing as mp

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as O

    O.build()
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle_o3.so")):
        import subprocess

        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "liboracle_o3.so"], check=True,
                       capture_output=True)
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        times = pool.map(_cpu_worker, [(q, nsteps) for q in range(cores)])
    wall = time.perf_counter() - t0
    member_years = cores * nsteps / STEPS_PER_YEAR
    per_member_year = [t * STEPS_PER_YEAR / nsteps for t in times]
    return dict(value=round(member_years / max(times), 3), unit="member-years/s", cores=cores, kind="port",
                nproc=os.cpu_count(), compiler=_compiler(),
                per_core_member_years_per_s=round(nsteps / STEPS_PER_YEAR / float(np.mean(times)), 4),
                p50_year_s=round(float(np.percentile(per_member_year, 50)), 4),
                p99_year_s=round(float(np.percentile(per_member_year, 99)), 4),
                sample=f"{cores} members x {nsteps} steps (36x36x16, 24 tracers, BIOGEM every step), one serial "
                       f"member per core on {cores} of the host's {os.cpu_count()} cores (the host cores one GPU's "
                       f"jobs may use); per-member {np.mean(times):.2f} s, wall {wall:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=960)
    ap.add_argument("--warmup", type=int, default=96)
    ap.add_argument("--latency-years", type=int, default=20,
                    help="model years replayed one at a time for the p50/p99 year latency (at least 10)")
    ap.add_argument("--members", type=int, default=256)
    ap.add_argument("--cpu-cores", type=int, default=int(os.environ.get("CGO_CPU_CORES", "16")))
    ap.add_argument("--cpu-steps", type=int, default=96)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-prognostic", action="store_true")
    ap.add_argument("--prognostic-only", action="store_true",
                    help="profile mode: the prognostic-ocean leg alone at the flux tile 4x256x2x1 (the autotuner's "
                         "usual pick), its figures on one line, no headline line")
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--probe", choices=("fp32c", "fp64"), default=None,
                    help="counter-run mode: eager steps of the ensemble and a copy calibration, no JSON line")
    ap.add_argument("--probe-tb", type=int, default=6)
    ap.add_argument("--probe-cols", type=int, default=256)
    ap.add_argument("--probe-ring", type=int, default=2)
    ap.add_argument("--probe-iso", type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              "(ensemble members shard across GPUs with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if args.probe:
        run_probe(args)
        return
    if args.prognostic_only:
        print(json.dumps(dict(prognostic=run_prognostic(args, dict(tracer_block=4, cols=256, ring=2,
                                                                   iso_prepass=1)))), flush=True)
        return
    args.latency_years = max(10, args.latency_years)
    cores = max(1, min(args.cpu_cores, os.cpu_count() or 1))
    cpu = None if args.no_cpu else run_cpu_baseline(cores, args.cpu_steps)
    print(f"# cpu baseline done: {cpu}", file=sys.stderr, flush=True)
    r = run_gpu(args)
    prognostic = run_prognostic(args, r["tuned_fp32c"]) if not args.no_prognostic else None
    serving = run_serving(args.serve_seconds) if args.serve_seconds > 0 else None
    serving_u = run_serving(args.serve_seconds, uniform=True) if args.serve_seconds > 0 else None
    serving_s = (run_serving(args.serve_seconds, rate=32.0, uniform=True, max_active=256)
                 if args.serve_seconds > 0 else None)
    main_ = r["fp32c"]
    single = r["fp64_single"]["member_years_per_s"]
    # the flux launch as it runs inside the step (the BIOGEM column step beside
    # it on the second stream); the same launch alone is reported next to it
    roof = dict(bound="hbm", achieved=round(main_["flux_gbs"], 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(main_["flux_gbs"] / HBM_PEAK_GBS, 4),
                traffic=measured_traffic(f'{main_["tracer_block"]}x{main_["flux_cols"]}x{main_["flux_ring"]}'
                                         f'x{main_["flux_iso_prepass"]}'),
                kernel=(f'k_flux<float, 2, {main_["flux_cols"]}, {main_["flux_ring"]}, 1> + ' if main_["flux_iso_prepass"]
                        else '') + f'k_flux<float, {main_["tracer_block"]}, {main_["flux_cols"]}, {main_["flux_ring"]}, '
                                   f'{2 if main_["flux_iso_prepass"] else 0}>',
                algorithmic_bytes=main_["flux_bytes"], algorithmic_bytes_with_carry=main_["flux_bytes_with_carry"],
                launch_ms=round(main_["flux_step_ms"], 4),
                alone_ms=round(main_["flux_ms"], 4), alone_achieved=round(main_["flux_alone_gbs"], 1))
    # the BIOGEM column kernel is FP64-VALU / latency bound: FP64 operations
    # per launch from the committed counter summary over its time alone
    bgc_roof = {}
    for tag in ("fp32c", "fp64"):
        fl = profile_figure(BGC_FLOPS_SUMMARY, f"fp64_flops_per_launch {tag}")
        bms = r[tag]["bgc_ms"]
        bgc_roof[tag] = dict(bound="fp64-valu", peak=78.6, unit="TFLOP/s", launch_ms=round(bms, 4),
                             flops=fl, achieved=None if fl is None else round(fl / (bms * 1e-3) / 1e12, 3),
                             frac=None if fl is None else round(fl / (bms * 1e-3) / 1e12 / 78.6, 4))
    line = dict(
        metric="member-years/s at 36x36x16x24 tracers (FP32-comp batched ensemble); p50/p99 year latency",
        value=round(main_["member_years_per_s"], 2), unit="member-years/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=round(main_["ms_per_step"], 4), higher_is_better=True, scaling="weak",
        vs_baseline=None, dtype="f32+carry/f64", data="synthetic",
        config=dict(workload=f"ensemble {args.members} members x 36x36x16 x 24 tracers, BIOGEM every step, "
                             "FP32-compensated advection-diffusion, FP64 co/chemistry/sums, replayed HIP graphs "
                             "(one model year, and the remainder of the step count)",
                    members=args.members, grid="36x36x16", tracers=24, steps_per_year=STEPS_PER_YEAR),
        p50_year_s=main_["p50_year_s"], p99_year_s=main_["p99_year_s"], latency_years=main_["latency_years"],
        fp64_batched_member_years_per_s=round(r["fp64"]["member_years_per_s"], 2),
        fp64_batched_p50_year_s=r["fp64"]["p50_year_s"], fp64_batched_p99_year_s=r["fp64"]["p99_year_s"],
        fp64_single_member_years_per_s=round(single, 3),
        speedup_vs_fp64_single=round(main_["member_years_per_s"] / single, 1),
        kernels_ms=dict(fp32c=dict(flux=main_["flux_ms"], flux_in_step=main_["flux_step_ms"], co=main_["co_ms"],
                                   bgc=main_["bgc_ms"], bgc_coupling=main_["bgc_coupling_ms"],
                                   tracer_block=main_["tracer_block"], flux_cols=main_["flux_cols"],
                                   flux_ring=main_["flux_ring"], flux_iso_prepass=main_["flux_iso_prepass"]),
                        fp64=dict(flux=r["fp64"]["flux_ms"], flux_in_step=r["fp64"]["flux_step_ms"],
                                  co=r["fp64"]["co_ms"], bgc=r["fp64"]["bgc_ms"],
                                  bgc_coupling=r["fp64"]["bgc_coupling_ms"], tracer_block=r["fp64"]["tracer_block"],
                                  flux_cols=r["fp64"]["flux_cols"], flux_ring=r["fp64"]["flux_ring"],
                                  flux_iso_prepass=r["fp64"]["flux_iso_prepass"])),
        roofline=roof, roofline_bgc=bgc_roof, cpu_baseline=cpu, prognostic=prognostic, serving=serving,
        serving_uniform=serving_u,
        serving_uniform_vs_fp64_single=(round(serving_u["member_years_per_s"] / single, 1) if serving_u else None),
        serving_saturated=serving_s,
        serving_saturated_vs_fp64_single=(round(serving_s["member_years_per_s"] / single, 1) if serving_s else None))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
