"""Continuous-batching service for ensemble runs of the ocean tracer step.

Requests (a grid, a tracer count, a run length in model years, a precision
form and a seed) arrive over time.  The server keeps every admitted run
resident on the GPU and steps all of them together: each round is one
batched call (one launch per kernel for the whole mixed set of grids, tracer
counts and precisions).  Between rounds it retires finished runs, frees
their memory and admits waiting requests while the HBM budget allows, so
new work joins runs already on the device instead of queueing behind them.

A run's model year is `nyear` ocean steps of its grid (48 at 18x18x8, 96
otherwise, as the reference's job configuration picks).  The server records
the wall time at which each model year of each run completes; the per-year
latency is the time between consecutive year completions of a run.

Each round is replayed from a HIP graph of the round's steps over the active
set, captured when the set (or the round length) changes, so a steady set of
runs is stepped with one graph launch per round instead of one launch per
kernel per step.

A run can be suspended between rounds (``suspend``): its exact checkpoint
(cgo_checkpoint) is kept on the host, its device memory is freed and it
waits to be admitted again like a new request; on re-admission it is
restored (cgo_restore) and continues to the bits it would have reached
uninterrupted.

Runs with at least seven tracers carry BIOGEM (T, S, DIC, ALK, PO4, O2,
DOM_P, then dyes) and run the full step (tstepo, step_biogem,
biogem_tracercoupling); smaller runs are T, S and dyes on GOLDSTEIN alone.
"""
from __future__ import annotations

import heapq
import math
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import engine as E
from . import grid as G

GRIDS = ((18, 18, 8), (36, 36, 16), (72, 72, 32))


def steps_per_year(dims) -> int:
    return 48 if dims == (18, 18, 8) else 96


@dataclass
class Request:
    rid: int
    dims: tuple
    L: int
    years: int
    precision: int
    seed: int
    arrival: float = 0.0          # seconds after the server starts
    # filled in by the server
    member: int = -1
    nbytes: int = 0
    steps_done: int = 0
    admitted: float = -1.0
    finished: float = -1.0
    year_done: list = field(default_factory=list)
    result: np.ndarray | None = None
    inventory: np.ndarray | None = None
    checkpoint: bytes | None = None   # held on the host while suspended
    suspensions: int = 0

    @property
    def nyear(self) -> int:
        return steps_per_year(self.dims)

    @property
    def total_steps(self) -> int:
        return self.years * self.nyear


class GridCache:
    """One engine grid (mask, metrics) and base velocity per grid shape."""

    def __init__(self, eng: E.Engine, seed: int = 2024):
        self.eng, self.seed, self.g = eng, seed, {}

    def get(self, dims):
        if dims not in self.g:
            g = G.make_grid(*dims, nyear=steps_per_year(dims))
            k1 = G.make_k1(g, self.seed)
            u = G.make_velocity(g, k1, self.seed + 1)
            gid = self.eng.add_grid(g, k1)
            self.g[dims] = (gid, g, k1, u)
        return self.g[dims]


def request_inputs(req: Request, g: G.Grid, k1: np.ndarray, base_u: np.ndarray):
    """Seeded inputs of one run: tracers, velocity scale, diffusivities, light."""
    rng = np.random.default_rng(req.seed)
    kind = "config2" if req.L >= 7 else "ts"
    tset = G.tracer_set(kind, req.L)
    ts = G.make_tracers(g, k1, tset, req.seed + 1)
    kh = rng.uniform(500.0, 5000.0)
    kv = math.exp(rng.uniform(math.log(1e-5), math.log(1e-4)))
    vs = rng.uniform(0.5, 1.5)
    light = G.make_light(g, req.seed + 2)
    mld = G.make_mld(g, req.seed + 3, 0.0)
    return tset, ts, base_u * vs, G.diffusivities(kh, kv), light, mld


class Server:
    def __init__(self, eng: E.Engine, hbm_budget: float = 200e9, round_steps: int = 48,
                 keep_results: bool = False, max_active: int | None = None):
        self.eng = eng
        self.budget = hbm_budget
        self.max_active = max_active  # at most this many runs on the GPU at once (None: memory decides)
        self.round_steps = round_steps
        self.keep = keep_results
        self.grids = GridCache(eng)
        self.waiting: list = []       # heap of (arrival, rid, Request)
        self.active: dict = {}        # member id -> Request
        self.done: list = []
        self.used = 0
        self.t0 = None
        self.rounds = 0
        # host-side inputs of arrived requests are prepared off the stepping loop
        self.pool = ThreadPoolExecutor(max_workers=2)
        self.prep: dict = {}
        self.graph, self.gkey = None, None   # the round graph and the (members, steps) it was captured for
        self.captures = 0
        self.t_admit = self.t_capture = self.t_gpu = 0.0   # wall seconds spent admitting, capturing, stepping
        self.t_calls: dict = {}       # wall seconds per engine call made while admitting and retiring
        self.on_round = None                 # called with the server after every round (a scheduling hook)

    # -- requests
    def submit(self, req: Request):
        heapq.heappush(self.waiting, (req.arrival, req.rid, req))

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def _bytes(self, req: Request) -> int:
        gid = self.grids.get(req.dims)[0]
        return self.eng.member_bytes(gid, req.L, req.precision, req.L >= 7)

    def _prepare(self, req: Request):
        if req.rid not in self.prep:
            gid, g, k1, base_u = self.grids.get(req.dims)
            self.prep[req.rid] = self.pool.submit(request_inputs, req, g, k1, base_u)
        return self.prep[req.rid]

    def _admit(self):
        t0 = time.perf_counter()
        try:
            self._admit_ready()
        finally:
            self.t_admit += time.perf_counter() - t0

    def _admit_ready(self):
        now = self.now()
        for _, _, r in self.waiting:
            if r.arrival <= now:
                self._prepare(r)
        while self.waiting and self.waiting[0][0] <= now:
            req = self.waiting[0][2]
            fut = self._prepare(req)
            if not fut.done():
                break                    # inputs still being prepared
            nb = self._bytes(req)
            if self.used + nb > self.budget:
                break                    # first come, first served: wait for memory
            if self.max_active is not None and len(self.active) >= self.max_active:
                break                    # the batch is full: wait for a run to finish
            heapq.heappop(self.waiting)
            gid = self.grids.get(req.dims)[0]
            tset, ts, u, diff, light, mld = self.prep[req.rid].result()
            m = self._call("add_member", self.eng.add_member, gid, req.L, req.precision, diff, diso=True,
                           carry=True)
            if req.checkpoint is not None:
                # a suspended run: its exact state, velocities and BIOGEM arrays included
                self._call("restore", self.eng.restore, m, req.checkpoint)
                req.checkpoint = None
            else:
                self._call("set_velocity", self.eng.set_velocity, m, u)
                self._call("upload", self.eng.upload, m, ts)
                if req.L >= 7:
                    p = E.default_bgc_params(1.0 / req.nyear, tset, carb_tol=1e-12, carb_maxit=1000)
                    self._call("bgc_init", self.eng.bgc_init, m, p, light, mld)
                req.admitted = self.now()
            req.member, req.nbytes = m, nb
            req._shape = ts.shape
            self.used += nb
            self.active[m] = req

    def _call(self, name, fn, *a, **kw):
        t0 = time.perf_counter()
        try:
            return fn(*a, **kw)
        finally:
            self.t_calls[name] = self.t_calls.get(name, 0.0) + time.perf_counter() - t0

    def suspend(self, req: Request):
        """Take an active run off the GPU between rounds: keep its exact
        checkpoint on the host, free its device memory and queue it again."""
        if req.member not in self.active:
            raise ValueError(f"request {req.rid} is not active")
        req.checkpoint = self.eng.checkpoint(req.member)
        self.eng.remove_member(req.member)
        self.used -= req.nbytes
        del self.active[req.member]
        req.member, req.suspensions = -1, req.suspensions + 1
        heapq.heappush(self.waiting, (self.now(), req.rid, req))

    def _retire(self, req: Request):
        if self.keep:
            req.result = self.eng.download(req.member, req._shape)
            req.inventory = self.eng.inventory(req.member)
        self._call("remove_member", self.eng.remove_member, req.member)
        self.used -= req.nbytes
        req.finished = self.now()
        del self.active[req.member]
        self.prep.pop(req.rid, None)
        self.done.append(req)

    def run(self, until_idle: bool = True, max_seconds: float | None = None, log=None):
        """Serve until every submitted request is done (or max_seconds pass)."""
        self.t0 = time.perf_counter() if self.t0 is None else self.t0
        while True:
            if not self.active:
                self._admit()
            if not self.active:
                if not self.waiting or (max_seconds is not None and self.now() > max_seconds):
                    break
                time.sleep(min(max(self.waiting[0][0] - self.now(), 0.002), 0.05))
                continue
            if max_seconds is not None and self.now() > max_seconds:
                break
            # one round: every active run advances by up to round_steps
            n = min(self.round_steps, min(r.total_steps - r.steps_done for r in self.active.values()))
            ids = list(self.active.keys())
            self._round(ids, n)
            # waiting runs are set up while the GPU steps this round (in the
            # room the last round's finished runs left): the engine waits for
            # its stream only for members with work in flight, and a new
            # member has none.  They join the next round.
            self._admit()
            t1 = time.perf_counter()
            self.eng.sync()
            self.t_gpu += time.perf_counter() - t1
            t = self.now()
            self.rounds += 1
            for r in [self.active[m] for m in ids]:
                before = r.steps_done // r.nyear
                r.steps_done += n
                for _ in range(r.steps_done // r.nyear - before):
                    r.year_done.append(t)
                if r.steps_done >= r.total_steps:
                    self._retire(r)
            if self.on_round is not None:
                self.on_round(self)
            if log and self.rounds % 20 == 0:
                log(f"t={t:.1f}s active={len(self.active)} waiting={len(self.waiting)} done={len(self.done)} "
                    f"hbm={self.used / 1e9:.2f} GB")
        self.close()
        return self.stats()

    def _capture(self, ids, n):
        t0 = time.perf_counter()
        try:
            self._capture_now(ids, n)
        finally:
            self.t_capture += time.perf_counter() - t0

    def _capture_now(self, ids, n):
        if self.graph is not None:
            self.eng.drop_graph(self.graph)
            self.graph = None
        self.graph, self.gkey = self.eng.capture(ids, n), (tuple(ids), n)
        self.captures += 1

    def _round(self, ids, n):
        """n ocean steps of every active run, from the round graph."""
        if self.gkey != (tuple(ids), n):
            self._capture(ids, n)
        try:
            self.eng.launch(self.graph, 1)
        except E.CgoError:
            # the same slots now hold other runs (a retired run's slot reused):
            # the engine refuses the old graph, so capture the set again
            self._capture(ids, n)
            self.eng.launch(self.graph, 1)

    def close(self):
        if self.graph is not None:
            self.eng.drop_graph(self.graph)
            self.graph, self.gkey = None, None

    def stats(self) -> dict:
        lat, member_years, weighted = [], 0.0, 0.0
        ref_work = 36 * 36 * 16 * 24
        for r in self.done + list(self.active.values()):
            prev = r.admitted
            for t in r.year_done:
                lat.append(t - prev)
                prev = t
            member_years += r.steps_done / r.nyear
            # member-years in units of a 36x36x16, 24-tracer member-year:
            # weighted by cells x tracers x steps per model year
            weighted += r.steps_done * (r.dims[0] * r.dims[1] * r.dims[2] * r.L) / (ref_work * 96)
        wall = self.now()
        return dict(wall_s=wall, member_years=member_years, member_years_per_s=member_years / max(wall, 1e-9),
                    equiv_36x36x16x24_member_years_per_s=weighted / max(wall, 1e-9), graph_captures=self.captures,
                    admit_s=self.t_admit, capture_s=self.t_capture, gpu_wait_s=self.t_gpu,
                    call_s={k: round(v, 3) for k, v in self.t_calls.items()},
                    requests_done=len(self.done), requests_active=len(self.active),
                    requests_waiting=len(self.waiting), rounds=self.rounds,
                    p50_year_s=float(np.percentile(lat, 50)) if lat else None,
                    p99_year_s=float(np.percentile(lat, 99)) if lat else None)


def poisson_requests(n: int, rate: float, seed: int, max_years: int = 1000, min_years: int = 10,
                     grids=GRIDS, L_range=(2, 40), precisions=None) -> list:
    """A seeded request stream: Poisson arrivals, grids uniform over `grids`,
    tracer counts uniform in L_range, run lengths log-uniform in
    [min_years, max_years], precision FP32C or FP64 with equal odds (or
    uniform over `precisions`)."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for q in range(n):
        t += rng.exponential(1.0 / rate)
        dims = grids[rng.integers(len(grids))]
        L = int(rng.integers(L_range[0], L_range[1] + 1))
        years = int(round(math.exp(rng.uniform(math.log(min_years), math.log(max_years)))))
        if precisions is None:
            prec = E.FP32C if rng.random() < 0.5 else E.FP64
        else:
            prec = precisions[int(rng.integers(len(precisions)))]
        out.append(Request(q, dims, L, years, prec, seed * 1000 + q, arrival=t))
    return out
