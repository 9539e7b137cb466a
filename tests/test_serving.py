"""The continuous-batching server (cgenie_amd/serving.py).

CPU: admission by HBM budget, slot recycling and per-year latency
bookkeeping, against a stand-in engine that only counts steps.
GPU: every run served in a changing mix of grids, tracer counts and
precisions gives the bits it gives when run alone."""
import numpy as np
import pytest

from cgenie_amd import engine as E
from cgenie_amd import serving as S


class CountingEngine:
    """Engine stand-in for the scheduler logic: members are step counters."""

    def __init__(self, bytes_per_member=100):
        self.bpm, self.members, self.nextm, self.calls = bytes_per_member, {}, 0, []
        self.graphs, self.captures, self.dropped = [], 0, 0

    def add_grid(self, g, k1):
        return 0

    def member_bytes(self, gid, L, prec, bgc):
        return self.bpm * L

    def add_member(self, gid, L, prec, diff, diso=True, carry=True):
        self.nextm += 1
        self.members[self.nextm] = 0
        return self.nextm

    def set_velocity(self, m, u):
        pass

    def upload(self, m, ts):
        pass

    def bgc_init(self, m, p, light, mld):
        pass

    def step(self, ids, n):
        self.calls.append((tuple(ids), n))
        for m in ids:
            self.members[m] += n

    def capture(self, ids, n):
        self.graphs.append((tuple(ids), n))
        self.captures += 1
        return len(self.graphs) - 1

    def launch(self, g, reps=1):
        ids, n = self.graphs[g]
        for _ in range(reps):
            self.step(ids, n)

    def drop_graph(self, g):
        self.dropped += 1

    def checkpoint(self, m):
        return self.members[m].to_bytes(8, "little")

    def restore(self, m, blob):
        self.members[m] = int.from_bytes(blob, "little")

    def sync(self):
        pass

    def remove_member(self, m):
        del self.members[m]


class SyncPool:
    """Runs submitted work at once (deterministic admission in the test)."""

    def submit(self, fn, *a):
        from concurrent.futures import Future
        f = Future()
        f.set_result(fn(*a))
        return f


def test_scheduler_admits_by_budget_and_recycles(monkeypatch):
    monkeypatch.setattr(S, "request_inputs", lambda req, g, k1, u: (None, np.zeros(1), None, (0, 0), None, None))
    monkeypatch.setattr(S.GridCache, "get", lambda self, dims: (0, None, None, None))
    eng = CountingEngine(bytes_per_member=100)
    srv = S.Server(eng, hbm_budget=1000, round_steps=48)
    srv.pool = SyncPool()
    # three runs of 4 tracers (400 B each): only two fit at once
    reqs = [S.Request(q, (36, 36, 16), 4, years=2 + q, precision=E.FP64, seed=q) for q in range(3)]
    for r in reqs:
        srv.submit(r)
    st = srv.run()
    assert st["requests_done"] == 3 and not eng.members
    first = eng.calls[0][0]
    assert len(first) == 2                       # the third waited for memory
    assert any(len(ids) == 2 and reqs[2].member in ids for ids, _ in eng.calls)  # joined a running one
    for r in reqs:
        assert r.steps_done == r.total_steps and len(r.year_done) == r.years
    assert st["member_years"] == sum(r.years for r in reqs)
    assert st["p50_year_s"] is not None
    # rounds replay a graph captured per (member set, round length); every graph is dropped
    assert eng.captures == len(set(eng.calls)) and eng.dropped == eng.captures
    assert st["equiv_36x36x16x24_member_years_per_s"] * 6 == pytest.approx(st["member_years_per_s"])


def test_batch_cap_limits_active_runs(monkeypatch):
    """max_active: with memory for all, at most that many runs share a round;
    the rest wait and join as runs finish, first come first served."""
    monkeypatch.setattr(S, "request_inputs", lambda req, g, k1, u: (None, np.zeros(1), None, (0, 0), None, None))
    monkeypatch.setattr(S.GridCache, "get", lambda self, dims: (0, None, None, None))
    eng = CountingEngine(bytes_per_member=1)
    srv = S.Server(eng, hbm_budget=1e9, round_steps=48, max_active=2)
    srv.pool = SyncPool()
    reqs = [S.Request(q, (36, 36, 16), 4, years=1 + q, precision=E.FP64, seed=q) for q in range(5)]
    for r in reqs:
        srv.submit(r)
    st = srv.run()
    assert st["requests_done"] == 5
    assert max(len(ids) for ids, _ in eng.calls) == 2
    assert [r.rid for r in sorted(reqs, key=lambda r: r.admitted)] == [0, 1, 2, 3, 4]


def test_suspended_run_resumes_where_it_stopped(monkeypatch):
    """A run suspended between rounds leaves the GPU (its memory is
    returned to the budget), waits like a new request and, restored from its
    checkpoint, finishes with exactly its own number of steps."""
    monkeypatch.setattr(S, "request_inputs", lambda req, g, k1, u: (None, np.zeros(1), None, (0, 0), None, None))
    monkeypatch.setattr(S.GridCache, "get", lambda self, dims: (0, None, None, None))
    eng = CountingEngine(bytes_per_member=100)
    srv = S.Server(eng, hbm_budget=1000, round_steps=48)
    srv.pool = SyncPool()
    reqs = [S.Request(q, (36, 36, 16), 4, years=3, precision=E.FP64, seed=q) for q in range(2)]
    for r in reqs:
        srv.submit(r)
    seen = []

    def hook(server):
        r = reqs[0]
        if r.member in server.active and r.suspensions == 0 and r.steps_done == 96:
            used = server.used
            server.suspend(r)
            seen.append((used - server.used, r.member, len(eng.members)))

    srv.on_round = hook
    st = srv.run()
    assert seen == [(400, -1, 1)]
    assert st["requests_done"] == 2 and not eng.members and srv.used == 0
    for r in reqs:
        assert r.steps_done == r.total_steps and len(r.year_done) == r.years
    assert reqs[0].suspensions == 1 and reqs[0].checkpoint is None


def test_poisson_stream_is_seeded():
    a = S.poisson_requests(20, 4.0, seed=3)
    b = S.poisson_requests(20, 4.0, seed=3)
    assert [(r.dims, r.L, r.years, r.precision, r.arrival) for r in a] == \
           [(r.dims, r.L, r.years, r.precision, r.arrival) for r in b]
    assert all(10 <= r.years <= 1000 and 2 <= r.L <= 40 for r in a)
    assert np.all(np.diff([r.arrival for r in a]) > 0)


@pytest.mark.gpu
def test_served_runs_equal_alone():
    eng = E.Engine(0)
    srv = S.Server(eng, hbm_budget=200e9, round_steps=48, keep_results=True)
    mix = [((18, 18, 8), 9, 2, E.FP64, 0.0), ((36, 36, 16), 24, 1, E.FP32C, 0.0), ((18, 18, 8), 3, 3, E.FP32C, 0.0),
           ((36, 36, 16), 12, 2, E.FP64, 0.3), ((18, 18, 8), 40, 1, E.FP64, 0.6), ((72, 72, 32), 5, 1, E.FP32C, 0.6)]
    reqs = [S.Request(q, d, L, y, p, 700 + q, arrival=t) for q, (d, L, y, p, t) in enumerate(mix)]
    for r in reqs:
        srv.submit(r)

    def suspend_once(server):
        # take two runs off the GPU part-way through; they resume from their checkpoints
        for r in (reqs[1], reqs[3]):
            if r.member in server.active and r.suspensions == 0 and r.steps_done >= 48:
                server.suspend(r)

    srv.on_round = suspend_once
    st = srv.run()
    print("served:", st, flush=True)
    assert st["requests_done"] == len(reqs)
    assert reqs[1].suspensions == 1 and reqs[3].suspensions == 1
    for r in reqs:
        gid, g, k1, base_u = srv.grids.get(r.dims)
        tset, ts, u, diff, light, mld = S.request_inputs(r, g, k1, base_u)
        m = eng.add_member(gid, r.L, r.precision, diff, diso=True, carry=True)
        eng.set_velocity(m, u)
        eng.upload(m, ts)
        if r.L >= 7:
            eng.bgc_init(m, E.default_bgc_params(1.0 / r.nyear, tset, carb_tol=1e-12, carb_maxit=1000), light, mld)
        eng.step([m], r.total_steps)
        alone = eng.download(m, ts.shape)
        eng.remove_member(m)
        same = np.array_equal(alone, r.result)
        print("request", r.rid, r.dims, "L", r.L, "years", r.years, "prec", r.precision, "alone == served:", same,
              flush=True)
        assert same
    eng.close()
