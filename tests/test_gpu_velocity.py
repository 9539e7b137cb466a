"""The velocity solve on the GPU (cgo_grid_set_momentum, cgo_set_wind)
against the oracle's (tests/test_velocity.py pins that one), through the C
ABI, on the reference's worbe2 world (36 x 36 x 8, one island).  Needs an
MI355X.

Bar: an FP64 member is the oracle's arithmetic, so its state, velocities and
psi match the oracle's coupled velocity + tracer step (og_mom_step, then
og_goldstein_step, with rho handed from one step to the next as tstepo hands
it to jbar) to 1e-12 relative per cell -- bit for bit in practice, the count
is printed.  An FP32 member solves its velocities in FP64 from its own rho by
the psi equation's dense inverse (one product for all of a grid's FP32
members) instead of the band substitution, so after its first step they equal
the FP64 member's to rounding (1e-12 relative; both start from the rho of the
uploaded state); its flux coefficients, formed on the device, equal the ones
cgo_set_velocity forms on the host from the same velocities."""
import os
import sys

import numpy as np
import pytest

from cgenie_amd import engine as E
from cgenie_amd import grid as G
import oracle_ffi as O

pytestmark = pytest.mark.gpu

MAXI, MAXJ, MAXK = 36, 36, 8
WORLD = os.path.join(os.path.dirname(__file__), "golden", "worbe2")
DIFF = G.diffusivities(2000.0, 2.7e-5)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = E.Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def world():
    return G.read_world(WORLD, MAXI, MAXJ)


def wind(scale=0.1, seed=1):
    """x stress varying with latitude and longitude, a random y stress (N m-2)."""
    rng = np.random.default_rng(seed)
    st = np.zeros((MAXJ, MAXI, 4))
    lat = (np.arange(MAXJ) + 0.5) / MAXJ
    lon = np.cos(2 * np.pi * np.arange(MAXI) / MAXI)
    st[..., 0] = scale * (np.sin(2 * np.pi * lat) + 0.5)[:, None] * (1 + 0.3 * lon)[None, :]
    st[..., 2] = scale * (np.sin(2 * np.pi * (lat + 0.5 / MAXJ)) + 0.5)[:, None] * (1 + 0.3 * lon)[None, :]
    st[..., 1] = rng.normal(0, 0.02, (MAXJ, MAXI))
    st[..., 3] = rng.normal(0, 0.02, (MAXJ, MAXI))
    return st


def state(k1, L=4, seed=3):
    """T stratified and varying in latitude and longitude, S anomalies, and
    L - 2 dyes; zero on land."""
    rng = np.random.default_rng(seed)
    ts = np.zeros((MAXK, MAXJ, MAXI, L))
    z = np.arange(MAXK)[:, None, None]
    jj = np.arange(MAXJ)[None, :, None]
    ii = np.arange(MAXI)[None, None, :]
    ts[..., 0] = 2 + 2 * z + 3 * np.cos(np.pi * jj / MAXJ) + 0.5 * np.sin(2 * np.pi * ii / MAXI) + \
        0.05 * rng.standard_normal((MAXK, MAXJ, MAXI))
    ts[..., 1] = 0.1 * rng.standard_normal((MAXK, MAXJ, MAXI))
    for l in range(2, L):
        ts[..., l] = rng.uniform(0.5, 1.5, (MAXK, MAXJ, MAXI))
    wet = np.arange(1, MAXK + 1)[:, None, None] >= k1[None, 1:-1, 1:-1]
    ts[~wet] = 0.0
    return ts


def wet_mask(k1):
    return np.broadcast_to((np.arange(1, MAXK + 1)[:, None, None] >= k1[None, 1:-1, 1:-1])[..., None],
                           (MAXK, MAXJ, MAXI, 1))[..., 0]


def new_member(eng, world, prec, L=4, seed=3, st=None):
    g = G.make_grid(MAXI, MAXJ, MAXK)
    gid = eng.add_grid(g, world.k1)
    eng.set_momentum(gid, g, world)
    m = eng.add_member(gid, L, prec, DIFF, diso=True)
    ts = state(world.k1, L, seed)
    eng.upload(m, ts)
    eng.set_wind(m, wind() if st is None else st)
    return g, gid, m, ts


class CoupledOracle:
    """og_mom_step then og_goldstein_step per ocean step, the order of
    step_goldstein (goldstein.f90:198-235); rho starts as the eos of the
    initial state (initialise_goldstein 1450) and is then tstepo's."""

    def __init__(self, ts, st):
        k1, labels, isles, npi, paths = O.read_world()
        self.og = O.grid(MAXI, MAXJ, MAXK)
        self.mo = O.MomOracle(self.og, k1, labels, isles, npi, paths)
        self.mo.set_wind(st)
        self.u = np.zeros((MAXK, MAXJ + 1, MAXI + 1, 3))
        self.gold = O.GoldsteinOracle(self.og, k1, ts, self.u, diso=True)
        self.gold.rho[...] = O.rho_from_ts(self.og, k1, ts)

    def step(self, n=1):
        for _ in range(n):
            self.mo.step(self.gold.rho, self.u)
            self.gold.step(1)


def rel(a, b, mask=None):
    if mask is not None:
        a, b = a[mask], b[mask]
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


@pytest.mark.parametrize("nsteps", [1, 24])
def test_fp64_matches_oracle(eng, world, nsteps):
    say("fp64 velocity solve + tstepo, steps", nsteps)
    g, gid, m, ts = new_member(eng, world, E.FP64)
    orc = CoupledOracle(ts, wind())
    eng.step([m], nsteps)
    orc.step(nsteps)
    u_dev = eng.velocity(m, g)
    u_ref = orc.u[:, 1:, 1:, :]
    psi_dev, psi_ref = eng.psi(m, g), orc.mo.psi()
    got = eng.download(m, ts.shape)
    w = wet_mask(world.k1)
    say("  velocity rel err", rel(u_dev, u_ref), "bit-identical", np.count_nonzero(u_dev == u_ref), "of", u_ref.size)
    say("  psi rel err", rel(psi_dev, psi_ref), "bit-identical", np.count_nonzero(psi_dev == psi_ref), "of", psi_ref.size)
    errs = [rel(got[..., l], orc.gold.ts[..., l], w) for l in range(ts.shape[-1])]
    say("  state rel err per tracer", errs, "bit-identical cells", np.count_nonzero(got[w] == orc.gold.ts[w]),
        "of", got[w].size)
    assert np.abs(u_ref).max() > 0 and np.abs(psi_ref).max() > 0
    assert rel(u_dev, u_ref) <= 1e-12
    assert rel(psi_dev, psi_ref) <= 1e-12
    assert max(errs) <= 1e-12
    # land faces and levels below the bottom stay zero
    assert np.all(u_dev[..., 2][~w] == 0.0)


def test_fp32_first_step_velocities_equal_fp64(eng, world):
    say("fp32 member (dense solve): first-step velocities = fp64 member's (band solve)")
    g, _, m64, ts = new_member(eng, world, E.FP64)
    _, _, m32, _ = new_member(eng, world, E.FP32C)
    eng.step([m64, m32], 1)
    u64, u32 = eng.velocity(m64, g), eng.velocity(m32, g)
    p64, p32 = eng.psi(m64, g), eng.psi(m32, g)
    eu, ep = rel(u32, u64), rel(p32, p64)
    say("  velocity rel err", eu, "psi rel err", ep)
    assert eu <= 1e-12 and ep <= 1e-12 and np.abs(u64).max() > 0


def test_fp32_device_coefficients_equal_host(eng, world):
    """One step of an FP32 member with the solve against one step of an FP32
    member given, through cgo_set_velocity, the velocities that solve
    produces: the same flux coefficients give the same state bit for bit."""
    say("fp32 member: device flux coefficients = cgo_set_velocity's")
    g, gid, m_solve, ts = new_member(eng, world, E.FP32C)
    eng.step([m_solve], 1)
    # the velocities that step solved, as cgo_set_velocity takes them
    u = np.zeros((MAXK, MAXJ + 1, MAXI + 1, 3))
    u[:, 1:, 1:, :] = eng.velocity(m_solve, g)
    m_set = eng.add_member(gid, ts.shape[-1], E.FP32C, DIFF, diso=True)
    eng.upload(m_set, ts)
    eng.set_velocity(m_set, u)
    eng.step([m_set], 1)
    a, b = eng.download(m_solve, ts.shape), eng.download(m_set, ts.shape)
    say("  max |diff|", float(np.abs(a - b).max()))
    assert np.array_equal(a, b)


def test_fp32_tracks_oracle(eng, world):
    """24 steps of an FP32 member with its own velocity solve against the
    FP64 oracle: the FP32 tracer step's rounding feeds back through rho into
    the velocities (measured: 2.3e-8 state, 3.8e-8 velocity); the bar, 1e-6,
    is the FP32 form's own."""
    say("fp32 member with the solve, 24 steps vs the fp64 oracle")
    g, _, m, ts = new_member(eng, world, E.FP32C)
    orc = CoupledOracle(ts, wind())
    eng.step([m], 24)
    orc.step(24)
    w = wet_mask(world.k1)
    got = eng.download(m, ts.shape)
    errs = [rel(got[..., l], orc.gold.ts[..., l], w) for l in range(ts.shape[-1])]
    eu = rel(eng.velocity(m, g), orc.u[:, 1:, 1:, :])
    say("  state rel err per tracer", errs, "velocity rel err", eu)
    assert max(errs) <= 1e-6 and eu <= 1e-6


def test_graph_replay_equals_eager(eng, world):
    say("captured steps with the solve = eager steps")
    g, _, ma, ts = new_member(eng, world, E.FP64)
    _, _, mb, _ = new_member(eng, world, E.FP64)
    gr = eng.capture([ma], 4)
    for _ in range(3):
        eng.launch(gr, 1)
        eng.sync()
    eng.step([mb], 12)
    assert np.array_equal(eng.download(ma, ts.shape), eng.download(mb, ts.shape))
    assert np.array_equal(eng.velocity(ma, g), eng.velocity(mb, g))
    eng.drop_graph(gr)


def test_fp32_dense_graph_replay_equals_eager(eng, world):
    """The dense solve's staging, product and slice sums under replay: three
    FP32 members of one grid captured (one product for the three), and three
    more of the same grid stepped eagerly."""
    say("captured fp32 steps (dense solve) = eager steps")
    g, gid, ma, ts = new_member(eng, world, E.FP32C)

    def more():
        m = eng.add_member(gid, ts.shape[-1], E.FP32C, DIFF, diso=True)
        eng.upload(m, ts)
        eng.set_wind(m, wind())
        return m

    mas = [ma, more(), more()]
    mbs = [more() for _ in range(3)]
    gr = eng.capture(mas, 4)
    for _ in range(3):
        eng.launch(gr, 1)
        eng.sync()
    eng.step(mbs, 12)
    for a, b in zip(mas, mbs):
        assert np.array_equal(eng.download(a, ts.shape), eng.download(b, ts.shape))
        assert np.array_equal(eng.velocity(a, g), eng.velocity(b, g))
    eng.drop_graph(gr)


def test_checkpoint_keeps_the_relaxation_memory(eng, world):
    say("checkpoint of a member with the solve continues bit for bit")
    g, gid, ma, ts = new_member(eng, world, E.FP64)
    eng.step([ma], 3)
    blob = eng.checkpoint(ma)
    eng.step([ma], 3)
    mb = eng.add_member(gid, ts.shape[-1], E.FP64, DIFF, diso=True)
    eng.restore(mb, blob)
    eng.step([mb], 3)
    assert np.array_equal(eng.download(ma, ts.shape), eng.download(mb, ts.shape))
    assert np.array_equal(eng.velocity(ma, g), eng.velocity(mb, g))


def test_set_velocity_switches_the_solve_off(eng, world):
    say("cgo_set_velocity after cgo_set_wind: prescribed velocities again")
    g, gid, m, ts = new_member(eng, world, E.FP64)
    eng.step([m], 2)
    u = np.zeros((MAXK, MAXJ + 1, MAXI + 1, 3))
    eng.set_velocity(m, u)
    eng.upload(m, ts)
    eng.step([m], 2)
    with pytest.raises(E.CgoError):
        eng.psi(m, g)
    assert np.abs(eng.velocity(m, g)).max() == 0.0
