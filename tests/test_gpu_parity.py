"""HIP engine vs the CPU oracle, through the C ABI (needs an MI355X).

Bar: the FP64 path within 1e-12 per cell, relative, with a floor of 1e-3 of
the tracer's largest magnitude for cells near zero (rel_err in cases.py);
land masks and indexing exact (land cells are never written); inventories
conserved to round-off.  The FP64 GOLDSTEIN step is also checked for bit
identity with the oracle (both round every operation as written)."""
import sys

import numpy as np
import pytest

from cgenie_amd import engine as E
from cgenie_amd import grid as G
import oracle_ffi as O
from cases import make_case, rel_err, wet

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = E.Engine(0)
    yield e
    e.close()


def new_member(eng, c, prec=E.FP64, carry=True, diso=True):
    gid = eng.add_grid(c.g, c.k1)
    m = eng.add_member(gid, c.tset.L, prec, c.diff, diso=diso, carry=carry)
    eng.set_velocity(m, c.u)
    eng.set_sflux(m, c.sflux)
    eng.upload(m, c.ts)
    return m


@pytest.mark.parametrize("dims,kind,L,random_u,sflux,diso,nsteps", [
    ((36, 36, 16), "ts", 2, False, False, True, 96),      # config 1 shape, one model year
    ((36, 36, 16), "ts", 5, True, True, True, 12),        # w, surface fluxes, odd tracer count
    ((18, 18, 8), "ts", 7, True, False, False, 12),       # isoneutral off
    ((72, 72, 32), "ts", 3, False, False, True, 3),       # the largest serving grid
])
def test_tstep_fp64_matches_oracle(eng, dims, kind, L, random_u, sflux, diso, nsteps):
    say("case", dims, kind, L, random_u, sflux, diso, nsteps)
    c = make_case(*dims, kind=kind, L=L, random_u=random_u, sflux=sflux)
    m = new_member(eng, c, diso=diso)
    gold = O.GoldsteinOracle(c.og, c.k1, c.ts, c.u, c.sflux, diso=diso)
    eng.step([m], nsteps)
    gold.step(nsteps)
    got = eng.download(m, c.ts.shape)
    w = wet(c)
    err = rel_err(got, gold.ts, w)
    nexact = np.count_nonzero(got[w] == gold.ts[w])
    say("  max rel err per tracer", err, "bit-identical cells", nexact, "of", got[w].size)
    assert np.all(err <= 1e-12)
    assert np.array_equal(got[~w], c.ts[~w])          # land untouched
    rho = eng.rho(m, c.g)
    rref = gold.rho[1:, 1:-1, 1:-1]
    assert np.abs(rho[w] - rref[w]).max() <= 1e-12 * np.abs(rref[w]).max()
    d = eng.diag(m)
    assert d.steps == nsteps
    assert abs(d.cost - gold.cost.sum()) < 0.5
    eng.remove_member(m)


def test_convection_of_many_tracers(eng):
    """100 tracers on 16 levels with surface fluxes and divergent flow:
    each of a convecting column's four lanes mixes 25 of its tracers, seven
    rounds of four loads issued together; the state stays the oracle's bit
    for bit and the filled levels are counted alike."""
    c = make_case(36, 36, 16, kind="ts", L=100, random_u=True, sflux=True)
    m = new_member(eng, c)
    gold = O.GoldsteinOracle(c.og, c.k1, c.ts, c.u, c.sflux)
    eng.step([m], 3)
    gold.step(3)
    got = eng.download(m, c.ts.shape)
    w = wet(c)
    diff = np.count_nonzero(got[w] != gold.ts[w])
    say("levels filled by convection", gold.cost.sum(), "cells differing", diff)
    assert gold.cost.sum() > 0
    assert diff == 0
    assert abs(eng.diag(m).cost - gold.cost.sum()) < 0.5
    eng.remove_member(m)


def test_tstep_fp64_bit_identical_to_oracle(eng):
    c = make_case(kind="ts", L=4, seed=5)
    m = new_member(eng, c)
    gold = O.GoldsteinOracle(c.og, c.k1, c.ts, c.u)
    eng.step([m], 8)
    gold.step(8)
    got = eng.download(m, c.ts.shape)
    w = wet(c)
    diff = np.count_nonzero(got[w] != gold.ts[w])
    say("cells differing from the oracle:", diff)
    assert diff == 0
    eng.remove_member(m)


def test_inventory_conserved_96_steps(eng):
    c = make_case(kind="ts", L=6, seed=9)
    m = new_member(eng, c)
    i0 = eng.inventory(m)
    gold = O.GoldsteinOracle(c.og, c.k1, c.ts, c.u)
    assert np.allclose(i0, gold.inventory(), rtol=1e-13, atol=0)
    eng.step([m], 96)
    i1 = eng.inventory(m)
    say("inventory rel change", (i1 - i0) / np.abs(i0))
    big = np.abs(i0) > 1e-3 * np.abs(i0).max()
    assert np.all(np.abs(i1 - i0)[big] <= 1e-12 * np.abs(i0)[big])
    eng.remove_member(m)


# (tracers per block, wet columns per block, level planes in LDS,
#  isoneutral coefficients from a pre-pass)
SHAPES = [(2, 128, 3, 0), (4, 256, 2, 1), (8, 128, 2, 1), (6, 256, 3, 0), (4, 128, 3, 1), (8, 256, 3, 0),
          (2, 256, 2, 1)]
DEFAULT_SHAPE = (8, 256, 2, 1)      # the engine's FP64 default
DEFAULT_SHAPE_32 = (6, 256, 2, 1)   # the engine's FP32-form default


@pytest.mark.parametrize("shape", SHAPES)
def test_tracer_block_shapes_agree(eng, shape):
    say("fp64 flux shape", shape)
    c = make_case(kind="ts", L=11, seed=21)
    eng.set_tile(E.FP64, *shape)
    m = new_member(eng, c)
    eng.step([m], 5)
    got = eng.download(m, c.ts.shape)
    eng.set_tile(E.FP64, *DEFAULT_SHAPE)
    gold = O.GoldsteinOracle(c.og, c.k1, c.ts, c.u)
    gold.step(5)
    w = wet(c)
    assert np.count_nonzero(got[w] != gold.ts[w]) == 0
    eng.remove_member(m)


@pytest.mark.parametrize("shape", [(6, 256, 2, 1), (4, 128, 3, 1), (8, 256, 2, 0)])
def test_tile_shapes_with_sflux_random_u_no_diso(eng, shape):
    """The isoneutral pre-pass and the tile shapes on a case with surface
    fluxes and a random (divergent) velocity field, with and without
    isoneutral diffusion: FP64 equal to the oracle bit for bit."""
    for diso in (True, False):
        say("fp64 flux shape", shape, "diso", diso)
        c = make_case(kind="ts", L=7, seed=27, random_u=True, sflux=True)
        eng.set_tile(E.FP64, *shape)
        m = new_member(eng, c, diso=diso)
        eng.step([m], 4)
        got = eng.download(m, c.ts.shape)
        eng.set_tile(E.FP64, *DEFAULT_SHAPE)
        gold = O.GoldsteinOracle(c.og, c.k1, c.ts, c.u, c.sflux, diso=diso)
        gold.step(4)
        w = wet(c)
        assert np.count_nonzero(got[w] != gold.ts[w]) == 0
        eng.remove_member(m)


def test_fp32_tile_shapes_identical(eng):
    """The FP32 form gives the same bits for every flux tile shape (each
    cell's arithmetic does not depend on how cells are grouped)."""
    c = make_case(kind="config2", seed=23)
    p = E.default_bgc_params(1.0 / 96, c.tset, carb_tol=1e-12, carb_maxit=1000)
    outs = []
    for shape in [DEFAULT_SHAPE_32] + SHAPES:
        say("fp32 flux shape", shape)
        eng.set_tile(E.FP32C, *shape)
        m = new_member(eng, c, prec=E.FP32C)
        eng.bgc_init(m, p, c.light, c.mld)
        eng.step([m], 4)
        outs.append(eng.download(m, c.ts.shape))
        eng.remove_member(m)
    eng.set_tile(E.FP32C, *DEFAULT_SHAPE_32)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


@pytest.mark.parametrize("kind,mld", [("config2", 0.0), ("domc", 150.0)])
def test_biogem_fp64_matches_oracle(eng, kind, mld):
    """GOLDSTEIN + BIOGEM + tracercoupling in the genie.f90 order, 24 tracers."""
    say("bgc case", kind, mld)
    c = make_case(kind=kind, seed=31, mld_max=mld)
    m = new_member(eng, c)
    p = E.default_bgc_params(1.0 / 96, c.tset, carb_tol=1e-12, carb_maxit=1000)
    eng.bgc_init(m, p, c.light, c.mld)
    gold = O.GoldsteinOracle(c.og, c.k1, c.ts, c.u)
    bio = O.BiogemOracle(gold, O.bgc_params(1.0 / 96, c.tset), c.light, c.mld)
    n = 12
    eng.step([m], n)
    bio.step(n)
    got = eng.download(m, c.ts.shape)
    w = wet(c)
    err = rel_err(got, gold.ts, w)
    say("  ts max rel err", err.max(), "per tracer", err[:8])
    assert np.all(err <= 1e-12)
    ocn = eng.bgc_download(m, 0, c.ts.shape)
    oref = bio.ocn()
    sl = [0, 1] + [q for q in (c.tset.DIC, c.tset.ALK, c.tset.PO4, c.tset.O2, c.tset.DOM_C, c.tset.DOM_P) if q >= 0]
    eo = rel_err(ocn[..., sl], oref[..., sl], w)
    say("  ocn max rel err", eo)
    assert np.all(eo <= 1e-12)
    part = eng.bgc_download(m, 1, (c.g.maxk, c.g.maxj, c.g.maxi, 5))
    pref = bio.part()
    scale = np.abs(pref).max(axis=(0, 1, 2)) + 1e-300
    ep = (np.abs(part - pref).max(axis=(0, 1, 2)) / scale)
    say("  part err", ep)
    assert np.all(ep <= 1e-12)
    carb = eng.bgc_download(m, 2, (c.g.maxj, c.g.maxi, 10))
    cref = bio.carb()
    wc = c.k1[1:-1, 1:-1] <= c.g.maxk
    ec = np.abs(carb[wc] - cref[wc]) / np.maximum(np.abs(cref[wc]), 1e-300)
    say("  carb max rel err", ec.max(axis=0))
    assert np.all(ec[:, :9] <= 1e-10)
    d = eng.diag(m)
    assert d.carb_fail == 0 and bio.st.carb_fail == 0
    eng.remove_member(m)


@pytest.mark.parametrize("sinkrate", [19200.0, 57600.0, 76800.0, 192000.0, 500000.0])
def test_biogem_stream_windows_match_oracle(eng, sinkrate):
    """Sinking speeds (m/yr) whose per-step distance on the 16-level grid
    selects each particle-stream window of the column march: 200, 600, 800
    and 2000 m cross 2, 5, 6 and 10 levels (windows 4, 6, 8 and 16), and
    5208 m leaves the column in one step (the single-source march)."""
    say("bgc sinkrate", sinkrate)
    c = make_case(kind="config2", seed=33)
    m = new_member(eng, c)
    p = E.default_bgc_params(1.0 / 96, c.tset, carb_tol=1e-12, carb_maxit=1000, sinkrate=sinkrate)
    eng.bgc_init(m, p, c.light, c.mld)
    gold = O.GoldsteinOracle(c.og, c.k1, c.ts, c.u)
    bio = O.BiogemOracle(gold, O.bgc_params(1.0 / 96, c.tset, sinkrate=sinkrate), c.light, c.mld)
    n = 6
    eng.step([m], n)
    bio.step(n)
    w = wet(c)
    err = rel_err(eng.download(m, c.ts.shape), gold.ts, w)
    say("  ts max rel err", err.max())
    assert np.all(err <= 1e-12)
    part = eng.bgc_download(m, 1, (c.g.maxk, c.g.maxj, c.g.maxi, 5))
    pref = bio.part()
    ep = np.abs(part - pref).max(axis=(0, 1, 2)) / (np.abs(pref).max(axis=(0, 1, 2)) + 1e-300)
    say("  part err", ep)
    assert np.all(ep <= 1e-12)
    eng.remove_member(m)


def test_batched_members_equal_alone(eng):
    """Members of different grids, tracer counts and parameters stepped in one
    launch per kernel give the bits each gives alone."""
    cases = [make_case(kind="config2", seed=41), make_case(18, 18, 8, kind="ts", L=9, seed=42, kappa_h=800.0),
             make_case(kind="domc", seed=43, kappa_v=8e-5, vscale=1.4)]
    ms, alone = [], []
    for q, c in enumerate(cases):
        m = new_member(eng, c)
        if q != 1:
            eng.bgc_init(m, E.default_bgc_params(1.0 / 96, c.tset, carb_tol=1e-12, carb_maxit=1000), c.light, c.mld)
        ms.append(m)
    eng.step(ms, 6)
    batched = [eng.download(m, c.ts.shape) for m, c in zip(ms, cases)]
    for m in ms:
        eng.remove_member(m)
    for q, c in enumerate(cases):
        m = new_member(eng, c)
        if q != 1:
            eng.bgc_init(m, E.default_bgc_params(1.0 / 96, c.tset, carb_tol=1e-12, carb_maxit=1000), c.light, c.mld)
        eng.step([m], 6)
        alone.append(eng.download(m, c.ts.shape))
        eng.remove_member(m)
    for a, b in zip(batched, alone):
        assert np.array_equal(a, b)


def test_graph_replay_equals_eager(eng):
    """A captured year replays to the eager result; replays are separated by
    eager work and synchronisation and every replay is checked."""
    c = make_case(kind="config2", seed=51)
    p = E.default_bgc_params(1.0 / 96, c.tset, carb_tol=1e-12, carb_maxit=1000)
    m1 = new_member(eng, c)
    eng.bgc_init(m1, p, c.light, c.mld)
    m2 = new_member(eng, c)
    eng.bgc_init(m2, p, c.light, c.mld)
    gid = eng.capture([m1], 4)
    for rep in range(3):
        eng.launch(gid, 1)
        eng.sync()
        eng.step([m2], 4)
        eng.sync()
        a = eng.download(m1, c.ts.shape)
        b = eng.download(m2, c.ts.shape)
        say("replay", rep, "identical:", np.array_equal(a, b))
        assert np.array_equal(a, b)
    eng.drop_graph(gid)
    eng.remove_member(m1)
    eng.remove_member(m2)


def test_graph_replay_refused_after_member_changes(eng):
    """A captured graph bakes in its members' launch extents, precision
    lists and stream windows.  Its replay is refused with CGO_E_STATE, before
    any kernel is enqueued, while a captured member is between bio_step and
    tracercoupling, and for good once a captured member was destroyed (also
    after an eager step has brought the member table up to date), its slot
    re-used by a new member, or re-initialised."""
    c = make_case(18, 18, 8, kind="config2", seed=31)
    p = E.default_bgc_params(1.0 / 96, c.tset, carb_tol=1e-12, carb_maxit=1000)
    a, b = new_member(eng, c), new_member(eng, c)
    for m in (a, b):
        eng.bgc_init(m, p, c.light, c.mld)
    g = eng.capture([a, b], 2)
    eng.launch(g, 1)
    eng.sync()
    eng.bio_step([a])
    with pytest.raises(E.CgoError, match="rc=-3"):
        eng.launch(g, 1)
    eng.tracercoupling([a])
    eng.launch(g, 1)
    eng.sync()

    def refused():
        eng.sync()
        before, steps = eng.download(a, c.ts.shape), eng.diag(a).steps
        with pytest.raises(E.CgoError, match="rc=-3"):
            eng.launch(g, 1)
        eng.sync()
        assert eng.diag(a).steps == steps
        assert np.array_equal(eng.download(a, c.ts.shape), before)

    eng.remove_member(b)
    eng.step([a], 1)          # an eager step syncs the member table
    refused()
    # the next members created fill the free slots, b's among them
    extra = []
    while True:
        extra.append(new_member(eng, c, prec=E.FP32C))
        if extra[-1] == b:
            break
        assert len(extra) < 4096
    refused()
    eng.drop_graph(g)
    g2 = eng.capture([a, b], 1)   # a fresh capture of the new member set replays
    eng.launch(g2, 1)
    eng.sync()
    eng.bgc_init(b, p, c.light, c.mld)
    with pytest.raises(E.CgoError, match="rc=-3"):
        eng.launch(g2, 1)
    eng.drop_graph(g2)
    for m in [a] + extra:
        eng.remove_member(m)


@pytest.mark.parametrize("carry", [True, False])
def test_fp32_form_tracks_fp64(eng, carry):
    c = make_case(kind="config2", seed=61)
    p = E.default_bgc_params(1.0 / 96, c.tset, carb_tol=1e-12, carb_maxit=1000)
    m64 = new_member(eng, c)
    eng.bgc_init(m64, p, c.light, c.mld)
    m32 = new_member(eng, c, prec=E.FP32C, carry=carry)
    eng.bgc_init(m32, p, c.light, c.mld)
    eng.step([m64, m32], 96)
    a = eng.download(m32, c.ts.shape)
    b = eng.download(m64, c.ts.shape)
    a[..., 1] += G.SALN0  # compare absolute salinity, not the anomaly GOLDSTEIN stores
    b[..., 1] += G.SALN0
    err = rel_err(a, b, wet(c))
    say("FP32 (carry=%s) vs FP64 after one year: max rel err" % carry, err.max(), err[:7])
    # the compensated form holds every tracer within 1e-6 of FP64; plain
    # FP32 (no carry) is reported for comparison at its own, looser bar
    assert np.all(err <= (1e-6 if carry else 5e-5))
    i32, i64 = eng.inventory(m32), eng.inventory(m64)
    big = np.abs(i64) > 1e-3 * np.abs(i64).max()
    say("  inventory rel diff", np.abs(i32 - i64)[big].max() / np.abs(i64)[big].max())
    eng.remove_member(m32)
    eng.remove_member(m64)


def test_fp32_form_carbonate_solve_reaches_the_same_root(eng):
    """FP32-form members solve the surface carbonate chemistry with the
    accelerated fixed point: same root as the reference iteration of the FP64
    member (to the solver tolerance; the inputs differ by the FP32 form's
    ~1e-12 rounding of the state), in fewer sweeps."""
    c = make_case(kind="config2", seed=65)
    p = E.default_bgc_params(1.0 / 96, c.tset, carb_tol=1e-12, carb_maxit=1000)
    m64 = new_member(eng, c)
    m32 = new_member(eng, c, prec=E.FP32C)
    for m in (m64, m32):
        eng.bgc_init(m, p, c.light, c.mld)
    for _ in range(3):
        eng.bio_step([m64, m32])
        eng.tracercoupling([m64, m32])
    shape = (c.g.maxj, c.g.maxi, 10)
    a, b = eng.bgc_download(m32, 2, shape), eng.bgc_download(m64, 2, shape)
    wc = G.wet_mask(c.g, c.k1)[-1]
    sl = [0, 1, 2, 3, 4, 5, 6]  # H, fCO2, CO2, CO3, HCO3, ohm calcite, ohm aragonite
    e = np.abs(a[wc][:, sl] - b[wc][:, sl]) / np.abs(b[wc][:, sl])
    d32, d64 = eng.diag(m32), eng.diag(m64)
    say("carbonate FP32-form vs FP64 max rel", e.max(axis=0), "sweeps", d32.carb_iters, d64.carb_iters)
    assert np.all(e <= 1e-9)
    assert d32.carb_fail == 0 and d64.carb_fail == 0
    assert d32.carb_iters < d64.carb_iters
    eng.remove_member(m32)
    eng.remove_member(m64)


def test_split_calls_equal_step(eng):
    """tstepo + bio_step + tracercoupling, called one by one in the genie.f90
    order, give the bits of cgo_step; the coupling is refused out of order."""
    c = make_case(kind="config2", seed=71)
    p = E.default_bgc_params(1.0 / 96, c.tset, carb_tol=1e-12, carb_maxit=1000)
    m1 = new_member(eng, c)
    eng.bgc_init(m1, p, c.light, c.mld)
    m2 = new_member(eng, c)
    eng.bgc_init(m2, p, c.light, c.mld)
    eng.step([m1], 3)
    for _ in range(3):
        eng.tstepo([m2])
        eng.bio_step([m2])
        eng.tracercoupling([m2])
    a = eng.download(m1, c.ts.shape)
    b = eng.download(m2, c.ts.shape)
    say("split == step:", np.array_equal(a, b))
    assert np.array_equal(a, b)
    with pytest.raises(E.CgoError):
        eng.tracercoupling([m2])          # no bio_step before it
    eng.bio_step([m2])
    with pytest.raises(E.CgoError):
        eng.tstepo([m2])                  # coupling still due
    with pytest.raises(E.CgoError):
        eng.step([m2], 1)
    eng.tracercoupling([m2])
    eng.tstepo([m2], 2)                   # transport alone is allowed again
    eng.remove_member(m1)
    eng.remove_member(m2)
