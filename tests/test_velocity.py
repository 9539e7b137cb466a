"""The velocity solve of the oracle pinned (CPU): GOLDSTEIN's barotropic
streamfunction with islands and its baroclinic velocities (goldstein.f90:
invert 3138-3204, jbar 3217-3315, ubarsolv 3500-3565, velc 3568-3679, wind
3685-3714; goldstein_lib.f90: island 186-241), on the reference's own world
data (worbe2: tests/golden/worbe2.k1, .psiles, .paths, the files
goldstein.f90:1113-1120, 1500-1575 read) and on a flat periodic channel whose
solution the mathematics fixes.

goldstein.f90 needs the netCDF Fortran module, absent here, so it cannot be
compiled in place; these properties pin the restatement instead:

* channel, zonal wind, uniform density: after the island correction the
  barotropic flow is the friction-wind balance drag * u = tau / h at every
  u point, with no meridional flow and a zonally uniform psi;
* the same channel, relaxed to a steady state: the top layer carries the
  frictional Ekman transport (r tau_x + f tau_y) / (f^2 + r^2) and
  (r tau_y - f tau_x) / (f^2 + r^2) relative to the layer below, and every
  column's depth-mean velocity is the barotropic one;
* worbe2 with wind and a stratified, horizontally varying density: the
  circulation round the island vanishes after the correction, psi's
  transports are non-divergent, the solve is linear in the forcing, and the
  three-dimensional velocity leaves a uniform tracer uniform under the
  oracle's own tracer step (its w is the continuity-consistent one);
* a density linear in depth: the bottom pressure is the exact hydrostatic
  integral (the trapezoid rule is exact for a linear integrand);
* the island path integral itself against the reference's own island()
  (goldstein_lib.f90 needs no other module and is compiled in place into
  oracle/_ref), on worbe2's path and on a path over u and v points.
"""
import numpy as np
import pytest

from cgenie_amd import grid as G
import oracle_ffi as O

MAXI, MAXJ, MAXK = 36, 36, 8


def channel(jp=3):
    """A flat periodic channel: every interior column 8 levels deep, land rows
    at j = 0 and maxj + 1; landmass 1 the northern wall, landmass 2 (the
    island) the southern one, whose path runs along the u points of row jp."""
    k1 = np.ones((MAXJ + 2, MAXI + 2), dtype=np.int32)
    k1[0, :] = 94
    k1[MAXJ + 1, :] = 94
    labels = np.zeros((MAXJ + 1, MAXI), dtype=np.int32)
    labels[0, :] = 2
    labels[MAXJ, :] = 1
    npi = np.array([MAXI], dtype=np.int32)
    paths = np.array([[1, i, jp] for i in range(1, MAXI + 1)], dtype=np.int32).ravel()
    return k1, labels, 1, npi, paths


def zonal_wind(scale=0.1):
    """x stress at u and v points varying with latitude, no y stress (N m-2)."""
    st = np.zeros((MAXJ, MAXI, 4))
    lat = (np.arange(MAXJ) + 0.5) / MAXJ
    st[..., 0] = scale * (np.sin(2 * np.pi * lat) + 0.5)[:, None]
    st[..., 2] = scale * (np.sin(2 * np.pi * (lat + 0.5 / MAXJ)) + 0.5)[:, None]
    return st


def state(k1, seed=3):
    rng = np.random.default_rng(seed)
    g = G.make_grid(MAXI, MAXJ, MAXK)
    ts = np.zeros((MAXK, MAXJ, MAXI, 2))
    z = np.arange(MAXK)[:, None, None]
    jj = np.arange(MAXJ)[None, :, None]
    ii = np.arange(MAXI)[None, None, :]
    ts[..., 0] = 2 + 2 * z + 3 * np.cos(np.pi * jj / MAXJ) + 0.5 * np.sin(2 * np.pi * ii / MAXI) + \
        0.05 * rng.standard_normal((MAXK, MAXJ, MAXI))
    ts[..., 1] = 0.1 * rng.standard_normal((MAXK, MAXJ, MAXI))
    wet = np.arange(1, MAXK + 1)[:, None, None] >= k1[None, 1:-1, 1:-1]
    ts[~wet] = 0.0
    return g, ts


def test_channel_barotropic_flow_is_the_friction_wind_balance():
    k1, labels, isles, npi, paths = channel()
    og = O.grid(MAXI, MAXJ, MAXK)
    m = O.MomOracle(og, k1, labels, isles, npi, paths)
    m.set_wind(zonal_wind())
    rho = np.zeros((MAXK + 1, MAXJ + 2, MAXI + 2))
    u = np.zeros((MAXK, MAXJ + 1, MAXI + 1, 3))
    m.step(rho, u)
    ub = m.ub()
    tau = np.ctypeslib.as_array(m.mo.tau, shape=(MAXJ, MAXI, 2))
    rh = m.field("rh")
    drag = m.field("drag")
    want = tau[:, :, 0] * rh[1:MAXJ + 1, 1:MAXI + 1, 0] / drag[1:MAXJ + 1, 1:MAXI + 1, 0]
    got = ub[1:MAXJ + 1, 1:MAXI + 1, 0]
    assert np.abs(got - want).max() <= 1e-9 * np.abs(want).max()
    assert np.abs(ub[:, 1:MAXI + 1, 1]).max() <= 1e-9 * np.abs(want).max()
    psi = m.psi()
    assert np.abs(psi - psi[:, :1]).max() <= 1e-9 * np.abs(psi).max()
    # a net transport between the walls: the island carries a psi of its own
    assert abs(psi[0, 1] - psi[MAXJ, 1]) > 1e-3 * np.abs(psi).max()


def test_channel_ekman_layer_and_depth_mean():
    k1, labels, isles, npi, paths = channel()
    og = O.grid(MAXI, MAXJ, MAXK)
    m = O.MomOracle(og, k1, labels, isles, npi, paths, rel=0.9)
    st = zonal_wind()
    st[..., 1] = 0.03      # y stress at u points
    st[..., 3] = 0.03      # and at v points
    m.set_wind(st)
    rho = np.zeros((MAXK + 1, MAXJ + 2, MAXI + 2))
    u = np.zeros((MAXK, MAXJ + 1, MAXI + 1, 3))
    for _ in range(300):   # the relaxation memory decays as 0.9^n
        m.step(rho, u)
    g = G.make_grid(MAXI, MAXJ, MAXK)
    drag = m.field("drag")
    dztau = np.ctypeslib.as_array(m.mo.dztau, shape=(MAXJ, MAXI, 2))
    dztav = np.ctypeslib.as_array(m.mo.dztav, shape=(MAXJ, MAXI, 2))
    dzz = g.dz[MAXK] * g.dza[MAXK - 1] / 2
    j = np.arange(1, MAXJ + 1)
    # u points: f = s(j), stress (tx, ty) = dztau * dzz
    f, r = g.s[j][:, None], drag[1:MAXJ + 1, 1:MAXI + 1, 0]
    tx, ty = dztau[..., 0] * dzz, dztau[..., 1] * dzz
    ek_u = (r * tx + f * ty) / (f * f + r * r)
    got_u = g.dz[MAXK] * (u[MAXK - 1, 1:, 1:, 0] - u[MAXK - 2, 1:, 1:, 0])
    assert np.abs(got_u - ek_u).max() <= 1e-8 * np.abs(ek_u).max()
    # v points (rows 1..maxj-1, the northern face of the last row is a wall): f = sv(j)
    jv = np.arange(1, MAXJ)
    fv, rv = g.sv[jv][:, None], drag[1:MAXJ, 1:MAXI + 1, 1]
    txv, tyv = dztav[:MAXJ - 1, :, 0] * dzz, dztav[:MAXJ - 1, :, 1] * dzz
    ek_v = (rv * tyv - fv * txv) / (fv * fv + rv * rv)
    got_v = g.dz[MAXK] * (u[MAXK - 1, 1:MAXJ, 1:, 1] - u[MAXK - 2, 1:MAXJ, 1:, 1])
    assert np.abs(got_v - ek_v).max() <= 1e-8 * np.abs(ek_v).max()
    # depth mean = barotropic velocity (flat bottom: H = sum dz)
    H = g.dz[1:MAXK + 1].sum()
    mean_u = np.tensordot(g.dz[1:MAXK + 1], u[..., 0], axes=(0, 0))[1:, 1:] / H
    ub = m.ub()
    assert np.abs(mean_u - ub[1:MAXJ + 1, 1:MAXI + 1, 0]).max() <= 1e-8 * np.abs(ub[..., 0]).max()


def _worbe2_step(stress, ts, seed=0, steps=1):
    k1, labels, isles, npi, paths = O.read_world()
    og = O.grid(MAXI, MAXJ, MAXK)
    m = O.MomOracle(og, k1, labels, isles, npi, paths)
    m.set_wind(stress)
    rho = O.rho_from_ts(og, k1, ts)
    u = np.zeros((MAXK, MAXJ + 1, MAXI + 1, 3))
    for _ in range(steps):
        m.step(rho, u)
    return m, u, k1, og


def test_worbe2_island_circulation_vanishes():
    k1, *_ = O.read_world()
    g, ts = state(k1)
    st = zonal_wind()
    st[..., 1] = np.random.default_rng(1).normal(0, 0.02, (MAXJ, MAXI))
    m, u, k1, og = _worbe2_step(st, ts)
    # the circulation round the island before the correction (erisl(1, 2)) and after it
    before = np.ctypeslib.as_array(m.mo.erisl, shape=(2,))[1]
    assert abs(before) > 0 and abs(m.island(1)) <= 1e-10 * abs(before)
    psi = m.psi()
    assert np.abs(psi).max() > 0
    # psi's transports are non-divergent: the net flux through every cell's
    # faces (h u across dsv, h v across dphi) cancels
    h = m.field("h")
    ub = m.ub()
    gg = G.make_grid(MAXI, MAXJ, MAXK)
    div = np.zeros((MAXJ, MAXI))
    for j in range(1, MAXJ + 1):
        for i in range(1, MAXI + 1):
            U = h[j, i, 0] * ub[j, i, 0] - h[j, i - 1, 0] * ub[j, i - 1, 0]
            V = gg.cv[j] * h[j, i, 1] * ub[j, i, 1] - gg.cv[j - 1] * h[j - 1, i, 1] * ub[j - 1, i, 1]
            div[j - 1, i - 1] = U * gg.rdphi * gg.rc[j] + V * gg.rds[j]
    wet = k1[1:-1, 1:-1] <= MAXK
    assert np.abs(div[wet]).max() <= 1e-10 * np.abs(ub).max()


def test_worbe2_solve_is_linear_in_the_forcing():
    k1, *_ = O.read_world()
    g, ts = state(k1)
    a, b = zonal_wind(0.1), zonal_wind(0.1)
    b[..., 0] = np.roll(b[..., 0], 5, axis=1) * 0.5
    ma, ua, *_ = _worbe2_step(a, np.zeros_like(ts))
    mb, ub_, *_ = _worbe2_step(b, np.zeros_like(ts))
    mab, uab, *_ = _worbe2_step(a + b, np.zeros_like(ts))
    assert np.abs(mab.ub() - ma.ub() - mb.ub()).max() <= 1e-12 * np.abs(mab.ub()).max()
    assert np.abs(uab - ua - ub_).max() <= 1e-12 * np.abs(uab).max()


def test_worbe2_velocity_leaves_a_uniform_tracer_uniform():
    """The solved (u, v, w) in the oracle's own tracer step: a uniform tracer
    stays uniform (to round-off) only if w closes every cell's volume budget
    in the tracer scheme's discretisation."""
    k1, *_ = O.read_world()
    g, ts = state(k1)
    m, u, k1, og = _worbe2_step(zonal_wind(), ts, steps=3)
    assert np.abs(u[..., 2]).max() > 0
    c = np.zeros((MAXK, MAXJ, MAXI, 3))
    c[..., :2] = ts
    wet = np.arange(1, MAXK + 1)[:, None, None] >= k1[None, 1:-1, 1:-1]
    c[..., 2] = np.where(wet, 1.0, 0.0)
    gold = O.GoldsteinOracle(og, k1, c, u, diso=False)
    gold.step(4)
    assert np.abs(gold.ts[..., 2][wet] - 1.0).max() <= 1e-12


def test_bad_island_paths_are_refused():
    k1, labels, isles, npi, paths = O.read_world()
    og = O.grid(MAXI, MAXJ, MAXK)
    bad = paths.copy()
    bad[2] = 0          # jpi = 0: a land row
    with pytest.raises(ValueError):
        O.MomOracle(og, k1, labels, isles, npi, bad)


def test_worbe2_no_wind_and_level_density_give_no_flow():
    """Density that varies with depth only and no wind: the hydrostatic
    pressure is level at every depth, so there is no torque and no flow --
    jbar's bottom-pressure terms of columns of different depths cancel."""
    k1, *_ = O.read_world()
    ts = np.zeros((MAXK, MAXJ, MAXI, 2))
    ts[..., 0] = (20.0 - 2.0 * np.arange(MAXK))[:, None, None]
    ts[..., 1] = (0.3 * np.arange(MAXK))[:, None, None]
    wet = np.arange(1, MAXK + 1)[:, None, None] >= k1[None, 1:-1, 1:-1]
    ts[~wet] = 0.0
    m, u, k1, og = _worbe2_step(np.zeros((MAXJ, MAXI, 4)), ts, steps=2)
    # the scale a 1 kg m-3 horizontal density difference would drive
    m2, u2, *_ = _worbe2_step(np.zeros((MAXJ, MAXI, 4)), state(k1)[1], steps=2)
    scale = np.abs(u2).max()
    assert scale > 0
    assert np.abs(u).max() <= 1e-12 * scale and np.abs(m.psi()).max() <= 1e-12 * np.abs(m2.psi()).max()


def test_drag_values():
    """drgset (goldstein.f90:2845-2885, adrag 2.5 days, drgf 3): the interior
    drag is 1 / (2.5 days x 2 Omega), nondimensional, and the equatorial band
    and the shallow-water rim take it times drgf^2 = 9."""
    k1, labels, isles, npi, paths = O.read_world()
    og = O.grid(MAXI, MAXJ, MAXK)
    m = O.MomOracle(og, k1, labels, isles, npi, paths)
    d = m.field("drag")[1:MAXJ + 1, 1:MAXI + 1, :]
    base = 1.0 / (2.5 * 86400 * 2 * 7.2921e-5)
    assert abs(d.min() - base) <= 1e-14 * base
    assert abs(d.max() - 9 * base) <= 1e-14 * base
    # the rows next to the equator are at the maximum everywhere
    assert np.allclose(d[MAXJ // 2, :, 0], 9 * base, rtol=1e-14, atol=0)


def test_bottom_pressure_is_the_hydrostatic_integral():
    """jbar's bp (goldstein.f90:3230-3236): bp(k) = -integral of rho over z
    from the column's bottom cell centre zro(k1) up to zro(k).  With rho =
    a + b z the trapezoid rule is exact: -(a dz + b (z^2 - z1^2) / 2)."""
    k1, *_ = O.read_world()
    g = G.make_grid(MAXI, MAXJ, MAXK)
    og = O.grid(MAXI, MAXJ, MAXK)
    a, b = 1.3, 0.7
    rho = np.zeros((MAXK + 1, MAXJ + 2, MAXI + 2))
    for k in range(1, MAXK + 1):
        rho[k] = a + b * g.zro[k]
    m, _, _, _ = _worbe2_step(np.zeros((MAXJ, MAXI, 4)), np.zeros((MAXK, MAXJ, MAXI, 2)))
    u = np.zeros((MAXK, MAXJ + 1, MAXI + 1, 3))
    m.step(rho, u)
    bp = np.ctypeslib.as_array(m.mo.bp, shape=(MAXK + 1, MAXJ + 2, MAXI + 2))
    z = np.array([g.zro[k] for k in range(MAXK + 1)])
    worst, n = 0.0, 0
    for j in range(1, MAXJ + 1):
        for i in range(1, MAXI + 1):
            kb = k1[j, i]
            for k in range(kb, MAXK + 1):
                want = -(a * (z[k] - z[kb]) + b * (z[k] ** 2 - z[kb] ** 2) / 2)
                worst = max(worst, abs(bp[k, j, i] - want))
                n += 1
    assert n > 1000 and worst <= 1e-13


def _ref_island_or_skip():
    if O.ref_gold() is None:
        pytest.skip("the reference's goldstein_lib is not built (oracle/_ref)")


def test_island_integral_matches_the_reference_worbe2():
    """The oracle's island() against goldstein_lib.f90's on worbe2 after a
    step with wind and density (its own ub, bp, sbp, tau), on the island
    basis velocities (indj = 0) and on a random ub."""
    _ref_island_or_skip()
    k1, *_ = O.read_world()
    g, ts = state(k1)
    st = zonal_wind()
    st[..., 1] = np.random.default_rng(1).normal(0, 0.02, (MAXJ, MAXI))
    m, u, k1, og = _worbe2_step(st, ts)
    ub = m.ub()
    got, want = m.island_of(ub, 1, 1), O.ref_island(m, 1, 1)
    assert got == want
    basis = np.ctypeslib.as_array(m.mo.ubisl, shape=(MAXJ + 1, MAXI + 2, 2))
    assert m.island_of(basis, 1, 0) == O.ref_island(m, 1, 0, basis)
    rnd = np.random.default_rng(5).standard_normal(ub.shape)
    for indj in (0, 1):
        assert m.island_of(rnd, 1, indj) == O.ref_island(m, 1, indj, rnd)


def test_island_integral_matches_the_reference_u_and_v_path():
    """A path over u and v points of both orientations (island() sums along
    whatever path it is given; this one need not be closed), random ub."""
    _ref_island_or_skip()
    k1, labels, isles, npi, _ = O.read_world()
    rng = np.random.default_rng(11)
    pts = []
    while len(pts) < 40:
        l = int(rng.choice([1, -1, 2, -2]))
        i, j = int(rng.integers(1, MAXI + 1)), int(rng.integers(1, MAXJ))
        if k1[j, i] <= MAXK and (abs(l) == 1 or j + 1 <= MAXJ):
            pts.append((l, i, j))
    paths = np.array(pts, dtype=np.int32).ravel()
    og = O.grid(MAXI, MAXJ, MAXK)
    m = O.MomOracle(og, k1, labels, isles, np.array([len(pts)], dtype=np.int32), paths)
    m.set_wind(zonal_wind() + 0.01)
    g, ts = state(k1)
    m.step(O.rho_from_ts(og, k1, ts), np.zeros((MAXK, MAXJ + 1, MAXI + 1, 3)))
    rnd = rng.standard_normal(m.ub().shape)
    for indj in (0, 1):
        assert m.island_of(rnd, 1, indj) == O.ref_island(m, 1, indj, rnd)
