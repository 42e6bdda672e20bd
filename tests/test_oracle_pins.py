"""Pins of the oracle's viscosity, minimum-pressure, Laplacian / quasi-
compressibility and higher-order-source-term stages against what the
mathematics fixes, independently of the oracle's own expressions:

* viscosity (src/mps.cpp:177-203): the MPS Laplacian model is exact for
  quadratic fields on a lattice, so v_x = C (y - yc)^2 gets a = 2 nu C at
  every particle whose re_v neighbourhood is complete, in 2D and 3D; a pair
  of particles gives nu 2D/(n0_v lambda_v) (v_j - v_i) w(r, re_v) with the
  lattice sums of the viscosity radius; a uniform field gives exactly 0;
* minimum pressure (src/mps.cpp:481-503): brute force over all pairs from
  the positions of the pressure stage, with the rule's three cases (a dummy
  wall holding the lowest p, p_i itself the minimum, a lower p between re_p
  and re_max) each shown to occur, so dropping any part of the rule fails;
* Laplacian coefficient and quasi compressibility (src/mps.cpp:322-360): for
  phi = |x|^2, sum_j a_ij (phi_j - phi_i) = 2D/rho exactly by the definition
  of lambda; a full lattice row has diagonal 2D/(lambda rho) + alpha/dt^2
  whether its neighbours are Inner or Surface; off-diagonals exist exactly
  between Inner pairs within re_p;
* higher-order source term (src/mps.cpp:288-308): a uniform velocity leaves
  gamma/dt^2 (n/n0 - 1) with n the closed-form lattice sum at the
  compressed spacing; a linear field u = E x gives the lattice sum
  re/(n0 dt) sum (r'.E r)/|r'|^3 over the lattice offsets, r' = (I + E dt) r.

A mistake of the kinds these stages invite (the wrong radius, a dropped
2D or 1/rho factor, walls left out of the viscosity, dummy walls kept in the
minimum, Surface neighbours dropped from the diagonal, u_j in place of u_ij)
fails at least one of them.
"""
import math

import numpy as np
import pytest

from mps_amd import configs
from oracle_lib import Oracle

RHO = 1000.0

# The lattice sums (n0, lambda) are recomputed here in closed form from the
# integer shells of the lattice offsets -4..4 (src/mps.cpp:132-163 loops over
# those offsets).


def _shells(dim):
    rng = range(-4, 5)
    out = {}
    for ix in rng:
        for iy in rng:
            for iz in (rng if dim == 3 else (0,)):
                q = ix * ix + iy * iy + iz * iz
                if q:
                    out[q] = out.get(q, 0) + 1
    return out


def _lattice_sums(dim, ratio, spacing=1.0):
    """(sum w, sum r^2 w / l0^2, sum 1/r * l0, count) over lattice neighbours
    at `spacing` l0 within ratio l0."""
    n = r2w = inv = 0.0
    cnt = 0
    for q, c in _shells(dim).items():
        r = spacing * math.sqrt(q)
        if r < ratio:
            w = ratio / r - 1.0
            n += c * w
            r2w += c * r * r * w
            inv += c / r
            cnt += c
    return n, r2w, inv, cnt


def _n0_lambda(dim, ratio):
    n0, r2w, _, _ = _lattice_sums(dim, ratio)
    return n0, r2w / n0  # lambda in units of l0^2


def _block(dim, l0, nx, ny, nz=1, spacing=1.0, origin=0.0):
    """A free fluid block on a square (cubic) lattice, id order x, y, z."""
    ax = [(np.arange(k) + 0.5) * l0 * spacing + origin for k in (nx, ny, nz if dim == 3 else 1)]
    g = np.meshgrid(*ax, indexing="ij")
    pos = np.zeros((g[0].size, 3))
    pos[:, 0] = g[0].ravel()
    pos[:, 1] = g[1].ravel()
    if dim == 3:
        pos[:, 2] = g[2].ravel()
    else:
        pos[:, 2] = 0.0
    typ = np.full(pos.shape[0], configs.FLUID, np.int8)
    return configs.Scene(typ, pos, np.zeros_like(pos))


def _settings(dim, l0, **kw):
    base = dict(dim=dim, particle_distance=l0, dt=1e-3, gravity=(0.0, 0.0, 0.0), kinematic_viscosity=0.0,
                ratio_pressure=2.1, ratio_viscosity=3.1, ratio_surface=2.1,
                domain_x=(-0.2, 0.6), domain_y=(-0.2, 0.6),
                domain_z=(-0.2, 0.6) if dim == 3 else (0.0, 0.0))
    base.update(kw)
    return configs.Settings(**base)


def _interior(sc, l0, margin, dim):
    lo = sc.pos.min(axis=0)
    hi = sc.pos.max(axis=0)
    m = np.ones(sc.n, bool)
    for d in range(dim):
        m &= (sc.pos[:, d] > lo[d] + margin * l0) & (sc.pos[:, d] < hi[d] - margin * l0)
    return m


# ---------------- viscosity (src/mps.cpp:177-203) ----------------

@pytest.mark.parametrize("dim,size", [(2, (24, 24, 1)), (3, (12, 12, 12))])
def test_viscosity_exact_for_quadratic_field(dim, size):
    l0, nu, C = 0.01, 1e-2, 10.0
    s = _settings(dim, l0, kinematic_viscosity=nu)
    sc = _block(dim, l0, *size)
    yc = sc.pos[:, 1].mean()
    sc.vel[:, 0] = C * (sc.pos[:, 1] - yc) ** 2
    o = Oracle(s, sc)
    o.load_initial_state()
    o.step()
    _, vel = o.snapshot()
    a = (vel[:, 0] - sc.vel[:, 0]) / s.dt
    full = _interior(sc, l0, 3.2, dim)  # complete re_v = 3.1 l0 neighbourhoods
    assert full.sum() >= 8
    np.testing.assert_allclose(a[full], 2.0 * nu * C, rtol=1e-8)
    # the edge rows see fewer neighbours, so the model is not exact there:
    # the test would not notice a radius mistake if every row passed
    edge = ~_interior(sc, l0, 1.0, dim)
    assert np.abs(a[edge] - 2.0 * nu * C).max() > 1e-3 * 2.0 * nu * C
    assert (vel[:, 1] == 0.0).all() and (vel[:, 2] == 0.0).all()


def test_viscosity_pair_with_wall_between_pressure_and_viscosity_radius():
    # a fluid particle 2.5 l0 from a (resting) wall particle: inside re_v =
    # 3.1 l0, outside re_p = 2.1 l0, so only the viscosity radius sees it
    l0, nu, v = 0.01, 1e-3, 0.2
    s = _settings(2, l0, kinematic_viscosity=nu)
    sc = configs.Scene(np.array([configs.FLUID, configs.WALL], np.int8),
                       np.array([[0.1, 0.1, 0.0], [0.1 + 2.5 * l0, 0.1, 0.0]]),
                       np.array([[0.0, v, 0.0], [0.0, 0.0, 0.0]]))
    o = Oracle(s, sc)
    o.load_initial_state()
    o.step()
    _, vel = o.snapshot()
    n0, lam = _n0_lambda(2, 3.1)
    assert n0 == pytest.approx(18.9764166318, abs=1e-9) and lam == pytest.approx(2.4484722940, abs=1e-9)
    coef = nu * (2.0 * 2) / (n0 * lam * l0 * l0)
    w = 3.1 / 2.5 - 1.0
    expect = v + s.dt * coef * (0.0 - v) * w
    assert vel[0, 1] == pytest.approx(expect, rel=1e-13)
    assert vel[0, 0] == 0.0 and (vel[1] == 0.0).all()


def test_viscosity_uniform_field_is_exactly_zero():
    l0 = 0.01
    s = _settings(2, l0, kinematic_viscosity=1e-1)
    sc = _block(2, l0, 16, 16)
    sc.vel[:, 0] = 0.3
    sc.vel[:, 1] = -0.7
    o = Oracle(s, sc)
    o.load_initial_state()
    o.step()
    _, vel = o.snapshot()
    np.testing.assert_array_equal(vel, sc.vel)


# ---------------- minimum pressure (src/mps.cpp:481-503) ----------------

def _brute_dist(pos):
    d = pos[None, :, :] - pos[:, None, :]
    return np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])


def test_minimum_pressure_rule_brute_force():
    s = configs.Settings(ratio_pressure=2.1, ratio_viscosity=3.1, ratio_surface=2.1)
    sc = configs.tutorial_dambreak()  # 1 wall layer, 2 dummy layers
    # jittered fluid: the pressure field gets local minima
    rng = np.random.default_rng(5)
    f = sc.type == configs.FLUID
    sc.pos[f, :2] += rng.uniform(-0.1, 0.1, size=(int(f.sum()), 2)) * s.particle_distance
    sc.vel[f, :2] = rng.normal(0.0, 0.05, size=(int(f.sum()), 2))
    o = Oracle(s, sc, tol=1e-12)
    o.load_initial_state()
    for _ in range(3):
        o.step()
    prev = o.state()["pmin"].copy()
    o.step()
    st = o.state()
    pos, _ = o.snapshot()  # the positions the pressure stage's neighbour build saw
    c = s.to_c()
    D = _brute_dist(pos)
    near = (D < c.re_pressure) & ~np.eye(sc.n, dtype=bool)
    p, typ, bc = st["p"], st["type"], st["bc"]
    dummy = typ == configs.DUMMY_WALL
    keep = near & ~dummy[None, :]
    masked = np.where(keep, p[None, :], np.inf)
    expect = np.minimum(p, masked.min(axis=1))
    act = bc != configs.BC_IGNORED
    np.testing.assert_array_equal(st["pmin"][act], expect[act])
    # Ignored particles keep the value they had (the loop skips them)
    np.testing.assert_array_equal(st["pmin"][~act], prev[~act])
    # each part of the rule decides some particle's value
    with_dummy = np.minimum(p, np.where(near, p[None, :], np.inf).min(axis=1))
    assert (with_dummy[act] < expect[act]).any(), "no dummy wall holds a lower pressure"
    assert ((expect[act] == p[act]) & (masked[act].min(axis=1) > p[act])).any(), "p_i never the strict minimum"
    wide = (D < c.re_max) & ~np.eye(sc.n, dtype=bool) & ~dummy[None, :]
    with_wide = np.minimum(p, np.where(wide, p[None, :], np.inf).min(axis=1))
    assert (with_wide[act] < expect[act]).any(), "re_max would give the same minimum"
    assert (expect[act] > 0).sum() > 20


# ---------------- Laplacian and quasi compressibility (src/mps.cpp:322-360) ----------------

def _system(dim, qc_on, size, spacing=1.0):
    l0 = 0.01
    s = _settings(dim, l0, qc_on=qc_on, dt=1e-3)
    sc = _block(dim, l0, *size, spacing=spacing)
    o = Oracle(s, sc, tol=1e-12)
    o.load_initial_state()
    o.step()
    return s, sc, o, o.pressure_system(), o.state()


@pytest.mark.parametrize("dim,size", [(2, (20, 20, 1)), (3, (11, 11, 11))])
@pytest.mark.parametrize("qc_on", [True, False])
def test_laplacian_coefficient_closed_forms(dim, size, qc_on):
    s, sc, o, ps, st = _system(dim, qc_on, size)
    c = s.to_c()
    l0 = c.particle_distance
    n0, lam = _n0_lambda(dim, 2.1)
    lam *= l0 * l0
    nrow = np.diff(ps["row_ptr"])
    inner = st["bc"] == configs.BC_INNER
    # rows whose whole re_p neighbourhood is lattice fluid (Inner or Surface):
    # two lattice rows in from the edge (offsets up to 2 l0 lie within 2.1 l0)
    full = _interior(sc, l0, 1.9, dim) & inner
    assert full.sum() >= 8
    qc = c.compressibility / (c.dt * c.dt) if qc_on else 0.0
    if qc_on:
        assert qc == pytest.approx(0.45e-9 / 1e-6, rel=1e-15)
    np.testing.assert_allclose(ps["diag"][full], 2.0 * dim / (lam * RHO) + qc, rtol=1e-12)
    # some of those rows have Surface neighbours: the diagonal sums them too
    pos = sc.pos
    D = _brute_dist(pos)
    near = (D < c.re_pressure) & ~np.eye(sc.n, dtype=bool)
    surf = st["bc"] == configs.BC_SURFACE
    assert (near[full][:, surf].any(axis=1)).any()
    # sum_j a_ij (phi_j - phi_i) = 2D / rho for phi = |x|^2 on rows with only
    # Inner neighbours (the off-diagonals are -a_ij)
    phi = (pos[:, 0] ** 2 + pos[:, 1] ** 2) + pos[:, 2] ** 2
    only_inner = full & ~(near[:, surf].any(axis=1))
    assert only_inner.sum() >= 4
    for i in np.flatnonzero(only_inner):
        cols = ps["col"][ps["row_ptr"][i]:ps["row_ptr"][i + 1]]
        vals = ps["val"][ps["row_ptr"][i]:ps["row_ptr"][i + 1]]
        lap = float(np.sum(-vals * (phi[cols] - phi[i])))
        assert lap == pytest.approx(2.0 * dim / RHO, rel=1e-9)
        # and the diagonal exceeds the off-diagonal magnitudes by alpha/dt^2 only
        assert ps["diag"][i] - float(np.sum(-vals)) == pytest.approx(qc, abs=1e-9 * ps["diag"][i])
    # sparsity: (i, j) stored exactly for Inner pairs within re_p
    want = near & inner[:, None] & inner[None, :]
    got = np.zeros_like(want)
    rows = np.repeat(np.arange(sc.n), nrow)
    got[rows, ps["col"]] = True
    np.testing.assert_array_equal(got, want)
    assert (ps["diag"][~inner] == 0).all()


def test_quasi_compressibility_shift_is_alpha_over_dt2():
    _, _, _, on, st_on = _system(2, True, (14, 14, 1))
    s, _, _, off, st_off = _system(2, False, (14, 14, 1))
    np.testing.assert_array_equal(st_on["bc"], st_off["bc"])
    inner = st_on["bc"] == configs.BC_INNER
    c = s.to_c()
    shift = 0.45e-9 / (c.dt * c.dt)
    np.testing.assert_allclose(on["diag"][inner] - off["diag"][inner], shift, rtol=1e-9)
    np.testing.assert_array_equal(on["val"], off["val"])


def test_source_term_closed_form_on_compressed_lattice():
    # b = relax / dt^2 (n - n0) / n0 with n the lattice sum at 0.96 l0
    l0, sp_ = 0.01, 0.96
    s = _settings(2, l0, relaxation=0.2, dt=2e-3)
    sc = _block(2, l0, 20, 20, spacing=sp_)
    o = Oracle(s, sc, tol=1e-12)
    o.load_initial_state()
    o.step()
    ps, st = o.pressure_system(), o.state()
    n0, _, _, _ = _lattice_sums(2, 2.1)
    n, _, _, _ = _lattice_sums(2, 2.1, sp_)
    full = _interior(sc, l0 * sp_, 2.2, 2) & (st["bc"] == configs.BC_INNER)
    assert full.sum() >= 8
    np.testing.assert_allclose(ps["b"][full], 0.2 * (1.0 / (2e-3 * 2e-3)) * (n - n0) / n0, rtol=1e-9)


# ---------------- higher-order source term (src/mps.cpp:288-308) ----------------

def test_hos_uniform_velocity_leaves_the_density_term():
    l0, sp_, gam, dt = 0.01, 0.96, 0.01, 1e-3
    s = _settings(2, l0, hos_on=True, hos_gamma=gam, dt=dt)
    sc = _block(2, l0, 20, 20, spacing=sp_)
    sc.vel[:, 0] = 0.25
    sc.vel[:, 1] = -0.4
    o = Oracle(s, sc, tol=1e-12)
    o.load_initial_state()
    ndr0 = o.state()["ndr"].copy()
    o.step()
    ps, st = o.pressure_system(), o.state()
    n0, _, _, _ = _lattice_sums(2, 2.1)
    n, _, _, _ = _lattice_sums(2, 2.1, sp_)
    inner = st["bc"] == configs.BC_INNER
    full = _interior(sc, l0 * sp_, 2.2, 2) & inner
    assert full.sum() >= 8
    np.testing.assert_allclose(ps["b"][full], gam * (1.0 / (dt * dt)) * (n / n0 - 1.0), rtol=1e-9)
    # u_ij = 0 on every row, also where the neighbourhood is one-sided (near
    # the surface a velocity term sum (r . u)/r^3 over u_j would not vanish)
    edge = inner & ~full
    assert edge.sum() >= 8
    np.testing.assert_allclose(ps["b"][inner], gam * (1.0 / (dt * dt)) * (ndr0[inner] - 1.0), rtol=1e-12)


@pytest.mark.parametrize("E", [((0.3, 0.0), (0.0, -0.1)), ((0.05, 0.4), (-0.2, 0.15))])
def test_hos_linear_velocity_divergence_sum(E):
    l0, gam, dt = 0.01, 0.01, 1e-3
    E = np.array(E)
    s = _settings(2, l0, hos_on=True, hos_gamma=gam, dt=dt)
    sc = _block(2, l0, 20, 20)
    sc.vel[:, :2] = sc.pos[:, :2] @ E.T
    o = Oracle(s, sc, tol=1e-12)
    o.load_initial_state()
    o.step()
    ps, st = o.pressure_system(), o.state()
    n0, _, _, _ = _lattice_sums(2, 2.1)
    re = 2.1 * l0
    M = np.eye(2) + E * dt  # positions after the explicit move: x' = (I + E dt) x
    acc = 0.0
    for ix in range(-4, 5):
        for iy in range(-4, 5):
            if ix == 0 and iy == 0:
                continue
            r = np.array([ix, iy], float) * l0
            rp = M @ r
            dist = math.sqrt(rp[0] ** 2 + rp[1] ** 2)
            if dist < re:
                acc += re * float(rp @ (E @ r)) / dist ** 3
    # interior rows: ndr = 1 up to rounding from the load, so the density
    # term is ~1e-13 of the divergence term
    expect = -(1.0 / (n0 * dt)) * acc
    full = _interior(sc, l0, 2.3, 2) & (st["bc"] == configs.BC_INNER)
    assert full.sum() >= 8
    np.testing.assert_allclose(ps["b"][full], expect, rtol=1e-8)
    assert abs(expect) > 1.0
