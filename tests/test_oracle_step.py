"""Pins of the oracle's whole step and loop: boundary maps, invariants, the Sod
shock tube against the exact solution, the CFL dt, and the initial conditions."""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib as O
import textbook as T

ID, IP, IU, IV = 0, 1, 2, 3


def ghost_source(i, n, lo, hi):
    """Interior index a ghost index copies from, and the normal-velocity sign:
    reflect mirrors about the wall, zero-gradient copies the edge cell,
    periodic wraps around (ghosts are 2 cells deep, interior is [2, n+1])."""
    if 2 <= i <= n + 1:
        return i, 1.0
    if i < 2:
        d = 2 - i  # distance from the wall, 1 or 2
        if lo == 1:
            return 1 + d, -1.0
        if lo == 2:
            return 2, 1.0
        return i + n, 1.0
    d = i - (n + 1)
    if hi == 1:
        return n + 2 - d, -1.0
    if hi == 2:
        return n + 1, 1.0
    return i - n, 1.0


@pytest.mark.parametrize("bcs", [(1, 1, 1, 1), (2, 2, 2, 2), (3, 3, 3, 3), (1, 2, 3, 1),
                                 (3, 1, 2, 2), (0, 4, 1, 3)])
def test_boundaries_closed_form(bcs):
    p = O.default_params(nx=5, ny=4, boundary_type_xmin=bcs[0], boundary_type_xmax=bcs[1],
                         boundary_type_ymin=bcs[2], boundary_type_ymax=bcs[3])
    r = O.OracleRun(p, init=False)
    rng = np.random.default_rng(1)
    a = rng.uniform(-1, 1, r.shape)
    r.set(0, a)
    r.make_boundaries(0)
    b = r.get(0)
    for i in range(p.isize):
        for j in range(p.jsize):
            si, sx = ghost_source(i, p.nx, bcs[0], bcs[1])
            sj, sy = ghost_source(j, p.ny, bcs[2], bcs[3])
            want = a[si, sj].copy()
            want[IU] *= sx
            want[IV] *= sy
            np.testing.assert_array_equal(b[i, j], want, err_msg=f"{i},{j}")


def _uniform_run(p, w):
    r = O.OracleRun(p, init=False)
    u = np.empty(r.shape)
    u[:] = T.cons(p.gamma0, w)
    r.set(0, u)
    r.set(1, u)
    return r


@pytest.mark.parametrize("bc", [2, 3])
def test_uniform_state_stays_uniform(bc):
    p = O.default_params(nx=24, ny=20, cfl=0.8, nStepmax=25, tEnd=1e9, boundary_type_xmin=bc,
                         boundary_type_xmax=bc, boundary_type_ymin=bc, boundary_type_ymax=bc)
    w = np.array([1.3, 0.7, 0.4, -0.25])
    r = _uniform_run(p, w)
    r.run()
    u = O.interior(r.current())
    np.testing.assert_allclose(u, np.broadcast_to(T.cons(p.gamma0, w), u.shape), rtol=1e-13)


def test_dt_closed_form():
    p = O.default_params(nx=16, ny=12, cfl=0.7, xmax=2.0, ymax=0.5)
    w = np.array([1.3, 0.7, 0.4, -0.25])
    r = _uniform_run(p, w)
    c = math.sqrt(p.gamma0 * w[1] / w[0])
    want = p.cfl / ((c + abs(w[2])) / p.dx + (c + abs(w[3])) / p.dy)
    assert r.compute_dt(0) == pytest.approx(want, rel=1e-14)


@pytest.mark.parametrize("bc,conserved", [(1, (ID, IP)), (3, (ID, IP, IU, IV))])
def test_conservation(bc, conserved):
    p = O.default_params(nx=40, ny=36, problemType=1, cfl=0.7, nStepmax=150, tEnd=1e9,
                         blast_radius=0.15, blast_center_x=0.4, blast_center_y=0.55,
                         boundary_type_xmin=bc, boundary_type_xmax=bc,
                         boundary_type_ymin=bc, boundary_type_ymax=bc)
    r = O.OracleRun(p)
    s0 = O.interior(r.get(0)).reshape(-1, 4).sum(0)
    r.run()
    s1 = O.interior(r.current()).reshape(-1, 4).sum(0)
    scale = O.interior(r.current()).reshape(-1, 4).__abs__().sum(0)
    for v in conserved:
        assert abs(s1[v] - s0[v]) <= 1e-13 * scale[v], (v, s0[v], s1[v])
    if bc == 1:  # walls exert pressure: momentum is not conserved
        assert abs(s1[IU] - s0[IU]) > 1e-8


def test_transpose_symmetry():
    """Swapping the x and y axes (and u with v) transposes the solution."""
    kw = dict(problemType=1, cfl=0.6, nStepmax=40, tEnd=1e9, blast_radius=0.2,
              blast_pressure_in=5.0)
    p1 = O.default_params(nx=30, ny=44, xmax=1.0, ymax=1.5, boundary_type_xmin=1,
                          boundary_type_xmax=2, boundary_type_ymin=3, boundary_type_ymax=3,
                          blast_center_x=0.35, blast_center_y=0.9, **kw)
    p2 = O.default_params(nx=44, ny=30, xmax=1.5, ymax=1.0, boundary_type_xmin=3,
                          boundary_type_xmax=3, boundary_type_ymin=1, boundary_type_ymax=2,
                          blast_center_x=0.9, blast_center_y=0.35, **kw)
    r1, r2 = O.OracleRun(p1), O.OracleRun(p2)
    u1 = r1.get(0)
    u2 = np.transpose(u1, (1, 0, 2))[:, :, [ID, IP, IV, IU]]
    r2.set(0, u2)
    r2.set(1, u2)
    r1.run()
    r2.run()
    assert r1.t == r2.t
    a = O.interior(r1.current())
    b = np.transpose(O.interior(r2.current()), (1, 0, 2))[:, :, [ID, IP, IV, IU]]
    assert np.all(O.rel_l1(b, a)[:2] < 1e-12)
    assert np.abs(b - a).max() < 1e-11


def test_mirror_symmetry():
    p = O.default_params(nx=48, ny=40, problemType=1, cfl=0.7, nStepmax=40, tEnd=1e9,
                         blast_radius=0.2, blast_center_x=0.5, blast_center_y=0.3)
    r = O.OracleRun(p)
    r.run()
    a = O.interior(r.current())
    m = a[::-1, :, :].copy()
    m[:, :, IU] *= -1
    assert np.abs(m - a).max() < 1e-11


def _sod(nx, axis):
    n_long, n_short = nx, 4
    kw = dict(cfl=0.8, tEnd=0.2, nStepmax=100000, gamma0=1.4, boundary_type_xmin=2,
              boundary_type_xmax=2, boundary_type_ymin=2, boundary_type_ymax=2)
    if axis == 0:
        p = O.default_params(nx=n_long, ny=n_short, ymax=n_short / n_long, **kw)
    else:
        p = O.default_params(nx=n_short, ny=n_long, xmax=n_short / n_long, **kw)
    r = O.OracleRun(p, init=False)
    u = np.empty(r.shape)
    coord = (np.arange(p.isize if axis == 0 else p.jsize) - 2 + 0.5) / n_long
    for k, x in enumerate(coord):
        w = [1.0, 1.0, 0.0, 0.0] if x < 0.5 else [0.125, 0.1, 0.0, 0.0]
        if axis == 0:
            u[k, :, :] = T.cons(1.4, w)
        else:
            u[:, k, :] = T.cons(1.4, w)
    r.set(0, u)
    r.set(1, u)
    r.run()
    assert r.t == pytest.approx(0.2, rel=1e-15)
    prof = O.interior(r.current())
    rho = prof[:, 1, ID] if axis == 0 else prof[1, :, ID]
    xc = (np.arange(n_long) + 0.5) / n_long
    exact = np.array([T.exact_sample(1.4, [1.0, 1.0, 0.0, 0.0], [0.125, 0.1, 0.0, 0.0],
                                     (x - 0.5) / 0.2)[0] for x in xc])
    return np.abs(rho - exact).mean()


@pytest.mark.parametrize("axis", [0, 1])
def test_sod_converges_to_exact(axis):
    e1, e2 = _sod(100, axis), _sod(400, axis)
    assert e1 < 0.02 and e2 < 0.008
    assert e2 < 0.5 * e1


def test_tend_clamp():
    p = O.default_params(nx=32, ny=32, cfl=0.8, tEnd=0.1125, nStepmax=100000)
    r = O.OracleRun(p)
    r.run()
    assert r.nStep > 5
    assert r.t == pytest.approx(0.1125, rel=1e-15, abs=0)


def _centres(p):
    x = [Fraction(p.xmin) + Fraction(p.dx) / 2 + (i - 2) * Fraction(p.dx) for i in range(p.isize)]
    y = [Fraction(p.ymin) + Fraction(p.dy) / 2 + (j - 2) * Fraction(p.dy) for j in range(p.jsize)]
    return x, y


def test_init_implode_region():
    p = O.default_params(nx=30, ny=26, xmin=-1.0, xmax=1.0, gamma0=1.666)
    u = O.OracleRun(p).get(0)
    x, y = _centres(p)
    for i in range(p.isize):
        for j in range(p.jsize):
            s = x[i] + y[j] * y[j]
            if abs(s - Fraction(1, 2)) < 1e-9 or abs(s - Fraction(3, 2)) < 1e-9:
                continue
            hi = Fraction(1, 2) < s < Fraction(3, 2)
            w = [1.0, 1.0, 0, 0] if hi else [0.125, 0.14, 0, 0]
            np.testing.assert_allclose(u[i, j], T.cons(p.gamma0, w), rtol=1e-15)


def test_init_four_quadrant_states():
    p = O.default_params(nx=20, ny=20, problemType=2)
    u = O.OracleRun(p).get(0)
    x, y = _centres(p)
    W = {(1, 1): [1.5, 1.5, 0, 0], (0, 1): [0.5323, 0.3, 1.206, 0],
         (0, 0): [0.138, 0.029, 1.206, 1.206], (1, 0): [0.5323, 0.3, 0, 1.206]}
    for i in range(p.isize):
        for j in range(p.jsize):
            w = W[(int(x[i] >= Fraction(4, 5)), int(y[j] >= Fraction(4, 5)))]
            q, _ = O.prim(p, u[i, j])
            np.testing.assert_allclose(q, w, rtol=1e-14, atol=1e-15)


def test_init_blast_disk_and_sedov_energy():
    p = O.default_params(nx=40, ny=40, problemType=1, blast_radius=0.2, blast_center_x=0.5,
                         blast_center_y=0.5, blast_density_in=2.0, blast_density_out=1.0)
    u = O.OracleRun(p).get(0)
    x, y = _centres(p)
    inside = np.zeros((p.isize, p.jsize), bool)
    for i in range(p.isize):
        for j in range(p.jsize):
            inside[i, j] = (x[i] - Fraction(1, 2)) ** 2 + (y[j] - Fraction(1, 2)) ** 2 < \
                Fraction(p.blast_radius) ** 2
    assert np.all(u[inside, ID] == 2.0) and np.all(u[~inside, ID] == 1.0)
    # total_energy_inside > 0: the disk holds exactly that energy
    p2 = O.default_params(nx=40, ny=40, problemType=1, blast_radius=0.2, blast_center_x=0.5,
                          blast_center_y=0.5, blast_total_energy_inside=3.0)
    u2 = O.OracleRun(p2).get(0)
    e_in = (u2[inside, IP] * p2.dx * p2.dy).sum()
    assert e_in == pytest.approx(3.0, rel=1e-13)


def test_shocked_bubble_rankine_hugoniot():
    """The deck's post-shock state is the Mach-3 Rankine-Hugoniot state."""
    g, M = 1.4, 3.0
    r2 = (g + 1) * M * M / ((g - 1) * M * M + 2)
    p2 = (2 * g * M * M - (g - 1)) / (g + 1)
    S = M * math.sqrt(g)
    u2 = S * (1 - 1 / r2)
    assert r2 == pytest.approx(3.857142857142857, rel=1e-15)
    assert p2 == pytest.approx(10.333333333333332, rel=1e-15)
    assert u2 == pytest.approx(2.6293687924887177, rel=1e-15)
    p = O.default_params(nx=40, ny=20, problemType=4, xmax=2.0, sb_preshock_density=1.0,
                         sb_preshock_pressure=1.0, sb_postshock_density=r2,
                         sb_postshock_pressure=p2, sb_postshock_velocity=u2, sb_shock_loc=0.1,
                         sb_bubble_density=0.1, sb_bubble_pressure=1.0, sb_bubble_radius=0.15,
                         sb_bubble_center_x=0.4, sb_bubble_center_y=0.5)
    u = O.OracleRun(p).get(0)
    q, _ = O.prim(p, u[2, 10])
    np.testing.assert_allclose(q, [r2, p2, u2, 0.0], rtol=1e-14)
    q, _ = O.prim(p, u[2 + 8, 12])  # x = 0.41, y = 0.525: inside the bubble
    np.testing.assert_allclose(q, [0.1, 1.0, 0, 0], rtol=1e-14)


def test_discontinuity_init():
    p = O.default_params(nx=16, ny=16, problemType=3)
    u = O.OracleRun(p).get(0)
    x, y = _centres(p)
    for i in range(p.isize):
        for j in range(p.jsize):
            if x[i] + y[j] < 1:
                assert u[i, j, ID] == pytest.approx(1 + float(x[i]) ** 2, rel=1e-15)
            else:
                assert u[i, j, ID] == 0.25
