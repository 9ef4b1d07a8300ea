"""Pins of the oracle's per-cell / per-face functions against mathematics that
is independent of the reference's code (tests/textbook.py)."""
import math

import numpy as np
import pytest

import oracle_lib as O
import textbook as T

ID, IP, IU, IV = 0, 1, 2, 3


@pytest.fixture(scope="module")
def P():
    return O.default_params(gamma0=1.4, smallr=1e-10, smallc=1e-10)


def random_states(n, seed, umax=2.0):
    rng = np.random.default_rng(seed)
    r = 10 ** rng.uniform(-1, 1, n)
    p = 10 ** rng.uniform(-1, 1, n)
    u = rng.uniform(-umax, umax, n)
    v = rng.uniform(-umax, umax, n)
    return np.stack([r, p, u, v], axis=1)


# ------------------------------------------------------------ primitives
def test_primitives_invert_conservative(P):
    g = P.gamma0
    for w in random_states(200, 1):
        q, c = O.prim(P, T.cons(g, w))
        np.testing.assert_allclose(q, w, rtol=1e-12, atol=1e-14)
        assert c == pytest.approx(math.sqrt(g * w[1] / w[0]), rel=1e-13)


def test_primitives_floors(P):
    # density floor smallr, pressure floor rho * smallp
    q, _ = O.prim(P, [0.0, 1.0, 0.0, 0.0])
    assert q[ID] == P.smallr
    q, _ = O.prim(P, [2.0, 0.5 * 2.0 * 9.0, 6.0, 0.0])  # all energy kinetic: e = 0
    assert q[IP] == 2.0 * P.smallp


# -------------------------------------------------------------- slopes
@pytest.mark.parametrize("st", [1.0, 2.0])
def test_slopes_closed_forms(st):
    P = O.default_params(slope_type=st)
    rng = np.random.default_rng(3)
    for _ in range(200):
        q = rng.uniform(-1, 1, 4)
        hl = rng.uniform(0.01, 1, 4)
        hr = rng.uniform(0.01, 1, 4)
        # monotone increasing in x, decreasing in y
        dx, dy = O.slopes(P, q, q + hr, q - hl, q - hr, q + hl)
        want = np.minimum(np.minimum(st * hl, st * hr), 0.5 * (hl + hr))
        np.testing.assert_allclose(dx, want, rtol=1e-14)
        np.testing.assert_allclose(dy, -want, rtol=1e-14)
        # extremum (both neighbours above) and flat data give zero slopes
        dx, dy = O.slopes(P, q, q + hr, q + hl, q, q)
        assert np.all(dx == 0) and np.all(dy == 0)
    # linear data: minmod and MC return the exact difference
    q = np.array([1.0, 2.0, 0.5, -0.25])
    h = np.array([0.125, 0.25, 0.5, 0.0625])
    dx, dy = O.slopes(P, q, q + h, q - h, q + 2 * h, q - 2 * h)
    np.testing.assert_array_equal(dx, h)
    np.testing.assert_array_equal(dy, 2 * h)


def test_slopes_type0_zero():
    P = O.default_params(slope_type=0.0)
    dx, dy = O.slopes(P, [1, 2, 3, 4], [2, 3, 4, 5], [0, 1, 2, 3], [3, 3, 3, 3], [0, 0, 0, 0])
    assert np.all(dx == 0) and np.all(dy == 0)


# --------------------------------------------------------------- trace
def test_trace_equals_hancock_predictor(P):
    """Face states equal W -/+ dW/2 + dt/2 * (-A dWx - B dWy)/dx (dx == dy)."""
    g = P.gamma0
    rng = np.random.default_rng(5)
    for w in random_states(100, 5):
        dwx = rng.uniform(-0.05, 0.05, 4) * np.abs(w).max()
        dwy = rng.uniform(-0.05, 0.05, 4) * np.abs(w).max()
        dtdx = dtdy = rng.uniform(0.05, 0.4)
        for face in range(4):
            got = O.trace(P, w, dwx, dwy, dtdx, dtdy, face)
            want = T.hancock_face(g, w, dwx, dwy, dtdx, dtdy, face)
            want[ID] = max(want[ID], P.smallr)
            np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-13)


# ------------------------------------------------------------- Riemann
def test_hllc_equals_toro(P):
    g = P.gamma0
    states = random_states(400, 11)
    for wl, wr in zip(states[:200], states[200:]):
        f = O.hllc(P, wl, wr)
        ref = T.hllc_flux(g, wl, wr)
        scale = np.abs(ref).max() + 1.0
        np.testing.assert_allclose(f, ref, rtol=0, atol=1e-12 * scale)


def test_hllc_supersonic_and_contact(P):
    g = P.gamma0
    wl = np.array([1.0, 1.0, 5.0, 0.3])
    wr = np.array([0.5, 0.4, 6.0, -0.2])
    np.testing.assert_allclose(O.hllc(P, wl, wr), T.xflux(g, wl), rtol=1e-15)
    # left-moving supersonic flow (SR < 0): the right state's flux
    ml = wl * [1, 1, -1, 1]
    mr = wr * [1, 1, -1, 1]
    np.testing.assert_allclose(O.hllc(P, mr, ml), T.xflux(g, ml), rtol=1e-15)
    np.testing.assert_allclose(O.hll(P, mr, ml), T.xflux(g, ml), rtol=1e-15)
    # stationary contact: HLLC resolves it exactly (no mass flux), HLL does not
    cl = np.array([1.0, 1.0, 0.0, 0.7])
    cr = np.array([0.125, 1.0, 0.0, -0.3])
    f = O.hllc(P, cl, cr)
    assert f[ID] == 0.0 and f[IV] == 0.0 and f[1] == 0.0
    assert f[IU] == pytest.approx(1.0, rel=1e-15)
    assert abs(O.hll(P, cl, cr)[ID]) > 1e-3


@pytest.mark.parametrize("form", ["10.21", "10.20"])
def test_hll_equals_toro(P, form):
    g = P.gamma0
    states = random_states(400, 13)
    for wl, wr in zip(states[:200], states[200:]):
        f = O.hll(P, wl, wr)
        ref = T.hll_flux(g, wl, wr, form)
        scale = np.abs(ref).max() + 1.0
        np.testing.assert_allclose(f, ref, rtol=0, atol=1e-12 * scale)


def test_fluxes_consistent(P):
    """F(W, W) is the physical flux for every solver."""
    g = P.gamma0
    for w in random_states(50, 17):
        ref = T.xflux(g, w)
        np.testing.assert_allclose(O.hll(P, w, w), ref, rtol=1e-12, atol=1e-13)
        np.testing.assert_allclose(O.hllc(P, w, w), ref, rtol=1e-12, atol=1e-13)
        np.testing.assert_allclose(O.approx(P, w, w)[1], ref, rtol=1e-9, atol=1e-10)


def test_approx_matches_exact_two_shock(P):
    """The two-shock approximation is exact when both waves are shocks."""
    g = P.gamma0
    rng = np.random.default_rng(19)
    checked = 0
    for _ in range(300):
        wl = np.array([10 ** rng.uniform(-0.5, 0.5), 10 ** rng.uniform(-0.5, 0.5),
                       rng.uniform(0.5, 3.0), rng.uniform(-1, 1)])
        wr = np.array([10 ** rng.uniform(-0.5, 0.5), 10 ** rng.uniform(-0.5, 0.5),
                       -rng.uniform(0.5, 3.0), rng.uniform(-1, 1)])
        ps, us = T.exact_star(g, wl, wr)
        if ps <= max(wl[IP], wr[IP]):
            continue  # not shock-shock
        qg, flux = O.approx(P, wl, wr)
        r, p, u = T.exact_sample(g, wl, wr, 0.0)
        np.testing.assert_allclose([qg[ID], qg[IP], qg[IU]], [r, p, u], rtol=1e-7, atol=1e-9)
        v = wl[IV] if us > 0 else wr[IV]
        np.testing.assert_allclose(flux, T.xflux(g, [r, p, u, v]), rtol=1e-6, atol=1e-8)
        checked += 1
    assert checked > 100


def test_approx_toro_test5(P):
    """Toro's test 5 (two strong shocks), a shock-shock problem with known p*."""
    P5 = O.default_params(gamma0=1.4)
    wl = [5.99924, 460.894, 19.5975, 0.0]
    wr = [5.99242, 46.0950, -6.19633, 0.0]
    ps, us = T.exact_star(1.4, wl, wr)
    assert ps == pytest.approx(1691.64, rel=1e-5) and us == pytest.approx(8.68975, rel=1e-5)
    qg, _ = O.approx(P5, wl, wr)
    r, p, u = T.exact_sample(1.4, wl, wr, 0.0)
    np.testing.assert_allclose([qg[ID], qg[IP], qg[IU]], [r, p, u], rtol=1e-7)
