"""The oracle's pointwise bodies pinned by mathematics, not by retyping them.

Each check is one a plausible transcription mistake (a dropped term, a wrong sign or index, a
swapped side) fails: consistency F(W, W) = F(W) . n, conservation F(L, R, n) = -F(R, L, -n),
rotation invariance, exact resolution of a stationary contact (Toro 2009, sections 10.4, 11.2),
pure upwinding of supersonic flow, the boundary states' closed forms, to_primitive as the inverse
of the conserved variables, the Venkatakrishnan function's published values, and exact sums
against math.fsum (correctly rounded).
"""

from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as orc

G = 1.4
FLUXES = {"roe": 0, "hllc": 1, "ausm+up": 2}


def prim_rng(rng, mach=0.5, p=1.0):
    rho = rng.uniform(0.5, 2.0)
    c = math.sqrt(G * p / rho)
    v = rng.normal(size=3)
    v *= mach * c / np.linalg.norm(v)
    return np.array([rho, *v, p])


def unit(rng):
    n = rng.normal(size=3)
    return n / np.linalg.norm(n)


def scale_of(w):
    rho, v, p = w[0], w[1:4], w[4]
    q = np.linalg.norm(v) + math.sqrt(G * p / rho)
    return np.array([rho * q, rho * q * q + p, rho * q * q + p, rho * q * q + p, (p / (G - 1) + 0.5 * rho * (v @ v) + p) * q])


@pytest.mark.parametrize("flux", list(FLUXES))
@pytest.mark.parametrize("pref", [0.0, 0.7])
def test_consistency(flux, pref):
    """F(W, W, n) is the physical flux (gauge pressure p - p_ref in the momentum)."""
    rng = np.random.default_rng(1)
    for _ in range(200):
        w = prim_rng(rng, mach=rng.uniform(0.05, 3.0), p=1.0)
        wg = w.copy()
        wg[4] -= pref
        n = unit(rng)
        f = orc.flux(FLUXES[flux], wg, wg, n, G, 0.5, pref)
        exact = orc.physical_flux(wg, n, G, pref)
        # The physical flux itself, from first principles (absolute pressure in the energy).
        rho, v, p = w[0], w[1:4], w[4]
        vn = v @ n
        e = p / (G - 1) + 0.5 * rho * (v @ v)
        closed = np.array([rho * vn, *(rho * vn * v + (p - pref) * n), (e + p) * vn])
        np.testing.assert_allclose(exact, closed, rtol=0, atol=1e-14 * scale_of(w).max())
        assert np.all(np.abs(f - exact) <= 1e-13 * scale_of(w)), (flux, f, exact)


@pytest.mark.parametrize("flux", list(FLUXES))
def test_conservation_antisymmetry(flux):
    """F(L, R, n) = -F(R, L, -n): the flux leaving one cell enters the other."""
    rng = np.random.default_rng(2)
    for _ in range(200):
        wl = prim_rng(rng, mach=rng.uniform(0.05, 2.0), p=rng.uniform(0.5, 2.0))
        wr = prim_rng(rng, mach=rng.uniform(0.05, 2.0), p=rng.uniform(0.5, 2.0))
        n = unit(rng)
        a = orc.flux(FLUXES[flux], wl, wr, n, G, 0.4)
        b = orc.flux(FLUXES[flux], wr, wl, -n, G, 0.4)
        assert np.all(np.abs(a + b) <= 1e-12 * scale_of(wl)), (flux, a, b)


@pytest.mark.parametrize("flux", list(FLUXES))
def test_rotation_invariance(flux):
    """F(Q W_L, Q W_R, Q n) = Q F(W_L, W_R, n) for a rotation Q of the velocities."""
    rng = np.random.default_rng(3)
    for _ in range(100):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        wl = prim_rng(rng, mach=rng.uniform(0.05, 2.0))
        wr = prim_rng(rng, mach=rng.uniform(0.05, 2.0))
        n = unit(rng)
        rot = lambda w: np.array([w[0], *(q @ w[1:4]), w[4]])  # noqa: E731
        a = orc.flux(FLUXES[flux], wl, wr, n, G, 0.4)
        b = orc.flux(FLUXES[flux], rot(wl), rot(wr), q @ n, G, 0.4)
        expect = np.array([a[0], *(q @ a[1:4]), a[4]])
        assert np.all(np.abs(b - expect) <= 1e-12 * scale_of(wl)), (flux, b, expect)


@pytest.mark.parametrize("flux", list(FLUXES))
def test_stationary_contact_is_exact(flux):
    """Gas at rest, one pressure, any densities: the flux is the pressure force alone."""
    rng = np.random.default_rng(4)
    for _ in range(100):
        p = rng.uniform(0.5, 2.0)
        wl = np.array([rng.uniform(0.1, 1.0), 0.0, 0.0, 0.0, p])
        wr = np.array([rng.uniform(1.0, 10.0), 0.0, 0.0, 0.0, p])
        n = unit(rng)
        f = orc.flux(FLUXES[flux], wl, wr, n, G, 0.1)
        expect = np.array([0.0, *(p * n), 0.0])
        assert np.all(np.abs(f - expect) <= 1e-14 * max(p, 1.0)), (flux, f)


@pytest.mark.parametrize("flux", list(FLUXES))
def test_supersonic_flow_is_upwind(flux):
    """Every wave leaves to the right when v . n > c on both sides: F = F(W_L) . n."""
    rng = np.random.default_rng(5)
    for _ in range(100):
        n = unit(rng)
        wl = np.array([1.0, *(2.5 * n * math.sqrt(G)), 1.0])
        wr = wl * np.array([1.05, 1.02, 1.02, 1.02, 1.04])
        f = orc.flux(FLUXES[flux], wl, wr, n, G, 0.5)
        assert np.all(np.abs(f - orc.physical_flux(wl, n, G)) <= 1e-12 * scale_of(wl)), flux


def test_roe_entropy_fix_only_near_sonic_points():
    """Harten's fix (delta = 0.1 c) changes the Roe flux only where an acoustic eigenvalue
    |v.n -+ c| is below delta: away from sonic points the flux is the unfixed Roe flux, which a
    one-dimensional restatement in numpy gives independently."""
    rng = np.random.default_rng(6)
    n = np.array([1.0, 0.0, 0.0])
    for _ in range(50):
        wl = np.array([rng.uniform(0.8, 1.2), rng.uniform(-0.3, 0.3), 0.0, 0.0, rng.uniform(0.8, 1.2)])
        wr = np.array([rng.uniform(0.8, 1.2), rng.uniform(-0.3, 0.3), 0.0, 0.0, rng.uniform(0.8, 1.2)])
        f = orc.flux(0, wl, wr, n, G)
        # Textbook 1D Roe (eigenvector form, Toro 2009 section 11.2) in numpy.
        def cons(w):
            return np.array([w[0], w[0] * w[1], w[0] * (w[4] / (G - 1) / w[0] + 0.5 * w[1] ** 2)])

        def fl(w):
            e = w[4] / (G - 1) + 0.5 * w[0] * w[1] ** 2
            return np.array([w[0] * w[1], w[0] * w[1] ** 2 + w[4], (e + w[4]) * w[1]])

        hl = (wl[4] / (G - 1) + 0.5 * wl[0] * wl[1] ** 2 + wl[4]) / wl[0]
        hr = (wr[4] / (G - 1) + 0.5 * wr[0] * wr[1] ** 2 + wr[4]) / wr[0]
        sl, sr = math.sqrt(wl[0]), math.sqrt(wr[0])
        u = (sl * wl[1] + sr * wr[1]) / (sl + sr)
        h = (sl * hl + sr * hr) / (sl + sr)
        c = math.sqrt((G - 1) * (h - 0.5 * u * u))
        du = cons(wr) - cons(wl)
        a2 = (G - 1) / c**2 * (du[0] * (h - u * u) + u * du[1] - du[2])
        a1 = (du[0] * (u + c) - du[1] - c * a2) / (2 * c)
        a3 = du[0] - a1 - a2
        lam = [abs(u - c), abs(u), abs(u + c)]
        k = [np.array([1, u - c, h - u * c]), np.array([1, u, 0.5 * u * u]), np.array([1, u + c, h + u * c])]
        diss = sum(a * l * kk for a, l, kk in zip((a1, a2, a3), lam, k))
        ref = 0.5 * (fl(wl) + fl(wr)) - 0.5 * diss
        if min(abs(u - c), abs(u + c)) > 0.1 * c:
            np.testing.assert_allclose(f[[0, 1, 4]], ref, rtol=1e-12, atol=1e-13)


def test_boundary_state_closed_forms():
    rng = np.random.default_rng(7)
    R = 287.0
    for _ in range(50):
        n = unit(rng)
        rho, p = rng.uniform(0.8, 1.5), rng.uniform(0.8e5, 1.2e5)
        pref = 1e5
        c = math.sqrt(G * p / rho)
        v = rng.normal(size=3) * 0.3 * c
        wg = np.array([rho, *v, p - pref])
        data = rng.normal(size=5)
        # state: the data as given (absolute pressure), returned gauge
        d = np.array([1.1, 1.0, 2.0, 3.0, 1.05e5])
        np.testing.assert_allclose(orc.boundary_value(wg, 0, d, n, G, R, pref), [*d[:4], d[4] - pref], rtol=1e-15)
        # slip: the mirror image of the velocity, same density and gauge pressure
        m = orc.boundary_value(wg, 2, data, n, G, R, pref)
        np.testing.assert_allclose(m[1:4], v - 2 * (v @ n) * n, rtol=1e-13, atol=1e-13 * c)
        assert m[0] == rho and m[4] == p - pref
        # no-slip wall: at rest, interior pressure; isothermal density p / (R T_w), adiabatic rho
        wa = orc.boundary_value(wg, 5, np.array([0.0, 0, 0, 0, 0]), n, G, R, pref)
        np.testing.assert_array_equal(wa, [rho, 0, 0, 0, p - pref])
        tw = 300.0
        wi = orc.boundary_value(wg, 5, np.array([tw, 0, 0, 0, 0]), n, G, R, pref)
        np.testing.assert_allclose(wi, [p / (R * tw), 0, 0, 0, p - pref], rtol=1e-14)
        # farfield with the freestream equal to the interior: the state itself
        ff = orc.boundary_value(wg, 1, np.array([rho, *v, p]), n, G, R, pref)
        np.testing.assert_allclose(ff, wg, rtol=1e-12, atol=1e-9 * c)
        # outlet at the interior's own pressure: the interior state
        ou = orc.boundary_value(wg, 4, np.array([0, 0, 0, 0, p]), n, G, R, pref)
        np.testing.assert_allclose(ou, wg, rtol=1e-13, atol=1e-12)


def test_farfield_supersonic_sides():
    n = np.array([1.0, 0.0, 0.0])
    inside = np.array([1.0, 2.0 * math.sqrt(G), 0.1, 0.0, 1.0])
    free = np.array([0.9, 0.5, 0.0, 0.0, 0.8])
    # supersonic outflow through the face: the interior state
    np.testing.assert_array_equal(orc.boundary_value(inside, 1, free, n, G, 1.0, 0.0), inside)
    # supersonic inflow: the freestream
    free2 = np.array([0.9, -3.0 * math.sqrt(G * 0.8 / 0.9), 0.0, 0.0, 0.8])
    inside2 = np.array([1.0, 0.0, 0.0, 0.0, 1.0])
    np.testing.assert_array_equal(orc.boundary_value(inside2, 1, free2, n, G, 1.0, 0.0), free2)


def test_inlet_reproduces_its_isentropic_state():
    """An interior state already on the inlet's total conditions, flowing in along -n, is the
    inlet's boundary state (Riemann invariant and total enthalpy both satisfied)."""
    R = 287.0
    n = np.array([0.0, 0.0, 1.0])
    for mach in (0.1, 0.3, 0.6):
        rho, p = 1.2, 1e5
        c = math.sqrt(G * p / rho)
        w = np.array([rho, 0.0, 0.0, -mach * c, p])
        t = p / (rho * R)
        t0 = t * (1 + 0.5 * (G - 1) * mach**2)
        p0 = p * (t0 / t) ** (G / (G - 1))
        out = orc.boundary_value(w, 3, np.array([p0, t0, 0, 0, 0]), n, G, R, 0.0)
        np.testing.assert_allclose(out, w, rtol=1e-12, atol=1e-9)


def test_to_primitive_inverts_the_conserved_variables():
    rng = np.random.default_rng(8)
    for _ in range(200):
        w = prim_rng(rng, mach=rng.uniform(0, 2), p=rng.uniform(0.5, 2))
        rho, v, p = w[0], w[1:4], w[4]
        u = np.array([rho, *(rho * v), p / (G - 1) + 0.5 * rho * (v @ v)])
        ref = np.array([1.0, 0.3, -0.1, 0.2, 2.6])
        pref = (G - 1) * (ref[4] - 0.5 * (ref[1] ** 2 + ref[2] ** 2 + ref[3] ** 2) / ref[0])
        out = orc.to_primitive(u - ref, ref, G)
        np.testing.assert_allclose(out[:4], w[:4], rtol=1e-13, atol=1e-14)
        assert abs(out[4] + pref - p) <= 1e-13 * p


def test_venkatakrishnan_published_values():
    """Venkatakrishnan (1995) eq. 5 with eps = 0 is phi(y) = (y^2 + 2y) / (y^2 + y + 2), y = dmax/d:
    phi(1) = 3/4, phi(2) = 1, phi = 1 + O(1/y) as y -> infinity; phi(d = 0) = 1 for any eps; and
    the side (dmax for d > 0, dmin for d < 0) follows the sign of d."""
    for y, expect in ((1.0, 0.75), (2.0, 1.0), (0.5, 1.25 / 2.75)):
        assert orc.venkat(1.0, y, -123.0, 0.0) == pytest.approx(expect, rel=1e-12)
        assert orc.venkat(-1.0, 123.0, -y, 0.0) == pytest.approx(expect, rel=1e-12)
    for y in (1e4, 1e8):
        assert abs(orc.venkat(1.0, y, -123.0, 0.0) - 1.0) <= 2.0 / y
    assert orc.venkat(0.0, 1.0, -1.0, 0.3) == 1.0
    assert orc.venkat(1.0, 1.0, -1.0, 1e30) == pytest.approx(1.0, rel=1e-12)


def test_warp_vanishes_without_moments_and_is_linear():
    rng = np.random.default_rng(9)
    w = np.array([1.1, 0.2, -0.1, 0.3, 0.05])
    gl, gr = rng.normal(size=15), rng.normal(size=15)
    pl, pr = rng.uniform(0, 1, 5), rng.uniform(0, 1, 5)
    zero = orc.warp(w, gl, pl, gr, pr, np.zeros(9), G, 0.7)
    assert np.all(zero == 0.0)
    m = rng.normal(size=9) * 1e-3
    a = orc.warp(w, gl, pl, gr, pr, m, G, 0.7)
    b = orc.warp(w, gl, pl, gr, pr, 2 * m, G, 0.7)
    np.testing.assert_allclose(b, 2 * a, rtol=1e-13, atol=1e-18)
    # The flux increment through e_j is the directional derivative of the physical flux along
    # e_j times the moment-weighted gradient: check one column against a finite difference.
    mm = np.zeros(9)
    mm[0] = 1e-3  # M[0, 0]: only d0 = 0.5 (pl * gl[:,0] + pr * gr[:,0]) * 1e-3 along x
    dw = 0.5 * (pl * gl[0::3] + pr * gr[0::3]) * 1e-3
    fx = lambda ww: orc.physical_flux(ww, np.array([1.0, 0.0, 0.0]), G, 0.7)  # noqa: E731
    h = 1e-4
    fd = (fx(w + h * dw) - fx(w - h * dw)) / (2 * h)
    np.testing.assert_allclose(orc.warp(w, gl, pl, gr, pr, mm, G, 0.7), fd, rtol=1e-7, atol=1e-12)


def test_exact_sum_is_correctly_rounded():
    rng = np.random.default_rng(10)
    cases = [
        rng.normal(size=10_000),
        rng.normal(size=10_000) * np.exp(rng.uniform(-300, 300, 10_000)),
        np.array([1e308, 1e308, -1e308, -1e308, 1.0, 1e-300, -1e-300]),
        np.array([1.0, 1e-16, 1e-16, 1e-16, -1.0]),
        np.array([5e-324, 5e-324, -1e-323]),
        np.square(rng.normal(size=50_000)) * 1e-20,
    ]
    for x in cases:
        # The exact rational sum, rounded once (Fraction -> float rounds to nearest, ties to even);
        # math.fsum agrees wherever it does not overflow in its partials.
        exact = float(sum(Fraction(v) for v in x.tolist()))
        assert orc.exact_sum(x) == exact
        assert orc.exact_sum(x[::-1]) == exact
        if np.abs(x).max() < 1e300:
            assert exact == math.fsum(x)
    assert math.isinf(orc.exact_sum(np.array([1.0, np.inf])))
    assert math.isnan(orc.exact_sum(np.array([np.inf, -np.inf])))
    assert math.isnan(orc.exact_sum(np.array([1.0, np.nan])))
    assert orc.exact_sum(np.zeros(0)) == 0.0
