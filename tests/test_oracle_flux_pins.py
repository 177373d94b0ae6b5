"""The HLLC star state, AUSM+-up's split polynomials and diffusion terms, and the subsonic
characteristic boundary states, pinned by what their derivations fix.

Consistency, antisymmetry, rotation invariance, the stationary contact and supersonic upwinding
(tests/test_oracle_bodies.py) all hold whatever the HLLC star energy or the AUSM+-up K_p / K_u
terms are, because those terms vanish or cancel there. The checks here do not:

* HLLC (Toro, Spruce and Speares 1994; Toro 2009 section 10.4): the star states are fixed by the
  Rankine-Hugoniot conditions across S_L and S_R together with the contact conditions
  u*_L = u*_R = S_M and p*_L = p*_R. The test derives U*_K from those conditions alone (mass,
  tangential momentum, normal momentum and energy jumps solved one by one) with Einfeldt's speed
  bounds (Batten et al. 1997 eq. 13 over Roe's averages), and requires the oracle's flux to equal
  the physical flux of that star state, F(U*_K) = (rho* S_M, rho* S_M v* + p* n, (E* + p*) S_M),
  in both S_M branches, with and without a gauge pressure.
* AUSM+-up (Liou 2006, J. Comput. Phys. 214:137-170): the split Mach and pressure polynomials in
  their published expanded form, M4(+-) = +-(M +- 1)^2/4 +- beta (M^2 - 1)^2 and
  P5(+-) = (M +- 1)^2 (2 -+ M)/4 +- alpha M (M^2 - 1)^2; closed forms at rest (the pressure
  diffusion alone carries mass, against the pressure difference) and in symmetric compression
  (no mass flux; the velocity diffusion raises the interface pressure); the full flux in the
  paper's notation on random pairs.
* Subsonic farfield, inlet and outlet states with an interior that differs from the data: the
  outgoing and incoming Riemann invariants, entropy, tangential velocity, total enthalpy,
  isentropic total pressure and the linearised outgoing characteristic, read back from the
  returned state.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as orc

G = 1.4
HLLC, AUSM = 1, 2


def unit(rng):
    n = rng.normal(size=3)
    return n / np.linalg.norm(n)


def prim(rng, mach, p):
    rho = rng.uniform(0.5, 2.0)
    c = math.sqrt(G * p / rho)
    v = rng.normal(size=3)
    v *= mach * c / np.linalg.norm(v)
    return np.array([rho, *v, p])


def scale(w):
    rho, v, p = w[0], w[1:4], w[4]
    q = np.linalg.norm(v) + math.sqrt(G * p / rho)
    return np.array([rho * q, rho * q * q + p, rho * q * q + p, rho * q * q + p, (p / (G - 1) + 0.5 * rho * (v @ v) + p) * q])


def gauge(w, pref):
    g = np.array(w, dtype=float)
    g[4] -= pref
    return g


# ------------------------------------------------------------------------------------------ HLLC
def einfeldt_bounds(wl, wr, n):
    """Batten et al. (1997) eq. 13: S_L = min(u_L - c_L, u~ - c~), S_R = max(u_R + c_R, u~ + c~),
    u~ and c~ from Roe's square-root-density averages of velocity and total enthalpy."""
    def enthalpy(w):
        return (G / (G - 1) * w[4] + 0.5 * w[0] * (w[1:4] @ w[1:4])) / w[0]

    kl, kr = math.sqrt(wl[0]), math.sqrt(wr[0])
    v = (kl * wl[1:4] + kr * wr[1:4]) / (kl + kr)
    h = (kl * enthalpy(wl) + kr * enthalpy(wr)) / (kl + kr)
    c = math.sqrt((G - 1) * (h - 0.5 * (v @ v)))
    ul, ur = wl[1:4] @ n, wr[1:4] @ n
    cl, cr = math.sqrt(G * wl[4] / wl[0]), math.sqrt(G * wr[4] / wr[0])
    return min(ul - cl, v @ n - c), max(ur + cr, v @ n + c)


def star_from_jumps(w, n, sk, sm):
    """U*_K from the Rankine-Hugoniot jumps across the wave of speed sk, given that the star region
    moves with the contact (normal velocity sm). Returns (rho*, v*, p*, E*) with absolute p*."""
    rho, v, p = w[0], w[1:4], w[4]
    u = v @ n
    e = p / (G - 1) + 0.5 * rho * (v @ v)
    # mass: sk (rho* - rho) = rho* sm - rho u
    rs = rho * (sk - u) / (sk - sm)
    # tangential momentum: sk (rho* t* - rho t) = rho* sm t* - rho u t  =>  t* = t
    vt = v - u * n
    vs = vt + sm * n
    # normal momentum: sk (rho* sm - rho u) = rho* sm^2 + p* - rho u^2 - p
    ps = sk * (rs * sm - rho * u) - rs * sm * sm + rho * u * u + p
    # energy: sk (E* - E) = (E* + p*) sm - (E + p) u
    es = (sk * e - (e + p) * u + ps * sm) / (sk - sm)
    return rs, vs, ps, es


def contact_speed(wl, wr, n, sl, sr):
    """S_M from p*_L = p*_R, with p*_K = p_K + rho_K (S_K - u_K)(S_M - u_K) (normal momentum jump)."""
    ul, ur = wl[1:4] @ n, wr[1:4] @ n
    al, ar = wl[0] * (sl - ul), wr[0] * (sr - ur)
    return (wr[4] - wl[4] + al * ul - ar * ur) / (al - ar)


def hllc_expected(wl, wr, n, pref):
    """The HLLC flux from its definition; wl, wr carry absolute pressures."""
    sl, sr = einfeldt_bounds(wl, wr, n)
    sm = contact_speed(wl, wr, n, sl, sr)

    def phys(rho, v, p, e):
        un = v @ n
        return np.array([rho * un, *(rho * un * v + (p - pref) * n), (e + p) * un])

    def state(w):
        return w[0], w[1:4], w[4], w[4] / (G - 1) + 0.5 * w[0] * (w[1:4] @ w[1:4])

    if sl >= 0:
        return phys(*state(wl)), (sl, sm, sr), None
    if sr <= 0:
        return phys(*state(wr)), (sl, sm, sr), None
    side, sk = (wl, sl) if sm >= 0 else (wr, sr)
    rs, vs, ps, es = star_from_jumps(side, n, sk, sm)
    # The contact condition the star pressures satisfy on both sides.
    other = star_from_jumps(wr if sm >= 0 else wl, n, sr if sm >= 0 else sl, sm)
    return phys(rs, vs, ps, es), (sl, sm, sr), (ps, other[2])


@pytest.mark.parametrize("pref", [0.0, 0.7, -0.4])
def test_hllc_is_the_physical_flux_of_the_rankine_hugoniot_star_state(pref):
    rng = np.random.default_rng(21)
    branches = {"left star": 0, "right star": 0}
    for _ in range(600):
        wl = prim(rng, rng.uniform(0.0, 0.9), rng.uniform(0.5, 2.0))
        wr = prim(rng, rng.uniform(0.0, 0.9), rng.uniform(0.5, 2.0))
        n = unit(rng)
        expect, (sl, sm, sr), pstar = hllc_expected(wl, wr, n, pref)
        if not (sl < 0 < sr):
            continue
        branches["left star" if sm >= 0 else "right star"] += 1
        # p*_L = p*_R holds for the S_M of Batten eq. 15: the contact is a pressure equilibrium.
        assert abs(pstar[0] - pstar[1]) <= 1e-12 * max(wl[4], wr[4])
        f = orc.flux(HLLC, gauge(wl, pref), gauge(wr, pref), n, G, 0.5, pref)
        tol = 1e-12 * np.maximum(scale(wl), scale(wr)) + 1e-12 * abs(pref)
        assert np.all(np.abs(f - expect) <= tol), (f, expect, sm)
    assert min(branches.values()) >= 100, branches


def test_hllc_star_energy_carries_the_absolute_pressure():
    """A pure gauge shift changes nothing physical: the mass and energy fluxes of the same flow with
    p_ref = 0 and p_ref != 0 (gauge states shifted accordingly) are equal, and the momentum flux
    differs by exactly p_ref n. Dropping p_ref from the star energy breaks the energy equality."""
    rng = np.random.default_rng(22)
    for _ in range(300):
        wl = prim(rng, rng.uniform(0.0, 0.9), rng.uniform(0.5, 2.0))
        wr = prim(rng, rng.uniform(0.0, 0.9), rng.uniform(0.5, 2.0))
        n = unit(rng)
        a = orc.flux(HLLC, wl, wr, n, G, 0.5, 0.0)
        for pref in (0.3, 0.9):
            b = orc.flux(HLLC, gauge(wl, pref), gauge(wr, pref), n, G, 0.5, pref)
            tol = 1e-12 * np.maximum(scale(wl), scale(wr))
            assert abs(a[0] - b[0]) <= tol[0] and abs(a[4] - b[4]) <= tol[4], (a, b)
            assert np.all(np.abs(a[1:4] - (b[1:4] + pref * n)) <= tol[1:4])


# ------------------------------------------------------------------------------------- AUSM+-up
BETA = 1.0 / 8.0
KP, KU, SIGMA = 0.25, 0.75, 1.0


def m4(m, sign):
    """Liou (2006) M_(4)^(+-): supersonic (M +- |M|)/2, else +-(M +- 1)^2/4 +- beta (M^2 - 1)^2."""
    if abs(m) >= 1.0:
        return 0.5 * (m + sign * abs(m))
    return sign * 0.25 * (m + sign) ** 2 + sign * BETA * (m * m - 1.0) ** 2


def p5(m, sign, alpha):
    """Liou (2006) P_(5)^(+-): supersonic (1 +- sign(M))/2, else
    (M +- 1)^2 (2 -+ M)/4 +- alpha M (M^2 - 1)^2."""
    if abs(m) >= 1.0:
        return 1.0 if sign * m > 0 else 0.0
    return 0.25 * (m + sign) ** 2 * (2.0 - sign * m) + sign * alpha * m * (m * m - 1.0) ** 2


def crit_speed(w):
    """c* with c*^2 = 2 (gamma - 1) / (gamma + 1) H (Liou 1996): the sonic speed of the total
    enthalpy."""
    h = (G / (G - 1) * w[4] + 0.5 * w[0] * (w[1:4] @ w[1:4])) / w[0]
    return math.sqrt(2 * (G - 1) / (G + 1) * h), h


def scaling(mbar2, minf):
    mo = math.sqrt(min(1.0, max(mbar2, minf * minf)))
    fa = mo * (2.0 - mo)
    return fa, 3.0 / 16.0 * (-4.0 + 5.0 * fa * fa)


def ausm_expected(wl, wr, n, minf, pref):
    """AUSM+-up in Liou (2006)'s notation, eqs. 73 and 75 with the interface speed of eq. 30;
    absolute pressures in, the momentum flux carrying p_1/2 - p_ref."""
    ul, ur = wl[1:4] @ n, wr[1:4] @ n
    csl, hl = crit_speed(wl)
    csr, hr = crit_speed(wr)
    a = min(csl * csl / max(csl, ul), csr * csr / max(csr, -ur))
    ml, mr = ul / a, ur / a
    mbar2 = (ul * ul + ur * ur) / (2.0 * a * a)
    fa, alpha = scaling(mbar2, minf)
    rho = 0.5 * (wl[0] + wr[0])
    mp = -KP / fa * max(1.0 - SIGMA * mbar2, 0.0) * (wr[4] - wl[4]) / (rho * a * a)
    mhalf = m4(ml, +1) + m4(mr, -1) + mp
    pu = -KU * p5(ml, +1, alpha) * p5(mr, -1, alpha) * (wl[0] + wr[0]) * (fa * a) * (ur - ul)
    phalf = p5(ml, +1, alpha) * wl[4] + p5(mr, -1, alpha) * wr[4] + pu
    up = wl if mhalf > 0 else wr
    mdot = a * mhalf * up[0]
    h = hl if mhalf > 0 else hr
    return np.array([mdot, *(mdot * up[1:4] + (phalf - pref) * n), mdot * h])


def test_liou_polynomials_published_values():
    """The expanded polynomials used here against hand-evaluated values of Liou's factored forms:
    M4+(0) = 3/8, M4+(1/2) = 81/128, P5+(0) = 1/2; M4+ + M4- = M and P5+ + P5- = 1."""
    assert m4(0.0, +1) == 0.375 and m4(0.0, -1) == -0.375
    assert m4(0.5, +1) == pytest.approx(81 / 128, rel=1e-15)
    assert p5(0.0, +1, 0.3) == 0.5 and p5(0.0, -1, 0.3) == 0.5
    for m in np.linspace(-1.5, 1.5, 31):
        assert m4(m, +1) + m4(m, -1) == pytest.approx(m, abs=1e-15)
        assert p5(m, +1, 0.1) + p5(m, -1, 0.1) == pytest.approx(1.0, abs=1e-15)


@pytest.mark.parametrize("pref", [0.0, 0.6])
def test_ausm_at_rest_only_the_pressure_diffusion_moves_mass(pref):
    """Both sides at rest: M_L = M_R = 0, so M_1/2 is the pressure diffusion term
    -K_p / f_a (p_R - p_L) / (rho_bar a^2) alone (f_a at M_o = M_inf); mass moves from high to low
    pressure, upwinded, and the interface pressure is (p_L + p_R)/2."""
    rng = np.random.default_rng(23)
    minf = 0.4
    fa = minf * (2.0 - minf)
    for _ in range(200):
        pl, pr_ = rng.uniform(0.5, 2.0, 2)
        wl = np.array([rng.uniform(0.5, 2.0), 0, 0, 0, pl])
        wr = np.array([rng.uniform(0.5, 2.0), 0, 0, 0, pr_])
        n = unit(rng)
        f = orc.flux(AUSM, gauge(wl, pref), gauge(wr, pref), n, G, minf, pref)
        cl = math.sqrt(2 * G * pl / ((G + 1) * wl[0]))  # c* at rest: 2 gamma p / ((gamma + 1) rho)
        cr = math.sqrt(2 * G * pr_ / ((G + 1) * wr[0]))
        a = min(cl, cr)
        m = -KP / fa * (pr_ - pl) / (0.5 * (wl[0] + wr[0]) * a * a)
        up = wl if m > 0 else wr
        mdot = a * m * up[0]
        assert (mdot > 0) == (pl > pr_)
        h = G / (G - 1) * up[4] / up[0]
        expect = np.array([mdot, *((0.5 * (pl + pr_) - pref) * n), mdot * h])
        assert np.all(np.abs(f - expect) <= 1e-13 * np.maximum(scale(wl), scale(wr))), (f, expect)


@pytest.mark.parametrize("pref", [0.0, 0.6])
def test_ausm_symmetric_compression_raises_the_interface_pressure(pref):
    """Equal densities and pressures, normal velocities +d and -d: no mass crosses the face, and
    the interface pressure is 2 P5+(M) p + 4 K_u P5+(M)^2 rho f_a a d with M = d / a - above the
    split pressure in compression (d > 0), below it in expansion (d < 0)."""
    rng = np.random.default_rng(24)
    minf = 0.3
    for _ in range(200):
        rho, p = rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0)
        c = math.sqrt(G * p / rho)
        d = rng.uniform(-0.8, 0.8) * c
        n = unit(rng)
        t = rng.normal(size=3)
        t = (t - (t @ n) * n) * rng.uniform(0, 0.5) * c / np.linalg.norm(t)
        wl = np.array([rho, *(t + d * n), p])
        wr = np.array([rho, *(t - d * n), p])
        f = orc.flux(AUSM, gauge(wl, pref), gauge(wr, pref), n, G, minf, pref)
        cs, _ = crit_speed(wl)
        a = cs * cs / max(cs, abs(d)) if d > 0 else cs
        m = d / a
        fa, alpha = scaling(m * m, minf)
        split = 2.0 * p5(m, +1, alpha) * p
        diffusion = 4.0 * KU * p5(m, +1, alpha) ** 2 * rho * fa * a * d
        # v.n of the two sides are opposite up to the rounding of the dot products
        assert abs(f[0]) <= 1e-14 * rho * c and abs(f[4]) <= 1e-14 * rho * c**3
        phalf = f[1:4] @ n + pref
        assert abs(phalf - (split + diffusion)) <= 1e-13 * (p + rho * c * c)
        assert (phalf > split) == (d > 0)


@pytest.mark.parametrize("pref", [0.0, 0.5])
def test_ausm_up_against_liou_2006(pref):
    rng = np.random.default_rng(25)
    for _ in range(800):
        wl = prim(rng, rng.uniform(0.0, 1.6), rng.uniform(0.5, 2.0))
        wr = prim(rng, rng.uniform(0.0, 1.6), rng.uniform(0.5, 2.0))
        n = unit(rng)
        minf = rng.uniform(0.05, 0.9)
        f = orc.flux(AUSM, gauge(wl, pref), gauge(wr, pref), n, G, minf, pref)
        expect = ausm_expected(wl, wr, n, minf, pref)
        tol = 1e-12 * np.maximum(scale(wl), scale(wr)) + 1e-12 * abs(pref)
        assert np.all(np.abs(f - expect) <= tol), (f, expect)


# ------------------------------------------------------------------------ characteristic states
FARFIELD, INLET, OUTLET = 1, 3, 4


def sound(w):
    return math.sqrt(G * w[4] / w[0])


@pytest.mark.parametrize("pref", [0.0, 0.8])
def test_subsonic_farfield_takes_each_invariant_from_its_side(pref):
    """Whitfield and Janus (1984): R+ = v.n + 2c/(gamma-1) from the interior, R- = v.n - 2c/(gamma-1)
    from the freestream; entropy p / rho^gamma and the tangential velocity from the upwind side."""
    rng = np.random.default_rng(26)
    seen = {"in": 0, "out": 0}
    for _ in range(400):
        n = unit(rng)
        wi = prim(rng, rng.uniform(0.0, 0.8), rng.uniform(0.6, 1.6))
        wf = prim(rng, rng.uniform(0.0, 0.8), rng.uniform(0.6, 1.6))
        if wi[1:4] @ n >= sound(wi) or wf[1:4] @ n <= -sound(wf):
            continue
        out = orc.boundary_value(gauge(wi, pref), FARFIELD, wf, n, G, 1.0, pref)
        wb = out.copy()
        wb[4] += pref
        rp_i = wi[1:4] @ n + 2 * sound(wi) / (G - 1)
        rm_f = wf[1:4] @ n - 2 * sound(wf) / (G - 1)
        vn = wb[1:4] @ n
        cb = sound(wb)
        assert vn + 2 * cb / (G - 1) == pytest.approx(rp_i, rel=1e-12, abs=1e-12)
        assert vn - 2 * cb / (G - 1) == pytest.approx(rm_f, rel=1e-12, abs=1e-12)
        up = wf if vn < 0 else wi
        seen["in" if vn < 0 else "out"] += 1
        assert wb[4] / wb[0] ** G == pytest.approx(up[4] / up[0] ** G, rel=1e-12)
        tb, tu = wb[1:4] - vn * n, up[1:4] - (up[1:4] @ n) * n
        np.testing.assert_allclose(tb, tu, rtol=0, atol=1e-12 * (1 + np.linalg.norm(tu)))
    assert min(seen.values()) >= 50, seen


def test_subsonic_inlet_holds_total_conditions_and_the_outgoing_invariant():
    """Carlson (2011): the state flows in along -n with total enthalpy cp T0
    (c_b^2/(gamma-1) + vn^2/2 = cp T0 with c_b^2 = gamma R T_b), total pressure p0 by the isentrope
    p_b/p0 = (T_b/T0)^(gamma/(gamma-1)), and keeps the interior's R+ = v.n + 2c/(gamma-1)."""
    rng = np.random.default_rng(27)
    rgas = 287.0
    cp = G * rgas / (G - 1)
    for _ in range(300):
        n = unit(rng)
        t0, p0 = rng.uniform(250, 350), rng.uniform(0.9e5, 1.3e5)
        # an interior somewhat below the total conditions, flowing in
        p = rng.uniform(0.7, 0.98) * p0
        rho = p / (rgas * t0 * rng.uniform(0.85, 0.99))
        wi = np.array([rho, *(rng.normal(size=3) * 20.0 - rng.uniform(10, 80) * n), p])
        out = orc.boundary_value(wi, INLET, np.array([p0, t0, 0, 0, 0]), n, G, rgas, 0.0)
        vn = out[1:4] @ n
        assert vn < 0
        np.testing.assert_allclose(out[1:4], vn * n, rtol=0, atol=1e-9)
        tb = out[4] / (rgas * out[0])
        assert cp * tb + 0.5 * vn * vn == pytest.approx(cp * t0, rel=1e-12)
        assert out[4] / p0 == pytest.approx((tb / t0) ** (G / (G - 1)), rel=1e-11)
        rp = wi[1:4] @ n + 2 * sound(wi) / (G - 1)
        assert vn + 2 * sound(out) / (G - 1) == pytest.approx(rp, rel=1e-11)


@pytest.mark.parametrize("pref", [0.0, 0.9])
def test_subsonic_outlet_keeps_entropy_and_the_outgoing_acoustic_invariant(pref):
    """Linearised characteristics about the interior (Blazek 2015, section 8.5): at the prescribed
    pressure, dp - c^2 drho = 0 and dp + rho c dvn = 0, tangential velocity unchanged."""
    rng = np.random.default_rng(28)
    for _ in range(300):
        n = unit(rng)
        wi = prim(rng, rng.uniform(0.0, 0.9), rng.uniform(0.6, 1.6))
        if wi[1:4] @ n >= sound(wi):
            continue
        pb = wi[4] * rng.uniform(0.7, 1.3)
        out = orc.boundary_value(gauge(wi, pref), OUTLET, np.array([0, 0, 0, 0, pb]), n, G, 1.0, pref)
        wb = out.copy()
        wb[4] += pref
        ci = sound(wi)
        dp = wb[4] - wi[4]
        assert wb[4] == pytest.approx(pb, rel=1e-14)
        assert abs(dp - ci * ci * (wb[0] - wi[0])) <= 1e-12 * wi[4]
        assert abs(dp + wi[0] * ci * ((wb[1:4] - wi[1:4]) @ n)) <= 1e-12 * wi[4]
        tb, ti = wb[1:4] - (wb[1:4] @ n) * n, wi[1:4] - (wi[1:4] @ n) * n
        np.testing.assert_allclose(tb, ti, rtol=0, atol=1e-13 * ci)
