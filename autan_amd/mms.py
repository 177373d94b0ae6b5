"""Manufactured solution of the steady Euler equations on the unit cube.

Smooth sinusoidal rho, u, v, w, p with amplitude 0.1 about rho = 1, p = 1 / gamma (sound speed 1)
and a mean velocity of Mach 0.3 along x. The volume source s = div F(W) makes W an exact steady
solution of R(U) = sum_f F . S - s V = 0 (the sign convention of kernels/euler.py:611-616: the
source enters the residual as -s V). Derived with sympy, evaluated with numpy.
"""

from __future__ import annotations

import functools

import numpy as np

GAMMA = 1.4
AMPLITUDE = 0.1
MACH = 0.3


def euler_source(rho, u, v, w, p, coords, gamma):
    """s = div F(W) of primitive fields given as sympy expressions of `coords` (x, y, z): the
    steady Euler fluxes rho v_d, rho v v_d + p e_d, (E + p) v_d with E = p/(gamma - 1) + rho |v|^2/2,
    each differentiated along its own coordinate and summed. Returns the five source expressions
    (mass, three momenta, energy)."""
    import sympy as sp

    g = sp.Float(gamma) if not isinstance(gamma, sp.Basic) else gamma
    e = p / (g - 1) + rho * (u * u + v * v + w * w) / 2
    vel = (u, v, w)
    fluxes = []
    for d in range(3):
        vd = vel[d]
        fluxes.append(
            [
                rho * vd,
                rho * u * vd + (p if d == 0 else 0),
                rho * v * vd + (p if d == 1 else 0),
                rho * w * vd + (p if d == 2 else 0),
                (e + p) * vd,
            ]
        )
    return [sum(sp.diff(fluxes[d][k], xd) for d, xd in enumerate(coords)) for k in range(5)]


@functools.cache
def _functions(gamma: float = GAMMA, amp: float = AMPLITUDE, mach: float = MACH):
    import sympy as sp

    x, y, z = sp.symbols("x y z", real=True)
    tau = 2 * sp.pi
    rho = 1 + amp * sp.sin(tau * x + sp.Rational(3, 10)) * sp.cos(tau * y + sp.Rational(1, 10)) * sp.sin(
        tau * z + sp.Rational(7, 10)
    )
    u = mach * (1 + amp * sp.cos(tau * x - sp.Rational(1, 2)) * sp.sin(tau * y + sp.Rational(1, 5)))
    v = mach * amp * sp.sin(tau * x + sp.Rational(1, 4)) * sp.cos(tau * z - sp.Rational(3, 10))
    w = mach * amp * sp.cos(tau * y - sp.Rational(2, 5)) * sp.sin(tau * z + sp.Rational(1, 10))
    p = (1 + amp * sp.cos(tau * x + sp.Rational(3, 5)) * sp.cos(tau * y - sp.Rational(1, 3)) * sp.cos(tau * z)) / gamma
    source = euler_source(rho, u, v, w, p, (x, y, z), gamma)
    prim = sp.lambdify((x, y, z), [rho, u, v, w, p], "numpy")
    src = sp.lambdify((x, y, z), source, "numpy")
    return prim, src


def primitive(points: np.ndarray) -> np.ndarray:
    """W = (rho, u, v, w, p) at points (n, 3): shape (5, n)."""
    prim, _ = _functions()
    out = prim(points[:, 0], points[:, 1], points[:, 2])
    return np.stack([np.broadcast_to(np.asarray(v, dtype=float), points.shape[:1]) for v in out])


def source(points: np.ndarray) -> np.ndarray:
    """s = div F(W) at points (n, 3): shape (5, n)."""
    _, src = _functions()
    out = src(points[:, 0], points[:, 1], points[:, 2])
    return np.stack([np.broadcast_to(np.asarray(v, dtype=float), points.shape[:1]) for v in out])


def conserved(w: np.ndarray, gamma: float = GAMMA) -> np.ndarray:
    rho, u, v, ww, p = w
    return np.stack([rho, rho * u, rho * v, rho * ww, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + ww * ww)])


def freestream() -> tuple[float, tuple[float, float, float], float]:
    """The reference state the scheme is built on: rho = 1, Mach 0.3 along x, p = 1 / gamma."""
    return 1.0, (MACH, 0.0, 0.0), 1.0 / GAMMA
