"""The manufactured-solution pipeline (autan_amd/mms.py) against the reference's own autan.mms/1
document for the Euler equations (py/tests/fixtures/euler.mms.json, kept as
tests/golden/mms/euler.mms.json): Roy et al. (2004)'s trigonometric fields, their gradients, the
conserved variables and the volume sources the reference derived for them.

The fields are read from the document and fed through this repository's own derivation
(`mms.euler_source`: div F(W) of the steady Euler fluxes) and conversion (`mms.conserved`); the
results must equal the document's gradients, conserved variables and sources at random points of
its domain. A dropped term, a wrong sign or a misplaced pressure in the derivation fails here."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

sp = pytest.importorskip("sympy")

from autan_amd import mms  # noqa: E402

DOC = Path(__file__).resolve().parent / "golden" / "mms" / "euler.mms.json"


def document():
    return json.loads(DOC.read_text())


def expressions(block: dict, symbols: dict) -> dict:
    """A block's outputs as sympy expressions, its subexpressions substituted in order."""
    local = dict(symbols)
    local.update({"pi": sp.pi, "pow": sp.Pow, "sin": sp.sin, "cos": sp.cos, "exp": sp.exp, "sqrt": sp.sqrt})
    for name, text in block.get("subexpressions", []):
        local[name] = sp.sympify(text, locals=local)
    return {k: sp.sympify(text, locals=local) for k, text in block["outputs"].items()}


@pytest.fixture(scope="module")
def setup():
    doc = document()
    assert doc["schema"] == "autan.mms/1" and doc["equations"] == "compressible Euler"
    x, y, z = sp.symbols("x y z", real=True)
    sym = {"x": x, "y": y, "z": z}
    fields = expressions(doc["fields"], sym)
    rng = np.random.default_rng(41)
    lo, hi = np.array(doc["domain"]["lower"]), np.array(doc["domain"]["upper"])
    pts = lo + (hi - lo) * rng.uniform(size=(64, 3))
    return doc, (x, y, z), fields, pts


def evaluate(exprs, coords, pts):
    f = sp.lambdify(coords, exprs, "numpy")
    out = f(pts[:, 0], pts[:, 1], pts[:, 2])
    return np.stack([np.broadcast_to(np.asarray(v, dtype=float), pts.shape[:1]) for v in out])


def test_sources_are_the_documents(setup):
    doc, coords, fw, pts = setup
    gamma = doc["parameters"]["gamma"]
    ours = mms.euler_source(fw["rho"], fw["u"], fw["v"], fw["w"], fw["p"], coords, gamma)
    theirs = expressions(doc["sources"], dict(zip("xyz", coords)))
    a = evaluate(ours, coords, pts)
    b = evaluate([theirs[k] for k in ("rho", "rhou", "rhov", "rhow", "rhoE")], coords, pts)
    scale = np.abs(b).max(axis=1, keepdims=True)
    np.testing.assert_allclose(a / scale, b / scale, rtol=0, atol=1e-12)


def test_conserved_variables_are_the_documents(setup):
    doc, coords, fw, pts = setup
    gamma = doc["parameters"]["gamma"]
    w = evaluate([fw[k] for k in ("rho", "u", "v", "w", "p")], coords, pts)
    ours = mms.conserved(w, gamma)
    block = doc["conserved"]
    theirs = expressions(block, dict(zip("xyz", coords)))
    names = list(block["outputs"])
    assert len(names) == 5
    b = evaluate([theirs[k] for k in names], coords, pts)
    np.testing.assert_allclose(ours, b, rtol=1e-13, atol=0)


def test_field_gradients_are_the_documents(setup):
    """The fields as read (their subexpressions substituted) differentiate to the document's
    gradients: the reading of the document is right before anything is derived from it."""
    doc, coords, fw, pts = setup
    theirs = expressions(doc["gradients"], dict(zip("xyz", coords)))
    for name in ("rho", "u", "v", "w", "p"):
        for axis, c in zip("xyz", coords):
            a = evaluate([sp.diff(fw[name], c)], coords, pts)[0]
            b = evaluate([theirs[f"{name}_{axis}"]], coords, pts)[0]
            np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-9 * max(1.0, np.abs(b).max()))


def test_a_wrong_energy_flux_fails():
    """The check has teeth: with the pressure left out of the energy flux the sources differ."""
    doc = document()
    x, y, z = sp.symbols("x y z", real=True)
    fw = expressions(doc["fields"], {"x": x, "y": y, "z": z})
    g = doc["parameters"]["gamma"]
    e = fw["p"] / (g - 1) + fw["rho"] * (fw["u"] ** 2 + fw["v"] ** 2 + fw["w"] ** 2) / 2
    wrong = sum(sp.diff(e * v, c) for v, c in zip((fw["u"], fw["v"], fw["w"]), (x, y, z)))
    theirs = expressions(doc["sources"], {"x": x, "y": y, "z": z})["rhoE"]
    pts = np.random.default_rng(42).uniform(size=(16, 3))
    a = evaluate([wrong], (x, y, z), pts)[0]
    b = evaluate([theirs], (x, y, z), pts)[0]
    assert np.abs(a - b).max() > 1e-3 * np.abs(b).max()
