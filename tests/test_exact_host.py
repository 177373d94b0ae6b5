"""The HIP library's host exact sums (autan_amd/csrc/device/exact.hpp: the norms', loads' and
residual scales' rounding) through the test program tests/cpp/exact_check.cpp, against the exact
rational value rounded once (fractions.Fraction -> float rounds to nearest, ties to even).

The words are checked directly too: random signed limbs (the device's slot sums are neither
normalised nor of one sign), sums whose carries run through every limb, halfway cases just above,
at and below a tie, results at the subnormal boundary and at overflow."""

from __future__ import annotations

import math
import subprocess
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
PROGRAM = ROOT / "tests" / "_build" / "exact_check"
LIMBS = 68


def run(mode: str, lines: list[str]) -> list[str]:
    if not PROGRAM.exists():
        raise RuntimeError(f"{PROGRAM} is not built: run autan_amd/build.py (or __graft_entry__.build())")
    out = subprocess.run([str(PROGRAM), mode], input="\n".join(lines) + "\n", capture_output=True, text=True,
                         check=True)
    return out.stdout.split()


def parse(text: str) -> float:
    if text in ("nan", "-nan"):
        return math.nan
    if text in ("inf", "-inf"):
        return float(text)
    return float.fromhex(text)


def words_value(w) -> float:
    """The words' exact value rounded once; past the largest double, an infinity."""
    v = sum(Fraction(int(x)) * Fraction(2) ** (32 * i - 1088) for i, x in enumerate(w[:LIMBS]))
    try:
        return float(v)
    except OverflowError:
        return math.inf if v > 0 else -math.inf


def check_sum(x: np.ndarray) -> None:
    got = parse(run("sum", [float(v).hex() for v in x])[0])
    exact = sum(Fraction(float(v)) for v in x)
    try:
        want = float(exact)
    except OverflowError:
        want = math.inf if exact > 0 else -math.inf
    assert got == want and math.copysign(1, got) == math.copysign(1, want) or (got == 0 and want == 0), (got, want)


def test_sums_of_doubles_are_correctly_rounded():
    rng = np.random.default_rng(31)
    cases = [
        rng.normal(size=3000),
        rng.normal(size=3000) * np.exp(rng.uniform(-700, 700, 3000)),
        np.array([1e308, 1e308, -1e308, -1e308, 1.0, 1e-300, -1e-300]),
        np.array([1.0, 1e-16, 1e-16, 1e-16, -1.0]),
        np.array([5e-324, 5e-324, -1e-323]),
        np.array([2.0**-1022, -5e-324]),  # just below the smallest normal: exact subnormal
        np.array([2.0**-1022, 2.0**-1074 * 3]),
        np.square(rng.normal(size=5000)) * 1e-20,
        np.array([1.0, 2.0**-53]),  # a tie, rounded to even (down)
        np.array([1.0, 2.0**-53, 2.0**-1074]),  # just above the tie: up
        np.array([1.0 + 2.0**-52, 2.0**-53]),  # a tie from an odd significand: up
        np.array([-1.0, -(2.0**-53), 2.0**-1074]),  # just inside the tie, negative: toward zero
        np.array([1.7e308, 1.7e308]),  # overflow to +inf
        np.array([-1.7e308, -1.7e308, 1e300]),  # overflow to -inf
        np.array([1.7976931348623157e308, 2.0**970]),  # rounds up past the largest double
        np.array([1.7976931348623157e308, 2.0**969]),  # the tie below it: stays (even)
        np.zeros(0),
        np.array([0.0, -0.0]),
    ]
    for x in cases:
        check_sum(x)
        check_sum(x[::-1])


def test_special_values():
    assert parse(run("sum", [(1.0).hex(), "inf"])[0]) == math.inf
    assert parse(run("sum", ["-inf", (5.0).hex()])[0]) == -math.inf
    assert math.isnan(parse(run("sum", ["inf", "-inf"])[0]))
    assert math.isnan(parse(run("sum", [(1.0).hex(), "nan"])[0]))


def test_words_of_any_sign_and_size_are_correctly_rounded():
    rng = np.random.default_rng(32)
    rows, want = [], []
    for trial in range(400):
        w = np.zeros(LIMBS + 3, dtype=np.int64)
        kind = trial % 5
        if kind == 0:  # dense signed limbs of up to 62 bits
            w[:LIMBS] = rng.integers(-(2**62), 2**62, LIMBS)
        elif kind == 1:  # a few limbs around a random position
            at = int(rng.integers(2, LIMBS - 3))
            w[at - 2 : at + 2] = rng.integers(-(2**40), 2**40, 4)
        elif kind == 2:  # carries through every limb: -1 everywhere above a positive low limb
            w[:LIMBS] = -1
            w[int(rng.integers(0, LIMBS))] = int(rng.integers(1, 2**33))
        elif kind == 3:  # a tie at a random leading position, optionally a far sticky bit
            lead = int(rng.integers(40, 32 * LIMBS - 40))
            for b in (lead, lead - 53):
                w[b // 32] += 1 << (b % 32)
            if rng.integers(0, 2):
                w[0] += 1
            if rng.integers(0, 2):
                w[:LIMBS] = -w[:LIMBS]
        else:  # normalised magnitudes near the top (overflow) and the bottom (subnormals)
            top = int(rng.integers(0, 2))
            span = slice(LIMBS - 3, LIMBS) if top else slice(0, 3)
            w[span] = rng.integers(0, 2**32, 3)
            if rng.integers(0, 2):
                w[:LIMBS] = -w[:LIMBS]
        rows.append(" ".join(str(int(v)) for v in w))
        want.append(words_value(w))
    # exact multiples of 2^1088 (every digit zero after the carries, the whole value above the last
    # limb) and the largest finite magnitudes just under it
    for top in (1 << 32, -(1 << 32), (1 << 31) - 1, -((1 << 31) - 1), -(1 << 40)):
        w = np.zeros(LIMBS + 3, dtype=np.int64)
        w[LIMBS - 1] = top
        rows.append(" ".join(str(int(v)) for v in w))
        want.append(words_value(w))
    got = [parse(t) for t in run("words", rows)]
    for g, e, r in zip(got, want, rows):
        assert g == e, (g, e, r[:80])


def test_words_counts_of_specials():
    base = ["0"] * LIMBS
    out = run("words", [" ".join(base + ["1", "0", "0"]), " ".join(base + ["0", "2", "0"]),
                        " ".join(base + ["1", "1", "0"]), " ".join(base + ["0", "0", "1"]),
                        " ".join(base + ["0", "0", "0"])])
    assert parse(out[0]) == math.inf and parse(out[1]) == -math.inf
    assert math.isnan(parse(out[2])) and math.isnan(parse(out[3]))
    assert parse(out[4]) == 0.0


@pytest.mark.slow
def test_many_equal_terms_sum_exactly():
    """2^21 copies of a term whose 53-bit significand straddles a limb boundary sum exactly."""
    x = np.full(1 << 21, (2.0**53 - 1) * 2.0**-20)
    got = parse(run("sum", [float(x[0]).hex()] * x.size)[0])
    assert got == float(Fraction(float(x[0])) * x.size)
