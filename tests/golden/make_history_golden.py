"""Writes the history fixtures from the reference runtime's own JSON library: oracle/_ref/json_dump
(oracle/build_ref.sh, built in a tree that has the reference) prints nlohmann::ordered_json's
dump() of doubles and of history lines. Run from the repository root:

    sh oracle/build_ref.sh && python tests/golden/make_history_golden.py

history_numbers.txt: "BITS TEXT" per line (the double's 16 hex bits, the library's text).
history_lines.txt: "REQUEST<TAB>LINE" per line, REQUEST as json_dump reads it."""

from __future__ import annotations

import struct
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
DUMP = ROOT / "oracle" / "_ref" / "json_dump"
HERE = Path(__file__).resolve().parent


def bits(x: float) -> str:
    return f"{struct.unpack('<Q', struct.pack('<d', float(x)))[0]:016x}"


def dump(requests: list[str]) -> list[str]:
    out = subprocess.run([str(DUMP)], input="\n".join(requests) + "\n", capture_output=True, text=True,
                         check=True).stdout.splitlines()
    assert len(out) == len(requests)
    return out


def numbers() -> list[float]:
    rng = np.random.default_rng(20261202)
    raw = rng.integers(0, 2**63 - 1, 1200, dtype=np.int64).astype(np.uint64).view(np.float64)
    raw = raw[np.isfinite(raw)]
    residual_like = 10 ** rng.uniform(-18, 6, 500)
    subnormal = rng.integers(1, 2**52, 60, dtype=np.int64).astype(np.uint64).view(np.float64)
    powers2 = [2.0**k for k in range(-1074, 1024, 23)]
    powers10 = [float(f"1e{k}") for k in range(-320, 309, 7)]
    near10 = [np.nextafter(p, d) for p in powers10[::3] for d in (0.0, np.inf)]
    layout = [0.0, -0.0, 1.0, -1.0, 0.1, 0.3, 2.0 / 3.0, 100.0, 1e-4, 1.2e-4, 9.99e-5, 1e-5, 1e15, 1e16,
              9.999999999999998e14, 123456789012345.0, 1234567890123456.0, 999999999999999.9,
              5e-324, 2.2250738585072014e-308, 1.7976931348623157e308, 4.35, 0.5, 1.5e300]
    vals = np.concatenate([raw, residual_like, -residual_like[:40], subnormal, powers2, powers10, near10, layout])
    return [float(v) for v in vals]


def _digits(text: str) -> str:
    """The significant digits of a number's text."""
    mantissa = text.lstrip("-").split("e")[0].replace(".", "")
    return mantissa.strip("0")


def not_shortest(count: int = 40) -> list[float]:
    """Doubles whose Grisu2 digits differ from the shortest-nearest ones (Python's repr): the
    cases a formatter built on repr gets wrong."""
    rng = np.random.default_rng(11)
    raw = rng.integers(0, 2**63 - 1, 40000, dtype=np.int64).astype(np.uint64).view(np.float64)
    raw = [float(v) for v in raw[np.isfinite(raw)]]
    text = dump([f"n {bits(v)}" for v in raw])
    found = [v for v, t in zip(raw, text) if _digits(t) != _digits(repr(v))]
    assert len(found) >= count
    return found[:count]


def lines() -> list[str]:
    rng = np.random.default_rng(7)
    req = []
    for i in range(40):
        r = 10 ** rng.uniform(-14, 2, 5)
        if i == 5:
            r[3] = float("nan")
        if i % 2 == 0:
            req.append(f"s {i + 1} {int(i < 10)} " + " ".join(bits(x) for x in r))
        else:
            req.append(f"u {i + 1} {bits(rng.uniform(0, 2))} " + " ".join(bits(x) for x in r))
    return req


def main() -> None:
    if not DUMP.exists():
        raise SystemExit(f"{DUMP} is missing: run sh oracle/build_ref.sh in a tree that has the reference")
    vals = numbers() + not_shortest()
    req = [f"n {bits(v)}" for v in vals]
    text = dump(req)
    (HERE / "history_numbers.txt").write_text("".join(f"{bits(v)} {t}\n" for v, t in zip(vals, text)))
    req = lines()
    out = dump(req)
    (HERE / "history_lines.txt").write_text("".join(f"{r}\t{o}\n" for r, o in zip(req, out)))
    print(f"{len(vals)} numbers, {len(req)} lines")


if __name__ == "__main__":
    main()
