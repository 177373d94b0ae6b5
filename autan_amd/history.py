"""history.jsonl in the reference runtime's format (docs/contracts/results.md, History;
runtime/src/run.cpp:2317-2342): one compact JSON object per recorded iteration, keys in the
runtime's order — `iteration`, `time` (null for steady runs, the accumulated time steps otherwise),
`residuals` by equation name (runtime/src/input.cpp:935: density, momentum-x, momentum-y,
momentum-z, energy; null where not finite), `monitors`, and for steady runs `provisional`.

Numbers are written as the runtime's JSON library writes a double (its `dump()`, through
nlohmann::ordered_json, runtime/src/run.cpp:63): the digits of Grisu2 (F. Loitsch, "Printing
Floating-Point Numbers Quickly and Accurately with Integers", PLDI 2010) with the rounding
interval of the double's neighbours and the target exponent window [-60, -32], then a fixed
layout — a decimal point when the point falls within 15 digits left or 4 zeros right of the digits,
`d.ddde±XX` otherwise, a trailing `.0` on whole numbers. Grisu2's digits round-trip but are not
always the shortest or nearest (about 0.5 % of doubles differ from Python's repr), so this module
generates them itself and the lines are the runtime's byte for byte."""

from __future__ import annotations

import json
import math
import struct
from functools import lru_cache
from pathlib import Path

EQUATIONS = ("density", "momentum-x", "momentum-y", "momentum-z", "energy")

_ALPHA, _GAMMA = -60, -32  # the exponent window of the scaled numbers (the paper's alpha, gamma)
_POW_MIN, _POW_STEP = -300, 8  # decimal exponents of the cached powers: -300, -292, ..., 324


@lru_cache(maxsize=None)
def _power(k: int) -> tuple[int, int]:
    """10^k as a 64-bit significand f in [2^63, 2^64) and binary exponent e, f 2^e rounded to
    nearest (ties up), from exact integer arithmetic."""
    if k >= 0:
        num, den = 10**k, 1
    else:
        num, den = 1, 10 ** (-k)
    e = num.bit_length() - den.bit_length() - 64
    while True:  # f = round(num / (den 2^e)) with 2^63 <= f < 2^64
        n, d = (num << -e, den) if e < 0 else (num, den << e)
        q, r = divmod(n, d)
        if q >= 1 << 64:
            e += 1
            continue
        if q < 1 << 63:
            e -= 1
            continue
        break
    if 2 * r >= d:
        q += 1
        if q == 1 << 64:
            q, e = 1 << 63, e + 1
    return q, e


def _cached_power(e: int) -> tuple[int, int, int]:
    """The cached power 10^k whose product with a normalised number of binary exponent e lands in
    [alpha, gamma]: k = ceil((alpha - e - 1) log10 2) with log10 2 ~ 78913 / 2^18, rounded up to the
    cache's grid."""
    f = _ALPHA - e - 1
    q = abs(f) * 78913 >> 18
    k = (q if f >= 0 else -q) + (1 if f > 0 else 0)  # C's truncating division, then + (f > 0)
    index = (-_POW_MIN + k + (_POW_STEP - 1)) // _POW_STEP
    kc = _POW_MIN + index * _POW_STEP
    cf, ce = _power(kc)
    return cf, ce, kc


def _normalize(f: int, e: int) -> tuple[int, int]:
    s = 64 - f.bit_length()
    return f << s, e - s


def _mul(a: tuple[int, int], b: tuple[int, int]) -> tuple[int, int]:
    """The 64 x 64-bit product's upper 64 bits, rounded half up."""
    return (a[0] * b[0] + (1 << 63)) >> 64, a[1] + b[1] + 64


def _round_down(digits: list[int], dist: int, delta: int, rest: int, ten_k: int) -> None:
    """Moves the last digit down while the result stays inside the interval and gets closer to w
    (the paper's round_weed, one sided)."""
    while rest < dist and delta - rest >= ten_k and (rest + ten_k < dist or dist - rest > rest + ten_k - dist):
        digits[-1] -= 1
        rest += ten_k


def _grisu2(x: float) -> tuple[str, int]:
    """Digits and decimal exponent of x > 0 (x = digits 10^exponent)."""
    bits = struct.unpack("<Q", struct.pack("<d", x))[0]
    be, bf = bits >> 52, bits & ((1 << 52) - 1)
    if be == 0:
        f, e = bf, 1 - 1075
    else:
        f, e = bf | (1 << 52), be - 1075
    closer_below = bf == 0 and be > 1
    plus = _normalize(2 * f + 1, e - 1)
    minus_f, minus_e = (4 * f - 1, e - 2) if closer_below else (2 * f - 1, e - 1)
    minus = (minus_f << (minus_e - plus[1]), plus[1])
    v = _normalize(f, e)
    cf, ce, kc = _cached_power(plus[1])
    c = (cf, ce)
    w, w_minus, w_plus = _mul(v, c), _mul(minus, c), _mul(plus, c)
    lo = w_minus[0] + 1
    hi, he = w_plus[0] - 1, w_plus[1]
    exponent = -kc

    delta = hi - lo
    dist = hi - w[0]
    shift = -he
    one = 1 << shift
    p1 = hi >> shift
    p2 = hi & (one - 1)
    n = len(str(p1)) if p1 > 0 else 1
    pow10 = 10 ** (n - 1)
    digits: list[int] = []
    while n > 0:
        d, p1 = divmod(p1, pow10)
        digits.append(d)
        n -= 1
        rest = (p1 << shift) + p2
        if rest <= delta:
            exponent += n
            _round_down(digits, dist, delta, rest, pow10 << shift)
            return "".join(map(str, digits)), exponent
        pow10 //= 10
    m = 0
    while True:
        p2 *= 10
        digits.append(p2 >> shift)
        p2 &= one - 1
        m += 1
        delta *= 10
        dist *= 10
        if p2 <= delta:
            break
    exponent -= m
    _round_down(digits, dist, delta, p2, one)
    return "".join(map(str, digits)), exponent


def format_double(x: float) -> str:
    """The text the runtime's JSON library writes for x (null when not finite)."""
    x = float(x)
    if not math.isfinite(x):
        return "null"
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    if x == 0.0:
        return sign + "0.0"
    digits, exponent = _grisu2(abs(x))
    k = len(digits)
    n = k + exponent  # the decimal point's place after the first digit
    if k <= n <= 15:
        text = digits + "0" * (n - k) + ".0"
    elif 0 < n <= 15:
        text = digits[:n] + "." + digits[n:]
    elif -4 < n <= 0:
        text = "0." + "0" * (-n) + digits
    else:
        e = n - 1
        mantissa = digits if k == 1 else digits[0] + "." + digits[1:]
        text = f"{mantissa}e{'-' if e < 0 else '+'}{abs(e):02d}"
    return sign + text


def _value(v) -> str:
    if v is None:
        return "null"
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, int):
        return str(v)
    if isinstance(v, float):
        return format_double(v)
    if isinstance(v, dict):
        return "{" + ",".join(json.dumps(str(k)) + ":" + _value(x) for k, x in v.items()) + "}"
    if isinstance(v, str):
        return json.dumps(v)
    return format_double(float(v))


def forces(force, moment, ref_area: float | None = None, ref_length: float | None = None, density: float = 1.0,
           speed: float = 1.0, aoa: float = 0.0, sideslip: float = 0.0) -> dict:
    """A forces monitor's record (runtime/src/run.cpp:1590-1614; docs/contracts/results.md, Forces
    monitor): fx fy fz mx my mz, then with a reference area cd cs cl (the force on the drag
    direction (cos a cos b, -sin b, sin a cos b), on the side direction lift x drag and on the lift
    direction (-sin a, 0, cos a), over q A with q = 1/2 rho V^2), and with a reference length also
    cmx cmy cmz (the moment over q A L); aoa and sideslip in radians, as the runtime holds them."""
    f = [float(x) for x in force]
    m = [float(x) for x in moment]
    r = {"fx": f[0], "fy": f[1], "fz": f[2], "mx": m[0], "my": m[1], "mz": m[2]}
    if ref_area is None:
        return r
    q = 0.5 * density * speed * speed
    a, b = aoa, sideslip
    drag = (math.cos(a) * math.cos(b), -math.sin(b), math.sin(a) * math.cos(b))
    lift = (-math.sin(a), 0.0, math.cos(a))
    side = (lift[1] * drag[2] - lift[2] * drag[1], lift[2] * drag[0] - lift[0] * drag[2],
            lift[0] * drag[1] - lift[1] * drag[0])

    def dot(u, v):
        return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]

    qa = q * ref_area
    r["cd"] = dot(f, drag) / qa
    r["cs"] = dot(f, side) / qa
    r["cl"] = dot(f, lift) / qa
    if ref_length is not None:
        qal = qa * ref_length
        r["cmx"] = m[0] / qal
        r["cmy"] = m[1] / qal
        r["cmz"] = m[2] / qal
    return r


def record(iteration: int, residuals, time: float | None = None, monitors: dict | None = None,
           provisional: bool | None = None) -> str:
    """One history line: steady when `provisional` is given (time null), unsteady otherwise."""
    if len(residuals) != len(EQUATIONS):
        raise ValueError("a history line carries the five equations' residuals")
    steady = provisional is not None
    if steady and time is not None:
        raise ValueError("a steady line has no time")
    if not steady and time is None:
        raise ValueError("an unsteady line has a time")
    fields = {
        "iteration": int(iteration),
        "time": None if steady else float(time),
        "residuals": {k: float(r) for k, r in zip(EQUATIONS, residuals)},
        "monitors": {} if monitors is None else monitors,
    }
    if steady:
        fields["provisional"] = bool(provisional)
    return _value(fields)


def steady_lines(result, first_iteration: int = 1, monitors: dict[int, dict] | None = None) -> list[str]:
    """The lines of a steady run (steady.solve's result: history and provisional per iteration),
    numbered from `first_iteration` (a continued run's start + 1); `monitors` maps an iteration to
    the monitors recorded there."""
    mon = monitors or {}
    return [
        record(first_iteration + i, r, monitors=mon.get(first_iteration + i), provisional=bool(p))
        for i, (r, p) in enumerate(zip(result.history, result.provisional))
    ]


def unsteady_lines(steps, first_iteration: int = 1, start_time: float = 0.0,
                   monitors: dict[int, dict] | None = None) -> list[str]:
    """The lines of unsteady steps (DeviceStepper.step results in order): each line's time is the
    previous one plus the step's time step, summed in FP64 as the runtime sums it."""
    mon = monitors or {}
    out, t = [], float(start_time)
    for i, s in enumerate(steps):
        t = t + float(s.time_step)
        it = first_iteration + i
        out.append(record(it, s.residuals, time=t, monitors=mon.get(it)))
    return out


def write(path, lines: list[str], append: bool = False) -> None:
    """Each line whole, LF-terminated, as the run directory appends them."""
    with Path(path).open("a" if append else "w", encoding="utf-8", newline="\n") as f:
        for line in lines:
            f.write(line + "\n")


def read(path) -> list[dict]:
    """The lines of a history file (an unterminated last line is skipped, as readers skip a live
    run's partial line)."""
    data = Path(path).read_bytes()
    complete, _, _ = data.rpartition(b"\n")
    return [json.loads(line) for line in complete.split(b"\n")] if complete else []
