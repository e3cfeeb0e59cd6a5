"""Plain-Python restatement of the reference's HITRAN .par reader.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and the cpu_baseline leg of
bench.py) as the checker of the device parser, never as the product.

Follows src/HITRAN/linelist.jl:90-103 (parse_hitran_par: fixed-width columns,
Julia's parse(Int8 / Float64 / Float32, field)), src/HITRAN/fetch.jl:12-61
(load_hitran_par: eachline, records shorter than 67 characters and records
whose fields do not parse are skipped, the [nu_min, nu_max] window) and
src/HITRAN/linelist.jl:51-56 (the list sorted stably by wavenumber).

Number semantics: a field is stripped of surrounding white space; Float64
fields are the correctly rounded value of the decimal (Python's float, as
strtod); Float32 fields are rounded from the exact decimal straight to the
nearest float32 (as strtof; not via a double), with exact rational arithmetic.
Int8 fields: optional sign and decimal digits in [-128, 127].  The grammar is
the one HITRAN files use (digits, point, E exponent); Python-only spellings
(underscores, "infinity") are not exercised by the tests.
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np

COLS = ((0, 2), (2, 3), (3, 15), (15, 25), (25, 35), (35, 40), (40, 45), (45, 55), (55, 59), (59, 67))
KINDS = ("i8", "i8", "f64", "f64", "f64", "f32", "f32", "f64", "f32", "f32")
NAMES = ("mol", "iso", "nu", "S", "A", "gamma_air", "gamma_self", "elower", "n_air", "delta_air")


def _space(c):
    return c in " \t\n\v\f\r"


def _strip(s):
    a, b = 0, len(s)
    while a < b and _space(s[a]):
        a += 1
    while b > a and _space(s[b - 1]):
        b -= 1
    return s[a:b]


def parse_int8(field: str) -> int:
    t = _strip(field)
    body = t[1:] if t[:1] in "+-" else t
    if not body or not all("0" <= c <= "9" for c in body):
        raise ValueError(field)
    v = int(t)
    if not -128 <= v <= 127:
        raise ValueError(field)
    return v


def parse_f64(field: str) -> float:
    t = _strip(field)
    if not t or "_" in t:
        raise ValueError(field)
    return float(t)


def round_f32(x: Fraction) -> np.float32:
    """The float32 nearest to x (ties to even), x in the normal range or 0."""
    if x == 0:
        return np.float32(0.0)
    neg = x < 0
    x = -x if neg else x
    e = x.numerator.bit_length() - x.denominator.bit_length()
    if Fraction(2) ** e > x:
        e -= 1
    if not -126 <= e <= 127:
        raise ValueError("outside the normal float32 range")
    q = x / Fraction(2) ** (e - 23)                      # in [2^23, 2^24)
    n = q.numerator // q.denominator
    r = q - n
    if r > Fraction(1, 2) or (r == Fraction(1, 2) and n % 2 == 1):
        n += 1
    v = np.float32(math.ldexp(n, e - 23))
    return np.float32(-v) if neg else v


def parse_f32(field: str) -> np.float32:
    t = _strip(field)
    if not t or "_" in t:
        raise ValueError(field)
    f = float(t)                                          # grammar check
    if f == 0.0:
        return np.float32(-0.0) if t.startswith("-") else np.float32(0.0)
    return round_f32(Fraction(t))


def parse_record(rec: str):
    """Ten fields of one record (newline removed); ValueError when malformed."""
    if len(rec) < 67:
        raise ValueError("short record")
    out = []
    for (a, b), k in zip(COLS, KINDS):
        s = rec[a:b]
        out.append(parse_int8(s) if k == "i8" else (parse_f64(s) if k == "f64" else parse_f32(s)))
    return tuple(out)


def load_par_text(text: str, nu_min: float = -math.inf, nu_max: float = math.inf):
    """(dict of field arrays sorted stably by wavenumber, number of skipped records)."""
    recs, skipped = [], 0
    lines = text.split("\n")
    if lines and lines[-1] == "":
        lines = lines[:-1]
    for rec in lines:
        if rec.endswith("\r"):
            rec = rec[:-1]
        try:
            r = parse_record(rec)
        except ValueError:
            skipped += 1
            continue
        if nu_min <= r[2] <= nu_max:
            recs.append(r)
    order = sorted(range(len(recs)), key=lambda i: recs[i][2])       # Python's sort is stable
    dt = {"i8": np.int8, "f64": np.float64, "f32": np.float32}
    out = {n: np.array([recs[i][j] for i in order], dt[k]) if recs else np.zeros(0, dt[k])
           for j, (n, k) in enumerate(zip(NAMES, KINDS))}
    return out, skipped
