"""The .par oracle (oracle/par_ref.py) pinned by the reference's own fixtures
and by exact arithmetic; no GPU."""
import math
from fractions import Fraction

import numpy as np

import par_ref as pr


def _mk(mol, iso, nu, S):
    """The 67-character record builder of test/runtests.jl:949-953 (Julia's
    string() of the floats: 700.0, 1.0e-21)."""
    def jl(v):
        s = repr(float(v))                        # shortest round-trip digits, as Julia prints
        if "e" in s:
            m, e = s.split("e")
            s = (m if "." in m else m + ".0") + "e" + e.lstrip("+")
        return s
    return (str(mol).rjust(2) + str(iso) + jl(nu).rjust(12) + jl(S).rjust(10) + "0.0".rjust(10) +
            "0.05".rjust(5) + "0.07".rjust(5) + "0.0".rjust(10) + "0.75".rjust(4) + "0.0".rjust(8))


def test_reference_fixture_records():
    """runtests.jl:955-980: three records parse to 700/710/720 cm^-1, and the
    705-715 window keeps the 710 line."""
    recs = [_mk(2, 1, 700.0, 1.0e-21), _mk(2, 1, 710.0, 2.0e-21), _mk(2, 1, 720.0, 3.0e-21)]
    assert all(len(r) == 67 for r in recs) and recs[0][15:25] == "   1.0e-21"
    text = "\n".join(recs) + "\n"
    ll, skipped = pr.load_par_text(text)
    assert skipped == 0 and len(ll["nu"]) == 3
    assert list(ll["nu"]) == [700.0, 710.0, 720.0] and list(ll["S"]) == [1e-21, 2e-21, 3e-21]
    assert np.all(ll["mol"] == 2) and np.all(ll["iso"] == 1)
    assert np.all(ll["gamma_air"] == np.float32(0.05)) and np.all(ll["gamma_self"] == np.float32(0.07))
    assert np.all(ll["n_air"] == np.float32(0.75)) and np.all(ll["delta_air"] == 0)
    w, _ = pr.load_par_text(text, 705.0, 715.0)
    assert len(w["nu"]) == 1 and w["nu"][0] == 710.0


def test_reference_160_character_record_is_not_in_parser_columns():
    """The record of runtests.jl:301-308 is written with single spaces between
    values, not in the parser's columns (the reference's test builds the struct
    instead of parsing it): columns 4:15 read ' 667.37986 2', which is not a
    number, so load_hitran_par skips the record."""
    rec = (" 11 667.37986 2.910E-19 3.510E-03 0.0602 0.4680  0.000000 0.75 0.000000" + " " * 15 + " " * 15 +
           " " * 6 + " " * 8 + " " * 4 + " " * 10)
    assert rec[3:15] == " 667.37986 2"
    ll, skipped = pr.load_par_text(rec + "\n")
    assert skipped == 1 and len(ll["nu"]) == 0


def test_julia_parse_rules():
    assert pr.parse_int8(" 2") == 2 and pr.parse_int8("+7") == 7 and pr.parse_int8("-3") == -3
    for bad in ("", " ", "A", "128", "1.0"):
        try:
            pr.parse_int8(bad)
            assert False, bad
        except ValueError:
            pass
    assert pr.parse_f64(" 667.379860") == 667.37986 and pr.parse_f64("2.910E-19") == 2.91e-19
    assert pr.parse_f64("  -.001234") == -0.001234
    for bad in ("", "    ", "1.0 2", "1.0-10"):
        try:
            pr.parse_f64(bad)
            assert False, bad
        except ValueError:
            pass


def test_float32_fields_round_directly():
    """Float32 fields round the exact decimal to float32 (strtof), which is
    not always float32(float64(decimal)): for x = 1 + 2^-24 + 2^-60 the double
    is the tie 1 + 2^-24 and rounds to even (1.0), x itself rounds up."""
    x = Fraction(1) + Fraction(1, 2 ** 24) + Fraction(1, 2 ** 60)
    s = format_decimal(x)
    assert Fraction(s) == x
    assert np.float32(float(s)) == np.float32(1.0)
    assert pr.round_f32(Fraction(s)) == np.float32(1.0 + 2.0 ** -23)
    # and agrees with numpy where no double rounding occurs
    rng = np.random.default_rng(0)
    for v in rng.uniform(-1, 1, 200):
        t = f"{v:.6f}"
        assert pr.parse_f32(t) == np.float32(float(t))


def format_decimal(x: Fraction) -> str:
    """Exact decimal expansion of a dyadic rational."""
    n, d = x.numerator, x.denominator
    k = d.bit_length() - 1
    assert d == 1 << k
    digits = n * 5 ** k
    s = str(digits).rjust(k + 1, "0")
    return s[:-k] + "." + s[-k:] if k else s


def test_crlf_short_blank_and_stable_order():
    a = _mk(2, 1, 710.0, 1.0e-21)
    b = _mk(1, 1, 700.0, 2.0e-21)
    c = _mk(3, 1, 710.0, 3.0e-21)                 # same wavenumber as a, later in the file
    text = a + "\r\n" + "short record\n" + "\n" + b + "\n" + c
    ll, skipped = pr.load_par_text(text)
    assert skipped == 2
    assert list(ll["nu"]) == [700.0, 710.0, 710.0] and list(ll["mol"]) == [1, 2, 3]
