"""HITRAN .par ingest on the device (irs_par_parse / irs_par_load) against the
oracle restatement of the reference's reader (oracle/par_ref.py): bit-exact
fields, order and skip counts; the cache semantics of test/runtests.jl:948-997;
a .par round trip through forward_model."""
import os
import time

import numpy as np
import pytest

import par_ref as pr
from irsounderlbl import synthetic as syn
from irsounderlbl.hitran import (clear_linelist_cache, format_hitran_par, linelist_cache_dir, load_hitran_par,
                                 load_linelist, parse_hitran_par, parse_hitran_par_bytes, write_hitran_par)
from irsounderlbl.types import GasSpecies, Sounder, filter_linelist

pytestmark = pytest.mark.gpu

FIELDS = (("mol", "mol_id"), ("iso", "iso_id"), ("nu", "wavenumber"), ("S", "intensity"), ("A", "a_coeff"),
          ("gamma_air", "air_broad"), ("gamma_self", "self_broad"), ("elower", "lower_energy"),
          ("n_air", "temp_depend"), ("delta_air", "pressure_shift"))


def _same(ll, ref):
    assert len(ll) == len(ref["nu"])
    for a, b in FIELDS:
        x, y = getattr(ll, b), ref[a]
        assert x.dtype == y.dtype, (a, x.dtype, y.dtype)
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), a        # bit-exact, -0.0 included


def _mk(mol, iso, nu, S, ga="0.05", gs="0.07", E="0.0", n="0.75", d="0.0", A="0.0"):
    return (str(mol).rjust(2) + str(iso) + str(nu).rjust(12) + str(S).rjust(10) + A.rjust(10) + ga.rjust(5) +
            gs.rjust(5) + E.rjust(10) + n.rjust(4) + d.rjust(8))


def _records(n_lines, seed=1):
    """A generated .par text from the synthetic line lists (HITRAN Fortran formats)."""
    lls = syn.linelists_config1(windows=True)
    parts = []
    rng = np.random.default_rng(seed)
    for sp, ll in lls.items():
        idx = np.sort(rng.choice(len(ll), size=min(n_lines, len(ll)), replace=False))
        sub = ll.subset(idx)
        sub.a_coeff = rng.uniform(1e-6, 10.0, len(sub)) * 10.0 ** rng.integers(-3, 2, len(sub))
        parts.extend(format_hitran_par(sub, i) for i in range(len(sub)))
    return parts


def test_generated_file_matches_oracle_bit_exact(ctx):
    recs = _records(4000)
    text = "\n".join(recs) + "\n"
    got, skipped = parse_hitran_par_bytes(text.encode())
    ref, rskip = pr.load_par_text(text)
    assert skipped == rskip == 0 and len(got) == len(recs)
    _same(got, ref)
    assert np.all(np.diff(got.wavenumber) >= 0)


def test_edge_records_match_oracle(ctx):
    """CRLF endings, a short record, a blank line, iso codes the parser rejects
    ('A'), an Int8 out of range, fields outside the device grammar that the host
    parser takes (inf, a 3-digit exponent), a negative zero, Fortran .0602-style
    fields, duplicate wavenumbers out of order (stable), no final newline."""
    recs = [
        _mk(2, 1, "710.000000", "1.000E-21"),
        _mk(1, 1, "700.500000", "2.500E-25", ga=".0602", gs=".468", d="-.001234", n="0.69"),
        "short",
        "",
        _mk(2, "A", "705.0", "1.0E-22"),                    # iso 'A': skipped
        _mk(200, 1, "705.0", "1.0E-22"),                    # 3 characters in the 2-wide field: malformed
        _mk(3, 1, "710.000000", "3.000E-21", A="inf"),      # host parser: strtod reads inf
        _mk(2, 2, "699.999999", "4.2E-101"),                # exponent beyond the device's exact range
        _mk(2, 3, "701.000000", "1.0E-23", d="-0.0"),
        _mk(2, 4, "710.000000", "5.000E-21", E="12345.6789"),
        _mk(2, 1, " 7.0e2", "1e-20"),                       # lower-case exponent, short forms
        _mk(2, 1, "702.0", "1.0E-20", ga="1 0.0"),          # space inside a field: skipped
    ]
    text = "\r\n".join(recs[:2]) + "\n" + "\n".join(recs[2:])
    got, skipped = parse_hitran_par_bytes(text.encode())
    ref, rskip = pr.load_par_text(text)
    print(f"edge records: {len(got)} kept, {skipped} skipped")
    assert skipped == rskip == 5
    _same(got, ref)
    assert list(got.mol_id[got.wavenumber == 710.0]) == [2, 3, 2]          # file order among equal nu
    assert np.isinf(got.a_coeff[(got.wavenumber == 710.0) & (got.mol_id == 3)][0])
    # the window and a single record
    w, _ = parse_hitran_par_bytes(text.encode(), 700.0, 705.0)
    wr, _ = pr.load_par_text(text, 700.0, 705.0)
    _same(w, wr)
    line = parse_hitran_par(recs[1])
    assert line.wavenumber == 700.5 and line.air_broad == np.float32(0.0602) and line.mol_id == 1
    with pytest.raises(ValueError):
        parse_hitran_par(recs[4])


def test_large_file_parse_rate(ctx, tmp_path):
    """300 000 records (the headline line lists as a .par file, 48 MB): every
    field equals Python's reading of it (float for Float64, float32(float) for
    the Float32 fields, exact at these widths); parse rate reported."""
    lls = syn.linelists_config1()
    path = str(tmp_path / "all.par")
    recs = []
    for ll in lls.values():
        ll.a_coeff = np.full(len(ll), 1.234e-2)
        recs.extend(format_hitran_par(ll, i) for i in range(len(ll)))
    data = ("\n".join(recs) + "\n").encode()
    with open(path, "wb") as f:
        f.write(data)
    parse_hitran_par_bytes(data[:100000])                   # warm-up
    t0 = time.perf_counter()
    got, skipped = parse_hitran_par_bytes(data)
    dt = time.perf_counter() - t0
    print(f"parsed {len(got)} records ({len(data) / 1e6:.1f} MB) in {dt * 1e3:.1f} ms from host memory: "
          f"{len(data) / dt / 1e9:.2f} GB/s, {len(got) / dt / 1e6:.1f} M records/s")
    assert skipped == 0 and len(got) == 300000
    cols = [(r[3:15], r[15:25], r[35:40], r[59:67]) for r in recs]
    order = np.argsort(np.array([float(c[0]) for c in cols]), kind="stable")
    nu = np.array([float(cols[i][0]) for i in order])
    S = np.array([float(cols[i][1]) for i in order])
    ga = np.array([np.float32(float(cols[i][2])) for i in order])
    da = np.array([np.float32(float(cols[i][3])) for i in order])
    assert np.array_equal(got.wavenumber, nu) and np.array_equal(got.intensity, S)
    assert np.array_equal(got.air_broad, ga) and np.array_equal(got.pressure_shift, da)
    ll2 = load_hitran_par(path, nu_min=700.0, nu_max=710.0)
    assert np.array_equal(ll2.wavenumber, nu[(nu >= 700.0) & (nu <= 710.0)])


def test_linelist_cache_like_the_reference(ctx, tmp_path, monkeypatch):
    """test/runtests.jl:955-997: first load parses and writes one cache file,
    the second reads it back exactly; windowing in memory; cache=False and
    rebuild=True agree; appending a record gives a new key; clearing removes both."""
    mk = lambda mol, iso, nu, S: _mk(mol, iso, repr(nu), repr(S).replace("e-", "e-"))
    par = str(tmp_path / "synthetic.par")
    with open(par, "w") as f:
        for nu, S in ((700.0, 1.0e-21), (710.0, 2.0e-21), (720.0, 3.0e-21)):
            f.write(mk(2, 1, nu, S) + "\n")
    cachedir = str(tmp_path / "cache")
    monkeypatch.setenv("IRSOUNDER_LINELIST_CACHE", cachedir)
    assert linelist_cache_dir() == cachedir
    n_cache = lambda: sum(f.endswith(".irsll") for f in os.listdir(cachedir))
    ll1 = load_linelist(par)
    assert len(ll1) == 3 and n_cache() == 1
    ll2 = load_linelist(par)
    _same(ll2, {a: getattr(ll1, b) for a, b in FIELDS})
    w = load_linelist(par, nu_min=705.0, nu_max=715.0)
    assert len(w) == 1 and w.wavenumber[0] == 710.0
    _same(load_linelist(par, cache=False), {a: getattr(ll1, b) for a, b in FIELDS})
    _same(load_linelist(par, rebuild=True), {a: getattr(ll1, b) for a, b in FIELDS})
    time.sleep(0.01)
    with open(par, "a") as f:
        f.write(mk(2, 1, 730.0, 4.0e-21) + "\n")
    assert len(load_linelist(par)) == 4 and n_cache() == 2
    assert clear_linelist_cache() == 2 and n_cache() == 0
    with pytest.raises(FileNotFoundError):
        load_linelist(str(tmp_path / "absent.par"))


def test_multi_isotopologue_merge_and_missing_file(ctx, tmp_path, monkeypatch):
    monkeypatch.setenv("IRSOUNDER_LINELIST_CACHE", str(tmp_path / "cache"))
    base = str(tmp_path / "co2_620_825")
    with open(base + ".par", "w") as f:
        f.write(_mk(2, 1, "700.0", "1.0E-21") + "\n" + _mk(2, 1, "720.0", "1.0E-21") + "\n")
    with open(base + "_iso2.par", "w") as f:
        f.write(_mk(2, 2, "710.0", "1.0E-22") + "\n")
    ll = load_linelist(base, [1, 2])
    assert list(ll.wavenumber) == [700.0, 710.0, 720.0] and list(ll.iso_id) == [1, 2, 1]
    with pytest.raises(FileNotFoundError, match="_iso3.par"):
        load_linelist(base, [1, 2, 3])


def test_par_roundtrip_through_forward_model(ctx, tmp_path):
    """Line lists written as .par, read on the device: the same fields as the
    oracle's reading and the same brightness temperatures, bit for bit."""
    from irsounderlbl.api import forward_model
    lo, hi = 1205.0, 1213.0
    lls = syn.linelists_config1(windows=True)
    sub = {k: filter_linelist(v, lo - 42.0, hi + 42.0) for k, v in lls.items()}
    read, oracle = {}, {}
    for sp, ll in sub.items():
        path = str(tmp_path / f"{sp.name}.par")
        write_hitran_par(path, ll)
        read[sp] = load_hitran_par(path)
        ref, _ = pr.load_par_text(open(path).read())
        _same(read[sp], ref)
        from irsounderlbl.types import HITRANLinelist
        oracle[sp] = HITRANLinelist(ref["mol"], ref["iso"], ref["nu"], ref["S"], ref["gamma_air"],
                                    ref["gamma_self"], ref["elower"], ref["n_air"], ref["delta_air"], a_coeff=ref["A"])
    snd = Sounder.make(lo, hi, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    prof = syn.base_profile(True)
    _, _, bt = forward_model(prof, read, sounder=snd, ctx=ctx, apply_continuum=False)
    _, _, bto = forward_model(prof, oracle, sounder=snd, ctx=ctx, apply_continuum=False)
    _, _, bt0 = forward_model(prof, sub, sounder=snd, ctx=ctx, apply_continuum=False)
    print(f".par round trip: BT {bt.min():.2f}-{bt.max():.2f} K, vs the unrounded lists {np.abs(bt - bt0).max():.2e} K")
    assert np.array_equal(bt, bto)
