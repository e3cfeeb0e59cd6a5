"""Pins the FP64 oracle (oracle/refcpu.c) before it is trusted as the checker.

Each check is independent of the formula it tests: library golden vectors
(mpmath, scipy), closed forms, brute force with scipy's w(z), and the
invariants of the reference test suite (test/runtests.jl line ranges cited).
"""
import math
import os

import numpy as np
import pytest
from scipy.special import dawsn, erfcx, wofz

import refcpu as o
from irsounderlbl import synthetic as syn
from irsounderlbl.types import (GasSpecies, HITRANLinelist, Sounder, filter_linelist,
                                wavenumber_grid)
from helpers import four_line_fixture

GOLD = os.path.join(os.path.dirname(__file__), "golden", "faddeeva_golden.npz")
L = o.lib()


# ---- Faddeeva -------------------------------------------------------------------
def test_faddeeva_matches_mpmath_golden():
    g = np.load(GOLD)
    w = np.array([o.faddeeva(x, y) for x, y in zip(g["x"], g["y"])])
    ref = g["re"] + 1j * g["im"]
    mod = np.abs(ref)
    re_rel = np.abs(w.real - ref.real) / np.maximum(np.abs(ref.real), 1e-300)
    ok_re = np.abs(ref.real) > 1e-12 * mod          # Re w not a tiny remnant of |w|
    assert re_rel[ok_re].max() < 1e-12
    im_ok = np.abs(ref.imag) > 1e-3 * mod
    im_rel = np.abs(w.imag - ref.imag) / np.maximum(np.abs(ref.imag), 1e-300)
    assert im_rel[im_ok].max() < 1e-11
    assert (np.abs(w - ref) / mod).max() < 1e-12
    # the library routine the reference wraps agrees with the same golden set
    sp = g["scipy_re"] + 1j * g["scipy_im"]
    assert (np.abs(sp - ref) / mod).max() < 1e-13


def test_faddeeva_closed_forms():
    for y in [1e-6, 1e-3, 0.3, 1.0, 5.0, 7.99, 8.01, 50.0, 1e3]:
        assert math.isclose(L.orc_erfcx(y), erfcx(y), rel_tol=1e-13)       # H(0, y) = erfcx(y)
    for x in [0.0, 0.5, 1.7, 3.0, 5.5, 7.9]:
        w = o.faddeeva(x, 0.0)
        assert math.isclose(w.real, math.exp(-x * x), rel_tol=1e-12, abs_tol=1e-300)
        assert math.isclose(w.imag, 2 / math.sqrt(math.pi) * dawsn(x), rel_tol=1e-12, abs_tol=1e-16)
    for x, y in [(200.0, 1.0), (1e4, 3.0)]:                                 # asymptote i/(sqrt(pi) z)
        z = complex(x, y)
        assert abs(o.faddeeva(x, y) - 1j / (math.sqrt(math.pi) * z)) < 1e-4 * abs(1 / z)
    w1, w2 = o.faddeeva(-3.3, 0.2), o.faddeeva(3.3, 0.2)                   # w(-x+iy) = conj w(x+iy)
    assert w1 == w2.conjugate()


# ---- spectroscopy (runtests.jl:321-349) ------------------------------------------
def test_partition_and_broadening():
    assert L.orc_q_ratio(1, 1, 296.0) == pytest.approx(1.0)
    assert L.orc_q_ratio(1, 1, 350.0) < 1.0 < L.orc_q_ratio(1, 1, 250.0)
    assert L.orc_q_ratio(99, 1, 250.0) == pytest.approx(296.0 / 250.0)    # power-law fallback
    line = HITRANLinelist(1, 1, [667.0], [1e-19], [0.060], [0.450], [0.0], [0.76], [0.0])
    c, S, gl, gd = o.line_params(line, 296.0, 1.0)
    assert gl[0] == pytest.approx(0.060, rel=1e-6)
    assert S[0] == pytest.approx(1e-19, rel=1e-6)
    _, _, gl2, _ = o.line_params(line, 296.0, 0.5)
    assert gl2[0] < gl[0]
    _, _, gl3, gd3 = o.line_params(line, 400.0, 1.0)
    assert gl3[0] < gl[0] and gd3[0] > gd[0]
    # gamma_D closed form (broadening.jl:50) and TIPS interpolation (partition.jl:10-18)
    assert gd[0] == pytest.approx(667.0 * 3.58126e-7 * math.sqrt(296.0 / 18.010), rel=1e-15)
    q250, q251 = L.orc_partition(2, 1, 250.0), L.orc_partition(2, 1, 251.0)
    assert L.orc_partition(2, 1, 250.25) == pytest.approx(0.75 * q250 + 0.25 * q251, rel=1e-15)
    assert L.orc_partition_deriv(2, 1, 250.5) == pytest.approx(q251 - q250, rel=1e-15)


# ---- Voigt line sum (runtests.jl:216-297) -----------------------------------------
def _single_line_profile(nu0, gl, gd, grid):
    f = math.sqrt(math.log(2)) / gd
    x = (grid - nu0) * f
    return f / math.sqrt(math.pi) * wofz(x + 1j * gl * f).real


def test_voigt_normalization_symmetry_lorentz():
    g = wavenumber_grid(995.0, 1005.0, 0.001)
    line = HITRANLinelist(2, 1, [1000.0], [1.0], [0.05], [0.0], [0.0], [0.0], [0.0])
    # T = 296, p = 1 -> gamma_L = 0.05; pedestal at cutoff 500 is negligible
    sig = o.voigt_xsec(line, g.nu_min, g.dnu, g.n, 296.0, 1.0, cutoff=500.0, x_far=np.inf)
    assert sig.sum() * g.dnu == pytest.approx(1.0, rel=3e-2)
    mid = (g.n - 1) // 2
    assert int(np.argmax(sig)) == mid
    for k in range(1, 11):
        assert sig[mid - k] == pytest.approx(sig[mid + k], rel=1e-10)
    _, _, gl, gd = o.line_params(line, 296.0, 1.0)
    ref = _single_line_profile(1000.0, gl[0], gd[0], g.nu)
    # parabolic pedestal (voigt.jl:269-271): the Lorentz value at the cutoff
    ref -= gl[0] / (math.pi * (500.0 ** 2 + gl[0] ** 2)) * (2 - ((g.nu - 1000.0) / 500.0) ** 2)
    assert np.max(np.abs(sig - ref)) < 1e-10 * ref.max()
    # Lorentz limit (gamma_D -> 0): V(nu0 + gamma_L) -> 1/(2 pi gamma_L)
    w = wofz(0.05 * 1e6 * math.sqrt(math.log(2)) + 1j * 0.05 * 1e6 * math.sqrt(math.log(2))).real
    v = 1e6 * math.sqrt(math.log(2)) / math.sqrt(math.pi) * w
    assert v == pytest.approx(1.0 / (2 * math.pi * 0.05), rel=2e-2)


def test_brute_force_three_lines_matches_double_loop():
    """Independent O(lines x grid) evaluation with scipy's w(z): pins windowing,
    pedestal, far branch, clamp and line parameters."""
    ll = HITRANLinelist(2, 1, [700.0, 700.4, 701.3], [1e-20, 3e-21, 5e-22], [0.07, 0.06, 0.08],
                        [0.09, 0.1, 0.11], [100.0, 900.0, 2500.0], [0.7, 0.75, 0.6],
                        [-0.002, -0.001, -0.004])
    g = wavenumber_grid(698.0, 703.0, 0.002)
    T, p, cut = 240.0, 0.3, 1.5
    got = o.voigt_xsec(ll, g.nu_min, g.dnu, g.n, T, p, cutoff=cut)
    c, S, gl, gd = o.line_params(ll, T, p)
    nu = g.nu
    ref = np.zeros(g.n)
    for j in range(3):
        f = math.sqrt(math.log(2)) / gd[j]
        y = gl[j] * f
        sn = S[j] * f / math.sqrt(math.pi)
        dv = nu - c[j]
        x = dv * f
        # far-wing Lorentz beyond |x| = 122 (voigt.jl:170-172); scipy w(z) elsewhere
        H = np.where(np.abs(x) > 122.0, y / math.sqrt(math.pi) / (x * x + y * y), wofz(x + 1j * y).real)
        xc = cut * f
        hc = y / math.sqrt(math.pi) / (xc * xc + y * y) if xc > 122 else wofz(xc + 1j * y).real
        contrib = sn * (H - (2 - (dv / cut) ** 2) * hc)
        ref += np.where(np.abs(dv) <= cut, contrib, 0.0)
    ref = np.maximum(ref, 0.0)
    m = ref > 1e-6 * ref.max()
    assert np.max(np.abs(got[m] - ref[m]) / ref[m]) < 1e-7


def test_x_far_gate_lossless_and_live():
    """runtests.jl:247-273."""
    ll = four_line_fixture()
    g = wavenumber_grid(2345.0, 2355.0, 0.001)
    ex = o.voigt_xsec(ll, g.nu_min, g.dnu, g.n, 230.0, 0.2, x_far=np.inf)
    a = o.voigt_xsec(ll, g.nu_min, g.dnu, g.n, 230.0, 0.2, x_far=122.0)
    z = o.voigt_xsec(ll, g.nu_min, g.dnu, g.n, 230.0, 0.2, x_far=0.0)
    pk = ex.max()
    assert np.max(np.abs(a - ex)) < 1e-5 * pk
    m = ex > 1e-6 * pk
    assert np.max(np.abs(a[m] - ex[m]) / ex[m]) < 3e-4
    assert np.max(np.abs(z - ex)) > 1e-3 * pk


def test_dptmin_strict_and_empty():
    ll = four_line_fixture()
    T, p = 230.0, 0.2
    c, S, gl, gd = o.line_params(ll, T, p)
    f = math.sqrt(math.log(2)) / gd
    peak = S * f / math.sqrt(math.pi) * erfcx(gl * f)
    coef = 1.0
    for fac, kept in [(1 - 1e-9, True), (1 + 1e-9, False)]:   # keep iff peak*coef > dptmn
        keep = o.reject_weak_lines(ll, T, p, coef, dptmn=float(peak[1]) * fac)
        assert bool(keep[1]) is kept
        assert list(keep) == list(peak > peak[1] * fac)
    none = o.reject_weak_lines(ll, T, p, coef, dptmn=10 * peak.max())
    assert not none.any()
    g = wavenumber_grid(2345.0, 2355.0, 0.01)
    assert np.all(o.voigt_xsec(ll, g.nu_min, g.dnu, g.n, T, p, keep=none) == 0.0)


def test_dispersive_single_line_closed_form():
    ll = HITRANLinelist(2, 1, [700.0], [1e-20], [0.07], [0.09], [300.0], [0.7], [-0.003], y_lm=[0.02])
    g = wavenumber_grid(699.0, 701.0, 0.001)
    T, p = 250.0, 0.5
    d = o.lm_dispersive(ll, g.nu_min, g.dnu, g.n, T, p, cutoff=25.0)
    qr = L.orc_q_ratio(2, 1, T)
    st = 1e-20 * qr * math.exp(-1.4388 * 300 * (1 / T - 1 / 296)) * \
        (1 - math.exp(-1.4388 * 700 / T)) / (1 - math.exp(-1.4388 * 700 / 296))
    gd = 3.5812e-7 * 700 * math.sqrt(T / 43.98983)
    gl = float(np.float32(0.07)) * (296 / T) ** float(np.float32(0.7)) * p
    f = math.sqrt(math.log(2)) / gd
    c0 = 700.0 + float(np.float32(-0.003)) * p
    x = (g.nu - c0) * f
    w = wofz(x + 1j * gl * f)
    imw = np.where(np.abs(x) > 122, x / math.sqrt(math.pi) / (x * x + (gl * f) ** 2), w.imag)
    ref = 0.02 * p * st * f / math.sqrt(math.pi) * imw
    assert np.max(np.abs(d - ref)) < 1e-9 * np.max(np.abs(ref))


# ---- solver (runtests.jl:192-213, 352-466) -----------------------------------------
def test_planck_roundtrip():
    for nu, T in [(1000.0, 300.0), (645.0, 190.0), (2760.0, 320.0)]:
        assert L.orc_bt(nu, L.orc_planck(nu, T)) == pytest.approx(T, rel=1e-8)
    assert L.orc_bt(700.0, 0.0) == 0.0 and L.orc_bt(700.0, -1.0) == 0.0
    # dB/dT against a central difference
    h = 1e-3
    fd = (L.orc_planck(900.0, 250.0 + h) - L.orc_planck(900.0, 250.0 - h)) / (2 * h)
    assert L.orc_dbdt(900.0, 250.0) == pytest.approx(fd, rel=1e-7)


def test_cim_correction_closed_forms():
    c = L.orc_cim_correction
    assert c(0.03) == pytest.approx(0.005)
    assert c(1e-9) == pytest.approx(1e-9 / 6, rel=1e-6)
    assert c(1.0) == pytest.approx(0.16395341373865302, rel=1e-12)
    assert not math.isclose(c(1.0), L.orc_lit_correction(1.0), rel_tol=1e-3)
    assert c(1000.0) == pytest.approx(1 - 2 / 1000, rel=1e-3)
    assert c(1e7) == pytest.approx(1.0, rel=1e-6)
    assert c(0.06 - 1e-9) == pytest.approx(c(0.06 + 1e-9), rel=1e-3)
    # derivatives vs central differences (runtests.jl:1113-1126)
    for s in [0.2, 1.0, 3.0]:
        h = 1e-6
        for f, fp in [(c, L.orc_cim_correction_deriv), (L.orc_lit_correction, L.orc_lit_correction_deriv)]:
            assert fp(s) == pytest.approx((f(s + h) - f(s - h)) / (2 * h), rel=1e-5)


def _cim_reference(nu, tau, T_lev, T_sfc, T_ave, mu=1.0):
    """Canonical LBLRTM EMIN per-layer form assembled before the T[k+1] weighting
    (an algebra path different from the folded loop under test)."""
    n, Ln = tau.shape
    tr = np.ones((n, Ln + 1))
    for k in range(Ln - 1, -1, -1):
        tr[:, k] = tr[:, k + 1] * np.exp(-tau[:, k] / mu)
    B = lambda T: 1.191042953e-5 * nu ** 3 / np.expm1(1.4387769 * nu / T)
    I = B(T_sfc) * tr[:, 0]
    for k in range(Ln):
        s = tau[:, k] / mu
        t = np.exp(-s)
        f = np.array([L.orc_cim_correction(v) for v in s])
        em = (1 - t) * (B(T_ave[k]) + (B(T_lev[k + 1]) - B(T_ave[k])) * f)
        I = I + em * tr[:, k + 1]
    return I


def test_rte_limits_and_cim_folded_equals_canonical():
    g = wavenumber_grid(645.0, 805.0, 10.0)
    T_lev = np.array([300.0, 290, 280, 270, 260, 250])
    T_ave = np.array([296.0, 285, 274, 263, 252])
    tau = np.full((g.n, 5), 0.7)
    I = o.schwarzschild_rte(g.nu_min, g.dnu, tau, T_lev, 305.0, 1.0, 1.0, 1, T_ave)
    assert np.allclose(I, _cim_reference(g.nu, tau, T_lev, 305.0, T_ave), rtol=1e-12, atol=0)
    B = lambda T: 1.191042953e-5 * g.nu ** 3 / np.expm1(1.4387769 * g.nu / T)
    for sf in (0, 1):
        # the Toon case of the reference test is isothermal above the surface (runtests.jl:357-367)
        Tl = T_lev if sf == 1 else np.full(6, 280.0)
        thick = o.schwarzschild_rte(g.nu_min, g.dnu, np.full((g.n, 5), 100.0), Tl, 305.0, 1.0, 1.0, sf, T_ave)
        assert np.allclose(thick, B(Tl[-1]), rtol=1e-3)
        thin = o.schwarzschild_rte(g.nu_min, g.dnu, np.full((g.n, 5), 1e-10), T_lev, 305.0, 1.0, 1.0, sf, T_ave)
        assert np.allclose(thin, B(305.0), rtol=1e-4)
        iso = o.schwarzschild_rte(g.nu_min, g.dnu, np.full((g.n, 5), 2.0), np.full(6, 280.0), 280.0,
                                  1.0, 1.0, sf, np.full(5, 280.0))
        assert np.allclose(iso, B(280.0), rtol=1e-6)
        # grey surface, isothermal column, no space-incident radiance: closed form
        # I = B (1 - (1 - eps) Tr^2) with Tr the surface-to-space transmittance
        iso2 = o.schwarzschild_rte(g.nu_min, g.dnu, np.full((g.n, 5), 0.4), np.full(6, 280.0), 280.0,
                                   0.7, 0.9, sf, np.full(5, 280.0))
        tr = math.exp(-2.0 / 0.7)
        assert np.allclose(iso2, B(280.0) * (1 - 0.1 * tr * tr), rtol=1e-12)
    with pytest.raises(ValueError):
        o.schwarzschild_rte(g.nu_min, g.dnu, tau, T_lev, 305.0, 0.0, 1.0, 1, T_ave)


# ---- ILS (runtests.jl:469-503, 1837-1852) -----------------------------------------
@pytest.mark.parametrize("apod", [0, 1, 2, 3])
def test_ils_kernel_properties(apod):
    dv, k = o.ils_kernel(0.25, 2.0, 0.5, apod)
    assert k.sum() * 0.25 == pytest.approx(1.0, rel=1e-2)
    assert np.allclose(k, k[::-1], rtol=1e-6)
    assert int(np.argmax(k)) == (len(k) - 1) // 2
    with pytest.raises(ValueError):
        o.ils_kernel(0.25, 2.0, 0.5, 7)


def test_norton_beer_values():
    for s, c0 in [(1, 0.384093), (2, 0.152442), (3, 0.045335)]:
        assert L.orc_norton_beer(0.0, s) == pytest.approx(1.0)
        assert L.orc_norton_beer(1.0, s) == pytest.approx(c0)
    assert L.orc_norton_beer(1.5, 2) == 0.0


def test_direct_ils_equals_fft():
    g = wavenumber_grid(699.0, 705.0, 0.002)
    rng = np.random.default_rng(1)
    spec = 80 + 5 * np.sin(g.nu * 3) + rng.normal(0, 1, g.n)
    _, k = o.ils_kernel(0.002, 2.0, 0.5, 0)
    idx = np.arange(g.n)
    direct = o.apply_ils_at(spec, 0.002, k, idx)
    half = (len(k) - 1) // 2
    n = g.n + len(k) - 1
    conv = np.fft.irfft(np.fft.rfft(spec, n) * np.fft.rfft(k, n), n)[half:half + g.n] * 0.002
    assert np.max(np.abs(direct - conv)) < 1e-12 * np.max(np.abs(conv))


# ---- grid, layers, geometry --------------------------------------------------------
def test_julia_range_nodes():
    n, first = o.grid(629.0, 0.001, 2776.0)
    assert n == 2147001
    assert L.orc_grid_node(629.0, 0.001, 16000) == 645.0
    assert L.orc_grid_node(629.0, 0.001, 16000 + 250 * 8460) == 2760.0
    assert L.orc_grid_node(629.0, 0.001, 123) == float("629.123")
    assert wavenumber_grid(629.0, 2776.0, 0.001).n == n


def test_layer_properties_limits():
    prof = syn.base_profile(True)
    lp = o.layer_properties(prof)
    assert len(lp["p_mid"]) == 49 and np.all(lp["dp"] > 0)
    # well-mixed limit: CG VMR equals the constant, CG pressure the arithmetic mean
    assert np.allclose(lp["vmr_cg"][GasSpecies.CO2], 4.15e-4, rtol=1e-15)
    assert np.allclose(lp["p_cg"][GasSpecies.CO2], lp["p_mid"], rtol=1e-15)
    # TBAR = Int T dp / dp for T linear in log p (layers.jl:93-106), by quadrature
    for k in [0, 5, 30]:
        p1, p2 = prof.pressure[k], prof.pressure[k + 1]
        T1, T2 = prof.temperature[k], prof.temperature[k + 1]
        pp = np.linspace(p2, p1, 20001)
        TT = T1 + (T2 - T1) * np.log(pp / p1) / np.log(p2 / p1)
        assert lp["T_ave"][k] == pytest.approx(np.trapezoid(TT, pp) / (p1 - p2), rel=1e-9)
    # CG column VMR = Int vmr dp / dp for the power-law VMR, by quadrature
    k = 3
    v1, v2 = prof.vmr[GasSpecies.H2O][k], prof.vmr[GasSpecies.H2O][k + 1]
    p1, p2 = prof.pressure[k], prof.pressure[k + 1]
    pp = np.linspace(p2, p1, 20001)
    vv = v1 * (pp / p1) ** (math.log(v2 / v1) / math.log(p2 / p1))
    assert lp["vmr_cg"][GasSpecies.H2O][k] == pytest.approx(np.trapezoid(vv, pp) / (p1 - p2), rel=1e-8)
    assert lp["p_cg"][GasSpecies.H2O][k] == pytest.approx(np.trapezoid(pp * vv, pp) / np.trapezoid(vv, pp), rel=1e-8)


def test_scan_angle_geometry():
    assert L.orc_scan_to_local_zenith(48.33, 817.0, 6371.0) == pytest.approx(57.46, abs=0.05)
    assert L.orc_scan_to_local_zenith(80.0, 817.0, 6371.0) < 0     # beyond the limb


def test_forward_small_semitransparent():
    """A weak-line forward run (sensitive to every layer) is finite, physical,
    and DPTMIN-lossless at 1e-6 within the 0.6 mK budget the reference cites."""
    ll = syn.linelist_config0()
    weak = HITRANLinelist(ll.mol_id, ll.iso_id, ll.wavenumber, ll.intensity * 1e-5, ll.air_broad,
                          ll.self_broad, ll.lower_energy, ll.temp_depend, ll.pressure_shift)
    sub = filter_linelist(weak, 680.0, 750.0)
    prof = syn.base_profile(False)
    snd = Sounder.make(705.0, 725.0, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    R, BT, _ = o.forward(prof, {GasSpecies.CO2: sub}, snd, internal_dnu=0.002)
    R0, BT0, _ = o.forward(prof, {GasSpecies.CO2: sub}, snd, internal_dnu=0.002, dptmn=0.0)
    assert np.all(np.isfinite(BT)) and BT.min() > 200 and BT.max() < 290
    assert np.max(np.abs(BT - BT0)) < 6e-4


def test_channel_od_helper_cim_closed_forms():
    """The FP64 channel-OD restatement's f_cim (tests/helpers.py) at the
    reference's closed-form values (runtests.jl:387-466)."""
    from helpers import _fcim
    assert abs(float(_fcim(0.03)) - 0.005) < 1e-15
    assert abs(float(_fcim(1.0)) - 0.16395341373865302) < 1e-12
    s = np.array([0.0599999, 0.0600001])
    assert abs(_fcim(s)[1] - _fcim(s)[0]) < 1e-5           # branch continuity
    assert abs(float(_fcim(200.0)) - (1.0 - 2.0 / 200.0)) < 1e-12


@pytest.mark.parametrize("T_method,tm_code", [("logp_at_pcg", 0), ("mass_weighted", 1)])
def test_library_layer_properties_match_oracle(T_method, tm_code):
    """The library's host layer code (irs_layer_properties, behind forward_model and
    the emulator features) against the oracle's restatement of layers.jl: p_mid, dp,
    TBAR and the Curtis-Godson vmr, p and T per species, on the config-0 profile with
    O3 and on perturbed profiles (level-by-level H2O included)."""
    from irsounderlbl.api import layer_properties
    profs = [syn.base_profile(True)] + [p for p, _, _ in syn.perturbed_profiles(
        6, seed=17, sigma_T=4.0, sigma_h2o=0.6, sigma_o3=0.2, co2_range=(390e-6, 430e-6), sigma_lnq=0.2)]
    for prof in profs:
        got = layer_properties(prof, T_method)
        ref = o.layer_properties(prof, tm_code)
        for k in ("p_mid", "dp", "T_ave"):
            np.testing.assert_allclose(got[k], ref[k], rtol=1e-13, atol=0, err_msg=k)
        for sp in prof.vmr:
            for k in ("vmr_cg", "p_cg", "T_cg"):
                np.testing.assert_allclose(got[k][sp], ref[k][sp], rtol=1e-12, atol=0, err_msg=f"{k} {sp}")
