"""Pins of the relaxation-matrix line-mixing oracle (oracle/relmat_ref.py): the
reference's own testsets (test/runtests.jl:714-945: no-coupling limits, the x_far
gate, the band-strength cutoff, CH4 plumbing), the sum rule and detailed balance
the W build imposes, and first-order perturbation theory (the eigenmode sum
reduces to the Rosenkranz dispersive term as p -> 0)."""
import math

import numpy as np
import pytest

import relmat_ref as R
from irsounderlbl.relmat import (HITRANRelmatData, RelmatBand, RelmatLine, W0B0Table, band_eff_strength,
                                 default_vpw_whitelist, lm_iso_abund, lm_mass_amu)
from irsounderlbl.types import GasSpecies

RL = RelmatLine


def _grid(a, b, d):
    n = int(round((b - a) / d)) + 1
    return a + np.arange(n) * d


def test_band_modes_no_coupling_limit():
    l1 = RL(1, 700.0, 0, 0, 0.07, 0.7, 0.0, 200.0, 1.0, 0.5, 1.0)
    l2 = RL(1, 710.0, 2, 0, 0.07, 0.7, 0.0, 220.0, 1.0, 0.45, 1.0)
    band = RelmatBand("TEST01", 1, 1, 1, 700.0, 710.0, [l1, l2])
    T, p = 296.0, 0.5
    lam, A, f = R.band_modes(band, W0B0Table(), T, p)
    assert len(lam) == 2 and len(A) == 2
    o = np.argsort(lam.real)
    lam, A = lam[o], A[o]
    assert lam[0].real == pytest.approx(700.0, abs=1e-10) and lam[1].real == pytest.approx(710.0, abs=1e-10)
    assert lam[0].imag == pytest.approx(p * 0.07, abs=1e-10) and lam[1].imag == pytest.approx(p * 0.07, abs=1e-10)
    for a, l in zip(A, (l1, l2)):
        S = l.DipoT ** 2 * l.PopuT0 * l.nu * (1 - math.exp(-1.4388 * l.nu / T))
        assert a.real == pytest.approx(S, rel=1e-12)
        assert abs(a.imag) < 1e-12 * abs(a.real)
    gd = 3.5812e-7 * 705.0 * math.sqrt(T / 43.98983)
    assert f == pytest.approx(math.sqrt(math.log(2.0)) / gd, rel=1e-12)


def test_vpw_evaluator_equals_voigt_without_coupling():
    from scipy.special import wofz
    l1 = RL(1, 700.00, 0, 0, 0.07, 0.7, 0.0, 200.0, 1.0, 0.5, 1.0)
    l2 = RL(1, 700.05, 2, 0, 0.07, 0.7, 0.0, 220.0, 1.0, 0.45, 1.0)
    band = RelmatBand("TEST02", 1, 1, 1, 700.0, 700.05, [l1, l2])
    T, p = 296.0, 0.5
    modes = R.band_modes(band, W0B0Table(), T, p)
    nu = _grid(699.5, 700.5, 0.005)
    s = R.vpw_band_xsec(nu, modes, cutoff=10.0, x_far=np.inf)
    f = modes[2]
    ref = np.zeros(len(nu))
    for l in (l1, l2):
        S = l.DipoT ** 2 * l.PopuT0 * l.nu * (1 - math.exp(-1.4388 * l.nu / T))
        ref += S * f / math.sqrt(math.pi) * wofz((nu - l.nu) * f + 1j * 0.07 * p * f).real
    assert np.max(np.abs(s - ref)) / np.max(np.abs(ref)) < 1e-12
    assert np.all(R.vpw_band_xsec(_grid(800.0, 800.5, 0.005), modes, cutoff=10.0) == 0.0)


def _coupled_q():
    l1 = RL(1, 667.00, 20, 0, 0.03, 0.7, 0.0, 300.0, 1.0, 1.0, 1.0)
    l2 = RL(1, 667.05, 18, 0, 0.03, 0.7, 0.0, 260.0, 1.0, 0.8, 1.0)
    band = RelmatBand("TESTQ", 1, 1, 1, 667.0, 667.05, [l1, l2])
    wt = W0B0Table()
    wt.qq[(20, 18)] = (math.log(0.03), 0.5)
    return band, wt


def test_far_wing_shortcut_is_lossless():
    band, wt = _coupled_q()
    T, p, cut = 250.0, 0.1, 25.0
    nu = _grid(665.0, 669.0, 0.002)
    modes = R.band_modes(band, wt, T, p)
    assert np.any(modes[1].imag != 0.0)
    ex = R.vpw_band_xsec(nu, modes, cut, np.inf)
    opt = R.vpw_band_xsec(nu, modes, cut, R.X_FAR)
    lor = R.vpw_band_xsec(nu, modes, cut, 0.0)
    pk = np.max(np.abs(ex))
    assert np.max(np.abs(opt - ex)) < 1e-5 * pk
    assert np.max(np.abs(lor - ex)) > 1e-3 * pk
    Y = R.calc_Y(band, wt, T)
    lines = R.dispersive_lines(band, Y, T, p)
    ey = R.dispersive_eval(nu, lines, cut, np.inf)
    oy = R.dispersive_eval(nu, lines, cut, R.X_FAR)
    ly = R.dispersive_eval(nu, lines, cut, 0.0)
    pky = np.max(np.abs(ey))
    assert pky > 0
    assert np.max(np.abs(oy - ey)) < 1e-5 * pky
    assert np.max(np.abs(ly - ey)) > 1e-3 * pky


def test_band_strength_cutoff():
    s1 = RL(1, 700.00, 20, 0, 0.03, 0.7, 0.0, 300.0, 1.0, 1.0, 1.0)
    s2 = RL(1, 700.05, 18, 0, 0.03, 0.7, 0.0, 260.0, 1.0, 0.8, 1.0)
    w1 = RL(1, 720.00, 4, 0, 0.03, 0.7, 0.0, 300.0, 1.0e-3, 1.0, 1.0e-3)
    w2 = RL(1, 720.05, 2, 0, 0.03, 0.7, 0.0, 260.0, 1.0e-3, 0.8, 1.0e-3)
    bS = RelmatBand("STRONG", 1, 1, 1, 700.0, 700.05, [s1, s2])
    bW = RelmatBand("WEAK", 1, 1, 1, 720.0, 720.05, [w1, w2])
    wt = W0B0Table()
    wt.qq[(20, 18)] = (math.log(0.03), 0.5)
    wt.qq[(4, 2)] = (math.log(0.03), 0.5)
    both = HITRANRelmatData([bS, bW], {(1, 1): wt})
    strong = HITRANRelmatData([bS], {(1, 1): wt})
    sS, sW = R.band_strength(bS), R.band_strength(bW)
    assert sS > sW and band_eff_strength(bS) == sS and band_eff_strength(bW) == sW
    thr = math.sqrt(sS * sW)
    T, p = 250.0, 0.5
    nu = _grid(695.0, 725.0, 0.01)
    full = R.lm_dispersive_correction(nu, both, T, p, min_band_strength=0.0)
    cut = R.lm_dispersive_correction(nu, both, T, p, min_band_strength=thr)
    st = R.lm_dispersive_correction(nu, strong, T, p, min_band_strength=0.0)
    assert np.array_equal(cut, st)
    assert np.max(np.abs(full - cut)) > 0.0
    bW2 = RelmatBand("WEAK2", 1, 1, 2, 720.0, 720.05, [w1, w2])
    assert R.band_strength(bW2) == sW


def test_species_generalisation_ch4():
    assert lm_mass_amu(6, 1) == 16.031300 and lm_iso_abund(6, 1) == 0.988274
    assert lm_mass_amu(2, 1) == 43.98983 and lm_mass_amu(99, 1) == 44.0
    l1 = RL(1, 1306.00, 6, 0, 0.05, 0.75, 0.0, 220.0, 1.0, 1.0, 1.0)
    l2 = RL(1, 1306.05, 5, 0, 0.05, 0.75, 0.0, 200.0, 1.0, 0.9, 1.0)
    b = RelmatBand("CH4Q", 1, 1, 1, 1306.0, 1306.05, [l1, l2], molecule=6)
    wt = W0B0Table()
    wt.qq[(6, 5)] = (math.log(0.04), 0.5)
    rm = HITRANRelmatData([b], {(1, 1): wt}, GasSpecies.CH4)
    s = R.lm_dispersive_correction(_grid(1304.0, 1308.0, 0.01), rm, 250.0, 0.5)
    assert np.all(np.isfinite(s)) and np.max(np.abs(s)) > 0.0


def _ladder(n=6, isot=1, seed=0):
    """A Q-branch ladder: n lines 0.02 cm^-1 apart, every pair coupled."""
    rng = np.random.default_rng(seed)
    Js = [2 * k + 2 for k in range(n)]
    lines = [RL(isot, 667.0 + 0.02 * k, Js[k], 0, 0.07 + 0.01 * rng.random(), 0.7, 0.0, 50.0 * k,
                1.0 + 0.2 * rng.random(), 0.5 + 0.3 * rng.random(), 1.0 + 0.3 * rng.random()) for k in range(n)]
    band = RelmatBand("LADDER", 1, 1, isot, 667.0, 667.0 + 0.02 * (n - 1), lines)
    wt = W0B0Table()
    for a in Js:
        for b in Js:
            if b < a:
                wt.qq[(a, b)] = (math.log(0.01 + 0.002 * (a - b)), 0.6)
    return band, wt


def test_w_matrix_sum_rule_and_detailed_balance():
    band, wt = _ladder()
    for T in (200.0, 296.0):
        W, rho = R.build_W(band, wt, T)
        D0 = np.abs(band.Dipo0)
        col = D0 @ W                           # sum_j |D0_j| W_ji for every i
        scale = D0 @ np.abs(W)
        S_eff = band.nu * rho * band.DipoT ** 2
        last = np.argmin(S_eff)                # the weakest line has no lower partner to renormalise
        for i in range(len(col)):
            if i != last:
                assert abs(col[i]) < 1e-12 * scale[i]
        off = ~np.eye(len(rho), dtype=bool)
        lhs = (W * rho[None, :])[off]
        rhs = (W.T * rho[:, None])[off]
        np.testing.assert_allclose(lhs, rhs, rtol=1e-12, atol=0)
        np.testing.assert_allclose(np.diag(W), band.gV_air * (296.0 / T) ** 0.7, rtol=1e-6)


def test_eigenmodes_reduce_to_first_order_mixing_at_low_pressure():
    """As p -> 0 the eigenmode perturbation approaches sum S_k p Y_k Im w(z_k)
    (first-order perturbation theory of D + i p W), with Y from the same W."""
    band, wt = _ladder(4)
    T = 250.0
    nu = _grid(666.5, 667.6, 0.0005)
    err = []
    for p in (1e-4, 1e-5):
        vpw = R.vpw_band_perturbation(nu, band, wt, T, p, keep_threshold=0.0, x_far=np.inf)
        Y = R.calc_Y(band, wt, T)
        vpy = R.dispersive_eval(nu, R.dispersive_lines(band, Y, T, p), 25.0, np.inf)
        err.append(np.max(np.abs(vpw - vpy)) / np.max(np.abs(vpy)))
    # the remainder is second order (pole shifts ~ p^2 W^2 / dnu): relative error ~ p
    assert err[0] < 0.01
    assert 0.05 < err[1] / err[0] < 0.2


def test_default_whitelist():
    bands = []
    for k, (iso, s) in enumerate(((1, 1.0), (1, 3.0), (2, 5.0), (1, 2.0), (2, 0.5))):
        bands.append(RelmatBand(f"B{k}", 1, 1, iso, 700.0 + 10 * k, 705.0 + 10 * k,
                                [RL(iso, 700.0 + 10 * k, 2, 0, 0.07, 0.7, 0.0, 0.0, 1.0, 1.0, s)]))
    rm = HITRANRelmatData(bands, {})
    assert default_vpw_whitelist(rm, n_top=2) == {"B1", "B3"}       # abundance-weighted
    assert default_vpw_whitelist(rm, n_top=1, isotopes=[1, 2]) == {"B1", "B2"}
    assert default_vpw_whitelist(rm, n_top=5, nu_window=(718.0, 726.0)) == {"B2"}
