"""Pins the oracle's continuum restatement (oracle/refcpu.c, continuum section)
against the reference's continuum tests (test/runtests.jl:640-688) and against
what the MT-CKD definitions fix independently of the code:

* at the reference state (296 K, 1013 hPa) the temperature and density
  ratios are 1, so k = C(nu) * rad(nu) * n with C read straight from the
  coefficient table (the bundled CSV rows, via the generated header);
* the self term scales as vmr^2 and the foreign term as vmr (1 - vmr);
* table values are linear between knots (a midpoint is the knot mean);
* the radiation term has closed forms in its three branches;
* dz is the hydrostatic R T / (M g) dp / p;
* CIA: linear in T between bracketing blocks, clamped outside, zero off a
  block's wavenumber range, negative sigma clamped to 0, and zero without a
  table (the reference's absent-data behaviour).
The continuum Jacobian terms are pinned by central differences of the oracle
forward model (the reference's own method for them).
"""
import math
import os
import re

import numpy as np
import pytest

import refcpu as o
from irsounderlbl import synthetic as syn
from irsounderlbl.types import AtmosphericProfile, GasSpecies, Sounder, filter_linelist

KB = 1.380649e-23
HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "irsounderlbl.jl_amd", "data",
                   "mtckd_table.h")


def _table(name):
    txt = open(HDR).read()
    body = txt.split(f"static const double {name}")[1].split("};")[0]
    rows = re.findall(r"\{([^{}]+)\}", body)
    return np.array([[float(v) for v in r.split(",")] for r in rows])


H2O = _table("irs_ckd_h2o")
CO2 = _table("irs_ckd_co2")


def rad(nu, T):
    x = 1.4387769 * nu / T
    return np.where(x <= 0.01, 0.5 * x * nu, np.where(x <= 10.0, nu * np.tanh(0.5 * x), nu))


def test_reference_continuum_invariants():
    """runtests.jl:640-688."""
    nu = 645.0 + 0.25 * np.arange(8461)
    k = o.h2o_continuum(nu, 0.01, 1013.25, 296.0)
    assert len(k) == 8461 and np.all(k >= 0)
    c = o.co2_continuum(nu, 4.15e-4, 1013.25, 296.0)
    assert np.all(c >= 0)
    i700, i790 = np.argmin(abs(nu - 700)), np.argmin(abs(nu - 790))
    assert c[i700] > c[i790] > 0
    T1, T2 = 296.0, 246.0
    k1 = o.co2_continuum(nu, 4.15e-4, 1013.25, T1)
    k2 = o.co2_continuum(nu, 4.15e-4, 1013.25, T2)
    base = lambda v, T: rad(v, T) / T ** 2
    i2400, i714 = np.argmin(abs(nu - 2400)), np.argmin(abs(nu - 714))
    r_bh = (k1[i2400] / base(nu[i2400], T1)) / (k2[i2400] / base(nu[i2400], T2))
    assert r_bh == pytest.approx((T1 / T2) ** 0.73, rel=1e-3)
    r_15 = (k1[i714] / base(nu[i714], T1)) / (k2[i714] / base(nu[i714], T2))
    assert r_15 == pytest.approx(1.0, rel=1e-3)
    # CIA absent: zeros (documented graceful degradation)
    o.cia_set(0, [])
    assert np.all(o.cia_k(0, nu[::100], 4.15e-4, 1013.25, 296.0) == 0.0)


def test_radiation_term_branches():
    L = o._declare_cont()
    for nu, T in ((0.001, 300.0), (1.0, 300.0), (700.0, 250.0), (2600.0, 200.0), (5000.0, 250.0)):
        x = 1.4387769 * nu / T
        want = 0.5 * x * nu if x <= 0.01 else (nu * math.tanh(0.5 * x) if x <= 10 else nu)
        assert L.orc_radfn(nu, T) == pytest.approx(want, rel=1e-14)


def test_h2o_continuum_at_reference_state_from_table():
    """At 296 K and 1013 hPa: k = (Cs v + Cf (1 - v)) rad n_H2O on a table knot."""
    p, T = 1013.0, 296.0
    for nu in (700.0, 1200.0, 2500.0):
        row = H2O[int(nu / 10)]
        assert row[0] == nu
        for v in (1e-3, 0.02):
            n = v * p * 100.0 / (KB * T) * 1e-6
            want = (row[1] * v + row[2] * (1 - v)) * rad(nu, T) * n
            assert o.h2o_continuum(nu, v, p, T)[0] == pytest.approx(want, rel=1e-13)


def test_h2o_self_and_foreign_scaling():
    """k(v)/v = A v + B (1 - v): self ~ v^2, foreign ~ v (1 - v), at any state."""
    nu, p, T = 1000.0, 500.0, 260.0
    vs = np.array([1e-4, 1e-3, 1e-2, 3e-2])
    k = np.array([o.h2o_continuum(nu, v, p, T)[0] for v in vs])
    # solve for A, B from two VMRs and predict the others
    M = np.array([[vs[0], 1 - vs[0]], [vs[3], 1 - vs[3]]])
    A, B = np.linalg.solve(M, [k[0] / vs[0], k[3] / vs[3]])
    for v, kk in zip(vs[1:3], k[1:3]):
        assert kk / v == pytest.approx(A * v + B * (1 - v), rel=1e-10)
    # pure self (v = 1, foreign vanishes): the T exponent of the self term
    T1, T2 = 296.0, 250.0
    row = H2O[100]                              # nu = 1000
    r = o.h2o_continuum(1000.0, 1.0, p, T1)[0] / o.h2o_continuum(1000.0, 1.0, p, T2)[0]
    want = (296 / T1) ** row[3] / (296 / T2) ** row[3] * (T2 / T1) ** 2 * rad(1000.0, T1) / rad(1000.0, T2)
    assert r == pytest.approx(want, rel=1e-12)


def test_tables_are_linear_between_knots():
    p, T, v = 1013.0, 296.0, 0.01
    a, b = o.h2o_continuum(700.0, v, p, T)[0], o.h2o_continuum(710.0, v, p, T)[0]
    m = o.h2o_continuum(705.0, v, p, T)[0]
    n = v * p * 100.0 / (KB * T) * 1e-6
    r0, r1 = H2O[70], H2O[71]
    Cs, Cf = 0.5 * (r0[1] + r1[1]), 0.5 * (r0[2] + r1[2])
    assert m == pytest.approx((Cs * v + Cf * (1 - v)) * rad(705.0, T) * n, rel=1e-13)
    assert min(a, b) * 0.5 < m < max(a, b) * 2
    # CO2: S x xfac on the 2 cm^-1 grid from -4; 2401 is midway between 2400 and 2402
    i = int((2400.0 + 4.0) / 2.0)
    assert CO2[i][0] == 2400.0
    S = 0.5 * (CO2[i][1] * CO2[i][2] + CO2[i + 1][1] * CO2[i + 1][2])
    td = 0.5 * (CO2[i][3] + CO2[i + 1][3])
    vc, pc, Tc = 4e-4, 700.0, 240.0
    nc = vc * pc * 100.0 / (KB * Tc) * 1e-6
    want = S * (Tc / 246.0) ** td * 1e-20 * nc * (pc / 1013.0) * (296.0 / Tc) * rad(2401.0, Tc)
    assert o.co2_continuum(2401.0, vc, pc, Tc)[0] == pytest.approx(want, rel=1e-12)
    # outside 2386-2434 the bandhead exponent is zero: no (T/246) factor
    j = int((714.0 + 4.0) / 2.0)
    want = CO2[j][1] * CO2[j][2] * 1e-20 * nc * (pc / 1013.0) * (296.0 / Tc) * rad(714.0, Tc)
    assert o.co2_continuum(714.0, vc, pc, Tc)[0] == pytest.approx(want, rel=1e-12)


def test_dp_to_dz_hydrostatic():
    L = o._declare_cont()
    dz = L.orc_dp_to_dz_cm(20.0, 500.0, 250.0)
    H = 8.314462 * 250.0 / (0.028964 * 9.80665)
    assert dz == pytest.approx(H * 20.0 / 500.0 * 100.0, rel=1e-15)
    assert 29000 < dz < 30000                # ~290 m for 20 hPa at 500 hPa and 250 K


def _copy(prof):
    return AtmosphericProfile(prof.pressure.copy(), prof.temperature.copy(), prof.altitude.copy(),
                              {k: v.copy() for k, v in prof.vmr.items()})


def cia_blocks():
    """Synthetic CIA table: three temperature blocks on a common grid, one
    narrower block, a negative value (clamped) and an out-of-order T."""
    nu = np.linspace(1000.0, 3000.0, 401)
    f = lambda T: 1e-46 * np.exp(-((nu - 2300.0) / 300.0) ** 2) * (300.0 / T) ** 1.5
    b = [(250.0, nu, f(250.0)), (200.0, nu, f(200.0)), (300.0, nu, f(300.0))]
    nn = np.linspace(2000.0, 2100.0, 51)
    s4 = 2e-46 * np.ones_like(nn)
    s4[3] = -1e-47
    b.append((275.0, nn, s4))
    return b


def test_cia_temperature_bracketing_and_clamps():
    o.cia_set(1, cia_blocks())
    L = o._declare_cont()
    nu0 = 1500.0                               # outside the narrow block: blocks 200, 250, 300
    s200, s250, s300 = (1e-46 * np.exp(-((nu0 - 2300.0) / 300.0) ** 2) * (300.0 / T) ** 1.5 for T in (200, 250, 300))
    assert L.orc_cia_sigma(1, nu0, 225.0) == pytest.approx(0.5 * (s200 + s250), rel=1e-12)
    assert L.orc_cia_sigma(1, nu0, 250.0) == pytest.approx(s250, rel=1e-12)
    assert L.orc_cia_sigma(1, nu0, 150.0) == pytest.approx(s200, rel=1e-12)      # clamped below
    assert L.orc_cia_sigma(1, nu0, 350.0) == pytest.approx(s300, rel=1e-12)      # clamped above
    assert L.orc_cia_sigma(1, 999.0, 250.0) == 0.0                               # off every block
    # inside the narrow block, 275 K lies between 250 and 300 and is a block itself
    assert L.orc_cia_sigma(1, 2050.0, 275.0) == pytest.approx(2e-46, rel=1e-12)
    nu3 = 2000.0 + 3 * 2.0                     # the negative knot is clamped to 0
    s250b = 1e-46 * np.exp(-((nu3 - 2300.0) / 300.0) ** 2) * (300.0 / 250.0) ** 1.5
    assert L.orc_cia_sigma(1, nu3, 262.5) == pytest.approx(0.5 * s250b, rel=1e-9)
    # number density scaling: N2-N2 ~ n_N2^2
    p, T, v = 800.0, 250.0, 0.78
    n = v * p * 100.0 / (KB * T) * 1e-6
    assert o.cia_k(1, nu0, v, p, T)[0] == pytest.approx(s250 * n * n, rel=1e-12)
    o.cia_set(1, [])
    assert o.cia_k(1, nu0, v, p, T)[0] == 0.0


def test_oracle_continuum_jacobian_matches_finite_differences():
    """K with continuum on (H2O and CO2 continua + an N2 CIA table) against
    central differences of the oracle forward model in T and ln q_H2O."""
    o.cia_set(1, cia_blocks())
    try:
        prof = syn.base_profile(True)
        ll = filter_linelist(syn.linelists_config1(windows=True)[GasSpecies.H2O], 1180.0, 1240.0)
        lls = {GasSpecies.H2O: ll}
        snd = Sounder.make(1205.0, 1212.0, 0.25, opd_max=2.0, fwhm_gauss=0.5)
        kw = dict(internal_dnu=0.005, continua=1 | 2 | 8)
        K, y0, x0 = o.jacobian(prof, lls, snd, retrieved=[GasSpecies.H2O], include_emissivity=False, **kw)
        L = prof.n_levels

        def bt(pr):
            return o.forward(pr, lls, snd, dptmn=0.0, **kw)[1]

        for j in (3, 20):                      # temperature levels
            h = 0.02
            pp, pm = _copy(prof), _copy(prof)
            pp.temperature[j] += h
            pm.temperature[j] -= h
            fd = (bt(pp) - bt(pm)) / (2 * h)
            assert np.max(np.abs(K[:, j] - fd)) < 2e-3 * np.max(np.abs(K[:, j])) + 1e-9
        for j in (2, 15):                      # log VMR levels of H2O
            h = 1e-3
            pp, pm = _copy(prof), _copy(prof)
            pp.vmr[GasSpecies.H2O][j] *= math.exp(h)
            pm.vmr[GasSpecies.H2O][j] *= math.exp(-h)
            fd = (bt(pp) - bt(pm)) / (2 * h)
            col = L + j
            assert np.max(np.abs(K[:, col] - fd)) < 2e-3 * np.max(np.abs(K[:, col])) + 1e-9
        # the continuum changes the Jacobian measurably (the test would pass vacuously otherwise)
        K0, _, _ = o.jacobian(prof, lls, snd, retrieved=[GasSpecies.H2O], include_emissivity=False,
                              internal_dnu=0.005, continua=0)
        assert np.max(np.abs(K0[:, L + 2] - K[:, L + 2])) > 1e-3 * np.max(np.abs(K[:, L + 2]))
    finally:
        o.cia_set(1, [])
