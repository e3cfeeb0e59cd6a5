"""Optimal-estimation driver, covariance builders and state unpacking on the host
(src/Estimation/optimal_estimation.jl, covariance.jl, sa_builder.jl, state_vector.jl).

The loop is pinned with models whose MAP solution is known in closed form (a linear
Gaussian problem: one Gauss-Newton step lands on x = xa + (Sa^-1 + K'Se^-1K)^-1 K'Se^-1
(y - F(xa)), and Rodgers' identities hold exactly) and by the stationarity of the cost
at the LM solution of a nonlinear model; the covariance builders at the bars of the
reference's own testsets (test/runtests.jl:1612-1720)."""
import math

import numpy as np
import pytest

from irsounderlbl.retrieval import (_oe_solve, apodized_measurement_covariance, build_sa, dB_dT,
                                    exclude_channels, scene_measurement_covariance, scene_nedt)
from irsounderlbl.types import AtmosphericProfile, ColumnScale, GasSpecies, StateVectorSpec, pack_state, unpack_state


def _linear_problem(seed=0, n=5, m=40, dense_se=False):
    rng = np.random.default_rng(seed)
    K0 = rng.normal(size=(m, n))
    c = rng.normal(size=m)
    xa = rng.normal(size=n)
    A = rng.normal(size=(n, n))
    Sa = A @ A.T + n * np.eye(n)
    if dense_se:
        B = rng.normal(size=(m, m)) * 0.1
        Se = B @ B.T + 0.5 * np.eye(m)
    else:
        Se = rng.uniform(0.2, 1.0, m)
    x_true = xa + rng.normal(size=n)
    y = K0 @ x_true + c
    nu = np.linspace(700.0, 710.0, m)

    def fwd(x):
        return K0 @ x + c

    def fj(x):
        return fwd(x), K0, nu
    return K0, c, xa, Sa, Se, y, fwd, fj


@pytest.mark.parametrize("dense_se", [False, True])
def test_linear_gaussian_map_and_diagnostics(dense_se):
    K0, c, xa, Sa, Se, y, fwd, fj = _linear_problem(dense_se=dense_se)
    n, m = K0.shape[1], K0.shape[0]
    Se_m = np.diag(Se) if not dense_se else Se
    Sei = np.linalg.inv(Se_m)
    Sai = np.linalg.inv(Sa)
    S_hat = np.linalg.inv(Sai + K0.T @ Sei @ K0)
    x_map = xa + S_hat @ K0.T @ Sei @ (y - fwd(xa))
    r = _oe_solve(y, n, fwd, fj, xa=xa, Sa=Sa, Se=Se, method="gauss_newton")
    # the first step lands on the MAP point; the second finds no decrease and stops
    assert r.n_iter <= 2 and len(r.cost) == 2
    np.testing.assert_allclose(r.x, x_map, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(r.S_hat, S_hat, rtol=1e-9, atol=1e-12)
    G = S_hat @ K0.T @ Sei
    np.testing.assert_allclose(r.G, G, rtol=1e-8, atol=1e-12)
    np.testing.assert_allclose(r.A, G @ K0, rtol=1e-8, atol=1e-12)
    assert r.dof == pytest.approx(np.trace(G @ K0), rel=1e-10)
    assert r.dof + r.dfn == pytest.approx(m, rel=1e-12)
    # Shannon information: -1/2 log2 |I - A| = 1/2 log2 |Sa S_hat^-1|
    assert r.H == pytest.approx(0.5 * np.linalg.slogdet(Sa @ np.linalg.inv(S_hat))[1] / math.log(2), rel=1e-8)
    # error budget closes exactly for a linear model (S_hat = G Se G' + (I-A) Sa (I-A)')
    np.testing.assert_allclose(r.S_noise + r.S_smoothing, r.S_hat, rtol=1e-8, atol=1e-12)
    res = y - fwd(r.x)
    assert r.chi2 == pytest.approx(res @ Sei @ res, rel=1e-9)
    assert r.cost[-1] < r.cost[0]
    # LM reaches the same fixed point
    r2 = _oe_solve(y, n, fwd, fj, xa=xa, Sa=Sa, Se=Se, max_iter=40, conv_factor=1e-12)
    np.testing.assert_allclose(r2.x, x_map, rtol=1e-6, atol=1e-6)


def test_lm_nonlinear_stationary_point():
    rng = np.random.default_rng(3)
    n, m = 4, 30
    B = rng.normal(size=(m, n)) * 0.4
    nu = np.arange(m, dtype=float)

    def fwd(x):
        return np.exp(B @ x) + 0.1 * (B @ x) ** 2

    def fj(x):
        u = B @ x
        return fwd(x), (np.exp(u) + 0.2 * u)[:, None] * B, nu
    x_true = rng.normal(size=n) * 0.5
    y = fwd(x_true)
    xa = np.zeros(n)
    Sa = np.eye(n) * 4.0
    Se = np.full(m, 1e-4)
    r = _oe_solve(y, n, fwd, fj, xa=xa, Sa=Sa, Se=Se, max_iter=60, conv_factor=1e-14)
    # grad J = -2 (K' Se^-1 (y - F) - Sa^-1 (x - xa)) vanishes at the MAP point
    F, K, _ = fj(r.x)
    g = K.T @ ((y - F) / Se) - np.linalg.inv(Sa) @ (r.x - xa)
    assert np.linalg.norm(g) < 1e-6 * np.linalg.norm(K.T @ (y / Se))
    assert np.all(np.diff(r.cost) < 0)
    np.testing.assert_allclose(r.x, x_true, atol=2e-3)


def test_channel_mask_semantics_and_errors():
    K0, c, xa, Sa, Se, y, fwd, fj = _linear_problem(seed=1)
    n, m = K0.shape[1], K0.shape[0]
    r = _oe_solve(y, n, fwd, fj, xa=xa, Sa=Sa, Se=Se)
    r0 = _oe_solve(y, n, fwd, fj, xa=xa, Sa=Sa, Se=Se, channel_mask=np.ones(m, bool))
    np.testing.assert_allclose(r0.x, r.x, atol=1e-10)
    assert r0.chi2 == pytest.approx(r.chi2, rel=1e-10)
    mask = exclude_channels(r.nu, (705.0, 710.0))
    assert 0 < mask.sum() < m and np.all(r.nu[~mask] >= 705.0)
    assert exclude_channels(r.nu).sum() == m
    rm = _oe_solve(y, n, fwd, fj, xa=xa, Sa=Sa, Se=Se, channel_mask=mask)
    assert len(rm.y_fit) == len(rm.nu) == m and rm.K.shape == (m, n)
    assert rm.G.shape == (n, mask.sum())
    np.testing.assert_allclose(rm.A, rm.G @ rm.K[mask], rtol=1e-10, atol=1e-13)
    assert rm.dof + rm.dfn == pytest.approx(mask.sum(), rel=1e-10)
    assert 0.0 <= rm.dof <= r.dof + 1e-8
    with pytest.raises(ValueError):
        _oe_solve(y, n, fwd, fj, xa=xa, Sa=Sa, Se=Se, channel_mask=np.zeros(m, bool))
    with pytest.raises(ValueError):
        _oe_solve(y, n, fwd, fj, xa=xa, Sa=Sa, Se=Se, channel_mask=np.ones(m + 1, bool))
    with pytest.raises(ValueError):
        _oe_solve(y, n, fwd, fj, xa=xa[:-1], Sa=Sa, Se=Se)
    with pytest.raises(ValueError):
        _oe_solve(y, n, fwd, fj, xa=xa, Sa=Sa, Se=Se, method="newton")


def test_apodized_measurement_covariance():
    nu = np.arange(645.0, 650.0 + 1e-9, 0.25)
    Se = apodized_measurement_covariance(nu, 0.2)
    assert np.allclose(Se, Se.T) and np.all(np.linalg.eigvalsh(Se) > 0)
    assert np.allclose(np.diag(Se), 0.04)
    s2 = 0.04
    assert Se[0, 1] / s2 == pytest.approx(0.705, abs=0.02)
    assert Se[0, 2] / s2 == pytest.approx(0.250, abs=0.02)
    assert Se[0, 3] / s2 == pytest.approx(0.044, abs=0.01)
    assert Se[0, 4] / s2 < 0.006
    assert Se[0, 5] == 0.0
    Ssp = apodized_measurement_covariance([645.0, 647.0, 649.0], 0.2)
    assert Ssp[0, 1] == 0.0 and Ssp[0, 2] == 0.0
    sv = np.array([0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7])
    Sv = apodized_measurement_covariance(np.arange(645.0, 646.5 + 1e-9, 0.25), sv)
    assert np.allclose(np.diag(Sv), sv ** 2)
    Sd = apodized_measurement_covariance(nu, 0.2, fwhm_gauss=0.0)
    assert Sd[0, 1] == 0.0 and np.allclose(np.diag(Sd), 0.04)
    # the lag correlation is the discrete autocorrelation of Gaussian weights: for a
    # sampling finer than the FWHM (and weights out to ~5 sigma) it is the continuous
    # exp(-d^2 / 4 s^2) up to exponentially small Riemann-sum and truncation errors
    s = 0.5 / (2 * math.sqrt(2 * math.log(2)))
    fine = apodized_measurement_covariance(np.arange(0, 5) * 0.1, 1.0, dnu_sample=0.1)
    for k in range(1, 5):
        assert fine[0, k] == pytest.approx(math.exp(-(0.1 * k) ** 2 / (4 * s * s)), rel=1e-8)


def test_scene_nedt():
    assert scene_nedt([720.0], [280.0], nedt_280K=0.25)[0] == pytest.approx(0.25, abs=1e-12)
    s = scene_nedt(np.full(4, 720.0), [215.0, 240.0, 270.0, 295.0], nedt_280K=0.25)
    assert np.all(np.diff(s) < 0) and s[0] > 0.4 and s[-1] < 0.25
    # dB/dT against a central difference of the Planck function
    B = lambda nu, T: 1.191042953e-5 * nu ** 3 / math.expm1(1.4387769 * nu / T)
    for nu, T in ((700.0, 230.0), (2400.0, 280.0)):
        fd = (B(nu, T + 1e-3) - B(nu, T - 1e-3)) / 2e-3
        assert dB_dT(nu, T) == pytest.approx(fd, rel=1e-7)
    nuv = np.arange(660.0, 661.0 + 1e-9, 0.25)
    Se = scene_measurement_covariance(nuv, np.full(len(nuv), 230.0), nedt_280K=0.25)
    assert np.allclose(Se, Se.T) and np.all(np.linalg.eigvalsh(Se) > 0)
    assert math.sqrt(Se[0, 0]) == pytest.approx(scene_nedt(nuv, np.full(len(nuv), 230.0))[0], rel=1e-12)


def _base8():
    nlev = 8
    p = np.linspace(1000.0, 100.0, nlev)
    T = 288.0 + (220.0 - 288.0) * (1000.0 - p) / 900.0
    z = (1000.0 - p) / 66.0
    return AtmosphericProfile(p, T, z, {GasSpecies.CO2: np.full(nlev, 4e-4), GasSpecies.H2O: np.full(nlev, 1e-3)})


def test_build_sa():
    base = _base8()
    spec = StateVectorSpec(8, [GasSpecies.CO2, GasSpecies.H2O])
    Sa = build_sa(spec, base, sigma_T=2.0, sigma_vmr=0.5, L_T=1.0, L_vmr=1.5, sigma_tsfc=3.0, sigma_emis=0.02)
    assert Sa.shape == (spec.n, spec.n) and np.allclose(Sa, Sa.T) and np.all(np.linalg.eigvalsh(Sa) > 0)
    t = spec.temp_range
    assert np.allclose(np.diag(Sa)[t.start:t.stop], 4.0)
    for _, r in spec.vmr_ranges:
        assert np.allclose(np.diag(Sa)[r.start:r.stop], 0.25)
    assert Sa[spec.tsfc_index, spec.tsfc_index] == pytest.approx(9.0)
    assert Sa[spec.emis_index, spec.emis_index] == pytest.approx(0.02 ** 2)
    rc = spec.vmr_ranges[0][1]
    rh = spec.vmr_ranges[1][1]
    assert np.all(Sa[t.start:t.stop, rc.start:rc.stop] == 0) and np.all(Sa[rc.start:rc.stop, rh.start:rh.stop] == 0)
    cors = Sa[t.start, t.start:t.stop] / 4.0
    assert cors[0] == pytest.approx(1.0) and np.all(np.diff(cors) < 1e-12) and 0 < cors[-1] < cors[1] < 1
    # Matern-5/2 at r = |ln(p0/p1)| / L in closed form
    r = abs(math.log(base.pressure[0] / base.pressure[1]))
    assert cors[1] == pytest.approx((1 + math.sqrt(5) * r + 5 * r * r / 3) * math.exp(-math.sqrt(5) * r), rel=1e-12)
    spec2 = StateVectorSpec(8, [GasSpecies.CO2, GasSpecies.O3], include_tsfc=False, include_emissivity=False)
    base2 = AtmosphericProfile(base.pressure, base.temperature, base.altitude,
                               {GasSpecies.CO2: np.full(8, 4e-4), GasSpecies.O3: np.full(8, 1e-6)})
    Sa2 = build_sa(spec2, base2, sigma_vmr={GasSpecies.CO2: 0.3, GasSpecies.O3: 0.8},
                   L_vmr={GasSpecies.CO2: 2.0, GasSpecies.O3: 0.5})
    c1 = spec2.vmr_ranges[0][1].start
    c2 = spec2.vmr_ranges[1][1].start
    assert Sa2[c1, c1] == pytest.approx(0.09) and Sa2[c2, c2] == pytest.approx(0.64)
    assert Sa2[c2, c2 + 1] / 0.64 < Sa2[c1, c1 + 1] / 0.09
    with pytest.raises(ValueError):
        build_sa(spec2, base2, sigma_vmr={GasSpecies.CO2: 0.3})
    with pytest.raises(ValueError):
        build_sa(spec, base, sigma_T=np.full(7, 2.0))
    with pytest.raises(ValueError):
        build_sa(spec, base, kernel="bogus")
    specC = StateVectorSpec(8, [ColumnScale(GasSpecies.CO2)], include_temperature=False,
                            include_tsfc=False, include_emissivity=False)
    SaC = build_sa(specC, base, sigma_col=0.3)
    assert SaC.shape == (1, 1) and SaC[0, 0] == pytest.approx(0.09)


def test_unpack_state_round_trip():
    base = _base8()
    spec = StateVectorSpec(8, [GasSpecies.CO2, GasSpecies.H2O])
    x = pack_state(spec, base, T_sfc=291.0, eps_sfc=0.97)
    p, ts, es = unpack_state(spec, x, base)
    np.testing.assert_allclose(p.temperature, base.temperature)
    np.testing.assert_allclose(p.vmr[GasSpecies.H2O], base.vmr[GasSpecies.H2O], rtol=1e-14)
    assert ts == 291.0 and es == 0.97
    specC = StateVectorSpec(8, [ColumnScale(GasSpecies.CO2)], include_temperature=False,
                            include_tsfc=False, include_emissivity=False)
    p2, ts2, es2 = unpack_state(specC, [math.log(1.2)], base)
    np.testing.assert_allclose(p2.vmr[GasSpecies.CO2], 1.2 * base.vmr[GasSpecies.CO2], rtol=1e-14)
    assert ts2 == base.temperature[0] and es2 == 1.0
    with pytest.raises(ValueError):
        unpack_state(spec, x[:-1], base)
