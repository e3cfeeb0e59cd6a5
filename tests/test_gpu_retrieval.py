"""Optimal estimation on the GPU forward model and Jacobian: the reference's synthetic
closed-loop retrieval (test/runtests.jl:1725-1834, cases A-E, same line list, sounder,
profile and bars), with the observations made by the FP64 oracle's forward model, and
the GPU retrieval compared with the same loop driven by the oracle's forward model and
analytic Jacobian."""
import numpy as np
import pytest

import refcpu
from irsounderlbl.api import Context
from irsounderlbl.retrieval import _oe_solve, apodized_measurement_covariance, exclude_channels, optimal_estimation
from irsounderlbl.types import (AtmosphericProfile, GasSpecies, HITRANLinelist, Sounder, StateVectorSpec, pack_state,
                                unpack_state)

pytestmark = pytest.mark.gpu
CO2 = GasSpecies.CO2


@pytest.fixture(scope="module")
def setup():
    ll = HITRANLinelist([2, 2], [1, 1], [700.5, 702.0], [2.0e-21, 8.0e-22], np.float32([0.07, 0.07]),
                        np.float32([0.08, 0.08]), [250.0, 600.0], np.float32([0.75, 0.75]),
                        np.float32([-0.003, -0.003]))
    lls = {CO2: ll}
    snd = Sounder.make(699.0, 705.0, 0.5, opd_max=2.0, fwhm_gauss=0.5, n_channels=13)
    nlev = 6
    p = np.linspace(1000.0, 200.0, nlev)
    T = 288.0 + (225.0 - 288.0) * (1000.0 - p) / 800.0
    z = (1000.0 - p) / 66.0
    base = AtmosphericProfile(p, T, z, {CO2: np.full(nlev, 4.0e-4)})
    fm = dict(sounder=snd, apply_continuum=False, with_ils=True)
    ctx = Context(0)
    yield lls, snd, base, fm, ctx
    ctx.close()


def _oracle_bt(prof, lls, snd, T_sfc, eps=1.0):
    return refcpu.forward(prof, lls, snd, dptmn=0.0, T_sfc=T_sfc, eps_sfc=eps)[1]


def _oracle_oe(y, spec, base, lls, snd, **kw):
    """The same retrieval loop with the FP64 oracle as the model."""
    species = [b.species for b in spec.vmr_blocks]

    def fwd(x):
        p, ts, es = unpack_state(spec, x, base)
        return _oracle_bt(p, lls, snd, ts, es)

    def fj(x):
        p, ts, es = unpack_state(spec, x, base)
        K, y0, _ = refcpu.jacobian(p, lls, snd, retrieved=species, include_temperature=spec.include_temperature,
                                   include_tsfc=spec.include_tsfc, include_emissivity=spec.include_emissivity,
                                   T_sfc=ts, eps_sfc=es)
        return y0, K, np.zeros(len(y0))
    return _oe_solve(y, spec.n, fwd, fj, spec=spec, **kw)


def test_retrieve_surface_temperature(setup):
    """(A) T_sfc from a 10 K-wrong guess, noiseless; (D) the same with an apodized
    (dense) Se."""
    lls, snd, base, fm, ctx = setup
    spec = StateVectorSpec(6, [], include_temperature=False, include_tsfc=True, include_emissivity=False)
    y = _oracle_bt(base, lls, snd, 298.0)
    Se = np.diag(np.full(len(y), 0.2 ** 2))
    r = optimal_estimation(y, spec, base, lls, xa=[288.0], Sa=np.array([[100.0]]), Se=Se, fm_kwargs=fm, ctx=ctx)
    print(r)
    assert r.converged
    assert r.x[0] == pytest.approx(298.0, abs=0.05)
    assert r.dof == pytest.approx(1.0, abs=0.05)
    assert r.cost[-1] < r.cost[0]
    nu = np.arange(699.0, 705.0 + 1e-9, 0.5)
    Se_ap = apodized_measurement_covariance(nu, 0.2, dnu_sample=0.5)
    assert np.all(np.linalg.eigvalsh(Se_ap) > 0)
    rD = optimal_estimation(y, spec, base, lls, xa=[288.0], Sa=np.array([[100.0]]), Se=Se_ap, fm_kwargs=fm, ctx=ctx)
    assert rD.converged and rD.x[0] == pytest.approx(298.0, abs=0.1)


def test_retrieve_profile_and_surface(setup):
    """(B) temperature profile + T_sfc from noisy observations: fit, diagnostics, error
    budget; (E) channel blacklist; and the GPU retrieval against the oracle-driven one."""
    lls, snd, base, fm, ctx = setup
    spec = StateVectorSpec(6, [], include_temperature=True, include_tsfc=True, include_emissivity=False)
    yclean = _oracle_bt(base, lls, snd, 296.0)
    sn = 0.2
    y = yclean + sn * np.sin(13.0 * np.arange(1, len(yclean) + 1))
    xt = pack_state(spec, base, T_sfc=296.0, eps_sfc=1.0)
    xa = xt.copy()
    xa[spec.temp_range.start:spec.temp_range.stop] += 4.0
    xa[spec.tsfc_index] = 290.0
    Sa = np.diag(np.r_[np.full(6, 25.0), 25.0])
    Se = np.diag(np.full(len(y), sn ** 2))
    r = optimal_estimation(y, spec, base, lls, xa=xa, Sa=Sa, Se=Se, fm_kwargs=fm, ctx=ctx)
    print(r)
    assert r.converged and r.cost[-1] < r.cost[0]
    assert np.max(np.abs(r.y_fit - y)) < 5 * sn
    assert r.x[spec.tsfc_index] == pytest.approx(296.0, abs=1.0)
    assert 0.0 < r.dof < spec.n
    assert np.allclose(r.S_hat, r.S_hat.T) and np.all(np.linalg.eigvalsh(r.S_hat) > 0)
    assert r.A.shape == (spec.n, spec.n) and r.G.shape == (spec.n, len(y))
    assert r.dof + r.dfn == pytest.approx(len(y), rel=1e-10)
    assert r.H >= 0.0 and np.isfinite(r.H)
    np.testing.assert_allclose(r.S_hat, r.S_noise + r.S_smoothing, rtol=1e-6)

    ro = _oracle_oe(y, spec, base, lls, snd, xa=xa, Sa=Sa, Se=Se)
    print("oracle:", ro, "max|x_gpu - x_oracle| =", np.max(np.abs(r.x - ro.x)))
    assert ro.n_iter == r.n_iter
    np.testing.assert_allclose(r.x, ro.x, atol=2e-3)
    np.testing.assert_allclose(r.y_fit, ro.y_fit, atol=5e-3)
    assert r.dof == pytest.approx(ro.dof, abs=1e-3)

    mask = exclude_channels(r.nu, (701.5, 705.0))
    assert 0 < mask.sum() < len(r.nu) and np.all(r.nu[~mask] >= 701.5)
    r0 = optimal_estimation(y, spec, base, lls, xa=xa, Sa=Sa, Se=Se, fm_kwargs=fm, ctx=ctx,
                            channel_mask=np.ones(len(y), bool))
    np.testing.assert_allclose(r0.x, r.x, atol=1e-10)
    assert r0.chi2 == pytest.approx(r.chi2, rel=1e-10)
    rE = optimal_estimation(y, spec, base, lls, xa=xa, Sa=Sa, Se=Se, fm_kwargs=fm, ctx=ctx, channel_mask=mask)
    assert rE.converged
    assert len(rE.y_fit) == len(rE.nu) == len(y) and rE.K.shape == (len(y), spec.n)
    assert rE.G.shape == (spec.n, mask.sum())
    np.testing.assert_allclose(rE.A, rE.G @ rE.K[mask], rtol=1e-10, atol=1e-14)
    assert rE.dof + rE.dfn == pytest.approx(mask.sum(), rel=1e-10)
    assert 0.0 <= rE.dof <= r.dof + 1e-8
    for bad in (np.zeros(len(y), bool), np.ones(len(y) + 1, bool)):
        with pytest.raises(ValueError):
            optimal_estimation(y, spec, base, lls, xa=xa, Sa=Sa, Se=Se, fm_kwargs=fm, ctx=ctx, channel_mask=bad)


def test_retrieve_co2_profile(setup):
    """(C) a +20% CO2 profile through the log-VMR state, noiseless."""
    lls, snd, base, fm, ctx = setup
    spec = StateVectorSpec(6, [CO2], include_temperature=False, include_tsfc=False, include_emissivity=False)
    base12 = AtmosphericProfile(base.pressure, base.temperature, base.altitude, {CO2: np.full(6, 4.8e-4)})
    y = _oracle_bt(base12, lls, snd, 288.0)
    xa = pack_state(spec, base)
    r = optimal_estimation(y, spec, base, lls, xa=xa, Sa=np.eye(spec.n), Se=np.full(len(y), 0.1 ** 2),
                           fm_kwargs=fm, max_iter=20, ctx=ctx)
    print(r)
    assert r.converged and r.cost[-1] < r.cost[0]
    assert np.max(np.abs(r.y_fit - y)) < 0.05
    assert np.sum(r.x) > np.sum(xa)
