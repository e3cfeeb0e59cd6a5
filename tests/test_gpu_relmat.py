"""Relaxation-matrix line mixing on the GPU against the FP64 oracle
(oracle/relmat_ref.py, oracle/refcpu.c): the perturbation of VPYLineMixing (Y(T)
from the sum-rule-corrected W of every band) and VPWLineMixing (eigenmodes of
D + i p W for whitelisted bands) on a 0.001 cm^-1 grid, the layer cross sections
max(sigma_voigt + dsigma, 0), and forward_model brightness temperatures."""
import numpy as np
import pytest

import refcpu
import relmat_ref as R
from irsounderlbl import synthetic as syn
from irsounderlbl.api import (Context, compute_voigt_lm_cross_sections, compute_voigt_vpw_cross_sections,
                              forward_model, lm_perturbation)
from irsounderlbl.relmat import VPWLineMixing, VPYLineMixing, linelist_isotopes
from irsounderlbl.types import AtmosphericProfile, GasSpecies, Sounder, wavenumber_grid

pytestmark = pytest.mark.gpu
WL = {"S626_01", "S628_12"}


@pytest.fixture(scope="module")
def env():
    ctx = Context(0)
    rm = syn.relmat_synthetic()
    ll = syn.linelist_from_relmat(rm)
    yield ctx, rm, ll
    ctx.close()


def _nodes(g):
    return refcpu.grid_nodes(g.nu_min, g.dnu, g.n)


@pytest.mark.parametrize("T,p", [(220.0, 0.05), (290.0, 0.8)])
def test_lm_perturbation_matches_oracle(env, T, p):
    ctx, rm, ll = env
    g = wavenumber_grid(640.0, 760.0, 0.001)
    nu = _nodes(g)
    for lm, ref in ((VPYLineMixing(rm), R.lm_dispersive_correction(nu, rm, T, p)),
                    (VPWLineMixing(rm, WL), R.vpw_perturbation(nu, rm, T, p, WL))):
        got = lm_perturbation(g, lm, T, p, ctx=ctx)
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        print(type(lm).__name__, T, p, "max|GPU - oracle| / peak =", err, flush=True)
        assert err < 1e-4


@pytest.mark.parametrize("T,p", [(230.0, 0.1), (285.0, 0.9)])
def test_lm_cross_sections_match_oracle(env, T, p):
    ctx, rm, ll = env
    g = wavenumber_grid(640.0, 760.0, 0.001)
    nu = _nodes(g)
    sv = refcpu.voigt_xsec(ll, g.nu_min, g.dnu, g.n, T, p)
    iso = linelist_isotopes(ll)
    for name, got, d in (
            ("vpy", compute_voigt_lm_cross_sections(g, ll, rm, T, p, ctx=ctx),
             R.lm_dispersive_correction(nu, rm, T, p, isotopes=iso)),
            ("vpw", compute_voigt_vpw_cross_sections(g, ll, rm, T, p, whitelist=WL, ctx=ctx),
             R.vpw_perturbation(nu, rm, T, p, WL, isotopes=iso))):
        ref = np.maximum(sv + d, 0.0)
        m = ref > 1e-6 * ref.max()
        rel = np.max(np.abs(got[m] - ref[m]) / ref[m])
        print(name, T, p, "max rel err where sigma > 1e-6 peak:", rel, flush=True)
        assert rel < 1e-4


def test_forward_model_relmat_matches_oracle(env):
    ctx, rm, ll = env
    # CO2 x 1e-3: BT 219-268 K, so every layer's optical depth shows (the full amount
    # saturates the window at the 217 K of the upper levels)
    p0 = syn.base_profile(False)
    prof = AtmosphericProfile(p0.pressure, p0.temperature, p0.altitude,
                              {GasSpecies.H2O: p0.vmr[GasSpecies.H2O], GasSpecies.CO2: p0.vmr[GasSpecies.CO2] * 1e-3})
    snd = Sounder.make(660.0, 680.0, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    lls = {GasSpecies.CO2: ll}
    iso = linelist_isotopes(ll)
    _, _, bt0 = forward_model(prof, lls, sounder=snd, ctx=ctx, apply_continuum=False)
    for lm, fn in ((VPYLineMixing(rm), lambda nu, T, p: R.lm_dispersive_correction(nu, rm, T, p, isotopes=iso)),
                   (VPWLineMixing(rm, WL), lambda nu, T, p: R.vpw_perturbation(nu, rm, T, p, WL, isotopes=iso))):
        _, _, bt = forward_model(prof, lls, sounder=snd, ctx=ctx, apply_continuum=False, line_mixing=lm)
        _, bto = refcpu.forward_lm(prof, lls, snd, 2, fn)
        err = np.max(np.abs(bt - bto))
        print(type(lm).__name__, "BT", bto.min(), bto.max(), "max|GPU - oracle| =", err,
              "LM effect", np.max(np.abs(bt - bt0)), flush=True)
        assert err < 0.005
        assert np.max(np.abs(bt - bt0)) > 0.5           # the perturbation is live


def test_batched_relmat_equals_single(env):
    ctx, rm, ll = env
    profs = [AtmosphericProfile(q.pressure, q.temperature, q.altitude,
                                {k: v * (1e-3 if k == GasSpecies.CO2 else 1.0) for k, v in q.vmr.items()})
             for q, _, _ in syn.perturbed_profiles(3, seed=11)]
    snd = Sounder.make(660.0, 670.0, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    lls = {GasSpecies.CO2: ll}
    lm = VPWLineMixing(rm, WL)
    _, _, bt = forward_model(profs, lls, sounder=snd, ctx=ctx, apply_continuum=False, line_mixing=lm)
    for q, p in enumerate(profs):
        _, _, b1 = forward_model(p, lls, sounder=snd, ctx=ctx, apply_continuum=False, line_mixing=lm)
        np.testing.assert_array_equal(bt[q], b1)


def _small_profile(nlev=10):
    p = np.geomspace(1000.0, 50.0, nlev)
    T = 288.0 - 60.0 * np.log(1000.0 / p) / np.log(20.0)
    z = 7.0 * np.log(1000.0 / p)
    return AtmosphericProfile(p, T, z, {GasSpecies.CO2: np.full(nlev, 4.15e-7)})


def test_analytic_jacobian_relmat_matches_finite_differences(env):
    """K with VPW line mixing (the perturbation central-differenced per layer, as the
    reference does) against central differences of the GPU forward model, and two
    columns against central differences of the FP64 oracle's forward model."""
    from irsounderlbl.api import analytic_jacobian
    from irsounderlbl.types import StateVectorSpec, pack_state, unpack_state
    ctx, rm, ll = env
    prof = _small_profile()
    snd = Sounder.make(665.0, 672.0, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    lls = {GasSpecies.CO2: ll}
    lm = VPWLineMixing(rm, WL)
    spec = StateVectorSpec(prof.n_levels, [GasSpecies.CO2], include_emissivity=False)
    kw = dict(sounder=snd, apply_continuum=False, line_mixing=lm, ctx=ctx)
    jac = analytic_jacobian(prof, lls, spec, **kw)
    x0 = pack_state(spec, prof)

    def fwd(x):
        p, ts, _ = unpack_state(spec, x, prof)
        return forward_model(p, lls, T_sfc=ts, dptmn=0.0, **kw)[2]
    np.testing.assert_allclose(jac.y0, fwd(x0), atol=1e-6)
    worst = 0.0
    for j in range(spec.n):
        h = 0.5 if (j < prof.n_levels or j == spec.tsfc_index) else 0.02
        xp, xm = x0.copy(), x0.copy()
        xp[j] += h
        xm[j] -= h
        fd = (fwd(xp) - fwd(xm)) / (2 * h)
        scale = max(np.max(np.abs(jac.K[:, j])), np.max(np.abs(fd)), 1e-12)
        worst = max(worst, np.max(np.abs(jac.K[:, j] - fd)) / scale)
    print("max column error of K vs FD of the GPU forward:", worst, flush=True)
    assert worst < 5e-3
    iso = linelist_isotopes(ll)
    fn = lambda nu, T, p: R.vpw_perturbation(nu, rm, T, p, WL, isotopes=iso)    # noqa: E731
    for j, h in ((3, 0.5), (prof.n_levels + 4, 0.02)):
        out = []
        for sgn in (1, -1):
            x = x0.copy()
            x[j] += sgn * h
            p, ts, _ = unpack_state(spec, x, prof)
            out.append(refcpu.forward_lm(p, lls, snd, 2, fn, dptmn=0.0, T_sfc=ts)[1])
        fd = (out[0] - out[1]) / (2 * h)
        err = np.max(np.abs(jac.K[:, j] - fd)) / np.max(np.abs(fd))
        print("column", j, "K vs FD of the FP64 oracle:", err, flush=True)
        assert err < 5e-3
