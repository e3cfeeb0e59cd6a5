"""GPU analytic Jacobian (irs_jacobian) against the FP64 oracle's analytic
Jacobian, which is itself pinned against finite differences of the oracle
forward model in test_oracle_jacobian.py.

Bar: every column within 2e-3 of that column's largest |K| (FP32 line sums
and derivative sums against FP64), y0 within 1e-3 K of the oracle's BT."""
import numpy as np
import pytest

import refcpu as o
from irsounderlbl import synthetic as syn
from irsounderlbl.api import analytic_jacobian, forward_model
from irsounderlbl.types import (ColumnScale, FullProfile, GasSpecies, PartialColumns, Sounder,
                                StateVectorSpec, ViewingGeometry, filter_linelist, pack_state)

from test_oracle_jacobian import KW, small_problem

pytestmark = pytest.mark.gpu

COL_TOL = 2e-3


def _geom(mu):
    return ViewingGeometry(float(np.degrees(np.arccos(mu))), 0.0, mu)


def _cmp(K, Kref, floor):
    """Largest per-column error relative to max(column max, floor)."""
    worst = 0.0
    for c in range(Kref.shape[1]):
        scale = max(np.abs(Kref[:, c]).max(), floor)
        diff = np.abs(K[:, c] - Kref[:, c]).max()
        if scale == 0.0:                     # an all-zero reference column (opaque surface)
            worst = max(worst, 0.0 if diff == 0.0 else np.inf)
            continue
        worst = max(worst, diff / scale)
    return worst


@pytest.mark.parametrize("sf,src", [(1, "cim"), (0, "toon")])
@pytest.mark.parametrize("log_vmr", [True, False])
def test_jacobian_small_problem(ctx, sf, src, log_vmr):
    prof, lls, snd = small_problem()
    T_sfc, eps, mu = 292.0, 0.97, 0.9
    spec = StateVectorSpec(prof.n_levels, [GasSpecies.H2O, GasSpecies.CO2], log_vmr=log_vmr)
    jac = analytic_jacobian(prof, lls, spec, sounder=snd, geom=_geom(mu), T_sfc=T_sfc, eps_sfc=eps,
                            source_function=src, ctx=ctx, **KW, apply_continuum=False)
    Kr, y0r, x0r = o.jacobian(prof, lls, snd, retrieved=[GasSpecies.H2O, GasSpecies.CO2],
                              T_sfc=T_sfc, eps_sfc=eps, mu=mu, source_function=sf, log_vmr=log_vmr, **KW)
    assert jac.K.shape == Kr.shape == (33, spec.n)
    assert np.max(np.abs(jac.y0 - y0r)) < 1e-3
    assert np.allclose(jac.x0, x0r, rtol=1e-14, atol=0)
    assert np.allclose(jac.x0, pack_state(spec, prof, T_sfc, eps), rtol=1e-14, atol=0)
    # floor: 1e-3 of the block's largest entry (upper levels carry ~no signal)
    n = prof.n_levels
    for blk in (slice(0, n), slice(n, 2 * n), slice(2 * n, 3 * n), slice(3 * n, 3 * n + 2)):
        err = _cmp(jac.K[:, blk], Kr[:, blk], 1e-3 * np.abs(Kr[:, blk]).max())
        assert err < COL_TOL, (blk, err)


def test_jacobian_radiance_and_reduced_blocks(ctx):
    """observable=:radiance; ColumnScale and PartialColumns equal the basis
    projection of the per-level log-VMR columns."""
    prof, lls, snd = small_problem()
    n = prof.n_levels
    B = np.zeros((n, 3))
    B[:15, 0], B[15:35, 1], B[35:, 2] = 1.0, 1.0, 1.0
    spec = StateVectorSpec(n, [ColumnScale(GasSpecies.CO2), PartialColumns(GasSpecies.H2O, B)],
                           include_temperature=False, include_tsfc=False, include_emissivity=False,
                           log_vmr=False)
    jac = analytic_jacobian(prof, lls, spec, sounder=snd, observable="radiance", ctx=ctx, **KW, apply_continuum=False)
    Kr, y0r, _ = o.jacobian(prof, lls, snd, retrieved=[GasSpecies.CO2, GasSpecies.H2O],
                            include_temperature=False, include_tsfc=False, include_emissivity=False,
                            log_vmr=True, observable="radiance", **KW)
    assert jac.K.shape == (33, 4)
    assert np.max(np.abs(jac.y0 - y0r) / y0r) < 1e-6
    ref_cs = Kr[:, :n].sum(axis=1, keepdims=True)
    ref_pc = Kr[:, n:2 * n] @ B
    assert _cmp(jac.K[:, :1], ref_cs, 0.0) < COL_TOL
    assert _cmp(jac.K[:, 1:], ref_pc, 0.0) < COL_TOL
    assert jac.spec.labels() == ["logVMRscale_CO2", "logVMRscale_H2O[1]", "logVMRscale_H2O[2]",
                                 "logVMRscale_H2O[3]"]


def test_jacobian_no_coupling_and_mass_weighted(ctx):
    prof, lls, snd = small_problem()
    spec = StateVectorSpec(prof.n_levels, [FullProfile(GasSpecies.H2O)])
    for tm, tmi in (("mass_weighted", 1), ("logp_at_pcg", 0)):
        for cpl in (False, True):
            jac = analytic_jacobian(prof, lls, spec, sounder=snd, T_method=tm, vmr_coupling=cpl,
                                    ctx=ctx, **KW, apply_continuum=False)
            Kr, _, _ = o.jacobian(prof, lls, snd, retrieved=[GasSpecies.H2O], T_method=tmi,
                                  vmr_coupling=cpl, **KW)
            err = _cmp(jac.K, Kr, 1e-3 * np.abs(Kr).max())
            assert err < COL_TOL, (tm, cpl, err)


def test_jacobian_line_mixing_window(ctx):
    """Lines of the full-size lists (VP_Y on CO2) in 698-705 cm^-1, channels
    700-703 cm^-1, 49 layers, CO2 + H2O retrieved, internal grid 0.002 cm^-1
    (the oracle's FP64 pair sum over every line bounds the window)."""
    lls = syn.linelists_config1()
    prof = syn.base_profile(True)
    sub = {k: filter_linelist(v, 698.0, 705.0) for k, v in lls.items()}
    snd = Sounder.make(700.0, 703.0, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    spec = StateVectorSpec(prof.n_levels, [GasSpecies.CO2, GasSpecies.H2O])
    jac = analytic_jacobian(prof, sub, spec, sounder=snd, line_mixing="vpy", internal_dnu=0.002,
                            ctx=ctx, apply_continuum=False)
    Kr, y0r, _ = o.jacobian(prof, sub, snd, retrieved=[GasSpecies.CO2, GasSpecies.H2O],
                            line_mixing=1, internal_dnu=0.002)
    assert np.max(np.abs(jac.y0 - y0r)) < 1e-3
    n = prof.n_levels
    for blk in (slice(0, n), slice(n, 2 * n), slice(2 * n, 3 * n), slice(3 * n, 3 * n + 2)):
        err = _cmp(jac.K[:, blk], Kr[:, blk], 1e-3 * np.abs(Kr[:, blk]).max())
        assert err < COL_TOL, (blk, err)
    # y0 is the forward model without DPTMIN
    _, _, bt = forward_model(prof, sub, sounder=snd, line_mixing="vpy", dptmn=0.0,
                             internal_dnu=0.002, ctx=ctx, apply_continuum=False)
    assert np.max(np.abs(bt - jac.y0)) < 1e-6


def test_jacobian_coarse_far_wing_matches_every_pair(ctx):
    """The shortcuts of the gradient line sum (the seven sums of lines beyond
    1.5 cm^-1 at coarse nodes, and of the cutoff-edge lines at four nodes per
    64-point slice, both with cubic interpolation) against the every-pair
    gradient sum (far_wing_cm = 0): 15 um CO2 band with VP_Y and
    H2O lines in 675-735 cm^-1, channels 700-710 cm^-1, internal grid
    0.001 cm^-1, 49 layers, CO2 + H2O retrieved.  The profile is a perturbed
    one: above 11 km the base profile is isothermal, where the VMR columns
    of this opaque band vanish and both sums return rounding noise."""
    lls = syn.linelists_config1()
    prof, T_sfc, mu = syn.perturbed_profiles(1, seed=7)[0]
    sub = {k: filter_linelist(v, 675.0, 735.0) for k, v in lls.items()}
    snd = Sounder.make(700.0, 710.0, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    spec = StateVectorSpec(prof.n_levels, [GasSpecies.CO2, GasSpecies.H2O])

    def run():
        return analytic_jacobian(prof, sub, spec, sounder=snd, geom=_geom(mu), T_sfc=T_sfc,
                                 line_mixing="vpy", internal_dnu=0.001, ctx=ctx, apply_continuum=False)
    try:
        ctx.set_schedule(4, far_wing_cm=0.0)
        exact = run()
        ctx.set_schedule(4, far_wing_cm=1.5)
        far = run()
    finally:
        ctx.set_schedule(4)
    assert np.any(far.K != exact.K)          # the coarse path ran
    assert np.max(np.abs(far.y0 - exact.y0)) < 1e-4
    n = prof.n_levels
    for blk in (slice(0, n), slice(n, 2 * n), slice(2 * n, 3 * n), slice(3 * n, 3 * n + 2)):
        err = _cmp(far.K[:, blk], exact.K[:, blk], 1e-3 * np.abs(exact.K[:, blk]).max())
        assert err < 2e-6, (blk, err)                 # measured 6.6e-8


def test_jacobian_errors(ctx):
    prof, lls, snd = small_problem()
    spec = StateVectorSpec(prof.n_levels, [GasSpecies.O3])
    with pytest.raises(ValueError):
        analytic_jacobian(prof, lls, spec, sounder=snd, ctx=ctx, **KW, apply_continuum=False)
    with pytest.raises(ValueError):
        analytic_jacobian(prof, lls, StateVectorSpec(prof.n_levels), sounder=snd, observable="tb",
                          ctx=ctx, **KW, apply_continuum=False)
    with pytest.raises(ValueError):
        analytic_jacobian(prof, lls, StateVectorSpec(prof.n_levels + 1), sounder=snd, ctx=ctx, **KW, apply_continuum=False)
