"""Continuum on the device (continuum_kernel, k_cont.hip) against the FP64
oracle: the layer functions of src/CrossSections/continuum.jl, CIA tables
set directly and read from a HITRAN .cia file, forward_model with the
reference's default apply_continuum=true, and the continuum terms of the
analytic Jacobian.  The oracle's continuum is pinned in
test_oracle_continuum.py."""
import numpy as np
import pytest

import refcpu as o
from irsounderlbl import synthetic as syn
from irsounderlbl.api import (analytic_jacobian, co2_cia, co2_continuum, forward_model, h2o_continuum,
                              n2_cia, o2_cia)
from irsounderlbl.types import GasSpecies, Sounder, StateVectorSpec, ViewingGeometry, filter_linelist, wavenumber_grid

from test_oracle_continuum import cia_blocks
from test_gpu_jacobian import COL_TOL, _cmp

pytestmark = pytest.mark.gpu

BT_TOL = 0.005


@pytest.mark.parametrize("p,T,v", [(1013.25, 296.0, 0.01), (300.0, 220.0, 2e-4), (20.0, 260.0, 4e-6)])
def test_layer_continua_match_oracle(ctx, p, T, v):
    g = wavenumber_grid(629.0, 2776.0, 0.25)
    got = h2o_continuum(g, v, p, T, ctx=ctx)
    ref = o.h2o_continuum(g.nu, v, p, T)
    assert np.all(got >= 0) and ref.max() > 0
    assert np.max(np.abs(got - ref) / ref.max()) < 1e-12
    got = co2_continuum(g, 4.15e-4, p, T, ctx=ctx)
    ref = o.co2_continuum(g.nu, 4.15e-4, p, T)
    m = ref > 0
    assert np.max(np.abs(got[m] - ref[m]) / ref[m]) < 1e-12
    # no CIA table loaded: zeros, as the reference without its data files
    for f in (co2_cia, n2_cia, o2_cia):
        assert np.all(f(g, 0.5, p, T, ctx=ctx) == 0.0)


def _write_cia(path, blocks, descending=()):
    with open(path, "w") as f:
        for b, (T, nu, s) in enumerate(blocks):
            f.write(f"{'N2-N2':>20s}{nu.min():10.3f}{nu.max():10.3f}{len(nu):7d}{T:7.1f}{s.max():10.3e}"
                    f"{0.0:10.3f}  synthetic\n")
            order = range(len(nu) - 1, -1, -1) if b in descending else range(len(nu))
            for i in order:
                f.write(f"{nu[i]:10.4f} {s[i]:12.5e}\n")


def test_cia_tables_match_oracle(ctx, tmp_path):
    blocks = cia_blocks()
    o.cia_set(1, blocks)
    try:
        g = wavenumber_grid(990.0, 3010.0, 0.1)
        for T in (190.0, 225.0, 262.5, 275.0, 300.0, 330.0):
            ctx.set_cia("N2-N2", blocks)
            got = n2_cia(g, 0.78, 600.0, T, ctx=ctx)
            ref = o.cia_k(1, g.nu, 0.78, 600.0, T)
            assert ref.max() > 0
            assert np.max(np.abs(got - ref)) < 1e-12 * ref.max(), T
            # the same table through the .cia text parser (one block stored descending)
            p = tmp_path / "N2-N2_synthetic.cia"
            _write_cia(p, blocks, descending=(1,))
            ctx.load_cia("N2-N2", str(p))
            got2 = n2_cia(g, 0.78, 600.0, T, ctx=ctx)
            # the text file carries 6 significant digits of sigma
            assert np.max(np.abs(got2 - ref)) < 2e-5 * ref.max(), T
        with pytest.raises(ValueError):
            ctx.set_cia("N2-N2", [(250.0, np.array([2.0, 1.0]), np.array([1.0, 1.0]))])
    finally:
        ctx.set_cia("N2-N2", [])
        o.cia_set(1, [])
    assert np.all(n2_cia(wavenumber_grid(1000.0, 1010.0, 1.0), 0.78, 600.0, 250.0, ctx=ctx) == 0.0)


def _window_problem():
    prof = syn.base_profile(True)
    lls = syn.linelists_config1(windows=True)
    lo, hi = 1205.0, 1215.0
    sub = {k: filter_linelist(v, lo - 42.0, hi + 42.0) for k, v in lls.items()}
    snd = Sounder.make(lo, hi, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    return prof, sub, snd


def test_forward_with_continuum_matches_oracle(ctx):
    """apply_continuum=true (the reference default) on a partly transparent H2O
    window, slant path, with an N2 CIA table loaded on both sides."""
    prof, sub, snd = _window_problem()
    mu = 0.8
    geom = ViewingGeometry(float(np.degrees(np.arccos(mu))), 0.0, mu)
    o.cia_set(1, cia_blocks())
    ctx.set_cia("N2-N2", cia_blocks())
    try:
        _, _, bt = forward_model(prof, sub, sounder=snd, geom=geom, line_mixing="vpy", ctx=ctx)
        _, bto, _ = o.forward(prof, sub, snd, line_mixing=1, mu=mu, continua=31)
        _, _, bt0 = forward_model(prof, sub, sounder=snd, geom=geom, line_mixing="vpy", ctx=ctx,
                                  apply_continuum=False)
        d = np.abs(bt - bto)
        print(f"continuum BT {bto.min():.2f}-{bto.max():.2f} K, max|d| {d.max():.2e} K, "
              f"continuum effect {np.abs(bt - bt0).max():.3f} K")
        assert d.max() < BT_TOL
        assert np.abs(bt - bt0).max() > 0.05          # the continuum is not negligible here
        # a subset of the continua
        _, _, bt2 = forward_model(prof, sub, sounder=snd, geom=geom, line_mixing="vpy", ctx=ctx,
                                  continua=("co2", "n2"))
        _, bto2, _ = o.forward(prof, sub, snd, line_mixing=1, mu=mu, continua=2 | 8)
        assert np.abs(bt2 - bto2).max() < BT_TOL
        with pytest.raises(ValueError):
            forward_model(prof, sub, sounder=snd, ctx=ctx, continua=("h2o", "ch4"))
        # batch of two profiles == single calls with continuum on
        pp = syn.perturbed_profiles(2, seed=77)
        _, _, btb = forward_model([x[0] for x in pp], sub, sounder=snd, ctx=ctx)
        for q in range(2):
            _, _, b1 = forward_model(pp[q][0], sub, sounder=snd, ctx=ctx)
            assert np.array_equal(btb[q], b1)
    finally:
        ctx.set_cia("N2-N2", [])
        o.cia_set(1, [])


def test_jacobian_with_continuum_matches_oracle(ctx):
    prof, sub, snd = _window_problem()
    lls = {GasSpecies.H2O: sub[GasSpecies.H2O]}
    kw = dict(internal_dnu=0.005)
    o.cia_set(1, cia_blocks())
    ctx.set_cia("N2-N2", cia_blocks())
    try:
        for cpl in (True, False):
            spec = StateVectorSpec(prof.n_levels, [GasSpecies.H2O], include_emissivity=False)
            jac = analytic_jacobian(prof, lls, spec, sounder=snd, vmr_coupling=cpl, ctx=ctx, **kw)
            Kr, y0r, _ = o.jacobian(prof, lls, snd, retrieved=[GasSpecies.H2O], include_emissivity=False,
                                    vmr_coupling=cpl, continua=31, **kw)
            assert np.max(np.abs(jac.y0 - y0r)) < 1e-3
            n = prof.n_levels
            for blk in (slice(0, n), slice(n, 2 * n)):
                err = _cmp(jac.K[:, blk], Kr[:, blk], 1e-3 * np.abs(Kr[:, blk]).max())
                assert err < COL_TOL, (cpl, blk, err)
    finally:
        ctx.set_cia("N2-N2", [])
        o.cia_set(1, [])
