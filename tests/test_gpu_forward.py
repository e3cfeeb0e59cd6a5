"""End-to-end forward_model parity on the GPU (BT vs the FP64 oracle)."""
import numpy as np
import pytest

import refcpu
from irsounderlbl import synthetic as syn
from irsounderlbl.types import GasSpecies, Sounder, filter_linelist, wavenumber_grid

pytestmark = pytest.mark.gpu

BT_TOL = 0.005   # K, plain LBL


def test_rte_matches_oracle(ctx):
    from irsounderlbl.api import schwarzschild_rte
    rng = np.random.default_rng(3)
    g = wavenumber_grid(645.0, 805.0, 0.01)
    L = 7
    tau = rng.uniform(0.0, 3.0, (g.n, L)) * rng.uniform(0, 1, (1, L))
    T_lev = np.linspace(295, 215, L + 1)
    T_ave = 0.5 * (T_lev[:-1] + T_lev[1:]) + 0.7
    for sf, eps in [("cim", 1.0), ("toon", 1.0), ("cim", 0.9), ("toon", 0.95)]:
        got = schwarzschild_rte(g, tau, T_lev, 300.0, mu=0.8, eps_sfc=eps, source_function=sf,
                                T_ave=T_ave if sf == "cim" else None, ctx=ctx)
        ref = refcpu.schwarzschild_rte(g.nu_min, g.dnu, tau, T_lev, 300.0, 0.8, eps,
                                       1 if sf == "cim" else 0, T_ave)
        assert np.max(np.abs(got - ref) / ref) < 1e-12


def test_ils_kernel_matches_oracle(ctx):
    from irsounderlbl.api import ils_kernel
    for apod in range(4):
        _, got = ils_kernel(0.001, 2.0, 0.5, apodization=apod, ctx=ctx)
        _, ref = refcpu.ils_kernel(0.001, 2.0, 0.5, apod)
        assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))


def test_forward_config0_bt(ctx):
    from irsounderlbl.api import forward_model
    prof = syn.base_profile(False)
    ll = syn.linelist_config0()
    snd = Sounder.make(645.0, 800.0, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    nu, R, BT = forward_model(prof, {GasSpecies.CO2: ll}, sounder=snd, ctx=ctx, apply_continuum=False)
    Ro, BTo, info = refcpu.forward(prof, {GasSpecies.CO2: ll}, snd)
    d = np.abs(BT - BTo)
    print(f"config0 BT max|d| {d.max():.2e} K rms {np.sqrt((d**2).mean()):.2e} K; "
          f"stats {ctx.stats()}")
    assert len(BT) == 621
    assert d.max() < BT_TOL


def test_forward_full_range_subset_bt(ctx):
    """Config-1 line lists (3 species, VP_Y) on a window sub-range with a nonblack
    surface and slant path, vs the oracle."""
    from irsounderlbl.api import forward_model
    lls = syn.linelists_config1()
    prof = syn.base_profile(True)
    snd = Sounder.make(700.0, 730.0, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    sub = {k: filter_linelist(v, 650.0, 780.0) for k, v in lls.items()}
    for lm in (None, "vpy"):
        nu, R, BT = forward_model(prof, sub, sounder=snd, line_mixing=lm, ctx=ctx, apply_continuum=False)
        Ro, BTo, info = refcpu.forward(prof, sub, snd, line_mixing=1 if lm else 0)
        d = np.abs(BT - BTo)
        print(f"subset lm={lm} BT max|d| {d.max():.2e} K")
        assert d.max() < BT_TOL


@pytest.mark.parametrize("lo,hi", [(1205.0, 1215.0), (778.0, 784.0), (2420.0, 2426.0)])
def test_forward_windows_variant_bt(ctx, lo, hi):
    """Brightness temperatures where the spectrum is partly transparent (the
    config-1 line lists with the H2O-like intensities of the windows variant:
    BT 240-275 K here, so every layer's optical depth matters), slant path and
    a non-black surface, with and without VP_Y (778-784 cm^-1 carries the wings
    of the line-mixing CO2 lines), vs the FP64 oracle at 0.005 K."""
    from irsounderlbl.api import forward_model
    from irsounderlbl.types import ViewingGeometry
    lls = syn.linelists_config1(windows=True)
    prof = syn.base_profile(True)
    snd = Sounder.make(lo, hi, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    sub = {k: filter_linelist(v, lo - 42.0, hi + 42.0) for k, v in lls.items()}
    mu, T_sfc, eps = 0.8, 295.0, 0.97
    geom = ViewingGeometry(float(np.degrees(np.arccos(mu))), 0.0, mu)
    for lm in ((None, "vpy") if lo < 800.0 else ("vpy",)):
        nu, R, BT = forward_model(prof, sub, sounder=snd, geom=geom, T_sfc=T_sfc, eps_sfc=eps, line_mixing=lm,
                                  ctx=ctx, apply_continuum=False)
        Ro, BTo, info = refcpu.forward(prof, sub, snd, line_mixing=1 if lm else 0, mu=mu, T_sfc=T_sfc,
                                       eps_sfc=eps)
        assert BTo.min() > 230.0 and BTo.max() - BTo.min() > 5.0        # not saturated, not flat
        d = np.abs(BT - BTo)
        print(f"windows {lo}-{hi} lm={lm}: BT {BTo.min():.1f}-{BTo.max():.1f} K, max|d| {d.max():.2e} K")
        assert d.max() < BT_TOL


@pytest.mark.parametrize("lo,hi,lm", [(1205.0, 1213.0, None), (778.0, 784.0, "vpy")])
def test_fp64_every_pair_forward_matches_oracle(ctx, lo, hi, lm):
    """The every-pair FP64 line sum (the reference kernel's arithmetic on the
    device, fp64_every_pair schedule) against the FP64 oracle: the two compute
    the same sums in the same precision, so BT agrees to ~1e-8 K (the bar is
    1e-6 K), and the FP32 engine lies within its own 0.005 K bar of both."""
    from irsounderlbl.api import forward_model
    from irsounderlbl.types import ViewingGeometry
    lls = syn.linelists_config1(windows=True)
    prof = syn.base_profile(True)
    snd = Sounder.make(lo, hi, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    sub = {k: filter_linelist(v, lo - 42.0, hi + 42.0) for k, v in lls.items()}
    mu, eps = 0.8, 0.97
    geom = ViewingGeometry(float(np.degrees(np.arccos(mu))), 0.0, mu)
    kw = dict(sounder=snd, geom=geom, eps_sfc=eps, line_mixing=lm, ctx=ctx)
    try:
        ctx.set_schedule(4, fp64_every_pair=True)
        _, R64, BT64 = forward_model(prof, sub, apply_continuum=False, **kw)
        st = ctx.stats()
        _, _, BT64c = forward_model(prof, sub, **kw)                    # with continuum
    finally:
        ctx.set_schedule(4)
    BT32 = forward_model(prof, sub, apply_continuum=False, **kw)[2]
    _, BTo, _ = refcpu.forward(prof, sub, snd, line_mixing=1 if lm else 0, mu=mu, eps_sfc=eps)
    _, BToc, _ = refcpu.forward(prof, sub, snd, line_mixing=1 if lm else 0, mu=mu, eps_sfc=eps, continua=31)
    d64 = np.abs(BT64 - BTo).max()
    print(f"FP64 every pair {lo}-{hi}: max|dBT| vs oracle {d64:.2e} K (with continuum "
          f"{np.abs(BT64c - BToc).max():.2e}), FP32 vs FP64 {np.abs(BT32 - BT64).max():.2e} K, "
          f"line sum {st['ms_linesum']:.1f} ms")
    assert d64 < 1e-6
    assert np.abs(BT64c - BToc).max() < 1e-6
    assert np.abs(BT32 - BT64).max() < BT_TOL
