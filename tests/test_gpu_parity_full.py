"""Parity of the device line sum and forward model against the FP64 oracle
where the earlier suites did not look:

* the O3-like band (980-1080 cm^-1): cross sections and brightness temperatures;
* the full internal grid of the headline workload (629-2776 cm^-1,
  2 147 001 points) at sampled layers (0, 24, 48), every species with its own
  CG state, VP_Y on the CO2 list: sigma within 1e-4 where sigma > 1e-6 peak;
* the DPTMIN decisions (src/CrossSections/voigt.jl:333-355): the device's
  active line count of every (layer, species) of configs 0 and 1 equals the
  oracle's, and forward_model's total equals the oracle forward's;
* forward_model option paths: with_ils=false, high_res_factor=4, the three
  Norton-Beer apodizations, source_function=:toon, and the interpolating
  channel resample of apply_ils (src/Sensor/sounder.jl:372-377) on a grid
  whose channels are not internal nodes.
"""
import numpy as np
import pytest

import refcpu as o
from irsounderlbl import synthetic as syn
from irsounderlbl.api import (apply_ils_channels, compute_voigt_cross_sections,
                              compute_voigt_lm_cross_sections, forward_model, ils_kernel, layer_properties)
from irsounderlbl.types import (GasSpecies, IASIInstrument, Sounder, filter_linelist, sounder_grid,
                                wavenumber_grid)

from helpers import rel_err_masked

pytestmark = pytest.mark.gpu

TOL = 1e-4
BT_TOL = 0.005


def _layer_state(prof, k, sp):
    lp = layer_properties(prof)
    vmr = lp["vmr_cg"][sp][k]
    return lp["T_cg"][sp][k], lp["p_cg"][sp][k] / 1013.25, vmr, vmr * lp["dp"][k] * 2.1209e22


@pytest.mark.parametrize("k", [0, 24, 48])
def test_o3_band_cross_sections(ctx, k):
    prof = syn.base_profile(True)
    o3 = syn.linelists_config1()[GasSpecies.O3]
    T, p, _, _ = _layer_state(prof, k, GasSpecies.O3)
    g = wavenumber_grid(1000.0, 1060.0, 0.001)
    sub = filter_linelist(o3, 974.0, 1086.0)
    assert len(sub) > 10000
    got = compute_voigt_cross_sections(g, sub, T, p, ctx=ctx)
    ref = o.voigt_xsec(sub, g.nu_min, g.dnu, g.n, T, p, 0.0)
    e, _ = rel_err_masked(got, ref)
    print(f"O3 band layer {k}: T={T:.1f} p={p:.3e} max rel {e:.2e}")
    assert e < TOL


def test_o3_window_brightness_temperature(ctx):
    """The O3 band over partly transparent H2O windows (windows variant), slant path."""
    from irsounderlbl.types import ViewingGeometry
    prof = syn.base_profile(True)
    lls = syn.linelists_config1(windows=True)
    lo, hi = 1020.0, 1040.0
    sub = {kk: filter_linelist(v, lo - 42.0, hi + 42.0) for kk, v in lls.items()}
    snd = Sounder.make(lo, hi, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    mu = 0.85
    geom = ViewingGeometry(float(np.degrees(np.arccos(mu))), 0.0, mu)
    _, _, bt = forward_model(prof, sub, sounder=snd, geom=geom, ctx=ctx, apply_continuum=False)
    _, bto, _ = o.forward(prof, sub, snd, mu=mu)
    _, _, bt_no3 = forward_model(prof, {kk: v for kk, v in sub.items() if kk != GasSpecies.O3}, sounder=snd,
                                 geom=geom, ctx=ctx, apply_continuum=False)
    d = np.abs(bt - bto)
    print(f"O3 window BT {bto.min():.2f}-{bto.max():.2f} K, max|d| {d.max():.2e} K, "
          f"O3 effect {np.abs(bt - bt_no3).max():.2f} K")
    assert np.abs(bt - bt_no3).max() > 0.1            # the O3 lines matter here
    assert d.max() < BT_TOL


@pytest.mark.parametrize("k", [0, 24, 48])
def test_full_range_sampled_layer_cross_sections(ctx, k):
    """Every species of the headline line lists on the whole internal IASI grid
    at one layer's CG state (VP_Y on CO2), against the oracle."""
    prof = syn.base_profile(True)
    lls = syn.linelists_config1()
    lo, d, n, _ = 629.0, 0.001, 2147001, 8461
    g = wavenumber_grid(lo, 2776.0, d)
    assert g.n == n
    for sp, ll in lls.items():
        T, p, vmr, _ = _layer_state(prof, k, sp)
        if sp == GasSpecies.CO2:
            got = compute_voigt_lm_cross_sections(g, ll, T, p, ctx=ctx)
            ref = o.voigt_lm_xsec(ll, g.nu_min, g.dnu, g.n, T, p)
        else:
            vs = vmr if sp == GasSpecies.H2O else 0.0
            got = compute_voigt_cross_sections(g, ll, T, p, vmr_self=vs, ctx=ctx)
            ref = o.voigt_xsec(ll, g.nu_min, g.dnu, g.n, T, p, vs)
        e, _ = rel_err_masked(got, ref)
        m = ref > 1e-6 * ref.max()
        print(f"full range layer {k} {sp.name}: T={T:.1f} p={p:.3e}, {m.sum()} points above 1e-6 peak, "
              f"max rel {e:.2e}")
        assert m.sum() > 20000                    # (O3 at the top layer: 4.9e4 points)
        assert e < TOL


def _dptmin_items(prof, lls):
    lp = layer_properties(prof)
    for k in range(prof.n_levels - 1):
        for sp, ll in lls.items():
            vmr = lp["vmr_cg"][sp][k]
            if vmr == 0.0:
                continue
            T, p = lp["T_cg"][sp][k], lp["p_cg"][sp][k] / 1013.25
            coef = vmr * lp["dp"][k] * 2.1209e22
            yield k, sp, ll, T, p, (vmr if sp == GasSpecies.H2O else 0.0), coef


@pytest.mark.parametrize("config", [0, 1])
def test_dptmin_active_sets_match_oracle(ctx, config):
    """Per (layer, species): the number of lines the device keeps (strict >,
    erfcx peak) equals the oracle's, at the reference threshold 1e-6."""
    if config == 0:
        prof, lls = syn.base_profile(False), {GasSpecies.CO2: syn.linelist_config0()}
    else:
        prof, lls = syn.base_profile(True), syn.linelists_config1(with_lm=False)
    g = wavenumber_grid(700.0, 700.0, 0.001)          # one point: the decisions do not depend on the grid
    n_items = n_diff = 0
    tot_dev = tot_ref = 0
    for k, sp, ll, T, p, vs, coef in _dptmin_items(prof, lls):
        compute_voigt_cross_sections(g, ll, T, p, vmr_self=vs, dptmn=1e-6, coef=coef, ctx=ctx)
        dev = ctx.stats()["active_lines"]
        ref = int(o.reject_weak_lines(ll, T, p, coef, vs, 1e-6).sum())
        n_items += 1
        n_diff += dev != ref
        tot_dev += dev
        tot_ref += ref
    print(f"config {config}: {n_items} items, {tot_ref} active lines (oracle), {tot_dev} (device), "
          f"{n_diff} items differ")
    assert 0 < tot_ref < sum(len(ll) for ll in lls.values()) * (prof.n_levels - 1)   # DPTMIN rejects some
    assert n_diff == 0


def test_forward_active_line_total_matches_oracle(ctx):
    """forward_model with DPTMIN live: total active lines over all items equals
    the oracle forward's (config 0, full profile)."""
    prof = syn.base_profile(False)
    ll = syn.linelist_config0()
    snd = Sounder.make(645.0, 800.0, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    forward_model(prof, {GasSpecies.CO2: ll}, sounder=snd, ctx=ctx, apply_continuum=False)
    dev = ctx.stats()["active_lines"]
    _, _, info = o.forward(prof, {GasSpecies.CO2: ll}, snd)
    print(f"active lines: device {dev}, oracle {info.active_lines}")
    assert dev == info.active_lines


def _small_forward_problem():
    prof = syn.base_profile(True)
    lls = syn.linelists_config1(windows=True)
    lo, hi = 1205.0, 1213.0
    sub = {kk: filter_linelist(v, lo - 42.0, hi + 42.0) for kk, v in lls.items()}
    return prof, sub, Sounder.make(lo, hi, 0.25, opd_max=2.0, fwhm_gauss=0.5)


@pytest.mark.parametrize("case", ["no_ils", "hrf4", "nb_weak", "nb_medium", "nb_strong", "toon"])
def test_forward_option_paths(ctx, case):
    prof, sub, snd = _small_forward_problem()
    kw, okw = {}, {}
    if case == "no_ils":
        kw, okw = dict(with_ils=False), dict(with_ils=False)
    elif case == "hrf4":
        kw, okw = dict(high_res_factor=4), dict(high_res_factor=4)
    elif case.startswith("nb_"):
        apod = {"nb_weak": 1, "nb_medium": 2, "nb_strong": 3}[case]
        kw, okw = dict(apodization="norton_beer_" + case[3:]), dict(apodization=apod)
    elif case == "toon":
        kw, okw = dict(source_function="toon"), dict(source_function=0)
    _, _, bt = forward_model(prof, sub, sounder=snd, ctx=ctx, apply_continuum=False, **kw)
    _, bto, _ = o.forward(prof, sub, snd, **okw)
    _, _, bt_def = forward_model(prof, sub, sounder=snd, ctx=ctx, apply_continuum=False)
    d = np.abs(bt - bto)
    print(f"{case}: BT {bto.min():.2f}-{bto.max():.2f} K, max|d| {d.max():.2e} K, "
          f"vs default path {np.abs(bt - bt_def).max():.3f} K")
    assert np.abs(bt - bt_def).max() > 1e-3          # the option changes the result
    assert d.max() < BT_TOL


def test_channel_resample_interpolates_between_nodes(ctx):
    """apply_ils on a grid whose channel wavenumbers are not internal nodes:
    the reference convolves every point and interpolates linearly
    (sounder.jl:372-377); the device evaluates the convolution at the two
    bracketing nodes and interpolates.  Checked against the oracle's direct
    convolution at those nodes, interpolated with the same weights."""
    rng = np.random.default_rng(5)
    g = wavenumber_grid(690.0, 740.0, 0.003)           # 705 + 0.25 c is not a node of 690 + 0.003 i
    spec = 80.0 + 10.0 * np.sin(g.nu * 3.1) + rng.uniform(0, 1, g.n)
    snd = Sounder.make(708.0, 722.0, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    got = apply_ils_channels(g, spec, snd, ctx=ctx)
    _, kern = o.ils_kernel(g.dnu, 2.0, 0.5, 0)
    ch = sounder_grid(snd).nu
    x = (ch - g.nu_min) / g.dnu
    k0 = np.floor(x).astype(np.int64)
    w = x - k0
    assert np.sum((w > 0.1) & (w < 0.9)) > len(ch) // 2              # most channels lie between nodes
    r0 = o.apply_ils_at(spec, g.dnu, kern, k0)
    r1 = o.apply_ils_at(spec, g.dnu, kern, k0 + 1)
    ref = (1.0 - w) * r0 + w * r1
    assert np.max(np.abs(got - ref) / np.abs(ref)) < 1e-12
