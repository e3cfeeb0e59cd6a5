"""GPU line-sum parity against the FP64 oracle (cross sections, one layer)."""
import numpy as np
import pytest

import refcpu
from helpers import four_line_fixture, rel_err_masked
from irsounderlbl import synthetic as syn
from irsounderlbl.types import GasSpecies, HITRANLinelist, filter_linelist, wavenumber_grid

pytestmark = pytest.mark.gpu

TOL = 1e-4   # relative, where sigma > 1e-6 * peak


def test_four_line_fixture_matches_oracle(ctx):
    from irsounderlbl.api import compute_voigt_cross_sections
    ll = four_line_fixture()
    g = wavenumber_grid(2345.0, 2355.0, 0.001)
    for T, p in [(230.0, 0.2), (296.0, 1.0), (200.0, 1e-4)]:
        got = compute_voigt_cross_sections(g, ll, T, p, cutoff=25.0, ctx=ctx)
        ref = refcpu.voigt_xsec(ll, g.nu_min, g.dnu, g.n, T, p, cutoff=25.0)
        e, _ = rel_err_masked(got, ref)
        print(f"4-line T={T} p={p}: max rel err {e:.2e}")
        assert e < TOL
        assert np.abs(got - ref).max() < 1e-5 * ref.max()


def test_x_far_branch_live_and_lossless(ctx):
    """runtests.jl:247-273 on the GPU path: x_far = 122 vs exact everywhere, and 0."""
    from irsounderlbl.api import compute_voigt_cross_sections
    ll = four_line_fixture()
    g = wavenumber_grid(2345.0, 2355.0, 0.001)
    exact = refcpu.voigt_xsec(ll, g.nu_min, g.dnu, g.n, 230.0, 0.2, x_far=np.inf)
    opt_a = compute_voigt_cross_sections(g, ll, 230.0, 0.2, x_far=122.0, ctx=ctx)
    allz = compute_voigt_cross_sections(g, ll, 230.0, 0.2, x_far=0.0, ctx=ctx)
    peak = exact.max()
    assert np.abs(opt_a - exact).max() < 1e-5 * peak
    m = exact > 1e-6 * peak
    assert (np.abs(opt_a[m] - exact[m]) / exact[m]).max() < 3e-4
    assert np.abs(allz - exact).max() > 1e-3 * peak


def _layers_config0():
    prof = syn.base_profile(False)
    lp = refcpu.layer_properties(prof)
    return prof, lp


@pytest.mark.parametrize("k", [0, 24, 48])
def test_config0_layer_matches_oracle(ctx, k):
    from irsounderlbl.api import compute_voigt_cross_sections
    prof, lp = _layers_config0()
    ll = syn.linelist_config0()
    T = lp["T_cg"][GasSpecies.CO2][k]
    p = lp["p_cg"][GasSpecies.CO2][k] / 1013.25
    g = wavenumber_grid(629.0, 816.0, 0.001)
    got = compute_voigt_cross_sections(g, ll, T, p, ctx=ctx)
    ref = refcpu.voigt_xsec(ll, g.nu_min, g.dnu, g.n, T, p)
    e, i = rel_err_masked(got, ref)
    print(f"layer {k} T={T:.2f} p={p:.3e}: max rel err {e:.2e} at {g.nu_min + i * g.dnu:.3f}")
    assert e < TOL


def test_dptmin_live_matches_oracle(ctx):
    from irsounderlbl.api import compute_voigt_cross_sections
    prof, lp = _layers_config0()
    ll = syn.linelist_config0()
    k = 45
    T = lp["T_cg"][GasSpecies.CO2][k]
    p = lp["p_cg"][GasSpecies.CO2][k] / 1013.25
    coef = lp["vmr_cg"][GasSpecies.CO2][k] * lp["dp"][k] * 2.1209e22
    keep = refcpu.reject_weak_lines(ll, T, p, coef, dptmn=1e-6)
    assert 0 < keep.sum() < len(ll)
    g = wavenumber_grid(629.0, 816.0, 0.001)
    got = compute_voigt_cross_sections(g, ll, T, p, dptmn=1e-6, coef=coef, ctx=ctx)
    ref = refcpu.voigt_xsec(ll, g.nu_min, g.dnu, g.n, T, p, keep=keep)
    e, _ = rel_err_masked(got, ref)
    print(f"DPTMIN layer {k}: kept {keep.sum()}/{len(ll)}, max rel err {e:.2e}")
    assert e < TOL


def test_vpy_line_mixing_matches_oracle(ctx):
    from irsounderlbl.api import compute_voigt_lm_cross_sections
    lls = syn.linelists_config1()
    co2 = filter_linelist(lls[GasSpecies.CO2], 680.0, 720.0)
    assert co2.y_lm is not None and np.count_nonzero(co2.y_lm) > 1000
    g = wavenumber_grid(695.0, 705.0, 0.001)
    for T, p in [(288.0, 0.9), (217.0, 0.05), (217.0, 1e-4)]:
        got = compute_voigt_lm_cross_sections(g, co2, T, p, ctx=ctx)
        ref = refcpu.voigt_lm_xsec(co2, g.nu_min, g.dnu, g.n, T, p)
        e, _ = rel_err_masked(got, ref)
        print(f"VP_Y T={T} p={p}: max rel err {e:.2e}")
        assert e < TOL


def test_empty_and_single_line(ctx):
    from irsounderlbl.api import compute_voigt_cross_sections
    g = wavenumber_grid(700.0, 701.0, 0.001)
    one = HITRANLinelist(1, 1, [700.5], [1e-20], [0.07], [0.3], [100.0], [0.7], [-0.001])
    got = compute_voigt_cross_sections(g, one, 250.0, 0.5, ctx=ctx)
    ref = refcpu.voigt_xsec(one, g.nu_min, g.dnu, g.n, 250.0, 0.5, vmr_self=0.0)
    assert rel_err_masked(got, ref)[0] < TOL
    far = HITRANLinelist(1, 1, [900.0], [1e-20], [0.07], [0.3], [100.0], [0.7], [0.0])
    assert np.all(compute_voigt_cross_sections(g, far, 250.0, 0.5, ctx=ctx) == 0.0)


@pytest.mark.parametrize("lm", [False, True])
def test_coarse_far_wing_matches_exact_sum(ctx, lm):
    """The coarse-grid far-wing sum (cubic interpolation beyond 1.5 cm^-1)
    against the every-pair sum on a dense 15 um band layer, and both against
    the FP64 oracle (1e-4 where sigma > 1e-6 peak)."""
    import refcpu as o
    from irsounderlbl import synthetic as syn
    from irsounderlbl.api import compute_voigt_cross_sections, compute_voigt_lm_cross_sections
    from irsounderlbl.types import GasSpecies, filter_linelist, wavenumber_grid
    co2 = filter_linelist(syn.linelists_config1()[GasSpecies.CO2], 640.0, 760.0)
    g = wavenumber_grid(690.0, 700.0, 0.001)
    sub = filter_linelist(co2, 664.0, 726.0)
    T, p = 240.0, 0.3
    def run():
        if lm:
            return compute_voigt_lm_cross_sections(g, sub, T, p, ctx=ctx)
        return compute_voigt_cross_sections(g, sub, T, p, ctx=ctx)
    try:
        ctx.set_schedule(4, far_wing_cm=0.0)
        exact = run()
        ctx.set_schedule(4, far_wing_cm=1.5)
        far = run()
        st = ctx.stats()
    finally:
        ctx.set_schedule(4)
    assert st["coarse_pairs"] > 0 and st["fine_pairs"] > 0
    m = exact > 1e-6 * exact.max()
    assert np.max(np.abs(far[m] - exact[m]) / exact[m]) < 5e-6
    ref = (o.voigt_lm_xsec(sub, g.nu_min, g.dnu, g.n, T, p) if lm
           else o.voigt_xsec(sub, g.nu_min, g.dnu, g.n, T, p, 0.0))
    mr = ref > 1e-6 * ref.max()
    assert np.max(np.abs(far[mr] - ref[mr]) / ref[mr]) < 1e-4


@pytest.mark.parametrize("lm", [False, True])
@pytest.mark.parametrize("cutoff", [25.0, 6.0])
def test_edge_lines_by_local_expansion_match_oracle(ctx, lm, cutoff):
    """A grid that only EDGE lines reach: CO2 lines within 1.5 cm^-1 of one
    cutoff below and above a 2 cm^-1 grid, so that nearly every line's window
    boundary crosses it.  The line sum evaluates them by the per-slice cubic
    expansion (boundary slices pair by pair); the FP64 oracle sums every pair.
    With cutoff 6 cm^-1 the lines are ~5000 grid steps away, where the cubic
    terms of the expansion matter at the 1e-6 level.  Bar: 1e-4 relative where
    sigma > 1e-6 of the peak (observed <= 5.5e-5, at points where the pedestal
    cancels the accumulated wings to a few %) and 2e-5 of the peak anywhere."""
    import refcpu as o
    from irsounderlbl.api import compute_voigt_cross_sections, compute_voigt_lm_cross_sections
    co2 = syn.linelists_config1()[GasSpecies.CO2]
    g = wavenumber_grid(700.0, 702.0, 0.001)
    lo, hi = 701.0 - cutoff, 701.0 + cutoff
    m = (((co2.wavenumber >= lo - 1.5) & (co2.wavenumber <= lo + 1.5)) |
         ((co2.wavenumber >= hi - 1.5) & (co2.wavenumber <= hi + 1.5)))
    sub = co2.subset(m)
    assert len(sub) > 100
    for T, p in [(240.0, 0.3), (290.0, 1.0)]:
        if lm:
            got = compute_voigt_lm_cross_sections(g, sub, T, p, cutoff=cutoff, lm_cutoff=cutoff, ctx=ctx)
            ref = o.voigt_lm_xsec(sub, g.nu_min, g.dnu, g.n, T, p, cutoff=cutoff, lm_cutoff=cutoff)
        else:
            got = compute_voigt_cross_sections(g, sub, T, p, cutoff=cutoff, ctx=ctx)
            ref = o.voigt_xsec(sub, g.nu_min, g.dnu, g.n, T, p, 0.0, cutoff=cutoff)
        peak = np.abs(ref).max()
        assert peak > 0
        err = np.abs(got - ref).max() / peak
        e, _ = rel_err_masked(got, ref, 1e-6)
        print(f"EDGE-only lm={lm} cutoff={cutoff} T={T} p={p}: max abs err {err:.2e} of peak, rel {e:.2e} "
              f"(sigma > 1e-6 peak)")
        assert err < 2e-5
        assert e < 1e-4


@pytest.mark.parametrize("rule,re_tol,im_tol", [("core", 2.5e-6, 4e-5), ("pair", 6e-6, 1e-5)])
def test_faddeeva32_matches_wofz(ctx, rule, re_tol, im_tol):
    """The FP32 w(z) of the line sums (irs_faddeeva32) against scipy.special.wofz
    (FP64) on |x| <= 30, y in [1e-6, 30] at the FP32-rounded points: relative
    error of Re w (the Voigt H function) and of Im w against max(|Im w|,
    1e-2 |w|) (observed: core rule 1.5e-6 / 2.1e-5, pair rule 3.9e-6 / 5.4e-6,
    the latter at the |x| + y = 5.5 switch to the continued fraction); on the
    real axis Re w = exp(-x^2) (observed 4.4e-6 for |x| <= 5)."""
    from scipy.special import wofz
    from irsounderlbl.api import faddeeva32
    x = np.linspace(-30, 30, 2401).astype(np.float32).astype(np.float64)
    y = np.concatenate([[1e-6, 1e-4, 1e-3, 1e-2], np.linspace(0.02, 30, 400)]).astype(np.float32).astype(np.float64)
    X, Y = np.meshgrid(x, y)
    w = faddeeva32(X, Y, rule=rule, ctx=ctx)
    ref = wofz(X + 1j * Y)
    assert w.shape == X.shape
    assert np.max(np.abs(w.real - ref.real) / np.abs(ref.real)) < re_tol
    assert np.max(np.abs(w.imag - ref.imag) / np.maximum(np.abs(ref.imag), 1e-2 * np.abs(ref))) < im_tol
    x0 = np.linspace(-5, 5, 1001)
    w0 = faddeeva32(x0, 0.0, rule=rule, ctx=ctx)
    assert np.max(np.abs(w0.real - np.exp(-x0 ** 2)) / np.exp(-x0 ** 2)) < 1e-5
    xs = np.linspace(0.01, 5.0, 500)          # w(-x + iy) = conj w(x + iy), bit for bit off x = 0
    assert np.array_equal(faddeeva32(-xs, 0.5, rule=rule, ctx=ctx), np.conj(faddeeva32(xs, 0.5, rule=rule, ctx=ctx)))
    with pytest.raises(ValueError):
        faddeeva32([1.0], [-0.1], rule=rule, ctx=ctx)
