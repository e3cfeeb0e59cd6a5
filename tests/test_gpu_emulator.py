"""Emulator path: channel optical depths (training target), the BF16 MFMA
MLP, PCA compression, emulated forward model with LBL spot check and
fallback, emulator Jacobians."""
import numpy as np
import pytest

import refcpu as o
from helpers import channel_od_ref
from irsounderlbl import synthetic as syn
from irsounderlbl.api import channel_optical_depths, channel_transmittance
from irsounderlbl.types import GasSpecies, Sounder, ViewingGeometry, filter_linelist

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lo,mu,windows", [(700.0, 1.0, False), (700.0, 0.7, False), (1205.0, 0.7, True)])
def test_channel_od_matches_oracle(ctx, lo, mu, windows):
    """GPU channel optical depths and CIM weights vs the FP64 restatement on
    the oracle's tau (12 cm^-1 of channels, 49 layers, nadir and slant): in
    the opaque 15 um band with VP_Y, and in a partly transparent H2O window
    (the windows variant of the line lists, where the whole column transmits):
    1e-4 relative + 2e-5 absolute, away from the floor switch."""
    lls = syn.linelists_config1(windows=windows)
    sub = {k: filter_linelist(v, lo - 42.0, lo + 12.0 + 42.0) for k, v in lls.items()}
    snd = Sounder.make(lo, lo + 12.0, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    prof = syn.base_profile(True)
    floor = 1e-4
    geom = ViewingGeometry(float(np.degrees(np.arccos(mu))), 0.0, mu)
    got, g = channel_optical_depths(prof, sub, sounder=snd, geom=geom, floor=floor, line_mixing="vpy",
                                    cim=True, ctx=ctx, apply_continuum=False)
    _, _, info, tau = o.forward(prof, sub, snd, want_tau=True, line_mixing=1)
    _, kern = o.ils_kernel(info.dnu, snd.opd_max, snd.fwhm_gauss)
    ch = lo + 0.25 * np.arange(got.shape[1])
    ref, Tch, gref = channel_od_ref(tau, info.nu_lo, info.dnu, ch, kern, floor, mu=mu)
    assert got.shape == ref.shape == g.shape == (49, 49)
    near = np.abs(Tch[1:] - floor) < 1e-2 * floor          # branch switch at the floor
    err = np.abs(got - ref)
    ok = err <= 2e-5 + 1e-4 * np.abs(ref)
    assert np.all(ok | near), float(np.max(np.where(near, 0, err)))
    gerr = np.abs(g - gref)
    assert np.all((gerr <= 2e-5 + 1e-4 * np.abs(gref)) | near), float(np.max(np.where(near, 0, gerr)))
    # level transmittances and raw CIM sums (the emulator's target)
    Tl, G = channel_transmittance(prof, sub, sounder=snd, geom=geom, line_mixing="vpy", ctx=ctx, apply_continuum=False)
    assert np.max(np.abs(Tl - Tch)) < 2e-6 + 1e-4 * np.abs(Tch).max()
    Gr = channel_od_ref.last_G
    assert np.max(np.abs(G - Gr)) < 1e-6 + 1e-4 * np.abs(Gr).max()
    # telescoping: sum_k tau_ch_k = -ln T_0 where the whole column transmits
    full = Tch[0] >= floor
    if windows:
        assert full.all()
    if np.any(full):
        s = got[:, full].astype(np.float64).sum(axis=0)
        assert np.allclose(s, -np.log(Tch[0, full]), rtol=1e-4, atol=1e-4)


def test_channel_od_ils_moments(ctx):
    """ILS first moments of the level transmittance steps (irs_channel_od_moments)
    vs the FP64 restatement on the oracle's tau at the edge of the 4.3 um CO2
    band, where the +-16 cm^-1 ILS window spans opaque and transparent spectrum;
    and the channel-centre RTE with the moment term B'(nu_c) M_k reproduces the
    oracle's line-by-line brightness temperatures (to < 0.01 K) where the RTE
    without it is off by several hundredths of a kelvin."""
    from irsounderlbl.api import layer_properties
    lls = syn.linelists_config1(windows=True)
    lo, hi = 2406.0, 2412.0
    sub = {k: filter_linelist(v, lo - 42.0, hi + 42.0) for k, v in lls.items()}
    snd = Sounder.make(lo, hi, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    prof, T_sfc, mu = syn.perturbed_profiles(1, seed=4004, sigma_T=4.0, sigma_h2o=0.6, sigma_o3=0.2,
                                             co2_range=(390e-6, 430e-6))[0]
    floor = 1e-4
    geom = ViewingGeometry(float(np.degrees(np.arccos(mu))), 0.0, mu)
    od, g, m = channel_optical_depths(prof, sub, sounder=snd, geom=geom, floor=floor, line_mixing="vpy",
                                      moments=True, ctx=ctx, apply_continuum=False)
    od2, g2 = channel_optical_depths(prof, sub, sounder=snd, geom=geom, floor=floor, line_mixing="vpy", cim=True,
                                     ctx=ctx, apply_continuum=False)
    assert np.array_equal(od, od2) and np.array_equal(g, g2)
    _, BTo, info, tau = o.forward(prof, sub, snd, want_tau=True, line_mixing=1, mu=mu, T_sfc=T_sfc)
    _, kern = o.ils_kernel(info.dnu, snd.opd_max, snd.fwhm_gauss)
    ch = lo + 0.25 * np.arange(od.shape[1])
    ref, Tch, gref = channel_od_ref(tau, info.nu_lo, info.dnu, ch, kern, floor, mu=mu)
    Mref = channel_od_ref.last_M
    M = m.astype(np.float64) * Tch[1:]
    assert np.abs(Mref).max() > 1e-4
    assert np.max(np.abs(M - Mref)) < 1e-6 + 1e-4 * np.abs(Mref).max(), float(np.max(np.abs(M - Mref)))
    # channel RTE on the device's targets, Planck at the channel centre, with and without B' M
    C1, C2 = 1.191042953e-5, 1.4387769
    def B(T, v=ch):
        return C1 * v ** 3 / np.expm1(C2 * v / T)
    def dB(T):
        return (B(T, ch + 1e-3) - B(T, ch - 1e-3)) / 2e-3
    lp = layer_properties(prof)
    Ta, Tl = lp["T_ave"], prof.temperature
    top = np.ones(len(ch))
    R0, R1, Ms = B(T_sfc) * 0.0, B(T_sfc) * 0.0, np.zeros(len(ch))
    for k in range(od.shape[0] - 1, -1, -1):
        bot = top * np.exp(-od[k].astype(np.float64))
        G = g[k].astype(np.float64) * top
        term = B(Ta[k]) * (top - bot) + (B(Tl[k + 1]) - B(Ta[k])) * G
        R0 += term
        R1 += term + dB(Ta[k]) * m[k] * top
        Ms -= m[k] * top
        top = bot
    R0 += B(T_sfc) * top
    R1 += B(T_sfc) * top + dB(T_sfc) * Ms
    bt = lambda R: C2 * ch / np.log1p(C1 * ch ** 3 / R)
    e0, e1 = np.abs(bt(R0) - BTo).max(), np.abs(bt(R1) - BTo).max()
    print(f"channel RTE vs oracle LBL at {lo}-{hi}: max|dBT| {e0:.4f} K without the moment term, {e1:.5f} K with it")
    assert e0 > 0.02 and e1 < 0.01


def test_channel_od_transparent_and_flat(ctx):
    """No lines: every channel optical depth is 0.  Batched profiles give the
    same bits as single ones."""
    snd = Sounder.make(700.0, 703.0, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    prof = syn.base_profile(False)
    empty = {GasSpecies.CO2: filter_linelist(syn.linelist_config0(), 10.0, 11.0)}
    got = channel_optical_depths(prof, empty, sounder=snd, ctx=ctx, apply_continuum=False)
    assert np.all(got == 0.0)
    ll = {GasSpecies.CO2: filter_linelist(syn.linelist_config0(), 660.0, 745.0)}
    pp = [x[0] for x in syn.perturbed_profiles(3, seed=7, with_o3=False)]
    b = channel_optical_depths(pp, ll, sounder=snd, ctx=ctx, apply_continuum=False)
    for q in range(3):
        assert np.array_equal(b[q], channel_optical_depths(pp[q], ll, sounder=snd, ctx=ctx, apply_continuum=False))


# ---- BF16 MFMA MLP vs the FP64 reference (oracle/mlp_ref.py) -----------------
from mlp_ref import MLPRef  # noqa: E402
from irsounderlbl.emulator import (MLP, ChannelODEmulator, SpotChecker, forward_emulated,  # noqa: E402
                                   forward_emulated_steps)
from irsounderlbl.api import forward_model  # noqa: E402


def _mlp(seed=11, max_batch=512):
    m = MLP(5, 512, 256, max_batch, seed=seed)
    ref = MLPRef(m.get_params(), 5, m.n_in_pad, 512, 256)
    return m, ref


def test_mlp_forward_matches_reference():
    m, ref = _mlp()
    X = np.random.default_rng(1).normal(size=(200, 5)).astype(np.float32)   # not a multiple of 64
    y = m.forward(X)
    yr = ref.forward(X.astype(np.float64), emulate_bf16=True)
    y64 = ref.forward(X.astype(np.float64))
    scale = np.abs(yr).max()
    assert np.max(np.abs(y - yr)) < 2e-3 * scale
    assert np.max(np.abs(y - y64)) < 3e-2 * scale          # bf16 storage vs exact FP64
    m.close()


def test_mlp_loss_gradient_matches_reference():
    m, ref = _mlp(12)
    rng = np.random.default_rng(2)
    X = rng.normal(size=(128, 5)).astype(np.float32)
    T = rng.normal(size=(128, 256)).astype(np.float32)
    loss, g = m.loss_grad(X, T)
    lr, gr = ref.loss_grad(X.astype(np.float64), T.astype(np.float64), emulate_bf16=True)
    _, g64 = ref.loss_grad(X.astype(np.float64), T.astype(np.float64))
    assert abs(loss - lr) < 1e-4 * lr
    dims = ref.dims
    o = 0
    for l in range(4):
        for size in (dims[l] * dims[l + 1], dims[l + 1]):
            a, b, c = g[o:o + size], gr[o:o + size], g64[o:o + size]
            nb = np.linalg.norm(b)
            assert np.linalg.norm(a - b) < 1e-2 * nb + 1e-12, (l, size)
            assert np.linalg.norm(a - c) < 5e-2 * np.linalg.norm(c) + 1e-12, (l, size)
            o += size
    m.close()


def test_mlp_input_jacobian_matches_reference():
    m, ref = _mlp(13)
    X = np.random.default_rng(3).normal(size=(64, 5)).astype(np.float32)
    J = m.input_jacobian(X)
    Jr = ref.input_jacobian(X.astype(np.float64))
    assert np.linalg.norm(J - Jr) < 3e-2 * np.linalg.norm(Jr)
    m.close()


def test_mlp_training_reduces_loss():
    """Adam on a smooth 5 -> 256 regression (targets from a fixed random
    FP64 teacher network): the loss falls by more than 10x."""
    rng = np.random.default_rng(4)
    teacher = MLPRef(rng.normal(0, 0.1, MLPRef(np.zeros(10**7), 5, 64, 64, 256).n_params), 5, 64, 64, 256)
    X = rng.normal(size=(8192, 5))
    T = teacher.forward(X)
    T /= T.std()
    m = MLP(5, 512, 256, 1024, seed=5)
    m.set_data(X.astype(np.float32), T.astype(np.float32))
    losses = m.train(400, 512, 1e-3, seed=1)
    assert np.all(np.isfinite(losses))
    assert losses[-20:].mean() < 0.1 * losses[:5].mean()
    # the trained network evaluates to what its loss says
    y = m.forward(X[:512].astype(np.float32))
    assert np.mean((y - T[:512]) ** 2) < 2 * losses[-20:].mean() + 1e-4
    m.close()


# ---- emulator end to end on a 12 cm^-1 window --------------------------------
# A partly transparent H2O window of the windows variant (BT ~250-270 K), where
# the brightness temperatures depend on T_sfc, the temperature profile and H2O.
@pytest.fixture(scope="module")
def small_emu(ctx):
    lls = syn.linelists_config1(windows=True)
    sub = {k: filter_linelist(v, 1205.0 - 42.0, 1217.0 + 42.0) for k, v in lls.items()}
    snd = Sounder.make(1205.0, 1217.0, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    pp = syn.perturbed_profiles(264, seed=31, sigma_T=4.0, sigma_h2o=0.4)
    train, hold = pp[:200], pp[200:]
    emu, rep = ChannelODEmulator.train(train, sub, snd, n_pc=64, n_basis=128, chunk=64, steps=3000, batch=1024,
                                       lr=1e-3, n_hidden=512, ctx=ctx)
    return sub, snd, emu, rep, hold


def _hold(hold):
    return ([p[0] for p in hold], [ViewingGeometry(float(np.degrees(np.arccos(p[2]))), 0.0, p[2]) for p in hold],
            [p[1] for p in hold])


def test_emulator_trains_and_beats_regression_and_mean(ctx, small_emu):
    """Hold-out BT error of the trained emulator vs GPU LBL: the MLP on the
    regression residual gains over the regression alone (network off) and over
    the layer-mean predictor (regression and network off)."""
    sub, snd, emu, rep, hold = small_emu
    assert np.all(np.isfinite(rep.losses)) and rep.losses[-50:].mean() < 0.5 * rep.losses[:10].mean()
    assert rep.coef_rms_err_regression < 0.5 * rep.coef_rms
    profs, geoms, tsf = _hold(hold)
    _, _, bt_lbl = forward_model(profs, sub, sounder=snd, geom=geoms, T_sfc=tsf, line_mixing="vpy", ctx=ctx,
                                 apply_continuum=False)
    _, bt_emu = emu.forward(profs, geom=geoms, T_sfc=tsf)
    saved = emu.mlp.get_params()
    emu.mlp.set_params(np.zeros_like(saved))              # network off: the regression alone
    _, bt_reg = emu.forward(profs, geom=geoms, T_sfc=tsf)
    emu.mlp.set_params(saved)
    mean_only = ChannelODEmulator(emu.mlp, emu.x_mean, emu.x_std, np.zeros_like(emu.A), emu.coef_scale, emu.basis,
                                  emu.mean, emu.inv_w, emu.ch_nu, emu.context)
    emu.mlp.set_params(np.zeros_like(saved))
    _, bt_mean = mean_only.forward(profs, geom=geoms, T_sfc=tsf)
    emu.mlp.set_params(saved)
    mean_only.close()
    e = {k: float(np.sqrt(np.mean((b - bt_lbl) ** 2))) for k, b in
         (("trained", bt_emu), ("regression", bt_reg), ("mean", bt_mean))}
    print(f"hold-out RMS |dBT|: trained {e['trained']:.4f} K, regression alone {e['regression']:.4f} K, "
          f"layer mean {e['mean']:.4f} K; max trained {np.abs(bt_emu - bt_lbl).max():.3f} K")
    assert e["trained"] < e["regression"] < e["mean"]
    assert e["trained"] < 0.1


def test_emulator_calibrated_fallback_bounds_every_channel(ctx, small_emu):
    """Calibration on 32 held-out scenes without and with the spot-check correction.
    Uncorrected: spot blocks that flag the scenes over 0.8 x 0.05 K and the spot
    threshold under which every calibration scene is within that everywhere.
    Corrected: spot blocks whose errors predict every channel's error, the ridge map
    applied to every scene, and the threshold on the chosen statistic.  On 32 other
    held-out scenes no scene that the spot check lets through exceeds 0.05 K on any
    channel, fallen-back scenes are the LBL spectra, and the library's one-call path
    (irs_forward_emulated) equals the host-side restatement of the same steps."""
    sub, snd, emu, rep, hold = small_emu
    cal, tst = hold[:32], hold[32:]
    pc, gc, tc = _hold(cal)
    profs, geoms, tsf = _hold(tst)
    _, _, bt_lbl = forward_model(profs, sub, sounder=snd, geom=geoms, T_sfc=tsf, line_mixing="vpy", ctx=ctx,
                                 apply_continuum=False)
    _, bt_raw = emu.forward(profs, geom=geoms, T_sfc=tsf)
    for correct in (False, True):
        c = emu.calibrate(pc, sub, sounder=snd, geom=gc, T_sfc=tc, budget=0.05, n_blocks=2, block=8, ctx=ctx,
                          correct=correct)
        assert len(c.spot_channels) == 16 and c.ch_rms.shape == (49,)
        assert (c.W is not None) == correct
        r = forward_emulated(profs, sub, emu, sounder=snd, geom=geoms, T_sfc=tsf, ctx=ctx)
        d = np.abs(r.BT - bt_lbl).max(axis=1)
        print(f"correct={correct}: threshold {c.threshold:.4f} (weights {c.w_raw:.3g}, {c.w_lbo:.3g}) on blocks "
              f"{c.spot_blocks}: "
              f"fallback {r.fell_back.mean():.2f}, max |dBT| after fallback {d.max():.4f} K "
              f"(emulated scenes {np.sum(~r.fell_back)}; uncorrected max {np.abs(bt_raw - bt_lbl).max():.4f} K)")
        assert np.all(d[~r.fell_back] <= 0.05)
        assert np.array_equal(r.BT[r.fell_back], bt_lbl[r.fell_back])
        r2 = forward_emulated_steps(profs, sub, emu, sounder=snd, geom=geoms, T_sfc=tsf, ctx=ctx)
        assert np.array_equal(r.fell_back, r2.fell_back)
        np.testing.assert_allclose(r.spot_err, r2.spot_err, rtol=0, atol=1e-4)
        np.testing.assert_allclose(r.BT, r2.BT, rtol=0, atol=1e-4)
        if correct:
            # the spot channels carry their LBL values; the others moved by the fitted map
            np.testing.assert_allclose(r.BT[:, c.spot_channels], bt_lbl[:, c.spot_channels], rtol=0, atol=1e-4)
            # the library refuses spot blocks other than the correction's
            other = SpotChecker(sub, snd, blocks=[(float(snd.nu_min + (c0 - 1 if c0 > 0 else c0 + 1) * snd.dnu), n)
                                                  for c0, n in c.spot_blocks])
            with pytest.raises(ValueError, match="differ"):
                forward_emulated(profs[:2], sub, emu, sounder=snd, geom=geoms[:2], T_sfc=tsf[:2], spot=other,
                                 threshold=c.threshold, ctx=ctx, correction=True)
            # both statistics weighted (the leave-one-block-out residual computed by the
            # library): the library's statistic and decisions equal the host restatement's
            w_raw, w_lbo = c.w_raw, c.w_lbo
            c.w_raw, c.w_lbo = 1.0 / 0.03, 1.0 / 0.01
            emu._set_correction(c)
            try:
                r3 = forward_emulated(profs, sub, emu, sounder=snd, geom=geoms, T_sfc=tsf, threshold=1.0, ctx=ctx)
                r4 = forward_emulated_steps(profs, sub, emu, sounder=snd, geom=geoms, T_sfc=tsf, threshold=1.0,
                                            ctx=ctx)
                np.testing.assert_allclose(r3.spot_err, r4.spot_err, rtol=1e-4, atol=1e-4)
                assert np.array_equal(r3.fell_back, r4.fell_back)
                assert np.any(r3.spot_err > 0)
            finally:
                c.w_raw, c.w_lbo = w_raw, w_lbo
                emu._set_correction(c)


def test_emulator_spot_check_and_fallback(ctx, small_emu):
    sub, snd, emu, rep, hold = small_emu
    profs, geoms, tsf = _hold(hold[:16])
    spot = SpotChecker(sub, snd, blocks=((1206.0, 8), (1214.0, 8)))
    _, _, bt_lbl = forward_model(profs, sub, sounder=snd, geom=geoms, T_sfc=tsf, line_mixing="vpy", ctx=ctx,
                                 apply_continuum=False)
    # spot channels computed on sub-grids equal the full LBL at those channels
    bt_spot = spot(profs, geoms, tsf, [1.0] * len(profs), ctx, line_mixing="vpy", apply_continuum=False)
    assert np.max(np.abs(bt_spot - bt_lbl[:, spot.channels])) < 1e-4
    r_all = forward_emulated(profs, sub, emu, sounder=snd, geom=geoms, T_sfc=tsf, threshold=0.0, spot=spot, ctx=ctx)
    assert r_all.fell_back.all() and np.array_equal(r_all.BT, bt_lbl)
    r_none = forward_emulated(profs, sub, emu, sounder=snd, geom=geoms, T_sfc=tsf, threshold=1e9, spot=spot, ctx=ctx)
    assert not r_none.fell_back.any()
    _, bt_emu = emu.forward(profs, geom=geoms, T_sfc=tsf)
    # the library computes the scene features itself (float32 inputs equal to ~1 ulp)
    np.testing.assert_allclose(r_none.BT, bt_emu, rtol=0, atol=1e-5)
    thr = float(np.median(r_none.spot_err))
    r = forward_emulated(profs, sub, emu, sounder=snd, geom=geoms, T_sfc=tsf, threshold=thr, spot=spot, ctx=ctx)
    assert np.array_equal(r.fell_back, r_none.spot_err > thr)
    assert np.array_equal(r.BT[r.fell_back], bt_lbl[r.fell_back])
    assert np.all(r.spot_err[~r.fell_back] <= thr)
    # the one-call library path (irs_forward_emulated) and the host-side steps agree
    r2 = forward_emulated_steps(profs, sub, emu, sounder=snd, geom=geoms, T_sfc=tsf, threshold=thr, spot=spot,
                                ctx=ctx)
    np.testing.assert_allclose(r.spot_err, r2.spot_err, rtol=0, atol=1e-5)
    assert np.array_equal(r.fell_back, r2.fell_back)
    np.testing.assert_allclose(r.BT, r2.BT, rtol=0, atol=1e-5)


def test_emulator_save_load_roundtrip(ctx, small_emu, tmp_path):
    sub, snd, emu, rep, hold = small_emu
    profs, geoms, tsf = _hold(hold[:8])
    _, bt = emu.forward(profs, geom=geoms, T_sfc=tsf)
    path = str(tmp_path / "emu.npz")
    emu.save(path)
    emu2 = ChannelODEmulator.load(path)
    _, bt2 = emu2.forward(profs, geom=geoms, T_sfc=tsf)
    assert np.array_equal(bt, bt2)
    if emu.calibration is not None:
        assert emu2.calibration.threshold == emu.calibration.threshold
        assert np.array_equal(emu2.calibration.spot_channels, emu.calibration.spot_channels)
    emu2.close()
    emu2.mlp.close()


def test_emulator_jacobian_matches_finite_differences(ctx, small_emu):
    """Chain-rule emulator Jacobians (MLP input Jacobian + regression x decode x
    RTE adjoint x level->feature map, including the profile context) vs central
    differences of the emulator's own forward model, for level T, level ln q_H2O
    and T_sfc; then compared with the line-by-line analytic Jacobian (reported)."""
    sub, snd, emu, rep, hold = small_emu
    prof, T_sfc, mu = hold[0]
    geom = ViewingGeometry(float(np.degrees(np.arccos(mu))), 0.0, mu)
    jac = emu.jacobian([prof], geom=[geom], T_sfc=[T_sfc])[0]
    assert jac.K_T.shape == (49, 50) and jac.K_lnq_H2O.shape == (49, 50)
    def bt(p, ts=T_sfc):
        return emu.forward([p], geom=[geom], T_sfc=[ts])[1][0]
    # the emulator's response is a smooth function evaluated through BF16 inputs:
    # central differences at h = 0.5 K and 0.1 in ln q, compared on the whole
    # column (relative Frobenius norm) and channel by channel at 10% of the largest entry
    from irsounderlbl.types import AtmosphericProfile
    fdT = np.zeros_like(jac.K_T)
    fdq = np.zeros_like(jac.K_lnq_H2O)
    for j in range(50):
        h = 0.5
        Tp, Tm = prof.temperature.copy(), prof.temperature.copy()
        Tp[j] += h
        Tm[j] -= h
        fdT[:, j] = (bt(AtmosphericProfile(prof.pressure, Tp, prof.altitude, prof.vmr)) -
                     bt(AtmosphericProfile(prof.pressure, Tm, prof.altitude, prof.vmr))) / (2 * h)
        hq = 0.1
        vp, vm = dict(prof.vmr), dict(prof.vmr)
        q = prof.vmr[GasSpecies.H2O]
        vp[GasSpecies.H2O] = q * np.where(np.arange(len(q)) == j, np.exp(hq), 1.0)
        vm[GasSpecies.H2O] = q * np.where(np.arange(len(q)) == j, np.exp(-hq), 1.0)
        fdq[:, j] = (bt(AtmosphericProfile(prof.pressure, prof.temperature, prof.altitude, vp)) -
                     bt(AtmosphericProfile(prof.pressure, prof.temperature, prof.altitude, vm))) / (2 * hq)
    relT = np.linalg.norm(jac.K_T - fdT) / np.linalg.norm(fdT)
    relq = np.linalg.norm(jac.K_lnq_H2O - fdq) / np.linalg.norm(fdq)
    fds = (bt(prof, T_sfc + 0.5) - bt(prof, T_sfc - 0.5)) / 1.0
    print(f"chain rule vs FD of the emulator: relative Frobenius T {relT:.3f}, ln q {relq:.3f}; "
          f"T_sfc max {np.abs(fds - jac.K_Tsfc).max():.2e} of {np.abs(fds).max():.2e}")
    assert relT < 0.05 and relq < 0.1
    assert np.max(np.abs(jac.K_T - fdT)) < 0.1 * np.abs(fdT).max()
    assert np.max(np.abs(fds - jac.K_Tsfc)) < 1e-2 * max(np.abs(fds).max(), 1e-9)
    # vs the line-by-line analytic Jacobian (irs_jacobian, itself held to the
    # FP64 oracle in test_gpu_jacobian.py), reported
    from irsounderlbl.api import analytic_jacobian
    from irsounderlbl.types import StateVectorSpec
    ref = analytic_jacobian(prof, sub, StateVectorSpec(50, [GasSpecies.H2O], include_emissivity=False),
                            sounder=snd, geom=geom, T_sfc=T_sfc, line_mixing="vpy", vmr_coupling=False, ctx=ctx,
                            apply_continuum=False)
    KT_ref = ref.K[:, :50]
    rel = np.linalg.norm(jac.K_T - KT_ref) / np.linalg.norm(KT_ref)
    relq_lbl = np.linalg.norm(jac.K_lnq_H2O - ref.K[:, 50:100]) / np.linalg.norm(ref.K[:, 50:100])
    print(f"emulator vs LBL analytic Jacobian: relative Frobenius error T {rel:.3f}, ln q_H2O {relq_lbl:.3f}")
    assert rel < 0.5


def test_emulator_file_loaded_by_the_library_matches(ctx, small_emu, tmp_path):
    """export_binary -> irs_emu_load (the learned path's loader for C and Julia callers):
    the loaded emulator gives the same emulated spectra and the same corrected, spot-checked
    result of irs_forward_emulated as the emulator it was written from."""
    import ctypes as C
    from irsounderlbl import api
    from irsounderlbl.api import _p
    sub, snd, emu, rep, hold = small_emu
    profs, geoms, tsf = _hold(hold[:32])
    c = emu.calibrate(profs, sub, sounder=snd, geom=geoms, T_sfc=tsf, budget=0.05, n_blocks=2, block=8, ctx=ctx)
    path = str(tmp_path / "emu.bin")
    emu.export_binary(path)
    lib = emu.lib
    hm, he, thr, nbk = C.c_void_p(), C.c_void_p(), C.c_double(), C.c_int()
    c0 = np.zeros(8, np.int64)
    bn = np.zeros(8, np.int32)
    assert lib.irs_emu_load(path.encode(), 0, 4096, C.byref(hm), C.byref(he), C.byref(thr), 8, _p(c0), _p(bn),
                            C.byref(nbk)) == 0
    try:
        assert thr.value == c.threshold
        assert [(int(a), int(b)) for a, b in zip(c0[:nbk.value], bn[:nbk.value])] == \
            [(int(a), int(b)) for a, b in c.spot_blocks]
        tp, tg, tt = _hold(hold[32:48])
        scenes, eps = emu._scenes(tp, tg, tt, 1.0)
        X, tlev, tave, scal = emu._inputs(scenes, eps)
        n = len(tp)
        outs = []
        for h in (emu.h, he):
            R, BT = np.zeros((n, emu.n_ch)), np.zeros((n, emu.n_ch))
            assert lib.irs_emu_forward(h, _p(X), _p(tlev), _p(tave), _p(scal), n, api.SOURCE_FUNCTION["cim"],
                                       _p(R), _p(BT), None) == 0
            outs.append(BT)
        assert np.array_equal(outs[0], outs[1])
        r1 = forward_emulated(tp, sub, emu, sounder=snd, geom=tg, T_sfc=tt, ctx=ctx)

        class Loaded:                       # the library handle in the emulator's place
            h, lib, n_ch, calibration = he, emu.lib, emu.n_ch, emu.calibration

            def spot_checker(self, ll, s):
                return emu.spot_checker(ll, s)

        r2 = forward_emulated(tp, sub, Loaded(), sounder=snd, geom=tg, T_sfc=tt, ctx=ctx, correction=True)
        assert np.array_equal(r1.BT, r2.BT) and np.array_equal(r1.fell_back, r2.fell_back)
    finally:
        lib.irs_emu_destroy(he)
        lib.irs_mlp_destroy(hm)
