"""Benchmark: full-range IASI spectra/s on one MI355X (plain FP32 LBL).

Workload: one 50-level profile (49 layers) and 300,000 synthetic lines in
620–2785 cm^-1 (CO2-like 180k with first-order line mixing on the 15 um half,
H2O-like 100k, O3-like 20k), internal grid 0.001 cm^-1 (2,147,001 points),
IASI 645–2760 cm^-1 / 0.25 cm^-1 -> 8461 channels, Gaussian-apodized ILS.
A step is one forward_model call for one profile with the line lists already
resident in HBM.  Prints one JSON line.

A second figure, "batched", times one scan line of 120 perturbed profiles
(T GP(0, 3 K, 2 km), H2O x lognormal(0.3), T_sfc +-10 K, scan angles
U(0, 48.3 deg); DPTMIN live, so every profile has its own active line sets)
in one forward_model call with the autotuned schedule and hipGraph replay.

A third figure, "emulator", trains the channel-transmittance emulator on GPU
LBL targets (the windows variant of the line lists, whose spectrum has
partly transparent windows; the benchmark lists are opaque at every channel)
of --emu-train perturbed scenes (8192 by default, --emu-steps 96000 Adam steps;
T GP(0, 4 K), H2O x lognormal(0.6), O3 x lognormal(0.2), CO2 U(390, 430) ppm,
scan angles U(0, 48.3 deg)), calibrates it on --emu-cal held-out scenes (the
per-channel hold-out BT error against GPU LBL, the 4 x 16 spot channels where it
is largest and the spot threshold for a 0.05 K budget), then runs --emu-eval new
scenes (2048 by default) through emulator + spot check + full-LBL fallback and
measures the BT error after fallback against GPU LBL on every scene the emulator
answered (and against the FP64 oracle on two of them).  --emulator-only runs the
headline steps and this figure alone.

"lbl_jacobian" times one full-IASI analytic Jacobian (102 columns) of a
perturbed profile on the windows variant of the line lists.

"fp64_every_pair" times the same spectrum through the reference kernel's
arithmetic on the device (every pair, FP64 line parameters and sums, FP64
optical depths and RTE) and reports the BT difference of the FP32 engine to it.

"roofline" is the line-sum kernel: FP32 FLOP/s from the committed rocprofv3
counters of that kernel (SQ_INSTS_VALU_FLOPS_FP32 x 64 lanes per launch) over
the launch time measured here; the reference's every-pair FLOP count over the
same time is reported beside it.  "hbm" gives the RTE and channel-ILS kernels'
bytes per launch over their measured times against 8 TB/s.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-accuracy]
                    [--no-batch] [--batch N] [--no-exact] [--no-emulator] [--no-jacobian]
                    [--no-fp64] [--emu-train N] [--emu-cal N] [--emu-eval N] [--emu-steps N]
                    [--emulator-only] [--no-par]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "irsounderlbl.jl_amd"))

import numpy as np  # noqa: E402

PEAK_FP32_TFLOPS = 157.3       # MI355X published FP32 peak (packed FP32)
PEAK_FP32_NONPACKED_TFLOPS = 78.6
PEAK_HBM_GBPS = 8000.0
FLOP_PER_VOIGT_PAIR = 14       # nominal, far-wing expression of the reference kernel
FLOP_PER_DISP_PAIR = 9


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(prof, lls, snd):
    """The FP64 oracle (the reference algorithm restated in C, OpenMP over grid
    points inside each cross section, sequential layers) on this host:
    * config 0 (20k lines, 645-800 cm^-1): one full spectrum, with the
      reference's single-threaded full-grid direct ILS (measured);
    * config 1 (the headline workload): the line sums of the bottom, middle and
      top layers (all three species, VP_Y), x 49/3; the full RTE and channel ILS
      measured; the reference's single-threaded full-grid direct ILS timed on
      2000 points and scaled to the grid ("reference structure"); with the
      channel-only ILS ("optimized");
    * config 2: the middle layer of two perturbed profiles (per-profile rate)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refcpu
    from irsounderlbl import synthetic as syn
    from irsounderlbl.types import GasSpecies, Sounder
    L = prof.n_levels - 1
    threads = refcpu.lib().orc_num_threads()
    out = {"cores": threads, "host_threads": os.cpu_count(), "cpu_model": _cpu_model(), "kind": "port"}
    # config 0, full spectrum
    p0, ll0 = syn.base_profile(False), syn.linelist_config0()
    s0 = Sounder.make(645.0, 800.0, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    t0 = time.perf_counter()
    _, bt0, info0 = refcpu.forward(p0, {GasSpecies.CO2: ll0}, s0, ils_direct_full=True)
    ORACLE_CACHE["config0_bt"] = bt0
    out["config0_full_spectrum_s"] = round(time.perf_counter() - t0, 2)
    out["config0_seconds"] = {"linesum": round(info0.seconds_xsec, 2), "rte": round(info0.seconds_rte, 3),
                              "ils_full_grid_1thread": round(info0.seconds_ils, 2)}
    # config 1: three sampled layers
    layers = (0, L // 2, L - 1)
    t_ls = []
    for k in layers:
        mask = np.zeros(L, np.uint8)
        mask[k] = 1
        _, _, info = refcpu.forward(prof, lls, snd, layer_mask=mask, line_mixing=1)
        t_ls.append(info.seconds_xsec)
    ls = float(np.mean(t_ls)) * L
    rte, ils_ch = info.seconds_rte, info.seconds_ils
    # the reference's full-grid direct ILS, single-threaded, timed on 2000 points
    n_hi = info.n_hi
    idx = np.linspace(0, n_hi - 1, 2000).astype(np.int64)
    _, kern = refcpu.ils_kernel(info.dnu, snd.opd_max, snd.fwhm_gauss, 0)
    spec = np.ones(n_hi)
    refcpu.lib().orc_set_num_threads(1)
    try:
        t0 = time.perf_counter()
        refcpu.apply_ils_at(spec, info.dnu, kern, idx)
        t_ils_pt = (time.perf_counter() - t0) / len(idx)
    finally:
        refcpu.lib().orc_set_num_threads(threads)
    ils_full = t_ils_pt * n_hi
    ref_struct = ls + rte + ils_full
    optimized = ls + rte + ils_ch
    out.update({
        "value": 1.0 / ref_struct, "unit": "spectra/s",
        "sample": f"config 1: FP64 oracle line sums of layers {layers} (3 species, VP_Y) x {L}/3 = {ls:.1f} s, "
                  f"RTE {rte:.2f} s, the reference's single-threaded full-grid ILS {ils_full:.1f} s "
                  f"(2000 of {n_hi} points timed); value = 1 / their sum",
        "config1_seconds": {"linesum": round(ls, 2), "linesum_per_layer": [round(x, 3) for x in t_ls],
                            "rte": round(rte, 3), "ils_full_grid_1thread": round(ils_full, 2),
                            "ils_channels_only": round(ils_ch, 3)},
        "config1_reference_structure_spectra_per_s": 1.0 / ref_struct,
        "config1_optimized_spectra_per_s": 1.0 / optimized,
        "config1_linesum_only_spectra_per_s": 1.0 / ls})
    # config 2: two perturbed profiles, middle layer, per-profile rate
    pp = syn.perturbed_profiles(2, seed=2002)
    t2 = []
    for pr, _, mu in pp:
        mask = np.zeros(L, np.uint8)
        mask[L // 2] = 1
        _, _, info = refcpu.forward(pr, lls, snd, layer_mask=mask, line_mixing=1, mu=mu)
        t2.append(info.seconds_xsec * L + info.seconds_rte + ils_full)
    out["config2_per_profile_spectra_per_s"] = 2.0 / float(np.sum(t2))
    out["config2_sample"] = "2 perturbed profiles: middle-layer line sums x 49 + RTE + full-grid single-thread ILS"
    return out


# GPU results that the FP64 oracle checks in the CPU leg: (section, key, gpu
# value, callable that computes the oracle's value, kind).  Only the CPU leg
# runs the oracle, after every GPU figure is taken.
CHECKS = []
ORACLE_CACHE = {}


def check_bt(section, key, bt, oracle_bt):
    CHECKS.append((section, key, np.asarray(bt), oracle_bt, "bt"))


def accuracy_sample(ctx, prof, lls, lls_w):
    """GPU side of the accuracy samples: BT on a 10 cm^-1 channel window
    (1205-1215 cm^-1) of the windows variant of the line lists (the benchmark
    lists saturate every channel at 217 K, where BT would not test the optical
    depths; here BT is 250-270 K), with and without continuum, and the CO2
    cross section with line mixing of layer 24 on a 10 cm^-1 grid window in the
    15 um band.  The oracle values are computed in the CPU leg."""
    from irsounderlbl.api import compute_voigt_lm_cross_sections, forward_model, layer_properties
    from irsounderlbl.types import GasSpecies, Sounder, filter_linelist, wavenumber_grid
    lo, hi = 1205.0, 1215.0
    snd = Sounder.make(lo, hi, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    sub = {k: filter_linelist(v, lo - 42.0, hi + 42.0) for k, v in lls_w.items()}
    _, _, bt = forward_model(prof, sub, sounder=snd, line_mixing="vpy", ctx=ctx, apply_continuum=False)
    check_bt("accuracy_vs_fp64", "bt_1205-1215_windows_variant", bt,
             lambda o: o.forward(prof, sub, snd, line_mixing=1)[1])
    _, _, btc = forward_model(prof, sub, sounder=snd, line_mixing="vpy", ctx=ctx)
    check_bt("accuracy_vs_fp64", "bt_1205-1215_windows_variant_with_continuum", btc,
             lambda o: o.forward(prof, sub, snd, line_mixing=1, continua=31)[1])
    lp = layer_properties(prof)
    T = lp["T_cg"][GasSpecies.CO2][24]
    p = lp["p_cg"][GasSpecies.CO2][24] / 1013.25
    g = wavenumber_grid(700.0, 710.0, 0.001)
    co2 = filter_linelist(lls[GasSpecies.CO2], 674.0, 736.0)
    got = compute_voigt_lm_cross_sections(g, co2, T, p, ctx=ctx)
    CHECKS.append(("accuracy_vs_fp64", "sigma_co2_layer24_700-710", got,
                   lambda o: o.voigt_lm_xsec(co2, g.nu_min, g.dnu, g.n, T, p), "sigma"))
    return {}


def config0_figure(ctx):
    """Config 0 on the GPU (20k lines, 645-800 cm^-1, 621 channels): spectra/s,
    checked against the oracle's full config-0 spectrum of the CPU leg."""
    from irsounderlbl import synthetic as syn
    from irsounderlbl.api import forward_model
    from irsounderlbl.types import GasSpecies, Sounder
    p0, ll0 = syn.base_profile(False), {GasSpecies.CO2: syn.linelist_config0()}
    s0 = Sounder.make(645.0, 800.0, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    forward_model(p0, ll0, sounder=s0, ctx=ctx, apply_continuum=False)
    t0 = time.perf_counter()
    n = 20
    for _ in range(n):
        _, _, bt = forward_model(p0, ll0, sounder=s0, ctx=ctx, apply_continuum=False)
    dt = (time.perf_counter() - t0) / n
    check_bt("config0", "bt_err_vs_fp64_oracle_K", bt,
             lambda o: ORACLE_CACHE["config0_bt"] if "config0_bt" in ORACLE_CACHE else o.forward(p0, ll0, s0)[1])
    return {"workload": "config 0: one 50-level profile, 20k CO2-like lines, 645-800 cm^-1, 621 channels",
            "spectra_per_s": round(1.0 / dt, 2), "ms_per_spectrum": round(dt * 1e3, 3)}


def run_checks(out):
    """The CPU leg's oracle checks of the GPU results registered in CHECKS."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refcpu
    for section, key, got, oracle, kind in CHECKS:
        ref = oracle(refcpu)
        if kind == "sigma":
            m = ref > 1e-6 * ref.max()
            val = {"max_rel": float(np.max(np.abs(got[m] - ref[m]) / ref[m]))}
        else:
            d = np.abs(np.asarray(got) - np.asarray(ref))
            val = {"max_abs_K": float(d.max()), "rms_K": float(np.sqrt(np.mean(d * d))),
                   "bt_range_K": [float(np.min(ref)), float(np.max(ref))]}
        sec = out.setdefault(section, {})
        if isinstance(sec.get(key), dict):
            sec[key].update(val)
        else:
            sec[key] = val


def batched(ctx, lls, snd, n_prof):
    """One scan line of n_prof perturbed profiles in one call (autotuned
    schedule, graph replay); warm-up on the first group's shape."""
    from irsounderlbl import synthetic as syn
    from irsounderlbl.api import forward_model
    from irsounderlbl.types import ViewingGeometry
    pp = syn.perturbed_profiles(n_prof, seed=2002)
    profs = [x[0] for x in pp]
    geoms = [ViewingGeometry(float(np.degrees(np.arccos(x[2]))), 0.0, x[2]) for x in pp]
    tsf = [x[1] for x in pp]
    ctx.set_schedule(4, use_graph=True, autotune=True)
    log(f"bench: batched warm-up (autotune + graph capture)")
    forward_model(profs, lls, sounder=snd, geom=geoms, T_sfc=tsf, line_mixing="vpy", ctx=ctx, apply_continuum=False)
    log(f"bench: batched {n_prof} profiles")
    t0 = time.perf_counter()
    _, R, BT = forward_model(profs, lls, sounder=snd, geom=geoms, T_sfc=tsf, line_mixing="vpy", ctx=ctx, apply_continuum=False)
    dt = time.perf_counter() - t0
    st = ctx.stats()
    ctx.set_schedule(4)
    flops = FLOP_PER_VOIGT_PAIR * st["active_pairs"] + FLOP_PER_DISP_PAIR * st["dispersive_pairs"]
    # the FP64 oracle (CPU leg) on two of the profiles, on two channel windows
    # (each with the lines that reach its internal grid)
    from irsounderlbl.types import Sounder, filter_linelist, sounder_grid
    ch_all = sounder_grid(snd).nu
    got, calls = [], []
    for q in (0, n_prof // 2):
        for lo, hi in ((1205.0, 1208.0), (2410.0, 2413.0)):
            sw = Sounder.make(lo, hi, snd.dnu, opd_max=snd.opd_max, fwhm_gauss=snd.fwhm_gauss)
            sub = {k: filter_linelist(v, lo - 42.0, hi + 42.0) for k, v in lls.items()}
            c0 = int(round((lo - snd.nu_min) / snd.dnu))
            assert abs(ch_all[c0] - lo) < 1e-9
            n_w = int(round((hi - lo) / snd.dnu)) + 1
            got.append(BT[q, c0:c0 + n_w])
            calls.append((profs[q], sub, sw, pp[q][2], tsf[q]))
    check_bt("batched", "bt_err_vs_fp64_oracle_K", np.concatenate(got),
             lambda o: np.concatenate([o.forward(pr, sub, sw, line_mixing=1, mu=mu, T_sfc=ts)[1]
                                       for pr, sub, sw, mu, ts in calls]))
    return {"workload": f"{n_prof} perturbed profiles (one scan line), config-1 line lists + VP_Y, "
                        "DPTMIN live, one forward_model call",
            "value": round(n_prof / dt, 4), "unit": "spectra/s", "seconds": round(dt, 3),
            "ms_device": round(st["ms_device"], 1), "graph_launches": st["graph_launches"],
            "groups": st["linesum_launches"], "schedule": {"tile_points": st["tile_points"],
                                                         "persistent": st["persistent"]},
            "nominal_tflops": round(flops / (st["ms_device"] * 1e-3) / 1e12, 2),
            "bt_finite": bool(np.all(np.isfinite(BT))),
            "bt_err_vs_fp64_oracle_K": {"profiles": [0, n_prof // 2],
                                        "channels": "1205-1208 and 2410-2413 cm^-1 (13 + 13 channels per profile)"}}


def _bands(nu, v, fn):
    out = {}
    for lo, hi in ((645, 760), (760, 1000), (1000, 1100), (1100, 1300), (1300, 1600), (1600, 2000), (2000, 2250),
                   (2250, 2400), (2400, 2760)):
        m = (nu >= lo) & (nu < hi)
        out[f"{lo}-{hi}"] = round(float(fn(v[m])), 4)
    return out


def emulator_figure(ctx, lls, snd, n_train, n_cal, n_eval, steps, lbl_value, sigma_lnq=0.0):
    """Train the emulator on GPU LBL targets of n_train scenes, calibrate it on
    n_cal held-out scenes (per-channel hold-out BT error vs GPU LBL, spot blocks,
    spot threshold for a 0.05 K budget), then run n_eval new scenes through
    emulator + spot check + full-LBL fallback; the BT error after fallback is
    measured against GPU LBL on every scene the emulator answered (fallen-back
    scenes are LBL), and against the FP64 oracle on two of them (CPU leg)."""
    from irsounderlbl import synthetic as syn
    from irsounderlbl.api import analytic_jacobian, forward_model
    from irsounderlbl.emulator import ChannelODEmulator, forward_emulated
    from irsounderlbl.types import GasSpecies, Sounder, StateVectorSpec, ViewingGeometry, filter_linelist, sounder_grid
    kw = dict(sigma_T=4.0, sigma_h2o=0.6, sigma_o3=0.2, co2_range=(390e-6, 430e-6))
    def geoms(sc):
        return [ViewingGeometry(float(np.degrees(np.arccos(min(p[2], 1.0)))), 0.0, p[2]) for p in sc]
    train = syn.perturbed_profiles(n_train, seed=3003, sigma_lnq=sigma_lnq, **kw)
    log(f"bench: emulator training ({n_train} scenes of GPU LBL targets, per-layer PCA, regression, MLP)")
    t0 = time.perf_counter()
    emu, rep = ChannelODEmulator.train(train, lls, snd, n_pc=64, n_basis=min(2048, n_train), steps=steps,
                                       batch=4096, lr=1e-3, n_hidden=1024, ctx=ctx,
                                       progress=lambda m: log("bench: " + m))
    t_train = time.perf_counter() - t0
    del train
    nu = sounder_grid(snd).nu
    cal = syn.perturbed_profiles(n_cal, seed=5005, **kw)
    log(f"bench: emulator calibration ({n_cal} held-out scenes vs GPU LBL, spot blocks, correction, threshold)")
    t0 = time.perf_counter()
    margin = 0.2
    c = emu.calibrate([p[0] for p in cal], lls, sounder=snd, geom=geoms(cal), T_sfc=[p[1] for p in cal],
                      budget=0.05, margin=margin, n_blocks=6, block=64, ctx=ctx, correct=True)
    t_cal = time.perf_counter() - t0
    ev = syn.perturbed_profiles(n_eval, seed=4004, **kw)
    profs, gs, tsf = [p[0] for p in ev], geoms(ev), [p[1] for p in ev]
    log(f"bench: emulator evaluation ({n_eval} scenes: emulator + spot check + correction + fallback)")
    forward_emulated(profs[:8], lls, emu, sounder=snd, geom=gs[:8], T_sfc=tsf[:8], ctx=ctx)
    t0 = time.perf_counter()
    _, bt_emu_only = emu.forward(profs, geom=gs, T_sfc=tsf)
    t_emu = time.perf_counter() - t0
    t0 = time.perf_counter()
    res = forward_emulated(profs, lls, emu, sounder=snd, geom=gs, T_sfc=tsf, ctx=ctx)
    t_full = time.perf_counter() - t0
    kept = np.flatnonzero(~res.fell_back)
    log(f"bench: emulator reference LBL on all {n_eval} scenes")
    bt_ref = np.zeros_like(res.BT)
    for q0 in range(0, n_eval, 64):
        _, _, bt_ref[q0:q0 + 64] = forward_model(profs[q0:q0 + 64], lls, sounder=snd, geom=gs[q0:q0 + 64],
                                                 T_sfc=tsf[q0:q0 + 64], line_mixing="vpy", ctx=ctx,
                                                 apply_continuum=False)
    d_all = np.abs(res.BT - bt_ref)
    d_raw = np.abs(bt_emu_only - bt_ref)
    d_kept = d_all[kept]
    d_all_max = float(d_all.max())
    rms_all = float(np.sqrt(np.mean(d_all ** 2)))
    # FP64 oracle (CPU leg) on two answered scenes, two channel windows
    if len(kept):
        got, calls = [], []
        for q in (kept[0], kept[len(kept) // 2]):
            for lo, hi in ((1205.0, 1208.0), (2410.0, 2413.0)):
                sw = Sounder.make(lo, hi, snd.dnu, opd_max=snd.opd_max, fwhm_gauss=snd.fwhm_gauss)
                sub = {k: filter_linelist(v, lo - 42.0, hi + 42.0) for k, v in lls.items()}
                c0 = int(round((lo - snd.nu_min) / snd.dnu))
                n_w = int(round((hi - lo) / snd.dnu)) + 1
                got.append(res.BT[q, c0:c0 + n_w])
                calls.append((profs[q], sub, sw, ev[q][2], tsf[q]))
        check_bt("emulator", "bt_err_vs_fp64_oracle_K", np.concatenate(got),
                 lambda o: np.concatenate([o.forward(pr, sub, sw, line_mixing=1, mu=mu, T_sfc=ts)[1]
                                           for pr, sub, sw, mu, ts in calls]))
    log("bench: emulator Jacobians")
    t0 = time.perf_counter()
    jac = emu.jacobian(profs[:4], geom=gs[:4], T_sfc=tsf[:4])
    t_jac = (time.perf_counter() - t0) / 4
    ref = analytic_jacobian(profs[0], lls, StateVectorSpec(50, [GasSpecies.H2O], include_emissivity=False),
                            sounder=snd, geom=gs[0], T_sfc=tsf[0], line_mixing="vpy", vmr_coupling=False, ctx=ctx,
                            apply_continuum=False)
    rel_T = float(np.linalg.norm(jac[0].K_T - ref.K[:, :50]) / np.linalg.norm(ref.K[:, :50]))
    rel_q = float(np.linalg.norm(jac[0].K_lnq_H2O - ref.K[:, 50:100]) / np.linalg.norm(ref.K[:, 50:100]))
    sc_cal = c.scene_max
    out = {"workload": f"train {n_train} / calibrate {n_cal} / evaluate {n_eval} perturbed scenes (T GP 4 K, "
                       f"H2O x lognormal 0.6, O3 x lognormal 0.2, CO2 390-430 ppm, scan 0-48.3 deg; training "
                       f"scenes also ln q_H2O GP {sigma_lnq} per level), full IASI "
                       f"8461 channels, windows variant of the line lists; targets (t, g, ILS moment) x 49 layers, "
                       f"64 PCs per layer, per-layer quadratic + context regression on 13 layer features "
                       f"(incl. Curtis-Godson means of the path above), MLP 13 -> 3x1024 GELU -> 64 "
                       f"(BF16 MFMA) on its residual; spot check 6 x 64 channels, spot-error correction of "
                       f"every channel (ridge map fitted on the calibration scenes)",
           "training": {"seconds": round(t_train, 1), "lbl_targets_seconds": round(rep.seconds_lbl, 1),
                        "pca_seconds": round(rep.seconds_pca, 1), "regression_seconds": round(rep.seconds_fit, 1),
                        "mlp_seconds": round(rep.seconds_train, 1), "adam_steps": steps,
                        "basis_scenes": rep.n_basis_profiles, "final_loss": float(np.mean(rep.losses[-50:])),
                        "coef_rms_err_regression_over_coef_rms": round(rep.coef_rms_err_regression / rep.coef_rms, 4)},
           "holdout_per_channel_bt_err_K": {
               "scenes": n_cal, "what": "emulator alone (before the spot check) vs GPU LBL on the calibration scenes",
               "rms": round(float(np.sqrt(np.mean(c.ch_rms ** 2))), 4),
               "channel_rms_percentiles_10_50_90_99": [round(float(x), 4) for x in np.percentile(c.ch_rms, [10, 50, 90, 99])],
               "channel_max_percentiles_10_50_90_99": [round(float(x), 4) for x in np.percentile(c.ch_max, [10, 50, 90, 99])],
               "rms_by_band": _bands(nu, c.ch_rms, lambda v: np.sqrt(np.mean(v ** 2))),
               "scene_max_percentiles_10_50_90": [round(float(x), 4) for x in np.percentile(sc_cal, [10, 50, 90])],
               "scenes_within_0.05K": round(float(np.mean(sc_cal <= 0.05)), 3),
               "after_correction_scene_max_percentiles_10_50_90": [
                   round(float(x), 4) for x in np.percentile(c.scene_max_corrected, [10, 50, 90])],
               "after_correction_scenes_within_0.04K": round(float(np.mean(c.scene_max_corrected <= 0.04)), 3)},
           "calibration": {"seconds": round(t_cal, 1),
                           "spot_blocks_cm-1": [[round(float(nu[c0]), 2), round(float(nu[c0 + n - 1]), 2)]
                                                for c0, n in c.spot_blocks],
                           "kept_when": "largest spot error <= t_raw and leave-one-block-out residual <= t_lbo",
                           "t_raw_K": None if c.w_raw == 0 else round(1.0 / c.w_raw, 4),
                           "t_lbo_K": None if c.w_lbo == 0 else round(1.0 / c.w_lbo, 4),
                           "budget_K": c.budget, "margin": margin,
                           "calibration_fallback_rate": round(c.fallback_rate, 4)},
           "emulator_only_spectra_per_s": round(n_eval / t_emu, 1),
           "emulator_plus_fallback_spectra_per_s": round(n_eval / t_full, 3),
           "plain_lbl_spectra_per_s": lbl_value,
           "speedup_vs_plain_lbl": round(n_eval / t_full / lbl_value, 3),
           "fallback_rate": round(float(np.mean(res.fell_back)), 4),
           "seconds": {"emulator_only": round(t_emu, 3), "emulator_spot_check_fallback": round(t_full, 3)},
           "bt_err_after_fallback_K": {"max": round(d_all_max, 5), "rms": round(rms_all, 5),
                                       "max_on_answered_scenes": round(float(d_kept.max()), 5) if len(kept) else 0.0,
                                       "scenes_answered_by_emulator": int(len(kept)),
                                       "scenes_over_0.05K": int(np.sum(d_all.max(axis=1) > 0.05)),
                                       "measured_against": f"GPU LBL on all {n_eval} evaluation scenes"},
           "bt_err_emulator_alone_K": {"max": round(float(d_raw.max()), 4), "rms": round(float(np.sqrt(np.mean(d_raw ** 2))), 5),
                                       "scenes_within_0.05K": round(float(np.mean(d_raw.max(axis=1) <= 0.05)), 3)},
           "jacobian_seconds_per_profile": round(t_jac, 3), "jacobian_shape": list(jac[0].K_T.shape),
           "jacobian_vs_lbl_relative_frobenius": {"T": round(rel_T, 4), "lnq_H2O": round(rel_q, 4)}}
    emu.close()
    emu.mlp.close()
    return out


def par_figure(ctx, lls):
    """HITRAN .par ingest (load_hitran_par on the device): the headline line lists
    (300 000 lines) as one 160-character-record .par buffer in host memory, parsed,
    compacted and sorted by wavenumber, host to host.  The CPU leg checks the
    first 20 000 records bit for bit against the oracle reader and times it."""
    from irsounderlbl.hitran import format_hitran_par, parse_hitran_par_bytes
    recs = []
    for ll in lls.values():
        recs.extend(format_hitran_par(ll, i) for i in range(len(ll)))
    data = ("\n".join(recs) + "\n").encode()
    parse_hitran_par_bytes(data[:200000], ctx=ctx)
    t0 = time.perf_counter()
    n_rep = 5
    for _ in range(n_rep):
        got, skipped = parse_hitran_par_bytes(data, ctx=ctx)
    dt = (time.perf_counter() - t0) / n_rep
    sample = ("\n".join(recs[:20000]) + "\n").encode()
    got_s, _ = parse_hitran_par_bytes(sample, ctx=ctx)
    PAR_CHECK.append((sample, got_s))
    return {"workload": f"{len(recs)} records, {len(data) / 1e6:.1f} MB of .par text in host memory -> sorted SoA",
            "seconds": round(dt, 4), "records_per_s": round(len(got) / dt, 1),
            "MB_per_s": round(len(data) / dt / 1e6, 1), "lines": len(got), "skipped": skipped,
            "note": "host-to-host time: PCIe copy of the text, newline scan, per-record parse with exact "
                    "decimal rounding, compaction, sortedness check (the lists are in order), copy back"}


PAR_CHECK = []


def par_cpu_leg(out):
    """Oracle reader on the 20 000-record sample: bit-exact comparison and its rate."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import par_ref
    for sample, got in PAR_CHECK:
        t0 = time.perf_counter()
        ref, _ = par_ref.load_par_text(sample.decode())
        dt = time.perf_counter() - t0
        same = all(np.array_equal(getattr(got, b).view(np.uint8), ref[a].view(np.uint8)) for a, b in (
            ("mol", "mol_id"), ("iso", "iso_id"), ("nu", "wavenumber"), ("S", "intensity"), ("A", "a_coeff"),
            ("gamma_air", "air_broad"), ("gamma_self", "self_broad"), ("elower", "lower_energy"),
            ("n_air", "temp_depend"), ("delta_air", "pressure_shift")))
        sec = out.setdefault("par_ingest", {})
        sec["oracle_bit_exact_first_20000_records"] = bool(same)
        sec["cpu_oracle_records_per_s"] = round(len(ref["nu"]) / dt, 1)


def lbl_jacobian_figure(ctx, lls_w, snd):
    """Seconds per full-IASI analytic Jacobian (temperature, H2O log-VMR,
    surface temperature and emissivity columns; VP_Y on) of one perturbed
    profile, second call timed."""
    from irsounderlbl import synthetic as syn
    from irsounderlbl.api import analytic_jacobian
    from irsounderlbl.types import GasSpecies, StateVectorSpec, ViewingGeometry
    prof, T_sfc, mu = syn.perturbed_profiles(1, seed=5005)[0]
    spec = StateVectorSpec(prof.n_levels, [GasSpecies.H2O])
    kw = dict(sounder=snd, geom=ViewingGeometry(float(np.degrees(np.arccos(mu))), 0.0, mu), T_sfc=T_sfc,
              line_mixing="vpy", ctx=ctx)
    log("bench: LBL analytic Jacobian")
    analytic_jacobian(prof, lls_w, spec, **kw, apply_continuum=False)
    t0 = time.perf_counter()
    jac = analytic_jacobian(prof, lls_w, spec, **kw, apply_continuum=False)
    dt = time.perf_counter() - t0
    st = ctx.stats()
    return {"workload": "full IASI 8461 channels, windows variant of the line lists, 1 perturbed profile, "
                        "columns T(50) + H2O log-VMR(50) + T_sfc + emissivity",
            "seconds_per_profile": round(dt, 3), "K_shape": list(jac.K.shape),
            "K_finite": bool(np.all(np.isfinite(jac.K))),
            "kernels_ms": {"prep": round(st["ms_prep"], 2), "gradient_linesum": round(st["ms_linesum"], 2),
                           "rte_jacobian": round(st["ms_rte"], 2), "ils": round(st["ms_ils"], 2)}}


def exact_sum(ctx, prof, lls, snd):
    """One step with the coarse far-wing sum switched off (every far-wing pair
    on the internal grid; EDGE lines and the mid-core series as in the default
    path) for comparison with the default step."""
    from irsounderlbl.api import forward_model
    sched = ctx.schedule()
    ctx.set_schedule(sched["tile_points"], far_wing_cm=0.0)
    forward_model(prof, lls, sounder=snd, line_mixing="vpy", ctx=ctx, apply_continuum=False)
    t0 = time.perf_counter()
    _, _, bt0 = forward_model(prof, lls, sounder=snd, line_mixing="vpy", ctx=ctx, apply_continuum=False)
    dt = time.perf_counter() - t0
    st = ctx.stats()
    ctx.set_schedule(sched["tile_points"], far_wing_cm=sched["far_wing_cm"])
    _, _, bt1 = forward_model(prof, lls, sounder=snd, line_mixing="vpy", ctx=ctx, apply_continuum=False)
    return {"spectra_per_s": round(1.0 / dt, 4), "linesum_ms": round(st["ms_linesum"], 3),
            "max_abs_dBT_vs_default_K": float(np.max(np.abs(bt1 - bt0)))}


def fp64_figure(ctx, prof, lls, snd, bt32, value32):
    """One full-range spectrum through the every-pair FP64 path (the reference
    kernel's arithmetic on the device); second call timed."""
    from irsounderlbl.api import forward_model
    sched = ctx.schedule()
    ctx.set_schedule(sched["tile_points"], fp64_every_pair=True)
    try:
        log("bench: FP64 every-pair spectrum")
        forward_model(prof, lls, sounder=snd, line_mixing="vpy", ctx=ctx, apply_continuum=False)
        t0 = time.perf_counter()
        _, _, bt64 = forward_model(prof, lls, sounder=snd, line_mixing="vpy", ctx=ctx, apply_continuum=False)
        dt = time.perf_counter() - t0
        st = ctx.stats()
    finally:
        ctx.set_schedule(sched["tile_points"], far_wing_cm=sched["far_wing_cm"])
    d = np.abs(bt32 - bt64)
    # the same comparison where the spectrum is partly transparent (the headline
    # lists saturate every channel at 217 K): the windows variant of the lists
    from irsounderlbl import synthetic as syn
    lls_w = syn.linelists_config1(windows=True)
    ctx.set_schedule(sched["tile_points"], fp64_every_pair=True)
    try:
        _, _, bt64w = forward_model(prof, lls_w, sounder=snd, line_mixing="vpy", ctx=ctx, apply_continuum=False)
    finally:
        ctx.set_schedule(sched["tile_points"], far_wing_cm=sched["far_wing_cm"])
    _, _, bt32w = forward_model(prof, lls_w, sounder=snd, line_mixing="vpy", ctx=ctx, apply_continuum=False)
    dw = np.abs(bt32w - bt64w)
    return {"workload": "the headline spectrum, every pair in FP64 (FP64 line parameters, Faddeeva, sums, optical "
                        "depths and RTE)",
            "windows_variant": {"fp32_engine_max_abs_dBT_K": float(dw.max()),
                                "fp32_engine_rms_dBT_K": float(np.sqrt(np.mean(dw * dw))),
                                "bt_range_K": [float(bt64w.min()), float(bt64w.max())]},
            "spectra_per_s": round(1.0 / dt, 4), "linesum_ms": round(st["ms_linesum"], 1),
            "fp32_engine_max_abs_dBT_K": float(d.max()), "fp32_engine_rms_dBT_K": float(np.sqrt(np.mean(d * d))),
            "fp32_engine_speedup": round(value32 * dt, 1)}


def pmc_from_profiles(name):
    """Per-launch counter figures from a committed rocprofv3 summary under profiles/."""
    p = os.path.join(ROOT, "profiles", name)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return {}


def roofline_figure(stats, ls_avg):
    """Line-sum roofline: FP32 FLOP per launch from the committed counters of the
    same launch (SQ_INSTS_VALU_FLOPS_FP32 counts per wave instruction; x 64
    lanes) over the launch time measured here with HIP events on the context
    stream."""
    pmc = pmc_from_profiles("linesum_full_iasi_pmc.json")
    kernel = ("linesum_persistent_kernel" if stats["persistent"] else "linesum_kernel") + f"<{stats['tile_points']}>"
    nominal = FLOP_PER_VOIGT_PAIR * stats["active_pairs"] + FLOP_PER_DISP_PAIR * stats["dispersive_pairs"]
    t = ls_avg * 1e-3
    flop = 64.0 * pmc["valu_flops_fp32_per_launch"] if pmc.get("valu_flops_fp32_per_launch") else None
    achieved = flop / t / 1e12 if flop else None
    roof = {"bound": "valu", "achieved": round(achieved, 3) if achieved else None, "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4) if achieved else None,
            "traffic": pmc.get("hbm_bytes_per_launch"), "kernel": kernel, "flop_per_launch": flop,
            "launch_ms_avg": round(ls_avg, 3),
            "peak_non_packed": PEAK_FP32_NONPACKED_TFLOPS,
            "frac_non_packed": round(achieved / PEAK_FP32_NONPACKED_TFLOPS, 4) if achieved else None,
            "source": "flop_per_launch = 64 x SQ_INSTS_VALU_FLOPS_FP32 of this launch "
                      "(profiles/linesum_full_iasi_pmc.json); traffic = FETCH_SIZE + WRITE_SIZE scaled by a "
                      "device copy of known size",
            "valu_issue": pmc.get("valu_issue"),
            "algorithmic_equivalent": {
                "flop_per_launch": nominal, "tflops": round(nominal / t / 1e12, 2),
                "frac_of_peak": round(nominal / t / 1e12 / PEAK_FP32_TFLOPS, 3),
                "pairs_reference": {"voigt": stats["active_pairs"], "dispersive": stats["dispersive_pairs"]},
                "executed": {"fine_pairs": stats["fine_pairs"], "coarse_far_wing_node_pairs": stats["coarse_pairs"],
                             "fine_pairs_vpy": stats["fine_pairs_lm"],
                             "coarse_node_pairs_vpy": stats["coarse_pairs_lm"],
                             "line_slices_expanded": stats["expanded_line_slices"]},
                "note": "the reference's every-pair FLOP count (14 per Voigt pair, 9 per dispersive pair) over "
                        "the same launch time; above 1 because the kernel evaluates far fewer pairs (coarse far "
                        "wings, per-slice expansions, mid-core series), so it is not a kernel-efficiency figure"}}
    return roof


def hbm_figure(stats, n_hi, n_layers, n_ch, half):
    """RTE and channel-ILS kernels: algorithmic bytes per launch (tau read as
    FP32 + R_hi written; R_hi read + taps + R/BT written) and the committed
    counter bytes, over the launch times measured here, against 8 TB/s."""
    pmc = pmc_from_profiles("rte_ils_full_iasi_pmc.json")
    out = {}
    for key, ms, alg in (("rte", stats["ms_rte"], n_hi * n_layers * 4 + n_hi * 8),
                         ("ils", stats["ms_ils"], n_hi * 8 + (2 * half + 1) * 8 + n_ch * 16)):
        k = pmc.get(key, {})
        t = ms * 1e-3
        d = {"kernel": k.get("kernel"), "launch_ms": round(ms, 4), "algorithmic_bytes": alg,
             "achieved": round(alg / t / 1e9, 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
             "frac": round(alg / t / 1e9 / PEAK_HBM_GBPS, 4), "traffic": k.get("hbm_bytes_per_launch")}
        if k.get("hbm_bytes_per_launch"):
            d["traffic_GBps"] = round(k["hbm_bytes_per_launch"] / t / 1e9, 1)
        if k.get("bound_note"):
            d["note"] = k["bound_note"]
        out[key] = d
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch", type=int, default=120)
    ap.add_argument("--no-exact", action="store_true")
    ap.add_argument("--no-emulator", action="store_true")
    ap.add_argument("--no-jacobian", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-par", action="store_true")
    ap.add_argument("--emu-train", type=int, default=8192)
    ap.add_argument("--emu-cal", type=int, default=1024)
    ap.add_argument("--emu-eval", type=int, default=2048)
    ap.add_argument("--emu-steps", type=int, default=96000)
    ap.add_argument("--emu-sigma-lnq", type=float, default=0.05,
                    help="level-by-level ln q_H2O perturbation (GP sigma) of the emulator's training scenes")
    ap.add_argument("--emulator-only", action="store_true",
                    help="the headline steps and the emulator figure only (the other figures off)")
    args = ap.parse_args()
    if args.emulator_only:
        args.no_cpu = args.no_accuracy = args.no_batch = args.no_exact = args.no_jacobian = args.no_fp64 = True
        args.no_par = True
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this engine shards profiles across GPUs with no "
              "exchange, and every box here has one GPU; run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from irsounderlbl import synthetic as syn
    from irsounderlbl.api import Context, forward_model
    from irsounderlbl.types import GasSpecies, IASIInstrument

    log("bench: building workload (300k lines, 50 levels)")
    lls = syn.linelists_config1()
    lls_w = syn.linelists_config1(windows=True)     # accuracy checks and the emulator (see accuracy_sample)
    prof = syn.base_profile(True)
    snd = IASIInstrument()
    ctx = Context(0)
    for ll in lls.values():
        ctx.upload(ll)

    def step():
        return forward_model(prof, lls, sounder=snd, line_mixing="vpy", ctx=ctx, apply_continuum=False)

    for i in range(args.warmup):
        step()
        log(f"bench: warmup {i + 1}/{args.warmup} done")
    ls_ms = []
    stats = None
    t0 = time.perf_counter()
    for i in range(args.steps):
        nu, R, BT = step()
        st = ctx.stats()
        ls_ms.append(st["ms_linesum"])
        stats = st
    t1 = time.perf_counter()
    log(f"bench: {args.steps} steps in {t1 - t0:.3f} s")
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = 1e3 / ms_step
    ls_avg = float(np.mean(ls_ms))
    half = int(np.ceil(16.0 / 0.001))
    roof = roofline_figure(stats, ls_avg)
    out = {"metric": "IASI spectra/s on 1 GPU (LBL vs emulator) at BT err ≤0.005/0.05 K vs FP64",
           "value": round(value, 4), "unit": "spectra/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
           "data": "synthetic (seeded line lists and profile, see irsounderlbl/synthetic.py)",
           "config": {"workload": "full-range IASI, 1 profile, 300k lines (3 species, VP_Y), "
                                  "49 layers, 2147001-point grid, 8461 channels",
                      "path": "plain FP32 LBL", "profiles_per_step": 1},
           "roofline": roof,
           "far_wing_on_fine_grid": None if args.no_exact else exact_sum(ctx, prof, lls, snd),
           "hbm": hbm_figure(stats, stats["n_hi"], stats["n_layers"], len(BT), half),
           "kernels_ms": {"prep": round(stats["ms_prep"], 3), "linesum": round(ls_avg, 3),
                          "rte": round(stats["ms_rte"], 3), "ils": round(stats["ms_ils"], 3)},
           "host_ms_per_step": round(ms_step - stats["ms_device"], 3)}
    if not args.no_fp64:
        out["fp64_every_pair"] = fp64_figure(ctx, prof, lls, snd, BT, value)
    if not args.no_batch:
        out["batched"] = batched(ctx, lls, snd, args.batch)
    if not args.no_jacobian:
        out["lbl_jacobian"] = lbl_jacobian_figure(ctx, lls_w, snd)
    if not args.no_par:
        log("bench: .par ingest")
        out["par_ingest"] = par_figure(ctx, lls)
    if not args.no_emulator:
        out["emulator"] = emulator_figure(ctx, lls_w, snd, args.emu_train, args.emu_cal, args.emu_eval,
                                          args.emu_steps, round(value, 4), args.emu_sigma_lnq)
    if not args.no_accuracy:
        log("bench: accuracy samples and config 0 (GPU side)")
        accuracy_sample(ctx, prof, lls, lls_w)
        out["config0"] = config0_figure(ctx)
    if not args.no_cpu:
        log("bench: cpu baseline (FP64 oracle: config 0 full, config 1 sampled layers, config 2 two profiles)")
        out["cpu_baseline"] = cpu_baseline(prof, lls, snd)
        log("bench: FP64 oracle checks of the GPU results")
        run_checks(out)
        par_cpu_leg(out)
    elif args.emulator_only and CHECKS:
        log("bench: FP64 oracle checks of the emulator results")
        run_checks(out)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
