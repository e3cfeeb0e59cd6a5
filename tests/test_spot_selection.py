"""Spot-block selection and threshold of the emulator's fallback (host logic, no GPU),
on synthetic error matrices where the answer is known."""
import numpy as np

from irsounderlbl.emulator import select_spot_blocks, spot_threshold


def test_threshold_rule():
    se = np.array([0.01, 0.02, 0.03, 0.04])
    fe = np.array([0.02, 0.03, 0.05, 0.01])
    assert spot_threshold(se, fe, 0.04) == 0.02          # the third scene is bad
    assert spot_threshold(se, np.full(4, 0.01), 0.04) == 0.04
    assert spot_threshold(se, np.full(4, 0.09), 0.04) == 0.0


def test_greedy_blocks_find_the_channels_that_flag_bad_scenes():
    """Bad scenes carry their error at channels 400-415 and 800-815; good scenes have
    small errors everywhere, larger (but harmless) ones at 100-115.  The hold-out RMS
    points at 100-115, the pass count at the two informative blocks."""
    rng = np.random.default_rng(0)
    n, nch = 200, 1000
    d = rng.uniform(0.0, 0.01, (n, nch))
    d[:, 100:116] += 0.03                          # largest RMS, not informative
    bad = rng.random(n) < 0.3
    half = bad.copy()
    half[::2] = False
    d[np.flatnonzero(half), 400:416] += 0.08
    d[np.flatnonzero(bad & ~half), 800:816] += 0.08
    rms = np.sqrt((d ** 2).mean(axis=0))
    blocks = select_spot_blocks(d, rms, 2, 16, 0.04)
    assert blocks == [(400, 16), (800, 16)]
    spot = np.r_[400:416, 800:816]
    se, fe = d[:, spot].max(axis=1), d.max(axis=1)
    t = spot_threshold(se, fe, 0.04)
    assert np.all(~bad[se <= t]) and np.sum(se <= t) == np.sum(~bad)


# ---- spot-check correction ------------------------------------------------------
from irsounderlbl.api import brightness_temperature  # noqa: E402
from irsounderlbl.emulator import (correct_spectra, fit_spot_correction, select_correction_blocks,  # noqa: E402
                                   spot_statistics)


def _low_rank_errors(n, nch=900, seed=1):
    """Errors that are three fixed spectral patterns, each confined to its own third
    of the channels, with scene-dependent amplitudes (plus small noise)."""
    rng = np.random.default_rng(seed)
    V = np.zeros((3, nch))
    for j in range(3):
        V[j, 300 * j:300 * (j + 1)] = np.sin(np.linspace(0, 3 + j, 300)) + 1.5
    A = rng.normal(0.0, 0.03, (n, 3))
    return A @ V + rng.normal(0.0, 1e-4, (n, nch))


def test_correction_blocks_cover_every_pattern_and_remove_the_error():
    """Each pattern lives in one third of the spectrum, so a correction must see one
    block in every third; with them the ridge map predicts every channel's error from
    the spot errors of new scenes to the noise level."""
    E = _low_rank_errors(120)
    blocks = select_correction_blocks(E, 3, 16, lam=1e-6)
    spot = np.concatenate([np.arange(c, c + n) for c, n in blocks])
    assert set(spot // 300) == {0, 1, 2}
    W, lbo = fit_spot_correction(E, spot, [16, 16, 16], lam=1e-6)
    assert W.shape == (48, 900) and lbo.shape == (3 * 32 * 16,)
    Et = _low_rank_errors(50, seed=2)
    r = Et - Et[:, spot] @ W.astype(np.float64)
    assert np.abs(Et).max() > 0.1
    assert np.abs(r).max() < 3e-3


def test_leave_one_block_out_statistic_flags_an_unexplained_block():
    """Errors that the calibration's patterns explain give a small leave-one-block-out
    residual; an error confined to one block that no pattern carries gives a large one,
    while the largest spot error can stay small."""
    E = np.concatenate([_low_rank_errors(200, seed=3)[:, :600]] * 1, axis=1)     # two patterns over 600 channels
    # blocks inside both supports of pattern 0 so that blocks predict each other
    blocks = [(20, 16), (150, 16), (350, 16), (480, 16)]
    spot = np.concatenate([np.arange(c, c + n) for c, n in blocks])
    W, lbo = fit_spot_correction(E, spot, [16] * 4, lam=1e-6)
    es = _low_rank_errors(20, seed=4)[:, :600][:, spot]
    raw, res = spot_statistics(es, [16] * 4, lbo)
    assert np.all(res < 2e-3) and raw.max() > 0.03
    bad = es.copy()
    bad[0, 16:32] += 0.02                      # block 1 only
    raw_b, res_b = spot_statistics(bad, [16] * 4, lbo)
    assert res_b[0] > 0.015 and raw_b[0] < raw.max() + 0.03


def test_correct_spectra_sets_spot_channels_and_radiances():
    nu = np.linspace(700.0, 710.0, 41)
    bt = np.full((2, 41), 250.0) + np.linspace(0, 1, 41)
    spot = np.arange(10, 14)
    bt_spot = bt[:, spot] - 0.05
    W = np.zeros((4, 41), np.float32)
    W[:, :] = 0.25                                   # every channel moves by the mean spot error
    R, BTc, es = correct_spectra(bt, bt_spot, spot, W, nu)
    np.testing.assert_allclose(es, 0.05, atol=1e-6)
    np.testing.assert_allclose(np.delete(BTc, spot, axis=1), np.delete(bt, spot, axis=1) - 0.05, atol=1e-6)
    assert np.array_equal(BTc[:, spot], bt_spot)
    np.testing.assert_allclose(brightness_temperature(nu, R[0]), BTc[0], rtol=0, atol=1e-9)


def test_joint_thresholds_keep_the_most_scenes_without_a_bad_one():
    from irsounderlbl.emulator import fallback_thresholds
    # scene: (raw, lbo, bad).  The bad scenes are excluded by raw (scene 4) and by lbo (scene 5).
    s_raw = np.array([0.01, 0.02, 0.03, 0.04, 0.05, 0.02])
    s_lbo = np.array([0.01, 0.01, 0.02, 0.01, 0.01, 0.05])
    bad = np.array([False, False, False, False, True, True])
    t1, t2 = fallback_thresholds(s_raw, s_lbo, bad)
    assert (t1, t2) == (0.04, 0.02)
    keep = (s_raw <= t1) & (s_lbo <= t2)
    assert keep.tolist() == [True, True, True, True, False, False]
    # either statistic alone would keep fewer: raw <= 0.01 (scene 5 has raw 0.02), lbo < 0.01 (none)
    # no bad scene: every scene kept, the raw threshold at the largest calibration value
    assert fallback_thresholds(s_raw, s_lbo, np.zeros(6, bool)) == (0.05, np.inf)
    assert fallback_thresholds(s_raw, s_lbo, np.ones(6, bool)) == (-1.0, -1.0)
    # no bad scene among the small raw values: the lbo statistic is unused
    t1, t2 = fallback_thresholds(np.array([0.01, 0.02, 0.09]), np.array([0.5, 0.4, 0.1]),
                                 np.array([False, False, True]))
    assert (t1, t2) == (0.02, np.inf)
