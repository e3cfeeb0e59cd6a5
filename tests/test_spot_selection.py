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
