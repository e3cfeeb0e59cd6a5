"""irs_xsec_layer_noalloc, the entry point that allocates no device memory
(its buffers live in a fixed device array of the library), against
irs_xsec_layer through the C ABI: bit-identical sigma and identical work
counters on two calls in a row, the second with a smaller input, and a clear
IRS_EINVAL when the list and grid do not fit the array."""
import ctypes as C

import numpy as np
import pytest

from irsounderlbl import synthetic as syn
from irsounderlbl.types import GasSpecies, filter_linelist, wavenumber_grid

pytestmark = pytest.mark.gpu


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _call(ctx, fn, h, g, T, p, vs, coef, dptmn, lm):
    out = np.full(g.n, np.nan)
    rc = fn(ctx.h, h, g.nu_min, g.dnu, g.n, T, p, vs, coef, dptmn, 25.0, 25.0, 122.0, 1 if lm else 0, _p(out))
    return rc, out, ctx.stats()


def _same(ctx, h, g, T, p, vs=0.0, coef=1.0, dptmn=0.0, lm=False):
    L = ctx.lib
    rc1, a, sa = _call(ctx, L.irs_xsec_layer_noalloc, h, g, T, p, vs, coef, dptmn, lm)
    assert rc1 == 0, L.irs_last_error(ctx.h)
    rc0, b, sb = _call(ctx, L.irs_xsec_layer, h, g, T, p, vs, coef, dptmn, lm)
    assert rc0 == 0, L.irs_last_error(ctx.h)
    assert np.all(np.isfinite(a)) and a.max() > 0.0
    assert np.array_equal(a, b), np.max(np.abs(a - b))
    for k in ("active_lines", "active_pairs", "dispersive_pairs", "fine_pairs", "coarse_pairs",
              "expanded_line_slices"):
        assert sa[k] == sb[k], k
    return a, sa


def test_noalloc_matches_xsec_layer_two_calls(ctx):
    lls = syn.linelists_config1()
    # first call: the 180 000-line CO2 list with VP_Y on the full IASI grid (2 147 001 points)
    g1 = wavenumber_grid(629.0, 2776.0, 0.001)
    h1 = ctx.upload(lls[GasSpecies.CO2])
    a, st = _same(ctx, h1, g1, 250.0, 0.4, lm=True)
    print(f"call 1: {g1.n} points, {st['active_lines']} lines, line sum {st['ms_linesum']:.2f} ms")
    # second call, smaller: an H2O sub-list with self broadening and DPTMIN on 30 cm^-1
    g2 = wavenumber_grid(700.0, 730.0, 0.001)
    sub = filter_linelist(lls[GasSpecies.H2O], 650.0, 780.0)
    h2 = ctx.upload(sub)
    coef = 1e17                                           # DPTMIN keeps about 2/3 of the lines
    b, st = _same(ctx, h2, g2, 280.0, 0.8, vs=2e-3, coef=coef, dptmn=1e-6)
    assert 0 < st["active_lines"] < len(sub)             # DPTMIN rejects some lines
    print(f"call 2: {g2.n} points, {st['active_lines']} of {len(sub)} lines kept")
    # and back to a still smaller window of the first list after the second
    g3 = wavenumber_grid(2300.0, 2310.0, 0.002)
    _same(ctx, h1, g3, 220.0, 0.05, lm=True)


def test_noalloc_rejects_what_does_not_fit(ctx):
    from irsounderlbl.api import IRS_EINVAL
    ll = syn.linelists_config1()[GasSpecies.CO2]         # kept alive: its handle lives as long as it does
    h = ctx.upload(ll)
    n = 20_000_000                                         # 80 MB of FP32 output alone
    out = np.zeros(1)
    rc = ctx.lib.irs_xsec_layer_noalloc(ctx.h, h, 629.0, 1e-4, n, 250.0, 0.4, 0.0, 1.0, 0.0, 25.0, 25.0, 122.0,
                                        0, _p(out))
    assert rc == IRS_EINVAL
    msg = ctx.lib.irs_last_error(ctx.h).decode()
    assert "scratch" in msg and "irs_xsec_layer" in msg
    # the context still works after the rejection
    g = wavenumber_grid(700.0, 701.0, 0.001)
    _same(ctx, h, g, 250.0, 0.4)
