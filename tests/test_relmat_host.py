"""Host side of the relaxation-matrix line mixing in libirs (no GPU): the per-layer
terms irs_relmat_terms builds -- Y(T) from the sum-rule-corrected W, the S-file
line parameters, the VP_W eigenmodes from the library's own complex
eigen-solver -- against the oracle (oracle/relmat_ref.py, numpy's LAPACK eig),
and the sums of those terms on a grid against the oracle's perturbation."""
import numpy as np
import pytest

import relmat_ref as R
from irsounderlbl import synthetic as syn
from irsounderlbl.api import relmat_terms
from irsounderlbl.relmat import VPWLineMixing, VPYLineMixing, band_modes_for


@pytest.fixture(scope="module")
def rm():
    return syn.relmat_synthetic()


def _eval_terms(nu, terms, cutoff=25.0, x_far=R.X_FAR):
    out = np.zeros(len(nu))
    for c, f, y, a, b in terms:
        m = np.abs(nu - c) <= cutoff
        x = (nu[m] - c) * f
        far = np.abs(x) > x_far
        den = R.INV_SQRT_PI / (x * x + y * y)
        wz = R._w(np.where(far, 0.0, x), y)
        out[m] += a * np.where(far, y * den, wz.real) + b * np.where(far, x * den, wz.imag)
    return out


@pytest.mark.parametrize("T,p", [(215.0, 0.02), (288.0, 0.9)])
def test_vpy_terms_match_oracle(rm, T, p):
    modes = band_modes_for(VPYLineMixing(rm), None)
    assert list(modes) == [1, 1, 1, 1, 0]                 # (2, 3) has no W0/B0 table
    got = relmat_terms(VPYLineMixing(rm), T, p)
    ref = []
    for b, m in zip(rm.bands, modes):
        if m:
            ref += R.dispersive_lines(b, R.calc_Y(b, R._wt_of(rm, b), T), T, p)
    ref = np.array(sorted(ref, key=lambda t: t[0]))
    assert got.shape == (len(ref), 5)
    np.testing.assert_allclose(got[:, 0], ref[:, 0], rtol=1e-15, atol=0)
    np.testing.assert_allclose(got[:, 1:3], ref[:, 1:3], rtol=1e-13)
    assert np.all(got[:, 3] == 0.0)
    np.testing.assert_allclose(got[:, 4], ref[:, 3], rtol=1e-10, atol=1e-12 * np.max(np.abs(ref[:, 3])))


def test_isotopologue_and_strength_filters(rm):
    assert list(band_modes_for(VPYLineMixing(rm), {1})) == [1, 0, 0, 1, 0]
    thr = 0.5 * R.band_strength(rm.bands[0])
    assert list(band_modes_for(VPYLineMixing(rm, thr), None)) == [1, 0, 0, 0, 0]
    assert list(band_modes_for(VPWLineMixing(rm, {"S626_01", "S628_12"}), None)) == [2, 1, 2, 1, 0]


@pytest.mark.parametrize("T,p", [(220.0, 0.05), (290.0, 0.8)])
def test_vpw_eigenmodes_match_oracle(rm, T, p):
    """The library's eigen-solver gives the oracle's poles and amplitudes (numpy eig),
    and the term sum equals the oracle's VP_W + VP_Y perturbation on a grid."""
    lm = VPWLineMixing(rm, {"S626_01", "S628_12"})
    got = relmat_terms(lm, T, p, keep_threshold=0.0)
    for bi in (0, 2):
        b = rm.bands[bi]
        lam, A, f = R.band_modes(b, R._wt_of(rm, b), T, p, keep_threshold=0.0)
        sel = got[(np.abs(got[:, 1] - f) < 1e-12 * f) & (got[:, 4] != 0.0)]
        # match every oracle pole to the nearest library pole
        for k in range(len(lam)):
            d = np.abs(sel[:, 0] - lam[k].real) + np.abs(sel[:, 2] / f - lam[k].imag)
            j = int(np.argmin(d))
            assert d[j] < 1e-9
            amp = (sel[j, 3] + 1j * sel[j, 4]) / (f * R.INV_SQRT_PI)
            assert abs(amp - A[k]) <= 1e-8 * np.max(np.abs(A))
    nu = 640.0 + np.arange(0, 110001) * 0.001
    d_ref = R.vpw_perturbation(nu, rm, T, p, {"S626_01", "S628_12"}, keep_threshold=0.0)
    d_got = _eval_terms(nu, got)
    # the full and no-coupling sums nearly cancel: amplitudes equal to ~1e-9 of the
    # largest leave ~1e-8 of the perturbation's peak
    assert np.max(np.abs(d_got - d_ref)) < 1e-6 * np.max(np.abs(d_ref))


def test_keep_threshold_drops_weak_poles(rm):
    lm = VPWLineMixing(rm, {"S626_01"})
    all_ = relmat_terms(lm, 250.0, 0.3, keep_threshold=0.0)
    kept = relmat_terms(lm, 250.0, 0.3, keep_threshold=0.05)
    b = rm.bands[0]
    lam, A, f = R.band_modes(b, R._wt_of(rm, b), 250.0, 0.3, keep_threshold=0.05)
    assert len(kept) < len(all_)
    assert len(all_) - len(kept) == b.n - len(lam)
