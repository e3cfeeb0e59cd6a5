"""CPU checks of the asymptotic w(z) series that the line-sum kernel evaluates
for CORE pairs with |z| >= 8 (mid-core records written by the prep kernel,
k_prep.hip midcore_coeffs; evaluated by k_linesum.hip series_core).

The coefficients are those of w ~ (i/sqrt(pi)) sum_{n<=4} (2n-1)!!/(2^n z^(2n+1))
regrouped in powers of v = 1/(x^2 + y^2); this file restates them from the
kernel comment and checks them against scipy.special.wofz (an independent
implementation), plus the algebraic identities that fix the regrouping."""
import numpy as np
from scipy.special import wofz


def re_coeffs(Y):
    """k2..k9 of Re w = (y/sqrt(pi)) v (1 + sum_k k_k v^(k-1)), Y = y^2."""
    return [1.5, 3.75 - 2 * Y, 13.125 - 15 * Y, 59.0625 + Y * (12 * Y - 105), Y * (210 * Y - 787.5),
            Y * Y * (2835 - 120 * Y), -3780 * Y ** 3, 1680 * Y ** 4]


def im_coeffs(Y):
    """k2..k9 of Im w = (x/sqrt(pi)) v (1 + sum_k k_k v^(k-1))."""
    return [0.5, 0.75 - 2 * Y, 1.875 - 9 * Y, 6.5625 + Y * (12 * Y - 45), Y * (150 * Y - 262.5),
            Y * Y * (1575 - 120 * Y), -2940 * Y ** 3, 1680 * Y ** 4]


def series(x, y, im=False):
    Y = y * y
    v = 1.0 / (x * x + Y)
    k = im_coeffs(Y) if im else re_coeffs(Y)
    q = 0.0
    for c in reversed(k):
        q = q * v + c
    s = (q * v + 1.0) * v / np.sqrt(np.pi)
    return s * (x if im else y)


def test_series_matches_wofz_beyond_z1():
    worst_re = worst_im = 0.0
    for y in [1e-6, 1e-3, 0.05, 0.5, 1.0, 3.0, 7.9, 8.0, 12.0, 40.0, 150.0]:
        xmin = np.sqrt(max(64.0 - y * y, 0.0))
        x = np.concatenate([np.linspace(xmin, 130.0, 3000), -np.linspace(xmin, 130.0, 50)])
        x = x[x * x + y * y >= 64.0]
        w = wofz(x + 1j * y)
        worst_re = max(worst_re, np.max(np.abs(series(x, y) - w.real) / np.abs(w.real)))
        worst_im = max(worst_im, np.max(np.abs(series(x, y, True) - w.imag) / np.abs(w)))
    assert worst_re < 4e-7, worst_re
    assert worst_im < 4e-7, worst_im


def test_series_reduces_to_lorentz_and_leading_corrections():
    # y -> 0: Re w / (y/(sqrt(pi) x^2)) = 1 + 3/(2x^2) + 15/(4x^4) + 105/(8x^6) + ...
    x, y = 40.0, 1e-9
    lead = y / (np.sqrt(np.pi) * x * x)
    assert abs(series(x, y) / lead - (1 + 1.5 / x**2 + 3.75 / x**4 + 13.125 / x**6 + 59.0625 / x**8)) < 1e-15
    # Im w for y -> 0 is Dawson-like: (1/(sqrt(pi) x)) (1 + 1/(2x^2) + 3/(4x^4) + ...)
    assert abs(series(x, y, True) * np.sqrt(np.pi) * x - (1 + 0.5 / x**2 + 0.75 / x**4 + 1.875 / x**6
                                                          + 6.5625 / x**8)) < 1e-15
    # the truncation is a dropped term, not a wrong one: at |z| = 8 the error is
    # of the size of the first omitted series term, 945/32 * 11 / |z|^10
    x = np.linspace(8.0, 9.0, 101)
    err = np.abs(series(x, 1e-3) - wofz(x + 1e-3j).real) / wofz(x + 1e-3j).real
    assert err.max() < 945 / 32 * 11 / 8.0**10 * 1.2


def test_series_fp32_horner_is_accurate():
    # the kernel evaluates the polynomial in FP32 after one reciprocal
    f = np.float32
    for y in [1e-4, 0.7, 5.0, 60.0]:
        x = np.linspace(np.sqrt(max(64 - y * y, 0)) + 1e-3, 122.0, 2000)
        Y = f(y * y)
        v = (f(1) / (x.astype(f) * x.astype(f) + Y)).astype(f)
        q = np.zeros_like(v)
        for c in reversed([f(c) for c in re_coeffs(float(Y))]):
            q = (q * v + c).astype(f)
        s = ((q * v + f(1)) * v).astype(f) * f(y / np.sqrt(np.pi))
        w = wofz(x + 1j * y).real
        assert np.max(np.abs(s - w) / w) < 2e-6


def test_edge_expansion_coefficients():
    """The cubic (in the offset delta from a slice centre) that the line sum
    uses for EDGE lines (k_linesum.hip edge_expand), restated from its comment,
    against direct evaluation of the far wing + pedestal and of the dispersive
    wing at slice distances 4000 and 25000 grid steps."""
    A, B, P, C = 1.3, 0.7, 2.1e-9, 25000.0
    for d in (4000.0, -4000.0, 25000.0, -24000.0):
        for g in (0.5, 30.0, 90.0):
            D = d * d + g * g
            a = [A / D - P * (2 - d * d / C**2), -2 * d * A / D**2 + 2 * P * d / C**2,
                 A * (3 * d * d - g * g) / D**3 + P / C**2, -4 * d * A * (d * d - g * g) / D**4]
            b = [B * d / D, B * (g * g - d * d) / D**2, B * d * (d * d - 3 * g * g) / D**3,
                 -B * (d**4 - 6 * d * d * g * g + g**4) / D**4]
            x = np.arange(64) - 31.5
            f = A / ((d + x) ** 2 + g * g) - P * (2 - (d + x) ** 2 / C**2)
            h = B * (d + x) / ((d + x) ** 2 + g * g)
            fa = a[0] + x * (a[1] + x * (a[2] + x * a[3]))
            hb = b[0] + x * (b[1] + x * (b[2] + x * b[3]))
            L = A / ((d + x) ** 2 + g * g)
            assert np.max(np.abs(fa - f) / L) < 5 * (32 / abs(d)) ** 4
            assert np.max(np.abs(hb - h) / np.abs(h)) < 5 * (32 / abs(d)) ** 4
            # the cubic term is needed at d = 4000: dropping it costs ~(32/d)^3
            if abs(d) < 5000:
                f2 = a[0] + x * (a[1] + x * a[2])
                assert np.max(np.abs(f2 - f) / L) > 1e-7
