"""Shared comparison helpers for the parity tests."""
import numpy as np


def rel_err_masked(got, ref, frac=1e-6):
    """Max relative error where ref > frac * max(ref) (the cross-section criterion)."""
    ref = np.asarray(ref); got = np.asarray(got)
    peak = ref.max()
    m = ref > frac * peak
    if not m.any():
        return 0.0, 0
    e = np.abs(got[m] - ref[m]) / ref[m]
    return float(e.max()), int(np.argmax(e))


def four_line_fixture():
    """The 4-line CO2 fixture of the reference test suite (runtests.jl:252-256)."""
    from irsounderlbl.types import HITRANLinelist
    nu = np.array([2348.0, 2350.0, 2351.2, 2352.5])
    S = np.array([1e-19, 5e-20, 8e-20, 2e-19])
    return HITRANLinelist(2, 1, nu, S, np.full(4, 0.07), np.full(4, 0.09), np.full(4, 250.0),
                          np.full(4, 0.75), np.zeros(4))
