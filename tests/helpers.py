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


def _fcim(s):
    s = np.asarray(s, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        big = 1.0 - 2.0 * (np.exp(-s) / np.expm1(-s) + 1.0 / s)
    return np.where(s < 0.06, s / 6.0, big)


def channel_od_ref(tau, nu_lo, dnu, ch_nu, kern, floor, od_max=50.0, mu=1.0):
    """FP64 restatement of the per-layer channel optical depth along the path
    (vertical tau [L, n_hi], path factor 1/mu):
    T_k = ILS (x) exp(-sum_{j>=k} tau_j/mu), tau_ch_k = ln(T_{k+1}/T_k) where
    T_{k+1} >= floor, else -ln(ILS (x) exp(-tau_k/mu)); and the CIM channel
    weight g_k = ILS (x) [(T_{k+1} - T_k) f_cim(tau_k/mu)] / T^ch_{k+1} (layer-alone
    ILS (x) [(1 - e^{-s}) f_cim(s)] under an opaque column).
    Returns (tau_ch [L, n_ch], T^ch [L+1, n_ch], g [L, n_ch]); G [L, n_ch] (the
    unnormalised CIM sums) is left in channel_od_ref.last_G."""
    tau = np.asarray(tau, np.float64) / mu
    L, n = tau.shape
    half = (len(kern) - 1) // 2
    cum = np.cumsum(tau[::-1], axis=0)[::-1]          # cum[k] = sum_{j>=k} tau_j
    lev = np.vstack([np.exp(-cum), np.ones((1, n))])  # [L+1, n]
    alone = np.exp(-tau)
    f = _fcim(tau)
    G = (lev[1:] - lev[:-1]) * f
    Ga = (1.0 - alone) * f
    out = np.zeros((L, len(ch_nu)))
    gout = np.zeros((L, len(ch_nu)))
    Gall = np.zeros((L, len(ch_nu)))
    Tch_all = np.zeros((L + 1, len(ch_nu)))
    for c, v in enumerate(ch_nu):
        i = int(round((v - nu_lo) / dnu))
        lo, hi = i - half, i + half + 1
        klo, khi = max(0, -lo), len(kern) - max(0, hi - n)
        w = kern[klo:khi]
        sl = slice(max(lo, 0), min(hi, n))
        Tch = dnu * (lev[:, sl] @ w)
        Al = dnu * (alone[:, sl] @ w)
        Gc = dnu * (G[:, sl] @ w)
        Gac = dnu * (Ga[:, sl] @ w)
        Tch_all[:, c] = Tch
        Gall[:, c] = Gc
        for k in range(L):
            if Tch[k + 1] >= floor and Tch[k] > 0:
                od = np.log(Tch[k + 1] / Tch[k])
            else:
                od = -np.log(Al[k]) if Al[k] > 0 else od_max
            out[k, c] = min(max(od, 0.0), od_max)
            gv = Gc[k] / Tch[k + 1] if Tch[k + 1] >= floor else Gac[k]
            gout[k, c] = min(max(gv, 0.0), 1.0)
    channel_od_ref.last_G = Gall
    # ILS first moments of the level transmittance steps, M_k = ILS (x) [(nu - nu_c)(T_{k+1} - T_k)]
    mk = np.asarray(kern, np.float64) * (np.arange(len(kern)) - half) * dnu
    Mall = np.zeros((L, len(ch_nu)))
    for c, v in enumerate(ch_nu):
        i = int(round((v - nu_lo) / dnu))
        lo, hi = i - half, i + half + 1
        klo, khi = max(0, -lo), len(kern) - max(0, hi - n)
        Ml = dnu * (lev[:, max(lo, 0):min(hi, n)] @ mk[klo:khi])
        Mall[:, c] = Ml[1:] - Ml[:-1]
    channel_od_ref.last_M = Mall
    return out, Tch_all, gout
