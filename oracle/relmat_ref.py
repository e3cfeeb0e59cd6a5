"""FP64 restatement of the relaxation-matrix line mixing of
src/CrossSections/line_mixing.jl.  TEST INFRASTRUCTURE ONLY: imported by tests/ and
never by the product.

Bands are read through their attributes (name, li, lf, isot, molecule, nu, Ji,
branch, gV_air, n_air, shift, E_lower, Dipo0, PopuT0, DipoT arrays) and W0/B0 tables
through their nine sub-table dicts, so any container with those fields works.  The
partition sums come from the C oracle (TIPS-2024, partition.jl), w(z) from
scipy.special.wofz and the eigen-decomposition from numpy.linalg.eig (LAPACK zgeev),
library routines independent of the product.
"""
import ctypes as C
import math

import numpy as np
from scipy.special import wofz

import refcpu

T0 = 296.0                      # line_mixing.jl:30
CT = 1.4388                     # line_mixing.jl:31
CTGAMD = 3.5812e-7              # line_mixing.jl:755
SQRT_LN2 = math.sqrt(math.log(2.0))
INV_SQRT_PI = 1.0 / math.sqrt(math.pi)
X_FAR = 122.0                   # voigt.jl:13
MASS = {2: {1: 43.98983, 2: 44.99318, 3: 45.99398, 4: 44.99403, 5: 45.99706, 6: 44.99706,
            7: 47.99832, 8: 47.00134, 9: 48.00196, 10: 47.00427},
        6: {1: 16.031300, 2: 17.034655, 3: 17.037475, 4: 18.040830}}      # line_mixing.jl:743-753


def _lib():
    L = refcpu.lib()
    L.orc_partition.argtypes = [C.c_int, C.c_int, C.c_double]
    L.orc_partition.restype = C.c_double
    return L


def mass(mol, iso):
    """_lm_mass_amu (line_mixing.jl:772)."""
    return MASS.get(int(mol), {}).get(int(iso), 44.0)


def q_ratio(mol, iso, T):
    """Q(296)/Q(T) of the band's isotopologue (line_mixing.jl:615, 1198)."""
    L = _lib()
    return L.orc_partition(int(mol), int(iso), T0) / L.orc_partition(int(mol), int(iso), T)


def coupling(band, wt):
    """The ordered pairs (i, j) with Ji_j <= Ji_i that pass the isotopologue symmetry
    rule and have a W0/B0 entry, grouped by i in line order (line_mixing.jl:530-578)."""
    n = len(band.nu)
    sym = band.isot > 2 and band.isot != 7 and band.isot != 10
    Ji = [int(v) for v in band.Ji]
    br = [int(v) for v in band.branch]
    rows = []
    for i in range(n):
        jji, jjf = Ji[i], Ji[i] + br[i]
        row = []
        for j in range(n):
            if j == i:
                continue
            jjip, jjfp = Ji[j], Ji[j] + br[j]
            if jjip > jji:
                continue
            if sym and abs(jji - jjip) % 2 != 0:
                continue
            if band.li <= band.lf:
                a, b, c, d = jji, jjf, jjip, jjfp
            else:
                a, b, c, d = jjf, jji, jjfp, jjip
            ti = "p" if a > b else ("q" if a == b else "r")
            tip = "p" if c > d else ("q" if c == d else "r")
            v = getattr(wt, ti + tip).get((a, c))
            if v is None:
                continue
            row.append((j, v[0], v[1]))
        rows.append(row)
    return rows


def popu_T(band, T):
    """Lower-state populations at T (line_mixing.jl:614-619)."""
    return band.PopuT0 * q_ratio(band.molecule, band.isot, T) * np.exp(-CT * band.E_lower * (1.0 / T - 1.0 / T0))


def build_W(band, wt, T):
    """Sum-rule-corrected relaxation matrix and PopuT (line_mixing.jl:607-710)."""
    n = len(band.nu)
    rho = popu_T(band, T)
    gam = band.gV_air * (T0 / T) ** band.n_air.astype(np.float64)
    li, lf = band.li, band.lf
    if li > 8 or lf > 8 or abs(li - lf) > 1:
        return np.diag(gam), rho
    S_eff = band.nu * rho * band.DipoT ** 2
    order = np.argsort(-S_eff, kind="stable")
    rows = coupling(band, wt)
    W = np.zeros((n, n))
    dl = math.log(T0 / T)
    for i in order:
        for j, w0, b0 in rows[i]:
            y = math.exp(w0 - b0 * dl)
            W[j, i] = y
            W[i, j] = y * rho[i] / rho[j]
    off = ~np.eye(n, dtype=bool)
    W[off] = -np.abs(W[off])
    W[np.diag_indices(n)] = gam
    sym = band.isot > 2 and band.isot != 7 and band.isot != 10
    Ji = band.Ji.astype(np.int64)
    D0 = np.abs(band.Dipo0)
    for ir in range(n):
        i = order[ir]
        lw = up = 0.0
        for irp in range(n):
            j = order[irp]
            if sym and abs(Ji[i] - Ji[j]) % 2 != 0:
                continue
            if irp > ir:
                lw += D0[j] * W[j, i]
            else:
                up += D0[j] * W[j, i]
        if lw == 0.0:
            continue
        r = -up / lw
        for irp in range(ir + 1, n):
            j = order[irp]
            W[j, i] *= r
            W[i, j] = W[j, i] * rho[i] / rho[j]
    return W, rho


def calc_Y(band, wt, T):
    """Y_i = 2 sum_{j != i} |D_j|/|D_i| W_ji / (nu_i - nu_j), |dnu| >= 1e-4 (line_mixing.jl:712-738)."""
    n = len(band.nu)
    if band.li > 8 or band.lf > 8 or abs(band.li - band.lf) > 1:
        return np.zeros(n)
    W, _ = build_W(band, wt, T)
    D = np.abs(band.DipoT)
    Y = np.zeros(n)
    for i in range(n):
        s = 0.0
        for j in range(n):
            if j == i:
                continue
            d = band.nu[i] - band.nu[j]
            if abs(d) < 1e-4:
                d = 1e-4 if d >= 0.0 else -1e-4
            s += 2.0 * D[j] / D[i] * (1.0 / d) * W[j, i]
        Y[i] = s
    return Y


def band_strength(band):
    """_band_eff_strength (line_mixing.jl:804-811)."""
    return float(np.sum(band.DipoT ** 2 * band.PopuT0 * band.nu * (1.0 - np.exp(-CT * band.nu / T0))))


def dispersive_lines(band, Y, T, p):
    """(centre, f, y, weight) of the lines with Y != 0 and S(T) > 0 (line_mixing.jl:1193-1221)."""
    M = mass(band.molecule, band.isot)
    qr = q_ratio(band.molecule, band.isot, T)
    out = []
    for k in range(len(band.nu)):
        if Y[k] == 0.0:
            continue
        nu = band.nu[k]
        c = nu + float(band.shift[k]) * p
        gd = CTGAMD * nu * math.sqrt(T / M)
        gl = band.gV_air[k] * (T0 / T) ** float(band.n_air[k]) * p
        f = SQRT_LN2 / max(gd, 1e-10)
        s0t = 1.0 - math.exp(-CT * nu / T0)
        stt = 1.0 - math.exp(-CT * nu / T)
        S0 = band.DipoT[k] ** 2 * band.PopuT0[k] * nu * s0t
        ST = S0 * qr * math.exp(-CT * band.E_lower[k] * (1.0 / T - 1.0 / T0)) * stt / s0t
        if ST <= 0.0:
            continue
        out.append((c, f, gl * f, Y[k] * p * ST * f * INV_SQRT_PI))
    return out


def _w(x, y):
    return wofz(np.asarray(x) + 1j * np.asarray(y))


def dispersive_eval(nu, lines, cutoff=25.0, x_far=X_FAR):
    """Sum of weight Im w(z) over |nu - c| <= cutoff, far form beyond |x| > x_far
    (line_mixing.jl:1166-1186)."""
    nu = np.asarray(nu, np.float64)
    out = np.zeros(len(nu))
    for c, f, y, w in sorted(lines, key=lambda t: t[0]):
        m = np.abs(nu - c) <= cutoff
        if not m.any():
            continue
        x = (nu[m] - c) * f
        im = x * INV_SQRT_PI / (x * x + y * y)
        near = np.abs(x) <= x_far
        if near.any():
            im[near] = _w(x[near], y).imag
        out[m] += w * im
    return out


def _wt_of(relmat, band):
    return relmat.wtfit.get((min(band.li, band.lf), max(band.li, band.lf)))


def lm_dispersive_correction(nu, relmat, T, p, cutoff=25.0, min_band_strength=0.0, x_far=X_FAR, skip=None,
                             isotopes=None):
    """compute_lm_dispersive_correction (line_mixing.jl:1303-1349)."""
    lines = []
    for b in relmat.bands:
        if b.li > 8:
            continue
        if isotopes is not None and b.isot not in isotopes:
            continue
        if min_band_strength > 0.0 and band_strength(b) < min_band_strength:
            continue
        wt = _wt_of(relmat, b)
        if wt is None:
            continue
        if skip is not None and b.name in skip:
            continue
        lines += dispersive_lines(b, calc_Y(b, wt, T), T, p)
    return dispersive_eval(nu, lines, cutoff, x_far)


def band_modes(band, wt, T, p, keep_threshold=0.0):
    """Poles, amplitudes and band Doppler factor of D + i p W (line_mixing.jl:892-970)."""
    n = len(band.nu)
    W, rho = build_W(band, wt, T)
    d = band.DipoT * np.sqrt(band.nu * (1.0 - np.exp(-CT * band.nu / T)))
    M = np.diag(band.nu.astype(np.complex128)) + 1j * p * W * (~np.eye(n, dtype=bool))
    M[np.diag_indices(n)] = band.nu + 1j * p * np.diag(W)
    lam, X = np.linalg.eig(M)
    a = X.T @ d
    b = np.linalg.solve(X, rho * d)
    A = a * b
    gd = CTGAMD * (np.sum(band.nu) / n) * math.sqrt(T / mass(band.molecule, band.isot))
    f = SQRT_LN2 / max(gd, 1e-10)
    keep = lam.imag > 0
    if keep_threshold > 0 and keep.any():
        keep &= np.abs(A) >= keep_threshold * np.max(np.abs(A[keep]))
    return lam[keep], A[keep], f


def diagonal_modes(band, T, p, f):
    """No-coupling poles nu + i gamma_L and amplitudes S(T) (line_mixing.jl:1030-1063)."""
    gl = band.gV_air * (T0 / T) ** band.n_air.astype(np.float64) * p
    rho = popu_T(band, T)
    ST = band.DipoT ** 2 * rho * band.nu * (1.0 - np.exp(-CT * band.nu / T))
    return band.nu + 1j * gl, ST.astype(np.complex128), f


def vpw_band_xsec(nu, modes, cutoff=25.0, x_far=X_FAR):
    """(f/sqrt(pi)) Re sum conj(A_k) w(z_k) over |nu - Re lam_k| <= cutoff (line_mixing.jl:972-1028)."""
    lam, A, f = modes
    nu = np.asarray(nu, np.float64)
    out = np.zeros(len(nu))
    for k in range(len(lam)):
        m = np.abs(nu - lam[k].real) <= cutoff
        if not m.any():
            continue
        x = (nu[m] - lam[k].real) * f
        y = lam[k].imag * f
        den = INV_SQRT_PI / (x * x + y * y)
        re, im = y * den, x * den
        near = np.abs(x) <= x_far
        if near.any():
            wz = _w(x[near], y)
            re[near], im[near] = wz.real, wz.imag
        out[m] += A[k].real * re + A[k].imag * im
    return out * f * INV_SQRT_PI


def vpw_band_perturbation(nu, band, wt, T, p, cutoff=25.0, keep_threshold=1e-4, x_far=X_FAR):
    """Full eigenmode sum minus its no-coupling limit (line_mixing.jl:1065-1091)."""
    full = band_modes(band, wt, T, p, keep_threshold)
    return vpw_band_xsec(nu, full, cutoff, x_far) - vpw_band_xsec(nu, diagonal_modes(band, T, p, full[2]), cutoff,
                                                                  x_far)


def vpw_perturbation(nu, relmat, T, p, whitelist, cutoff=25.0, keep_threshold=1e-4, min_band_strength=0.0,
                     x_far=X_FAR, isotopes=None):
    """_vpw_perturbation (line_mixing.jl:1120-1154)."""
    d = lm_dispersive_correction(nu, relmat, T, p, cutoff, min_band_strength, x_far, skip=whitelist,
                                 isotopes=isotopes)
    for b in relmat.bands:
        if b.name not in whitelist or b.li > 8:
            continue
        if isotopes is not None and b.isot not in isotopes:
            continue
        if min_band_strength > 0.0 and band_strength(b) < min_band_strength:
            continue
        wt = _wt_of(relmat, b)
        if wt is None:
            continue
        d = d + vpw_band_perturbation(nu, b, wt, T, p, cutoff, keep_threshold, x_far)
    return d
