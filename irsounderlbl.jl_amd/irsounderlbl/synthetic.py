"""Synthetic workloads: profiles and line lists of the benchmark configurations.

Deterministic (numpy PCG64, fixed seeds) so the oracle and the HIP path see
bit-identical inputs.  Decisions where the workload text is open:

* levels: p_k = 1013.25 * (0.01/1013.25)**((k-1)/49), k = 1..50 (hPa)
* T(z) = max(288 - 6.5 z, 217) K, z(p) from hypsometric integration
  (R_d = 287.05 J/kg/K, g = 9.80665 m/s^2), RK4 in ln p
* H2O: q_k = 1e-2 * (4e-4)**((k-1)/49);  CO2: 4.15e-4
* O3 (full-range workloads): 8e-6 * exp(-((z-33)/9)**2) + 3e-8 (an analytic
  stratospheric layer; the workload text gives no O3 profile)
* FP32 line fields are rounded to FP32 before anyone uses them
* Y ~ N(0, 1e-2) atm^-1 on the 15 um CO2-like lines only
"""
from __future__ import annotations

import math

import numpy as np

from .types import AtmosphericProfile, GasSpecies, HITRANLinelist, scan_angle_to_local_zenith

R_D = 287.05
G0 = 9.80665
N_LEVELS = 50


def pressure_levels(n_levels: int = N_LEVELS) -> np.ndarray:
    k = np.arange(n_levels, dtype=np.float64)
    return 1013.25 * (0.01 / 1013.25) ** (k / (n_levels - 1))


def _t_of_z(z_km):
    return max(288.0 - 6.5 * z_km, 217.0)


def hypsometric_altitudes(p: np.ndarray, t_of_z=_t_of_z, substeps: int = 64) -> np.ndarray:
    """z(p) in km by integrating dz/dlnp = -R_d T(z)/g with RK4."""
    z = np.zeros(len(p))
    for k in range(len(p) - 1):
        h = (math.log(p[k + 1]) - math.log(p[k])) / substeps
        zz = z[k]
        f = lambda zk: -R_D * t_of_z(zk) / G0 / 1000.0
        for _ in range(substeps):
            k1 = f(zz); k2 = f(zz + 0.5 * h * k1); k3 = f(zz + 0.5 * h * k2); k4 = f(zz + h * k3)
            zz += h / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
        z[k + 1] = zz
    return z


def base_profile(with_o3: bool = False, n_levels: int = N_LEVELS) -> AtmosphericProfile:
    p = pressure_levels(n_levels)
    z = hypsometric_altitudes(p)
    T = np.array([_t_of_z(zz) for zz in z])
    k = np.arange(n_levels, dtype=np.float64)
    vmr = {GasSpecies.H2O: 1e-2 * (4e-4) ** (k / (n_levels - 1)),
           GasSpecies.CO2: np.full(n_levels, 4.15e-4)}
    if with_o3:
        vmr[GasSpecies.O3] = 8e-6 * np.exp(-((z - 33.0) / 9.0) ** 2) + 3e-8
    return AtmosphericProfile(p, T, z, vmr)


def _lines(rng, n, lo, hi, mol, g_air=(0.05, 0.10), band=None):
    if band is None:
        nu = rng.uniform(lo, hi, n)
    else:
        nu = np.concatenate([rng.uniform(a, b, m) for (a, b, m) in band])
        n = len(nu)
    S = 10.0 ** rng.uniform(-27.0, -18.0, n)
    ga = rng.uniform(*g_air, n).astype(np.float32)
    gs = rng.uniform(0.08, 0.15, n).astype(np.float32)
    na = rng.uniform(0.5, 0.8, n).astype(np.float32)
    el = rng.uniform(0.0, 3000.0, n)
    da = rng.uniform(-0.005, 0.0, n).astype(np.float32)
    return nu, S, ga, gs, na, el, da


def linelist_config0(seed: int = 1000) -> HITRANLinelist:
    """20,000 CO2-like lines, nu ~ U(620, 825) cm^-1 (the 645–800 cm^-1 workload)."""
    rng = np.random.default_rng(seed)
    nu, S, ga, gs, na, el, da = _lines(rng, 20000, 620.0, 825.0, 2)
    return HITRANLinelist(2, 1, nu, S, ga, gs, el, na, da)


def linelists_config1(seed: int = 1001, with_lm: bool = True, windows: bool = False):
    """300,000 lines in 620–2785 cm^-1: CO2-like 90k U(640,760) + 90k U(2250,2400),
    H2O-like 100k U(620,2785) (gamma_air U(.02,.10)), O3-like 20k U(980,1080);
    Y ~ N(0, 1e-2) atm^-1 on the 640–760 CO2-like lines.

    With these intensities (S = 10^U(-27,-18) for every species) the H2O-like
    wings alone make the base profile opaque across IASI: every channel's
    brightness temperature is the 217 K of the upper levels.  That is the
    performance workload (the work per spectrum does not depend on the values).
    windows=True maps the H2O-like intensity exponents linearly onto
    U(-27, -23), with every other field (positions, widths, CO2, O3) drawn
    identically, so that the spectrum has partially transparent windows
    (brightness temperatures 217-282 K; 244-281 K between the bands): the
    variant used for brightness-temperature accuracy checks and the emulator."""
    rng = np.random.default_rng(seed)
    n15 = 90000
    nu, S, ga, gs, na, el, da = _lines(rng, 0, 0, 0, 2, band=[(640.0, 760.0, n15), (2250.0, 2400.0, 90000)])
    y = np.zeros(len(nu))
    if with_lm:
        y[:n15] = rng.normal(0.0, 1e-2, n15)
    co2 = HITRANLinelist(2, 1, nu, S, ga, gs, el, na, da, y_lm=y if with_lm else None)
    nu, S, ga, gs, na, el, da = _lines(rng, 100000, 620.0, 2785.0, 1, g_air=(0.02, 0.10))
    if windows:
        S = 10.0 ** (-27.0 + (np.log10(S) + 27.0) * (4.0 / 9.0))
    h2o = HITRANLinelist(1, 1, nu, S, ga, gs, el, na, da)
    nu, S, ga, gs, na, el, da = _lines(rng, 20000, 980.0, 1080.0, 3)
    o3 = HITRANLinelist(3, 1, nu, S, ga, gs, el, na, da)
    return {GasSpecies.CO2: co2, GasSpecies.H2O: h2o, GasSpecies.O3: o3}


def _matern52(z, L):
    r = np.abs(z[:, None] - z[None, :]) / L
    return (1.0 + math.sqrt(5.0) * r + 5.0 * r * r / 3.0) * np.exp(-math.sqrt(5.0) * r)


def perturbed_profiles(n_prof: int, seed: int = 2002, sigma_T: float = 3.0, sigma_h2o: float = 0.3,
                       sigma_o3: float = 0.0, co2_range=None, dT_sfc: float = 10.0,
                       corr_km: float = 2.0, with_o3: bool = True, max_scan_deg: float = 48.3,
                       sigma_lnq: float = 0.0):
    """Scan-line profiles: base profile with T + GP(0, sigma_T^2, Matern-5/2, L = corr_km),
    H2O x lognormal(sigma_h2o), O3 x lognormal(sigma_o3), CO2 ~ U(co2_range),
    T_sfc = T_1 + U(-dT_sfc, dT_sfc), scan ~ U(0, max_scan_deg) -> mu.  sigma_lnq > 0
    also multiplies H2O level by level by exp(GP(0, sigma_lnq^2, Matern-5/2, L = corr_km))
    (drawn after the other perturbations of the profile, so sigma_lnq = 0 keeps every
    seed's profiles).  Returns a list of (AtmosphericProfile, T_sfc, mu)."""
    rng = np.random.default_rng(seed)
    base = base_profile(with_o3=with_o3)
    z = base.altitude
    C = sigma_T ** 2 * _matern52(z, corr_km) + 1e-10 * np.eye(len(z))
    Lc = np.linalg.cholesky(C)
    out = []
    for _ in range(n_prof):
        T = base.temperature + Lc @ rng.standard_normal(len(z))
        vmr = dict(base.vmr)
        vmr[GasSpecies.H2O] = base.vmr[GasSpecies.H2O] * math.exp(sigma_h2o * rng.standard_normal())
        if with_o3 and sigma_o3 > 0:
            vmr[GasSpecies.O3] = base.vmr[GasSpecies.O3] * math.exp(sigma_o3 * rng.standard_normal())
        if co2_range is not None:
            vmr[GasSpecies.CO2] = np.full(len(z), rng.uniform(*co2_range))
        T_sfc = T[0] + rng.uniform(-dT_sfc, dT_sfc)
        scan = rng.uniform(0.0, max_scan_deg)
        if sigma_lnq > 0.0:
            vmr[GasSpecies.H2O] = vmr[GasSpecies.H2O] * np.exp((sigma_lnq / sigma_T) * (Lc @ rng.standard_normal(len(z))))
        mu = math.cos(math.radians(scan_angle_to_local_zenith(scan)))
        out.append((AtmosphericProfile(base.pressure.copy(), T, z.copy(), vmr), float(T_sfc), float(mu)))
    return out


def relmat_synthetic(seed: int = 3001):
    """A CO2-like relaxation matrix (no HITRAN relaxation data travels with this
    package): P, Q and R branches of five bands around 15 um --
      "S626_01" isot 1, l 0 -> 1 at 667.38 (J even, every branch, 21 J values),
      "S636_11" isot 2, l 1 -> 1 at 648.48,
      "S628_12" isot 3, l 1 -> 2 at 720.80 (the isotopologue symmetry rule applies),
      "S626_21" isot 1, l 2 -> 1 at 741.72 (li > lf: the swapped W0/B0 lookup),
      "S626_23" isot 1, l 2 -> 3 at 618.03, whose (2, 3) table is absent --
    with W0/B0 tables (0, 1), (1, 1), (1, 2) over every branch pair: W(296 K)
    0.02 exp(-|dJ| / 8) cm^-1/atm, temperature exponent 0.7 + 0.1 dJ / 40.  Line
    intensities 1e-22..1e-19 at 296 K, gamma_air 0.075 - 0.0004 J, n 0.75, shifts
    -0.002 - 2e-5 J."""
    from .relmat import HITRANRelmatData, RelmatBand, RelmatLine, W0B0Table
    rng = np.random.default_rng(seed)
    Bc = 0.39
    specs = (("S626_01", 1, 0, 1, 667.38, 2, 1.0), ("S636_11", 2, 1, 1, 648.48, 1, 0.05),
             ("S628_12", 3, 1, 2, 720.80, 1, 0.01), ("S626_21", 1, 2, 1, 741.72, 2, 0.1),
             ("S626_23", 1, 2, 3, 618.03, 2, 0.02))
    bands = []
    for name, iso, li, lf, nu0, jstep, scale in specs:
        lines = []
        for br in (-1, 0, 1):
            if br == 0 and li == 0 and lf == 0:
                continue
            for Ji in range(1 if br == -1 else 0, 41, jstep):
                if br == -1 and Ji == 0:
                    continue
                Jf = Ji + br
                nu = nu0 + Bc * (Jf * (Jf + 1) - Ji * (Ji + 1)) - 1e-4 * Ji * (Ji + 1) * (br == 0)
                E = Bc * Ji * (Ji + 1)
                pop = (2 * Ji + 1) * math.exp(-1.4388 * E / 296.0) / 290.0
                S = scale * 1e-19 * (2 * Ji + 1) * math.exp(-1.4388 * E / 296.0) * 10 ** rng.uniform(-0.3, 0.3)
                stim = 1.0 - math.exp(-1.4388 * nu / 296.0)
                dipoT = math.sqrt(S / (pop * nu * stim))
                lines.append(RelmatLine(iso, nu, Ji, br, 0.075 - 0.0004 * Ji, float(np.float32(0.75)),
                                        float(np.float32(-0.002 - 2e-5 * Ji)), E, dipoT * (1.0 + 0.05 * rng.random()),
                                        pop, dipoT))
        lines.sort(key=lambda l: l.nu)
        bands.append(RelmatBand(name, li, lf, iso, min(l.nu for l in lines), max(l.nu for l in lines), lines))
    wt = {}
    for key in ((0, 1), (1, 1), (1, 2)):
        t = W0B0Table()
        for sub in ("pp", "pq", "pr", "qp", "qq", "qr", "rp", "rq", "rr"):
            d = getattr(t, sub)
            for a in range(0, 42):
                for b in range(0, 42):
                    dj = abs(a - b)
                    d[(a, b)] = (math.log(0.02 * math.exp(-dj / 8.0)), 0.7 + 0.1 * dj / 40.0)
        wt[key] = t
    return HITRANRelmatData(bands, wt, GasSpecies.CO2)


def linelist_from_relmat(rm) -> HITRANLinelist:
    """The HITRAN-style line list of a relaxation matrix's S-file lines (the Voigt
    baseline the line-mixing perturbation is added to): S(296) = DipoT^2 PopuT0 nu
    (1 - exp(-1.4388 nu / 296)), gamma_air, n_air, the pressure shift and E'' of the
    S-file, gamma_self 0.1; isotopologue 10 written as HITRAN's 0."""
    mol, iso, nu, S, ga, gs, E, n, d = ([] for _ in range(9))
    for b in rm.bands:
        stim = 1.0 - np.exp(-1.4388 * b.nu / 296.0)
        mol += [b.molecule] * b.n
        iso += [0 if b.isot == 10 else b.isot] * b.n
        nu.append(b.nu)
        S.append(b.DipoT ** 2 * b.PopuT0 * b.nu * stim)
        ga.append(b.gV_air)
        gs.append(np.full(b.n, 0.1))
        E.append(b.E_lower)
        n.append(b.n_air)
        d.append(b.shift)
    cat = np.concatenate
    return HITRANLinelist(np.array(mol, np.int8), np.array(iso, np.int8), cat(nu), cat(S),
                          np.float32(cat(ga)), np.float32(cat(gs)), cat(E), np.float32(cat(n)), np.float32(cat(d)))
