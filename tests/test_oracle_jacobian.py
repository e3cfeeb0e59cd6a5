"""Pins the oracle's analytic Jacobians against finite differences of the
oracle's own forward model (the reference's ground-truth discipline,
src/Estimation/jacobian_fd.jl; tolerances of test/runtests.jl:1128-1478)."""
import math

import numpy as np
import pytest

import refcpu as o
from irsounderlbl import synthetic as syn
from irsounderlbl.types import AtmosphericProfile, GasSpecies, HITRANLinelist, Sounder, wavenumber_grid


def small_problem():
    """Semi-transparent 50-level problem: weak CO2-like and H2O-like lines near 720 cm^-1."""
    rng = np.random.default_rng(5)
    def lines(mol, n, lo, hi, smin, smax):
        return HITRANLinelist(mol, 1, rng.uniform(lo, hi, n), 10 ** rng.uniform(smin, smax, n),
                              rng.uniform(0.05, 0.1, n), rng.uniform(0.08, 0.15, n),
                              rng.uniform(0, 1500, n), rng.uniform(0.5, 0.8, n),
                              rng.uniform(-0.005, 0, n))
    lls = {GasSpecies.CO2: lines(2, 300, 700.0, 740.0, -25.0, -21.0),
           GasSpecies.H2O: lines(1, 150, 700.0, 740.0, -24.0, -21.5)}
    prof = syn.base_profile(False)
    snd = Sounder.make(716.0, 724.0, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    return prof, lls, snd


KW = dict(internal_dnu=0.005)


def fd_column(prof, lls, snd, kind, j, T_sfc, eps, step):
    def run(dx):
        T = prof.temperature.copy()
        vmr = {k: v.copy() for k, v in prof.vmr.items()}
        ts, e = T_sfc, eps
        if kind == "T":
            T[j] += dx
        elif kind == "H2O":
            vmr[GasSpecies.H2O][j] *= math.exp(dx)
        elif kind == "Tsfc":
            ts += dx
        elif kind == "eps":
            e += dx
        p2 = AtmosphericProfile(prof.pressure, T, prof.altitude, vmr)
        _, bt, _ = o.forward(p2, lls, snd, dptmn=0.0, T_sfc=ts, eps_sfc=e, mu=0.9, **KW)
        return bt
    return (run(step) - run(-step)) / (2 * step)


@pytest.mark.parametrize("sf", [1, 0])
def test_analytic_jacobian_matches_fd(sf):
    prof, lls, snd = small_problem()
    T_sfc, eps = 292.0, 0.97
    K, y0, x0 = o.jacobian(prof, lls, snd, retrieved=[GasSpecies.H2O], T_sfc=T_sfc, eps_sfc=eps,
                           mu=0.9, source_function=sf, **KW)
    nlev = prof.n_levels
    assert K.shape == (33, 2 * nlev + 2)
    _, bt0, _ = o.forward(prof, lls, snd, dptmn=0.0, T_sfc=T_sfc, eps_sfc=eps, mu=0.9,
                          source_function=sf, **KW)
    assert np.allclose(y0, bt0, rtol=1e-12, atol=0)
    checks = [("T", j, j, 0.5, 3e-3) for j in (0, 8, 20, 35, 49)]
    checks += [("H2O", j, nlev + j, 1e-3, 1e-4) for j in (0, 4, 12, 30)]
    checks += [("Tsfc", 0, 2 * nlev, 0.5, 1e-4), ("eps", 0, 2 * nlev + 1, 1e-3, 1e-4)]
    for kind, j, col, step, tol in checks:
        if sf == 0:
            fd = fd_column_toon(prof, lls, snd, kind, j, T_sfc, eps, step)
        else:
            fd = fd_column(prof, lls, snd, kind, j, T_sfc, eps, step)
        scale = max(np.abs(fd).max(), 1e-6)
        err = np.abs(K[:, col] - fd).max() / scale
        print(kind, j, f"|K| {np.abs(K[:, col]).max():.3e} rel err {err:.2e}")
        assert err < tol, (kind, j, err)


def fd_column_toon(prof, lls, snd, kind, j, T_sfc, eps, step):
    def run(dx):
        T = prof.temperature.copy()
        vmr = {k: v.copy() for k, v in prof.vmr.items()}
        ts, e = T_sfc, eps
        if kind == "T":
            T[j] += dx
        elif kind == "H2O":
            vmr[GasSpecies.H2O][j] *= math.exp(dx)
        elif kind == "Tsfc":
            ts += dx
        elif kind == "eps":
            e += dx
        p2 = AtmosphericProfile(prof.pressure, T, prof.altitude, vmr)
        _, bt, _ = o.forward(p2, lls, snd, dptmn=0.0, T_sfc=ts, eps_sfc=e, mu=0.9, source_function=0, **KW)
        return bt
    return (run(step) - run(-step)) / (2 * step)


def test_rte_jacobian_vs_fd_and_kirchhoff():
    """runtests.jl:1128-1189: dI/dtau, dI/dT_lev, dI/dT_sfc, dI/deps vs FD; isothermal -> dI/dtau = 0."""
    g = wavenumber_grid(700.0, 760.0, 5.0)
    rng = np.random.default_rng(2)
    Ln = 6
    tau = rng.uniform(0.05, 1.5, (g.n, Ln))
    T_lev = np.linspace(290, 230, Ln + 1)
    p_lev = np.geomspace(1000, 100, Ln + 1)
    T_ave = np.array([o.lib().orc_cg_temperature_mass(T_lev[k], T_lev[k + 1], p_lev[k], p_lev[k + 1])
                      for k in range(Ln)])
    for sf in (0, 1):
        for eps in (1.0, 0.9):
            I, dtau, dT, dts, de = o.rte_jacobian(g.nu_min, g.dnu, tau, T_lev, p_lev, 295.0, 0.8, eps, sf)
            f = lambda tau_, Tl, Ts, e: o.schwarzschild_rte(
                g.nu_min, g.dnu, tau_, Tl, Ts, 0.8, e, sf,
                np.array([o.lib().orc_cg_temperature_mass(Tl[k], Tl[k + 1], p_lev[k], p_lev[k + 1])
                          for k in range(Ln)]))
            assert np.allclose(I, f(tau, T_lev, 295.0, eps), rtol=1e-12, atol=0)
            for m in range(Ln):
                h = 1e-6
                tp = tau.copy(); tp[:, m] += h
                tm = tau.copy(); tm[:, m] -= h
                fd = (f(tp, T_lev, 295.0, eps) - f(tm, T_lev, 295.0, eps)) / (2 * h)
                assert np.max(np.abs(dtau[:, m] - fd)) < 1e-6 * np.max(np.abs(I))
            for j in range(Ln + 1):
                h = 1e-4
                Tp = T_lev.copy(); Tp[j] += h
                Tm = T_lev.copy(); Tm[j] -= h
                fd = (f(tau, Tp, 295.0, eps) - f(tau, Tm, 295.0, eps)) / (2 * h)
                assert np.max(np.abs(dT[:, j] - fd)) < 1e-7 * np.max(np.abs(I))
            fd = (f(tau, T_lev, 295.0 + 1e-4, eps) - f(tau, T_lev, 295.0 - 1e-4, eps)) / 2e-4
            assert np.max(np.abs(dts - fd)) < 1e-7 * np.max(np.abs(I))
    # Kirchhoff: isothermal black column, dI/dtau == 0
    I, dtau, *_ = o.rte_jacobian(g.nu_min, g.dnu, tau, np.full(Ln + 1, 270.0), p_lev, 270.0, 1.0, 1.0, 1)
    assert np.max(np.abs(dtau)) < 1e-9 * np.max(I)


def test_cg_gradients_vs_fd():
    """runtests.jl:1192-1219: CG column-VMR gradient vs FD; well-mixed limit = TBAR frac."""
    L = o._declare_jac()
    import ctypes as C
    for v1, v2, p1, p2 in [(1e-2, 6e-3, 1013.25, 800.0), (5e-6, 4e-6, 10.0, 7.9), (4e-4, 4e-4, 500.0, 400.0)]:
        a, b = C.c_double(), C.c_double()
        L.orc_cg_column_vmr_grad(v1, v2, p1, p2, C.byref(a), C.byref(b))
        h1, h2 = v1 * 1e-6, v2 * 1e-6
        fd1 = (L.orc_cg_column_vmr(v1 + h1, v2, p1, p2) - L.orc_cg_column_vmr(v1 - h1, v2, p1, p2)) / (2 * h1)
        fd2 = (L.orc_cg_column_vmr(v1, v2 + h2, p1, p2) - L.orc_cg_column_vmr(v1, v2 - h2, p1, p2)) / (2 * h2)
        if v1 != v2:
            assert a.value == pytest.approx(fd1, rel=1e-6) and b.value == pytest.approx(fd2, rel=1e-6)
        else:
            fr = L.orc_cg_mass_frac(p1, p2)
            assert a.value == pytest.approx(1 - fr) and b.value == pytest.approx(fr)


def test_xsec_grad_vs_fd():
    """runtests.jl:1285-1343: d sigma/dT, dp, dvmr_self vs FD (T step inside one 1 K TIPS cell)."""
    ll = HITRANLinelist(1, 1, [700.0, 700.3, 701.1], [1e-20, 2e-21, 6e-21], [0.07, 0.05, 0.09],
                        [0.3, 0.35, 0.4], [200.0, 900.0, 50.0], [0.7, 0.6, 0.75], [-0.002, 0.001, -0.004])
    g = wavenumber_grid(699.0, 702.0, 0.002)
    T, p, vs = 250.4, 0.4, 0.01
    s, dT, dp, dv = o.species_grad(ll, g.nu_min, g.dnu, g.n, T, p, vs, cutoff=25.0)
    hT, hp, hv = 0.05, 1e-5, 1e-6
    fdT = (o.voigt_xsec(ll, g.nu_min, g.dnu, g.n, T + hT, p, vs) - o.voigt_xsec(ll, g.nu_min, g.dnu, g.n, T - hT, p, vs)) / (2 * hT)
    fdp = (o.voigt_xsec(ll, g.nu_min, g.dnu, g.n, T, p + hp, vs) - o.voigt_xsec(ll, g.nu_min, g.dnu, g.n, T, p - hp, vs)) / (2 * hp)
    fdv = (o.voigt_xsec(ll, g.nu_min, g.dnu, g.n, T, p, vs + hv) - o.voigt_xsec(ll, g.nu_min, g.dnu, g.n, T, p, vs - hv)) / (2 * hv)
    for a, b in [(dT, fdT), (dp, fdp), (dv, fdv)]:
        assert np.max(np.abs(a - b)) < 1e-5 * np.max(np.abs(b))
