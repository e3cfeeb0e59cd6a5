"""ctypes bindings for the FP64 CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the product.
Build with `make -C oracle`.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")

_lib = None

i64, f64, i32 = C.c_int64, C.c_double, C.c_int
P = C.c_void_p


class OrcLinelist(C.Structure):
    _fields_ = [("n", i64), ("mol", P), ("iso", P), ("nu", P), ("S", P), ("gamma_air", P),
                ("gamma_self", P), ("elower", P), ("n_air", P), ("delta_air", P), ("y_lm", P)]


class OrcProfile(C.Structure):
    _fields_ = [("n_levels", i32), ("p_hpa", P), ("T", P), ("n_species", i32), ("species", P),
                ("vmr", P), ("lines", P)]


class OrcSounder(C.Structure):
    _fields_ = [("nu_min", f64), ("nu_max", f64), ("dnu", f64), ("opd_max", f64), ("fwhm", f64)]


class OrcOptions(C.Structure):
    _fields_ = [("internal_dnu", f64), ("high_res_factor", i32), ("cutoff", f64),
                ("lm_cutoff", f64), ("dptmn", f64), ("x_far", f64), ("with_ils", i32),
                ("apodization", i32), ("line_mixing", i32), ("T_method", i32),
                ("source_function", i32), ("T_sfc", f64), ("eps_sfc", f64), ("mu", f64),
                ("ils_direct_full", i32), ("continua", i32)]


class OrcInfo(C.Structure):
    _fields_ = [("nu_lo", f64), ("dnu", f64), ("n_hi", i64), ("n_ch", i64),
                ("seconds_xsec", f64), ("seconds_rte", f64), ("seconds_ils", f64),
                ("active_pairs", i64), ("active_lines", i64)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_faddeeva.argtypes = [f64, f64, C.POINTER(f64), C.POINTER(f64)]
        for name in ("orc_erfcx",):
            getattr(L, name).argtypes = [f64]
            getattr(L, name).restype = f64
        for name in ("orc_partition", "orc_q_ratio", "orc_partition_deriv"):
            getattr(L, name).argtypes = [i32, i32, f64]
            getattr(L, name).restype = f64
        for name in ("orc_planck", "orc_bt", "orc_dbdt"):
            getattr(L, name).argtypes = [f64, f64]
            getattr(L, name).restype = f64
        for name in ("orc_cim_correction", "orc_lit_correction", "orc_cim_correction_deriv",
                     "orc_lit_correction_deriv"):
            getattr(L, name).argtypes = [f64]
            getattr(L, name).restype = f64
        for name in ("orc_cg_column_vmr", "orc_cg_pressure", "orc_cg_temperature_mass"):
            getattr(L, name).argtypes = [f64, f64, f64, f64]
            getattr(L, name).restype = f64
        L.orc_scaled_intensity.argtypes = [C.POINTER(OrcLinelist), i64, f64]
        L.orc_scaled_intensity.restype = f64
        L.orc_broadened_width.argtypes = [C.POINTER(OrcLinelist), i64, f64, f64, f64,
                                          C.POINTER(f64), C.POINTER(f64)]
        L.orc_pressure_shift.argtypes = [C.POINTER(OrcLinelist), i64, f64, f64]
        L.orc_pressure_shift.restype = f64
        L.orc_reject_weak_lines.argtypes = [C.POINTER(OrcLinelist), f64, f64, f64, f64, f64, P]
        L.orc_reject_weak_lines.restype = i64
        L.orc_voigt_xsec.argtypes = [C.POINTER(OrcLinelist), P, f64, f64, i64, f64, f64, f64,
                                     f64, f64, P]
        L.orc_lm_dispersive.argtypes = [C.POINTER(OrcLinelist), f64, f64, i64, f64, f64, f64,
                                        f64, P]
        L.orc_voigt_lm_xsec.argtypes = [C.POINTER(OrcLinelist), f64, f64, i64, f64, f64, f64,
                                        f64, f64, P]
        L.orc_grid_len.argtypes = [f64, f64, f64]
        L.orc_grid_len.restype = i64
        L.orc_grid_node.argtypes = [f64, f64, i64]
        L.orc_grid_node.restype = f64
        L.orc_interp_logp.argtypes = [P, P, i32, P, i32, P]
        L.orc_schwarzschild_rte.argtypes = [f64, f64, i64, P, i32, P, f64, f64, f64, i32, P, P]
        L.orc_schwarzschild_rte.restype = i32
        L.orc_ils_half.argtypes = [f64]
        L.orc_ils_half.restype = i64
        L.orc_ils_kernel.argtypes = [f64, f64, f64, i32, i32, P]
        L.orc_ils_kernel.restype = i32
        L.orc_norton_beer.argtypes = [f64, i32]
        L.orc_norton_beer.restype = f64
        L.orc_apply_ils_at.argtypes = [P, i64, f64, P, i64, P, i64, P]
        L.orc_scan_to_local_zenith.argtypes = [f64, f64, f64]
        L.orc_scan_to_local_zenith.restype = f64
        L.orc_forward.argtypes = [C.POINTER(OrcProfile), C.POINTER(OrcSounder),
                                  C.POINTER(OrcOptions), P, P, P, P, C.POINTER(OrcInfo)]
        L.orc_forward.restype = i32
        L.orc_forward_lm.argtypes = [C.POINTER(OrcProfile), C.POINTER(OrcSounder), C.POINTER(OrcOptions),
                                     i32, P, P, P, C.POINTER(OrcInfo)]
        L.orc_forward_lm.restype = i32
        L.orc_forward_dims.argtypes = [C.POINTER(OrcSounder), C.POINTER(OrcOptions),
                                       C.POINTER(OrcInfo)]
        L.orc_layer_properties.argtypes = [C.POINTER(OrcProfile), i32, P, P, P, P, P, P, P]
        L.orc_num_threads.restype = i32
        L.orc_set_num_threads.argtypes = [i32]
        L.orc_set_num_threads.restype = None
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(P)


def make_linelist(ll):
    """OrcLinelist view of an irsounderlbl HITRANLinelist (keeps arrays alive)."""
    s = OrcLinelist(len(ll), _ptr(ll.mol_id), _ptr(ll.iso_id), _ptr(ll.wavenumber),
                    _ptr(ll.intensity), _ptr(ll.air_broad), _ptr(ll.self_broad),
                    _ptr(ll.lower_energy), _ptr(ll.temp_depend), _ptr(ll.pressure_shift),
                    _ptr(ll.y_lm))
    s._keep = ll
    return s


def faddeeva(x, y):
    re, im = f64(), f64()
    lib().orc_faddeeva(float(x), float(y), C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def grid(nu0, dnu, nu1):
    n = lib().orc_grid_len(nu0, dnu, nu1)
    return n, np.array([lib().orc_grid_node(nu0, dnu, i) for i in range(min(n, 3))])


def voigt_xsec(ll, nu0, dnu, n, T, p_atm, vmr_self=0.0, cutoff=25.0, x_far=122.0, keep=None):
    sig = np.zeros(n)
    k = None if keep is None else np.ascontiguousarray(keep, np.uint8)
    s = make_linelist(ll)
    lib().orc_voigt_xsec(C.byref(s), _ptr(k), nu0, dnu, n, T, p_atm, vmr_self, cutoff, x_far,
                         _ptr(sig))
    return sig


def lm_dispersive(ll, nu0, dnu, n, T, p_atm, cutoff=25.0, x_far=122.0):
    d = np.zeros(n)
    s = make_linelist(ll)
    lib().orc_lm_dispersive(C.byref(s), nu0, dnu, n, T, p_atm, cutoff, x_far, _ptr(d))
    return d


def voigt_lm_xsec(ll, nu0, dnu, n, T, p_atm, cutoff=25.0, lm_cutoff=25.0, x_far=122.0):
    sig = np.zeros(n)
    s = make_linelist(ll)
    lib().orc_voigt_lm_xsec(C.byref(s), nu0, dnu, n, T, p_atm, cutoff, lm_cutoff, x_far, _ptr(sig))
    return sig


def reject_weak_lines(ll, T, p_atm, coef, vmr_self=0.0, dptmn=1e-6):
    keep = np.zeros(len(ll), np.uint8)
    s = make_linelist(ll)
    lib().orc_reject_weak_lines(C.byref(s), T, p_atm, coef, vmr_self, dptmn, _ptr(keep))
    return keep.astype(bool)


def line_params(ll, T, p_atm, vmr_self=0.0):
    """(nu_shifted, S(T), gamma_L, gamma_D) per line."""
    s = make_linelist(ll)
    L = lib()
    n = len(ll)
    c = np.empty(n); S = np.empty(n); gl = np.empty(n); gd = np.empty(n)
    a, b = f64(), f64()
    for j in range(n):
        c[j] = L.orc_pressure_shift(C.byref(s), j, p_atm, vmr_self)
        S[j] = L.orc_scaled_intensity(C.byref(s), j, T)
        L.orc_broadened_width(C.byref(s), j, p_atm, T, vmr_self, C.byref(a), C.byref(b))
        gl[j] = a.value; gd[j] = b.value
    return c, S, gl, gd


def schwarzschild_rte(nu0, dnu, tau, T_lev, T_sfc, mu=1.0, eps=1.0, source_function=1, T_ave=None):
    tau = np.asfortranarray(tau, np.float64)
    n, L = tau.shape
    out = np.zeros(n)
    T_lev = np.ascontiguousarray(T_lev, np.float64)
    ta = None if T_ave is None else np.ascontiguousarray(T_ave, np.float64)
    rc = lib().orc_schwarzschild_rte(nu0, dnu, n, _ptr(tau), L, _ptr(T_lev), T_sfc, mu, eps,
                                     source_function, _ptr(ta), _ptr(out))
    if rc != 0:
        raise ValueError(f"orc_schwarzschild_rte error {rc}")
    return out


def ils_kernel(dnu, opd, fwhm, apodization=0, n_pts=2048):
    nh = lib().orc_ils_half(dnu)
    out = np.zeros(2 * nh + 1)
    rc = lib().orc_ils_kernel(dnu, opd, fwhm, apodization, n_pts, _ptr(out))
    if rc != 0:
        raise ValueError("bad apodization")
    return (np.arange(-nh, nh + 1) * dnu), out


def apply_ils_at(spec, dnu, kern, idx):
    spec = np.ascontiguousarray(spec, np.float64)
    kern = np.ascontiguousarray(kern, np.float64)
    idx = np.ascontiguousarray(idx, np.int64)
    out = np.zeros(len(idx))
    lib().orc_apply_ils_at(_ptr(spec), len(spec), dnu, _ptr(kern), len(kern), _ptr(idx), len(idx),
                           _ptr(out))
    return out


SPECIES_ORDER = None


class Problem:
    """Holds ctypes structures for one profile + line lists (keeps buffers alive)."""

    def __init__(self, prof, linelists):
        sp = [int(s) for s in prof.vmr.keys()]
        self.species = np.array(sp, np.int32)
        self.vmr = np.ascontiguousarray(np.stack([prof.vmr[s] for s in prof.vmr.keys()]), np.float64)
        self.lls = (OrcLinelist * len(sp))()
        self._keep = []
        for i, s in enumerate(prof.vmr.keys()):
            ll = linelists.get(s)
            if ll is not None and len(ll):
                self._keep.append(ll)
                self.lls[i] = make_linelist(ll)
            else:
                self.lls[i] = OrcLinelist(0)
        self.p = np.ascontiguousarray(prof.pressure)
        self.T = np.ascontiguousarray(prof.temperature)
        self.prof = OrcProfile(len(self.p), _ptr(self.p), _ptr(self.T), len(sp), _ptr(self.species),
                               _ptr(self.vmr), C.cast(self.lls, P))


def options(internal_dnu=0.001, high_res_factor=0, cutoff=25.0, lm_cutoff=None, dptmn=1e-6,
            x_far=122.0, with_ils=True, apodization=0, line_mixing=0, T_method=0,
            source_function=1, T_sfc=0.0, eps_sfc=1.0, mu=1.0, ils_direct_full=False, continua=0):
    """continua: bit mask (1 h2o, 2 co2, 4 co2_cia, 8 n2, 16 o2); 0 = no continuum."""
    return OrcOptions(internal_dnu, high_res_factor, cutoff, cutoff if lm_cutoff is None else lm_cutoff,
                      dptmn, x_far, int(with_ils), apodization, line_mixing, T_method,
                      source_function, T_sfc, eps_sfc, mu, int(ils_direct_full), int(continua))


# ---- continuum (src/CrossSections/continuum.jl) ------------------------------------
def _declare_cont():
    L = lib()
    if getattr(L, "_cont_declared", False):
        return L
    for nm in ("orc_h2o_continuum", "orc_co2_continuum"):
        getattr(L, nm).argtypes = [f64, f64, f64, f64]
        getattr(L, nm).restype = f64
    L.orc_radfn.argtypes = [f64, f64]
    L.orc_radfn.restype = f64
    L.orc_cia_set.argtypes = [i32, i32, P, P, P, P]
    L.orc_cia_set.restype = i32
    L.orc_cia_sigma.argtypes = [i32, f64, f64]
    L.orc_cia_sigma.restype = f64
    L.orc_cia_k.argtypes = [i32, f64, f64, f64, f64]
    L.orc_cia_k.restype = f64
    L.orc_dp_to_dz_cm.argtypes = [f64, f64, f64]
    L.orc_dp_to_dz_cm.restype = f64
    L.orc_continuum_layer.argtypes = [f64, f64, i64, i32, f64, f64, f64, f64, f64, P]
    L.orc_continuum_layer.restype = None
    L._cont_declared = True
    return L


def h2o_continuum(nu, vmr, p_hpa, T):
    L = _declare_cont()
    return np.array([L.orc_h2o_continuum(float(x), vmr, p_hpa, T) for x in np.atleast_1d(nu)])


def co2_continuum(nu, vmr, p_hpa, T):
    L = _declare_cont()
    return np.array([L.orc_co2_continuum(float(x), vmr, p_hpa, T) for x in np.atleast_1d(nu)])


def cia_k(pair, nu, vmr, p_hpa, T):
    L = _declare_cont()
    return np.array([L.orc_cia_k(pair, float(x), vmr, p_hpa, T) for x in np.atleast_1d(nu)])


def cia_set(pair, blocks):
    """blocks: list of (T, nu[], sigma[]); [] removes the table."""
    L = _declare_cont()
    T = np.array([float(b[0]) for b in blocks] or [0.0])
    off = np.zeros(len(blocks) + 1, np.int64)
    for i, b in enumerate(blocks):
        off[i + 1] = off[i] + len(b[1])
    nu = np.concatenate([np.asarray(b[1], np.float64) for b in blocks]) if blocks else np.zeros(1)
    sg = np.concatenate([np.asarray(b[2], np.float64) for b in blocks]) if blocks else np.zeros(1)
    if L.orc_cia_set(pair, len(blocks), _ptr(T), _ptr(off), _ptr(nu), _ptr(sg)) != 0:
        raise ValueError("orc_cia_set failed")


def continuum_layer(nu0, dnu, n, continua, vmr_h2o, vmr_co2, p_mid, T_mid, dp):
    L = _declare_cont()
    out = np.zeros(n)
    L.orc_continuum_layer(nu0, dnu, n, continua, vmr_h2o, vmr_co2, p_mid, T_mid, dp, _ptr(out))
    return out


def forward(prof, linelists, sounder, opts=None, layer_mask=None, want_tau=False, **kw):
    """FP64 forward model; returns (R, BT, info[, tau])."""
    opts = opts or options(**kw)
    pb = Problem(prof, linelists)
    snd = OrcSounder(sounder.nu_min, sounder.nu_max, sounder.dnu, sounder.opd_max, sounder.fwhm_gauss)
    info = OrcInfo()
    lib().orc_forward_dims(C.byref(snd), C.byref(opts), C.byref(info))
    R = np.zeros(info.n_ch); BT = np.zeros(info.n_ch)
    tau = np.zeros((prof.n_levels - 1, info.n_hi)) if want_tau else None
    m = None if layer_mask is None else np.ascontiguousarray(layer_mask, np.uint8)
    rc = lib().orc_forward(C.byref(pb.prof), C.byref(snd), C.byref(opts), _ptr(R), _ptr(BT),
                           _ptr(tau), _ptr(m), C.byref(info))
    if rc != 0:
        raise ValueError(f"orc_forward error {rc}")
    return (R, BT, info, tau) if want_tau else (R, BT, info)


def grid_nodes(nu0, dnu, n):
    """The n nodes of the Julia range nu0:dnu:... (wavenumber_grid.jl:23-27)."""
    L = lib()
    return np.array([L.orc_grid_node(nu0, dnu, i) for i in range(n)])


def forward_lm(prof, linelists, sounder, lm_species, lm_fn, opts=None, **kw):
    """FP64 forward model with a relaxation-matrix line-mixing perturbation:
    lm_fn(nu_nodes, T, p_atm) is the perturbation of species lm_species at one
    layer's T_cg, p_cg (the caller's restatement); returns (R, BT)."""
    opts = opts or options(**kw)
    pb = Problem(prof, linelists)
    snd = OrcSounder(sounder.nu_min, sounder.nu_max, sounder.dnu, sounder.opd_max, sounder.fwhm_gauss)
    info = OrcInfo()
    lib().orc_forward_dims(C.byref(snd), C.byref(opts), C.byref(info))
    nodes = grid_nodes(info.nu_lo, info.dnu, info.n_hi)
    lp = layer_properties(prof, opts.T_method)
    sp = [k for k in prof.vmr.keys() if int(k) == int(lm_species)]
    L = prof.n_levels - 1
    add = np.zeros((L, info.n_hi))
    if sp:
        s = sp[0]
        for k in range(L):
            if lp["vmr_cg"][s][k] != 0.0:
                add[k] = lm_fn(nodes, lp["T_cg"][s][k], lp["p_cg"][s][k] / 1013.25)
    R = np.zeros(info.n_ch)
    BT = np.zeros(info.n_ch)
    rc = lib().orc_forward_lm(C.byref(pb.prof), C.byref(snd), C.byref(opts), int(lm_species), _ptr(add),
                              _ptr(R), _ptr(BT), C.byref(info))
    if rc != 0:
        raise ValueError(f"orc_forward_lm error {rc}")
    return R, BT


def layer_properties(prof, T_method=0):
    pb = Problem(prof, {})
    L = prof.n_levels - 1
    ns = len(pb.species)
    out = {k: np.zeros(L) for k in ("p_mid", "T_mid", "dp", "T_ave")}
    vc = np.zeros((ns, L)); pc = np.zeros((ns, L)); tc = np.zeros((ns, L))
    lib().orc_layer_properties(C.byref(pb.prof), T_method, _ptr(out["p_mid"]), _ptr(out["T_mid"]),
                               _ptr(out["dp"]), _ptr(vc), _ptr(pc), _ptr(tc), _ptr(out["T_ave"]))
    keys = list(prof.vmr.keys())
    out["vmr_cg"] = {k: vc[i] for i, k in enumerate(keys)}
    out["p_cg"] = {k: pc[i] for i, k in enumerate(keys)}
    out["T_cg"] = {k: tc[i] for i, k in enumerate(keys)}
    return out


# ---- Jacobians -------------------------------------------------------------------
class OrcStateSpec(C.Structure):
    _fields_ = [("include_temperature", i32), ("n_retrieved", i32), ("retrieved", P),
                ("column_scale", P), ("include_tsfc", i32), ("include_emissivity", i32),
                ("log_vmr", i32), ("vmr_coupling", i32), ("observable", i32)]


def _declare_jac():
    L = lib()
    if getattr(L, "_jac_declared", False):
        return L
    L.orc_species_grad.argtypes = [C.POINTER(OrcLinelist), i32, f64, f64, i64, f64, f64, f64, f64,
                                   f64, f64, P, P, P, P]
    L.orc_rte_jacobian.argtypes = [f64, f64, i64, P, i32, P, P, f64, f64, f64, i32, P, P, P, P, P]
    L.orc_rte_jacobian.restype = i32
    L.orc_cg_column_vmr_grad.argtypes = [f64, f64, f64, f64, C.POINTER(f64), C.POINTER(f64)]
    L.orc_cg_mass_frac.argtypes = [f64, f64]
    L.orc_cg_mass_frac.restype = f64
    L.orc_dbt_dr.argtypes = [f64, f64]
    L.orc_dbt_dr.restype = f64
    L.orc_state_size.argtypes = [C.POINTER(OrcStateSpec), i32]
    L.orc_state_size.restype = i32
    L.orc_jacobian.argtypes = [C.POINTER(OrcProfile), C.POINTER(OrcSounder), C.POINTER(OrcOptions),
                               C.POINTER(OrcStateSpec), P, P, P]
    L.orc_jacobian.restype = i32
    L._jac_declared = True
    return L


def species_grad(ll, nu0, dnu, n, T, p_atm, vmr_self=0.0, line_mixing=0, cutoff=25.0,
                 lm_cutoff=25.0, x_far=122.0):
    L = _declare_jac()
    out = [np.zeros(n) for _ in range(4)]
    s = make_linelist(ll)
    L.orc_species_grad(C.byref(s), line_mixing, nu0, dnu, n, T, p_atm, vmr_self, cutoff, lm_cutoff,
                       x_far, *[_ptr(a) for a in out])
    return out


def rte_jacobian(nu0, dnu, tau, T_lev, p_lev, T_sfc, mu=1.0, eps=1.0, source_function=1):
    L = _declare_jac()
    tau = np.asfortranarray(tau, np.float64)
    n, Ln = tau.shape
    I = np.zeros(n); dtau = np.zeros((Ln, n)); dT = np.zeros((Ln + 1, n))
    dts = np.zeros(n); de = np.zeros(n)
    T_lev = np.ascontiguousarray(T_lev, np.float64)
    pl = None if p_lev is None else np.ascontiguousarray(p_lev, np.float64)
    rc = L.orc_rte_jacobian(nu0, dnu, n, _ptr(tau), Ln, _ptr(T_lev), _ptr(pl), T_sfc, mu, eps,
                            source_function, _ptr(I), _ptr(dtau), _ptr(dT), _ptr(dts), _ptr(de))
    if rc != 0:
        raise ValueError(f"orc_rte_jacobian error {rc}")
    return I, dtau.T, dT.T, dts, de


def jacobian(prof, linelists, sounder, retrieved=None, include_temperature=True,
             include_tsfc=True, include_emissivity=True, log_vmr=True, vmr_coupling=True,
             observable="bt", column_scale=None, **kw):
    """Analytic Jacobian (FP64 oracle): returns (K [n_ch x n_state], y0, x0)."""
    L = _declare_jac()
    retrieved = list(retrieved or [])
    r = np.array([int(s) for s in retrieved] or [0], np.int32)
    cs = None if column_scale is None else np.array([int(bool(c)) for c in column_scale], np.int32)
    spec = OrcStateSpec(int(include_temperature), len(retrieved), _ptr(r), _ptr(cs), int(include_tsfc),
                        int(include_emissivity), int(log_vmr), int(vmr_coupling),
                        0 if observable == "bt" else 1)
    opts = options(dptmn=0.0, **kw)
    pb = Problem(prof, linelists)
    snd = OrcSounder(sounder.nu_min, sounder.nu_max, sounder.dnu, sounder.opd_max, sounder.fwhm_gauss)
    info = OrcInfo()
    L.orc_forward_dims(C.byref(snd), C.byref(opts), C.byref(info))
    ns = L.orc_state_size(C.byref(spec), prof.n_levels)
    K = np.zeros((ns, info.n_ch)); y0 = np.zeros(info.n_ch); x0 = np.zeros(ns)
    rc = L.orc_jacobian(C.byref(pb.prof), C.byref(snd), C.byref(opts), C.byref(spec), _ptr(K),
                        _ptr(y0), _ptr(x0))
    if rc != 0:
        raise ValueError(f"orc_jacobian error {rc}")
    return K.T.copy(), y0, x0
