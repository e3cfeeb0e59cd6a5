"""Host mirror of the reference operators on top of libirs (include/irs.h).

Names, argument meaning, defaults and error behaviour follow the reference:
  forward_model                   src/Sensor/sounder.jl:233-360
  compute_voigt_cross_sections    src/CrossSections/voigt.jl:375-462 (+ DPTMIN :333-355)
  compute_voigt_lm_cross_sections src/CrossSections/line_mixing.jl:1351-1369
  schwarzschild_rte               src/Solver/schwarzschild.jl:130-209,228-274
  ils_kernel / apply_ils          src/Sensor/ils.jl:80-153
  h2o_continuum, co2_continuum, co2_cia, n2_cia, o2_cia
                                  src/CrossSections/continuum.jl:163-327
  planck_radiance / brightness_temperature   src/Solver/planck.jl:11-57
  analytic_jacobian               src/Estimation/vmr_jacobian.jl:152-505
Every numeric call runs in the HIP library; there is no CPU fallback.  A
missing library or GPU raises IRSError.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import weakref
from typing import Dict, Optional

import numpy as np

from .relmat import HITRANRelmatData, VPWLineMixing, VPYLineMixing, linelist_isotopes
from .relmat import band_modes_for as relmat_band_modes
from .types import (AtmosphericProfile, GasSpecies, HITRANLinelist, Jacobian, Sounder,
                    StateVectorSpec, ViewingGeometry, WavenumberGrid, IASIInstrument,
                    nadir_geometry, sounder_grid)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(HERE, "..", "csrc", "libirs.so"))

APODIZATION = {"gaussian": 0, "norton_beer_weak": 1, "norton_beer_medium": 2, "norton_beer_strong": 3}
SOURCE_FUNCTION = {"toon": 0, "cim": 1}
T_METHOD = {"logp_at_pcg": 0, "mass_weighted": 1}
CONTINUA = {"h2o": 1, "co2": 2, "co2_cia": 4, "n2": 8, "o2": 16}
DEFAULT_CONTINUA = ("h2o", "co2", "co2_cia", "n2", "o2")      # sounder.jl:246
CIA_PAIRS = {"CO2-CO2": 0, "N2-N2": 1, "O2-O2": 2}

IRS_OK, IRS_EINVAL, IRS_EHIP, IRS_ENOMEM, IRS_ENOTFOUND, IRS_ENODEV = 0, -1, -2, -3, -4, -5


class IRSError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


i64, f64, i32 = C.c_int64, C.c_double, C.c_int
P = C.c_void_p


class IrsLinelist(C.Structure):
    _fields_ = [("n", i64), ("mol", P), ("iso", P), ("nu", P), ("S", P), ("gamma_air", P),
                ("gamma_self", P), ("elower", P), ("n_air", P), ("delta_air", P), ("y_lm", P)]


class IrsSounder(C.Structure):
    _fields_ = [("nu_min", f64), ("nu_max", f64), ("dnu", f64), ("opd_max", f64),
                ("fwhm_gauss", f64), ("n_channels", i64)]


class IrsProfile(C.Structure):
    _fields_ = [("n_levels", i32), ("p_hpa", P), ("T", P), ("n_species", i32), ("species", P),
                ("linelist", P), ("vmr", P), ("T_sfc", f64), ("eps_sfc", f64), ("mu", f64)]


class IrsOptions(C.Structure):
    _fields_ = [("internal_dnu", f64), ("high_res_factor", i32), ("cutoff", f64),
                ("lm_cutoff", f64), ("dptmn", f64), ("x_far", f64), ("with_ils", i32),
                ("apodization", i32), ("line_mixing", i32), ("T_method", i32),
                ("source_function", i32), ("continua", i32), ("relmat", i32), ("lm_species", i32),
                ("relmat_band_mode", P), ("keep_threshold", f64)]


class IrsRelmat(C.Structure):
    _fields_ = [("n_bands", i32), ("li", P), ("lf", P), ("isot", P), ("molecule", P), ("line_off", P),
                ("Ji", P), ("branch", P), ("nu", P), ("gV_air", P), ("E_lower", P), ("Dipo0", P),
                ("PopuT0", P), ("DipoT", P), ("n_air", P), ("shift", P), ("n_tables", i32),
                ("table_lli", P), ("table_llf", P), ("entry_off", P), ("entry_sub", P), ("entry_ji", P),
                ("entry_jip", P), ("entry_w0", P), ("entry_b0", P)]


def relmat_struct(data):
    """ctypes view of a HITRANRelmatData (irs_relmat) and the arrays it points into."""
    bands = data.bands
    keys = sorted(data.wtfit.keys())
    cat = (lambda xs, dt: np.ascontiguousarray(np.concatenate(xs) if xs else np.zeros(0), dt))
    arr = dict(
        li=np.array([b.li for b in bands], np.int8), lf=np.array([b.lf for b in bands], np.int8),
        isot=np.array([b.isot for b in bands], np.int8), molecule=np.array([b.molecule for b in bands], np.int8),
        line_off=np.concatenate([[0], np.cumsum([b.n for b in bands])]).astype(np.int64),
        Ji=cat([b.Ji for b in bands], np.int16), branch=cat([b.branch for b in bands], np.int8),
        nu=cat([b.nu for b in bands], np.float64), gV_air=cat([b.gV_air for b in bands], np.float64),
        E_lower=cat([b.E_lower for b in bands], np.float64), Dipo0=cat([b.Dipo0 for b in bands], np.float64),
        PopuT0=cat([b.PopuT0 for b in bands], np.float64), DipoT=cat([b.DipoT for b in bands], np.float64),
        n_air=cat([b.n_air for b in bands], np.float32), shift=cat([b.shift for b in bands], np.float32))
    ents = [list(data.wtfit[k].entries()) for k in keys]
    flat = [e for es in ents for e in es]
    arr.update(table_lli=np.array([k[0] for k in keys], np.int8), table_llf=np.array([k[1] for k in keys], np.int8),
               entry_off=np.concatenate([[0], np.cumsum([len(e) for e in ents])]).astype(np.int64),
               entry_sub=np.array([e[0] for e in flat], np.int8), entry_ji=np.array([e[1] for e in flat], np.int32),
               entry_jip=np.array([e[2] for e in flat], np.int32),
               entry_w0=np.array([e[3] for e in flat], np.float64), entry_b0=np.array([e[4] for e in flat], np.float64))
    arr = {k: np.ascontiguousarray(v) for k, v in arr.items()}
    st = IrsRelmat(len(bands), *[_p(arr[k]) for k in ("li", "lf", "isot", "molecule", "line_off", "Ji", "branch",
                                                        "nu", "gV_air", "E_lower", "Dipo0", "PopuT0", "DipoT",
                                                        "n_air", "shift")],
                   len(keys), *[_p(arr[k]) for k in ("table_lli", "table_llf", "entry_off", "entry_sub", "entry_ji",
                                                      "entry_jip", "entry_w0", "entry_b0")])
    return st, arr


class IrsStateSpec(C.Structure):
    _fields_ = [("include_temperature", i32), ("n_retrieved", i32), ("retrieved", P),
                ("column_scale", P), ("include_tsfc", i32), ("include_emissivity", i32),
                ("log_vmr", i32), ("vmr_coupling", i32), ("observable", i32)]


class IrsStats(C.Structure):
    _fields_ = [("ms_total", f64), ("ms_prep", f64), ("ms_linesum", f64), ("ms_rte", f64),
                ("ms_ils", f64), ("n_hi", i64), ("n_layers", i64), ("linesum_launches", i64),
                ("active_lines", i64), ("active_pairs", i64), ("dispersive_pairs", i64),
                ("core_pairs_bound", i64), ("ms_device", f64), ("graph_launches", i64),
                ("tile_points", i32), ("persistent", i32), ("autotuned", i32), ("reserved1", i32),
                ("fine_pairs", i64), ("coarse_pairs", i64), ("fine_pairs_lm", i64),
                ("coarse_pairs_lm", i64), ("expanded_line_slices", i64), ("ms_continuum", f64)]


class IrsSchedule(C.Structure):
    _fields_ = [("tile_points", i32), ("persistent", i32), ("use_graph", i32), ("autotune", i32),
                ("far_wing_cm", f64), ("fp64_every_pair", i32), ("reserved", i32)]


EXPORTS = ("irs_version", "irs_device_info", "irs_create", "irs_destroy", "irs_last_error",
           "irs_default_options", "irs_last_stats", "irs_linelist_upload", "irs_linelist_free",
           "irs_xsec_layer", "irs_faddeeva32", "irs_rte", "irs_ils_half", "irs_ils_kernel", "irs_ils_channels",
           "irs_brightness_temperature", "irs_forward_dims", "irs_forward", "irs_state_size",
           "irs_jacobian", "irs_set_schedule", "irs_get_schedule", "irs_channel_od",
           "irs_layer_properties", "irs_mlp_create", "irs_mlp_destroy", "irs_mlp_last_error",
           "irs_mlp_n_params", "irs_mlp_input_pad", "irs_mlp_get_params", "irs_mlp_set_params",
           "irs_mlp_forward", "irs_mlp_loss_grad", "irs_mlp_set_data", "irs_mlp_train",
           "irs_mlp_input_jacobian", "irs_emu_create", "irs_emu_destroy", "irs_emu_last_error",
           "irs_emu_forward", "irs_channel_transmittance", "irs_emu_jacobian", "irs_cia_set",
           "irs_cia_load_file", "irs_continuum_layer", "irs_xsec_layer_noalloc",
           "irs_channel_od_moments", "irs_par_parse", "irs_par_load", "irs_relmat_upload",
           "irs_relmat_free", "irs_lm_perturbation", "irs_relmat_terms", "irs_emu_set_context",
           "irs_forward_emulated", "irs_emu_set_correction", "irs_emu_load")

_lib = None


def load_library(path: str = LIB_PATH):
    """Load libirs.so and declare the C ABI.  Raises IRSError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise IRSError(IRS_ENODEV, f"libirs.so not built at {path} (run __graft_entry__.build())")
    L = C.CDLL(path)
    L.irs_version.restype = i32
    L.irs_device_info.argtypes = [i32, C.c_char_p, i32]
    L.irs_create.argtypes = [i32, C.POINTER(P)]
    L.irs_destroy.argtypes = [P]
    L.irs_destroy.restype = None
    L.irs_last_error.argtypes = [P]
    L.irs_last_error.restype = C.c_char_p
    L.irs_default_options.argtypes = [C.POINTER(IrsOptions)]
    L.irs_default_options.restype = None
    L.irs_last_stats.argtypes = [P, C.POINTER(IrsStats)]
    L.irs_linelist_upload.argtypes = [P, C.POINTER(IrsLinelist), C.POINTER(C.c_int32)]
    L.irs_linelist_free.argtypes = [P, C.c_int32]
    L.irs_xsec_layer.argtypes = [P, C.c_int32, f64, f64, i64, f64, f64, f64, f64, f64, f64, f64,
                                 f64, i32, P]
    L.irs_xsec_layer_noalloc.argtypes = L.irs_xsec_layer.argtypes
    L.irs_rte.argtypes = [P, f64, f64, i64, P, i32, P, f64, f64, f64, i32, P, P]
    L.irs_faddeeva32.argtypes = [P, P, P, i64, i32, P, P]
    L.irs_ils_half.argtypes = [f64, C.POINTER(i64)]
    L.irs_ils_kernel.argtypes = [P, f64, f64, f64, i32, i32, P]
    L.irs_ils_channels.argtypes = [P, P, f64, f64, i64, C.POINTER(IrsSounder), i32, P]
    L.irs_brightness_temperature.argtypes = [P, P, i64, P]
    L.irs_forward_dims.argtypes = [C.POINTER(IrsSounder), C.POINTER(IrsOptions), C.POINTER(f64),
                                   C.POINTER(f64), C.POINTER(i64), C.POINTER(i64)]
    L.irs_forward.argtypes = [P, C.POINTER(IrsProfile), i32, C.POINTER(IrsSounder),
                              C.POINTER(IrsOptions), P, P]
    L.irs_state_size.argtypes = [C.POINTER(IrsStateSpec), i32]
    L.irs_set_schedule.argtypes = [P, C.POINTER(IrsSchedule)]
    L.irs_channel_od.argtypes = [P, C.POINTER(IrsProfile), i32, C.POINTER(IrsSounder), C.POINTER(IrsOptions),
                                 f64, P, P]
    L.irs_channel_od_moments.argtypes = L.irs_channel_od.argtypes + [P]
    L.irs_get_schedule.argtypes = [P, C.POINTER(IrsSchedule)]
    L.irs_jacobian.argtypes = [P, C.POINTER(IrsProfile), C.POINTER(IrsSounder), C.POINTER(IrsOptions),
                               C.POINTER(IrsStateSpec), P, P, P]
    L.irs_channel_transmittance.argtypes = [P, C.POINTER(IrsProfile), i32, C.POINTER(IrsSounder),
                                            C.POINTER(IrsOptions), P, P]
    L.irs_layer_properties.argtypes = [C.POINTER(IrsProfile), i32, P, P, P, P, P, P]
    L.irs_mlp_create.argtypes = [i32, i32, i32, i32, i32, C.c_uint64, C.POINTER(P)]
    L.irs_mlp_destroy.argtypes = [P]
    L.irs_mlp_destroy.restype = None
    L.irs_mlp_last_error.argtypes = [P]
    L.irs_mlp_last_error.restype = C.c_char_p
    L.irs_mlp_n_params.argtypes = [P]
    L.irs_mlp_n_params.restype = i64
    L.irs_mlp_input_pad.argtypes = [P]
    L.irs_mlp_get_params.argtypes = [P, P]
    L.irs_mlp_set_params.argtypes = [P, P]
    L.irs_mlp_forward.argtypes = [P, P, i64, P]
    L.irs_mlp_loss_grad.argtypes = [P, P, P, i32, P, P]
    L.irs_mlp_set_data.argtypes = [P, P, P, i64]
    L.irs_mlp_train.argtypes = [P, i32, i32, C.c_float, C.c_uint64, P]
    L.irs_mlp_input_jacobian.argtypes = [P, P, i64, P]
    L.irs_emu_create.argtypes = [P, i32, i32, i32, i32, i32, P, P, P, C.c_float, P, P, P, P, C.POINTER(P)]
    L.irs_emu_destroy.argtypes = [P]
    L.irs_emu_destroy.restype = None
    L.irs_emu_last_error.argtypes = [P]
    L.irs_emu_last_error.restype = C.c_char_p
    L.irs_emu_forward.argtypes = [P, P, P, P, P, i32, i32, P, P, P]
    L.irs_emu_jacobian.argtypes = [P, P, P, P, P, i32, P, P, P, P, P]
    L.irs_cia_set.argtypes = [P, i32, i32, P, P, P, P]
    L.irs_cia_load_file.argtypes = [P, i32, C.c_char_p]
    L.irs_continuum_layer.argtypes = [P, f64, f64, i64, i32, f64, f64, f64, f64, f64, f64, f64, P]
    L.irs_relmat_upload.argtypes = [P, C.POINTER(IrsRelmat), C.POINTER(C.c_int32)]
    L.irs_relmat_free.argtypes = [P, C.c_int32]
    L.irs_lm_perturbation.argtypes = [P, C.c_int32, P, f64, f64, i64, f64, f64, f64, f64, f64, P]
    L.irs_relmat_terms.argtypes = [C.POINTER(IrsRelmat), P, f64, f64, f64, i64, P, C.POINTER(i64)]
    L.irs_emu_set_context.argtypes = [P, i32, i32, P, P]
    L.irs_emu_set_correction.argtypes = [P, i32, P, P, P, P, f64, f64]
    L.irs_emu_load.argtypes = [C.c_char_p, i32, i32, C.POINTER(P), C.POINTER(P), C.POINTER(f64), i32, P, P,
                               C.POINTER(i32)]
    L.irs_forward_emulated.argtypes = [P, P, C.POINTER(IrsProfile), i32, C.POINTER(IrsSounder),
                                       C.POINTER(IrsOptions), i32, P, P, P, f64, P, P, P, P]
    for name in EXPORTS:
        if name not in ("irs_destroy", "irs_default_options", "irs_last_error", "irs_mlp_destroy",
                        "irs_mlp_last_error", "irs_mlp_n_params", "irs_emu_destroy", "irs_emu_last_error"):
            getattr(L, name).restype = getattr(L, name).restype or i32
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


def _sounder(s: Sounder) -> IrsSounder:
    return IrsSounder(s.nu_min, s.nu_max, s.dnu, s.opd_max, s.fwhm_gauss, s.n_channels)


def _apod(apodization) -> int:
    key = apodization.lstrip(":") if isinstance(apodization, str) else apodization
    if isinstance(key, int):
        return key
    if key not in APODIZATION:
        raise ValueError(f"Unknown apodization {apodization}; expected :gaussian or one of "
                         f"{[k for k in APODIZATION if k != 'gaussian']}")
    return APODIZATION[key]


class Context:
    """One device context (HIP stream + device buffers + uploaded line lists)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = P()
        rc = self.lib.irs_create(device, C.byref(h))
        if rc != IRS_OK:
            raise IRSError(rc, f"irs_create failed ({rc}): no usable gfx950 device {device}")
        self.h = h
        self._lists: Dict[int, tuple] = {}     # id(list) -> (handle, weakref, finalizer)

    def close(self):
        for _, _, fin in list(getattr(self, "_lists", {}).values()):
            fin.detach()
        self._lists = {}
        if getattr(self, "h", None):
            self.lib.irs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc != IRS_OK:
            msg = self.lib.irs_last_error(self.h).decode(errors="replace")
            if rc == IRS_EINVAL:
                raise ValueError(msg)
            raise IRSError(rc, msg)

    def stats(self) -> dict:
        s = IrsStats()
        self.check(self.lib.irs_last_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in IrsStats._fields_}

    def schedule(self) -> dict:
        sc = IrsSchedule()
        self.check(self.lib.irs_get_schedule(self.h, C.byref(sc)))
        return {k: getattr(sc, k) for k, _ in IrsSchedule._fields_}

    def set_schedule(self, tile_points: int = 4, persistent: bool = True, use_graph: bool = False,
                     autotune: bool = False, far_wing_cm: float = 1.5, fp64_every_pair: bool = False):
        """Line-sum schedule of forward_model: points per thread (4/8/16), the
        persistent heaviest-first work-queue kernel, hipGraph capture and
        replay of each profile group, autotuning on the first call of a shape,
        and the distance beyond which far wings are summed on a coarse grid
        and interpolated (0: every pair on the internal grid).
        fp64_every_pair runs the reference's every-pair sum in FP64 instead."""
        sc = IrsSchedule(int(tile_points), int(bool(persistent)), int(bool(use_graph)), int(bool(autotune)),
                         float(far_wing_cm), int(bool(fp64_every_pair)), 0)
        self.check(self.lib.irs_set_schedule(self.h, C.byref(sc)))

    def upload(self, ll: HITRANLinelist) -> int:
        """Device handle of a line list, uploaded on first use.  The cache holds
        the list weakly: when the list is garbage collected its device copy is
        freed (irs_linelist_free)."""
        key = id(ll)
        hit = self._lists.get(key)
        if hit is not None and hit[1]() is ll:
            return hit[0]
        s = IrsLinelist(len(ll), _p(ll.mol_id), _p(ll.iso_id), _p(ll.wavenumber), _p(ll.intensity),
                        _p(ll.air_broad), _p(ll.self_broad), _p(ll.lower_energy), _p(ll.temp_depend),
                        _p(ll.pressure_shift), _p(ll.y_lm))
        h = C.c_int32()
        self.check(self.lib.irs_linelist_upload(self.h, C.byref(s), C.byref(h)))
        fin = weakref.finalize(ll, Context._release_list, weakref.ref(self), key, h.value)
        self._lists[key] = (h.value, weakref.ref(ll), fin)
        return h.value

    @staticmethod
    def _release_list(ctx_ref, key, handle):
        ctx = ctx_ref()
        if ctx is None or not getattr(ctx, "h", None):
            return
        ent = ctx._lists.get(key)
        if ent is not None and ent[0] == handle:
            del ctx._lists[key]
        ctx.lib.irs_linelist_free(ctx.h, handle)

    def upload_relmat(self, data) -> int:
        """Handle of a relaxation matrix (HITRANRelmatData), uploaded on first use and
        released when the data object is garbage collected (irs_relmat_free)."""
        key = ("relmat", id(data))
        hit = self._lists.get(key)
        if hit is not None and hit[1]() is data:
            return hit[0]
        st, _keep = relmat_struct(data)
        h = C.c_int32()
        self.check(self.lib.irs_relmat_upload(self.h, C.byref(st), C.byref(h)))
        fin = weakref.finalize(data, Context._release_relmat, weakref.ref(self), key, h.value)
        self._lists[key] = (h.value, weakref.ref(data), fin)
        return h.value

    @staticmethod
    def _release_relmat(ctx_ref, key, handle):
        ctx = ctx_ref()
        if ctx is None or not getattr(ctx, "h", None):
            return
        ent = ctx._lists.get(key)
        if ent is not None and ent[0] == handle:
            del ctx._lists[key]
        ctx.lib.irs_relmat_free(ctx.h, handle)

    def set_cia(self, pair, blocks):
        """CIA table of one collision pair ("CO2-CO2", "N2-N2", "O2-O2"): blocks
        is a list of (T, nu[], sigma[]) (cm^5 molec^-2); an empty list removes
        it, so the pair adds zero (the reference's behaviour without the data
        file, src/CrossSections/continuum.jl:254-258)."""
        pr = CIA_PAIRS[pair] if isinstance(pair, str) else int(pair)
        T = np.array([float(b[0]) for b in blocks] or [0.0])
        nus = [np.ascontiguousarray(b[1], np.float64) for b in blocks]
        sig = [np.ascontiguousarray(b[2], np.float64) for b in blocks]
        off = np.zeros(len(blocks) + 1, np.int64)
        for i, x in enumerate(nus):
            off[i + 1] = off[i] + len(x)
        nu = np.concatenate(nus) if nus else np.zeros(1)
        sg = np.concatenate(sig) if sig else np.zeros(1)
        self.check(self.lib.irs_cia_set(self.h, pr, len(blocks), _p(T), _p(off), _p(nu), _p(sg)))

    def load_cia(self, pair, path: str):
        """Reads a HITRAN .cia file into the pair's table (continuum.jl:106-148)."""
        pr = CIA_PAIRS[pair] if isinstance(pair, str) else int(pair)
        self.check(self.lib.irs_cia_load_file(self.h, pr, os.fsencode(path)))


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(int(os.environ.get("IRS_DEVICE", "0")))
    return _default_ctx


def compute_voigt_cross_sections(nu_grid: WavenumberGrid, linelist: HITRANLinelist, T: float,
                                 p_atm: float, *, vmr_self: float = 0.0, cutoff: float = 25.0,
                                 x_far: float = 122.0, dptmn: float = 0.0, coef: float = 1.0,
                                 ctx: Optional[Context] = None) -> np.ndarray:
    """Voigt cross sections (cm^2/molec) on nu_grid; dptmn > 0 applies DPTMIN with coef."""
    ctx = ctx or default_context()
    h = ctx.upload(linelist)
    out = np.zeros(nu_grid.n)
    ctx.check(ctx.lib.irs_xsec_layer(ctx.h, h, nu_grid.nu_min, nu_grid.dnu, nu_grid.n, float(T),
                                     float(p_atm), float(vmr_self), float(coef), float(dptmn),
                                     float(cutoff), float(cutoff), float(x_far), 0, _p(out)))
    return out


def faddeeva32(x, y, *, rule: str = "core", ctx: Optional[Context] = None) -> np.ndarray:
    """w(x + iy) (complex, y >= 0) from the FP32 device routines of the line
    sums: rule "core" is the CORE-range rule, "pair" the per-pair and gradient
    rule.  Re w is faddeeva_voigt (voigt.jl:158-160)."""
    ctx = ctx or default_context()
    shape = np.broadcast(np.asarray(x), np.asarray(y)).shape
    x = np.ascontiguousarray(np.broadcast_to(np.asarray(x, np.float64), shape)).ravel()
    y = np.ascontiguousarray(np.broadcast_to(np.asarray(y, np.float64), shape)).ravel()
    if rule not in ("core", "pair"):
        raise ValueError("rule must be 'core' or 'pair'")
    re, im = np.zeros(len(x)), np.zeros(len(x))
    ctx.check(ctx.lib.irs_faddeeva32(ctx.h, _p(x), _p(y), len(x), 0 if rule == "core" else 1, _p(re), _p(im)))
    return (re + 1j * im).reshape(shape)


def compute_voigt_lm_cross_sections(nu_grid: WavenumberGrid, ll_co2: HITRANLinelist, *args,
                                     cutoff: float = 25.0, lm_cutoff: Optional[float] = None, x_far: float = 122.0,
                                     min_band_strength: float = 0.0, isotopes=None,
                                     ctx: Optional[Context] = None) -> np.ndarray:
    """max(max(sigma_voigt, 0) + sigma_disp, 0) (src/CrossSections/line_mixing.jl:1351-1369).

    compute_voigt_lm_cross_sections(nu_grid, ll_co2, relmat, T, p_atm; ...) takes the
    dispersive term from the relaxation matrix (Y(T) per band, S-file lines, the
    isotopologues of ll_co2 unless `isotopes` is given);
    compute_voigt_lm_cross_sections(nu_grid, ll_co2, T, p_atm; ...) from the per-line
    Y of ll_co2 (y_lm)."""
    ctx = ctx or default_context()
    lc = cutoff if lm_cutoff is None else lm_cutoff
    if args and isinstance(args[0], HITRANRelmatData):
        relmat, T, p_atm = args
        sv = compute_voigt_cross_sections(nu_grid, ll_co2, T, p_atm, cutoff=cutoff, x_far=x_far, ctx=ctx)
        iso = linelist_isotopes(ll_co2) if isotopes is None else set(isotopes)
        d = lm_perturbation(nu_grid, VPYLineMixing(relmat, min_band_strength), T, p_atm, cutoff=lc, x_far=x_far,
                            isotopes=iso, ctx=ctx)
        return np.maximum(sv + d, 0.0)
    T, p_atm = args
    h = ctx.upload(ll_co2)
    out = np.zeros(nu_grid.n)
    ctx.check(ctx.lib.irs_xsec_layer(ctx.h, h, nu_grid.nu_min, nu_grid.dnu, nu_grid.n, float(T),
                                     float(p_atm), 0.0, 1.0, 0.0, float(cutoff), float(lc),
                                     float(x_far), 1, _p(out)))
    return out


def compute_voigt_vpw_cross_sections(nu_grid: WavenumberGrid, ll_co2: HITRANLinelist, relmat, T: float,
                                     p_atm: float, *, whitelist, cutoff: float = 25.0,
                                     lm_cutoff: Optional[float] = None, keep_threshold: float = 1e-4,
                                     min_band_strength: float = 0.0, isotopes=None, x_far: float = 122.0,
                                     ctx: Optional[Context] = None) -> np.ndarray:
    """max(sigma_voigt + dsigma, 0) with the eigenmode (VP_W) perturbation of the
    whitelisted bands and the VP_Y term of the others (line_mixing.jl:1093-1118)."""
    ctx = ctx or default_context()
    sv = compute_voigt_cross_sections(nu_grid, ll_co2, T, p_atm, cutoff=cutoff, x_far=x_far, ctx=ctx)
    iso = linelist_isotopes(ll_co2) if isotopes is None else set(isotopes)
    d = lm_perturbation(nu_grid, VPWLineMixing(relmat, set(whitelist), min_band_strength), T, p_atm,
                        cutoff=cutoff if lm_cutoff is None else lm_cutoff, x_far=x_far, isotopes=iso,
                        keep_threshold=keep_threshold, ctx=ctx)
    return np.maximum(sv + d, 0.0)


def lm_perturbation(nu_grid: WavenumberGrid, line_mixing, T: float, p_atm: float, *, cutoff: float = 25.0,
                    x_far: float = 122.0, isotopes=None, keep_threshold: float = 1e-4,
                    ctx: Optional[Context] = None) -> np.ndarray:
    """The additive line-mixing term dsigma of a VPYLineMixing / VPWLineMixing
    (_lm_perturbation, line_mixing.jl:283-303), evaluated on the device."""
    ctx = ctx or default_context()
    h = ctx.upload_relmat(line_mixing.data)
    modes = np.ascontiguousarray(relmat_band_modes(line_mixing, None if isotopes is None else set(isotopes)),
                                 np.uint8)
    out = np.zeros(nu_grid.n)
    ctx.check(ctx.lib.irs_lm_perturbation(ctx.h, h, _p(modes) if len(modes) else None, nu_grid.nu_min,
                                          nu_grid.dnu, nu_grid.n, float(T), float(p_atm), float(cutoff),
                                          float(x_far), float(keep_threshold), _p(out)))
    return out


def relmat_terms(line_mixing, T: float, p_atm: float, *, isotopes=None, keep_threshold: float = 1e-4) -> np.ndarray:
    """Host-side terms [n, 5] = (c, f, y, a, b) the device sums for one layer (no GPU)."""
    lib = load_library()
    st, _keep = relmat_struct(line_mixing.data)
    modes = np.ascontiguousarray(relmat_band_modes(line_mixing, None if isotopes is None else set(isotopes)),
                                 np.uint8)
    n = C.c_int64()
    mp = _p(modes) if len(modes) else None
    if lib.irs_relmat_terms(C.byref(st), mp, float(T), float(p_atm), float(keep_threshold), 0, None,
                            C.byref(n)) != IRS_OK:
        raise ValueError("irs_relmat_terms: bad arguments")
    out = np.zeros((max(n.value, 1), 5))
    if lib.irs_relmat_terms(C.byref(st), mp, float(T), float(p_atm), float(keep_threshold), n.value, _p(out),
                            C.byref(n)) != IRS_OK:
        raise ValueError("irs_relmat_terms: bad arguments")
    return out[:n.value]


def _continuum(bit, nu_grid, vmr, p_hPa, T, ctx):
    ctx = ctx or default_context()
    out = np.zeros(nu_grid.n)
    v = [0.0, 0.0, 0.0, 0.0]
    v[{1: 0, 2: 1, 4: 1, 8: 2, 16: 3}[bit]] = float(vmr)
    ctx.check(ctx.lib.irs_continuum_layer(ctx.h, nu_grid.nu_min, nu_grid.dnu, nu_grid.n, bit, v[0], v[1], v[2],
                                          v[3], float(p_hPa), float(T), 1.0, _p(out)))
    return out


def h2o_continuum(nu_grid: WavenumberGrid, vmr_h2o: float, p_hPa: float, T: float, *,
                  ctx: Optional[Context] = None) -> np.ndarray:
    """MT-CKD 4.3 H2O self + foreign continuum absorption coefficient (cm^-1)
    (continuum.jl:163-190), evaluated on the device."""
    return _continuum(1, nu_grid, vmr_h2o, p_hPa, T, ctx)


def co2_continuum(nu_grid: WavenumberGrid, vmr_co2: float, p_hPa: float, T: float, *,
                  ctx: Optional[Context] = None) -> np.ndarray:
    """MT-CKD CO2 line-coupling continuum (cm^-1) (continuum.jl:209-232)."""
    return _continuum(2, nu_grid, vmr_co2, p_hPa, T, ctx)


def co2_cia(nu_grid: WavenumberGrid, vmr_co2: float, p_hPa: float, T: float, *,
            ctx: Optional[Context] = None) -> np.ndarray:
    """CO2-CO2 CIA scaled by n_CO2 n_air (cm^-1); zeros without a loaded table
    (continuum.jl:249-270)."""
    return _continuum(4, nu_grid, vmr_co2, p_hPa, T, ctx)


def n2_cia(nu_grid: WavenumberGrid, vmr_n2: float, p_hPa: float, T: float, *,
           ctx: Optional[Context] = None) -> np.ndarray:
    """N2-N2 CIA (cm^-1) (continuum.jl:279-298)."""
    return _continuum(8, nu_grid, vmr_n2, p_hPa, T, ctx)


def o2_cia(nu_grid: WavenumberGrid, vmr_o2: float, p_hPa: float, T: float, *,
           ctx: Optional[Context] = None) -> np.ndarray:
    """O2-O2 CIA (cm^-1) (continuum.jl:307-327)."""
    return _continuum(16, nu_grid, vmr_o2, p_hPa, T, ctx)


def _continua_mask(apply_continuum, continua) -> int:
    if not apply_continuum:
        return 0
    m = 0
    for c in continua:
        key = c.lstrip(":") if isinstance(c, str) else c
        if key not in CONTINUA:
            raise ValueError(f"unknown continuum :{key}; expected one of {list(CONTINUA)}")
        m |= CONTINUA[key]
    return m


def schwarzschild_rte(nu_grid: WavenumberGrid, tau_layers: np.ndarray, T_levels, T_sfc: float, *,
                      mu: float = 1.0, eps_sfc: float = 1.0, source_function: str = "toon",
                      T_ave=None, ctx: Optional[Context] = None) -> np.ndarray:
    """TOA radiance; tau_layers is (n_nu, n_layers) (schwarzschild.jl:130-209)."""
    tau = np.asfortranarray(tau_layers, np.float64)
    n_nu, L = tau.shape
    if n_nu != nu_grid.n:
        raise ValueError(f"τ_layers rows {n_nu} ≠ ν_grid.n {nu_grid.n}")
    T_levels = np.ascontiguousarray(T_levels, np.float64)
    if len(T_levels) != L + 1:
        raise ValueError(f"T_levels must have n_layers+1 = {L + 1} entries")
    if not (0.0 < mu <= 1.0):
        raise ValueError("μ must be in (0, 1]")
    sf = source_function.lstrip(":")
    if sf not in SOURCE_FUNCTION:
        raise ValueError(f"source_function must be :toon or :cim, got :{sf}")
    ta = None
    if sf == "cim":
        if T_ave is None:
            raise ValueError(":cim source_function requires T_ave keyword")
        ta = np.ascontiguousarray(T_ave, np.float64)
        if len(ta) != L:
            raise ValueError(f"T_ave must have n_layers = {L} entries, got {len(ta)}")
    ctx = ctx or default_context()
    out = np.zeros(n_nu)
    ctx.check(ctx.lib.irs_rte(ctx.h, nu_grid.nu_min, nu_grid.dnu, n_nu, _p(tau), L, _p(T_levels),
                              float(T_sfc), float(mu), float(eps_sfc), SOURCE_FUNCTION[sf], _p(ta),
                              _p(out)))
    return out


def ils_kernel(dnu: float, opd_max: float, fwhm_gauss: float, *, apodization="gaussian",
               n_pts: int = 2048, ctx: Optional[Context] = None):
    """(delta_nu array, taps) area-normalised (ils.jl:80-114)."""
    a = _apod(apodization)
    ctx = ctx or default_context()
    half = i64()
    ctx.check(ctx.lib.irs_ils_half(float(dnu), C.byref(half)))
    taps = np.zeros(2 * half.value + 1)
    ctx.check(ctx.lib.irs_ils_kernel(ctx.h, float(dnu), float(opd_max), float(fwhm_gauss), a, n_pts,
                                     _p(taps)))
    return np.arange(-half.value, half.value + 1) * dnu, taps


def apply_ils_channels(nu_grid: WavenumberGrid, spectrum, sounder: Sounder, *,
                       apodization="gaussian", ctx: Optional[Context] = None) -> np.ndarray:
    """ILS-convolved spectrum sampled at the sounder channels (ils.jl:131-153 +
    sounder.jl:372-377), evaluated only at the channel nodes."""
    ctx = ctx or default_context()
    spec = np.ascontiguousarray(spectrum, np.float64)
    if len(spec) != nu_grid.n:
        raise ValueError("spectrum length mismatch")
    snd = _sounder(sounder)
    out = np.zeros(sounder_grid(sounder).n)
    ctx.check(ctx.lib.irs_ils_channels(ctx.h, _p(spec), nu_grid.nu_min, nu_grid.dnu, nu_grid.n,
                                       C.byref(snd), _apod(apodization), _p(out)))
    return out


apply_ils = apply_ils_channels


def planck_radiance(nu, T):
    """B(nu, T) = C1 nu^3 / expm1(C2 nu / T) (planck.jl:22-24); host arithmetic."""
    nu = np.asarray(nu, np.float64)
    return 1.191042953e-5 * nu * nu * nu / np.expm1(1.4387769 * nu / T)


def brightness_temperature(nu, R):
    """Inverse Planck through the C ABI (planck.jl:43-57)."""
    lib = load_library()
    nu = np.ascontiguousarray(np.atleast_1d(nu), np.float64)
    R = np.ascontiguousarray(np.atleast_1d(R), np.float64)
    if nu.shape != R.shape:
        raise ValueError(f"R length {len(R)} ≠ grid size {len(nu)}")
    out = np.zeros_like(R)
    rc = lib.irs_brightness_temperature(_p(nu), _p(R), len(R), _p(out))
    if rc != IRS_OK:
        raise ValueError("bad arguments")
    return out


def _options(internal_dnu=0.001, high_res_factor=None, cutoff=25.0, lm_cutoff=None, dptmn=1e-6,
             x_far=122.0, with_ils=True, apodization="gaussian", line_mixing=None,
             T_method="logp_at_pcg", source_function="cim", continua_mask=0):
    o = IrsOptions()
    load_library().irs_default_options(C.byref(o))
    o.internal_dnu = 0.0 if internal_dnu is None else float(internal_dnu)
    o.high_res_factor = 0 if high_res_factor is None else int(high_res_factor)
    o.cutoff = float(cutoff)
    o.lm_cutoff = float(cutoff if lm_cutoff is None else lm_cutoff)
    o.dptmn = float(dptmn)
    o.x_far = float(x_far)
    o.with_ils = int(bool(with_ils))
    o.apodization = _apod(apodization)
    lm = line_mixing
    if isinstance(lm, (VPYLineMixing, VPWLineMixing)):
        o.line_mixing = 2            # completed by _relmat_options
    elif lm in (None, False, 0, "none"):
        o.line_mixing = 0
    elif lm in ("vpy", "VPY", 1, True):
        o.line_mixing = 1
    else:
        raise ValueError(f"unsupported line_mixing {line_mixing!r}")
    tm = T_method.lstrip(":")
    if tm not in T_METHOD:
        raise ValueError(f"T_method must be :logp_at_pcg or :mass_weighted, got :{tm}")
    o.T_method = T_METHOD[tm]
    sf = source_function.lstrip(":")
    if sf not in SOURCE_FUNCTION:
        raise ValueError(f"source_function must be :toon or :cim, got :{sf}")
    o.source_function = SOURCE_FUNCTION[sf]
    o.continua = int(continua_mask)
    return o


def _relmat_options(o, line_mixing, ctx, linelists):
    """Relaxation-matrix fields of the options for a VPYLineMixing / VPWLineMixing:
    the uploaded matrix, its species and the per-band treatment, restricted to the
    isotopologues of that species' line list (line_mixing.jl:199-206,222-241).
    Returns the band-mode array (kept alive by the caller)."""
    if not isinstance(line_mixing, (VPYLineMixing, VPWLineMixing)):
        return None
    data = line_mixing.data
    ll = linelists.get(data.species)
    iso = linelist_isotopes(ll) if ll is not None else None
    modes = np.ascontiguousarray(relmat_band_modes(line_mixing, iso), np.uint8)
    o.relmat = ctx.upload_relmat(data)
    o.lm_species = int(data.species)
    o.relmat_band_mode = _p(modes) if len(modes) else None
    return modes


class _Batch:
    """ctypes view of a batch of profiles (keeps the numpy buffers alive)."""

    def __init__(self, ctx, profs, linelists, T_sfcs, eps, mus):
        species = list(profs[0].vmr.keys())
        self.species = np.array([int(s) for s in species], np.int32)
        self.handles = np.array([ctx.upload(linelists[s]) if s in linelists and len(linelists[s]) else -1
                                 for s in species], np.int32)
        self.arr = (IrsProfile * len(profs))()
        self.keep = []
        for q, pr in enumerate(profs):
            if list(pr.vmr.keys()) != species:
                raise ValueError("all profiles of a batch need the same species")
            vmr = np.ascontiguousarray(np.stack([pr.vmr[s] for s in species]), np.float64)
            p = np.ascontiguousarray(pr.pressure, np.float64)
            T = np.ascontiguousarray(pr.temperature, np.float64)
            self.keep += [vmr, p, T]
            self.arr[q] = IrsProfile(len(p), _p(p), _p(T), len(species), _p(self.species),
                                     _p(self.handles), _p(vmr),
                                     float(T_sfcs[q]) if T_sfcs[q] is not None else 0.0,
                                     float(eps[q]), float(mus[q]))


def forward_model(prof, linelists: Dict[GasSpecies, HITRANLinelist], *,
                  sounder: Optional[Sounder] = None, iasi: Optional[Sounder] = None,
                  geom: Optional[ViewingGeometry] = None, T_sfc=None, eps_sfc: float = 1.0,
                  internal_dnu: Optional[float] = 0.001, high_res_factor: Optional[int] = None,
                  cutoff: float = 25.0, lm_cutoff: Optional[float] = None, dptmn: float = 1e-6,
                  apply_continuum: bool = True, continua=DEFAULT_CONTINUA, with_ils: bool = True,
                  apodization="gaussian", line_mixing=None, T_method="logp_at_pcg",
                  source_function="cim", x_far: float = 122.0, ctx: Optional[Context] = None):
    """forward_model(prof, linelists; ...) -> (nu, R, BT)   (src/Sensor/sounder.jl:233-360).

    `prof` may be one AtmosphericProfile or a list of them (a batch on one
    device launch sequence); geom, T_sfc and eps_sfc may then be lists too.
    apply_continuum / continua add the MT-CKD H2O and CO2 continua and the
    CIA pairs whose tables are loaded (Context.load_cia) per layer, as the
    reference does by default (sounder.jl:245-246,310-322)."""
    cmask = _continua_mask(apply_continuum, continua)
    inst = iasi if iasi is not None else (sounder if sounder is not None else IASIInstrument())
    batch = isinstance(prof, (list, tuple))
    profs = list(prof) if batch else [prof]
    nb = len(profs)
    geoms = geom if isinstance(geom, (list, tuple)) else [geom] * nb
    geoms = [g if g is not None else nadir_geometry() for g in geoms]
    tsf = T_sfc if isinstance(T_sfc, (list, tuple, np.ndarray)) else [T_sfc] * nb
    eps = eps_sfc if isinstance(eps_sfc, (list, tuple, np.ndarray)) else [eps_sfc] * nb
    ctx = ctx or default_context()
    opt = _options(internal_dnu, high_res_factor, cutoff, lm_cutoff, dptmn, x_far, with_ils,
                   apodization, line_mixing, T_method, source_function, cmask)
    _modes = _relmat_options(opt, line_mixing, ctx, linelists)     # noqa: F841 (kept alive for the call)
    b = _Batch(ctx, profs, linelists, tsf, eps, [g.mu for g in geoms])
    snd = _sounder(inst)
    grid = sounder_grid(inst)
    R = np.zeros((nb, grid.n))
    BT = np.zeros((nb, grid.n))
    ctx.check(ctx.lib.irs_forward(ctx.h, b.arr, nb, C.byref(snd), C.byref(opt), _p(R), _p(BT)))
    if batch:
        return grid, R, BT
    return grid, R[0], BT[0]


def analytic_jacobian(prof: AtmosphericProfile, linelists: Dict[GasSpecies, HITRANLinelist],
                      spec: StateVectorSpec, *, sounder: Optional[Sounder] = None,
                      iasi: Optional[Sounder] = None, geom: Optional[ViewingGeometry] = None,
                      T_sfc=None, eps_sfc: float = 1.0, observable: str = "bt",
                      internal_dnu: Optional[float] = 0.001, high_res_factor: Optional[int] = None,
                      cutoff: float = 25.0, lm_cutoff: Optional[float] = None,
                      apply_continuum: bool = True, continua=DEFAULT_CONTINUA, with_ils: bool = True,
                      apodization="gaussian",
                      line_mixing=None, T_method="logp_at_pcg", source_function="cim",
                      vmr_coupling: bool = True, x_far: float = 122.0,
                      ctx: Optional[Context] = None) -> Jacobian:
    """analytic_jacobian(prof, linelists, spec; ...) -> Jacobian
    (src/Estimation/vmr_jacobian.jl:152-505).

    K = dy/dx about the state packed from prof (+ T_sfc, eps_sfc), y = BT
    (observable="bt") or radiance.  Full line set (no DPTMIN), as the
    reference runs it.  Full-profile and column-scale blocks are assembled on
    the device; a PartialColumns block is the device's per-level log-VMR
    columns times its basis (the ILS and the channel map are linear, so this
    equals projecting before the convolution)."""
    if observable not in ("bt", "radiance"):
        raise ValueError(f"observable must be :bt or :radiance, got :{observable}")
    cmask = _continua_mask(apply_continuum, continua)
    if prof.n_levels != spec.n_levels:
        raise ValueError(f"profile has {prof.n_levels} levels, spec expects {spec.n_levels}")
    inst = iasi if sounder is None else sounder
    inst = inst if inst is not None else IASIInstrument()
    geom = geom if geom is not None else nadir_geometry()
    ctx = ctx or default_context()
    opt = _options(internal_dnu, high_res_factor, cutoff, lm_cutoff, 0.0, x_far, with_ils,
                   apodization, line_mixing, T_method, source_function, cmask)
    _modes = _relmat_options(opt, line_mixing, ctx, linelists)     # noqa: F841 (kept alive for the call)
    # device blocks: ColumnScale natively, PartialColumns as full log-VMR columns
    blocks = spec.vmr_blocks
    if len(blocks) > 4:
        raise ValueError("at most 4 retrieved species per call")
    for b in blocks:
        if b.species not in prof.vmr:
            raise ValueError(f"profile is missing retrieved species {b.species.name}")
    is_cs = [b.basis is not None and b.m == 1 and bool(np.all(b.basis == 1.0)) for b in blocks]
    retrieved = np.array([int(b.species) for b in blocks] or [0], np.int32)
    colscale = np.array([1 if c else 0 for c in is_cs] or [0], np.int32)
    batch = _Batch(ctx, [prof], linelists, [T_sfc], [eps_sfc], [geom.mu])
    L = ctx.lib
    n_l = prof.n_levels

    ss = IrsStateSpec(int(spec.include_temperature), len(blocks), _p(retrieved), _p(colscale),
                      int(spec.include_tsfc), int(spec.include_emissivity), int(spec.log_vmr),
                      int(bool(vmr_coupling)), 0 if observable == "bt" else 1)
    n_dev = L.irs_state_size(C.byref(ss), n_l)
    grid = sounder_grid(inst)
    Kdev = np.zeros((n_dev, grid.n))
    y0 = np.zeros(grid.n)
    xd = np.zeros(n_dev)
    ctx.check(L.irs_jacobian(ctx.h, batch.arr, C.byref(_sounder(inst)), C.byref(opt), C.byref(ss),
                             _p(Kdev), _p(y0), _p(xd)))
    Kd = Kdev.T
    K = np.zeros((grid.n, spec.n))
    x0 = np.zeros(spec.n)
    c = 0
    if spec.include_temperature:
        K[:, :n_l] = Kd[:, :n_l]
        x0[:n_l] = xd[:n_l]
        c = n_l
    for b, cs in zip(blocks, is_cs):
        r = slice(b.range.start, b.range.stop)
        if b.basis is None:
            K[:, r] = Kd[:, c:c + n_l]
            x0[r] = xd[c:c + n_l]
            c += n_l
        elif cs:
            K[:, r] = Kd[:, c:c + 1]
            c += 1
        else:
            cols = Kd[:, c:c + n_l]
            if not spec.log_vmr:      # reduced blocks are always multiplicative-log
                cols = cols * prof.vmr[b.species][None, :]
            K[:, r] = cols @ b.basis
            c += n_l
    if spec.include_tsfc:
        K[:, spec.tsfc_index] = Kd[:, c]
        x0[spec.tsfc_index] = xd[c]
        c += 1
    if spec.include_emissivity:
        K[:, spec.emis_index] = Kd[:, c]
        x0[spec.emis_index] = xd[c]
    return Jacobian(K, spec, grid.nu, x0, y0, observable)


def channel_optical_depths(prof, linelists: Dict[GasSpecies, HITRANLinelist], *,
                           sounder: Optional[Sounder] = None, geom=None, floor: float = 1e-4,
                           internal_dnu: Optional[float] = 0.001, high_res_factor: Optional[int] = None,
                           cutoff: float = 25.0, lm_cutoff: Optional[float] = None, dptmn: float = 1e-6,
                           with_ils: bool = True, apodization="gaussian", line_mixing=None,
                           T_method="logp_at_pcg", x_far: float = 122.0, cim: bool = False,
                           moments: bool = False, apply_continuum: bool = True, continua=DEFAULT_CONTINUA,
                           ctx: Optional[Context] = None):
    """Per-layer channel optical depths along each profile's viewing path
    (geom, default nadir), [n_prof, n_layers, n_channels] float32
    (irs_channel_od): tau_k = ln(T_{k+1}/T_k) of the ILS-convolved
    level-to-space transmittances T_k, or the layer-alone channel optical depth
    where T_{k+1} < floor.  The emulator's training target.  With cim=True also
    returns the CIM channel weights g (same shape): (tau, g); with moments=True
    also the normalised ILS first moments m of each layer's transmittance step
    (irs_channel_od_moments): (tau, g, m)."""
    inst = sounder if sounder is not None else IASIInstrument()
    batch = isinstance(prof, (list, tuple))
    profs = list(prof) if batch else [prof]
    ctx = ctx or default_context()
    opt = _options(internal_dnu, high_res_factor, cutoff, lm_cutoff, dptmn, x_far, with_ils,
                   apodization, line_mixing, T_method, "cim", _continua_mask(apply_continuum, continua))
    _modes = _relmat_options(opt, line_mixing, ctx, linelists)     # noqa: F841 (kept alive for the call)
    nb = len(profs)
    geoms = geom if isinstance(geom, (list, tuple)) else [geom] * nb
    mus = [1.0 if g is None else g.mu for g in geoms]
    b = _Batch(ctx, profs, linelists, [None] * nb, [1.0] * nb, mus)
    grid = sounder_grid(inst)
    out = np.zeros((nb, profs[0].n_levels - 1, grid.n), np.float32)
    g = np.zeros_like(out) if (cim or moments) else None
    if moments:
        m = np.zeros_like(out)
        ctx.check(ctx.lib.irs_channel_od_moments(ctx.h, b.arr, nb, C.byref(_sounder(inst)), C.byref(opt),
                                                 float(floor), _p(out), _p(g), _p(m)))
        return (out, g, m) if batch else (out[0], g[0], m[0])
    ctx.check(ctx.lib.irs_channel_od(ctx.h, b.arr, nb, C.byref(_sounder(inst)), C.byref(opt),
                                     float(floor), _p(out), _p(g)))
    if cim:
        return (out, g) if batch else (out[0], g[0])
    return out if batch else out[0]


def layer_properties(prof: AtmosphericProfile, T_method: str = "logp_at_pcg") -> dict:
    """layer_properties(prof; T_method) (src/Atmosphere/layers.jl:10-178): p_mid, dp,
    T_ave (TBAR) per layer and vmr_cg, p_cg, T_cg per species (host C code)."""
    tm = T_method.lstrip(":")
    if tm not in T_METHOD:
        raise ValueError(f"T_method must be :logp_at_pcg or :mass_weighted, got :{tm}")
    species = list(prof.vmr.keys())
    sp = np.array([int(s) for s in species] or [0], np.int32)
    hs = np.full(max(len(species), 1), -1, np.int32)
    vmr = np.ascontiguousarray(np.stack([prof.vmr[s] for s in species]) if species else np.zeros((1, 1)))
    p, T = prof.pressure, prof.temperature
    pr = IrsProfile(len(p), _p(p), _p(T), len(species), _p(sp), _p(hs), _p(vmr), 0.0, 1.0, 1.0)
    L = prof.n_levels - 1
    out = {k: np.zeros(L) for k in ("p_mid", "dp", "T_ave")}
    vc, pc, tc = (np.zeros((max(len(species), 1), L)) for _ in range(3))
    rc = load_library().irs_layer_properties(C.byref(pr), T_METHOD[tm], _p(out["p_mid"]), _p(out["dp"]),
                                             _p(out["T_ave"]), _p(vc), _p(pc), _p(tc))
    if rc != IRS_OK:
        raise ValueError("bad profile for layer_properties")
    out["vmr_cg"] = {s: vc[i] for i, s in enumerate(species)}
    out["p_cg"] = {s: pc[i] for i, s in enumerate(species)}
    out["T_cg"] = {s: tc[i] for i, s in enumerate(species)}
    return out


def channel_transmittance(prof, linelists: Dict[GasSpecies, HITRANLinelist], *,
                          sounder: Optional[Sounder] = None, geom=None,
                          internal_dnu: Optional[float] = 0.001, high_res_factor: Optional[int] = None,
                          cutoff: float = 25.0, lm_cutoff: Optional[float] = None, dptmn: float = 1e-6,
                          with_ils: bool = True, apodization="gaussian", line_mixing=None,
                          T_method="logp_at_pcg", x_far: float = 122.0, apply_continuum: bool = True,
                          continua=DEFAULT_CONTINUA, ctx: Optional[Context] = None):
    """ILS-convolved level-to-space transmittances along each profile's path,
    T_lev [n_prof, n_layers+1, n_channels], and the CIM sums
    G [n_prof, n_layers, n_channels] (irs_channel_transmittance): the
    emulator's training target."""
    inst = sounder if sounder is not None else IASIInstrument()
    batch = isinstance(prof, (list, tuple))
    profs = list(prof) if batch else [prof]
    ctx = ctx or default_context()
    opt = _options(internal_dnu, high_res_factor, cutoff, lm_cutoff, dptmn, x_far, with_ils,
                   apodization, line_mixing, T_method, "cim", _continua_mask(apply_continuum, continua))
    _modes = _relmat_options(opt, line_mixing, ctx, linelists)     # noqa: F841 (kept alive for the call)
    nb = len(profs)
    geoms = geom if isinstance(geom, (list, tuple)) else [geom] * nb
    mus = [1.0 if g is None else g.mu for g in geoms]
    b = _Batch(ctx, profs, linelists, [None] * nb, [1.0] * nb, mus)
    grid = sounder_grid(inst)
    L = profs[0].n_levels - 1
    Tl = np.zeros((nb, L + 1, grid.n), np.float32)
    G = np.zeros((nb, L, grid.n), np.float32)
    ctx.check(ctx.lib.irs_channel_transmittance(ctx.h, b.arr, nb, C.byref(_sounder(inst)), C.byref(opt),
                                                _p(Tl), _p(G)))
    return (Tl, G) if batch else (Tl[0], G[0])


def forward_dims(sounder: Sounder, **kw):
    o = _options(**kw)
    snd = _sounder(sounder)
    lo, d, n, nch = f64(), f64(), i64(), i64()
    rc = load_library().irs_forward_dims(C.byref(snd), C.byref(o), C.byref(lo), C.byref(d),
                                         C.byref(n), C.byref(nch))
    if rc != IRS_OK:
        raise ValueError("bad sounder/options")
    return lo.value, d.value, n.value, nch.value
