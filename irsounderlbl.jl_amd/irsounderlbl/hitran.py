"""HITRAN .par ingest and the parsed-linelist cache.

Mirrors src/HITRAN/linelist.jl:90-103 (parse_hitran_par), src/HITRAN/fetch.jl:12-61
(load_hitran_par) and src/HITRAN/cache.jl:1-198 (linelist_cache_dir,
set_linelist_cache_dir!, clear_linelist_cache, load_linelist).  The records are
parsed on the device (irs_par_parse / irs_par_load in libirs); the cache is a
structure-of-arrays binary file per .par, keyed by the file's absolute path,
size, modification time and LINELIST_CACHE_VERSION, so an edited .par or a
version bump is simply a different key.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import warnings
from dataclasses import dataclass
from typing import Iterable, Optional, Union

import numpy as np

from .api import IRSError, IRS_ENOTFOUND, Context, _p, default_context, load_library
from .types import HITRANLinelist

LINELIST_CACHE_VERSION = 1
CACHE_SUFFIX = ".irsll"
P = C.c_void_p


class IrsParLines(C.Structure):
    _fields_ = [("mol", P), ("iso", P), ("nu", P), ("S", P), ("A", P), ("elower", P),
                ("gamma_air", P), ("gamma_self", P), ("n_air", P), ("delta_air", P)]


@dataclass(frozen=True)
class HITRANLine:
    """One HITRAN line (src/HITRAN/linelist.jl:20-31)."""
    mol_id: int
    iso_id: int
    wavenumber: float
    intensity: float
    a_coeff: float
    air_broad: np.float32
    self_broad: np.float32
    lower_energy: float
    temp_depend: np.float32
    pressure_shift: np.float32


def _declare(L):
    if getattr(L, "_par_declared", False):
        return
    L.irs_par_parse.argtypes = [P, C.c_char_p, C.c_int64, C.c_double, C.c_double, C.c_int64,
                                C.POINTER(IrsParLines), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.irs_par_load.argtypes = [P, C.c_char_p, C.c_char_p, C.c_int, C.c_double, C.c_double, C.c_int64,
                               C.POINTER(IrsParLines), C.POINTER(C.c_int64)]
    L._par_declared = True


class _Out:
    def __init__(self, cap: int):
        cap = max(int(cap), 1)
        self.a = {"mol": np.zeros(cap, np.int8), "iso": np.zeros(cap, np.int8), "nu": np.zeros(cap),
                  "S": np.zeros(cap), "A": np.zeros(cap), "elower": np.zeros(cap),
                  "gamma_air": np.zeros(cap, np.float32), "gamma_self": np.zeros(cap, np.float32),
                  "n_air": np.zeros(cap, np.float32), "delta_air": np.zeros(cap, np.float32)}
        self.s = IrsParLines(*[_p(self.a[f]) for f, _ in IrsParLines._fields_])

    def linelist(self, n: int) -> HITRANLinelist:
        a = {k: v[:n].copy() for k, v in self.a.items()}
        return HITRANLinelist(a["mol"], a["iso"], a["nu"], a["S"], a["gamma_air"], a["gamma_self"], a["elower"],
                              a["n_air"], a["delta_air"], a_coeff=a["A"], sort=False)


def _bounds(nu_min, nu_max):
    return (-math.inf if nu_min is None else float(nu_min)), (math.inf if nu_max is None else float(nu_max))


def parse_hitran_par_bytes(data: bytes, nu_min: Optional[float] = None, nu_max: Optional[float] = None,
                           ctx: Optional[Context] = None):
    """Parse the records of a .par buffer on the device: (HITRANLinelist sorted
    by wavenumber, number of records skipped as short or malformed)."""
    ctx = ctx or default_context()
    _declare(ctx.lib)
    lo, hi = _bounds(nu_min, nu_max)
    cap = len(data) // 67 + 1
    out = _Out(cap)
    n, skipped = C.c_int64(), C.c_int64()
    ctx.check(ctx.lib.irs_par_parse(ctx.h, data, len(data), lo, hi, cap, C.byref(out.s), C.byref(n),
                                    C.byref(skipped)))
    return out.linelist(n.value), skipped.value


def parse_hitran_par(record: str, ctx: Optional[Context] = None) -> HITRANLine:
    """One 160-character .par record (parse_hitran_par, linelist.jl:90-103);
    ValueError when its fields do not parse."""
    rec = record.rstrip("\n")
    if rec.endswith("\r"):
        rec = rec[:-1]
    if len(rec) < 67:
        raise ValueError("a .par record has at least 67 characters")
    ll, _ = parse_hitran_par_bytes((rec + "\n").encode("latin-1"), ctx=ctx)
    if len(ll) != 1:
        raise ValueError("malformed .par record")
    return HITRANLine(int(ll.mol_id[0]), int(ll.iso_id[0]), float(ll.wavenumber[0]), float(ll.intensity[0]),
                      float(ll.a_coeff[0]), ll.air_broad[0], ll.self_broad[0], float(ll.lower_energy[0]),
                      ll.temp_depend[0], ll.pressure_shift[0])


def load_hitran_par(filepath: str, nu_min: Optional[float] = None, nu_max: Optional[float] = None,
                    ctx: Optional[Context] = None) -> HITRANLinelist:
    """Lines of a local .par file in [nu_min, nu_max] (load_hitran_par, fetch.jl:12-61)."""
    if not os.path.isfile(filepath):
        raise FileNotFoundError(f"File not found: {filepath}")
    return _load(filepath, None, False, nu_min, nu_max, ctx)


def _load(path, cache_dir, rebuild, nu_min, nu_max, ctx):
    ctx = ctx or default_context()
    _declare(ctx.lib)
    lo, hi = _bounds(nu_min, nu_max)
    cap = os.path.getsize(path) // 67 + 1
    out = _Out(cap)
    n = C.c_int64()
    rc = ctx.lib.irs_par_load(ctx.h, os.fsencode(path), None if cache_dir is None else os.fsencode(cache_dir),
                              1 if rebuild else 0, lo, hi, cap, C.byref(out.s), C.byref(n))
    if rc == IRS_ENOTFOUND:
        raise FileNotFoundError(ctx.lib.irs_last_error(ctx.h).decode(errors="replace"))
    ctx.check(rc)
    if n.value == 0:
        warnings.warn(f"No lines loaded from {path}")
    return out.linelist(n.value)


# ---- cache directory (cache.jl:25-76) -------------------------------------------
def _prefs_path() -> str:
    base = os.environ.get("XDG_CONFIG_HOME") or os.path.join(os.path.expanduser("~"), ".config")
    return os.path.join(base, "irsounderlbl_amd", "preferences.json")


def _read_prefs() -> dict:
    try:
        with open(_prefs_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def linelist_cache_dir() -> str:
    """Where parsed-linelist caches live: IRSOUNDER_LINELIST_CACHE if set and
    non-empty, else the linelist_cache_dir preference (set_linelist_cache_dir),
    else a per-user cache directory.  Created if missing."""
    env = os.environ.get("IRSOUNDER_LINELIST_CACHE", "")
    if env:
        d = env
    else:
        d = _read_prefs().get("linelist_cache_dir")
        if not d:
            base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
            d = os.path.join(base, "irsounderlbl_amd", "linelists")
    os.makedirs(d, exist_ok=True)
    return d


def set_linelist_cache_dir(path: Optional[str]) -> None:
    """Persist a cache directory (None clears it); the environment variable still wins."""
    prefs = _read_prefs()
    if path is None:
        prefs.pop("linelist_cache_dir", None)
    else:
        prefs["linelist_cache_dir"] = str(path)
    os.makedirs(os.path.dirname(_prefs_path()), exist_ok=True)
    with open(_prefs_path(), "w") as f:
        json.dump(prefs, f)


def clear_linelist_cache() -> int:
    """Delete the cache files in linelist_cache_dir(); returns how many.  The
    .par files are untouched."""
    d = linelist_cache_dir()
    n = 0
    for f in os.listdir(d):
        if f.endswith(CACHE_SUFFIX):
            os.remove(os.path.join(d, f))
            n += 1
    return n


def _linelist_file(base: str, iso: int) -> str:
    return f"{base}.par" if iso == 1 else f"{base}_iso{iso}.par"


def load_linelist(parpath: str, isotopologues: Optional[Iterable[int]] = None, *,
                  nu_min: Optional[float] = None, nu_max: Optional[float] = None, cache: bool = True,
                  rebuild: bool = False, ctx: Optional[Context] = None) -> HITRANLinelist:
    """load_linelist(parpath; nu_min, nu_max, cache, rebuild) (cache.jl:153-175):
    the whole .par parsed once and cached, the window applied in memory.  With
    `isotopologues`, parpath is a base path and the files "<base>.par" (iso 1)
    and "<base>_iso<n>.par" are loaded and merged (cache.jl:177-198); a missing
    file is an error naming it."""
    if isotopologues is None:
        if not os.path.isfile(parpath):
            raise FileNotFoundError(f"File not found: {parpath}")
        return _load(parpath, linelist_cache_dir() if cache else None, rebuild, nu_min, nu_max, ctx)
    isos = list(isotopologues)
    missing = [_linelist_file(parpath, i) for i in isos if not os.path.isfile(_linelist_file(parpath, i))]
    if missing:
        raise FileNotFoundError("load_linelist: requested isotopologue(s) with no .par file on disk: " +
                                ", ".join(missing) + f" (base={parpath}, isotopologues={isos})")
    parts = [load_linelist(_linelist_file(parpath, i), nu_min=nu_min, nu_max=nu_max, cache=cache,
                           rebuild=rebuild, ctx=ctx) for i in isos]
    if sum(len(p) for p in parts) == 0:
        raise ValueError(f"No lines found for base={parpath}, isotopologues={isos} "
                         "(files present but empty in the requested window)")
    cat = {f: np.concatenate([getattr(p, f) for p in parts]) for f in
           HITRANLinelist.FIELDS_F64 + HITRANLinelist.FIELDS_F32 + ("mol_id", "iso_id")}
    return HITRANLinelist(cat["mol_id"], cat["iso_id"], cat["wavenumber"], cat["intensity"], cat["air_broad"],
                          cat["self_broad"], cat["lower_energy"], cat["temp_depend"], cat["pressure_shift"],
                          a_coeff=cat["a_coeff"])


# ---- writer ----------------------------------------------------------------------
def _fortran_f(v: float, w: int, d: int) -> str:
    s = f"{v:{w}.{d}f}"
    if len(s) > w:                       # Fortran Fw.d drops the leading zero: .0602, -.001234
        s = s.replace("0.", ".", 1)
    if len(s) > w:
        raise ValueError(f"{v} does not fit F{w}.{d}")
    return s.rjust(w)


def _fortran_e(v: float, w: int, d: int) -> str:
    s = f"{v:.{d}E}"
    if len(s) > w:
        raise ValueError(f"{v} does not fit E{w}.{d}")
    return s.rjust(w)


def format_hitran_par(ll: HITRANLinelist, i: int) -> str:
    """Line i of ll as a 160-character .par record (the fixed-width layout
    parse_hitran_par reads; the quantum-number and reference columns blank)."""
    rec = (f"{int(ll.mol_id[i]):2d}{int(ll.iso_id[i]) % 10:1d}{_fortran_f(ll.wavenumber[i], 12, 6)}"
           f"{_fortran_e(ll.intensity[i], 10, 3)}{_fortran_e(ll.a_coeff[i], 10, 3)}"
           f"{_fortran_f(float(ll.air_broad[i]), 5, 4)}{_fortran_f(float(ll.self_broad[i]), 5, 3)}"
           f"{_fortran_f(ll.lower_energy[i], 10, 4)}{_fortran_f(float(ll.temp_depend[i]), 4, 2)}"
           f"{_fortran_f(float(ll.pressure_shift[i]), 8, 6)}")
    return rec.ljust(160)


def write_hitran_par(path: str, ll: HITRANLinelist) -> None:
    """Write ll as a .par file (one 160-character record per line)."""
    with open(path, "w", newline="\n") as f:
        for i in range(len(ll)):
            f.write(format_hitran_par(ll, i) + "\n")
