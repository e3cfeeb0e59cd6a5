"""Relaxation-matrix line mixing: data, loaders and the line-mixing models.

Mirrors the data side of src/CrossSections/line_mixing.jl: the HITRAN 2020
relaxation-matrix files (BandInfo.dat, S*.dat, WTfit*.dat) parsed into bands of
S-file lines and W0/B0 tables (`load_hitran_relmat`), and the two models a
forward model takes as `line_mixing`:

* `VPYLineMixing(relmat)` -- first-order Rosenkranz mixing with Y(T) from the
  sum-rule-corrected relaxation matrix of every band (line_mixing.jl:144-175);
* `VPWLineMixing(relmat; whitelist)` -- the eigenmodes of D + i p W for the
  whitelisted bands and the VP_Y term for the rest (line_mixing.jl:177-206).

The per-layer work (W(T), Y(T), the eigen-decompositions and the pole and
dispersive sums over the internal grid) runs in libirs: the relaxation data is
uploaded once (`irs_relmat_upload`), the host side of the library builds the
per-layer terms and the GPU evaluates them on the grid.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import Dict, Iterable, List, Optional, Sequence, Set, Tuple

import numpy as np

from .types import GasSpecies

T0_LM = 296.0          # line_mixing.jl:30
CT_LM = 1.4388         # line_mixing.jl:31
SUBTABLES = ("pp", "pq", "pr", "qp", "qq", "qr", "rp", "rq", "rr")

# isotopologue masses (amu) and natural abundances (line_mixing.jl:743-766)
LM_MASS_AMU = {2: {1: 43.98983, 2: 44.99318, 3: 45.99398, 4: 44.99403, 5: 45.99706, 6: 44.99706,
                   7: 47.99832, 8: 47.00134, 9: 48.00196, 10: 47.00427},
               6: {1: 16.031300, 2: 17.034655, 3: 17.037475, 4: 18.040830}}
LM_ISO_ABUND = {2: {1: 0.984204, 2: 0.011057, 3: 0.003947, 4: 0.000734, 5: 4.434e-5, 6: 1.625e-5,
                    7: 3.957e-6, 8: 1.4717e-7, 9: 6.554e-8, 10: 0.0},
                6: {1: 0.988274, 2: 0.0111031, 3: 0.000615751, 4: 6.91785e-6}}


def lm_mass_amu(molecule: int, iso: int) -> float:
    """Isotopologue mass; 44.0 for an unknown molecule or isotopologue (line_mixing.jl:772)."""
    return LM_MASS_AMU.get(int(molecule), {}).get(int(iso), 44.0)


def lm_iso_abund(molecule: int, iso: int) -> float:
    """Natural abundance; 1e-8 when unknown (line_mixing.jl:773)."""
    return LM_ISO_ABUND.get(int(molecule), {}).get(int(iso), 1e-8)


@dataclass
class RelmatLine:
    """One S-file line (line_mixing.jl:35-47)."""
    isot: int
    nu: float
    Ji: int
    branch: int              # -1 P, 0 Q, +1 R
    gV_air: float
    n_air: float             # stored as Float32
    shift: float             # stored as Float32
    E_lower: float
    Dipo0: float
    PopuT0: float
    DipoT: float


class RelmatBand:
    """A band of S-file lines (line_mixing.jl:72-89), the lines kept as arrays in
    their file order."""

    def __init__(self, name: str, li: int, lf: int, isot: int, nu_min: float, nu_max: float,
                 lines: Sequence[RelmatLine], molecule: int = 2):
        self.name = str(name)
        self.li, self.lf, self.isot = int(li), int(lf), int(isot)
        self.nu_min, self.nu_max = float(nu_min), float(nu_max)
        self.molecule = int(molecule)
        ls = list(lines)
        self.line_isot = np.array([l.isot for l in ls], np.int8)
        self.nu = np.array([l.nu for l in ls], np.float64)
        self.Ji = np.array([l.Ji for l in ls], np.int16)
        self.branch = np.array([l.branch for l in ls], np.int8)
        self.gV_air = np.array([l.gV_air for l in ls], np.float64)
        self.n_air = np.array([l.n_air for l in ls], np.float32)
        self.shift = np.array([l.shift for l in ls], np.float32)
        self.E_lower = np.array([l.E_lower for l in ls], np.float64)
        self.Dipo0 = np.array([l.Dipo0 for l in ls], np.float64)
        self.PopuT0 = np.array([l.PopuT0 for l in ls], np.float64)
        self.DipoT = np.array([l.DipoT for l in ls], np.float64)

    @property
    def n(self) -> int:
        return len(self.nu)

    def __repr__(self):
        return f"RelmatBand({self.name!r}, li={self.li}, lf={self.lf}, isot={self.isot}, {self.n} lines)"


class W0B0Table:
    """W0/B0 parameters per branch-pair sub-table, keyed by (Ji, Jip)
    (line_mixing.jl:92-104): W(T) = exp(W0) (T/T0)^B0."""

    def __init__(self):
        for s in SUBTABLES:
            setattr(self, s, {})

    def entries(self):
        for si, s in enumerate(SUBTABLES):
            for (ji, jip), (w0, b0) in getattr(self, s).items():
                yield si, int(ji), int(jip), float(w0), float(b0)


@dataclass
class HITRANRelmatData:
    """Bands + W0/B0 tables keyed by (min(li, lf), max(li, lf)) (line_mixing.jl:106-117)."""
    bands: List[RelmatBand]
    wtfit: Dict[Tuple[int, int], W0B0Table]
    species: GasSpecies = GasSpecies.CO2
    _handles: Dict[int, int] = field(default_factory=dict, repr=False, compare=False)


def band_eff_strength(band: RelmatBand) -> float:
    """Integrated band intensity at 296 K, sum DipoT^2 PopuT0 nu stim
    (line_mixing.jl:804-811); the S-file data already carry the abundance."""
    stim = 1.0 - np.exp(-CT_LM * band.nu / T0_LM)
    return float(np.sum(band.DipoT ** 2 * band.PopuT0 * band.nu * stim))


def default_vpw_whitelist(data: HITRANRelmatData, n_top: int = 5, isotopes: Optional[Iterable[int]] = None,
                          nu_window: Optional[Tuple[float, float]] = None) -> Set[str]:
    """Top-n_top bands by S(296) x abundance, or per isotopologue by S(296)
    (line_mixing.jl:813-853); nu_window keeps bands overlapping it."""
    def in_window(b):
        return nu_window is None or (b.nu_max >= nu_window[0] and b.nu_min <= nu_window[1])
    if isotopes is None:
        scored = [(b.name, band_eff_strength(b) * lm_iso_abund(b.molecule, b.isot)) for b in data.bands
                  if in_window(b)]
        scored.sort(key=lambda t: t[1], reverse=True)     # stable, as Julia's sort!
        return {nm for nm, _ in scored[:n_top]}
    out = set()
    for iso in isotopes:
        scored = [(b.name, band_eff_strength(b)) for b in data.bands if b.isot == int(iso) and in_window(b)]
        scored.sort(key=lambda t: t[1], reverse=True)
        out |= {nm for nm, _ in scored[:n_top]}
    return out


@dataclass
class VPYLineMixing:
    """First-order (VP_Y) mixing with Y(T) from the relaxation matrix (line_mixing.jl:144-175)."""
    data: HITRANRelmatData
    min_band_strength: float = 0.0


@dataclass
class VPWLineMixing:
    """Full-matrix mixing for the whitelisted bands, VP_Y for the others (line_mixing.jl:177-206)."""
    data: HITRANRelmatData
    whitelist: Set[str]
    min_band_strength: float = 0.0

    @staticmethod
    def make(data: HITRANRelmatData, n_top: int = 5, isotopes=None, nu_window=None,
             whitelist: Optional[Set[str]] = None, min_band_strength: float = 0.0) -> "VPWLineMixing":
        wl = default_vpw_whitelist(data, n_top, isotopes, nu_window) if whitelist is None else set(whitelist)
        return VPWLineMixing(data, wl, min_band_strength)


def linelist_isotopes(ll) -> Set[int]:
    """Isotopologues present in a line list, HITRAN's 0 read as 10 (line_mixing.jl:199-206)."""
    return {10 if int(i) == 0 else int(i) for i in np.unique(ll.iso_id)}


def band_modes_for(lm, isotopes: Optional[Set[int]]) -> np.ndarray:
    """Per band: 0 skipped, 1 VP_Y dispersive, 2 VP_W eigenmodes -- the band filters of
    compute_lm_dispersive_correction and _vpw_perturbation (line_mixing.jl:1120-1154,
    1303-1349): li > 8 (the BandInfo skip), an isotopologue without Voigt baseline,
    min_band_strength, no W0/B0 table."""
    data = lm.data
    wl = lm.whitelist if isinstance(lm, VPWLineMixing) else set()
    out = np.zeros(len(data.bands), np.int32)
    for i, b in enumerate(data.bands):
        if b.li > 8:
            continue
        if isotopes is not None and b.isot not in isotopes:
            continue
        if lm.min_band_strength > 0.0 and band_eff_strength(b) < lm.min_band_strength:
            continue
        if (min(b.li, b.lf), max(b.li, b.lf)) not in data.wtfit:
            continue
        out[i] = 2 if b.name in wl else 1
    return out


# ---- file parsers (line_mixing.jl:305-528) ---------------------------------------
def _fortran_d(s: str) -> float:
    return float(s.strip().replace("D", "e").replace("d", "e"))


def parse_bandinfo(basedir: str, nu_min: float, nu_max: float, stot_min: float = 1e-28) -> list:
    """BandInfo.dat entries overlapping [nu_min, nu_max] with Stot >= stot_min
    (line_mixing.jl:338-381).  Bands with any branch of fewer than 40 J values get
    li = 55 (the Fortran skip)."""
    out = []
    with open(os.path.join(basedir, "BandInfo.dat"), encoding="latin-1") as fh:
        for raw in fh:
            line = raw.rstrip("\r\n ")
            if len(line) < 71:
                continue
            try:
                isot = int(line[0:1])
                if isot == 0:
                    isot = 10
                lf = int(line[3:4])
                li = int(line[8:9])
                stot = float(line[13:25])
                if stot < stot_min:
                    continue
                sg_min = float(line[26:38])
                sg_max = float(line[39:51])
                if sg_min > nu_max or sg_max < nu_min:
                    continue
                jp, jq, jr = int(line[59:63]), int(line[63:67]), int(line[67:71])
                band_id = line[0:13].rstrip()
                li_eff = 55 if (jp < 40 or jr < 40 or (0 <= jq < 40)) else li
                out.append(dict(sfile="S" + band_id, li=li_eff, lf=lf, isot=isot, nu_min=sg_min, nu_max=sg_max,
                                Stot=stot, jmxP=jp, jmxQ=jq, jmxR=jr))
            except ValueError:
                continue
    return out


def parse_sfile(path: str) -> List[RelmatLine]:
    """S*.dat lines (line_mixing.jl:393-432); DipoT = sqrt(|S| / (|PopuT0| nu stim(296)))."""
    lines = []
    with open(path, encoding="latin-1") as fh:
        for raw in fh:
            line = raw.rstrip("\r\n")
            if len(line) < 215:
                continue
            try:
                c = line[2]
                isot = 10 if c == "0" else int(c)
                nu = float(line[3:15])
                S = float(line[15:25])
                gV = float(line[35:40])
                E = float(line[63:73])
                n_air = float(np.float32(line[73:77]))
                shift = float(np.float32(line[81:89]))
                br = line[139]
                Ji = int(line[140:143].strip())
                dipo0 = _fortran_d(line[183:199])
                popu0 = _fortran_d(line[199:215])
                branch = -1 if br == "P" else (1 if br == "R" else 0)
                stim = 1.0 - np.exp(-CT_LM * nu / T0_LM)
                dipoT = float(np.sqrt(max(0.0, abs(S) / (abs(popu0) * nu * stim))))
                lines.append(RelmatLine(isot, nu, Ji, branch, gV, n_air, shift, E, dipo0, popu0, dipoT))
            except (ValueError, IndexError):
                continue
    return lines


def load_wtfit(basedir: str, lli: int, llf: int) -> W0B0Table:
    """WTfit{lli}{llf}.dat (line_mixing.jl:434-478): (2d20.12, 2f14.6, 4I4)."""
    tbl = W0B0Table()
    fn = os.path.join(basedir, f"WTfit{lli}{llf}.dat")
    if not os.path.isfile(fn):
        return tbl
    with open(fn, encoding="latin-1") as fh:
        for raw in fh:
            line = raw.rstrip("\r\n")
            if len(line) < 84:
                continue
            try:
                w0 = _fortran_d(line[0:20])
                b0 = _fortran_d(line[20:40])
                jic, jfc = int(line[68:72].strip()), int(line[72:76].strip())
                jipc, jfpc = int(line[76:80].strip()), int(line[80:84].strip())
            except ValueError:
                continue
            ti = "p" if jic > jfc else ("q" if jic == jfc else "r")
            tip = "p" if jipc > jfpc else ("q" if jipc == jfpc else "r")
            getattr(tbl, ti + tip)[(jic, jipc)] = (w0, b0)
    return tbl


def load_hitran_relmat(basedir: str, nu_min: float, nu_max: float, stot_min: float = 1e-28) -> HITRANRelmatData:
    """Bands overlapping [nu_min, nu_max] with their S-files and W0/B0 tables
    (line_mixing.jl:480-528); bands with li or lf > 8 or |li - lf| > 1 are not loaded."""
    bands, wt = [], {}
    for bi in parse_bandinfo(basedir, nu_min, nu_max, stot_min):
        li, lf = bi["li"], bi["lf"]
        if li > 8 or lf > 8 or abs(li - lf) > 1:
            continue
        sf = os.path.join(basedir, bi["sfile"] + "  .dat")
        if not os.path.isfile(sf):
            sf = os.path.join(basedir, bi["sfile"] + ".dat")
        if not os.path.isfile(sf):
            continue
        lines = parse_sfile(sf)
        if not lines:
            continue
        bands.append(RelmatBand(bi["sfile"], li, lf, bi["isot"], bi["nu_min"], bi["nu_max"], lines))
        key = (min(li, lf), max(li, lf))
        if key not in wt:
            wt[key] = load_wtfit(basedir, *key)
    return HITRANRelmatData(bands, wt)
