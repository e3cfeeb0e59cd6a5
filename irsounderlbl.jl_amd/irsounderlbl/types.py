"""Host-side data types mirroring the reference's public types.

Each class follows the reference struct it stands in for; field names keep the
reference's meaning (src/HITRAN/linelist.jl:20-56, src/Atmosphere/profiles.jl:12-17,
src/Atmosphere/species.jl:6-52, src/Sensor/sounder.jl:33-140,
src/Sensor/geometry.jl:18-113, src/Utils/wavenumber_grid.jl:1-41).
"""
from __future__ import annotations

import enum
import math
from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np


class GasSpecies(enum.IntEnum):
    """HITRAN molecule ids (src/Atmosphere/species.jl:6-16)."""
    H2O = 1
    CO2 = 2
    O3 = 3
    N2O = 4
    CO = 5
    CH4 = 6
    O2 = 7
    SO2 = 9
    NH3 = 11


class HITRANLinelist:
    """Line list in structure-of-arrays form, sorted by unshifted wavenumber.

    Mirrors `HITRANLine`/`HITRANLinelist` (src/HITRAN/linelist.jl:20-56): FP64
    wavenumber/intensity/a_coeff/lower_energy, FP32 air_broad/self_broad/
    temp_depend/pressure_shift; the constructor sorts by wavenumber
    (linelist.jl:52).  `y_lm` (optional, FP64, atm^-1) is the first-order
    line-mixing coefficient per line used by the VP_Y dispersive term.
    """

    FIELDS_F64 = ("wavenumber", "intensity", "a_coeff", "lower_energy")
    FIELDS_F32 = ("air_broad", "self_broad", "temp_depend", "pressure_shift")

    def __init__(self, mol_id, iso_id, wavenumber, intensity, air_broad, self_broad,
                 lower_energy, temp_depend, pressure_shift, a_coeff=None, y_lm=None,
                 sort=True):
        n = len(wavenumber)
        self.mol_id = np.ascontiguousarray(np.broadcast_to(np.asarray(mol_id, np.int8), (n,)))
        self.iso_id = np.ascontiguousarray(np.broadcast_to(np.asarray(iso_id, np.int8), (n,)))
        self.wavenumber = np.ascontiguousarray(wavenumber, np.float64)
        self.intensity = np.ascontiguousarray(intensity, np.float64)
        self.a_coeff = np.ascontiguousarray(np.zeros(n) if a_coeff is None else a_coeff, np.float64)
        self.air_broad = np.ascontiguousarray(air_broad, np.float32)
        self.self_broad = np.ascontiguousarray(self_broad, np.float32)
        self.lower_energy = np.ascontiguousarray(lower_energy, np.float64)
        self.temp_depend = np.ascontiguousarray(temp_depend, np.float32)
        self.pressure_shift = np.ascontiguousarray(pressure_shift, np.float32)
        self.y_lm = None if y_lm is None else np.ascontiguousarray(y_lm, np.float64)
        for name in self.FIELDS_F64 + self.FIELDS_F32 + ("mol_id", "iso_id"):
            if getattr(self, name).shape != (n,):
                raise ValueError(f"field {name} has shape {getattr(self, name).shape}, expected ({n},)")
        if self.y_lm is not None and self.y_lm.shape != (n,):
            raise ValueError("y_lm length mismatch")
        if sort and n > 1 and np.any(np.diff(self.wavenumber) < 0):
            order = np.argsort(self.wavenumber, kind="stable")
            for name in self.FIELDS_F64 + self.FIELDS_F32 + ("mol_id", "iso_id"):
                setattr(self, name, np.ascontiguousarray(getattr(self, name)[order]))
            if self.y_lm is not None:
                self.y_lm = np.ascontiguousarray(self.y_lm[order])

    def __len__(self):
        return len(self.wavenumber)

    @property
    def nu_min(self):
        return float(self.wavenumber[0]) if len(self) else 0.0

    @property
    def nu_max(self):
        return float(self.wavenumber[-1]) if len(self) else 0.0

    @property
    def molecules(self):
        return set(int(m) for m in np.unique(self.mol_id))

    def subset(self, mask):
        return HITRANLinelist(self.mol_id[mask], self.iso_id[mask], self.wavenumber[mask],
                              self.intensity[mask], self.air_broad[mask], self.self_broad[mask],
                              self.lower_energy[mask], self.temp_depend[mask],
                              self.pressure_shift[mask], self.a_coeff[mask],
                              None if self.y_lm is None else self.y_lm[mask], sort=False)

    def __repr__(self):
        return (f"HITRANLinelist({len(self)} lines, mol_ids={sorted(self.molecules)}, "
                f"{self.nu_min}–{self.nu_max} cm⁻¹)")


def filter_linelist(ll: HITRANLinelist, nu_min: float, nu_max: float) -> HITRANLinelist:
    """Lines with nu in [nu_min, nu_max] (src/HITRAN/linelist.jl:61-67)."""
    m = (ll.wavenumber >= nu_min) & (ll.wavenumber <= nu_max)
    return ll.subset(m)


@dataclass
class AtmosphericProfile:
    """State on pressure levels, surface first (src/Atmosphere/profiles.jl:12-17)."""
    pressure: np.ndarray          # hPa
    temperature: np.ndarray       # K
    altitude: np.ndarray          # km
    vmr: Dict[GasSpecies, np.ndarray] = field(default_factory=dict)

    def __post_init__(self):
        self.pressure = np.ascontiguousarray(self.pressure, np.float64)
        self.temperature = np.ascontiguousarray(self.temperature, np.float64)
        self.altitude = np.ascontiguousarray(self.altitude, np.float64)
        self.vmr = {GasSpecies(k): np.ascontiguousarray(v, np.float64) for k, v in self.vmr.items()}
        n = len(self.pressure)
        for a in (self.temperature, self.altitude, *self.vmr.values()):
            if len(a) != n:
                raise ValueError("profile arrays must all have n_levels entries")

    @property
    def n_levels(self):
        return len(self.pressure)


@dataclass
class Sounder:
    """FTS sounder (src/Sensor/sounder.jl:33-56)."""
    nu_min: float
    nu_max: float
    dnu: float
    n_channels: int
    opd_max: float
    fwhm_gauss: float

    @staticmethod
    def make(nu_min, nu_max, dnu, opd_max=None, fwhm_gauss=0.0, n_channels=None):
        if n_channels is None:
            n_channels = int(round((nu_max - nu_min) / dnu)) + 1
        if opd_max is None:
            opd_max = 1.0 / (2.0 * dnu)
        return Sounder(float(nu_min), float(nu_max), float(dnu), int(n_channels),
                       float(opd_max), float(fwhm_gauss))


def IASIInstrument(fwhm_gauss: float = 0.5) -> Sounder:
    """IASI L1C: 645–2760 cm⁻¹, 0.25 cm⁻¹, OPD 2 cm, Gaussian FWHM 0.5 (sounder.jl:65-68)."""
    return Sounder.make(645.0, 2760.0, 0.25, opd_max=2.0, fwhm_gauss=fwhm_gauss)


def IASINGInstrument(fwhm_gauss: float = 0.25) -> Sounder:
    """IASI-NG (sounder.jl:86-89)."""
    return Sounder.make(645.0, 2760.0, 0.125, opd_max=4.0, fwhm_gauss=fwhm_gauss)


def CrISInstrument(band: str = "full", fwhm_gauss: float = 0.0) -> Sounder:
    """CrIS FSR (sounder.jl:102-110)."""
    bands = {"full": (650.0, 2550.0), "lwir": (650.0, 1095.0), "mwir": (1210.0, 1750.0),
             "swir": (2155.0, 2550.0)}
    if band not in bands:
        raise ValueError(f"CrIS band must be full, lwir, mwir, or swir; got {band}")
    lo, hi = bands[band]
    return Sounder.make(lo, hi, 0.625, opd_max=0.8, fwhm_gauss=fwhm_gauss)


def MTGIRSInstrument(band: str = "full", fwhm_gauss: float = 0.0) -> Sounder:
    """MTG-IRS (sounder.jl:121-128)."""
    bands = {"full": (700.0, 2175.0), "lwir": (700.0, 1210.0), "mwir": (1600.0, 2175.0)}
    if band not in bands:
        raise ValueError(f"MTG-IRS band must be full, lwir, or mwir; got {band}")
    lo, hi = bands[band]
    return Sounder.make(lo, hi, 0.625, opd_max=0.8, fwhm_gauss=fwhm_gauss)


@dataclass
class ViewingGeometry:
    """Viewing geometry (src/Sensor/geometry.jl:18-34)."""
    zenith_angle: float
    azimuth_angle: float = 0.0
    mu: float = 1.0

    @staticmethod
    def from_zenith(zenith_angle: float, azimuth_angle: float = 0.0) -> "ViewingGeometry":
        if not (0.0 <= zenith_angle < 90.0):
            raise ValueError(f"Viewing zenith angle must be in [0, 90); got {zenith_angle}°")
        return ViewingGeometry(float(zenith_angle), float(azimuth_angle),
                               math.cos(math.radians(zenith_angle)))


def nadir_geometry() -> ViewingGeometry:
    return ViewingGeometry.from_zenith(0.0)


def scan_angle_to_local_zenith(scan_angle: float, h_sat_km: float = 817.0,
                               r_earth_km: float = 6371.0) -> float:
    """Law-of-sines local zenith angle (src/Sensor/geometry.jl:92-100)."""
    s = (r_earth_km + h_sat_km) / r_earth_km * math.sin(math.radians(abs(scan_angle)))
    if s > 1.0:
        raise ValueError(f"scan_angle {scan_angle}° exceeds limb angle from h={h_sat_km} km")
    return math.degrees(math.asin(s))


def airmass_factor(geom: ViewingGeometry) -> float:
    """Plane-parallel airmass 1/mu (src/Sensor/geometry.jl:48)."""
    return 1.0 / geom.mu


def iasi_scan_angles():
    """The 30 nominal IASI scan angles (src/Sensor/geometry.jl:61-68)."""
    half = np.linspace(48.33 / 15 / 2, 48.33, 15)
    return np.concatenate([-half[::-1], half])


def iasi_zenith_angles(h_sat_km: float = 817.0):
    """Local zenith angles of the 30 IASI scan positions, sign of the scan angle
    kept (src/Sensor/geometry.jl:110-113)."""
    return np.array([math.copysign(scan_angle_to_local_zenith(s, h_sat_km=h_sat_km), s)
                     for s in iasi_scan_angles()])


class WavenumberGrid:
    """Uniform FP64 grid with Julia float-range semantics (wavenumber_grid.jl:23-27).

    When the step is a rational n/d and nu_min*d is an integer (every grid on
    this path), node i is the correctly rounded (nu_min*d + i*n)/d as Julia's
    `nu_min:step:nu_max` produces."""

    def __init__(self, nu_min: float, nu_max: float, dnu: float):
        self.nu_min = float(nu_min)
        self.dnu = float(dnu)
        num, den = _rational(self.dnu)
        an = self.nu_min * den if den else None
        if den and abs(an - round(an)) <= 1e-7 * max(1.0, abs(an)):
            bn = float(nu_max) * den
            self._rat = (float(round(an)), float(num), float(den))
            self.n = int(math.floor((round(bn) - round(an)) / num)) + 1
        else:
            self._rat = None
            self.n = int(math.floor((nu_max - nu_min) / dnu * (1 + 1e-14))) + 1

    @property
    def nu(self) -> np.ndarray:
        i = np.arange(self.n, dtype=np.float64)
        if self._rat:
            an, sn, den = self._rat
            return (an + i * sn) / den
        return self.nu_min + i * self.dnu

    def __len__(self):
        return self.n

    def __repr__(self):
        nu = self.nu
        return f"WavenumberGrid({nu[0]}–{nu[-1]} cm⁻¹, Δν={self.dnu} cm⁻¹, n={self.n})"


def _rational(x: float):
    """Continued-fraction convergent n/d with n/d == x to 2 ulp, d <= 1e9."""
    h0, h1, k0, k1 = 0.0, 1.0, 1.0, 0.0
    y = x
    for _ in range(40):
        a = math.floor(y)
        h0, h1 = h1, a * h1 + h0
        k0, k1 = k1, a * k1 + k0
        if k1 > 1e9:
            return 0.0, 0.0
        if abs(h1 / k1 - x) <= 2.220446049250313e-16 * abs(x):
            return h1, k1
        fr = y - a
        if fr == 0:
            break
        y = 1.0 / fr
    return 0.0, 0.0


def wavenumber_grid(nu_min: float, nu_max: float, dnu: float) -> WavenumberGrid:
    return WavenumberGrid(nu_min, nu_max, dnu)


def sounder_grid(s: Sounder) -> WavenumberGrid:
    return WavenumberGrid(s.nu_min, s.nu_max, s.dnu)


# ---------------------------------------------------------------------------
# Retrieval state layout (src/Estimation/state_vector.jl:45-118, 253-300)

@dataclass(frozen=True)
class FullProfile:
    """One (log-)VMR per level (state_vector.jl:48-50)."""
    species: GasSpecies


@dataclass(frozen=True)
class ColumnScale:
    """One multiplicative log-scale of the whole profile (state_vector.jl:60-62)."""
    species: GasSpecies


@dataclass(frozen=True, eq=False)
class PartialColumns:
    """Log-VMR on a basis [n_levels x m] (state_vector.jl:87-104)."""
    species: GasSpecies
    basis: np.ndarray

    def __post_init__(self):
        b = np.asarray(self.basis, np.float64)
        if b.ndim != 2 or b.shape[1] < 1:
            raise ValueError("PartialColumns basis must be a non-empty matrix")
        object.__setattr__(self, "basis", b)


def _normalize_param(entry):
    if isinstance(entry, (FullProfile, ColumnScale, PartialColumns)):
        return entry
    return FullProfile(GasSpecies(entry))


@dataclass
class VMRBlock:
    species: GasSpecies
    basis: Optional[np.ndarray]      # None = identity (full profile)
    m: int
    range: range


class StateVectorSpec:
    """StateVectorSpec(n_levels, species; include_temperature, include_tsfc,
    include_emissivity, log_vmr) (state_vector.jl:253-300): the state is
    [T levels][one block per retrieved gas][T_sfc][emissivity], 0-based ranges."""

    def __init__(self, n_levels: int, species=(), *, include_temperature: bool = True,
                 include_tsfc: bool = True, include_emissivity: bool = True, log_vmr: bool = True):
        if n_levels <= 0:
            raise ValueError("n_levels must be positive")
        self.n_levels = int(n_levels)
        self.include_temperature = bool(include_temperature)
        self.include_tsfc = bool(include_tsfc)
        self.include_emissivity = bool(include_emissivity)
        self.log_vmr = bool(log_vmr)
        idx = 0
        self.temp_range = range(0, 0)
        if include_temperature:
            self.temp_range = range(idx, idx + n_levels)
            idx += n_levels
        self.vmr_blocks = []
        for entry in species:
            p = _normalize_param(entry)
            if isinstance(p, FullProfile):
                B, m = None, n_levels
            elif isinstance(p, ColumnScale):
                B, m = np.ones((n_levels, 1)), 1
            else:
                if p.basis.shape[0] != n_levels:
                    raise ValueError(f"PartialColumns basis has {p.basis.shape[0]} levels, spec expects {n_levels}")
                B, m = p.basis, p.basis.shape[1]
            self.vmr_blocks.append(VMRBlock(p.species, B, m, range(idx, idx + m)))
            idx += m
        if len({b.species for b in self.vmr_blocks}) != len(self.vmr_blocks):
            raise ValueError("a species may appear only once in the state vector")
        self.tsfc_index = -1
        if include_tsfc:
            self.tsfc_index = idx
            idx += 1
        self.emis_index = -1
        if include_emissivity:
            self.emis_index = idx
            idx += 1
        self.n = idx

    @property
    def species(self):
        return [b.species for b in self.vmr_blocks]

    @property
    def vmr_ranges(self):
        return [(b.species, b.range) for b in self.vmr_blocks]

    def __len__(self):
        return self.n

    def vmr_range(self, s) -> range:
        for b in self.vmr_blocks:
            if b.species == GasSpecies(s):
                return b.range
        raise KeyError(f"species {GasSpecies(s).name} is not in the state vector")

    def __repr__(self):
        blocks = [f"T({self.n_levels})"] if self.include_temperature else []
        for b in self.vmr_blocks:
            if b.basis is None:
                blocks.append(f"{'logVMR' if self.log_vmr else 'VMR'}_{b.species.name}({self.n_levels})")
            else:
                blocks.append(f"col_{b.species.name}({b.m})")
        if self.include_tsfc:
            blocks.append("T_sfc")
        if self.include_emissivity:
            blocks.append("ε")
        return f"StateVectorSpec(n={self.n}: " + " + ".join(blocks) + ")"

    def labels(self):
        """state_labels (state_vector.jl:335-350), 1-based level indices."""
        out = [f"T[{i + 1}]" for i in self.temp_range]
        pre = "logVMR" if self.log_vmr else "VMR"
        for b in self.vmr_blocks:
            if b.basis is None:
                out += [f"{pre}_{b.species.name}[{i + 1}]" for i in range(self.n_levels)]
            elif b.m == 1:
                out.append(f"logVMRscale_{b.species.name}")
            else:
                out += [f"logVMRscale_{b.species.name}[{m + 1}]" for m in range(b.m)]
        if self.include_tsfc:
            out.append("T_sfc")
        if self.include_emissivity:
            out.append("emissivity")
        return out


VMR_FLOOR = 1e-30   # state_vector.jl:26


def pack_state(spec: StateVectorSpec, prof: AtmosphericProfile, T_sfc: Optional[float] = None,
               eps_sfc: float = 1.0) -> np.ndarray:
    """pack_state (state_vector.jl:350-370)."""
    if prof.n_levels != spec.n_levels:
        raise ValueError(f"profile has {prof.n_levels} levels, spec expects {spec.n_levels}")
    x = np.zeros(spec.n)
    if spec.include_temperature:
        x[spec.temp_range.start:spec.temp_range.stop] = prof.temperature
    for b in spec.vmr_blocks:
        if b.species not in prof.vmr:
            raise ValueError(f"profile is missing retrieved species {b.species.name}")
        if b.basis is None:
            v = prof.vmr[b.species]
            x[b.range.start:b.range.stop] = np.log(np.maximum(v, VMR_FLOOR)) if spec.log_vmr else v
    if spec.include_tsfc:
        x[spec.tsfc_index] = prof.temperature[0] if T_sfc is None else T_sfc
    if spec.include_emissivity:
        x[spec.emis_index] = eps_sfc
    return x


def unpack_state(spec: StateVectorSpec, x, base_prof: AtmosphericProfile):
    """unpack_state (state_vector.jl:380-403) -> (profile, T_sfc, eps_sfc): fields outside
    the state come from base_prof; T_sfc falls back to the lowest level temperature and
    eps_sfc to 1.0 when they are not retrieved."""
    x = np.asarray(x, np.float64)
    if len(x) != spec.n:
        raise ValueError(f"state vector length {len(x)} ≠ spec.n {spec.n}")
    if spec.include_temperature:
        T = x[spec.temp_range.start:spec.temp_range.stop].copy()
    else:
        T = base_prof.temperature.copy()
    vmr = {s: v.copy() for s, v in base_prof.vmr.items()}
    for b in spec.vmr_blocks:
        xv = x[b.range.start:b.range.stop]
        if b.basis is None:
            vmr[b.species] = np.exp(xv) if spec.log_vmr else xv.copy()
        else:
            vmr[b.species] = base_prof.vmr[b.species] * np.exp(b.basis @ xv)
    prof = AtmosphericProfile(base_prof.pressure.copy(), T, base_prof.altitude.copy(), vmr)
    T_sfc = float(x[spec.tsfc_index]) if spec.include_tsfc else float(T[0])
    eps_sfc = float(x[spec.emis_index]) if spec.include_emissivity else 1.0
    return prof, T_sfc, eps_sfc


@dataclass
class Jacobian:
    """Jacobian struct (src/Estimation/jacobian_fd.jl:43-50): K is [n_y x n_x]."""
    K: np.ndarray
    spec: StateVectorSpec
    nu: np.ndarray
    x0: np.ndarray
    y0: np.ndarray
    observable: str
