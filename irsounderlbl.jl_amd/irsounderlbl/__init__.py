"""IRSounderLBL on MI355X: host-side mirror of the reference's public interface.

`forward_model(prof, linelists; sounder, geom, ...)` and the layer-level
operators call hand-written gfx950 HIP kernels through the C ABI declared in
include/irs.h (library: csrc/libirs.so next to this package).  There is no CPU
fallback: if the library or a GPU is missing, calls raise `IRSError`.
"""
from .types import (AtmosphericProfile, ColumnScale, CrISInstrument, FullProfile, GasSpecies,
                    HITRANLinelist, IASIInstrument, IASINGInstrument, Jacobian, MTGIRSInstrument,
                    PartialColumns, Sounder, StateVectorSpec, ViewingGeometry, WavenumberGrid,
                    airmass_factor, filter_linelist, iasi_scan_angles, iasi_zenith_angles,
                    nadir_geometry, pack_state, scan_angle_to_local_zenith, sounder_grid,
                    wavenumber_grid)

__all__ = ["AtmosphericProfile", "ColumnScale", "CrISInstrument", "FullProfile", "GasSpecies",
           "HITRANLinelist", "IASIInstrument", "IASINGInstrument", "Jacobian", "MTGIRSInstrument",
           "PartialColumns", "Sounder", "StateVectorSpec", "pack_state",
           "ViewingGeometry", "WavenumberGrid", "airmass_factor", "filter_linelist",
           "iasi_scan_angles", "iasi_zenith_angles", "nadir_geometry", "scan_angle_to_local_zenith", "sounder_grid", "wavenumber_grid"]

from .api import (IRSError, Context, analytic_jacobian, apply_ils_channels,  # noqa: E402
                  brightness_temperature, co2_cia, co2_continuum, compute_voigt_cross_sections,
                  compute_voigt_lm_cross_sections, forward_model, h2o_continuum, ils_kernel,
                  n2_cia, o2_cia, planck_radiance, schwarzschild_rte)

__all__ += ["IRSError", "Context", "analytic_jacobian", "apply_ils_channels",
            "brightness_temperature", "co2_cia", "co2_continuum", "compute_voigt_cross_sections",
            "compute_voigt_lm_cross_sections", "forward_model", "h2o_continuum", "ils_kernel",
            "n2_cia", "o2_cia", "planck_radiance", "schwarzschild_rte"]

from .hitran import (HITRANLine, clear_linelist_cache, linelist_cache_dir, load_hitran_par,  # noqa: E402
                     load_linelist, parse_hitran_par, set_linelist_cache_dir, write_hitran_par)

__all__ += ["HITRANLine", "clear_linelist_cache", "linelist_cache_dir", "load_hitran_par", "load_linelist",
            "parse_hitran_par", "set_linelist_cache_dir", "write_hitran_par"]

from .retrieval import (RetrievalResult, apodized_measurement_covariance, build_sa, dB_dT,  # noqa: E402
                        exclude_channels, optimal_estimation, scene_measurement_covariance, scene_nedt)
from .types import unpack_state  # noqa: E402

__all__ += ["RetrievalResult", "apodized_measurement_covariance", "build_sa", "dB_dT", "exclude_channels",
            "optimal_estimation", "scene_measurement_covariance", "scene_nedt", "unpack_state"]
