"""IRSounderLBL on MI355X: host-side mirror of the reference's public interface.

`forward_model(prof, linelists; sounder, geom, ...)` and the layer-level
operators call hand-written gfx950 HIP kernels through the C ABI declared in
include/irs.h (library: csrc/libirs.so next to this package).  There is no CPU
fallback: if the library or a GPU is missing, calls raise `IRSError`.
"""
from .types import (AtmosphericProfile, CrISInstrument, GasSpecies, HITRANLinelist,
                    IASIInstrument, IASINGInstrument, MTGIRSInstrument, Sounder,
                    ViewingGeometry, WavenumberGrid, filter_linelist, iasi_scan_angles,
                    nadir_geometry, scan_angle_to_local_zenith, sounder_grid, wavenumber_grid)

__all__ = ["AtmosphericProfile", "CrISInstrument", "GasSpecies", "HITRANLinelist",
           "IASIInstrument", "IASINGInstrument", "MTGIRSInstrument", "Sounder",
           "ViewingGeometry", "WavenumberGrid", "filter_linelist", "iasi_scan_angles",
           "nadir_geometry", "scan_angle_to_local_zenith", "sounder_grid", "wavenumber_grid"]

try:  # the GPU operators need the built library; importing the types never does
    from .api import (IRSError, Context, apply_ils, brightness_temperature,  # noqa: F401
                      compute_voigt_cross_sections, compute_voigt_lm_cross_sections,
                      forward_model, ils_kernel, planck_radiance, schwarzschild_rte)
    __all__ += ["IRSError", "Context", "apply_ils", "brightness_temperature",
                "compute_voigt_cross_sections", "compute_voigt_lm_cross_sections",
                "forward_model", "ils_kernel", "planck_radiance", "schwarzschild_rte"]
except ImportError:  # pragma: no cover - api.py not yet present during bootstrap
    pass
