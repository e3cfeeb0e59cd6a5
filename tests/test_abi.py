"""The C ABI library loads and exports every symbol include/irs.h declares
(no GPU needed); host-only entry points behave like the reference."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "irs.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(irs_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from irsounderlbl.api import EXPORTS, load_library
    lib = load_library()
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(EXPORTS)


def test_host_entry_points():
    from irsounderlbl.api import forward_dims, load_library, brightness_temperature
    from irsounderlbl.types import IASIInstrument, Sounder
    lib = load_library()
    assert lib.irs_version() == 1
    lo, d, n, nch = forward_dims(IASIInstrument())
    assert (lo, d, n, nch) == (629.0, 0.001, 2147001, 8461)
    lo, d, n, nch = forward_dims(Sounder.make(645.0, 800.0, 0.25, opd_max=2.0, fwhm_gauss=0.5))
    assert (n, nch) == (187001, 621)
    nu = np.array([1000.0, 700.0])
    B = 1.191042953e-5 * nu ** 3 / np.expm1(1.4387769 * nu / 280.0)
    assert np.allclose(brightness_temperature(nu, B), 280.0, rtol=1e-12)
    assert brightness_temperature(np.array([700.0]), np.array([0.0]))[0] == 0.0
    h = C.c_int64()
    assert lib.irs_ils_half(0.001, C.byref(h)) == 0 and h.value == 16000


def test_errors_without_device_or_bad_input():
    from irsounderlbl.api import IRSError, Context, load_library, _options
    lib = load_library()
    with pytest.raises(ValueError):
        _options(apodization="bogus")
    with pytest.raises(ValueError):
        _options(T_method="bogus")
    with pytest.raises(ValueError):
        _options(source_function="bogus")
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if not has_gpu:
        with pytest.raises(IRSError):
            Context(0)
