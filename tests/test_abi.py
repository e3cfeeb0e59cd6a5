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


def declared_arity():
    """Parameter count of every function declared in irs.h."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"\b(irs_[a-z_0-9]+)\s*\(([^;{)]*)\)\s*;", text):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_ctypes_bindings_match_the_header_arity():
    """Every ctypes prototype that declares argtypes has as many as the header's
    declaration, so a call with the wrong count fails in ctypes, not in the library."""
    from irsounderlbl.api import load_library
    lib = load_library()
    ar = declared_arity()
    assert len(ar) >= 40
    checked = 0
    for name, n in ar.items():
        at = getattr(lib, name).argtypes
        if at is not None:
            assert len(at) == n, (name, len(at), n)
            checked += 1
    assert checked >= 40


def test_emulator_correction_call_sites_match_the_binding():
    """The emulator's calls of irs_emu_set_correction reach the library (which
    rejects the null handle) with the argument list the binding declares."""
    from irsounderlbl.api import IRSError, load_library
    from irsounderlbl.emulator import Calibration, ChannelODEmulator

    class Fake:
        lib = load_library()
        h = None

    with pytest.raises(IRSError):
        ChannelODEmulator._set_correction(Fake(), None)
    cal = Calibration(np.arange(4), [(0, 2), (2, 2)], 1.0, 0.05, np.zeros(4), np.zeros(4), np.zeros(1), np.zeros(1), 1,
                      W=np.zeros((4, 4), np.float32), lbo_W=np.zeros(8, np.float32), w_raw=1.0, w_lbo=2.0)
    with pytest.raises(IRSError):
        ChannelODEmulator._set_correction(Fake(), cal)


def test_emulator_file_loader_rejects_bad_files(tmp_path):
    """irs_emu_load checks the whole file before it touches a device: a missing file is
    IRS_ENOTFOUND, a wrong magic, inconsistent dimensions or a truncated file IRS_EINVAL."""
    from irsounderlbl.api import IRS_EINVAL, IRS_ENOTFOUND, load_library
    lib = load_library()

    def load(path):
        hm, he, thr, nb = C.c_void_p(), C.c_void_p(), C.c_double(), C.c_int()
        c0 = np.zeros(4, np.int64)
        bn = np.zeros(4, np.int32)
        rc = lib.irs_emu_load(str(path).encode(), 0, 64, C.byref(hm), C.byref(he), C.byref(thr), 4,
                              c0.ctypes.data, bn.ctypes.data, C.byref(nb))
        assert not hm.value and not he.value
        return rc

    assert load(tmp_path / "none.bin") == IRS_ENOTFOUND
    p = tmp_path / "bad.bin"
    p.write_bytes(b"NOTANEMU" + bytes(64))
    assert load(p) == IRS_EINVAL
    # n_in n_hidden n_out L n_pc n_ch n_feat n_local n_ctx n_vec n_blocks corrected
    hdr = np.array([13, 32, 8, 3, 8, 5, 21, 13, 8, 11, 1, 0], np.int32)
    bad = hdr.copy()
    bad[2] = 7                                           # n_out must equal n_pc
    p.write_bytes(b"IRSEMU01" + bad.tobytes() + np.float32(1.0).tobytes() + np.int64(10).tobytes())
    assert load(p) == IRS_EINVAL
    p.write_bytes(b"IRSEMU01" + hdr.tobytes() + np.float32(1.0).tobytes() + np.int64(10).tobytes() +
                  np.zeros(10, np.float32).tobytes())     # the arrays after the parameters are missing
    assert load(p) == IRS_EINVAL
