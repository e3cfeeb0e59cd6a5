"""Viewing geometry of the host mirror (src/Sensor/geometry.jl), at the bars of the
reference's "Viewing Geometry" testset (test/runtests.jl:609-637), and against the
oracle's law-of-sines restatement on a sweep of scan angles."""
import math

import numpy as np
import pytest

import refcpu
from irsounderlbl.types import (ViewingGeometry, airmass_factor, iasi_scan_angles, iasi_zenith_angles,
                                nadir_geometry, scan_angle_to_local_zenith)

L = refcpu.lib()


def test_viewing_geometry_constructor():
    nadir = nadir_geometry()
    assert nadir.mu == pytest.approx(1.0)
    assert airmass_factor(nadir) == pytest.approx(1.0)
    g48 = ViewingGeometry.from_zenith(48.33)
    assert g48.mu == pytest.approx(math.cos(math.radians(48.33)), rel=1e-10)
    assert airmass_factor(g48) == pytest.approx(1.0 / math.cos(math.radians(48.33)), rel=1e-10)
    for bad in (90.0, -1.0):
        with pytest.raises(ValueError):
            ViewingGeometry.from_zenith(bad)


def test_scan_angles_and_local_zenith():
    ang = iasi_scan_angles()
    assert len(ang) == 30
    assert np.max(np.abs(ang)) == pytest.approx(48.33, rel=1e-2)
    assert scan_angle_to_local_zenith(0.0) == pytest.approx(0.0, abs=1e-12)
    assert scan_angle_to_local_zenith(48.33) == pytest.approx(57.46, abs=0.05)
    assert scan_angle_to_local_zenith(-48.33) == pytest.approx(57.46, abs=0.05)
    with pytest.raises(ValueError):
        scan_angle_to_local_zenith(80.0)
    zen = iasi_zenith_angles()
    assert len(zen) == 30
    assert np.all(np.abs(zen) > np.abs(ang))
    assert zen[0] == pytest.approx(-zen[-1], abs=1e-12)


def test_local_zenith_matches_oracle_sweep():
    """Product and oracle agree over the whole scan range and at other altitudes;
    both refuse angles beyond the limb (the oracle returns -1 there)."""
    for h in (817.0, 705.0, 35786.0):
        limb = math.degrees(math.asin(6371.0 / (6371.0 + h)))
        for a in np.linspace(-limb * 0.999, limb * 0.999, 41):
            assert scan_angle_to_local_zenith(a, h_sat_km=h) == pytest.approx(
                L.orc_scan_to_local_zenith(a, h, 6371.0), abs=1e-12)
        assert L.orc_scan_to_local_zenith(limb * 1.01, h, 6371.0) < 0
        with pytest.raises(ValueError):
            scan_angle_to_local_zenith(limb * 1.01, h_sat_km=h)
