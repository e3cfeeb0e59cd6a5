"""Batched forward model: one call over a scan line of perturbed profiles
(per-profile DPTMIN active sets), the persistent heaviest-first line-sum
kernel, hipGraph capture/replay and the autotuner.

Bars: results of one schedule at fixed tile_points are bit-identical whether
the profiles run alone or batched, in one group or several, through the grid
kernel or the persistent kernel, eagerly or from a replayed graph; every
schedule (including the autotuned one) stays within 0.005 K of the FP64
oracle."""
import os

import numpy as np
import pytest

import refcpu as o
from irsounderlbl import synthetic as syn
from irsounderlbl.api import Context, forward_model
from irsounderlbl.types import GasSpecies, Sounder, ViewingGeometry, filter_linelist

pytestmark = pytest.mark.gpu

LO, HI = 700.0, 712.0


@pytest.fixture(scope="module")
def scene():
    lls = syn.linelists_config1()
    sub = {k: filter_linelist(v, LO - 42.0, HI + 42.0) for k, v in lls.items()}
    snd = Sounder.make(LO, HI, 0.25, opd_max=2.0, fwhm_gauss=0.5)
    profs = syn.perturbed_profiles(5, seed=2002)
    return sub, snd, profs


def _run(ctx, scene, idx=None, **kw):
    sub, snd, profs = scene
    sel = [profs[i] for i in (idx if idx is not None else range(len(profs)))]
    p = [x[0] for x in sel]
    geoms = [ViewingGeometry(float(np.degrees(np.arccos(x[2]))), 0.0, x[2]) for x in sel]
    return forward_model(p, sub, sounder=snd, geom=geoms, T_sfc=[x[1] for x in sel],
                         line_mixing="vpy", ctx=ctx, **kw, apply_continuum=False)


@pytest.fixture()
def bctx():
    c = Context(0)
    yield c
    c.close()


def test_batch_equals_single_and_oracle(bctx, scene):
    bctx.set_schedule(4, persistent=False)
    _, Rb, BTb = _run(bctx, scene)
    st = bctx.stats()
    assert st["tile_points"] == 4 and st["persistent"] == 0
    for q in range(Rb.shape[0]):
        _, R1, BT1 = _run(bctx, scene, [q])
        assert np.array_equal(R1[0], Rb[q]) and np.array_equal(BT1[0], BTb[q])
    sub, snd, profs = scene
    for q in (0, 3):
        prof, T_sfc, mu = profs[q]
        _, bto, _ = o.forward(prof, sub, snd, line_mixing=1, T_sfc=T_sfc, mu=mu)
        assert np.max(np.abs(BTb[q] - bto)) < 0.005


def test_groups_persistent_and_tile_points(bctx, scene, monkeypatch):
    bctx.set_schedule(8, persistent=False)
    _, R0, BT0 = _run(bctx, scene)
    monkeypatch.setenv("IRS_PROFILE_GROUP", "2")
    _, R1, BT1 = _run(bctx, scene)
    assert bctx.stats()["linesum_launches"] == 3
    assert np.array_equal(R0, R1) and np.array_equal(BT0, BT1)
    monkeypatch.delenv("IRS_PROFILE_GROUP")
    bctx.set_schedule(8, persistent=True)
    _, R2, BT2 = _run(bctx, scene)
    assert bctx.stats()["persistent"] == 1
    assert np.array_equal(R0, R2) and np.array_equal(BT0, BT2)
    for P in (4, 16):
        bctx.set_schedule(P, persistent=(P == 16))
        _, _, BTp = _run(bctx, scene)
        assert np.max(np.abs(BTp - BT0)) < 1e-3


def test_graph_replays_equal_eager(bctx, scene):
    bctx.set_schedule(4, persistent=True)
    _, Re, BTe = _run(bctx, scene)
    bctx.set_schedule(4, persistent=True, use_graph=True)
    for rep in range(3):
        _, Rg, BTg = _run(bctx, scene)
        assert bctx.stats()["graph_launches"] == 1
        assert np.array_equal(Rg, Re) and np.array_equal(BTg, BTe), rep
        # eager work on the same stream between replays
        bctx.set_schedule(4, persistent=True)
        _, Rx, _ = _run(bctx, scene, [1, 2])
        assert np.array_equal(Rx, Re[1:3])
        bctx.set_schedule(4, persistent=True, use_graph=True)
    # a different batch shape captures a second graph; the first still replays
    _, R2, _ = _run(bctx, scene, [0, 4])
    assert np.array_equal(R2, Re[[0, 4]])
    _, R3, _ = _run(bctx, scene)
    assert np.array_equal(R3, Re)


def test_autotune_keeps_parity(bctx, scene):
    bctx.set_schedule(4, autotune=True)
    _, Ra, BTa = _run(bctx, scene)
    st = bctx.stats()
    assert st["autotuned"] == 1 and st["tile_points"] in (4, 8, 16)
    P, pers = st["tile_points"], st["persistent"]
    _, Rb, BTb = _run(bctx, scene)
    assert bctx.stats()["autotuned"] == 0          # cached per shape
    assert np.array_equal(Ra, Rb)
    bctx.set_schedule(P, persistent=bool(pers))
    _, Rc, _ = _run(bctx, scene)
    assert np.array_equal(Ra, Rc)
    sub, snd, profs = scene
    prof, T_sfc, mu = profs[2]
    _, bto, _ = o.forward(prof, sub, snd, line_mixing=1, T_sfc=T_sfc, mu=mu)
    assert np.max(np.abs(BTa[2] - bto)) < 0.005


def test_autotune_persists_in_tune_file(scene, tmp_path, monkeypatch):
    """With IRS_TUNE_FILE set, the first context times the candidates and appends the
    result (keyed by device, grid and line-list content); a second context reuses it
    without timing and produces the same spectra; an unset variable writes nothing."""
    path = tmp_path / "tune.tsv"
    monkeypatch.setenv("IRS_TUNE_FILE", str(path))
    a = Context(0)
    a.set_schedule(4, autotune=True)
    _, Ra, _ = _run(a, scene)
    sa = a.stats()
    assert sa["autotuned"] == 1
    lines = path.read_text().splitlines()
    assert len(lines) == 1 and lines[0].split("\t")[1:] == [str(sa["tile_points"]), str(sa["persistent"])]
    b = Context(0)
    b.set_schedule(4, autotune=True)
    _, Rb, _ = _run(b, scene)
    sb = b.stats()
    assert sb["autotuned"] == 0 and (sb["tile_points"], sb["persistent"]) == (sa["tile_points"], sa["persistent"])
    assert np.array_equal(Ra, Rb)
    assert len(path.read_text().splitlines()) == 1
    a.close()
    b.close()
