"""Host logic of the emulator's Jacobian chain (no GPU): the level -> feature
derivatives recompute only the two layers a level bounds, which must give the
same central differences as recomputing the whole profile."""
import numpy as np

from irsounderlbl import synthetic as syn
from irsounderlbl.api import layer_properties
from irsounderlbl.emulator import (N_LOCAL, ProfileContext, _feature_level_jacobian, _features_from_layers,
                                   _layer_quantities, layer_features)
from irsounderlbl.types import AtmosphericProfile, GasSpecies


def _full_profile_fd(p, mu, hT=1e-2, hq=1e-3):
    def feats(T=None, vmr=None):
        pp = AtmosphericProfile(p.pressure, p.temperature if T is None else T, p.altitude,
                                p.vmr if vmr is None else vmr)
        return layer_features(pp, mu), layer_properties(pp)["T_ave"]
    nl = p.n_levels
    L = nl - 1
    dXdT, dAdT, dXdq = np.zeros((nl, L, N_LOCAL)), np.zeros((nl, L)), np.zeros((nl, L, N_LOCAL))
    for j in range(nl):
        Tp, Tm = p.temperature.copy(), p.temperature.copy()
        Tp[j] += hT
        Tm[j] -= hT
        (xp, ap), (xm, am) = feats(T=Tp), feats(T=Tm)
        dXdT[j], dAdT[j] = (xp - xm) / (2 * hT), (ap - am) / (2 * hT)
        vp, vm = dict(p.vmr), dict(p.vmr)
        q = p.vmr[GasSpecies.H2O]
        qp, qm = q.copy(), q.copy()
        qp[j] *= np.exp(hq)
        qm[j] *= np.exp(-hq)
        vp[GasSpecies.H2O], vm[GasSpecies.H2O] = qp, qm
        dXdq[j] = (feats(vmr=vp)[0] - feats(vmr=vm)[0]) / (2 * hq)
    return dXdT, dAdT, dXdq


def test_level_local_feature_jacobian_equals_full_recompute():
    sc = syn.perturbed_profiles(16, seed=7, sigma_T=4.0, sigma_h2o=0.6, sigma_o3=0.2, co2_range=(390e-6, 430e-6),
                                sigma_lnq=0.05)
    ctx = ProfileContext.fit(sc)
    p, ts, mu = sc[3]
    dXdT, dAdT, dXdq, dXdTs = _feature_level_jacobian(p, ts, mu, ctx, "logp_at_pcg")
    rT, rA, rq = _full_profile_fd(p, mu)
    assert np.array_equal(dXdT[:, :, :N_LOCAL], rT)
    assert np.array_equal(dAdT, rA)
    assert np.array_equal(dXdq[:, :, :N_LOCAL], rq)
    assert np.abs(rq).max() > 0.1 and np.abs(rT).max() > 0.1          # both chains are live
    # the profile context is linear: its columns are the projection's
    nl = p.n_levels
    np.testing.assert_allclose(dXdT[1, 0, N_LOCAL:], ctx.proj[:, 1], rtol=0, atol=0)
    np.testing.assert_allclose(dXdq[2, 0, N_LOCAL:], ctx.proj[:, nl + 2], rtol=0, atol=0)


def test_vectorised_features_equal_per_profile_features():
    sc = syn.perturbed_profiles(4, seed=9, sigma_T=4.0, sigma_h2o=0.6, sigma_o3=0.2)
    Q = np.stack([np.stack(_layer_quantities(layer_properties(p), p.n_levels - 1)) for p, _, _ in sc])   # [4, 6, L]
    X = _features_from_layers(*(Q[:, i] for i in range(6)), 0.8)
    for q, (p, _, _) in enumerate(sc):
        assert np.array_equal(X[q], layer_features(p, 0.8))
