"""Learned emulator of per-layer channel optical depths, with LBL spot check
and fallback.  No counterpart in the reference.

Pipeline (all numeric work on the GPU through libirs except the PCA, which is
an FP64 host SVD):
  1. training target: for every layer k of GPU line-by-line spectra along
     each scene's viewing path (`channel_optical_depths`), the layer channel
     transmittance t_k = exp(-tau_ch_k) (ratio of the ILS-convolved
     level-to-space transmittances) and the CIM weight g_k, with which the
     channel RTE reproduces the ILS-convolved LBL radiance of the scene (to
     2e-3 K, the variation of Planck across the ILS);
  2. PCA: [t_k, g_k] minus its per-layer training mean, projected on the n_pc
     leading right singular vectors (randomized range finder + FP64 SVD);
  3. MLP [ln p, T, ln q_H2O, ln q_O3, q_CO2 (ppm), 1/mu, and the overlying
     column: ln H2O and ln O3 amounts above, T of the layer above]
     -> 3 x 512 GELU -> n_pc,
     trained with the BF16 MFMA kernels (FP32 accumulation) and Adam on the
     coefficients (one global scale, so the loss is the transmittance MSE);
  4. inference: MLP -> decode -> channel-space RTE (irs_emu_forward);
  5. per-scene LBL spot check on blocks of channels; scenes whose largest
     |BT_emulator - BT_LBL| exceeds the budget are recomputed with full LBL.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import api
from .api import (IRS_OK, Context, IRSError, _p, channel_optical_depths, forward_model,
                  layer_properties, load_library)
from .types import GasSpecies, HITRANLinelist, Sounder, ViewingGeometry, filter_linelist, sounder_grid

N_FEATURES = 9
FEATURE_NAMES = ("ln p_mid [hPa]", "T_ave [K]", "ln vmr H2O", "ln vmr O3", "vmr CO2 [ppm]", "1/mu",
                 "ln H2O column above [hPa]", "ln O3 column above [hPa]", "T_ave of the layer above [K]")


class MLP:
    """Device-resident MLP n_in -> 3 x n_hidden (GELU) -> n_out (irs_mlp_*)."""

    def __init__(self, n_in: int = N_FEATURES, n_hidden: int = 512, n_out: int = 256,
                 max_batch: int = 4096, seed: int = 0, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.irs_mlp_create(device, n_in, n_hidden, n_out, max_batch, seed, C.byref(h))
        if rc != IRS_OK:
            raise IRSError(rc, f"irs_mlp_create failed ({rc})")
        self.h = h
        self.n_in, self.n_hidden, self.n_out, self.max_batch = n_in, n_hidden, n_out, max_batch
        self.n_params = self.lib.irs_mlp_n_params(h)
        self.n_in_pad = self.lib.irs_mlp_input_pad(h)

    def _check(self, rc):
        if rc != IRS_OK:
            msg = self.lib.irs_mlp_last_error(self.h).decode(errors="replace")
            raise (ValueError(msg) if rc == api.IRS_EINVAL else IRSError(rc, msg))

    def close(self):
        if getattr(self, "h", None):
            self.lib.irs_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        self._check(self.lib.irs_mlp_get_params(self.h, _p(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.n_params:
            raise ValueError("parameter vector has the wrong length")
        self._check(self.lib.irs_mlp_set_params(self.h, _p(p)))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((len(x), self.n_out), np.float32)
        self._check(self.lib.irs_mlp_forward(self.h, _p(x), len(x), _p(y)))
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = np.zeros(1, np.float32)
        self._check(self.lib.irs_mlp_loss_grad(self.h, _p(x), _p(t), len(x), _p(g), _p(loss)))
        return float(loss[0]), g

    def set_data(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        self._check(self.lib.irs_mlp_set_data(self.h, _p(X), _p(T), len(X)))

    def train(self, steps: int, batch: int = 4096, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        losses = np.zeros(max(steps, 1), np.float32)
        self._check(self.lib.irs_mlp_train(self.h, int(steps), int(batch), float(lr), int(seed), _p(losses)))
        return losses[:steps]

    def input_jacobian(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        J = np.zeros((len(x), self.n_out, self.n_in), np.float32)
        self._check(self.lib.irs_mlp_input_jacobian(self.h, _p(x), len(x), _p(J)))
        return J


def layer_features(prof, mu: float = 1.0, T_method: str = "logp_at_pcg") -> np.ndarray:
    """Per-layer emulator inputs [L, 9]: ln p_mid, T_ave (TBAR), ln vmr_cg of H2O
    and O3 (1e-30 floor), CO2 vmr_cg in ppm, the path factor 1/mu, and the
    overlying column (the channel optical depth of a layer is defined through
    the transmittance above it): ln of the H2O and O3 amounts above the layer
    (sum vmr_cg dp over the layers above, + a small floor) and T_ave of the
    layer above (its own at the top)."""
    lp = layer_properties(prof, T_method)
    L = prof.n_levels - 1
    def v(s):
        return lp["vmr_cg"].get(GasSpecies(s), np.zeros(L))
    dp = lp["dp"]
    def above(q):
        a = np.cumsum((q * dp)[::-1])[::-1]                  # sum_{j>=k}
        return np.concatenate([a[1:], [0.0]])                # sum_{j>k}
    Ta = lp["T_ave"]
    T_up = np.concatenate([Ta[1:], Ta[-1:]])
    return np.stack([np.log(lp["p_mid"]), Ta, np.log(np.maximum(v(1), 1e-30)),
                     np.log(np.maximum(v(3), 1e-30)), v(2) * 1e6, np.full(L, 1.0 / mu),
                     np.log(above(v(1)) + 1e-9), np.log(above(v(3)) + 1e-12), T_up], axis=1)


def _scene(p):
    """(profile, T_sfc, mu) from a profile or a (profile, T_sfc, mu) tuple."""
    if isinstance(p, tuple):
        return p[0], p[1], p[2]
    return p, None, 1.0


def _randomized_pca(dev: np.ndarray, k: int, oversample: int = 32, power: int = 2, seed: int = 0):
    """Leading k right singular vectors of dev [n, m] (float32 data): range
    finder with power iterations (Halko, Martinsson & Tropp 2011), FP64 SVD of
    the small projected matrix."""
    rng = np.random.default_rng(seed)
    n, m = dev.shape
    kk = min(k + oversample, m, n)
    Om = rng.standard_normal((m, kk)).astype(np.float32)
    Y = dev @ Om
    for _ in range(power):
        Q, _ = np.linalg.qr(Y.astype(np.float64))
        Z = dev.T @ Q.astype(np.float32)
        Q2, _ = np.linalg.qr(Z.astype(np.float64))
        Y = dev @ Q2.astype(np.float32)
    Q, _ = np.linalg.qr(Y.astype(np.float64))
    Bm = (Q.T.astype(np.float32) @ dev).astype(np.float64)       # [kk, m]
    _, S, Vt = np.linalg.svd(Bm, full_matrices=False)
    return Vt[:k], S[:k]


@dataclass
class TrainReport:
    n_profiles: int
    n_samples: int
    seconds_lbl: float
    seconds_pca: float
    seconds_train: float
    pca_rms_t: float          # PCA truncation error of the relevance-weighted targets
    pca_max_t: float
    losses: np.ndarray = field(repr=False, default=None)


class ChannelODEmulator:
    """Trained emulator: MLP + PCA basis + normalisation, on the device."""

    def __init__(self, mlp: MLP, x_mean, x_std, coef_scale, basis, t_mean, ch_nu, T_method="logp_at_pcg",
                 n_tpc: int = 2, inv_w=None):
        self.mlp = mlp
        self.lib = mlp.lib
        self.x_mean = np.ascontiguousarray(x_mean, np.float32)
        self.x_std = np.ascontiguousarray(x_std, np.float32)
        self.coef_scale = np.ascontiguousarray(coef_scale, np.float32)
        self.basis = np.ascontiguousarray(basis, np.float32)
        self.t_mean = np.ascontiguousarray(t_mean, np.float32)
        self.ch_nu = np.ascontiguousarray(ch_nu, np.float64)
        self.inv_w = None if inv_w is None else np.ascontiguousarray(inv_w, np.float32)
        self.n_tpc = int(n_tpc)            # target kind: 2 (t_k, g_k) ratios, 3 (T_k, G_k) levels
        self.n_pc = self.basis.shape[0]
        self.n_ch = self.basis.shape[1] // 2
        if len(self.ch_nu) != self.n_ch:
            raise ValueError("basis width does not match the channels")
        self.L = self.t_mean.shape[0]
        self.T_method = T_method
        h = C.c_void_p()
        rc = self.lib.irs_emu_create(mlp.h, self.L, self.n_pc, self.n_ch, self.n_tpc, _p(self.x_mean), _p(self.x_std),
                                     _p(self.coef_scale), _p(self.basis), _p(self.t_mean), _p(self.inv_w),
                                     _p(self.ch_nu), C.byref(h))
        if rc != IRS_OK:
            raise IRSError(rc, f"irs_emu_create failed ({rc})")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.irs_emu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- training -----------------------------------------------------------
    @staticmethod
    def train(profiles: Sequence, linelists: Dict[GasSpecies, HITRANLinelist], sounder: Sounder, *,
              n_pc: int = 256, steps: int = 3000, batch: int = 4096, lr: float = 1e-3, seed: int = 0,
              floor: float = 1e-4, line_mixing="vpy", ctx: Optional[Context] = None, targets=None,
              relevance_floor: Optional[float] = None, apply_continuum: bool = False, **kw):
        """Train on scenes `profiles` (AtmosphericProfile or (profile, T_sfc, mu)
        tuples); the targets along each scene's path come from the GPU LBL
        engine unless `targets` = (tau_ch [N, L, n_ch], g [N, L, n_ch]) is given.
        The LBL runs without continuum unless apply_continuum is True; the spot
        check and fallback of forward_emulated must use the same setting."""
        scenes = [_scene(p) for p in profiles]
        profs = [s[0] for s in scenes]
        mus = [s[2] for s in scenes]
        geoms = [ViewingGeometry(float(np.degrees(np.arccos(min(m, 1.0)))), 0.0, m) for m in mus]
        t0 = time.perf_counter()
        if targets is None:
            targets = channel_optical_depths(profs, linelists, sounder=sounder, geom=geoms, floor=floor,
                                             line_mixing=line_mixing, cim=True, ctx=ctx,
                                             apply_continuum=apply_continuum, **kw)
        tau_ch, g = targets
        t1 = time.perf_counter()
        N, L, nch = g.shape
        # target weights W (PCA and loss run on W x deviation, the decode divides by
        # W): the CIM weights enter the radiance through a Planck difference across
        # half a layer, ~10x less than the transmittances, so they carry 0.1 W.
        # Optional relevance weights (relevance_floor not None): a layer's channel
        # values reach the top of the atmosphere through the transmittance above it,
        # W[k, c] = max over scenes of T_{k+1}(c) (floor relevance_floor); scenes more
        # transparent than every training scene then see errors amplified by up to
        # 1/W (measured: without them the maximum held-out error is smaller).
        if relevance_floor is None:
            W = np.ones((L, nch), np.float32)
        else:
            above = np.zeros((L, nch))
            for q in range(N):
                cum = np.cumsum(tau_ch[q, ::-1].astype(np.float64), axis=0)[::-1]     # sum_{j>=k}
                Ta = np.exp(-np.vstack([cum[1:], np.zeros((1, nch))]))               # T_{k+1}
                np.maximum(above, Ta, out=above)
            W = np.maximum(above, relevance_floor).astype(np.float32)
        Wv = np.concatenate([W, 0.1 * W], axis=1)                                  # [L, 2 nch]
        t = np.concatenate([np.exp(-tau_ch.astype(np.float32)), g.astype(np.float32)], axis=2)
        del targets, tau_ch, g
        t_mean = t.mean(axis=0)
        t -= t_mean[None]
        t *= Wv[None]
        dev = t.reshape(N * L, 2 * nch)
        V, _ = _randomized_pca(dev, n_pc, seed=seed)
        Vf = np.zeros((n_pc, 2 * nch), np.float32)           # fewer columns than n_pc: zero rows
        Vf[:len(V)] = V
        coef = dev @ Vf.T                                           # [N L, n_pc]
        rec = coef @ Vf
        err = dev - rec                                             # relevance-weighted units
        pca_rms, pca_max = float(np.sqrt(np.mean(err * err))), float(np.abs(err).max())
        del err, rec
        t2 = time.perf_counter()
        feats = np.stack([layer_features(p, m) for p, m in zip(profs, mus)]).astype(np.float64)   # [N, L, F]
        # per-layer standardisation keeps the bf16 inputs fine-grained (feature 0,
        # ln p, is the same grid for every scene: global statistics keep the layer identity)
        x_mean = feats.mean(axis=0)
        x_std = feats.std(axis=0)
        g_mean, g_std = feats.reshape(-1, feats.shape[2]).mean(axis=0), feats.reshape(-1, feats.shape[2]).std(axis=0)
        x_mean[:, 0], x_std[:, 0] = g_mean[0], max(g_std[0], 1e-6)
        x_std = np.where(x_std > 1e-3 * np.maximum(g_std, 1e-12)[None], x_std, np.maximum(g_std, 1e-6)[None])
        feats_n = ((feats - x_mean[None]) / x_std[None]).reshape(N * L, -1)
        scale = float(max(coef[:, 0].std(), 1e-12))
        coef_scale = np.full(n_pc, scale, np.float32)
        mlp = MLP(N_FEATURES, 512, n_pc, max(batch, 64), seed=seed)
        mlp.set_data(feats_n.astype(np.float32), (coef / scale).astype(np.float32))
        # Adam at lr for half the steps, then lr/3 and lr/10 for a quarter each
        n1, n2 = steps // 2, steps // 4
        losses = np.concatenate([mlp.train(n1, batch, lr, seed), mlp.train(n2, batch, lr / 3.0, seed + 1),
                                 mlp.train(steps - n1 - n2, batch, lr / 10.0, seed + 2)])
        t3 = time.perf_counter()
        sg = sounder_grid(sounder)
        # the decode undoes the weights: target_k = mean_k + (sum_m a_{k,m} V_m) / W_k
        emu = ChannelODEmulator(mlp, x_mean, x_std, coef_scale, Vf, t_mean, sg.nu, n_tpc=2, inv_w=1.0 / Wv)
        rep = TrainReport(N, N * L, t1 - t0, t2 - t1, t3 - t2, pca_rms, pca_max, losses)
        return emu, rep

    # ---- persistence --------------------------------------------------------
    def save(self, path: str):
        np.savez_compressed(path, params=self.mlp.get_params(), x_mean=self.x_mean, x_std=self.x_std,
                            coef_scale=self.coef_scale, basis=self.basis.astype(np.float32),
                            t_mean=self.t_mean, ch_nu=self.ch_nu,
                            inv_w=self.inv_w if self.inv_w is not None else np.ones_like(self.t_mean),
                            dims=np.array([self.mlp.n_in, self.mlp.n_hidden, self.mlp.n_out, self.n_tpc]))

    @staticmethod
    def load(path: str, max_batch: int = 4096):
        z = np.load(path, allow_pickle=False)
        n_in, H, n_out, n_tpc = (int(v) for v in z["dims"])
        mlp = MLP(n_in, H, n_out, max_batch)
        mlp.set_params(z["params"])
        return ChannelODEmulator(mlp, z["x_mean"], z["x_std"], z["coef_scale"], z["basis"], z["t_mean"],
                                 z["ch_nu"], n_tpc=n_tpc, inv_w=z["inv_w"])

    # ---- Jacobians ----------------------------------------------------------
    def jacobian(self, profs, *, geom=None, T_sfc=None, eps_sfc=1.0) -> List["EmulatorJacobian"]:
        """Emulator Jacobians per scene (levels: T and ln q_H2O; T_sfc); the
        reflected term (eps < 1) is held fixed."""
        nb = len(profs)
        geoms = geom if isinstance(geom, (list, tuple)) else [geom] * nb
        geoms = [g if g is not None else ViewingGeometry(0.0, 0.0, 1.0) for g in geoms]
        tsf = T_sfc if isinstance(T_sfc, (list, tuple, np.ndarray)) else [T_sfc] * nb
        eps = eps_sfc if isinstance(eps_sfc, (list, tuple, np.ndarray)) else [eps_sfc] * nb
        return _emu_jacobian(self, list(profs), geoms, list(tsf), list(eps))

    # ---- inference ----------------------------------------------------------
    def _inputs(self, profs, T_sfcs, eps, mus):
        feats, tlev, tave, scal = [], [], [], []
        for q, p in enumerate(profs):
            lp = layer_properties(p, self.T_method)
            feats.append(layer_features(p, mus[q], self.T_method))
            tlev.append(p.temperature)
            tave.append(lp["T_ave"])
            scal.append([p.temperature[0] if T_sfcs[q] is None else T_sfcs[q], mus[q], eps[q]])
        return (np.ascontiguousarray(np.stack(feats), np.float32), np.ascontiguousarray(np.stack(tlev)),
                np.ascontiguousarray(np.stack(tave)), np.ascontiguousarray(np.array(scal, np.float64)))

    def forward(self, profs, *, geom=None, T_sfc=None, eps_sfc=1.0, source_function="cim",
                want_t: bool = False):
        """Emulated (R, BT) [n_prof, n_ch] for a list of profiles."""
        nb = len(profs)
        geoms = geom if isinstance(geom, (list, tuple)) else [geom] * nb
        mus = [1.0 if g is None else g.mu for g in geoms]
        tsf = T_sfc if isinstance(T_sfc, (list, tuple, np.ndarray)) else [T_sfc] * nb
        eps = eps_sfc if isinstance(eps_sfc, (list, tuple, np.ndarray)) else [eps_sfc] * nb
        X, tlev, tave, scal = self._inputs(profs, tsf, eps, mus)
        if X.shape[1] != self.L:
            raise ValueError(f"profiles have {X.shape[1]} layers, the emulator {self.L}")
        R = np.zeros((nb, self.n_ch))
        BT = np.zeros((nb, self.n_ch))
        tout = np.zeros((nb, self.L, self.n_ch), np.float32) if want_t else None
        if source_function.lstrip(":") != "cim" and self.n_tpc == 2:
            raise ValueError("the emulator is trained for source_function=:cim")
        sf = api.SOURCE_FUNCTION[source_function.lstrip(":")]
        rc = self.lib.irs_emu_forward(self.h, _p(X), _p(tlev), _p(tave), _p(scal), nb, sf, _p(R), _p(BT),
                                      _p(tout))
        if rc != IRS_OK:
            raise IRSError(rc, self.lib.irs_emu_last_error(self.h).decode(errors="replace"))
        return (R, BT, tout) if want_t else (R, BT)


def _feature_level_jacobian(prof, mu, T_method, hT=1e-2, hq=1e-3):
    """d features / d (level T) and d features / d (level ln q_H2O), and
    d T_ave / d (level T), by central differences of the host layer code
    (smooth, O(levels x layers) work)."""
    from .types import AtmosphericProfile
    nl = prof.n_levels
    def feats(T=None, vmr=None):
        p = AtmosphericProfile(prof.pressure, prof.temperature if T is None else T, prof.altitude,
                               prof.vmr if vmr is None else vmr)
        return layer_features(p, mu, T_method), layer_properties(p, T_method)["T_ave"]
    X0, _ = feats()
    L, nf = X0.shape
    dXdT = np.zeros((nl, L, nf))
    dAdT = np.zeros((nl, L))
    dXdq = np.zeros((nl, L, nf))
    for j in range(nl):
        Tp, Tm = prof.temperature.copy(), prof.temperature.copy()
        Tp[j] += hT
        Tm[j] -= hT
        (xp, ap), (xm, am) = feats(T=Tp), feats(T=Tm)
        dXdT[j] = (xp - xm) / (2 * hT)
        dAdT[j] = (ap - am) / (2 * hT)
        if GasSpecies.H2O in prof.vmr:
            vp = dict(prof.vmr)
            vm = dict(prof.vmr)
            q = prof.vmr[GasSpecies.H2O]
            qp, qm = q.copy(), q.copy()
            qp[j] *= np.exp(hq)
            qm[j] *= np.exp(-hq)
            vp[GasSpecies.H2O], vm[GasSpecies.H2O] = qp, qm
            dXdq[j] = (feats(vmr=vp)[0] - feats(vmr=vm)[0]) / (2 * hq)
    return dXdT, dAdT, dXdq


@dataclass
class EmulatorJacobian:
    """Per scene: BT [n_ch], K_T [n_ch, n_levels] (dBT/dT_level), K_lnq_H2O
    [n_ch, n_levels] (dBT/d ln q_H2O level), K_Tsfc [n_ch], and the layer
    form dBT/dfeature [L, n_features, n_ch]."""
    BT: np.ndarray
    K_T: np.ndarray
    K_lnq_H2O: np.ndarray
    K_Tsfc: np.ndarray
    dfeat: np.ndarray = field(repr=False, default=None)


def _emu_jacobian(emu, profs, geoms, tsf, eps):
    nb = len(profs)
    mus = [g.mu for g in geoms]
    X, tlev, tave, scal = emu._inputs(profs, tsf, eps, mus)
    L, nin, nch = emu.L, emu.mlp.n_in, emu.n_ch
    BT = np.zeros((nb, nch))
    dfeat = np.zeros((nb, L, nin, nch), np.float32)
    dTave = np.zeros((nb, L, nch), np.float32)
    dTlev = np.zeros((nb, L + 1, nch), np.float32)
    dTsfc = np.zeros((nb, nch), np.float32)
    rc = emu.lib.irs_emu_jacobian(emu.h, _p(X), _p(tlev), _p(tave), _p(scal), nb, _p(BT), _p(dfeat), _p(dTave),
                                  _p(dTlev), _p(dTsfc))
    if rc != IRS_OK:
        raise IRSError(rc, emu.lib.irs_emu_last_error(emu.h).decode(errors="replace"))
    out = []
    for q in range(nb):
        dXdT, dAdT, dXdq = _feature_level_jacobian(profs[q], mus[q], emu.T_method)
        F = dfeat[q].astype(np.float64)
        KT = np.einsum("kic,jki->cj", F, dXdT) + np.einsum("kc,jk->cj", dTave[q].astype(np.float64), dAdT)
        KT += dTlev[q].T.astype(np.float64)
        Kq = np.einsum("kic,jki->cj", F, dXdq)
        out.append(EmulatorJacobian(BT[q], KT, Kq, dTsfc[q].astype(np.float64), dfeat[q]))
    return out


# ---- spot check + fallback -----------------------------------------------------
DEFAULT_SPOT_BLOCKS = ((700.0, 16), (960.0, 16), (1050.0, 16), (2300.0, 16))


class SpotChecker:
    """LBL of a few blocks of consecutive channels, each on its own internal
    sub-grid (channel block +- the 16 cm^-1 ILS pad) with only the lines that
    reach it (+- 25 cm^-1 + 1): the same channel radiances as the full-range
    LBL at a fraction of its work."""

    def __init__(self, linelists, sounder: Sounder, blocks=DEFAULT_SPOT_BLOCKS):
        sg = sounder_grid(sounder)
        self.sounder = sounder
        self.blocks = []
        for lo, n in blocks:
            c0 = int(round((lo - sounder.nu_min) / sounder.dnu))
            c0 = max(0, min(c0, sg.n - n))
            b_lo = sounder.nu_min + c0 * sounder.dnu
            b_hi = sounder.nu_min + (c0 + n - 1) * sounder.dnu
            snd = Sounder.make(b_lo, b_hi, sounder.dnu, opd_max=sounder.opd_max, fwhm_gauss=sounder.fwhm_gauss)
            sub = {k: filter_linelist(v, b_lo - 16.0 - 26.0, b_hi + 16.0 + 26.0) for k, v in linelists.items()}
            self.blocks.append((np.arange(c0, c0 + n), snd, sub))
        self.channels = np.concatenate([b[0] for b in self.blocks])

    def __call__(self, profs, geoms, T_sfcs, eps, ctx, **kw):
        out = np.zeros((len(profs), len(self.channels)))
        o = 0
        for ch, snd, sub in self.blocks:
            _, _, bt = forward_model(list(profs), sub, sounder=snd, geom=list(geoms), T_sfc=list(T_sfcs),
                                     eps_sfc=list(eps), ctx=ctx, **kw)
            out[:, o:o + len(ch)] = bt
            o += len(ch)
        return out


@dataclass
class EmulatedResult:
    R: np.ndarray
    BT: np.ndarray
    fell_back: np.ndarray
    spot_err: np.ndarray
    seconds: Dict[str, float]


def forward_emulated(profs, linelists, emulator: ChannelODEmulator, *, sounder: Sounder,
                     geom=None, T_sfc=None, eps_sfc=1.0, threshold: float = 0.05,
                     spot: Optional[SpotChecker] = None, ctx: Optional[Context] = None,
                     line_mixing="vpy", apply_continuum: bool = False, **kw) -> EmulatedResult:
    """Emulator + per-scene LBL spot check + full-LBL fallback (BT budget
    `threshold` K on the spot channels).  apply_continuum as in training."""
    kw["apply_continuum"] = apply_continuum
    nb = len(profs)
    geoms = geom if isinstance(geom, (list, tuple)) else [geom] * nb
    geoms = [g if g is not None else ViewingGeometry(0.0, 0.0, 1.0) for g in geoms]
    tsf = T_sfc if isinstance(T_sfc, (list, tuple, np.ndarray)) else [T_sfc] * nb
    eps = eps_sfc if isinstance(eps_sfc, (list, tuple, np.ndarray)) else [eps_sfc] * nb
    ctx = ctx or api.default_context()
    spot = spot or SpotChecker(linelists, sounder)
    t0 = time.perf_counter()
    R, BT = emulator.forward(list(profs), geom=geoms, T_sfc=list(tsf), eps_sfc=list(eps))
    t1 = time.perf_counter()
    bt_spot = spot(profs, geoms, tsf, eps, ctx, line_mixing=line_mixing, **kw)
    err = np.max(np.abs(BT[:, spot.channels] - bt_spot), axis=1)
    t2 = time.perf_counter()
    fb = err > threshold
    if np.any(fb):
        idx = np.flatnonzero(fb)
        _, Rf, BTf = forward_model([profs[i] for i in idx], linelists, sounder=sounder,
                                   geom=[geoms[i] for i in idx], T_sfc=[tsf[i] for i in idx],
                                   eps_sfc=[eps[i] for i in idx], ctx=ctx, line_mixing=line_mixing, **kw)
        R[idx] = Rf
        BT[idx] = BTf
    t3 = time.perf_counter()
    return EmulatedResult(R, BT, fb, err, {"emulator": t1 - t0, "spot_check": t2 - t1, "fallback": t3 - t2})
