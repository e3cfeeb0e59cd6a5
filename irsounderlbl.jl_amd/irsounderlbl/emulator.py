"""Learned emulator of per-layer channel optical depths, with LBL spot check
and fallback.  No counterpart in the reference.

Pipeline (numeric work on the GPU through libirs, except the PCA and the
per-layer least-squares fits, which are FP64/FP32 host linear algebra):
  1. training targets: for every layer k of GPU line-by-line spectra along each
     scene's viewing path (irs_channel_od_moments) the layer channel transmittance
     t_k = T_k / T_{k+1} of the ILS-convolved level-to-space transmittances, the
     CIM weight g_k and the normalised ILS first moment m_k of the layer's
     transmittance step; the channel RTE with Planck linear in nu across the ILS
     reproduces the ILS-convolved LBL radiance from them (0.02 K at worst on the
     IASI grid, against 0.14 K with Planck at the channel centre alone);
  2. compression: per layer, the relevance-weighted deviations from the layer's
     training mean (weight: the largest training transmittance above the layer
     times the BT sensitivity to radiance, so the basis spends itself where the
     top of the atmosphere sees the layer) projected on n_pc leading right
     singular vectors of that layer (randomized range finder + FP64 SVD); the
     bases come from the first scenes, later scenes are projected as their
     targets are produced, so the training set is never held in memory whole;
  3. model of the coefficients: a per-layer quadratic regression in the layer's
     own 13 features plus a linear term in 8 profile-context features (leading
     principal components of the standardised T, ln q_H2O, ln q_O3 profiles,
     T_sfc and 1/mu), and an MLP 13 -> 3 x n_hidden GELU -> n_pc trained with the BF16
     MFMA kernels (FP32 accumulation, Adam) on the regression's residual;
  4. inference: features -> coefficients (regression + MLP) -> per-layer decode ->
     channel RTE (irs_emu_forward), batched on the device;
  5. calibration on held-out scenes against GPU LBL: the per-channel hold-out BT
     error; spot blocks whose measured errors predict every channel's error (the
     emulator's errors are low-rank across channels); a ridge map from the spot
     errors to every channel's error; and the fallback rule (thresholds on the
     largest spot error and on the map's leave-one-block-out residual) under which
     every calibration scene is within (1 - margin) x budget after correction,
     judged out of fold;
  6. the per-scene LBL spot check (irs_forward_emulated): every channel corrected
     from the scene's spot errors, the spot channels set to their LBL values, and
     the scenes the rule flags recomputed with full LBL.
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
from .types import (AtmosphericProfile, GasSpecies, HITRANLinelist, Sounder, ViewingGeometry, filter_linelist,
                    sounder_grid)

N_LOCAL = 13
N_CTX = 8
N_FEATURES = N_LOCAL + N_CTX
FEATURE_NAMES = ("ln p_mid [hPa]", "T_ave [K]", "ln vmr H2O", "ln vmr O3", "vmr CO2 [ppm]", "1/mu",
                 "ln H2O column above [hPa]", "ln O3 column above [hPa]", "T_ave of the layer above [K]",
                 "H2O-weighted ln p above", "H2O-weighted T above [K]", "O3-weighted T above [K]",
                 "mass-weighted T above [K]") + \
                tuple(f"profile context {i}" for i in range(N_CTX))
W_G, W_M = 0.1, 0.01          # decode weights of the CIM weight and the ILS moment relative to t_k


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
    """Per-layer emulator inputs [L, 13]: ln p_mid, T_ave (TBAR), ln vmr_cg of H2O
    and O3 (1e-30 floor), CO2 vmr_cg in ppm, the path factor 1/mu, and the
    overlying path (the channel transmittance of a layer is the ratio of the
    ILS-convolved transmittances below and above it, so it depends on the spectral
    structure of the path above): ln of the H2O and O3 amounts above the layer
    (sum vmr_cg dp over the layers above, + a small floor), T_ave of the layer
    above (its own at the top), and the Curtis-Godson scaling of the path above --
    the H2O-weighted mean ln p and T, the O3-weighted mean T and the mass-weighted
    mean T of the layers above (u-weighted means sum u_j x_j / sum u_j, which fall
    back to the layer's own x where nothing is above)."""
    return _features_from_layers(*_layer_quantities(layer_properties(prof, T_method), prof.n_levels - 1), mu)


def _layer_quantities(lp, L):
    """(p_mid, dp, T_ave, vmr_cg H2O, O3, CO2) of layer_properties, zero for an absent species."""
    def v(sp):
        return lp["vmr_cg"].get(GasSpecies(sp), np.zeros(L))
    return lp["p_mid"], lp["dp"], lp["T_ave"], v(1), v(3), v(2)


def _features_from_layers(p_mid, dp, Ta, vh, vo, vc, mu):
    """layer_features from the layer quantities, vectorised over any leading axes
    (the last axis is the layer)."""
    p_mid, dp, Ta, vh, vo, vc = np.broadcast_arrays(*(np.asarray(x, np.float64) for x in (p_mid, dp, Ta, vh, vo, vc)))
    def above(q):
        a = np.cumsum(q[..., ::-1], axis=-1)[..., ::-1]            # sum_{j>=k}
        return np.concatenate([a[..., 1:], np.zeros(a.shape[:-1] + (1,))], axis=-1)   # sum_{j>k}
    def cg_mean(u, x, eps):
        return (above(u * x) + eps * x) / (above(u) + eps)
    lnp = np.log(p_mid)
    T_up = np.concatenate([Ta[..., 1:], Ta[..., -1:]], axis=-1)
    uh, uo = vh * dp, vo * dp
    return np.stack([lnp, Ta, np.log(np.maximum(vh, 1e-30)), np.log(np.maximum(vo, 1e-30)), vc * 1e6,
                     np.full(Ta.shape, 1.0 / mu), np.log(above(uh) + 1e-9), np.log(above(uo) + 1e-12), T_up,
                     cg_mean(uh, lnp, 1e-9), cg_mean(uh, Ta, 1e-9), cg_mean(uo, Ta, 1e-12),
                     cg_mean(dp, Ta, 1e-6)], axis=-1)


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


def _scene(p):
    """(profile, T_sfc, mu) from a profile or a (profile, T_sfc, mu) tuple."""
    if isinstance(p, tuple):
        return p[0], p[1], p[2]
    return p, None, 1.0


def _geom(mu):
    return ViewingGeometry(float(np.degrees(np.arccos(min(mu, 1.0)))), 0.0, mu)


class ProfileContext:
    """Profile-context features: the leading principal components of the
    standardised profile vector [T levels, ln q_H2O levels, ln q_O3 levels, T_sfc,
    1/mu] of the training scenes, each scaled to unit variance."""

    def __init__(self, p_mean, p_std, proj):
        self.p_mean = np.asarray(p_mean, np.float64)
        self.p_std = np.asarray(p_std, np.float64)
        self.proj = np.asarray(proj, np.float64)            # [n_ctx, n_vec], includes the 1/std of each PC

    @staticmethod
    def vector(prof, T_sfc, mu):
        v = prof.vmr
        nl = prof.n_levels
        def lnq(sp):
            return np.log(np.maximum(v[sp], 1e-30)) if sp in v else np.full(nl, np.log(1e-30))
        ts = prof.temperature[0] if T_sfc is None else T_sfc
        return np.concatenate([prof.temperature, lnq(GasSpecies.H2O), lnq(GasSpecies.O3), [ts, 1.0 / mu]])

    @staticmethod
    def fit(scenes, n_ctx: int = N_CTX):
        P = np.stack([ProfileContext.vector(*s) for s in scenes])
        pm, ps = P.mean(0), P.std(0)
        ps = np.where(ps > 1e-12, ps, 1.0)
        Z = (P - pm) / ps
        _, S, Vt = np.linalg.svd(Z, full_matrices=False)
        V = Vt[:n_ctx]
        sd = (Z @ V.T).std(0)
        sd = np.where(sd > 1e-12, sd, 1.0)
        if len(V) < n_ctx:                                     # fewer scenes than context features
            V = np.vstack([V, np.zeros((n_ctx - len(V), V.shape[1]))])
            sd = np.concatenate([sd, np.ones(n_ctx - len(sd))])
        return ProfileContext(pm, ps, V / sd[:, None] / ps[None, :])

    def __call__(self, prof, T_sfc, mu):
        return self.proj @ (self.vector(prof, T_sfc, mu) - self.p_mean)


def emulator_features(scenes, context: ProfileContext, T_method: str = "logp_at_pcg") -> np.ndarray:
    """Raw emulator inputs [N, L, N_FEATURES]: layer_features and the profile
    context (the same for every layer of a scene)."""
    out = []
    for prof, T_sfc, mu in scenes:
        lf = layer_features(prof, mu, T_method)
        cx = context(prof, T_sfc, mu)
        out.append(np.concatenate([lf, np.repeat(cx[None], lf.shape[0], 0)], axis=1))
    return np.stack(out)


def _phi(z: np.ndarray, n_local: int = N_LOCAL) -> np.ndarray:
    """Regression terms [..., n_phi]: 1, z_i (i < n_local), z_a z_b (a <= b < n_local), the rest of z."""
    zl = z[..., :n_local]
    iu = np.triu_indices(n_local)
    q = (zl[..., :, None] * zl[..., None, :])[..., iu[0], iu[1]]
    return np.concatenate([np.ones(z.shape[:-1] + (1,)), zl, q, z[..., n_local:]], axis=-1)


def _bt_sensitivity(nu, T=250.0):
    """B / (dB/dT) at T (K per unit relative radiance), normalised to mean 1."""
    x = 1.4387769 * nu / T
    s = (T / x) * (1.0 - np.exp(-x))
    return s / s.mean()


@dataclass
class TrainReport:
    n_profiles: int
    n_samples: int
    n_basis_profiles: int
    seconds_lbl: float
    seconds_pca: float
    seconds_fit: float
    seconds_train: float
    coef_rms_err_regression: float     # training-set coefficient error of the regression alone
    coef_rms: float
    losses: np.ndarray = field(repr=False, default=None)


@dataclass
class Calibration:
    """Hold-out statistics of an emulator against GPU LBL on calibration scenes."""
    spot_channels: np.ndarray          # channel indices of the spot check
    spot_blocks: list                  # [(first channel, n)] of consecutive spot channels
    threshold: float                   # spot-error threshold (K) of the fallback decision
    budget: float                      # BT budget (K) the threshold guarantees on the calibration scenes
    ch_rms: np.ndarray                 # per-channel hold-out BT error, RMS over the scenes
    ch_max: np.ndarray                 # per-channel largest hold-out error
    scene_max: np.ndarray              # per-scene largest error over all channels
    scene_spot: np.ndarray             # per-scene fallback statistic (largest spot error, or the correction's)
    n_scenes: int
    # spot-check correction (None: the emulator's spectra are used as they come)
    W: Optional[np.ndarray] = field(repr=False, default=None)          # [n_spot, n_ch]
    lbo_W: Optional[np.ndarray] = field(repr=False, default=None)      # leave-one-block-out maps, concatenated
    w_raw: float = 1.0                 # fallback statistic max(w_raw max|e_S|, w_lbo leave-one-block-out residual)
    w_lbo: float = 0.0
    scene_max_corrected: Optional[np.ndarray] = field(repr=False, default=None)   # out-of-fold, per scene
    fallback_rate: float = float("nan")                                # calibration scenes over the threshold


class ChannelODEmulator:
    """Trained emulator: regression + MLP + per-layer PCA bases, on the device."""

    def __init__(self, mlp: MLP, x_mean, x_std, A, coef_scale, basis, mean, inv_w, ch_nu, context: ProfileContext,
                 T_method="logp_at_pcg"):
        self.mlp = mlp
        self.lib = mlp.lib
        self.x_mean = np.ascontiguousarray(x_mean, np.float32)
        self.x_std = np.ascontiguousarray(x_std, np.float32)
        self.A = np.ascontiguousarray(A, np.float32)
        self.coef_scale = float(coef_scale)
        self.basis = np.ascontiguousarray(basis, np.float32)
        self.mean = np.ascontiguousarray(mean, np.float32)
        self.inv_w = np.ascontiguousarray(inv_w, np.float32)
        self.ch_nu = np.ascontiguousarray(ch_nu, np.float64)
        self.context = context
        self.L, self.n_pc, w3 = self.basis.shape
        self.n_ch = w3 // 3
        if len(self.ch_nu) != self.n_ch:
            raise ValueError("basis width does not match the channels")
        self.T_method = T_method
        self.calibration: Optional[Calibration] = None
        self._spot = None
        h = C.c_void_p()
        rc = self.lib.irs_emu_create(mlp.h, self.L, self.n_pc, self.n_ch, N_FEATURES, N_LOCAL, _p(self.x_mean),
                                     _p(self.x_std), _p(self.A), self.coef_scale, _p(self.basis), _p(self.mean),
                                     _p(self.inv_w), _p(self.ch_nu), C.byref(h))
        if rc != IRS_OK:
            raise IRSError(rc, f"irs_emu_create failed ({rc})")
        self.h = h
        cm = np.ascontiguousarray(context.p_mean, np.float64)
        cp = np.ascontiguousarray(context.proj, np.float64)
        rc = self.lib.irs_emu_set_context(h, cp.shape[0], cp.shape[1], _p(cm), _p(cp))
        if rc != IRS_OK:
            raise IRSError(rc, self.lib.irs_emu_last_error(h).decode(errors="replace"))

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
              n_pc: int = 64, n_basis: int = 2048, steps: int = 24000, batch: int = 4096, lr: float = 1e-3,
              n_hidden: int = 1024,
              seed: int = 0, floor: float = 1e-8, line_mixing="vpy", ctx: Optional[Context] = None,
              relevance_floor: float = 1e-2, chunk: int = 256, apply_continuum: bool = False, progress=None,
              T_method: str = "logp_at_pcg", **kw):
        """Train on scenes `profiles` (AtmosphericProfile or (profile, T_sfc, mu)
        tuples) from GPU LBL targets along each scene's path.  The LBL runs
        without continuum unless apply_continuum is True; the spot check and
        fallback of forward_emulated must use the same setting."""
        scenes = [_scene(p) for p in profiles]
        N = len(scenes)
        n_basis = min(n_basis, N)
        sg = sounder_grid(sounder)
        nch = sg.n
        t0 = time.perf_counter()
        t_lbl = 0.0

        def targets(sc):
            nonlocal t_lbl
            ta = time.perf_counter()
            od, g, m = channel_optical_depths([s[0] for s in sc], linelists, sounder=sounder,
                                              geom=[_geom(s[2]) for s in sc], floor=floor, moments=True,
                                              line_mixing=line_mixing, ctx=ctx, apply_continuum=apply_continuum,
                                              T_method=T_method, **kw)
            t_lbl += time.perf_counter() - ta
            return np.concatenate([np.exp(-od), g, m], axis=2)       # [n, L, 3 nch] float32

        Yb = []
        for c0 in range(0, n_basis, chunk):
            Yb.append(targets(scenes[c0:min(c0 + chunk, n_basis)]))
            if progress:
                progress(f"emulator targets {min(c0 + chunk, n_basis)}/{N}")
        Yb = np.concatenate(Yb)
        L = Yb.shape[1]
        tp = time.perf_counter()
        mean = Yb.mean(axis=0)                                        # [L, 3 nch]
        top = np.ones((n_basis, nch), np.float32)
        rel = np.zeros((L, nch), np.float32)
        for k in range(L - 1, -1, -1):
            rel[k] = top.max(axis=0)
            top *= np.clip(Yb[:, k, :nch], 0.0, 1.0)
        del top
        wr = np.maximum(rel, relevance_floor) * _bt_sensitivity(sg.nu).astype(np.float32)[None]
        WR = np.concatenate([wr, W_G * wr, W_M * wr], axis=1).astype(np.float32)     # [L, 3 nch]
        basis = np.zeros((L, n_pc, 3 * nch), np.float32)
        coef = np.zeros((N, L, n_pc), np.float32)
        for l in range(L):
            d = ((Yb[:, l] - mean[l]) * WR[l]).astype(np.float32)
            V, _ = _randomized_pca(d, n_pc, seed=seed + l)
            basis[l, :len(V)] = V
            coef[:n_basis, l] = d @ basis[l].T
        del Yb, d
        t_pca = time.perf_counter() - tp
        for c0 in range(n_basis, N, chunk):
            sc = scenes[c0:c0 + chunk]
            Y = targets(sc)
            for l in range(L):
                coef[c0:c0 + len(sc), l] = ((Y[:, l] - mean[l]) * WR[l]).astype(np.float32) @ basis[l].T
            del Y
            if progress:
                progress(f"emulator targets {c0 + len(sc)}/{N}")
        tf = time.perf_counter()
        context = ProfileContext.fit(scenes)
        X = emulator_features(scenes, context, T_method)                      # [N, L, F]
        x_mean, x_std = X.mean(axis=0), X.std(axis=0)
        g_mean, g_std = X.reshape(-1, N_FEATURES).mean(axis=0), X.reshape(-1, N_FEATURES).std(axis=0)
        # feature 0 (ln p) is the same grid for every scene: global statistics keep the layer identity
        x_mean[:, 0], x_std[:, 0] = g_mean[0], max(g_std[0], 1e-6)
        x_std = np.where(x_std > 1e-3 * np.maximum(g_std, 1e-12)[None], x_std, np.maximum(g_std, 1e-6)[None])
        Z = (X - x_mean[None]) / x_std[None]
        Phi = _phi(Z)
        n_phi = Phi.shape[-1]
        A = np.zeros((L, n_phi, n_pc))
        res = np.empty_like(coef)
        for l in range(L):
            P = Phi[:, l]
            G = P.T @ P
            G += 1e-8 * np.trace(G) / n_phi * np.eye(n_phi)
            A[l] = np.linalg.solve(G, P.T @ coef[:, l].astype(np.float64))
            res[:, l] = coef[:, l] - (P @ A[l]).astype(np.float32)
        reg_err = float(np.sqrt(np.mean(res.astype(np.float64) ** 2)))
        coef_rms = float(np.sqrt(np.mean(coef.astype(np.float64) ** 2)))
        del Phi
        t_fit = time.perf_counter() - tf
        tt = time.perf_counter()
        scale = float(max(res.std(), 1e-12))
        mlp = MLP(N_LOCAL, n_hidden, n_pc, max(batch, 64), seed=seed)
        mlp.set_data(Z[..., :N_LOCAL].reshape(-1, N_LOCAL).astype(np.float32),
                     (res / scale).reshape(-1, n_pc).astype(np.float32))
        # Adam at lr for half the steps, then lr/3 and lr/10 for a quarter each (in
        # segments of at most 3000 steps, each reshuffled with its own seed)
        n1, n2 = steps // 2, steps // 4
        parts, done = [], 0
        for n_ph, lr_ph in ((n1, lr), (n2, lr / 3.0), (steps - n1 - n2, lr / 10.0)):
            k0 = 0
            while k0 < n_ph:
                k = min(3000, n_ph - k0)
                parts.append(mlp.train(k, batch, lr_ph, seed + 7919 * (done + k0)))
                k0 += k
                if progress:
                    progress(f"emulator MLP {done + k0}/{steps} Adam steps, loss {parts[-1][-20:].mean():.3e}")
            done += n_ph
        losses = np.concatenate(parts) if parts else np.zeros(0, np.float32)
        t_train = time.perf_counter() - tt
        emu = ChannelODEmulator(mlp, x_mean, x_std, A.astype(np.float32), scale, basis, mean, 1.0 / WR, sg.nu,
                                context, T_method)
        rep = TrainReport(N, N * L, n_basis, t_lbl, t_pca, t_fit, t_train, reg_err, coef_rms, losses)
        return emu, rep

    # ---- persistence --------------------------------------------------------
    def save(self, path: str):
        cal = self.calibration
        extra = {}
        if cal is not None:
            extra = dict(cal_spot=cal.spot_channels, cal_blocks=np.array(cal.spot_blocks, np.int64),
                         cal_thr=np.array([cal.threshold, cal.budget, cal.n_scenes, cal.w_raw, cal.w_lbo,
                                           cal.fallback_rate]),
                         cal_ch_rms=cal.ch_rms, cal_ch_max=cal.ch_max, cal_scene_max=cal.scene_max,
                         cal_scene_spot=cal.scene_spot)
            if cal.W is not None:
                extra.update(cal_W=cal.W, cal_lbo=cal.lbo_W, cal_scene_max_corr=cal.scene_max_corrected)
        np.savez_compressed(path, params=self.mlp.get_params(), x_mean=self.x_mean, x_std=self.x_std, A=self.A,
                            coef_scale=np.array([self.coef_scale]), basis=self.basis, mean=self.mean,
                            inv_w=self.inv_w, ch_nu=self.ch_nu, ctx_mean=self.context.p_mean,
                            ctx_std=self.context.p_std, ctx_proj=self.context.proj,
                            dims=np.array([self.mlp.n_in, self.mlp.n_hidden, self.mlp.n_out]), **extra)

    def export_binary(self, path: str):
        """The emulator, its profile context and calibration (with the spot-check
        correction, when fitted) in the library's file format (irs_emu_load), for callers
        of the C ABI such as the Julia package."""
        cal = self.calibration
        if cal is None:
            raise ValueError("calibrate the emulator before exporting it")
        L, n_feat = self.x_mean.shape
        cx = self.context
        corrected = cal.W is not None
        blocks = list(cal.spot_blocks)
        w_raw, w_lbo = (cal.w_raw, cal.w_lbo) if corrected else (1.0, 0.0)
        with open(path, "wb") as f:
            f.write(b"IRSEMU01")
            f.write(np.array([self.mlp.n_in, self.mlp.n_hidden, self.mlp.n_out, L, self.n_pc, self.n_ch, n_feat,
                              N_LOCAL, cx.proj.shape[0], cx.proj.shape[1], len(blocks), int(corrected)],
                             np.int32).tobytes())
            params = self.mlp.get_params()
            f.write(np.float32(self.coef_scale).tobytes())
            f.write(np.int64(params.size).tobytes())
            for a in (params, self.x_mean, self.x_std, self.A, self.basis, self.mean, self.inv_w):
                f.write(np.ascontiguousarray(a, np.float32).tobytes())
            for a in (self.ch_nu, cx.p_mean, cx.proj):
                f.write(np.ascontiguousarray(a, np.float64).tobytes())
            f.write(np.array([cal.threshold, w_raw, w_lbo], np.float64).tobytes())
            f.write(np.array([c for c, _ in blocks], np.int64).tobytes())
            f.write(np.array([n for _, n in blocks], np.int32).tobytes())
            if corrected:
                f.write(np.ascontiguousarray(cal.W, np.float32).tobytes())
                f.write(np.ascontiguousarray(cal.lbo_W, np.float32).tobytes())

    @staticmethod
    def load(path: str, max_batch: int = 4096):
        z = np.load(path, allow_pickle=False)
        n_in, H, n_out = (int(v) for v in z["dims"])
        mlp = MLP(n_in, H, n_out, max_batch)
        mlp.set_params(z["params"])
        emu = ChannelODEmulator(mlp, z["x_mean"], z["x_std"], z["A"], float(z["coef_scale"][0]), z["basis"],
                                z["mean"], z["inv_w"], z["ch_nu"],
                                ProfileContext(z["ctx_mean"], z["ctx_std"], z["ctx_proj"]))
        if "cal_spot" in z:
            thr, budget, n, w_raw, w_lbo, rate = z["cal_thr"]
            corr = "cal_W" in z
            emu.calibration = Calibration(z["cal_spot"], [tuple(int(v) for v in b) for b in z["cal_blocks"]],
                                          float(thr), float(budget), z["cal_ch_rms"], z["cal_ch_max"],
                                          z["cal_scene_max"], z["cal_scene_spot"], int(n),
                                          W=z["cal_W"] if corr else None, lbo_W=z["cal_lbo"] if corr else None,
                                          w_raw=float(w_raw), w_lbo=float(w_lbo),
                                          scene_max_corrected=z["cal_scene_max_corr"] if corr else None,
                                          fallback_rate=float(rate))
            emu._set_correction(emu.calibration)
        return emu

    # ---- inference ----------------------------------------------------------
    def _inputs(self, scenes, eps):
        X = emulator_features(scenes, self.context, self.T_method)
        tlev, tave, scal = [], [], []
        for q, (p, ts, mu) in enumerate(scenes):
            tlev.append(p.temperature)
            tave.append(layer_properties(p, self.T_method)["T_ave"])
            scal.append([p.temperature[0] if ts is None else ts, mu, eps[q]])
        return (np.ascontiguousarray(X, np.float32), np.ascontiguousarray(np.stack(tlev)),
                np.ascontiguousarray(np.stack(tave)), np.ascontiguousarray(np.array(scal, np.float64)))

    @staticmethod
    def _scenes(profs, geom, T_sfc, eps_sfc):
        nb = len(profs)
        geoms = geom if isinstance(geom, (list, tuple)) else [geom] * nb
        mus = [1.0 if g is None else g.mu for g in geoms]
        tsf = T_sfc if isinstance(T_sfc, (list, tuple, np.ndarray)) else [T_sfc] * nb
        eps = eps_sfc if isinstance(eps_sfc, (list, tuple, np.ndarray)) else [eps_sfc] * nb
        return [(profs[q], tsf[q], mus[q]) for q in range(nb)], list(eps)

    def forward(self, profs, *, geom=None, T_sfc=None, eps_sfc=1.0, source_function="cim",
                want_t: bool = False):
        """Emulated (R, BT) [n_prof, n_ch] for a list of profiles."""
        scenes, eps = self._scenes(list(profs), geom, T_sfc, eps_sfc)
        nb = len(scenes)
        if source_function.lstrip(":") != "cim":
            raise ValueError("the emulator is trained for source_function=:cim")
        X, tlev, tave, scal = self._inputs(scenes, eps)
        if X.shape[1] != self.L:
            raise ValueError(f"profiles have {X.shape[1]} layers, the emulator {self.L}")
        R = np.zeros((nb, self.n_ch))
        BT = np.zeros((nb, self.n_ch))
        tout = np.zeros((nb, self.L, self.n_ch), np.float32) if want_t else None
        rc = self.lib.irs_emu_forward(self.h, _p(X), _p(tlev), _p(tave), _p(scal), nb, api.SOURCE_FUNCTION["cim"],
                                      _p(R), _p(BT), _p(tout))
        if rc != IRS_OK:
            raise IRSError(rc, self.lib.irs_emu_last_error(self.h).decode(errors="replace"))
        return (R, BT, tout) if want_t else (R, BT)

    # ---- Jacobians ----------------------------------------------------------
    def jacobian(self, profs, *, geom=None, T_sfc=None, eps_sfc=1.0) -> List["EmulatorJacobian"]:
        """Emulator Jacobians per scene (levels: T and ln q_H2O; T_sfc); the
        reflected term (eps < 1) is held fixed."""
        scenes, eps = self._scenes(list(profs), geom, T_sfc, eps_sfc)
        return _emu_jacobian(self, scenes, eps)

    # ---- calibration --------------------------------------------------------
    def calibrate(self, profs, linelists, *, sounder: Sounder, geom=None, T_sfc=None, eps_sfc=1.0,
                  budget: float = 0.05, margin: float = 0.2, n_blocks: int = 6, block: int = 64,
                  ctx: Optional[Context] = None, line_mixing="vpy", apply_continuum: bool = False,
                  bt_lbl=None, correct: bool = True, lam: float = 1e-3, folds: int = 4, **kw) -> Calibration:
        """Hold-out statistics on calibration scenes against GPU LBL (the per-channel
        error) and the fallback rule that keeps every answered scene within
        cap = budget * (1 - margin) on every calibration scene.

        correct=True: the emulator's errors are low-rank across channels, so the errors
        measured at the spot channels predict them everywhere.  n_blocks spot blocks of
        `block` channels are picked greedily (select_correction_blocks), a ridge map W from
        the spot errors to every channel's error is fitted on the calibration scenes, and a
        scene is kept when both its largest spot error and the largest leave-one-block-out
        residual of the map are within their thresholds, chosen together
        (fallback_thresholds) to keep the most calibration scenes with none above cap after
        correction, judged out of fold (`folds` folds); the library tests
        max(raw / t_raw, residual / t_lbo) against 1.  correct=False: spot blocks by
        select_spot_blocks and the threshold on the uncorrected errors."""
        scenes, eps = self._scenes(list(profs), geom, T_sfc, eps_sfc)
        _, bt = self.forward([s[0] for s in scenes], geom=[_geom(s[2]) for s in scenes], T_sfc=[s[1] for s in scenes],
                             eps_sfc=eps)
        if bt_lbl is None:
            _, _, bt_lbl = forward_model([s[0] for s in scenes], linelists, sounder=sounder,
                                         geom=[_geom(s[2]) for s in scenes], T_sfc=[s[1] for s in scenes],
                                         eps_sfc=eps, ctx=ctx, line_mixing=line_mixing,
                                         apply_continuum=apply_continuum, **kw)
        E = bt - bt_lbl
        d = np.abs(E)
        ch_rms = np.sqrt(np.mean(d ** 2, axis=0))
        ch_max = d.max(axis=0)
        cap = budget * (1.0 - margin)
        fe = d.max(axis=1)
        self._set_correction(None)
        if not correct:
            blocks = select_spot_blocks(d, ch_rms, n_blocks, block, cap)
            spot = np.concatenate([np.arange(c0, c0 + n) for c0, n in blocks])
            se = d[:, spot].max(axis=1)
            thr = spot_threshold(se, fe, cap)
            self.calibration = Calibration(spot, blocks, thr, budget, ch_rms, ch_max, fe, se, len(scenes),
                                           fallback_rate=float(np.mean(se > thr)))
        else:
            blocks = select_correction_blocks(E, n_blocks, block, lam)
            spot = np.concatenate([np.arange(c0, c0 + n) for c0, n in blocks])
            bn = [n for _, n in blocks]
            res, st = _out_of_fold(E, spot, bn, lam, folds)
            t_raw, t_lbo = fallback_thresholds(st[0], st[1], res > cap)
            if t_raw < 0.0:                     # no calibration scene can be kept
                thr, w_raw, w_lbo = -1.0, 1.0, 0.0
            else:
                thr = 1.0
                w_raw = 0.0 if np.isinf(t_raw) else 1.0 / max(t_raw, 1e-300)
                w_lbo = 0.0 if np.isinf(t_lbo) else 1.0 / max(t_lbo, 1e-300)
            stat = np.maximum(w_raw * st[0], w_lbo * st[1])
            W, lbo = fit_spot_correction(E, spot, bn, lam)
            self.calibration = Calibration(spot, blocks, thr, budget, ch_rms, ch_max, fe, stat, len(scenes),
                                           W=W, lbo_W=lbo, w_raw=w_raw, w_lbo=w_lbo, scene_max_corrected=res,
                                           fallback_rate=float(np.mean(stat > thr)))
            self._set_correction(self.calibration)
        self._spot = None
        return self.calibration

    def _set_correction(self, cal: Optional[Calibration]):
        if cal is None or cal.W is None:
            rc = self.lib.irs_emu_set_correction(self.h, 0, None, None, None, None, 0.0, 0.0)
        else:
            c0 = np.ascontiguousarray([c for c, _ in cal.spot_blocks], np.int64)
            bn = np.ascontiguousarray([n for _, n in cal.spot_blocks], np.int32)
            W = np.ascontiguousarray(cal.W, np.float32)
            lbo = np.ascontiguousarray(cal.lbo_W if cal.lbo_W is not None and len(cal.lbo_W) else np.zeros(1),
                                       np.float32)
            rc = self.lib.irs_emu_set_correction(self.h, len(bn), _p(c0), _p(bn), _p(W), _p(lbo), float(cal.w_raw),
                                                 float(cal.w_lbo))
        if rc != IRS_OK:
            raise IRSError(rc, self.lib.irs_emu_last_error(self.h).decode(errors="replace"))

    def spot_checker(self, linelists, sounder: Sounder) -> "SpotChecker":
        if self.calibration is None:
            raise ValueError("the emulator has no calibration (run calibrate)")
        if self._spot is None or self._spot.linelists is not linelists:
            sg = sounder_grid(sounder)
            self._spot = SpotChecker(linelists, sounder,
                                     blocks=[(float(sg.nu[c0]), n) for c0, n in self.calibration.spot_blocks])
        return self._spot


def spot_threshold(se, fe, cap) -> float:
    """Largest spot error t such that every scene with spot error <= t has its largest
    error over all channels <= cap (0 when the scene with the smallest spot error
    already exceeds it)."""
    thr = 0.0
    for t in np.sort(se):
        if np.all(fe[se <= t] <= cap):
            thr = float(t)
        else:
            break
    return thr


def fallback_thresholds(s_raw, s_lbo, bad):
    """Thresholds (t_raw, t_lbo) of the two fallback statistics, chosen together: a
    scene is kept when s_raw <= t_raw and s_lbo <= t_lbo.  Among the rules that keep
    no bad scene, the one that keeps the most scenes, with each threshold at the
    largest value of a kept scene (inf: that statistic unused).  (-1, -1) when no
    scene can be kept."""
    s_raw, s_lbo, bad = np.asarray(s_raw, float), np.asarray(s_lbo, float), np.asarray(bad, bool)
    best, best_n, best_t = None, 0, None
    for t1 in np.concatenate([np.unique(s_raw), [np.inf]]):
        cand = s_raw <= t1
        m = bad & cand
        lim = s_lbo[m].min() if m.any() else np.inf
        keep = cand & (s_lbo < lim)
        n = int(keep.sum())
        if n > best_n:
            best, best_n, best_t = keep, n, (t1, lim)
    if best is None:
        return -1.0, -1.0
    t_raw = np.inf if np.isinf(best_t[0]) else float(s_raw[best].max())
    t_lbo = np.inf if np.isinf(best_t[1]) else float(s_lbo[best].max())
    return t_raw, t_lbo


def select_spot_blocks(d, ch_rms, n_blocks: int, block: int, cap: float, stride: int = 4, q: float = 0.9):
    """Spot blocks on an error matrix d [scenes, channels] as a greedy cover of the bad
    scenes (those exceeding cap on some channel): each step adds the non-overlapping
    block of `block` channels (starts every `stride` channels) that flags the most bad
    scenes not flagged yet -- a block flags a scene when the scene's largest error in
    it exceeds the q-quantile of the good scenes' -- ties broken by the block's summed
    hold-out RMS error.  Returns [(first, block)] sorted."""
    n, nch = d.shape
    if nch < block:
        return [(0, nch)]
    from numpy.lib.stride_tricks import sliding_window_view
    starts = np.arange(0, nch - block + 1, stride)
    dc = sliding_window_view(d, block, axis=1)[:, starts].max(axis=2)       # [n, m]
    score = np.convolve(ch_rms, np.ones(block), mode="valid")[starts]
    bad = d.max(axis=1) > cap
    good = ~bad
    lim = np.quantile(dc[good], q, axis=0) if good.any() else np.full(len(starts), np.inf)
    flags = dc > lim[None, :]                                                  # [n, m]
    open_ = bad.copy()
    chosen = []
    free = np.ones(len(starts), bool)
    for _ in range(n_blocks):
        if not free.any():
            break
        gain = np.where(free, (flags & open_[:, None]).sum(axis=0), -1)
        best = gain.max()
        pick = np.flatnonzero(gain == best)
        k = int(pick[np.argmax(score[pick])])
        chosen.append((int(starts[k]), block))
        open_ &= ~flags[:, k]
        free &= np.abs(starts - starts[k]) >= block
    chosen.sort()
    return chosen


def _ridge(X, Y, lam):
    """argmin_W |X W - Y|^2 + lam tr(X^T X)/n |W|^2 (FP64)."""
    G = X.T @ X
    return np.linalg.solve(G + lam * max(np.trace(G), 1e-30) / len(G) * np.eye(len(G)), X.T @ Y)


def fit_spot_correction(E, spot, block_n, lam: float = 1e-3):
    """Ridge maps from the spot errors E[:, spot] to every channel's error (W
    [n_spot, n_ch]) and, per spot block b, from the other blocks' errors to block
    b's (leave-one-block-out maps M_b [n_spot - n_b, n_b], concatenated row-major)."""
    E = np.asarray(E, np.float64)
    Es = E[:, spot]
    W = _ridge(Es, E, lam).astype(np.float32)
    lbo, off = [], 0
    for n in block_n:
        inb = np.zeros(len(spot), bool)
        inb[off:off + n] = True
        lbo.append(_ridge(Es[:, ~inb], Es[:, inb], lam).astype(np.float32).ravel() if len(block_n) > 1
                   else np.zeros(0, np.float32))
        off += n
    return W, np.concatenate(lbo)


def spot_statistics(es, block_n, lbo):
    """Fallback statistics of scenes from their spot errors es [n, n_spot]: (largest
    |error|, largest leave-one-block-out residual) -- the library's irs_forward_emulated
    computes the same."""
    es = np.asarray(es, np.float64)
    raw = np.abs(es).max(axis=1)
    res = np.zeros(len(es))
    if len(block_n) > 1:
        off, o = 0, 0
        ns = es.shape[1]
        for n in block_n:
            inb = np.zeros(ns, bool)
            inb[off:off + n] = True
            M = np.asarray(lbo[o:o + (ns - n) * n], np.float64).reshape(ns - n, n)
            res = np.maximum(res, np.abs(es[:, inb] - es[:, ~inb] @ M).max(axis=1))
            off += n
            o += (ns - n) * n
    return raw, res


def _out_of_fold(E, spot, block_n, lam, folds):
    """Per calibration scene, out of fold: the largest corrected error over all
    channels (spot channels exact) and both fallback statistics."""
    E = np.asarray(E, np.float64)
    n = len(E)
    k = max(2, min(folds, n // 4)) if n >= 8 else 1
    parts = np.array_split(np.random.default_rng(0).permutation(n), k) if k > 1 else [np.arange(n)]
    res = np.zeros(n)
    st = [np.zeros(n), np.zeros(n)]
    for f in parts:
        tr = np.setdiff1d(np.arange(n), f) if k > 1 else f
        W, lbo = fit_spot_correction(E[tr], spot, block_n, lam)
        r = E[f] - E[f][:, spot] @ W.astype(np.float64)
        r[:, spot] = 0.0
        res[f] = np.abs(r).max(axis=1)
        raw, lb = spot_statistics(E[f][:, spot], block_n, lbo)
        st[0][f], st[1][f] = raw, lb
    return res, st


def select_correction_blocks(E, n_blocks: int, block: int, lam: float = 1e-3, gap: Optional[int] = None,
                             seed: int = 0, stride: Optional[int] = None, max_scenes: int = 512):
    """Spot blocks for the correction: greedily add the block of `block` channels
    (starts every `stride` channels, by default `block`; at least `gap` channels from the
    chosen ones, by default 64, less on a window too narrow for n_blocks such gaps) whose
    errors, with those already chosen, best predict every channel's error -- the
    smallest mean over scenes of the largest corrected error, fitted on half of (at most
    max_scenes of) the calibration scenes and judged on the other half.  Returns
    [(first, block)] sorted."""
    E = np.asarray(E, np.float64)
    n, nch = E.shape
    if nch <= block:
        return [(0, nch)]
    if gap is None:
        gap = int(min(64, max(0, nch // (2 * max(n_blocks, 1)) - block)))
    idx = np.random.default_rng(seed).permutation(n)[:max_scenes]
    n = len(idx)
    A, B = (idx[: n // 2], idx[n // 2:]) if n >= 4 else (idx, idx)
    starts = np.arange(0, nch - block + 1, max(stride or block, 1))
    chosen = []
    for _ in range(n_blocks):
        best, bv = None, np.inf
        for s0 in starts:
            if any(abs(int(s0) - c) < block + gap for c in chosen):
                continue
            S = np.concatenate([np.arange(c, c + block) for c in chosen + [int(s0)]])
            W = _ridge(E[A][:, S], E[A], lam)
            r = E[B] - E[B][:, S] @ W
            r[:, S] = 0.0
            v = float(np.mean(np.abs(r).max(axis=1)))
            if v < bv:
                bv, best = v, int(s0)
        if best is None:
            break
        chosen.append(best)
    return sorted((c, block) for c in chosen)


def correct_spectra(BT, bt_spot, spot, W, nu):
    """Host restatement of the library's correction: BT -= (BT[:, spot] - bt_spot) W,
    spot channels set to bt_spot, R the Planck radiance of the corrected BT."""
    es = (BT[:, spot] - bt_spot).astype(np.float32)
    BTc = BT - (es @ np.asarray(W, np.float32)).astype(np.float64)
    BTc[:, spot] = bt_spot
    R = api.planck_radiance(nu, BTc)
    return R, BTc, es


def _feature_level_jacobian(prof, T_sfc, mu, context: ProfileContext, T_method, hT=1e-2, hq=1e-3):
    """d features / d (level T), d features / d (level ln q_H2O), d T_ave / d (level T),
    d features / d T_sfc: central differences of the host layer code for the layer
    features, analytic for the (linear) profile context.  A level bounds two layers
    and the layer quantities of layer_properties depend on a layer's own two levels
    only, so each perturbed level recomputes those layers from its three-level
    neighbourhood; the features of every perturbed profile (the columns and
    Curtis-Godson means above each layer) are then evaluated together."""
    nl = prof.n_levels
    L = nl - 1
    q0 = np.stack(_layer_quantities(layer_properties(prof, T_method), L))          # [6, L]
    has_q = GasSpecies.H2O in prof.vmr
    steps = [("T", hT), ("T", -hT)] + ([("q", hq), ("q", -hq)] if has_q else [])
    Q = np.repeat(q0[None, None], nl, axis=0).repeat(len(steps), axis=1)           # [nl, n_steps, 6, L]
    for j in range(nl):
        lo, hi = max(j - 1, 0), min(j + 1, nl - 1)
        for m, (kind, h) in enumerate(steps):
            T = prof.temperature[lo:hi + 1].copy()
            vmr = {sp: v[lo:hi + 1].copy() for sp, v in prof.vmr.items()}
            if kind == "T":
                T[j - lo] += h
            else:
                vmr[GasSpecies.H2O][j - lo] *= np.exp(h)
            sub = AtmosphericProfile(prof.pressure[lo:hi + 1], T, prof.altitude[lo:hi + 1], vmr)
            Q[j, m, :, lo:hi] = np.stack(_layer_quantities(layer_properties(sub, T_method), hi - lo))
    X = _features_from_layers(*(Q[:, :, i] for i in range(6)), mu)                 # [nl, n_steps, L, 13]
    nf = X.shape[-1]
    dXdT = np.zeros((nl, L, N_FEATURES))
    dXdq = np.zeros((nl, L, N_FEATURES))
    dXdT[:, :, :nf] = (X[:, 0] - X[:, 1]) / (2 * hT)
    dAdT = (Q[:, 0, 2] - Q[:, 1, 2]) / (2 * hT)
    if has_q:
        dXdq[:, :, :nf] = (X[:, 2] - X[:, 3]) / (2 * hq)
    # the context is linear in [T levels, ln q_H2O levels, ln q_O3 levels, T_sfc, 1/mu]
    P = context.proj
    dXdT[:, :, nf:] = P[:, :nl].T[:, None, :]
    dXdq[:, :, nf:] = P[:, nl:2 * nl].T[:, None, :]
    dXdTs = np.zeros((L, N_FEATURES))
    dXdTs[:, nf:] = P[:, 3 * nl][None, :]
    if T_sfc is None:                      # T_sfc defaults to the bottom level temperature
        dXdT[0, :, nf:] += P[:, 3 * nl][None, :]
    return dXdT, dAdT, dXdq, dXdTs


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


def _emu_jacobian(emu, scenes, eps):
    nb = len(scenes)
    X, tlev, tave, scal = emu._inputs(scenes, eps)
    L, nch = emu.L, emu.n_ch
    BT = np.zeros((nb, nch))
    dfeat = np.zeros((nb, L, N_FEATURES, nch), np.float32)
    dTave = np.zeros((nb, L, nch), np.float32)
    dTlev = np.zeros((nb, L + 1, nch), np.float32)
    dTsfc = np.zeros((nb, nch), np.float32)
    rc = emu.lib.irs_emu_jacobian(emu.h, _p(X), _p(tlev), _p(tave), _p(scal), nb, _p(BT), _p(dfeat), _p(dTave),
                                  _p(dTlev), _p(dTsfc))
    if rc != IRS_OK:
        raise IRSError(rc, emu.lib.irs_emu_last_error(emu.h).decode(errors="replace"))
    out = []
    for q, (prof, ts, mu) in enumerate(scenes):
        dXdT, dAdT, dXdq, dXdTs = _feature_level_jacobian(prof, ts, mu, emu.context, emu.T_method)
        F = dfeat[q].astype(np.float64).reshape(L * N_FEATURES, nch)                 # [(k, i), c]
        nl = dXdT.shape[0]
        KT = (dXdT.reshape(nl, -1) @ F).T + (dAdT @ dTave[q].astype(np.float64)).T
        KT += dTlev[q].T.astype(np.float64)
        Kq = (dXdq.reshape(nl, -1) @ F).T
        Ks = dXdTs.reshape(-1) @ F
        if ts is None:                      # T_sfc is the bottom level temperature: its Planck term goes there
            KT[:, 0] += dTsfc[q].astype(np.float64)
        else:
            Ks = Ks + dTsfc[q].astype(np.float64)
        out.append(EmulatorJacobian(BT[q], KT, Kq, Ks, dfeat[q]))
    return out


# ---- spot check + fallback -----------------------------------------------------
class SpotChecker:
    """LBL of a few blocks of consecutive channels, each on its own internal
    sub-grid (channel block +- the 16 cm^-1 ILS pad) with only the lines that
    reach it (+- 25 cm^-1 + 1): the same channel radiances as the full-range
    LBL at a fraction of its work."""

    def __init__(self, linelists, sounder: Sounder, blocks=((700.0, 16), (960.0, 16), (1050.0, 16), (2300.0, 16))):
        sg = sounder_grid(sounder)
        self.sounder = sounder
        self.linelists = linelists
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
    threshold: float = 0.0


def _uses_correction(emulator, spot) -> bool:
    """The calibration's spot-check correction applies when the spot check runs on
    the calibration's own blocks."""
    cal = emulator.calibration
    if cal is None or cal.W is None:
        return False
    if spot is None:
        return True
    return [(int(ch[0]), len(ch)) for ch, _, _ in spot.blocks] == [(int(a), int(b)) for a, b in cal.spot_blocks]


def forward_emulated(profs, linelists, emulator: ChannelODEmulator, *, sounder: Sounder,
                     geom=None, T_sfc=None, eps_sfc=1.0, threshold: Optional[float] = None,
                     spot: Optional[SpotChecker] = None, ctx: Optional[Context] = None,
                     line_mixing="vpy", apply_continuum: bool = False, internal_dnu: float = 0.001,
                     cutoff: float = 25.0, dptmn: float = 1e-6, T_method: str = "logp_at_pcg",
                     correction: Optional[bool] = None) -> EmulatedResult:
    """Emulator + per-scene LBL spot check (+ the calibration's correction of every
    channel) + full-LBL fallback in one library call (irs_forward_emulated).  The spot
    channels and the threshold come from the emulator's calibration (calibrate) unless
    `spot` / `threshold` are given; a scene whose fallback statistic exceeds the
    threshold is recomputed with full LBL.  correction=None applies the calibration's
    correction when the spot check runs on the calibration's blocks, False never, True
    always (the library then refuses other blocks).  apply_continuum as in training."""
    scenes, eps = ChannelODEmulator._scenes(list(profs), geom, T_sfc, eps_sfc)
    ctx = ctx or api.default_context()
    cal = emulator.calibration
    use_corr = _uses_correction(emulator, spot) if correction is None else bool(correction)
    if spot is None:
        spot = emulator.spot_checker(linelists, sounder)
    if threshold is None:
        if cal is None:
            raise ValueError("no threshold given and the emulator has no calibration")
        if cal.W is not None and not use_corr:
            raise ValueError("the calibrated threshold belongs to the calibration's spot blocks; pass a threshold")
        threshold = cal.threshold
    opt = api._options(internal_dnu, None, cutoff, None, dptmn, 122.0, True, "gaussian", line_mixing, T_method,
                       "cim", api._continua_mask(apply_continuum, api.DEFAULT_CONTINUA))
    b = api._Batch(ctx, [sc[0] for sc in scenes], linelists, [sc[1] for sc in scenes], eps, [sc[2] for sc in scenes])
    species = [GasSpecies(int(v)) for v in b.species]
    c0 = np.array([int(ch[0]) for ch, _, _ in spot.blocks], np.int64)
    nb = np.array([len(ch) for ch, _, _ in spot.blocks], np.int32)
    lists = np.array([[ctx.upload(sub[sp]) if sp in sub and len(sub[sp]) else -1 for sp in species]
                      for _, _, sub in spot.blocks], np.int32).reshape(-1)
    n = len(scenes)
    R = np.zeros((n, emulator.n_ch))
    BT = np.zeros((n, emulator.n_ch))
    fb = np.zeros(n, np.uint8)
    err = np.zeros(n)
    snd = api._sounder(sounder)
    t0 = time.perf_counter()
    if cal is not None and cal.W is not None and not use_corr:
        emulator._set_correction(None)              # other spot blocks: the uncorrected emulator
    try:
        rc = emulator.lib.irs_forward_emulated(ctx.h, emulator.h, b.arr, n, C.byref(snd), C.byref(opt), len(c0),
                                               _p(c0), _p(nb), _p(lists), float(threshold), _p(R), _p(BT), _p(fb),
                                               _p(err))
    finally:
        if cal is not None and cal.W is not None and not use_corr:
            emulator._set_correction(cal)
    if rc != IRS_OK:
        msg = emulator.lib.irs_emu_last_error(emulator.h).decode(errors="replace") or \
            emulator.lib.irs_last_error(ctx.h).decode(errors="replace")
        raise (ValueError(msg) if rc == api.IRS_EINVAL else IRSError(rc, msg))
    return EmulatedResult(R, BT, fb.astype(bool), err, {"total": time.perf_counter() - t0}, float(threshold))


def forward_emulated_steps(profs, linelists, emulator: ChannelODEmulator, *, sounder: Sounder,
                           geom=None, T_sfc=None, eps_sfc=1.0, threshold: Optional[float] = None,
                           spot: Optional[SpotChecker] = None, ctx: Optional[Context] = None,
                           line_mixing="vpy", apply_continuum: bool = False, **kw) -> EmulatedResult:
    """The same decision as forward_emulated from host-side steps (emulator forward,
    spot-check forward_model calls, fallback forward_model), timing each step."""
    kw["apply_continuum"] = apply_continuum
    scenes, eps = ChannelODEmulator._scenes(list(profs), geom, T_sfc, eps_sfc)
    geoms = [_geom(s[2]) for s in scenes]
    tsf = [s[1] for s in scenes]
    ctx = ctx or api.default_context()
    if spot is None:
        spot = emulator.spot_checker(linelists, sounder)
    if threshold is None:
        if emulator.calibration is None:
            raise ValueError("no threshold given and the emulator has no calibration")
        threshold = emulator.calibration.threshold
    t0 = time.perf_counter()
    R, BT = emulator.forward(list(profs), geom=geoms, T_sfc=tsf, eps_sfc=eps)
    t1 = time.perf_counter()
    bt_spot = spot(profs, geoms, tsf, eps, ctx, line_mixing=line_mixing, **kw)
    err = np.max(np.abs(BT[:, spot.channels] - bt_spot), axis=1)
    cal = emulator.calibration
    if _uses_correction(emulator, spot):
        R, BT, es = correct_spectra(BT, bt_spot, spot.channels, cal.W, emulator.ch_nu)
        raw, res = spot_statistics(es, [len(b[0]) for b in spot.blocks], cal.lbo_W)
        err = np.maximum(cal.w_raw * raw, cal.w_lbo * res)
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
    return EmulatedResult(R, BT, fb, err, {"emulator": t1 - t0, "spot_check": t2 - t1, "fallback": t3 - t2},
                          float(threshold))
