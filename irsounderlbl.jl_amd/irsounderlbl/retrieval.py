"""Optimal-estimation retrieval driven by the GPU forward model and analytic Jacobian.

Mirrors src/Estimation/optimal_estimation.jl (the Levenberg-Marquardt / Gauss-Newton
MAP loop, Rodgers 2000 Eq. 5.36, and its diagnostics), src/Estimation/covariance.jl
(measurement-error covariances) and src/Estimation/sa_builder.jl (a-priori
covariances).  Every forward evaluation and every Jacobian is one call into libirs
(`forward_model` with dptmn=0, `analytic_jacobian`); the state-space linear algebra
(n_state x n_state, n_state ~ 100) is FP64 host work that takes well under a
millisecond per iteration next to a 0.2 s full-IASI Jacobian.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Dict, Optional, Sequence

import numpy as np

from .types import AtmosphericProfile, GasSpecies, HITRANLinelist, StateVectorSpec, unpack_state

C1 = 1.191042953e-5     # planck.jl:11
C2 = 1.4387769          # planck.jl:12


@dataclass
class RetrievalResult:
    """RetrievalResult (optimal_estimation.jl:57-76).  K, y_fit and nu are on the full
    channel grid; G, A, S_hat, S_noise, S_smoothing, dof, dfn, H and chi2 cover the
    channels kept by channel_mask."""
    x: np.ndarray
    spec: object
    converged: bool
    n_iter: int
    cost: np.ndarray
    S_hat: np.ndarray
    A: np.ndarray
    G: np.ndarray
    K: np.ndarray
    dof: float
    dfn: float
    H: float
    S_noise: np.ndarray
    S_smoothing: np.ndarray
    chi2: float
    y_fit: np.ndarray
    nu: np.ndarray
    channel_mask: np.ndarray

    def __repr__(self):
        return (f"RetrievalResult({'converged' if self.converged else 'NOT converged'} in {self.n_iter} "
                f"iter, DOF={self.dof:.2f}, H={self.H:.2f} bits, chi2={self.chi2:.3f}, {self.spec})")


class _SeFactor:
    """Se on the kept channels, factored once: a 1-D Se is a diagonal (O(n) solves), a
    2-D Se is Cholesky-factored on its kept sub-block (optimal_estimation.jl:178-179)."""

    def __init__(self, Se, kidx):
        Se = np.asarray(Se, np.float64)
        if Se.ndim == 1:
            self.diag = Se[kidx].copy()
            self.dense = np.diag(self.diag)
            self.L = None
        else:
            self.diag = None
            self.dense = Se[np.ix_(kidx, kidx)].copy()
            self.L = np.linalg.cholesky(0.5 * (self.dense + self.dense.T))

    def solve(self, b):
        if self.L is None:
            return b / (self.diag if b.ndim == 1 else self.diag[:, None])
        z = np.linalg.solve(self.L, b)
        return np.linalg.solve(self.L.T, z)


def _oe_cost(resid, sei_r, dx, sa_inv_dx) -> float:
    """J = r' Se^-1 r + dx' Sa^-1 dx (optimal_estimation.jl:92-96)."""
    return float(resid @ sei_r + dx @ sa_inv_dx)


def _oe_solve(y, n: int, forward: Callable, forward_jac: Callable, *, xa, Sa, Se, channel_mask=None,
              x0=None, method: str = "levenberg_marquardt", max_iter: int = 15, max_linesearch: int = 8,
              gamma0: float = 1.0, gamma_factor: float = 3.0, gamma_min: float = 1e-4,
              conv_factor: float = 1e-1, verbose: bool = False, spec=None) -> RetrievalResult:
    """The retrieval loop and diagnostics of optimal_estimation (optimal_estimation.jl:154-295)
    over a model given as forward(x) -> F and forward_jac(x) -> (F, K, nu)."""
    if method not in ("levenberg_marquardt", "gauss_newton"):
        raise ValueError(f"method must be :levenberg_marquardt or :gauss_newton, got :{method}")
    xa = np.asarray(xa, np.float64)
    Sa = np.asarray(Sa, np.float64)
    y = np.asarray(y, np.float64)
    if len(xa) != n:
        raise ValueError(f"xa length {len(xa)} ≠ spec.n {n}")
    if Sa.shape != (n, n):
        raise ValueError(f"Sa must be {n}×{n}")
    Se_a = np.asarray(Se, np.float64)
    ny = Se_a.shape[0]
    if not (len(y) == ny and (Se_a.ndim == 1 or Se_a.shape == (ny, ny))):
        raise ValueError(f"Se must be n_y×n_y with n_y = length(y) = {len(y)}")
    keep = np.ones(len(y), bool) if channel_mask is None else np.asarray(channel_mask, bool)
    if len(keep) != len(y):
        raise ValueError(f"channel_mask length {len(keep)} ≠ length(y) {len(y)}")
    if not keep.any():
        raise ValueError("channel_mask excludes every channel")
    kidx = np.flatnonzero(keep)

    Sa_inv = np.linalg.inv(Sa)
    y_k = y[kidx]
    se = _SeFactor(Se_a, kidx)

    x = np.array(xa if x0 is None else x0, np.float64)
    F, Kfull, nu = forward_jac(x)
    K = Kfull[kidx, :]
    resid = y_k - F[kidx]
    sei_r = se.solve(resid)
    dx_a = x - xa
    J = _oe_cost(resid, sei_r, dx_a, Sa_inv @ dx_a)
    cost = [J]

    gamma = 0.0 if method == "gauss_newton" else gamma0
    converged = False
    it = 0
    while it < max_iter:
        it += 1
        seik = se.solve(K)
        ktseik = K.T @ seik
        ktseik = 0.5 * (ktseik + ktseik.T)
        g = K.T @ sei_r - Sa_inv @ (x - xa)
        accepted = False
        for _ in range(max_linesearch):
            H = (1.0 + gamma) * Sa_inv + ktseik
            dx = np.linalg.solve(0.5 * (H + H.T), g)
            x_try = x + dx
            F_try = forward(x_try)
            r_try = y_k - F_try[kidx]
            sei_rt = se.solve(r_try)
            dxa_t = x_try - xa
            J_try = _oe_cost(r_try, sei_rt, dxa_t, Sa_inv @ dxa_t)
            if J_try < J:
                # Rodgers step size d^2 = dx' S_hat^-1 dx, per parameter
                d2n = float(dx @ (Sa_inv @ dx + ktseik @ dx)) / n
                x = x_try
                F, Kfull, nu = forward_jac(x)
                K = Kfull[kidx, :]
                resid = y_k - F[kidx]
                sei_r = se.solve(resid)
                J = J_try
                cost.append(J)
                if method != "gauss_newton":
                    gamma = max(gamma / gamma_factor, gamma_min)
                accepted = True
                if verbose:
                    kk = K.T @ se.solve(K)
                    trA = float(np.trace(np.linalg.solve(Sa_inv + kk, kk)))
                    shat = float(resid @ sei_r) / (len(resid) - trA)
                    print(f"  iter {it}: J={J:.6g} d²/n={d2n:.3g} ŝ={shat:.3g} γ={gamma}", flush=True)
                if d2n < conv_factor:
                    converged = True
                break
            else:
                if method == "gauss_newton":
                    break
                gamma *= gamma_factor
        if converged or not accepted:
            break

    seik = se.solve(K)
    ktseik = K.T @ seik
    ktseik = 0.5 * (ktseik + ktseik.T)
    S_hat = np.linalg.inv(Sa_inv + ktseik)
    S_hat = 0.5 * (S_hat + S_hat.T)
    G = S_hat @ seik.T
    A = G @ K
    dof = float(np.trace(A))
    m = len(resid)
    sign, ld = np.linalg.slogdet(np.eye(n) - A)
    Hbits = -0.5 * ld / math.log(2.0)
    S_noise = G @ (se.dense @ G.T)
    S_noise = 0.5 * (S_noise + S_noise.T)
    ImA = np.eye(n) - A
    S_smooth = ImA @ (Sa @ ImA.T)
    S_smooth = 0.5 * (S_smooth + S_smooth.T)
    chi2 = float(resid @ se.solve(resid))
    return RetrievalResult(x, spec, converged, it, np.array(cost), S_hat, A, G, Kfull, dof, m - dof, Hbits,
                           S_noise, S_smooth, chi2, np.asarray(F, np.float64), np.asarray(nu, np.float64), keep)


def optimal_estimation(y, spec: StateVectorSpec, base_prof: AtmosphericProfile,
                       linelists: Dict[GasSpecies, HITRANLinelist], *, xa, Sa, Se, channel_mask=None,
                       x0=None, method: str = "levenberg_marquardt", max_iter: int = 15,
                       max_linesearch: int = 8, gamma0: float = 1.0, gamma_factor: float = 3.0,
                       gamma_min: float = 1e-4, conv_factor: float = 1e-1, observable: str = "bt",
                       eps_fixed: float = 1.0, fm_kwargs: Optional[dict] = None, verbose: bool = False,
                       ctx=None) -> RetrievalResult:
    """optimal_estimation(y, spec, base_prof, linelists; xa, Sa, Se, ...) -> RetrievalResult
    (src/Estimation/optimal_estimation.jl:134-296).

    F(x) is `forward_model(..., dptmn=0)` and K(x) `analytic_jacobian(...)`, both on the
    GPU; `fm_kwargs` goes to both (sounder/iasi, apply_continuum, with_ils, ...), never
    T_sfc/eps_sfc (they come from the state or `eps_fixed`) or dptmn.  Se is a 1-D vector
    (a diagonal) or a dense n_y x n_y matrix; channel_mask keeps the channels marked True."""
    from .api import analytic_jacobian, default_context, forward_model
    if observable not in ("bt", "radiance"):
        raise ValueError(f"observable must be :bt or :radiance, got :{observable}")
    fm = dict(fm_kwargs or {})
    for k in ("T_sfc", "eps_sfc", "dptmn"):
        if k in fm:
            raise ValueError(f"fm_kwargs must not set {k}")
    c = ctx if ctx is not None else default_context()

    def eps(e):
        return e if spec.include_emissivity else eps_fixed

    def forward(x):
        p, ts, es = unpack_state(spec, x, base_prof)
        _, R, BT = forward_model(p, linelists, T_sfc=ts, eps_sfc=eps(es), dptmn=0.0, ctx=c, **fm)
        return BT if observable == "bt" else R

    def forward_jac(x):
        p, ts, es = unpack_state(spec, x, base_prof)
        jac = analytic_jacobian(p, linelists, spec, T_sfc=ts, eps_sfc=eps(es), observable=observable,
                                ctx=c, **fm)
        return jac.y0, jac.K, jac.nu

    return _oe_solve(y, spec.n, forward, forward_jac, xa=xa, Sa=Sa, Se=Se, channel_mask=channel_mask,
                     x0=x0, method=method, max_iter=max_iter, max_linesearch=max_linesearch,
                     gamma0=gamma0, gamma_factor=gamma_factor, gamma_min=gamma_min,
                     conv_factor=conv_factor, verbose=verbose, spec=spec)


def exclude_channels(nu, *ranges) -> np.ndarray:
    """Keep-mask dropping every channel inside any closed (lo, hi) range
    (optimal_estimation.jl:298-312)."""
    nu = np.asarray(nu, np.float64)
    keep = np.ones(len(nu), bool)
    for lo, hi in ranges:
        keep &= ~((nu >= lo) & (nu <= hi))
    return keep


def ils_noise_autocorr(dnu_sample: float, fwhm_gauss: float, max_lag: int = 4) -> np.ndarray:
    """Channel-lag autocorrelation of the Gaussian apodization weights sampled on the
    channel grid (covariance.jl:27-35); diagonal for fwhm_gauss <= 0."""
    if fwhm_gauss <= 0.0:
        return np.array([1.0] + [0.0] * max_lag)
    s = fwhm_gauss / (2.0 * math.sqrt(2.0 * math.log(2.0)))
    K = max_lag + 6
    w = np.exp(-(np.arange(-K, K + 1) * dnu_sample) ** 2 / (2 * s * s))
    r0 = float(np.sum(w * w))
    return np.array([1.0] + [float(np.sum(w[:len(w) - k] * w[k:])) / r0 for k in range(1, max_lag + 1)])


def apodized_measurement_covariance(nu, sigma, *, dnu_sample: float = 0.25, fwhm_gauss: float = 0.5,
                                    max_lag: int = 4) -> np.ndarray:
    """Se[i,j] = s_i s_j rho(|round((nu_i - nu_j)/dnu_sample)|), zero past max_lag
    (covariance.jl:59-75)."""
    nu = np.asarray(nu, np.float64)
    n = len(nu)
    sv = np.full(n, float(sigma)) if np.ndim(sigma) == 0 else np.asarray(sigma, np.float64)
    if len(sv) != n:
        raise ValueError(f"σ must be a scalar or length(ν)={n}; got {len(sv)}")
    rho = ils_noise_autocorr(dnu_sample, fwhm_gauss, max_lag)
    lag = np.rint(np.abs(nu[:, None] - nu[None, :]) / dnu_sample).astype(np.int64)
    band = lag <= max_lag
    Se = np.zeros((n, n))
    Se[band] = (sv[:, None] * sv[None, :])[band] * rho[lag[band]]
    return Se


def dB_dT(nu, T):
    """dB/dT = C1 C2 nu^4 e^x / (T^2 (e^x - 1)^2), x = C2 nu / T (planck.jl:67-71)."""
    nu = np.asarray(nu, np.float64)
    T = np.asarray(T, np.float64)
    x = C2 * nu / T
    ex = np.exp(x)
    return C1 * C2 * nu ** 4 * ex / (T * T * (ex - 1.0) ** 2)


def scene_nedt(nu, Tb_scene, *, nedt_280K=0.25, T_ref: float = 280.0) -> np.ndarray:
    """NEdT_ref * dB/dT(nu, T_ref) / dB/dT(nu, Tb_scene) (covariance.jl:103-113)."""
    nu = np.asarray(nu, np.float64)
    tb = np.asarray(Tb_scene, np.float64)
    if len(tb) != len(nu):
        raise ValueError(f"Tb_scene must match length(ν)={len(nu)}; got {len(tb)}")
    n0 = np.full(len(nu), float(nedt_280K)) if np.ndim(nedt_280K) == 0 else np.asarray(nedt_280K, np.float64)
    if len(n0) != len(nu):
        raise ValueError(f"nedt_280K must be scalar or length(ν)={len(nu)}; got {len(n0)}")
    return n0 * dB_dT(nu, T_ref) / dB_dT(nu, tb)


def scene_measurement_covariance(nu, Tb_scene, *, nedt_280K=0.25, T_ref: float = 280.0, **kw) -> np.ndarray:
    """Scene NEdT diagonal with the apodization off-diagonals (covariance.jl:122-130)."""
    return apodized_measurement_covariance(nu, scene_nedt(nu, Tb_scene, nedt_280K=nedt_280K, T_ref=T_ref), **kw)


_KERNELS = {
    "exponential": lambda r: np.exp(-r),
    "matern32": lambda r: (1.0 + math.sqrt(3.0) * r) * np.exp(-math.sqrt(3.0) * r),
    "matern52": lambda r: (1.0 + math.sqrt(5.0) * r + 5.0 * r * r / 3.0) * np.exp(-math.sqrt(5.0) * r),
    "sqexp": lambda r: np.exp(-0.5 * r * r),
}


def _resolve_sigma(s, n: int, name: str) -> np.ndarray:
    if np.ndim(s) == 0:
        if not s > 0:
            raise ValueError(f"{name} must be positive")
        return np.full(n, float(s))
    v = np.asarray(s, np.float64)
    if len(v) != n:
        raise ValueError(f"{name} must be a scalar or length n_levels={n}; got {len(v)}")
    if not np.all(v > 0):
        raise ValueError(f"{name} entries must all be positive")
    return v


def _per_species(v, sp, name: str):
    if isinstance(v, dict):
        if sp not in v:
            raise ValueError(f"{name} has no entry for {GasSpecies(sp).name}")
        return v[sp]
    return v


def build_sa(spec: StateVectorSpec, base_prof: AtmosphericProfile, *, sigma_T=2.0, L_T: float = 1.0,
             sigma_vmr=0.5, L_vmr=1.0, sigma_col=0.2, sigma_tsfc: float = 2.0, sigma_emis: float = 0.02,
             kernel: str = "matern52", jitter: float = 0.0) -> np.ndarray:
    """A-priori covariance with log-pressure correlated profile blocks and independent
    scalars (sa_builder.jl:99-153): Sa[i,j] = s_i s_j c(|ln(p_i/p_j)| / L)."""
    if kernel not in _KERNELS:
        raise ValueError(f"kernel must be :matern52, :matern32, :exponential, or :sqexp; got :{kernel}")
    if len(base_prof.pressure) != spec.n_levels:
        raise ValueError(f"base_prof has {len(base_prof.pressure)} levels, spec expects {spec.n_levels}")
    if not np.all(base_prof.pressure > 0):
        raise ValueError("pressures must be positive for log-pressure correlation")
    if jitter < 0:
        raise ValueError("jitter must be ≥ 0")
    logp = np.log(base_prof.pressure)
    corr = _KERNELS[kernel]
    r = np.abs(logp[:, None] - logp[None, :])
    Sa = np.zeros((spec.n, spec.n))

    def block(sig, L):
        return sig[:, None] * sig[None, :] * corr(r / L)

    if spec.include_temperature:
        t = slice(spec.temp_range.start, spec.temp_range.stop)
        Sa[t, t] = block(_resolve_sigma(sigma_T, spec.n_levels, "σ_T"), float(L_T))
    for b in spec.vmr_blocks:
        s = slice(b.range.start, b.range.stop)
        if b.basis is None:
            L = float(_per_species(L_vmr, b.species, "L_vmr"))
            if not L > 0:
                raise ValueError("L_vmr must be positive")
            Sa[s, s] = block(_resolve_sigma(_per_species(sigma_vmr, b.species, "σ_vmr"), spec.n_levels,
                                            f"σ_vmr[{b.species.name}]"), L)
        else:
            sc = _resolve_sigma(_per_species(sigma_col, b.species, "σ_col"), b.m, f"σ_col[{b.species.name}]")
            idx = np.arange(b.range.start, b.range.stop)
            Sa[idx, idx] = sc ** 2
    if spec.include_tsfc:
        Sa[spec.tsfc_index, spec.tsfc_index] = float(sigma_tsfc) ** 2
    if spec.include_emissivity:
        Sa[spec.emis_index, spec.emis_index] = float(sigma_emis) ** 2
    if jitter > 0:
        Sa[np.diag_indices(spec.n)] *= 1.0 + jitter
    return Sa
