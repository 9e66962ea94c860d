"""The host side of the 8-coordinate (Tian-Yau, Schoen) Donaldson solve around the GPU
contracts of :mod:`papers_amd.donaldson_gpu`, restating ``route34/ty_metric.rs``:

* :func:`donaldson_iteration_route34` is ``donaldson_iteration_impl`` (``ty_metric.rs:1063-1343``)
  after its T-operator call: h_pair / sum w, Hermitian symmetrisation, trace normalisation to
  N, the Tikhonov shift on the 2N x 2N real block, LU inversion with the near-singular sentinel,
  Hermitian unpacking, trace renormalisation of the inverse, and the damped update
  ``h <- alpha h_new + (1 - alpha) h`` with residual ``|h_new - h|_F``;
* :class:`AdaptiveDamping` is ``AdaptiveDampingState`` (``schoen_metric.rs:576-625``);
* :func:`solve_route34` is the iteration loop of ``solve_ty_metric`` (``ty_metric.rs:573-708``):
  sigma at the starting metric, the iteration cap scaled by ``max(ceil(2 / alpha_0), 2)`` when
  the first alpha is below 1, the minimum-residual snapshot and the catastrophe guard
  (restore after 5 consecutive residuals above 100x a minimum below 1e-2, or on a non-finite
  residual), and the ``residual < tol`` stop.

The T-operator and sigma are callables, normally ``Cy3DonaldsonGpu.t_operator_raw`` and
``Cy3SigmaGpu.compute_sigma``. Only a constant Tikhonov shift is supported (the reference's
residual-dependent ``TikhonovShift`` schedules and their stall-band gating are not restated).
"""
from __future__ import annotations

import math

import numpy as np
import scipy.linalg as sla

__all__ = ["AdaptiveDamping", "donaldson_iteration_route34", "solve_route34"]

CATASTROPHE_FLOOR = 1.0e-2
CATASTROPHE_MULTIPLIER = 100.0
CATASTROPHE_STREAK = 5


class AdaptiveDamping:
    """AdaptiveDampingState (schoen_metric.rs:576-625): alpha x 1.5 (capped) after 3 residuals
    each below 0.95 x the previous, x 0.7 (floored) after one above 1.05 x the previous."""

    def __init__(self, alpha_initial: float, alpha_min: float, alpha_max: float):
        lo = min(max(alpha_min, 1e-6), 1.0)
        hi = min(max(alpha_max, lo), 1.0)
        self.current_alpha = min(max(alpha_initial, lo), hi)
        self.alpha_min, self.alpha_max = lo, hi
        self.last_residual = math.inf
        self.monotone_streak = 0

    def alpha(self) -> float:
        return self.current_alpha

    def update(self, residual: float):
        if not math.isfinite(residual):
            self.last_residual = residual
            return
        if math.isfinite(self.last_residual):
            if residual < 0.95 * self.last_residual:
                self.monotone_streak += 1
                if self.monotone_streak >= 3:
                    self.current_alpha = min(self.current_alpha * 1.5, self.alpha_max)
                    self.monotone_streak = 0
            elif residual > 1.05 * self.last_residual:
                self.current_alpha = max(self.current_alpha * 0.7, self.alpha_min)
                self.monotone_streak = 0
        self.last_residual = residual


def donaldson_iteration_route34(t_operator, h_re, h_im, weights, damping: float = 1.0,
                                tikhonov: float = 0.0):
    """One iteration; returns (h_re, h_im, residual). On a non-finite weight sum, a singular or
    non-finite inverse, h comes back unchanged with residual NaN (the reference returns NaN
    before its damped update)."""
    h_re = np.asarray(h_re, dtype=np.float64)
    h_im = np.asarray(h_im, dtype=np.float64)
    n = h_re.shape[0]
    pr, pi = t_operator(h_re, h_im)
    pr = np.asarray(pr, dtype=np.float64).reshape(n, n)
    pi = np.asarray(pi, dtype=np.float64).reshape(n, n)
    wsum = float(np.sum(np.asarray(weights, dtype=np.float64)))
    if not wsum > 1e-30:
        return h_re, h_im, float("nan")
    tr_, ti_ = pr * (1.0 / wsum), pi * (1.0 / wsum)
    tr_ = 0.5 * (tr_ + tr_.T)
    ti_ = 0.5 * (ti_ - ti_.T)
    np.fill_diagonal(ti_, 0.0)
    tr = float(np.trace(tr_))
    if abs(tr) > 1e-30:
        tr_, ti_ = tr_ * (n / tr), ti_ * (n / tr)
    block = np.zeros((2 * n, 2 * n))
    block[0::2, 0::2], block[0::2, 1::2] = tr_, -ti_
    block[1::2, 0::2], block[1::2, 1::2] = ti_, tr_
    if math.isfinite(tikhonov) and tikhonov > 0.0:
        block[np.diag_indices(2 * n)] += tikhonov
    try:
        lu, piv = sla.lu_factor(block, check_finite=False)
    except (ValueError, np.linalg.LinAlgError):
        return h_re, h_im, float("nan")
    d = np.abs(np.diag(lu))
    if not np.all(np.isfinite(d)) or np.any(d < np.finfo(np.float64).tiny):
        return h_re, h_im, float("nan")
    inv = sla.lu_solve((lu, piv), np.eye(2 * n), check_finite=False)
    if not np.all(np.isfinite(inv)):
        return h_re, h_im, float("nan")
    nr = 0.5 * (inv[0::2, 0::2] + inv[1::2, 1::2])
    ni = 0.5 * (inv[1::2, 0::2] - inv[0::2, 1::2])
    tri = float(np.trace(nr))
    if abs(tri) > 1e-30:
        nr, ni = nr * (n / tri), ni * (n / tri)
    alpha = min(max(damping, 1e-6), 1.0)
    resid = float(math.sqrt(np.sum((nr - h_re) ** 2) + np.sum((ni - h_im) ** 2)))
    return alpha * nr + (1.0 - alpha) * h_re, alpha * ni + (1.0 - alpha) * h_im, resid


def solve_route34(t_operator, sigma, h_re, h_im, weights, max_iter: int, tol: float,
                  damping=None, tikhonov: float = 0.0) -> dict:
    """The Donaldson loop of solve_ty_metric. damping: None or a float (static alpha, clamped to
    [1e-6, 1]) or an AdaptiveDamping. sigma: callable(h_re, h_im) or None. Returns a dict with
    h_re, h_im, iterations_run, final_donaldson_residual, donaldson_history, sigma_history,
    sigma_fs_identity."""
    h_re = np.array(h_re, dtype=np.float64)
    h_im = np.array(h_im, dtype=np.float64)
    adaptive = damping if isinstance(damping, AdaptiveDamping) else None
    static = 1.0 if damping is None or adaptive is not None else min(max(float(damping), 1e-6),
                                                                        1.0)
    alpha0 = adaptive.alpha() if adaptive is not None else static
    eff_max = max_iter
    if alpha0 < 1.0:
        eff_max = max_iter * max(int(math.ceil(2.0 / alpha0)), 2)
    sig_fn = sigma if sigma is not None else (lambda a, b: float("nan"))
    out = {"sigma_fs_identity": sig_fn(h_re, h_im), "donaldson_history": [], "sigma_history": []}
    min_res, min_h, min_it, bad = math.inf, (h_re.copy(), h_im.copy()), 0, 0
    final_res, iters = math.inf, 0
    for it in range(eff_max):
        alpha = adaptive.alpha() if adaptive is not None else static
        h_re_n, h_im_n, r = donaldson_iteration_route34(t_operator, h_re, h_im, weights, alpha,
                                                        tikhonov)
        if adaptive is not None:
            adaptive.update(r)
        if not math.isfinite(r):
            if math.isfinite(min_res):
                h_re, h_im = min_h[0].copy(), min_h[1].copy()
                out["donaldson_history"].append(min_res)
                out["sigma_history"].append(sig_fn(h_re, h_im))
                final_res, iters = min_res, min_it
                break
            raise FloatingPointError("Donaldson residual non-finite before any min snapshot")
        h_re, h_im = h_re_n, h_im_n
        out["donaldson_history"].append(r)
        out["sigma_history"].append(sig_fn(h_re, h_im))
        final_res, iters = r, it + 1
        if r < min_res:
            min_res, min_h, min_it, bad = r, (h_re.copy(), h_im.copy()), iters, 0
        elif (math.isfinite(min_res) and min_res < CATASTROPHE_FLOOR
              and r > CATASTROPHE_MULTIPLIER * min_res):
            bad += 1
            if bad >= CATASTROPHE_STREAK:
                h_re, h_im = min_h[0].copy(), min_h[1].copy()
                out["donaldson_history"].append(min_res)
                out["sigma_history"].append(sig_fn(h_re, h_im))
                final_res, iters = min_res, min_it
                break
        else:
            bad = 0
        if r < tol:
            break
    out.update(h_re=h_re, h_im=h_im, iterations_run=iters, final_donaldson_residual=final_res)
    return out
