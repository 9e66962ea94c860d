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

* :class:`TikhonovShift`, :class:`TikhonovGating` and :class:`GatingState` are the
  residual-dependent Tikhonov schedule and its stall-band gate (``schoen_metric.rs:200-500``),
  with the auto-default helpers of both callers (``ty_metric.rs:262-310``).

The T-operator and sigma are callables, normally ``Cy3DonaldsonGpu.t_operator_raw`` and
``Cy3SigmaGpu.compute_sigma``. The Tikhonov shift is a constant or a :class:`TikhonovShift`.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.linalg as sla

import ctypes as C

from .quintic import ERR_NOT_PD, Cy3qError, _check, _load

__all__ = ["AdaptiveDamping", "GatingState", "TikhonovGating", "TikhonovShift",
           "auto_schoen_gated_tikhonov", "auto_schoen_tikhonov", "auto_ty_gated_tikhonov",
           "auto_ty_tikhonov", "donaldson_iteration_route34", "resolve_tikhonov", "solve_route34"]

CATASTROPHE_FLOOR = 1.0e-2
CATASTROPHE_MULTIPLIER = 100.0
CATASTROPHE_STREAK = 5


# Largest lambda as a fraction of the current residual (schoen_metric.rs:420-441).
TIKHONOV_STEP_FRACTION = 1.0e-2


class TikhonovGating:
    """When the schedule engages (``TikhonovGating``, schoen_metric.rs:245-300): on every
    iteration (``always_on``), or only once the residual has sat in [residual_lo, residual_hi]
    with per-iteration ratios in [ratio_lo, ratio_hi] for min_stuck_iters consecutive
    iterations (``stall_band``)."""

    def __init__(self, stall_band: bool, residual_lo: float = 0.0, residual_hi: float = 0.0,
                 min_stuck_iters: int = 0, ratio_lo: float = 0.95, ratio_hi: float = 1.05):
        self.stall_band = bool(stall_band)
        self.residual_lo, self.residual_hi = float(residual_lo), float(residual_hi)
        self.min_stuck_iters = int(min_stuck_iters)
        self.ratio_lo, self.ratio_hi = float(ratio_lo), float(ratio_hi)

    @classmethod
    def always_on(cls) -> "TikhonovGating":
        return cls(False)

    @classmethod
    def schoen_stall_band_default(cls) -> "TikhonovGating":
        """Band [1e-7, 1e-3], ratios [0.95, 1.05], streak 3 (schoen_metric.rs:288-299)."""
        return cls(True, 1.0e-7, 1.0e-3, 3, 0.95, 1.05)


class GatingState:
    """The streak counter of the stall-band gate (``GatingState``, schoen_metric.rs:300-375).
    ``update(r)`` runs after each iteration; ``is_open(r_prev, gating)`` before each inversion."""

    def __init__(self):
        self.last_residual = None
        self.in_band_streak = 0

    def update(self, residual: float, ratio_lo: float, ratio_hi: float):
        if not math.isfinite(residual) or residual <= 0.0:
            self.in_band_streak = 0
            self.last_residual = None
            return
        prev = self.last_residual
        if prev is not None and prev > 0.0 and math.isfinite(prev):
            ratio = residual / prev
            self.in_band_streak = self.in_band_streak + 1 if ratio_lo <= ratio <= ratio_hi else 0
        else:
            self.in_band_streak = 0
        self.last_residual = residual

    def is_open(self, r_curr: float, gating: TikhonovGating) -> bool:
        if not gating.stall_band:
            return True
        if not math.isfinite(r_curr) or r_curr <= 0.0:
            return False
        if r_curr < gating.residual_lo or r_curr > gating.residual_hi:
            return False
        return self.in_band_streak >= gating.min_stuck_iters


class TikhonovShift:
    """T(G) -> T(G) + lambda I before the inversion, with lambda from a residual-dependent
    schedule (``TikhonovShift``, schoen_metric.rs:200-245, 375-445)."""

    def __init__(self, lambda_max: float, lambda_min: float, schedule_exponent: float,
                 gating: TikhonovGating | None = None):
        self.lambda_max, self.lambda_min = float(lambda_max), float(lambda_min)
        self.schedule_exponent = float(schedule_exponent)
        self.gating = gating if gating is not None else TikhonovGating.always_on()

    @classmethod
    def k4_default(cls) -> "TikhonovShift":
        return cls(1.0e-3, 1.0e-9, 1.0, TikhonovGating.always_on())

    @classmethod
    def k4_gated_default(cls) -> "TikhonovShift":
        return cls(1.0e-3, 1.0e-9, 1.0, TikhonovGating.schoen_stall_band_default())

    def lambda_at(self, residual_curr: float, residual_init: float) -> float:
        """max(lambda_max (r / r_init)^p, lambda_min), capped at 1% of r and kept within
        [lambda_min, lambda_max]; lambda_max while either residual is missing (NaN, <= 0)."""
        lo = max(self.lambda_min, 0.0)
        hi = max(self.lambda_max, lo)
        if (not math.isfinite(residual_init) or residual_init <= 0.0
                or not math.isfinite(residual_curr) or residual_curr <= 0.0):
            return hi
        ratio = min(max(residual_curr / residual_init, 0.0), 1.0)
        p = max(self.schedule_exponent, 0.0)
        geom = hi * math.pow(ratio, p)
        lam = min(geom, residual_curr * TIKHONOV_STEP_FRACTION)
        return min(max(lam, lo), hi)


def resolve_tikhonov(override):
    """The solvers never engage a schedule on their own: None stays None, a given schedule is
    used as is (``resolve_ty_tikhonov``, ``resolve_schoen_tikhonov``)."""
    return override


def auto_ty_tikhonov(k_degree: int):
    """k4_default() for Tian-Yau k >= 4, else None (ty_metric.rs:284-293)."""
    return TikhonovShift.k4_default() if k_degree >= 4 else None


def auto_ty_gated_tikhonov(k_degree: int):
    return TikhonovShift.k4_gated_default() if k_degree >= 4 else None


def auto_schoen_tikhonov(d_x: int, d_y: int, d_t: int):
    """k4_default() for Schoen bidegrees with d_x + d_y + d_t >= 10, else None
    (schoen_metric.rs:475-481)."""
    return TikhonovShift.k4_default() if d_x + d_y + d_t >= 10 else None


def auto_schoen_gated_tikhonov(d_x: int, d_y: int, d_t: int):
    return TikhonovShift.k4_gated_default() if d_x + d_y + d_t >= 10 else None

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


def gpu_inverse(device: int = 0):
    """An `inverse` for the loops below: (T + shift I)^-1 of the Hermitian T = tr + i ti on the
    GPU (cy3q_hermitian_inverse), as (Re, Im), or None when T + shift I is not numerically
    positive definite (the loop then inverts the real block by LU, as the reference does)."""
    L = _load()
    if not getattr(L, "_cy3q_inv_bound", False):
        L.cy3q_hermitian_inverse.argtypes = [C.c_void_p, C.c_int64, C.c_double, C.c_int,
                                             C.c_void_p]
        L._cy3q_inv_bound = True

    def inverse(tr_, ti_, shift):
        n = tr_.shape[0]
        t = np.empty((n, n, 2))
        t[..., 0], t[..., 1] = tr_, ti_
        out = np.empty((n, n, 2))
        rc = L.cy3q_hermitian_inverse(t.ctypes.data, n, float(shift), device, out.ctypes.data)
        if rc == ERR_NOT_PD:
            return None
        _check(rc)
        return out[..., 0], out[..., 1]

    return inverse


def donaldson_iteration_route34(t_operator, h_re, h_im, weights, damping: float = 1.0,
                                tikhonov: float = 0.0, inverse=None):
    """One iteration; returns (h_re, h_im, residual). On a non-finite weight sum, a singular or
    non-finite inverse, h comes back unchanged with residual NaN (the reference returns NaN
    before its damped update). inverse: None (LU of the 2N x 2N real block, as the reference)
    or a callable(tr, ti, shift) -> (Re, Im) of (T + shift I)^-1, or None where it cannot
    invert (then the LU runs), e.g. gpu_inverse(device)."""
    h_re = np.asarray(h_re, dtype=np.float64)
    h_im = np.asarray(h_im, dtype=np.float64)
    n = h_re.shape[0]
    try:
        pr, pi = t_operator(h_re, h_im)
    except Cy3qError as e:
        # The GPU T-operator forms K_p = |W s_p|^2 with W^H W = conj(h), so it rejects an h that
        # is not positive definite, where the reference clamps K at 1e-30 and goes on
        # (cy3_donaldson_gpu.rs:84). A NaN residual hands the iterate to the caller's
        # minimum-residual snapshot and catastrophe guard.
        if e.code != ERR_NOT_PD:
            raise
        return h_re, h_im, float("nan")
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
    shift = tikhonov if math.isfinite(tikhonov) and tikhonov > 0.0 else 0.0
    # A positive definite T + shift I: its inverse's real block has the unpacked parts
    # nr = Re, ni = Im exactly.
    got = inverse(tr_, ti_, shift) if inverse is not None else None
    if got is not None:
        nr, ni = got
        if not (np.all(np.isfinite(nr)) and np.all(np.isfinite(ni))):
            return h_re, h_im, float("nan")
    else:
        block = np.zeros((2 * n, 2 * n))
        block[0::2, 0::2], block[0::2, 1::2] = tr_, -ti_
        block[1::2, 0::2], block[1::2, 1::2] = ti_, tr_
        if shift > 0.0:
            block[np.diag_indices(2 * n)] += shift
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
                  damping=None, tikhonov=0.0, inverse=None) -> dict:
    """The Donaldson loop of solve_ty_metric. damping: None or a float (static alpha, clamped to
    [1e-6, 1]) or an AdaptiveDamping. tikhonov: a constant shift (0 disables it) or a
    TikhonovShift; a schedule's lambda for an iteration is lambda_at(previous residual, first
    finite residual) while its gate is open, else 0 (ty_metric.rs:620-668). sigma:
    callable(h_re, h_im) or None. Returns a dict with h_re, h_im, iterations_run,
    final_donaldson_residual, donaldson_history, sigma_history, sigma_fs_identity and
    tikhonov_history (the lambda of each iteration). inverse: as in
    donaldson_iteration_route34."""
    h_re = np.array(h_re, dtype=np.float64)
    h_im = np.array(h_im, dtype=np.float64)
    adaptive = damping if isinstance(damping, AdaptiveDamping) else None
    static = 1.0 if damping is None or adaptive is not None else min(max(float(damping), 1e-6),
                                                                        1.0)
    alpha0 = adaptive.alpha() if adaptive is not None else static
    eff_max = max_iter
    if alpha0 < 1.0:
        eff_max = max_iter * max(int(math.ceil(2.0 / alpha0)), 2)
    def sig_fn(a, b):
        # The GPU sigma contract contracts the sections with W, W^H W = h, so it rejects an h
        # that is not positive definite; the reference clamps K at 1e-30 and returns a (large)
        # number for such an iterate (cy3_sigma_gpu.rs:80-107). sigma never steers the loop,
        # so the iterate is recorded with sigma NaN and the loop goes on.
        if sigma is None:
            return float("nan")
        try:
            return sigma(a, b)
        except Cy3qError as e:
            if e.code != ERR_NOT_PD:
                raise
            return float("nan")

    out = {"sigma_fs_identity": sig_fn(h_re, h_im), "donaldson_history": [], "sigma_history": []}
    min_res, min_h, min_it, bad = math.inf, (h_re.copy(), h_im.copy()), 0, 0
    final_res, iters = math.inf, 0
    shift = tikhonov if isinstance(tikhonov, TikhonovShift) else None
    gate = GatingState()
    residual_init, prev_residual = math.nan, math.nan
    out["tikhonov_history"] = []
    for it in range(eff_max):
        alpha = adaptive.alpha() if adaptive is not None else static
        if shift is None:
            lam = float(tikhonov) if tikhonov is not None else 0.0
        elif gate.is_open(prev_residual, shift.gating):
            lam = shift.lambda_at(prev_residual, residual_init)
        else:
            lam = 0.0
        out["tikhonov_history"].append(lam)
        h_re_n, h_im_n, r = donaldson_iteration_route34(t_operator, h_re, h_im, weights, alpha,
                                                        lam, inverse=inverse)
        if adaptive is not None:
            adaptive.update(r)
        if shift is not None and shift.gating.stall_band:
            gate.update(r, shift.gating.ratio_lo, shift.gating.ratio_hi)
        else:
            gate.update(r, 0.95, 1.05)
        if not math.isfinite(residual_init) and math.isfinite(r) and r > 0.0:
            residual_init = r
        if math.isfinite(r):
            prev_residual = r
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
