"""Python mirror of the reference's QuinticSolver, over the cy3q C ABI (include/cy3q.h).

Names, argument meaning and error behaviour follow cy3/src/quintic.rs:
  QuinticSolver.new_with_sampler(k, n_points, seed, newton_tol, sampler)   quintic.rs:281-322
  QuinticSolver.new_with_psi_sz(k, n_points, seed, psi)                     quintic.rs:241-279
  sigma(), donaldson_step(), donaldson_solve(max_iter, tol),
  orthonormalise_basis_fs_gram()                                            quintic.rs:423-443, 3140-3165
Construction returns None where the reference returns None: no point was sampled. A
not-positive-definite FS Gram raises ValueError, as the reference's Err does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libcy3q.so")

SAMPLER_NEWTON = 0
SAMPLER_SZ = 1
SAMPLER_SZ_DWORK = 2
SAMPLER_NEWTON_DWORK = 3
PREC_FP64 = 0
PREC_SPLIT16 = 1
REFINE_STEP_FP64 = 0
REFINE_T_FP64 = 1

OK, ERR_ARG, ERR_HIP, ERR_NOT_PD, ERR_EMPTY, ERR_STATE = 0, -1, -2, -3, -4, -5


class Cy3qError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cy3q error {code}: {msg}")
        self.code = code


class Options(C.Structure):
    _fields_ = [("quotient", C.c_int), ("precision", C.c_int), ("device", C.c_int),
                ("dk_iters", C.c_int), ("newton_polish", C.c_int), ("newton_tol", C.c_double)]


def lib_path() -> str:
    return _LIB


_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB):
        raise ImportError(f"{_LIB} is not built; run papers_amd/build.py (or __graft_entry__.build())")
    L = C.CDLL(_LIB)
    i64, u64, u32, dbl, vp, ip = C.c_int64, C.c_uint64, C.c_uint32, C.c_double, C.c_void_p, C.c_int
    L.cy3q_version.restype = C.c_char_p
    L.cy3q_last_error.restype = C.c_char_p
    L.cy3q_stage_name.restype = C.c_char_p
    L.cy3q_stage_name.argtypes = [ip]
    L.cy3q_options_default.argtypes = [C.POINTER(Options)]
    L.cy3q_n_sections.argtypes = [u32, ip, C.POINTER(i64)]
    L.cy3q_monomials.argtypes = [u32, ip, vp, i64, C.POINTER(i64)]
    L.cy3q_solver_new.argtypes = [u32, i64, u64, dbl, ip, C.POINTER(Options), C.POINTER(vp)]
    L.cy3q_solver_new_from_points.argtypes = [u32, vp, i64, dbl, C.POINTER(Options), C.POINTER(vp)]
    L.cy3q_solver_free.argtypes = [vp]
    L.cy3q_solver_free.restype = None
    L.cy3q_solver_info.argtypes = [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(u32), C.POINTER(dbl)]
    L.cy3q_solver_set_precision.argtypes = [vp, ip]
    L.cy3q_solver_set_refine.argtypes = [vp, i64]
    L.cy3q_solver_set_refine_kind.argtypes = [vp, ip]
    L.cy3q_solver_donaldson_step_legacy.argtypes = [vp, ip, C.POINTER(dbl)]
    L.cy3q_solver_donaldson_solve_legacy.argtypes = [vp, ip, i64, dbl, C.POINTER(i64), vp]
    L.cy3q_solver_orthonormalise_fs_gram.argtypes = [vp]
    L.cy3q_solver_donaldson_step.argtypes = [vp, C.POINTER(dbl)]
    L.cy3q_solver_donaldson_solve.argtypes = [vp, i64, dbl, C.POINTER(i64), vp]
    L.cy3q_solver_sigma.argtypes = [vp, C.POINTER(dbl)]
    L.cy3q_solver_sigma_ext.argtypes = [vp, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(dbl), vp]
    L.cy3q_solver_sigma_at_sample.argtypes = [vp, i64, u64, ip, C.POINTER(dbl), C.POINTER(dbl),
                                              C.POINTER(dbl), C.POINTER(i64), vp, vp]
    L.cy3q_solver_sigma_sq_grad.argtypes = [vp, C.POINTER(dbl), vp, C.POINTER(dbl)]
    L.cy3q_solver_sigma_refine_analytic.argtypes = [vp, i64, dbl, vp, C.POINTER(i64)]
    L.cy3q_solver_set_adam.argtypes = [vp, dbl, dbl, dbl]
    L.cy3q_solver_sigma_refine_lbfgs.argtypes = [vp, i64, i64, vp, C.POINTER(i64)]
    L.cy3q_solver_sigma_refine_analytic_restarts.argtypes = [vp, i64, i64, dbl, dbl, u64,
                                                             C.POINTER(dbl), vp]
    L.cy3q_solver_sigma_refine_lbfgs_restarts.argtypes = [vp, i64, i64, i64, dbl, u64,
                                                          C.POINTER(dbl), vp]
    L.cy3q_solver_get_h_block.argtypes = [vp, vp]
    L.cy3q_solver_set_h_block.argtypes = [vp, vp]
    L.cy3q_solver_get_t.argtypes = [vp, vp]
    L.cy3q_solver_get_points.argtypes = [vp, vp]
    L.cy3q_solver_get_weights.argtypes = [vp, vp, vp]
    L.cy3q_solver_get_sections.argtypes = [vp, vp, vp]
    L.cy3q_solver_set_weights.argtypes = [vp, vp]
    L.cy3q_monge_ampere_residual.argtypes = [u32, vp, i64, vp, vp, ip, C.POINTER(dbl)]
    L.cy3q_solver_set_profiling.argtypes = [vp, ip]
    L.cy3q_solver_stage_times.argtypes = [vp, vp, vp, ip, C.POINTER(ip)]
    L.cy3q_solver_synchronize.argtypes = [vp]
    _lib = L
    return L


def library_version() -> str:
    """cy3q_version(): the library version, the compiler that built it and its GPU target."""
    return _load().cy3q_version().decode()


def _check(rc: int):
    if rc != OK:
        raise Cy3qError(rc, _load().cy3q_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def n_sections_quintic(k: int, quotient: bool = False) -> int:
    """n_sections_quintic (quintic.rs:3185-3188); with quotient, the basis without e0 >= 5."""
    n = C.c_int64(0)
    _check(_load().cy3q_n_sections(k, int(quotient), C.byref(n)))
    return n.value


def build_degree_k_quintic_monomials(k: int, quotient: bool = False) -> np.ndarray:
    """build_degree_k_quintic_monomials (quintic.rs:3192-3209) as a [N, 5] uint32 table."""
    n = n_sections_quintic(k, quotient)
    out = np.zeros((n, 5), np.uint32)
    got = C.c_int64(0)
    _check(_load().cy3q_monomials(k, int(quotient), _ptr(out), n, C.byref(got)))
    return out


def _options(quotient=False, precision=PREC_FP64, device=0, dk_iters=50, newton_polish=0,
             newton_tol=1e-12):
    o = Options()
    _load().cy3q_options_default(C.byref(o))
    o.quotient, o.precision, o.device = int(quotient), int(precision), int(device)
    o.dk_iters, o.newton_polish = int(dk_iters), int(newton_polish)
    o.newton_tol = float(newton_tol)
    return o


class QuinticSolver:
    """Workspace for the Fermat/Dwork quintic metric pipeline, resident on one GPU."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        n, nb, k, psi = C.c_int64(), C.c_int64(), C.c_uint32(), C.c_double()
        _check(_load().cy3q_solver_info(self._h, C.byref(n), C.byref(nb), C.byref(k), C.byref(psi)))
        self.n_points, self.n_basis, self.k_degree, self.psi = n.value, nb.value, k.value, psi.value
        self.donaldson_residuals: list[float] = []

    # ---- constructors -------------------------------------------------------------------
    @classmethod
    def new(cls, k_degree: int, n_points: int, seed: int, newton_tol: float = 1e-12, **opts):
        """QuinticSolver::new (quintic.rs:174-182): the Newton-projection sampler."""
        return cls._new(k_degree, n_points, seed, 0.0, SAMPLER_NEWTON, newton_tol=newton_tol,
                        **opts)

    @classmethod
    def new_with_psi(cls, k_degree: int, n_points: int, seed: int, newton_tol: float, psi: float,
                     **opts):
        """QuinticSolver::new_with_psi (quintic.rs:184-221): Newton projection onto the Dwork
        quintic at psi."""
        return cls._new(k_degree, n_points, seed, psi, SAMPLER_NEWTON_DWORK,
                        newton_tol=newton_tol, **opts)

    @classmethod
    def new_with_sampler(cls, k_degree: int, n_points: int, seed: int, newton_tol: float = 1e-12,
                         sampler: int = SAMPLER_SZ, **opts):
        """QuinticSolver::new_with_sampler (quintic.rs:281-322): SAMPLER_NEWTON or SAMPLER_SZ."""
        return cls._new(k_degree, n_points, seed, 0.0, sampler, newton_tol=newton_tol, **opts)

    @classmethod
    def new_with_psi_sz(cls, k_degree: int, n_points: int, seed: int, psi: float, **opts):
        """QuinticSolver::new_with_psi_sz (quintic.rs:241-279)."""
        return cls._new(k_degree, n_points, seed, psi, SAMPLER_SZ_DWORK, **opts)

    @classmethod
    def _new(cls, k, n, seed, psi, sampler, **opts):
        h = C.c_void_p()
        o = _options(**opts)
        rc = _load().cy3q_solver_new(k, n, seed, psi, sampler, C.byref(o), C.byref(h))
        if rc == ERR_EMPTY:
            return None
        _check(rc)
        return cls(h.value)

    @classmethod
    def from_points(cls, k_degree: int, points: np.ndarray, psi: float = 0.0, **opts):
        pts = np.ascontiguousarray(points, np.float64)
        h = C.c_void_p()
        o = _options(**opts)
        rc = _load().cy3q_solver_new_from_points(k_degree, _ptr(pts), pts.shape[0], psi, C.byref(o),
                                                 C.byref(h))
        if rc == ERR_EMPTY:
            return None
        _check(rc)
        return cls(h.value)

    def close(self):
        if self._h:
            _load().cy3q_solver_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- reference methods --------------------------------------------------------------
    def set_precision(self, precision: int):
        _check(_load().cy3q_solver_set_precision(self._h, precision))

    def set_refine(self, sweeps: int):
        """FP64 T-operator sweeps that end a split-precision donaldson_solve (default 1)."""
        _check(_load().cy3q_solver_set_refine(self._h, sweeps))

    def set_refine_kind(self, kind: int):
        """REFINE_T_FP64 (split D pass, FP64 T pass: the default) or REFINE_STEP_FP64 (a
        whole FP64 step)."""
        _check(_load().cy3q_solver_set_refine_kind(self._h, kind))

    def orthonormalise_basis_fs_gram(self):
        rc = _load().cy3q_solver_orthonormalise_fs_gram(self._h)
        if rc == ERR_NOT_PD:
            raise ValueError("FS Gram matrix not positive-definite")
        _check(rc)

    def donaldson_step(self) -> float:
        r = C.c_double()
        _check(_load().cy3q_solver_donaldson_step(self._h, C.byref(r)))
        return r.value

    def donaldson_solve(self, max_iter: int, tol: float) -> int:
        res = np.zeros(max(max_iter, 1))
        it = C.c_int64()
        _check(_load().cy3q_solver_donaldson_solve(self._h, max_iter, tol, C.byref(it), _ptr(res)))
        self.donaldson_residuals = list(res[: it.value])
        return it.value

    def donaldson_iteration_legacy(self, weighted: bool = True) -> float:
        """donaldson_iteration_quintic(_weighted) (quintic.rs:4140-4331) on this solver's h and
        sections, in place: h <- T(h) without inversion. Returns the block residual."""
        r = C.c_double()
        _check(_load().cy3q_solver_donaldson_step_legacy(self._h, int(weighted), C.byref(r)))
        return r.value

    def donaldson_solve_legacy(self, max_iter: int, tol: float, weighted: bool = True) -> list:
        """donaldson_solve_quintic(_weighted) (quintic.rs:4224-4357): from h = I; residuals."""
        res = np.zeros(max(max_iter, 1))
        it = C.c_int64()
        _check(_load().cy3q_solver_donaldson_solve_legacy(self._h, int(weighted), max_iter, tol,
                                                          C.byref(it), _ptr(res)))
        return list(res[: it.value])

    def sigma(self) -> float:
        s = C.c_double()
        _check(_load().cy3q_solver_sigma(self._h, C.byref(s)))
        return s.value

    def sigma_ext(self, want_eta: bool = False):
        s, k, l2 = C.c_double(), C.c_double(), C.c_double()
        eta = np.zeros(self.n_points) if want_eta else None
        _check(_load().cy3q_solver_sigma_ext(self._h, C.byref(s), C.byref(k), C.byref(l2), _ptr(eta)))
        return s.value, k.value, l2.value, eta

    def sigma_at_sample(self, n_points: int, seed: int, sampler: int = SAMPLER_SZ,
                        want_eta: bool = False):
        """sigma_at_metric_closure / sigma_and_etas_at_metric_closure (quintic.rs:1520-1700)
        with this solver's Bergman metric as the callback: sigma on a fresh sample of n_points
        points drawn with `sampler` and `seed`. Returns (sigma_L1, kappa, sigma_L2^2, n_actual,
        eta, points); eta and points are None unless want_eta."""
        s, k, l2 = C.c_double(), C.c_double(), C.c_double()
        na = C.c_int64()
        eta = np.zeros(n_points) if want_eta else None
        pts = np.zeros((n_points, 10)) if want_eta else None
        _check(_load().cy3q_solver_sigma_at_sample(self._h, n_points, seed, sampler, C.byref(s),
                                                   C.byref(k), C.byref(l2), C.byref(na),
                                                   _ptr(eta), _ptr(pts)))
        m = na.value
        return (s.value, k.value, l2.value, m, None if eta is None else eta[:m],
                None if pts is None else pts[:m])

    def sigma_squared_and_gradient(self):
        """sigma_squared_and_gradient (quintic.rs:2714-2860): (sigma_L2^2, grad [2 N^2], sigma_L1).
        grad[:N^2] = d/dh_re(a, b), diagonal halved; grad[N^2:] = d/dh_im(a, b) for a < b."""
        s2, s1 = C.c_double(), C.c_double()
        g = np.zeros(2 * self.n_basis * self.n_basis)
        _check(_load().cy3q_solver_sigma_sq_grad(self._h, C.byref(s2), _ptr(g), C.byref(s1)))
        return s2.value, g, s1.value

    def sigma_refine_analytic(self, n_iter: int, lr: float) -> list:
        """QuinticSolver::sigma_refine_analytic (quintic.rs:453-514); returns the sigma_L1
        history (one value per Adam step, before the step)."""
        hist = np.zeros(max(n_iter, 1))
        done = C.c_int64()
        _check(_load().cy3q_solver_sigma_refine_analytic(self._h, n_iter, lr, _ptr(hist),
                                                         C.byref(done)))
        return list(hist[: done.value])

    def set_adam(self, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
        _check(_load().cy3q_solver_set_adam(self._h, beta1, beta2, eps))

    def sigma_refine_lbfgs_gpu(self, n_iter: int, m_history: int) -> list:
        """QuinticSolver::sigma_refine_lbfgs_gpu (quintic.rs:678-1094): L-BFGS on sigma_L2^2 with
        Armijo backtracking and a steepest-descent retry; the lowest-sigma_L1 point is kept.
        Returns the sigma_L1 history (the start, then one value per accepted step)."""
        hist = np.zeros(n_iter + 1)
        done = C.c_int64()
        _check(_load().cy3q_solver_sigma_refine_lbfgs(self._h, n_iter, m_history, _ptr(hist),
                                                      C.byref(done)))
        return list(hist[: done.value])

    def sigma_refine_analytic_with_restarts(self, n_restarts: int, n_iter_per_restart: int,
                                            lr: float, perturb_scale: float, seed: int):
        """QuinticSolver::sigma_refine_analytic_with_restarts (quintic.rs:545-582): Adam, then
        n_restarts Adam runs from LCG-perturbed copies of the best h. Returns (best sigma_L1,
        history of n_restarts + 1 values); the best h stays in the solver."""
        hist = np.zeros(n_restarts + 1)
        best = C.c_double()
        _check(_load().cy3q_solver_sigma_refine_analytic_restarts(
            self._h, n_restarts, n_iter_per_restart, lr, perturb_scale, seed, C.byref(best),
            _ptr(hist)))
        return best.value, list(hist)

    # The reference's GPU form (quintic.rs:1145-1182) runs the same loop.
    sigma_refine_analytic_with_restarts_gpu = sigma_refine_analytic_with_restarts

    def sigma_refine_lbfgs_with_restarts_gpu(self, n_restarts: int, n_iter_per_restart: int,
                                             m_history: int, perturb_scale: float, seed: int):
        """QuinticSolver::sigma_refine_lbfgs_with_restarts_gpu (quintic.rs:1100-1139)."""
        hist = np.zeros(n_restarts + 1)
        best = C.c_double()
        _check(_load().cy3q_solver_sigma_refine_lbfgs_restarts(
            self._h, n_restarts, n_iter_per_restart, m_history, perturb_scale, seed,
            C.byref(best), _ptr(hist)))
        return best.value, list(hist)

    # ---- workspace access ---------------------------------------------------------------
    @property
    def h_block(self) -> np.ndarray:
        tn = 2 * self.n_basis
        out = np.zeros((tn, tn))
        _check(_load().cy3q_solver_get_h_block(self._h, _ptr(out)))
        return out

    @h_block.setter
    def h_block(self, h: np.ndarray):
        h = np.ascontiguousarray(h, np.float64)
        _check(_load().cy3q_solver_set_h_block(self._h, _ptr(h)))

    def last_t(self) -> np.ndarray:
        out = np.zeros((self.n_basis, self.n_basis, 2))
        _check(_load().cy3q_solver_get_t(self._h, _ptr(out)))
        return out[..., 0] + 1j * out[..., 1]

    @property
    def points(self) -> np.ndarray:
        out = np.zeros((self.n_points, 10))
        _check(_load().cy3q_solver_get_points(self._h, _ptr(out)))
        return out

    @property
    def weights(self) -> np.ndarray:
        out = np.zeros(self.n_points)
        _check(_load().cy3q_solver_get_weights(self._h, _ptr(out), None))
        return out

    @property
    def log_omega_sq(self) -> np.ndarray:
        out = np.zeros(self.n_points)
        _check(_load().cy3q_solver_get_weights(self._h, None, _ptr(out)))
        return out

    def sections(self, derivs: bool = False):
        sv = np.zeros((self.n_points, 2 * self.n_basis))
        sd = np.zeros((self.n_points, 5, 2 * self.n_basis)) if derivs else None
        _check(_load().cy3q_solver_get_sections(self._h, _ptr(sv), _ptr(sd)))
        return sv, sd

    # ---- profiling ------------------------------------------------------------------------
    def set_profiling(self, on: bool):
        _check(_load().cy3q_solver_set_profiling(self._h, int(on)))

    def stage_times(self) -> dict:
        ms = np.zeros(32)
        cnt = np.zeros(32, np.int64)
        ns = C.c_int()
        _check(_load().cy3q_solver_stage_times(self._h, _ptr(ms), _ptr(cnt), 32, C.byref(ns)))
        L = _load()
        return {L.cy3q_stage_name(i).decode(): (float(ms[i]), int(cnt[i])) for i in range(ns.value)}

    def set_weights(self, weights: np.ndarray):
        """Replace QuinticSolver::weights (later steps, sigma and gradients use them)."""
        w = np.ascontiguousarray(weights, np.float64)
        if w.shape != (self.n_points,):
            raise ValueError(f"weights must have shape ({self.n_points},)")
        _check(_load().cy3q_solver_set_weights(self._h, _ptr(w)))

    def synchronize(self):
        _check(_load().cy3q_solver_synchronize(self._h))


def monge_ampere_residual_quintic(points: np.ndarray, h_block: np.ndarray, k_degree: int,
                                  weights: np.ndarray | None = None, device: int = 0) -> float:
    """monge_ampere_residual_quintic(_weighted) (quintic.rs:4961-5049): sigma_L1 of the Fermat
    quintic metric h_block over the raw degree-k monomials at the given points, with the CY
    weights 1/|grad f|^2 or the given weights."""
    pts = np.ascontiguousarray(points, np.float64)
    h = np.ascontiguousarray(h_block, np.float64)
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    out = C.c_double()
    _check(_load().cy3q_monge_ampere_residual(k_degree, _ptr(pts), pts.shape[0], _ptr(h), _ptr(w),
                                              device, C.byref(out)))
    return out.value
