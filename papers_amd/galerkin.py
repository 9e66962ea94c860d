"""Galerkin assembly on the GPU: metric Laplacian spectra and Yukawa triple overlaps.

Python mirror of the reference functions, over the cy3q C ABI (include/cy3q.h):
  compute_metric_laplacian_spectrum(metric, config)     route34/metric_laplacian.rs:296-418
  build_test_basis(max_total_degree)                     route34/metric_laplacian.rs:87-115
  hermitian_jacobi_n(a, n, max_sweeps, tol)              route34/metric_laplacian.rs:616-693
  sz_weights(metric)                                     route34/yukawa_overlap_real.rs:222-268
  triple_overlap(modes, omega, weights, h_v_diag)        route34/yukawa_overlap_real.rs:271-315
  compute_yukawa_couplings(metric, h_v_diag, modes, cfg) route34/yukawa_overlap_real.rs:372-525
A metric background is given as (points [n, 8] complex, weights [n], omega [n] complex), the
three accessors of MetricBackground (route34/hym_hermitian.rs:257-270).
Every call runs on the GPU. The library is required and there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from .quintic import _check, _load, _ptr

_bound = False


def _lib():
    global _bound
    L = _load()
    if not _bound:
        i64, u64, u32, dbl, vp, ip = (C.c_int64, C.c_uint64, C.c_uint32, C.c_double, C.c_void_p,
                                      C.c_int)
        L.cy3q_laplacian_config_default.argtypes = [C.POINTER(LaplacianConfig)]
        L.cy3q_laplacian_basis.argtypes = [u32, vp, i64, C.POINTER(i64)]
        L.cy3q_metric_laplacian_spectrum.argtypes = [vp, vp, vp, i64, C.POINTER(LaplacianConfig),
                                                     ip, vp, vp, vp, vp, C.POINTER(i64)]
        L.cy3q_solver_laplacian_spectrum.argtypes = [vp, u32, i64, dbl, i64, vp, vp, vp,
                                                     C.POINTER(i64), C.POINTER(i64)]
        L.cy3q_hermitian_eigen.argtypes = [vp, i64, i64, dbl, ip, vp, vp, C.POINTER(i64)]
        L.cy3q_solver_set_galerkin_precision.argtypes = [vp, ip]
        L.cy3q_yukawa_sz_weights.argtypes = [vp, vp, i64, ip, vp, C.POINTER(dbl)]
        L.cy3q_yukawa_triple_overlap.argtypes = [vp, vp, vp, vp, i64, i64, ip, vp]
        L.cy3q_yukawa_couplings.argtypes = [vp, vp, vp, vp, i64, i64, i64, u64, ip, vp, vp,
                                            C.POINTER(dbl), C.POINTER(dbl)]
        L.cy3q_solver_yukawa_random_modes.argtypes = [vp, u32, i64, u64, vp, vp, C.POINTER(i64)]
        L.cy3q_galerkin_stage_times.argtypes = [vp, vp, ip, C.POINTER(ip)]
        L.cy3q_galerkin_stage_name.argtypes = [ip]
        L.cy3q_galerkin_stage_name.restype = C.c_char_p
        _bound = True
    return L


class LaplacianConfig(C.Structure):
    """MetricLaplacianConfig (metric_laplacian.rs:224-271); defaults from :277-290."""
    _fields_ = [("max_total_degree", C.c_uint32), ("n_low_eigenvalues", C.c_int64),
                ("jacobi_max_sweeps", C.c_int64), ("jacobi_tol", C.c_double),
                ("mass_regularisation", C.c_double)]

    @classmethod
    def default(cls, **kw):
        c = cls()
        _lib().cy3q_laplacian_config_default(C.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        return c


@dataclass
class MetricLaplacianSpectrum:
    """MetricLaplacianSpectrum (metric_laplacian.rs:177-221), without the optional vectors."""
    eigenvalues: np.ndarray
    eigenvalues_full: np.ndarray
    basis_dim: int
    n_points: int
    mass: np.ndarray | None = None
    stiffness: np.ndarray | None = None


def _c128(a, shape=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.complex128)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def build_test_basis(max_total_degree: int) -> np.ndarray:
    n = C.c_int64()
    _check(_lib().cy3q_laplacian_basis(max_total_degree, None, 0, C.byref(n)))
    out = np.zeros((n.value, 8), np.uint32)
    _check(_lib().cy3q_laplacian_basis(max_total_degree, _ptr(out), n.value, C.byref(n)))
    return out


def compute_metric_laplacian_spectrum(points, weights, omega, config: LaplacianConfig | None = None,
                                      device: int = 0, want_matrices: bool = False
                                      ) -> MetricLaplacianSpectrum:
    pts = _c128(points)
    n = pts.shape[0]
    if pts.shape != (n, 8):
        raise ValueError("points must be [n, 8] complex")
    w = np.ascontiguousarray(weights, np.float64)
    om = _c128(omega, (n,))
    cfg = config if config is not None else LaplacianConfig.default()
    nb = build_test_basis(cfg.max_total_degree).shape[0]
    n_low = min(max(int(cfg.n_low_eigenvalues), 0), nb)
    low = np.zeros(max(n_low, 1))
    full = np.zeros(max(nb, 1))
    M = np.zeros((nb, nb), np.complex128) if want_matrices else None
    K = np.zeros((nb, nb), np.complex128) if want_matrices else None
    bd = C.c_int64()
    _check(_lib().cy3q_metric_laplacian_spectrum(_ptr(pts), _ptr(w), _ptr(om), n, C.byref(cfg),
                                                 device, _ptr(low), _ptr(full), _ptr(M), _ptr(K),
                                                 C.byref(bd)))
    if n == 0:
        return MetricLaplacianSpectrum(np.zeros(0), np.zeros(0), nb, 0, M, K)
    return MetricLaplacianSpectrum(low[:n_low].copy(), full[:nb].copy(), nb, n, M, K)


def hermitian_eigen(a, max_sweeps: int = 128, tol: float = 1e-14, device: int = 0,
                    vectors: bool = False):
    """(eigenvalues ascending, eigenvector columns or None, sweeps) of a Hermitian matrix."""
    a = _c128(a)
    n = a.shape[0]
    ev = np.zeros(n)
    vec = np.zeros((n, n), np.complex128) if vectors else None
    sw = C.c_int64()
    _check(_lib().cy3q_hermitian_eigen(_ptr(a), n, max_sweeps, tol, device, _ptr(ev), _ptr(vec),
                                       C.byref(sw)))
    return ev, vec, sw.value


def solver_laplacian_spectrum(solver, section_degree: int = 3, n_low: int = 50,
                              deflation_tol: float = 1e-10, max_sweeps: int = 64,
                              want_matrices: bool = False, precision: int | None = None):
    """Intrinsic-metric Laplacian of the solver's current h on phi_ab = s_a conj(s_b)/|z|^(2d):
    (lowest eigenvalues, rank kept, M, K). precision: PREC_FP64 or PREC_SPLIT16 for the rank-k
    assembly (None keeps the solver's setting, FP64 by default); the eigensolve is FP64."""
    if precision is not None:
        _check(_lib().cy3q_solver_set_galerkin_precision(solver._h, int(precision)))
    ns = (section_degree + 1) * (section_degree + 2) * (section_degree + 3) * (section_degree + 4) // 24
    B = ns * ns
    low = np.zeros(max(n_low, 1))
    M = np.zeros((B, B), np.complex128) if want_matrices else None
    K = np.zeros((B, B), np.complex128) if want_matrices else None
    bd, rk = C.c_int64(), C.c_int64()
    _check(_lib().cy3q_solver_laplacian_spectrum(solver._h, section_degree, n_low, deflation_tol,
                                                 max_sweeps, _ptr(low), _ptr(M), _ptr(K),
                                                 C.byref(bd), C.byref(rk)))
    return low[:min(n_low, rk.value)].copy(), rk.value, M, K


def sz_weights(weights, omega, device: int = 0):
    w = np.ascontiguousarray(weights, np.float64)
    om = _c128(omega, (w.shape[0],))
    sz = np.zeros_like(w)
    sc = C.c_double()
    _check(_lib().cy3q_yukawa_sz_weights(_ptr(w), _ptr(om), w.shape[0], device, _ptr(sz),
                                         C.byref(sc)))
    return sz, sc.value


def triple_overlap(modes, omega, weights, h_v_diag=None, device: int = 0) -> np.ndarray:
    md = _c128(modes)
    nm, n = md.shape
    om = _c128(omega, (n,))
    w = np.ascontiguousarray(weights, np.float64)
    hd = None if h_v_diag is None else np.ascontiguousarray(h_v_diag, np.float64)
    y = np.zeros((nm, nm, nm), np.complex128)
    _check(_lib().cy3q_yukawa_triple_overlap(_ptr(md), _ptr(om), _ptr(w), _ptr(hd), nm, n, device,
                                             _ptr(y)))
    return y


@dataclass
class YukawaResult:
    """YukawaResult (yukawa_overlap_real.rs:179-193) without the run metadata."""
    couplings: np.ndarray
    couplings_uncertainty: np.ndarray
    n_quadrature_points: int
    quadrature_uniformity_score: float
    convergence_ratio: float
    extra: dict = field(default_factory=dict)


def compute_yukawa_couplings(modes, omega, weights, h_v_diag=None, n_bootstrap: int = 64,
                             seed: int = 0xBADDCAFEF00D, device: int = 0) -> YukawaResult:
    md = _c128(modes)
    nm, n = md.shape
    om = _c128(omega, (n,))
    w = np.ascontiguousarray(weights, np.float64)
    hd = None if h_v_diag is None else np.ascontiguousarray(h_v_diag, np.float64)
    y = np.zeros((nm, nm, nm), np.complex128)
    unc = np.zeros((nm, nm, nm))
    sc, cv = C.c_double(), C.c_double()
    _check(_lib().cy3q_yukawa_couplings(_ptr(md), _ptr(om), _ptr(w), _ptr(hd), nm, n, n_bootstrap,
                                        seed, device, _ptr(y), _ptr(unc), C.byref(sc),
                                        C.byref(cv)))
    return YukawaResult(y, unc, n, sc.value, cv.value)


def solver_yukawa_random_modes(solver, degree: int = 5, n_modes: int = 20, coeff_seed: int = 42):
    """Y_abc over a <= b <= c for random modes psi = S_degree C with measure w |Omega|^2:
    (triples [T] complex, coefficients [nmono, n_modes] complex)."""
    nmono = (degree + 1) * (degree + 2) * (degree + 3) * (degree + 4) // 24
    T = n_modes * (n_modes + 1) * (n_modes + 2) // 6
    y = np.zeros(T, np.complex128)
    cf = np.zeros((nmono, n_modes), np.complex128)
    nt = C.c_int64()
    _check(_lib().cy3q_solver_yukawa_random_modes(solver._h, degree, n_modes, coeff_seed, _ptr(y),
                                                  _ptr(cf), C.byref(nt)))
    return y[:nt.value], cf


def last_stage_times() -> dict:
    """{stage: (device ms, launches)} of this thread's last Galerkin call."""
    ms = np.zeros(16)
    cnt = np.zeros(16, np.int64)
    ns = C.c_int()
    L = _lib()
    _check(L.cy3q_galerkin_stage_times(_ptr(ms), _ptr(cnt), 16, C.byref(ns)))
    return {L.cy3q_galerkin_stage_name(i).decode(): (float(ms[i]), int(cnt[i]))
            for i in range(ns.value) if cnt[i]}


def triple_index(n_modes: int) -> np.ndarray:
    """The (a, b, c), a <= b <= c, order of the triples returned above."""
    return np.array([(a, b, c) for a in range(n_modes) for b in range(a, n_modes)
                     for c in range(b, n_modes)], np.int64)
