"""The route34 GPU contracts of the 8-coordinate (Tian-Yau, Schoen) metric solvers: the
T-operator ``Cy3DonaldsonGpu`` and the sigma evaluator ``Cy3SigmaGpu``.

T-operator: ``Cy3DonaldsonGpu`` of the reference
(``route34/cy3_donaldson_gpu.rs:164-333``), the T-operator the Tian-Yau and Schoen metric
solvers call once per Donaldson iteration, over ``cy3q_top_*`` of ``libcy3q.so``.

The static inputs are the caller's section values (``n_points x n_basis`` complex, stored as
``(re, im)`` pairs) and weights; ``t_operator_raw(h_re, h_im)`` returns the unnormalised
``h_pair[a][b] = sum_p (w_p / max(K_p, 1e-30)) conj(s_a) s_b`` with
``K_p = sum_ab h_re Re(conj(s_a) s_b) + h_im Im(conj(s_a) s_b)`` (the route34 h convention,
``h = conj(G)`` of :class:`papers_amd.QuinticSolver`). ``h`` must be Hermitian positive
definite.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .quintic import ERR_NOT_PD, Cy3qError, _check, _load, _ptr

__all__ = ["Cy3DonaldsonGpu", "Cy3SigmaGpu"]

_bound = False


def _lib():
    global _bound
    L = _load()
    if not _bound:
        vp, i64, ip = C.c_void_p, C.c_int64, C.c_int
        L.cy3q_top_new.argtypes = [i64, i64, ip, C.POINTER(vp)]
        L.cy3q_top_free.argtypes = [vp]
        L.cy3q_top_free.restype = None
        L.cy3q_top_upload_static.argtypes = [vp, vp, vp]
        L.cy3q_top_t_operator_raw.argtypes = [vp, vp, vp, vp, vp]
        L.cy3q_sig8_new.argtypes = [i64, i64, ip, C.POINTER(vp)]
        L.cy3q_sig8_free.argtypes = [vp]
        L.cy3q_sig8_free.restype = None
        L.cy3q_sig8_upload_static.argtypes = [vp, vp, vp, vp, vp, vp]
        L.cy3q_sig8_compute_sigma.argtypes = [vp, vp, vp, C.POINTER(C.c_double)]
        _bound = True
    return L


class Cy3DonaldsonGpu:
    """Cy3DonaldsonGpu::{new, upload_static, t_operator_raw} (cy3_donaldson_gpu.rs:182-320)."""

    def __init__(self, n_points: int, n_basis: int, device: int = 0):
        h = C.c_void_p()
        _check(_lib().cy3q_top_new(n_points, n_basis, device, C.byref(h)))
        self._h = h
        self.n_points, self.n_basis = int(n_points), int(n_basis)

    def close(self):
        if self._h:
            _lib().cy3q_top_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_static(self, section_values: np.ndarray, weights: np.ndarray):
        sv = np.ascontiguousarray(section_values, np.float64)
        w = np.ascontiguousarray(weights, np.float64)
        if sv.size != self.n_points * 2 * self.n_basis:
            raise ValueError(f"section_values has {sv.size} values; expected "
                             f"{self.n_points * 2 * self.n_basis}")
        if w.size != self.n_points:
            raise ValueError(f"weights has {w.size} values; expected {self.n_points}")
        _check(_lib().cy3q_top_upload_static(self._h, _ptr(sv), _ptr(w)))

    def t_operator_raw(self, h_re: np.ndarray, h_im: np.ndarray):
        nb = self.n_basis
        hr = np.ascontiguousarray(h_re, np.float64).reshape(nb * nb)
        hi = np.ascontiguousarray(h_im, np.float64).reshape(nb * nb)
        out_re, out_im = np.zeros(nb * nb), np.zeros(nb * nb)
        rc = _lib().cy3q_top_t_operator_raw(self._h, _ptr(hr), _ptr(hi), _ptr(out_re),
                                            _ptr(out_im))
        if rc == ERR_NOT_PD:
            raise Cy3qError(rc, "h is not Hermitian positive definite")
        _check(rc)
        return out_re, out_im


class Cy3SigmaGpu:
    """Cy3SigmaGpu::{new, upload_static, compute_sigma} (route34/cy3_sigma_gpu.rs:314-547):
    sigma = sum w (eta / kappa - 1)^2 / sum w on 8 ambient coordinates with caller-supplied
    3 x 8 tangent frames; h is used as given (K = s^H h s) and must be Hermitian positive
    definite."""

    def __init__(self, n_points: int, n_basis: int, device: int = 0):
        h = C.c_void_p()
        _check(_lib().cy3q_sig8_new(n_points, n_basis, device, C.byref(h)))
        self._h = h
        self.n_points, self.n_basis = int(n_points), int(n_basis)

    def close(self):
        if self._h:
            _lib().cy3q_sig8_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_static(self, section_values, section_derivs, frames, omega_sq, weights):
        n, two_n = self.n_points, 2 * self.n_basis
        arrs = [np.ascontiguousarray(a, np.float64) for a in
                (section_values, section_derivs, frames, omega_sq, weights)]
        for a, size, name in zip(arrs, (n * two_n, n * 8 * two_n, n * 48, n, n),
                                 ("section_values", "section_derivs", "frames", "omega_sq",
                                  "weights")):
            if a.size != size:
                raise ValueError(f"{name} has {a.size} values; expected {size}")
        _check(_lib().cy3q_sig8_upload_static(self._h, *[_ptr(a) for a in arrs]))

    def compute_sigma(self, h_re, h_im) -> float:
        nb = self.n_basis
        hr = np.ascontiguousarray(h_re, np.float64).reshape(nb * nb)
        hi = np.ascontiguousarray(h_im, np.float64).reshape(nb * nb)
        out = C.c_double()
        rc = _lib().cy3q_sig8_compute_sigma(self._h, _ptr(hr), _ptr(hi), C.byref(out))
        if rc == ERR_NOT_PD:
            raise Cy3qError(rc, "h is not Hermitian positive definite")
        _check(rc)
        return out.value

