"""ctypes binding of the CPU FP64 oracle (oracle/cy3q_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` as the checker. The product (``papers_amd``) never
imports it.

Every wrapper mirrors one reference function (cited in the C source) and returns numpy
arrays in the reference layouts: points [n, 10] (re/im interleaved), section values
[n, 2N], derivatives [n, 5, 2N], h as the 2N x 2N real block.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


def build() -> str:
    """Compiles liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    i64, u64, u32, dbl, vp = C.c_int64, C.c_uint64, C.c_uint32, C.c_double, C.c_void_p
    L.orc_lcg_stream.argtypes = [u64, i64, vp, vp, vp]
    L.orc_n_sections.argtypes = [u32]
    L.orc_n_sections.restype = i64
    L.orc_monomials.argtypes = [u32, C.c_int, vp]
    L.orc_monomials.restype = i64
    L.orc_quintic_eval.argtypes = [_dp, i64, dbl, vp, vp]
    L.orc_sample_sz.argtypes = [i64, u64, C.c_int, dbl, C.c_int, _dp, C.POINTER(i64)]
    L.orc_sample_sz.restype = i64
    L.orc_sample_newton.argtypes = [i64, u64, C.c_int, dbl, dbl, _dp, C.POINTER(i64), C.c_void_p]
    L.orc_sample_newton.restype = i64
    L.orc_newton_project.argtypes = [_dp, dbl, dbl, C.c_int]
    L.orc_newton_project.restype = C.c_int
    L.orc_sz_lines.argtypes = [i64, u64, C.c_int, dbl, C.c_int, _dp, _i32p]
    L.orc_weights_logomega.argtypes = [_dp, i64, dbl, vp, vp, vp]
    L.orc_sections.argtypes = [_dp, i64, _u32p, i64, _dp, vp]
    L.orc_fs_gram.argtypes = [_dp, _dp, i64, i64, _dp, _dp]
    L.orc_cholesky.argtypes = [_dp, _dp, i64, _dp, _dp]
    L.orc_cholesky.restype = C.c_int
    L.orc_apply_inv_l.argtypes = [_dp, _dp, i64, i64, _dp, vp]
    L.orc_lu_invert.argtypes = [_dp, i64, _dp]
    L.orc_lu_invert.restype = C.c_int
    L.orc_donaldson_step.argtypes = [_dp, _dp, i64, i64, _dp, _dp, vp, vp, vp]
    L.orc_donaldson_step.restype = dbl
    L.orc_eta.argtypes = [_dp, _dp, _dp, _dp, _dp, i64, i64, dbl, _dp]
    L.orc_sigma_reduce.argtypes = [_dp, _dp, i64, vp, vp]
    L.orc_sigma_reduce.restype = dbl
    L.orc_num_threads.restype = C.c_int
    L.orc_sigma_sq_grad.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, i64, i64, dbl, _dp, vp]
    L.orc_sigma_sq_grad.restype = dbl
    L.orc_adam_step.argtypes = [_dp, i64, _dp, _dp, _dp, i64, dbl, dbl, dbl, dbl]
    L.orc_adam_step.restype = i64
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def lcg_stream(seed: int, n: int):
    states = np.zeros(n, np.uint64)
    f64s = np.zeros(n, np.float64)
    normals = np.zeros(n, np.float64)
    lib().orc_lcg_stream(seed, n, _ptr(states), _ptr(f64s), _ptr(normals))
    return states, f64s, normals


def n_sections(k: int) -> int:
    return int(lib().orc_n_sections(k))


def monomials(k: int, quotient: bool = False) -> np.ndarray:
    n = lib().orc_monomials(k, int(quotient), None)
    out = np.zeros((n, 5), np.uint32)
    lib().orc_monomials(k, int(quotient), _ptr(out))
    return out


def quintic_eval(pts: np.ndarray, psi: float = 0.0):
    pts = np.ascontiguousarray(pts, np.float64)
    n = pts.shape[0]
    f = np.zeros((n, 2))
    g = np.zeros((n, 10))
    lib().orc_quintic_eval(pts, n, psi, _ptr(f), _ptr(g))
    return f, g


def sample_sz(n: int, seed: int, psi: float = 0.0, dwork: bool | None = None, dk_iters: int = 50,
              polish: int = 0):
    """dwork=False: sample_quintic_points_sz; dwork=True: sample_deformed_quintic_points_sz.
    Default: Fermat form at psi == 0, Dwork form otherwise. polish > 0 adds that many FP64
    Newton passes per root after the Durand-Kerner sweeps (not in the reference)."""
    if dwork is None:
        dwork = psi != 0.0
    out = np.zeros((n, 10))
    tried = C.c_int64(0)
    L = lib()
    if polish:
        if not getattr(L, "_polish_bound", False):
            L.orc_sample_sz_polish.argtypes = [C.c_int64, C.c_uint64, C.c_int, C.c_double,
                                               C.c_int, C.c_int, _dp, C.POINTER(C.c_int64)]
            L.orc_sample_sz_polish.restype = C.c_int64
            L._polish_bound = True
        got = L.orc_sample_sz_polish(n, seed, int(dwork), psi, dk_iters, polish, out,
                                     C.byref(tried))
    else:
        got = L.orc_sample_sz(n, seed, int(dwork), psi, dk_iters, out, C.byref(tried))
    return out[:got].copy(), int(tried.value)


def sample_newton(n: int, seed: int, psi: float = 0.0, tol: float = 1e-12,
                  dwork: bool | None = None):
    """dwork=False: sample_quintic_points (QuinticSolver::new); dwork=True:
    sample_deformed_quintic_points (QuinticSolver::new_with_psi). Returns (points, tries,
    per-try accept flags)."""
    if dwork is None:
        dwork = psi != 0.0
    out = np.zeros((n, 10))
    tried = C.c_int64(0)
    acc = np.zeros(n * (8 if dwork else 4), np.int32)
    got = lib().orc_sample_newton(n, seed, int(dwork), psi, tol, out, C.byref(tried), _ptr(acc))
    return out[:got].copy(), int(tried.value), acc[: tried.value].copy()


def newton_project(z, psi: float = 0.0, tol: float = 1e-12, max_iter: int = 50):
    """newton_project_to_(deformed_)quintic (quintic.rs:3334-3443): (accepted, projected z)."""
    zz = np.array(z, dtype=np.float64)
    ok = lib().orc_newton_project(zz, psi, tol, max_iter)
    return bool(ok), zz


def sz_lines(n_lines: int, seed: int, psi: float = 0.0, dwork: bool = False, dk_iters: int = 50):
    zs = np.zeros((n_lines, 5, 10))
    acc = np.zeros((n_lines, 5), np.int32)
    lib().orc_sz_lines(n_lines, seed, int(dwork), psi, dk_iters, zs, acc)
    return zs, acc


def weights_logomega(pts: np.ndarray, psi: float = 0.0):
    pts = np.ascontiguousarray(pts, np.float64)
    n = pts.shape[0]
    w = np.zeros(n)
    lo = np.zeros(n)
    ce = np.zeros((n, 2), np.int32)
    lib().orc_weights_logomega(pts, n, psi, _ptr(w), _ptr(lo), _ptr(ce))
    return w, lo, ce


def sections(pts: np.ndarray, mono: np.ndarray, derivs: bool = True):
    pts = np.ascontiguousarray(pts, np.float64)
    mono = np.ascontiguousarray(mono, np.uint32)
    n, nb = pts.shape[0], mono.shape[0]
    sv = np.zeros((n, 2 * nb))
    sd = np.zeros((n, 5, 2 * nb)) if derivs else None
    lib().orc_sections(pts, n, mono, nb, sv, _ptr(sd))
    return sv, sd


def fs_gram(sv: np.ndarray, w: np.ndarray):
    n, two_n = sv.shape
    nb = two_n // 2
    gre = np.zeros((nb, nb))
    gim = np.zeros((nb, nb))
    lib().orc_fs_gram(np.ascontiguousarray(sv), np.ascontiguousarray(w), n, nb, gre, gim)
    return gre, gim


def cholesky(gre: np.ndarray, gim: np.ndarray):
    nb = gre.shape[0]
    lre = np.zeros((nb, nb))
    lim = np.zeros((nb, nb))
    rc = lib().orc_cholesky(np.ascontiguousarray(gre), np.ascontiguousarray(gim), nb, lre, lim)
    if rc != 0:
        raise ValueError("FS Gram matrix not positive-definite")
    return lre, lim


def orthonormalise(sv: np.ndarray, sd, w: np.ndarray):
    """orthonormalise_basis_fs_gram (quintic.rs:3140-3165), in place on copies."""
    sv = sv.copy()
    sd = None if sd is None else sd.copy()
    gre, gim = fs_gram(sv, w)
    lre, lim = cholesky(gre, gim)
    n, two_n = sv.shape
    lib().orc_apply_inv_l(lre, lim, n, two_n // 2, sv, _ptr(sd))
    return sv, sd, (lre, lim)


def lu_invert(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float64)
    out = np.zeros_like(a)
    if lib().orc_lu_invert(a, a.shape[0], out) != 0:
        raise ValueError("singular")
    return out


def identity_block(nb: int) -> np.ndarray:
    return np.eye(2 * nb)


def donaldson_step(sv: np.ndarray, w: np.ndarray, h_block: np.ndarray, want_t: bool = False):
    n, two_n = sv.shape
    nb = two_n // 2
    h_new = np.zeros((two_n, two_n))
    tre = np.zeros((nb, nb)) if want_t else None
    tim = np.zeros((nb, nb)) if want_t else None
    kinv = np.zeros(n) if want_t else None
    r = lib().orc_donaldson_step(np.ascontiguousarray(sv), np.ascontiguousarray(w), n, nb,
                                 np.ascontiguousarray(h_block), h_new, _ptr(tre), _ptr(tim),
                                 _ptr(kinv))
    if want_t:
        return h_new, r, (tre, tim, kinv)
    return h_new, r


def t_operator_raw_route34(section_values, weights, h_re, h_im):
    """Cy3DonaldsonGpu::t_operator_raw (route34/cy3_donaldson_gpu.rs:249-320) with the kernels'
    rules: K_p = sum_ab h_re Re(conj(s_a) s_b) + h_im Im(conj(s_a) s_b) (:62-92), raised to 1e-30
    when not above it (a NaN included, :88-89); h_pair[a][b] = sum (w_p / K_p) conj(s_a) s_b over
    the points with finite w_p > 0 and finite K_p > 0 (:100-160). cpu_t_operator_reference
    (:338-386) is the same sum without the weight test. Returns (h_pair_re, h_pair_im)."""
    sv = np.asarray(section_values, dtype=np.float64)
    n = sv.shape[0]
    sv = sv.reshape(n, -1)
    s = sv[:, 0::2] + 1j * sv[:, 1::2]
    h = np.asarray(h_re, dtype=np.float64) + 1j * np.asarray(h_im, dtype=np.float64)
    # Re(conj(h_ab) conj(s_a) s_b) = h_re Re(conj(s_a) s_b) + h_im Im(conj(s_a) s_b)
    k = np.einsum("pa,ab,pb->p", s.conj(), h.conj(), s).real
    k = np.where(k > 1e-30, k, 1e-30)
    w = np.asarray(weights, dtype=np.float64)
    ok = np.isfinite(w) & (w > 0) & np.isfinite(k) & (k > 0)
    f = np.where(ok, w / np.where(ok, k, 1.0), 0.0)
    hp = (s.conj() * f[:, None]).T @ s
    return hp.real.copy(), hp.imag.copy()


def sigma_route34(section_values, section_derivs, frames, omega_sq, weights, h_re, h_im):
    """cpu_sigma_reference (route34/cy3_sigma_gpu.rs:585-721): per point K = s^H H s (raised to
    1e-30), dk_i = s^H H ds_i, g_amb_ij = (ds_j^H H ds_i) / K - dk_i conj(dk_j) / K^2,
    g_tan = conj(T) g_amb T^T with the 3 x 8 frame T, eta = |Re det g_tan| / |Omega|^2 for points
    with finite w > 0 and |Omega|^2 > 0 and a finite det of magnitude >= 1e-30; then
    kappa = sum w eta / sum w and sigma = sum w (eta / kappa - 1)^2 / sum w (NaN when the weight
    sum or kappa is below 1e-30). Returns (sigma, eta)."""
    sv = np.asarray(section_values, dtype=np.float64)
    n = sv.shape[0]
    sv = sv.reshape(n, -1)
    nb = sv.shape[1] // 2
    s = sv[:, 0::2] + 1j * sv[:, 1::2]
    d = np.asarray(section_derivs, dtype=np.float64).reshape(n, 8, 2 * nb)
    ds = d[..., 0::2] + 1j * d[..., 1::2]
    H = np.asarray(h_re, dtype=np.float64) + 1j * np.asarray(h_im, dtype=np.float64)
    hs = s @ H.T
    hds = ds @ H.T
    k = np.einsum("pa,pa->p", s.conj(), hs).real
    ks = np.maximum(k, 1e-30)
    dk = np.einsum("pa,pia->pi", s.conj(), hds)
    m = np.einsum("pja,pia->pij", ds.conj(), hds)
    g = m / ks[:, None, None] - dk[:, :, None] * dk.conj()[:, None, :] / (ks * ks)[:, None, None]
    f = np.asarray(frames, dtype=np.float64).reshape(n, 3, 8, 2)
    t = f[..., 0] + 1j * f[..., 1]
    gt = np.einsum("pai,pij,pbj->pab", t.conj(), g, t)
    m1 = gt[:, 1, 1] * gt[:, 2, 2] - gt[:, 1, 2] * gt[:, 2, 1]
    m2 = gt[:, 1, 0] * gt[:, 2, 2] - gt[:, 1, 2] * gt[:, 2, 0]
    m3 = gt[:, 1, 0] * gt[:, 2, 1] - gt[:, 1, 1] * gt[:, 2, 0]
    det = (gt[:, 0, 0] * m1 - gt[:, 0, 1] * m2 + gt[:, 0, 2] * m3).real
    w = np.asarray(weights, dtype=np.float64)
    om = np.asarray(omega_sq, dtype=np.float64)
    ok = (np.isfinite(w) & (w > 0) & np.isfinite(om) & (om > 0) & np.isfinite(det)
          & (np.abs(det) >= 1e-30))
    eta = np.where(ok, np.abs(det) / np.where(ok, om, 1.0), np.nan)
    tw = float(np.sum(w[ok]))
    if tw < 1e-30:
        return float("nan"), eta
    kap = float(np.sum(w[ok] * eta[ok])) / tw
    if abs(kap) < 1e-30:
        return float("nan"), eta
    return float(np.sum(w[ok] * (eta[ok] / kap - 1.0) ** 2)) / tw, eta


def donaldson_solve(sv, w, h_block, max_iter: int, tol: float):
    """donaldson_solve (quintic.rs:433-443): returns (h, residual history)."""
    res = []
    h = h_block
    for _ in range(max_iter):
        h, r = donaldson_step(sv, w, h)
        res.append(r)
        if r < tol:
            break
    return h, res


def eta(pts, sv, sd, log_om, h_block, psi: float = 0.0):
    n, two_n = sv.shape
    out = np.zeros(n)
    lib().orc_eta(np.ascontiguousarray(pts), np.ascontiguousarray(sv), np.ascontiguousarray(sd),
                  np.ascontiguousarray(log_om), np.ascontiguousarray(h_block), n, two_n // 2,
                  psi, out)
    return out


def sigma_reduce(eta_v, w):
    kap = C.c_double(0)
    l2 = C.c_double(0)
    s = lib().orc_sigma_reduce(np.ascontiguousarray(eta_v), np.ascontiguousarray(w), len(w),
                               C.byref(kap), C.byref(l2))
    return s, kap.value, l2.value


def sigma(pts, sv, sd, log_om, w, h_block, psi: float = 0.0) -> float:
    return sigma_reduce(eta(pts, sv, sd, log_om, h_block, psi), w)[0]


def sigma_sq_grad(pts, sv, sd, log_om, w, h_block, psi: float = 0.0):
    """sigma_squared_and_gradient (quintic.rs:2714-2860): (sigma_L2^2, grad [2 N^2], sigma_L1).
    grad[:N^2] = d/dh_re (a, b) with the diagonal halved; grad[N^2:] = d/dh_im (a, b), a < b."""
    n, two_n = sv.shape
    nb = two_n // 2
    g = np.zeros(2 * nb * nb)
    l1 = C.c_double(0)
    s2 = lib().orc_sigma_sq_grad(np.ascontiguousarray(pts), np.ascontiguousarray(sv),
                                 np.ascontiguousarray(sd), np.ascontiguousarray(log_om),
                                 np.ascontiguousarray(w), np.ascontiguousarray(h_block), n, nb, psi,
                                 g, C.byref(l1))
    return s2, g, l1.value


def per_point(pts, sv, sd, log_om, w, h_block, psi: float = 0.0):
    """cpu_compute_per_point_reference (gpu_adam.rs:2677-2739): [n, 2 + 2 N^2] rows
    (eta_p, w_p, d log eta_p / dh)."""
    L = lib()
    if not getattr(L, "_pp_bound", False):
        L.orc_per_point.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, C.c_int64, C.c_int64,
                                    C.c_double, _dp]
        L._pp_bound = True
    n, two_n = sv.shape
    nb = two_n // 2
    out = np.zeros((n, 2 + 2 * nb * nb))
    L.orc_per_point(np.ascontiguousarray(pts), np.ascontiguousarray(sv), np.ascontiguousarray(sd),
                    np.ascontiguousarray(log_om), np.ascontiguousarray(w),
                    np.ascontiguousarray(h_block), n, nb, psi, out)
    return out


def adam_step(h_block, grad, m, v, t: int, lr: float, beta1: float = 0.9, beta2: float = 0.999,
              eps: float = 1e-8):
    """One Adam step of sigma_refine_analytic (quintic.rs:453-514, 2071-2161, 3169-3182) on a
    copy of h_block; m, v are updated in place. Returns (h_new, t_new)."""
    h = np.ascontiguousarray(h_block, dtype=np.float64).copy()
    nb = h.shape[0] // 2
    t2 = lib().orc_adam_step(h, nb, np.ascontiguousarray(grad), m, v, t, lr, beta1, beta2, eps)
    return h, t2


def sigma_refine_analytic(prob, h_block, n_iter: int, lr: float):
    """QuinticSolver::sigma_refine_analytic (quintic.rs:453-514): Adam with zeroed moments.
    Returns (history of sigma_L1, final h_block)."""
    nb = h_block.shape[0] // 2
    m = np.zeros(2 * nb * nb)
    v = np.zeros(2 * nb * nb)
    t = 0
    h = np.array(h_block, dtype=np.float64)
    hist = []
    for _ in range(n_iter):
        _, g, s1 = sigma_sq_grad(prob.pts, prob.sv, prob.sd, prob.log_om, prob.w, h, prob.psi)
        if not np.isfinite(s1):
            break
        hist.append(s1)
        h, t = adam_step(h, g, m, v, t, lr)
    return hist, h


def sigma_at_sample(prob, h_block, n_pts: int, seed: int, sampler: str = "sz",
                    want_eta: bool = False):
    """sigma_at_metric_closure (quintic.rs:1520-1616) with the Bergman metric of h_block in the
    basis of `prob` as the callback: a fresh sample (sample_quintic_points_with, :4006-4016;
    sampler "sz" or "newton", Fermat at psi = 0, the Dwork forms at prob.psi != 0), its CY
    weights and log|Omega|^2, the raw sections carried through prob's FS-Gram basis change
    (apply_inverse_l_to_*, :3095-3133), then eta and the two-pass reduction of
    compute_sigma_from_workspace (:1301-1499). Returns (sigma_L1, kappa, sigma_L2^2, n_actual)
    and, with want_eta, also eta and the points."""
    if sampler == "sz":
        pts, _ = sample_sz(n_pts, seed, prob.psi)
    elif sampler == "newton":
        pts, _, _ = sample_newton(n_pts, seed, prob.psi)
    else:
        raise ValueError(sampler)
    n = pts.shape[0]
    if n == 0:
        out = (float("nan"), float("nan"), float("nan"), 0)
        return out + (np.zeros(0), pts) if want_eta else out
    w, lo, _ = weights_logomega(pts, prob.psi)
    sv, sd = sections(pts, prob.mono, True)
    if prob.L is not None:
        lre, lim = prob.L
        lib().orc_apply_inv_l(lre, lim, n, prob.nb, sv, _ptr(sd))
    e = eta(pts, sv, sd, lo, h_block, prob.psi)
    s1, kap, l2 = sigma_reduce(e, w)
    out = (s1, kap, l2, n)
    return out + (e, pts) if want_eta else out


class Problem:
    """The static inputs of QuinticSolver::new_with_sampler(.., ShiffmanZelditch) or
    ::new_with_psi_sz (quintic.rs:241-322), computed by the oracle."""

    def __init__(self, k: int, n_pts: int, seed: int, psi: float = 0.0, dwork: bool | None = None,
                 quotient: bool = False, derivs: bool = True, fs_gram: bool = True):
        self.k, self.seed, self.psi = k, seed, psi
        self.mono = monomials(k, quotient)
        self.nb = self.mono.shape[0]
        self.pts, self.lines_tried = sample_sz(n_pts, seed, psi, dwork)
        self.n = self.pts.shape[0]
        self.w, self.log_om, self.chart_elim = weights_logomega(self.pts, psi)
        self.sv_raw, self.sd_raw = sections(self.pts, self.mono, derivs)
        if fs_gram:
            self.sv, self.sd, self.L = orthonormalise(self.sv_raw, self.sd_raw, self.w)
        else:
            self.sv, self.sd, self.L = self.sv_raw, self.sd_raw, None
        self.h = identity_block(self.nb)

    def sigma(self, h=None) -> float:
        return sigma(self.pts, self.sv, self.sd, self.log_om, self.w, self.h if h is None else h,
                     self.psi)

    def sigma_l2sq(self, h=None) -> float:
        e = eta(self.pts, self.sv, self.sd, self.log_om, self.h if h is None else h, self.psi)
        return sigma_reduce(e, self.w)[2]

    def sigma_sq_grad(self, h=None):
        return sigma_sq_grad(self.pts, self.sv, self.sd, self.log_om, self.w,
                             self.h if h is None else h, self.psi)


# ----------------------------------------------------------------------------------------------
# Metric Laplacian (route34/metric_laplacian.rs) and Yukawa overlaps (route34/yukawa_overlap_real.rs)
# ----------------------------------------------------------------------------------------------
def _bind_galerkin(L):
    i64, u64, u32, dbl, vp = C.c_int64, C.c_uint64, C.c_uint32, C.c_double, C.c_void_p
    L.orc_lap_basis.argtypes = [u32, vp]
    L.orc_lap_basis.restype = i64
    L.orc_lap_ambient_mk.argtypes = [_dp, _dp, _dp, i64, _u32p, i64, _dp, _dp]
    L.orc_herm_inv_reg.argtypes = [_dp, i64, dbl, _dp]
    L.orc_herm_jacobi.argtypes = [_dp, i64, i64, dbl, _dp, vp]
    L.orc_herm_jacobi.restype = i64
    L.orc_lap_spectrum_from_mk.argtypes = [_dp, _dp, i64, i64, dbl, dbl, _dp]
    L.orc_lap_quintic_mk.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, i64, i64, dbl, _u32p, i64,
                                     C.c_int, _dp, _dp]
    L.orc_lap_quintic_mk.restype = dbl
    L.orc_point_metric.argtypes = [_dp, _dp, _dp, _dp, i64, i64, dbl, _dp, _dp, _i32p]
    L.orc_yuk_sz_weights.argtypes = [_dp, _dp, i64, _dp]
    L.orc_yuk_sz_weights.restype = dbl
    L.orc_yuk_triple.argtypes = [_dp, _dp, _dp, vp, i64, i64, vp, i64, _dp]
    L.orc_yuk_bootstrap_indices.argtypes = [u64, i64, i64, vp]
    L.orc_yuk_couplings.argtypes = [_dp, _dp, _dp, vp, i64, i64, i64, u64, _dp, _dp,
                                    C.POINTER(dbl), C.POINTER(dbl)]


def _glib():
    L = lib()
    if not getattr(L, "_galerkin_bound", False):
        _bind_galerkin(L)
        L._galerkin_bound = True
    return L


def _c128(a):
    """complex array -> contiguous float64 view [..., 2]."""
    a = np.ascontiguousarray(a, np.complex128)
    return a.view(np.float64).reshape(a.shape + (2,))


def lap_basis(max_deg: int) -> np.ndarray:
    """build_test_basis (metric_laplacian.rs:87-115): [nb, 8] exponents."""
    n = _glib().orc_lap_basis(max_deg, None)
    out = np.zeros((n, 8), np.uint32)
    _glib().orc_lap_basis(max_deg, _ptr(out))
    return out


def lap_ambient_mk(pts8: np.ndarray, w: np.ndarray, omega: np.ndarray, exps: np.ndarray):
    """build_mass_matrix / build_stiffness_matrix (metric_laplacian.rs:424-524).
    pts8: complex [n, 8]; omega: complex [n]. Returns complex (M, K)."""
    n, nb = pts8.shape[0], exps.shape[0]
    M = np.zeros((nb, nb, 2))
    K = np.zeros((nb, nb, 2))
    _glib().orc_lap_ambient_mk(_c128(pts8).reshape(n, 16), np.ascontiguousarray(w, np.float64),
                               _c128(omega).reshape(n, 2), n, np.ascontiguousarray(exps, np.uint32),
                               nb, M, K)
    return M[..., 0] + 1j * M[..., 1], K[..., 0] + 1j * K[..., 1]


def herm_inv_reg(h: np.ndarray, eps: float) -> np.ndarray:
    """invert_hermitian_regularised (metric_laplacian.rs:547-613)."""
    n = h.shape[0]
    out = np.zeros((n, n, 2))
    _glib().orc_herm_inv_reg(_c128(h).reshape(n, n, 2), n, eps, out)
    return out[..., 0] + 1j * out[..., 1]


def herm_jacobi(a: np.ndarray, max_sweeps: int = 128, tol: float = 1e-12, vectors: bool = False):
    """hermitian_jacobi_n (metric_laplacian.rs:616-693): (eigenvalues ascending, vectors, sweeps)."""
    n = a.shape[0]
    ev = np.zeros(n)
    vec = np.zeros((n, n, 2)) if vectors else None
    sw = _glib().orc_herm_jacobi(_c128(a).reshape(n, n, 2), n, max_sweeps, tol, ev, _ptr(vec))
    return ev, (None if vec is None else vec[..., 0] + 1j * vec[..., 1]), int(sw)


def lap_spectrum_from_mk(M, K, max_sweeps: int = 128, tol: float = 1e-12, reg: float = 1e-10):
    """The spectrum step of compute_metric_laplacian_spectrum (metric_laplacian.rs:342-389):
    |eigenvalues| of the Hermitian part of (M + reg ||M||_F)^-1 K, ascending."""
    nb = M.shape[0]
    full = np.zeros(nb)
    _glib().orc_lap_spectrum_from_mk(_c128(M).reshape(nb, nb, 2), _c128(K).reshape(nb, nb, 2), nb,
                                     max_sweeps, tol, reg, full)
    return full


def lap_spectrum_ambient(pts8, w, omega, max_deg: int = 4, n_low: int = 15,
                         max_sweeps: int = 128, tol: float = 1e-12, reg: float = 1e-10):
    """compute_metric_laplacian_spectrum (metric_laplacian.rs:296-418): (low, full, M, K)."""
    exps = lap_basis(max_deg)
    M, K = lap_ambient_mk(pts8, w, omega, exps)
    full = lap_spectrum_from_mk(M, K, max_sweeps, tol, reg)
    return full[:min(n_low, len(full))], full, M, K


def lap_quintic_mk(prob, h_block, sdeg: int = 3):
    """Intrinsic-metric Galerkin matrices on phi_ab = s_a conj(s_b)/|z|^(2 sdeg) (raw degree-sdeg
    monomials s) at the balanced h of `prob`. Returns complex (M, K, sum w)."""
    smono = monomials(sdeg, False)
    ns = smono.shape[0]
    B = ns * ns
    M = np.zeros((B, B, 2))
    K = np.zeros((B, B, 2))
    tw = _glib().orc_lap_quintic_mk(np.ascontiguousarray(prob.pts), np.ascontiguousarray(prob.w),
                                    np.ascontiguousarray(prob.log_om), np.ascontiguousarray(prob.sv),
                                    np.ascontiguousarray(prob.sd), np.ascontiguousarray(h_block),
                                    prob.n, prob.nb, prob.psi, smono, ns, sdeg, M, K)
    return M[..., 0] + 1j * M[..., 1], K[..., 0] + 1j * K[..., 1], tw


def point_metric(prob, h_block):
    """Per-point ambient metric g_amb (complex [n,5,5]), affine frame (complex [n,3,5]), ok flags."""
    n = prob.n
    ga = np.zeros((n, 5, 5, 2))
    fr = np.zeros((n, 3, 10))
    ok = np.zeros(n, np.int32)
    _glib().orc_point_metric(np.ascontiguousarray(prob.pts), np.ascontiguousarray(prob.sv),
                             np.ascontiguousarray(prob.sd), np.ascontiguousarray(h_block), n,
                             prob.nb, prob.psi, ga, fr, ok)
    frc = fr.reshape(n, 3, 5, 2)
    return ga[..., 0] + 1j * ga[..., 1], frc[..., 0] + 1j * frc[..., 1], ok.astype(bool)


def yuk_sz_weights(w, omega):
    """sz_weights (yukawa_overlap_real.rs:222-268): (sz, uniformity score)."""
    n = len(w)
    sz = np.zeros(n)
    score = _glib().orc_yuk_sz_weights(np.ascontiguousarray(w, np.float64),
                                       _c128(omega).reshape(n, 2), n, sz)
    return sz, score


def yuk_triple(modes, omega, weights, hdiag=None, idx=None):
    """triple_overlap_cpu (yukawa_overlap_real.rs:271-315), or bootstrap_one (:318-357) over
    resample indices idx. modes: complex [n_modes, n]. Returns complex [nm, nm, nm]."""
    nm, n = modes.shape
    y = np.zeros((nm, nm, nm, 2))
    hd = None if hdiag is None else np.ascontiguousarray(hdiag, np.float64)
    ix = None if idx is None else np.ascontiguousarray(idx, np.int64)
    _glib().orc_yuk_triple(_c128(modes).reshape(nm, n, 2), _c128(omega).reshape(n, 2),
                           np.ascontiguousarray(weights, np.float64), _ptr(hd), nm, n, _ptr(ix),
                           0 if ix is None else len(ix), y)
    return y[..., 0] + 1j * y[..., 1]


def yuk_bootstrap_indices(seed: int, b: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.int64)
    _glib().orc_yuk_bootstrap_indices(seed, b, n, _ptr(out))
    return out


def yuk_couplings(modes, omega, w, hdiag=None, n_bootstrap: int = 64, seed: int = 0xBADDCAFEF00D):
    """compute_yukawa_couplings (yukawa_overlap_real.rs:372-525) given h_v_diag:
    (Y, uncertainty, uniformity score, convergence ratio)."""
    nm, n = modes.shape
    y = np.zeros((nm, nm, nm, 2))
    unc = np.zeros((nm, nm, nm))
    sc = C.c_double(0)
    cv = C.c_double(0)
    hd = None if hdiag is None else np.ascontiguousarray(hdiag, np.float64)
    _glib().orc_yuk_couplings(_c128(modes).reshape(nm, n, 2), _c128(omega).reshape(n, 2),
                              np.ascontiguousarray(w, np.float64), _ptr(hd), nm, n, n_bootstrap,
                              seed, y, unc, C.byref(sc), C.byref(cv))
    return y[..., 0] + 1j * y[..., 1], unc, sc.value, cv.value


def legacy_step(sv, h_block, w=None):
    """donaldson_iteration_quintic (quintic.rs:4140-4222), or the _weighted form (:4256-4331)
    when w is given: the un-inverted update h <- T(h), trace-normalised."""
    L = lib()
    if not getattr(L, "_legacy_bound", False):
        L.orc_legacy_step.argtypes = [_dp, C.c_void_p, _dp, C.c_int64, C.c_int64, _dp]
        L._legacy_bound = True
    n, two_n = sv.shape
    out = np.zeros((two_n, two_n))
    wv = None if w is None else np.ascontiguousarray(w, np.float64)
    L.orc_legacy_step(np.ascontiguousarray(sv), _ptr(wv), np.ascontiguousarray(h_block), n,
                      two_n // 2, out)
    return out


def legacy_solve(sv, max_iter: int, tol: float, w=None):
    """donaldson_solve_quintic(_weighted) (quintic.rs:4224-4248, 4333-4357): from h = I."""
    h = identity_block(sv.shape[1] // 2)
    res = []
    for _ in range(max_iter):
        hn = legacy_step(sv, h, w)
        r = float(np.sqrt(np.sum((hn - h) ** 2)))
        res.append(r)
        h = hn
        if r < tol:
            break
    return h, res


# ----------------------------------------------------------------------------------------------
# L-BFGS sigma refinement and random restarts (quintic.rs:545-582, 678-880, 1100-1223, 2003-2048)
# ----------------------------------------------------------------------------------------------
def _unpack_h(h_block):
    """h_re[a][b] = h[2a][2b], h_im[a][b] = -h[2a][2b+1] (quintic.rs:5128-5140)."""
    h = np.asarray(h_block, dtype=np.float64)
    return h[0::2, 0::2].copy(), -h[0::2, 1::2].copy()


def _pack_h(hre, him):
    """pack_h_block_workspace (quintic.rs:2148-2161)."""
    nb = hre.shape[0]
    h = np.zeros((2 * nb, 2 * nb))
    h[0::2, 0::2] = hre
    h[0::2, 1::2] = -him
    h[1::2, 0::2] = him
    h[1::2, 1::2] = hre
    return h


def renorm_trace(h_block):
    """renormalise_h_trace_workspace (quintic.rs:3169-3182): sequential trace of the even
    diagonal, scale to N when it exceeds 1e-10."""
    h = np.array(h_block, dtype=np.float64)
    nb = h.shape[0] // 2
    tr = 0.0
    for a in range(nb):
        tr += h[2 * a, 2 * a]
    if tr > 1e-10:
        h *= nb / tr
    return h


def lbfgs_apply_step(h_block, d, step: float):
    """apply_lbfgs_step (quintic.rs:2003-2048) on a copy: h_re += step d_re for every (a, b),
    then the (a, b)/(b, a) average; h_im(a, b) += step d_im(a, b), h_im(b, a) -= the same, a < b;
    zero h_im diagonal; pack; trace renormalisation."""
    hre, him = _unpack_h(h_block)
    nb = hre.shape[0]
    n2 = nb * nb
    d = np.asarray(d, dtype=np.float64)
    hre += step * d[:n2].reshape(nb, nb)
    for a in range(nb):
        for b in range(a + 1, nb):
            avg = 0.5 * (hre[a, b] + hre[b, a])
            hre[a, b] = avg
            hre[b, a] = avg
    for a in range(nb):
        for b in range(a + 1, nb):
            upd = step * d[n2 + a * nb + b]
            him[a, b] += upd
            him[b, a] -= upd
    for a in range(nb):
        him[a, a] = 0.0
    return renorm_trace(_pack_h(hre, him))


def lbfgs_direction(g, pairs):
    """The two-loop recursion of quintic.rs:746-776: d = -H g from the curvature pairs
    (s, y, rho = 1/(s.y)) listed oldest first, with H0 = gamma I, gamma = (s.y)/(y.y) of the
    newest pair (1 when y.y <= 1e-30 or rho is not finite, or without pairs)."""
    q = np.array(g, dtype=np.float64)
    m = len(pairs)
    alphas = [0.0] * m
    for k in range(m - 1, -1, -1):
        s, y, rho = pairs[k]
        a = rho * float(np.dot(s, q))
        alphas[k] = a
        q -= a * y
    gamma = 1.0
    if m > 0:
        _, y_l, rho_l = pairs[-1]
        yy = float(np.dot(y_l, y_l))
        if yy > 1e-30 and np.isfinite(rho_l):
            gamma = 1.0 / (rho_l * yy)
    r = q * gamma
    for k in range(m):
        s, y, rho = pairs[k]
        beta = rho * float(np.dot(y, r))
        r += s * (alphas[k] - beta)
    return -r


def lbfgs_minimize(x0, fgs, apply_step, n_iter: int, m_history: int, c1: float = 1e-4,
                   max_backtracks: int = 30):
    """The optimiser of sigma_refine_lbfgs_gpu_legacy (quintic.rs:678-880) on a generic
    objective: fgs(x) -> (f, grad, score), apply_step(x, d, step) -> the trial point. Armijo
    backtracking from step 1 (or 1/|g| for steepest descent) with halving, a steepest-descent
    retry when the L-BFGS line search fails, pairs kept when s.y > 1e-12, and the point with the
    lowest score kept. Returns (score history, best x, trial points that were evaluated)."""
    f, g, score = fgs(x0)
    hist = []
    trials = []
    if not np.isfinite(score):
        return hist, x0, trials
    hist.append(score)
    best, best_x = score, x0
    x = x0
    pairs = []
    it = 0
    while it < n_iter:
        d = lbfgs_direction(g, pairs)
        gd = float(np.dot(g, d))
        used = len(pairs) > 0
        if not np.isfinite(gd) or gd >= 0.0:
            pairs = []
            d = -g
            gd = float(np.dot(g, d))
            used = False
            if gd >= 0.0:
                break
        g_norm = float(np.sqrt(np.dot(g, g)))
        step = 1.0 if used else (1.0 / g_norm if g_norm > 1e-30 else 1.0)
        x_save = x
        accepted = False
        for _ in range(max_backtracks):
            x = apply_step(x_save, d, step)
            f2, g2, s2 = fgs(x)
            trials.append(x)
            if np.isfinite(f2) and f2 <= f + c1 * step * gd:
                accepted = True
                break
            step *= 0.5
        if not accepted and used:
            pairs = []
            d = -g
            gd = float(np.dot(g, d))
            sd = 1.0 / g_norm if g_norm > 1e-30 else 1.0
            for _ in range(max_backtracks):
                x = apply_step(x_save, d, sd)
                f2, g2, s2 = fgs(x)
                trials.append(x)
                if np.isfinite(f2) and f2 <= f + c1 * sd * gd:
                    accepted = True
                    step = sd
                    break
                sd *= 0.5
        if not accepted:
            x = x_save
            break
        s_k = step * d
        y_k = g2 - g
        sy = float(np.dot(s_k, y_k))
        if sy > 1e-12:
            pairs.append((s_k, y_k, 1.0 / sy))
            if len(pairs) > m_history:
                pairs.pop(0)
        if np.isfinite(s2) and s2 < best:
            best, best_x = s2, x
        f, g = f2, g2
        hist.append(s2)
        it += 1
    return hist, best_x, trials


def sigma_refine_lbfgs(prob, h_block, n_iter: int, m_history: int):
    """QuinticSolver::sigma_refine_lbfgs_gpu(_legacy) (quintic.rs:678-880): L-BFGS on sigma_L2^2
    over (h_re, h_im), scored by sigma_L1. Returns (history, best h_block, min over every
    evaluated trial point of the smallest eigenvalue of its complex G)."""
    def fgs(h):
        s2, g, s1 = prob.sigma_sq_grad(h)
        return s2, g, s1
    hist, h_best, trials = lbfgs_minimize(np.array(h_block, dtype=np.float64), fgs,
                                          lbfgs_apply_step, n_iter, m_history)
    min_eig = np.inf
    for h in trials:
        hre, him = _unpack_h(h)
        min_eig = min(min_eig, float(np.linalg.eigvalsh(hre + 1j * him).min()))
    return hist, h_best, min_eig


def perturb_h_hermitian(h_block, scale: float, draws):
    """perturb_h_block_hermitian (quintic.rs:1195-1223): for each a, the diagonal h_re gets
    scale (2u - 1); for each b > a, h_re(a, b) and h_re(b, a) get scale (2u - 1) and h_im(a, b)
    gets +scale (2u' - 1), h_im(b, a) the opposite, with u, u' the next LCG draws (`draws`
    iterates over next_f64 values)."""
    hre, him = _unpack_h(h_block)
    nb = hre.shape[0]
    for a in range(nb):
        hre[a, a] += scale * (next(draws) * 2.0 - 1.0)
        for b in range(a + 1, nb):
            dr = scale * (next(draws) * 2.0 - 1.0)
            di = scale * (next(draws) * 2.0 - 1.0)
            hre[a, b] += dr
            hre[b, a] += dr
            him[a, b] += di
            him[b, a] -= di
    return _pack_h(hre, him)


def refine_with_restarts(prob, h_block, n_restarts: int, refine, perturb_scale: float,
                         seed: int):
    """The restart loop of sigma_refine_analytic_with_restarts (quintic.rs:545-582) and
    sigma_refine_lbfgs_with_restarts_gpu (:1100-1139): refine(h) -> h after one run. Returns
    (best sigma_L1, history of sigma_L1 after each run, best h_block)."""
    nb = h_block.shape[0] // 2
    h = refine(np.array(h_block, dtype=np.float64))
    best = prob.sigma(h)
    best_h = h
    history = [best]
    _, u, _ = lcg_stream(seed, max(1, n_restarts * nb * nb))
    draws = iter(u.tolist())
    for _ in range(n_restarts):
        h = renorm_trace(perturb_h_hermitian(best_h, perturb_scale, draws))
        h = refine(h)
        trial = prob.sigma(h)
        history.append(trial)
        if np.isfinite(trial) and trial < best:
            best, best_h = trial, h
    return best, history, best_h
