"""Pins the oracle's per-point buffer (cpu_compute_per_point_reference, gpu_adam.rs:2677-2739).

Row p is (eta_p, w_p, d log eta_p / dh). The gradient part is pinned by central finite
differences of log eta_p (the oracle's eta, a separate code path) along random Hermitian
directions, and the whole buffer by recombining it into sigma_L2^2 and its gradient as the
docstring of sigma_squared_and_gradient states (quintic.rs:2697-2713) against the oracle's own
sigma_squared_and_gradient.
"""
import numpy as np


def _direction(rng, nb):
    X = rng.standard_normal((nb, nb))
    X = 0.5 * (X + X.T)
    Y = rng.standard_normal((nb, nb))
    Y = 0.5 * (Y - Y.T)
    D = np.zeros((2 * nb, 2 * nb))
    D[0::2, 0::2] = X
    D[0::2, 1::2] = -Y
    D[1::2, 0::2] = Y
    D[1::2, 1::2] = X
    return X, Y, D


def _dir_deriv(gre, gim, X, Y):
    nb = X.shape[0]
    iu, iu1 = np.triu_indices(nb, 0), np.triu_indices(nb, 1)
    return np.sum(gre[iu] * X[iu]) + np.sum(gim[iu1] * Y[iu1])


def test_per_point_gradient_is_d_log_eta(oracle):
    O = oracle
    prob = O.Problem(2, 300, 42)
    h, _ = O.donaldson_step(prob.sv, prob.w, prob.h)
    nb = prob.nb
    buf = O.per_point(prob.pts, prob.sv, prob.sd, prob.log_om, prob.w, h)
    assert buf.shape == (prob.n, 2 + 2 * nb * nb)
    rng = np.random.default_rng(3)
    eps = 1e-6
    errs = []
    for _ in range(4):
        X, Y, D = _direction(rng, nb)
        ep = O.eta(prob.pts, prob.sv, prob.sd, prob.log_om, h + eps * D)
        em = O.eta(prob.pts, prob.sv, prob.sd, prob.log_om, h - eps * D)
        fd = (np.log(ep) - np.log(em)) / (2 * eps)
        for p in range(0, prob.n, 7):
            if not np.isfinite(buf[p, 0]):
                continue
            g = buf[p, 2:]
            an = _dir_deriv(g[: nb * nb].reshape(nb, nb), g[nb * nb:].reshape(nb, nb), X, Y)
            errs.append(abs(fd[p] - an) / max(abs(an), 1e-3))
    assert len(errs) > 100
    assert np.median(errs) < 1e-7 and max(errs) < 1e-4


def test_per_point_recombines_into_sigma_sq_gradient(oracle):
    O = oracle
    prob = O.Problem(3, 800, 9)
    h, _ = O.donaldson_step(prob.sv, prob.w, prob.h)
    buf = O.per_point(prob.pts, prob.sv, prob.sd, prob.log_om, prob.w, h)
    eta, w, gl = buf[:, 0], buf[:, 1], buf[:, 2:]
    ok = np.isfinite(eta) & (eta > 0) & (w != 0)
    tw = np.sum(w[ok])
    kappa = np.sum(w[ok] * eta[ok]) / tw
    dkappa = (w[ok] * eta[ok]) @ gl[ok] / tw
    en = eta[ok] / kappa
    grad = (w[ok] * 2 * (en - 1) * en) @ (gl[ok] - dkappa / kappa) / tw
    l2 = np.sum(w[ok] * (en - 1) ** 2) / tw
    s2, g_ref, _ = O.sigma_sq_grad(prob.pts, prob.sv, prob.sd, prob.log_om, prob.w, h)
    assert abs(l2 / s2 - 1) < 1e-12
    assert np.max(np.abs(grad - g_ref)) < 1e-12 * np.max(np.abs(g_ref))


def test_per_point_invalid_rows(oracle):
    O = oracle
    prob = O.Problem(2, 50, 1)
    w = prob.w.copy()
    lo = prob.log_om.copy()
    w[3] = 0.0
    w[5] = np.nan
    lo[8] = np.nan
    buf = O.per_point(prob.pts, prob.sv, prob.sd, lo, w, prob.h)
    for p in (3, 5, 8):
        assert np.isnan(buf[p, 0]) and buf[p, 1] == 0.0 and not np.any(buf[p, 2:])
    assert np.all(np.isfinite(buf[[0, 1, 2, 4], 0]))
