"""Pins the oracle's sigma on a fresh sample (sigma_at_metric_closure, quintic.rs:1520-1616).

The Bergman metric of G = diag(multinomial(e)) over the raw degree-k monomials is
k * g_FS, since sum_e multinomial(e) |z^e|^2 = |z|^(2k). The closure with the analytic
Fubini-Study metric (fs_metric_analytic, nn_metric.rs:309-335) is restated here in numpy, with
the reference's chart, frame, projection and determinant (quintic.rs:4541-4584, 4739-4809,
4870-4916), so eta of that G must equal k^3 times the numpy eta point by point, and sigma must
agree. In the FS-Gram basis s' = conj(L)^-1 s the same metric is G' = L^T D conj(L), which pins
the basis change carried to the fresh points.
"""
from math import factorial

import numpy as np
import pytest


def _fs_closure_eta(pts, psi=0.0):
    """eta of the analytic Fubini-Study metric at each point, as sigma_at_metric_closure."""
    z = pts[:, 0::2] + 1j * pts[:, 1::2]
    n = z.shape[0]
    prod = np.prod(z, axis=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        grad = 5 * z ** 4 - 5 * psi * prod[:, None] / z
    out = np.full(n, np.nan)
    for p in range(n):
        zp, gp = z[p], grad[p]
        zsq = np.abs(zp) ** 2
        chart, mz = 0, 0.0
        for k in range(5):
            if zsq[k] > mz:
                mz, chart = zsq[k], k
        gsq = np.abs(gp) ** 2
        elim, mg = (1 if chart == 0 else 0), 0.0
        for k in range(5):
            if k != chart and gsq[k] > mg:
                mg, elim = gsq[k], k
        if mg < 1e-20 or mz < 1e-20:
            continue
        log_om = np.log(mz) - np.log(mg)
        nz = np.sum(zsq)
        g = np.eye(5) / nz - np.outer(np.conj(zp), zp) / nz ** 2
        free = [k for k in range(5) if k not in (chart, elim)]
        T = np.zeros((3, 5), complex)
        for i, fi in enumerate(free):
            T[i, fi] = 1.0
            T[i, elim] = -gp[fi] / gp[elim]
        gt = T @ g @ np.conj(T).T  # g_tan[a][b] = sum T_a[i] g_ij conj(T_b[j])
        det = np.real(np.linalg.det(gt))
        if not np.isfinite(det) or abs(det) < 1e-30:
            continue
        out[p] = np.exp(np.log(abs(det)) - log_om)
    return out


def _sigma_l1(eta, w):
    ok = np.isfinite(eta) & (eta > 0) & np.isfinite(w) & (w > 0)
    kappa = np.sum(w[ok] * eta[ok]) / np.sum(w[ok])
    return np.sum(w[ok] * np.abs(eta[ok] / kappa - 1)) / np.sum(w[ok]), kappa


def _multinom_diag(mono):
    return np.array([factorial(int(sum(e))) / np.prod([factorial(int(x)) for x in e])
                     for e in mono])


def test_fresh_sample_with_the_problem_seed_is_the_in_sample_sigma(oracle):
    O = oracle
    prob = O.Problem(3, 2500, 42)
    h, _ = O.donaldson_step(prob.sv, prob.w, prob.h)
    s1, kap, l2, n = O.sigma_at_sample(prob, h, 2500, 42)
    assert n == prob.n
    e = O.eta(prob.pts, prob.sv, prob.sd, prob.log_om, h)
    ref = O.sigma_reduce(e, prob.w)
    assert (s1, kap, l2) == ref


@pytest.mark.parametrize("k", [2, 3])
def test_fresh_sample_fubini_study_closed_form_raw_basis(oracle, k):
    O = oracle
    prob = O.Problem(k, 500, 42, fs_gram=False)
    d = _multinom_diag(prob.mono)
    h = O._pack_h(np.diag(d), np.zeros((prob.nb, prob.nb)))
    s1, kap, l2, n, eta, pts = O.sigma_at_sample(prob, h, 1500, 7, want_eta=True)
    assert n == 1500
    eta_fs = _fs_closure_eta(pts)
    np.testing.assert_allclose(eta, k ** 3 * eta_fs, rtol=1e-10)
    w, _, _ = O.weights_logomega(pts)
    s_fs, kap_fs = _sigma_l1(eta_fs, w)
    assert abs(s1 - s_fs) < 1e-11
    assert abs(kap / kap_fs - k ** 3) < 1e-10


def test_fresh_sample_fubini_study_through_the_fs_gram_basis(oracle):
    O = oracle
    k = 3
    prob = O.Problem(k, 3000, 42)
    lre, lim = prob.L
    L = lre + 1j * lim
    D = np.diag(_multinom_diag(prob.mono))
    G = L.T @ D @ np.conj(L)  # s_raw^H D s_raw with s_raw = conj(L) s'
    h = O._pack_h(G.real, G.imag)
    s1, kap, l2, n, eta, pts = O.sigma_at_sample(prob, h, 1200, 99, want_eta=True)
    eta_fs = _fs_closure_eta(pts)
    np.testing.assert_allclose(eta, k ** 3 * eta_fs, rtol=1e-8)
    w, _, _ = O.weights_logomega(pts)
    assert abs(s1 - _sigma_l1(eta_fs, w)[0]) < 1e-9
    # A transposed or unconjugated basis change is not the FS metric.
    Gbad = L.conj().T @ D @ L
    hb = O._pack_h(Gbad.real, Gbad.imag)
    eb = O.sigma_at_sample(prob, hb, 1200, 99, want_eta=True)[4]
    assert np.nanmax(np.abs(eb / (k ** 3 * eta_fs) - 1)) > 1e-3


def test_fresh_sample_newton_sampler_and_dwork(oracle):
    O = oracle
    prob = O.Problem(2, 400, 42, psi=0.3, fs_gram=False)
    d = _multinom_diag(prob.mono)
    h = O._pack_h(np.diag(d), np.zeros((prob.nb, prob.nb)))
    for sampler in ("sz", "newton"):
        s1, kap, l2, n, eta, pts = O.sigma_at_sample(prob, h, 300, 5, sampler, want_eta=True)
        assert n > 0
        f, _ = O.quintic_eval(pts, 0.3)
        assert np.max(np.abs(f[:, 0] + 1j * f[:, 1])) < 1e-2
        np.testing.assert_allclose(eta, 8 * _fs_closure_eta(pts, 0.3), rtol=1e-9)
