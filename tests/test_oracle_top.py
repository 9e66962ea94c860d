"""The route34 T-operator contract (route34/cy3_donaldson_gpu.rs) in the oracle, pinned against
the quintic oracle's T (the C restatement of quintic.rs, itself pinned on the golden vectors)
through h_route34 = conj(G), against a literal loop restatement of cpu_t_operator_reference on
a tiny case, and against the closed-form scaling K(lambda h) = lambda K(h)."""
import numpy as np
import pytest


def _rand_herm_pd(rng, nb):
    a = rng.standard_normal((nb, nb)) + 1j * rng.standard_normal((nb, nb))
    return a @ a.conj().T + nb * np.eye(nb)


def test_route34_t_operator_is_conjugate_quintic_t(oracle):
    # T(G) of the quintic step, divided by sum w and trace-normalised, is the route34 h_pair
    # at h = conj(G), divided and normalised the same way (quintic.rs:1745-1812).
    P = oracle.Problem(3, 3000, 21)
    rng = np.random.default_rng(4)
    G = _rand_herm_pd(rng, P.nb) / P.nb
    hb = np.zeros((2 * P.nb, 2 * P.nb))
    hb[0::2, 0::2] = G.real
    hb[0::2, 1::2] = -G.imag
    hb[1::2, 0::2] = G.imag
    hb[1::2, 1::2] = G.real
    _, _, (tre, tim, _) = oracle.donaldson_step(P.sv, P.w, hb, want_t=True)
    pr, pi = oracle.t_operator_raw_route34(P.sv, P.w, G.real, -G.imag)
    t = (pr + 1j * pi) / np.sum(P.w[np.isfinite(P.w) & (P.w > 0)])
    t *= P.nb / np.trace(t).real
    assert np.abs(t - (tre + 1j * tim)).max() <= 1e-12 * np.abs(tre + 1j * tim).max()


def test_route34_t_operator_literal_loops(oracle):
    rng = np.random.default_rng(7)
    n, nb = 9, 3
    sv = rng.standard_normal((n, 2 * nb))
    w = rng.uniform(0.1, 2.0, n)
    w[2], w[5] = np.nan, -1.0  # skipped by the GPU kernel's weight test
    h = _rand_herm_pd(rng, nb)
    hr, hi = h.real, h.imag
    kv = np.zeros(n)
    for p in range(n):
        k = 0.0
        for a in range(nb):
            for b in range(nb):
                sar, sai, sbr, sbi = sv[p, 2 * a], sv[p, 2 * a + 1], sv[p, 2 * b], sv[p, 2 * b + 1]
                k += hr[a, b] * (sar * sbr + sai * sbi) + hi[a, b] * (sar * sbi - sai * sbr)
        kv[p] = k if k > 1e-30 else 1e-30
    ref_re, ref_im = np.zeros((nb, nb)), np.zeros((nb, nb))
    for a in range(nb):
        for b in range(nb):
            for p in range(n):
                if not (np.isfinite(w[p]) and w[p] > 0):
                    continue
                sar, sai, sbr, sbi = sv[p, 2 * a], sv[p, 2 * a + 1], sv[p, 2 * b], sv[p, 2 * b + 1]
                f = w[p] / kv[p]
                ref_re[a, b] += f * (sar * sbr + sai * sbi)
                ref_im[a, b] += f * (sar * sbi - sai * sbr)
    pr, pi = oracle.t_operator_raw_route34(sv, w, hr, hi)
    assert np.allclose(pr, ref_re, rtol=1e-13, atol=0)
    assert np.allclose(pi, ref_im, rtol=1e-13, atol=1e-15)


def test_route34_t_operator_scaling_and_hermitian(oracle):
    rng = np.random.default_rng(11)
    sv = rng.standard_normal((400, 10))
    w = rng.uniform(0.1, 1.0, 400)
    h = _rand_herm_pd(rng, 5)
    r1, i1 = oracle.t_operator_raw_route34(sv, w, h.real, h.imag)
    r2, i2 = oracle.t_operator_raw_route34(sv, w, 4.0 * h.real, 4.0 * h.imag)
    assert np.allclose(r2, r1 / 4.0, rtol=1e-14, atol=0)
    assert np.allclose(i2, i1 / 4.0, rtol=1e-14, atol=1e-16)
    assert np.allclose(r1, r1.T, rtol=1e-14, atol=0)
    assert np.allclose(i1, -i1.T, rtol=1e-14, atol=1e-16)


def _fs_case(rng, n):
    # Linear sections s_a = z_a (N = 8, ds_i = e_i), H = I and frames T orthonormal with
    # T_a . z = 0: g_amb is the Fubini-Study metric delta/|z|^2 - conj(z_i) z_j / |z|^4, its
    # restriction to the frame is I / |z|^2, so eta = 1 / (|z|^6 |Omega|^2) exactly.
    z = rng.standard_normal((n, 8)) + 1j * rng.standard_normal((n, 8))
    sv = np.zeros((n, 16))
    sv[:, 0::2], sv[:, 1::2] = z.real, z.imag
    sd = np.zeros((n, 8, 16))
    for i in range(8):
        sd[:, i, 2 * i] = 1.0
    frames = np.zeros((n, 3, 8, 2))
    for p in range(n):
        # orthonormal complement of conj(z) (Hermitian), so that T_a . z = <conj(z), T_a> = 0
        q, _ = np.linalg.qr(np.column_stack([z[p].conj(), rng.standard_normal((8, 7))
                                             + 1j * rng.standard_normal((8, 7))]))
        t = q[:, 1:4].T
        frames[p, ..., 0], frames[p, ..., 1] = t.real, t.imag
    om = rng.uniform(0.5, 2.0, n)
    w = rng.uniform(0.1, 1.0, n)
    return sv, sd, frames, om, w, z


def test_route34_sigma_fubini_study_closed_form(oracle):
    rng = np.random.default_rng(5)
    sv, sd, frames, om, w, z = _fs_case(rng, 60)
    sig, eta = oracle.sigma_route34(sv, sd, frames, om, w, np.eye(8), np.zeros((8, 8)))
    k = np.sum(np.abs(z) ** 2, axis=1)
    eta_ref = 1.0 / (k ** 3 * om)
    assert np.allclose(eta, eta_ref, rtol=1e-11, atol=0)
    kap = np.sum(w * eta_ref) / np.sum(w)
    assert abs(sig - np.sum(w * (eta_ref / kap - 1) ** 2) / np.sum(w)) < 1e-11 * sig
    # H -> c H scales K, dk and M by c, which leaves g_amb, eta and sigma unchanged.
    sig2, eta2 = oracle.sigma_route34(sv, sd, frames, om, w, 3.0 * np.eye(8), np.zeros((8, 8)))
    assert np.allclose(eta2, eta_ref, rtol=1e-11, atol=0)
    assert abs(sig2 - sig) < 1e-12 * sig
