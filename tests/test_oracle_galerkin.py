"""Pins of the oracle's Laplacian-Galerkin and Yukawa restatements (CPU only).

The reference's own tests for these modules need its TY/Schoen solvers (metric_laplacian.rs
tests are #[ignore]d, yukawa_overlap_real.rs tests build HYM/harmonic-mode pipelines), so the
restatement is pinned here by what the mathematics fixes: direct numpy evaluation of the
defining sums, library eigen/inverse routines, the kernel of the Laplacian (constants),
independence of the tangent frame, finite differences of the test functions, and the
symmetries of the triple overlap.
"""
import itertools
import math

import numpy as np
import pytest

from oracle import oracle as O


def _rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


# ------------------------------------------------------------------------------------------
# Ambient (Bochner) Laplacian: metric_laplacian.rs
# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("deg", [1, 2, 3, 4])
def test_lap_basis_closed_form(deg):
    ex = O.lap_basis(deg)
    c = lambda d: math.comb(d + 3, 3)
    expect = sum(c(dz) * c(d - dz) for d in range(1, deg + 1) for dz in range(d + 1))
    assert ex.shape == (expect, 8)
    # Same set as a brute-force enumeration, no constant, no duplicates.
    brute = {e for e in itertools.product(range(deg + 1), repeat=8) if 1 <= sum(e) <= deg}
    got = [tuple(int(v) for v in r) for r in ex]
    assert set(got) == brute and len(set(got)) == len(got)
    # Order: (dz, dw) loops, then a, b, c ascending inside each 4-block (monomials_4).
    keys = [(sum(r[:4]), sum(r[4:]), tuple(r[:3]), tuple(r[4:7])) for r in got]
    assert keys == sorted(keys)
    assert got[0] == (0, 0, 0, 0, 0, 0, 0, 1)


def _ambient_inputs(rng, n):
    pts = _rand_c(rng, n, 8) * 0.5
    w = rng.uniform(0.2, 2.0, n)
    om = _rand_c(rng, n)
    w[3] = 0.0
    w[7] = np.nan
    om[11] = np.nan
    return pts, w, om


def test_lap_ambient_mk_matches_direct_sums():
    rng = np.random.default_rng(5)
    pts, w, om = _ambient_inputs(rng, 300)
    ex = O.lap_basis(2).astype(np.int64)
    M, K = O.lap_ambient_mk(pts, w, om, ex)
    ok = np.isfinite(w) & (w > 0) & np.isfinite(om)
    pw = np.where(ok, np.abs(np.nan_to_num(om)) ** 2 * np.nan_to_num(w), 0.0)
    phi = np.prod(pts[:, None, :] ** ex[None, :, :], axis=2)
    dphi = np.zeros(phi.shape + (8,), complex)
    for i in range(8):
        e2 = ex.copy()
        e2[:, i] = np.maximum(e2[:, i] - 1, 0)
        dphi[:, :, i] = ex[None, :, i] * np.prod(pts[:, None, :] ** e2[None], axis=2)
    tw = np.nansum(np.where(ok, w, 0.0))
    Mn = np.einsum("p,pa,pb->ab", pw, phi.conj(), phi) / tw
    Kn = np.einsum("p,pai,pbi->ab", pw, dphi.conj(), dphi) / tw
    np.fill_diagonal(Mn, Mn.diagonal().real)
    np.fill_diagonal(Kn, np.maximum(Kn.diagonal().real, 0.0))
    assert np.abs(M - Mn).max() <= 1e-12 * np.abs(Mn).max()
    assert np.abs(K - Kn).max() <= 1e-12 * np.abs(Kn).max()
    assert np.abs(M - M.conj().T).max() == 0.0 and np.abs(K - K.conj().T).max() == 0.0


def test_herm_inv_reg_matches_library_inverse():
    rng = np.random.default_rng(1)
    a = _rand_c(rng, 24, 24)
    h = a @ a.conj().T + 0.5 * np.eye(24)
    eps = 1e-3
    lam = eps * np.linalg.norm(h)
    inv = O.herm_inv_reg(h, eps)
    ref = np.linalg.inv(h + lam * np.eye(24))
    assert np.abs(inv - ref).max() <= 1e-11 * np.abs(ref).max()


def test_herm_jacobi_matches_eigh():
    rng = np.random.default_rng(2)
    a = _rand_c(rng, 30, 30)
    h = (a + a.conj().T) / 2
    ev, vec, sw = O.herm_jacobi(h, vectors=True)
    ref = np.linalg.eigvalsh(h)
    assert np.abs(ev - ref).max() <= 1e-11 * np.abs(ref).max()
    assert np.abs(h @ vec - vec * ev[None, :]).max() <= 1e-10 * np.abs(ref).max()
    assert np.abs(vec.conj().T @ vec - np.eye(30)).max() <= 1e-12
    # Entries just under the per-pair threshold are never rotated, so the absolute stopping
    # test (sqrt(off) < tol) is seldom met and the reference runs all max_sweeps.
    assert sw >= 1


def test_lap_spectrum_pipeline_matches_numpy():
    rng = np.random.default_rng(3)
    pts, w, om = _ambient_inputs(rng, 400)
    low, full, M, K = O.lap_spectrum_ambient(pts, w, om, max_deg=2, n_low=10)
    lam = 1e-10 * np.linalg.norm(M)
    L = np.linalg.inv(M + lam * np.eye(M.shape[0])) @ K
    Lh = (L + L.conj().T) / 2
    ref = np.sort(np.abs(np.linalg.eigvalsh(Lh)))
    assert np.abs(full - ref).max() <= 1e-8 * ref.max()
    assert np.array_equal(low, full[:10])


# ------------------------------------------------------------------------------------------
# Intrinsic Laplacian on the quintic
# ------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def quintic_small():
    prob = O.Problem(2, 300, 42, fs_gram=True)
    h, _ = O.donaldson_solve(prob.sv, prob.w, prob.h, 3, 0.0)
    return prob, h


def _multinomial(e):
    return math.factorial(int(sum(e))) / np.prod([math.factorial(int(v)) for v in e])


@pytest.mark.parametrize("sdeg", [1, 2])
def test_lap_quintic_constant_is_in_kernel(quintic_small, sdeg):
    prob, h = quintic_small
    M, K, tw = O.lap_quintic_mk(prob, h, sdeg)
    mono = O.monomials(sdeg, False)
    ns = mono.shape[0]
    v = np.zeros(ns * ns, complex)
    for a in range(ns):
        v[a * ns + a] = _multinomial(mono[a])  # sum_a mult(a) |z^a|^2 = |z|^(2d)
    assert np.linalg.norm(K @ v) <= 1e-11 * np.abs(K).max() * np.linalg.norm(v)
    ok = np.isfinite(prob.w) & (prob.w > 0) & np.isfinite(prob.log_om)
    expect = np.sum(prob.w[ok] * np.exp(prob.log_om[ok])) / np.sum(prob.w[ok])
    assert abs((v.conj() @ M @ v).real - expect) <= 1e-12 * expect
    assert abs(tw - np.sum(prob.w[ok])) <= 1e-12 * tw
    # Hermitian positive semi-definite pencil, M definite
    assert np.abs(K - K.conj().T).max() <= 1e-13 * np.abs(K).max()
    assert np.linalg.eigvalsh((K + K.conj().T) / 2).min() >= -1e-12 * np.abs(K).max()
    assert np.linalg.eigvalsh((M + M.conj().T) / 2).min() > 0


def test_lap_quintic_matches_fd_in_another_frame(quintic_small):
    """K recomputed with finite-difference d/dconj(z) of phi and a random other basis of the
    tangent space (T' = A T + b z): pins the derivative (with the |z|^-2d term), the frame
    independence and the orientation of the inverse metric."""
    prob, h = quintic_small
    sdeg = 1
    M, K, tw = O.lap_quintic_mk(prob, h, sdeg)
    ga, fr, ok = O.point_metric(prob, h)
    mono = O.monomials(sdeg, False).astype(np.int64)
    ns = mono.shape[0]
    z = prob.pts[:, 0::2] + 1j * prob.pts[:, 1::2]
    keep = ok & np.isfinite(prob.w) & (prob.w > 0) & np.isfinite(prob.log_om)
    rng = np.random.default_rng(9)

    def phi_all(zz):
        s = np.prod(zz[:, None, :] ** mono[None], axis=2)
        r2 = np.sum(np.abs(zz) ** 2, axis=1)
        return (s[:, :, None] * s.conj()[:, None, :]).reshape(len(zz), -1) / r2[:, None] ** sdeg

    step = 1e-6
    F = np.zeros((prob.n, ns * ns, 5), complex)
    for i in range(5):
        e = np.zeros(5)
        e[i] = 1.0
        dx = (phi_all(z + step * e) - phi_all(z - step * e)) / (2 * step)
        dy = (phi_all(z + 1j * step * e) - phi_all(z - 1j * step * e)) / (2 * step)
        F[:, :, i] = 0.5 * (dx + 1j * dy)
    Kn = np.zeros_like(K)
    Mn = np.zeros_like(M)
    ph = phi_all(z)
    for p in np.nonzero(keep)[0]:
        A = _rand_c(rng, 3, 3)
        b = _rand_c(rng, 3)
        T = A @ fr[p] + b[:, None] * z[p][None, :]
        gt = T @ ga[p] @ T.conj().T
        Y = T.conj().T @ np.linalg.inv(gt) @ T
        c = prob.w[p] * np.exp(prob.log_om[p])
        Kn += c * np.einsum("aj,ij,bi->ab", F[p].conj(), Y, F[p])
        Mn += c * np.outer(ph[p].conj(), ph[p])
    Kn /= tw
    Mn /= tw
    assert np.abs(Mn - M).max() <= 1e-12 * np.abs(M).max()
    assert np.abs(Kn - K).max() <= 1e-6 * np.abs(K).max()


# ------------------------------------------------------------------------------------------
# Yukawa: yukawa_overlap_real.rs
# ------------------------------------------------------------------------------------------
def _yuk_inputs(rng, nm=5, n=200):
    modes = _rand_c(rng, nm, n)
    om = _rand_c(rng, n)
    w = rng.uniform(0.1, 1.0, n)
    return modes, om, w


def test_yuk_triple_matches_einsum_and_skips():
    rng = np.random.default_rng(4)
    modes, om, w = _yuk_inputs(rng)
    hd = rng.uniform(0.5, 2.0, modes.shape[0])
    w[3] = np.nan
    om[8] = np.nan
    modes[2, 5] = np.nan
    y = O.yuk_triple(modes, om, w, hd)
    nm, n = modes.shape
    ref = np.zeros((nm, nm, nm), complex)
    for i, j, k in itertools.product(range(nm), repeat=3):
        m = np.isfinite(w) & np.isfinite(om) & np.isfinite(modes[i]) & np.isfinite(
            modes[j]) & np.isfinite(modes[k])
        ref[i, j, k] = np.sum((w * om.conj() * modes[i] * modes[j] * modes[k])[m])
        ref[i, j, k] *= np.sqrt(hd[i] * hd[j] * hd[k])
    assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max()
    # symmetric tensor up to rounding of the product order
    for perm in itertools.permutations(range(3)):
        assert np.abs(y - y.transpose(perm)).max() <= 1e-14 * np.abs(y).max()


def test_yuk_sz_weights_and_score():
    rng = np.random.default_rng(6)
    _, om, w = _yuk_inputs(rng, n=100)
    w[4] = -1.0
    om[9] = 0.0
    sz, score = O.yuk_sz_weights(w, om)
    ok = (w > 0) & (np.abs(om) > 0)
    ref = np.where(ok, w / np.where(ok, np.abs(om) ** 2, 1.0), 0.0)
    ref /= ref.sum()
    assert np.abs(sz - ref).max() <= 1e-15
    cv = ref.std() / ref.mean()
    assert abs(score - 1.0 / (1.0 + cv)) <= 1e-12


def test_yuk_bootstrap_and_convergence():
    rng = np.random.default_rng(7)
    modes, om, w = _yuk_inputs(rng, nm=3, n=64)
    y, unc, score, conv = O.yuk_couplings(modes, om, w, None, n_bootstrap=4, seed=11)
    sz, _ = O.yuk_sz_weights(w, om)
    assert np.abs(y - O.yuk_triple(modes, om, sz)).max() == 0.0
    # the resample path with the identity resample is the central value
    assert np.abs(O.yuk_triple(modes, om, sz, idx=np.arange(64)) - y).max() <= 1e-15
    dev = np.zeros(y.shape)
    for b in range(4):
        idx = O.yuk_bootstrap_indices(11, b, 64)
        assert idx.min() >= 0 and idx.max() < 64
        yb = np.einsum("p,pi,pj,pk->ijk", (sz * om.conj())[idx], modes.T[idx], modes.T[idx],
                       modes.T[idx])
        dev += np.abs(yb - y) ** 2
    assert np.abs(np.sqrt(dev / 4) - unc).max() <= 1e-12 * np.abs(y).max()
    hw = sz[:32] / sz[:32].sum()
    yh = np.einsum("p,pi,pj,pk->ijk", hw * om[:32].conj(), modes.T[:32], modes.T[:32],
                   modes.T[:32])
    assert abs(conv - np.abs(y - yh).max() / np.abs(y).max()) <= 1e-12
