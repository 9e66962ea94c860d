"""Pins of the CPU oracle against the reference's own published values, tests and closed forms.

The oracle (oracle/cy3q_oracle.c) is the checker for every GPU parity test, so each part of it
is pinned here by something independent of its own formulas:
  * published sigma values of the reference's p5_9 run (output/p5_9_ako_apples_to_apples.json);
  * the reference's unit tests (quintic.rs:5467-5555);
  * closed forms, invariants and brute force on tiny inputs.
"""
import itertools
import math

import numpy as np
import pytest


# --------------------------------------------------------------------------------------------
# LCG (cy3/src/lib.rs:222-245)
# --------------------------------------------------------------------------------------------
def test_lcg_integer_recurrence(oracle):
    states, f64s, normals = oracle.lcg_stream(42, 1000)
    M = 2 ** 64
    s = (42 * 6364136223846793005 + 1) % M
    for i in range(1000):
        s = (s * 6364136223846793005 + 1442695040888963407) % M
        assert int(states[i]) == s
        assert f64s[i] == ((s >> 11) & ((1 << 53) - 1)) / float(1 << 53)
    # Box-Muller from the same stream: u1, u2 pairs.
    u = f64s
    for i in range(0, 20, 2):
        u1 = max(u[i] + 1e-12, 1e-12)
        z = math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u[i + 1])
        assert normals[i // 2] == z


# --------------------------------------------------------------------------------------------
# Index tables (quintic.rs:3185-3209, 5472-5484)
# --------------------------------------------------------------------------------------------
def test_monomial_counts_and_order(oracle):
    assert oracle.n_sections(2) == 15 and oracle.n_sections(3) == 35 and oracle.n_sections(4) == 70
    for k in range(1, 9):
        m = oracle.monomials(k)
        assert len(m) == math.comb(k + 4, 4)
        assert np.all(m.sum(1) == k)
        # Brute force of the reference's nested loop order: lexicographic in (e0, e1, e2, e3).
        brute = [e for e in itertools.product(range(k + 1), repeat=4) if sum(e) <= k]
        brute = [list(e) + [k - sum(e)] for e in sorted(brute)]
        assert np.array_equal(m, np.array(brute, np.uint32))
        q = oracle.monomials(k, True)
        assert len(q) == math.comb(k + 4, 4) - (math.comb(k - 1, 4) if k >= 5 else 0)
        assert np.array_equal(q, m[m[:, 0] < 5])
    assert len(oracle.monomials(6, True)) == 205
    assert len(oracle.monomials(5, True)) == 125
    assert len(oracle.monomials(8, True)) == 460


def test_library_index_tables_bit_exact(oracle, cy3q):
    for k in range(1, 10):
        for q in (False, True):
            assert np.array_equal(cy3q.build_degree_k_quintic_monomials(k, q), oracle.monomials(k, q))


# --------------------------------------------------------------------------------------------
# Quintic, sampler (quintic.rs:3213-3323, 3633-3983, 5487-5494)
# --------------------------------------------------------------------------------------------
def test_known_root(oracle):
    z = np.array([[1.0, 0, 0, 0, 0, 0, 0, 0, -1.0, 0]])
    f, g = oracle.quintic_eval(z)
    assert abs(f[0, 0]) < 1e-12 and abs(f[0, 1]) < 1e-12
    assert np.allclose(g[0], [5, 0, 0, 0, 0, 0, 0, 0, 5, 0])


def test_sz_points_on_variety_and_normalised(oracle):
    pts, tried = oracle.sample_sz(5000, 42)
    assert pts.shape == (5000, 10) and tried == 1000
    assert np.allclose(np.sum(pts ** 2, 1), 1.0, atol=1e-14)
    f, _ = oracle.quintic_eval(pts)
    assert np.max(np.hypot(f[:, 0], f[:, 1])) < 1e-3


def test_sz_roots_match_companion_eigenvalues(oracle):
    # Each line's 5 candidates are a + t_k b for the 5 roots t_k of f(a + t b). Rebuild a, b from
    # the LCG stream and compare t_k with numpy's companion-matrix roots (brute force).
    zs, acc = oracle.sz_lines(30, 123)
    _, _, normals = oracle.lcg_stream(123, 20 * 30)
    for l in range(30):
        nrm = normals[20 * l: 20 * l + 20]
        a = nrm[0::4] + 1j * nrm[1::4]
        b = nrm[2::4] + 1j * nrm[3::4]
        coef = [sum(math.comb(5, j) * a[i] ** (5 - j) * b[i] ** j for i in range(5)) for j in range(6)]
        roots = np.roots(coef[::-1])
        for k in range(5):
            if not acc[l, k]:
                continue
            z = zs[l, k, 0::2] + 1j * zs[l, k, 1::2]
            # z is (a + t b)/|a + t b|: recover t from the projective ratio with the best pair.
            best = min(abs((z * np.linalg.norm(a + r * b) - (a + r * b))).max() for r in roots)
            assert best < 1e-9


def test_dwork_at_psi0_equals_fermat_points(oracle):
    # quintic.rs:6048-6073: psi = 0 Dwork == Fermat (same points; only the acceptance bound
    # differs, 1e-2 vs 1e-3, and no Fermat point sits between the two bounds here).
    p0, _ = oracle.sample_sz(3000, 5, psi=0.0, dwork=False)
    p1, _ = oracle.sample_sz(3000, 5, psi=0.0, dwork=True)
    assert np.array_equal(p0, p1)


def test_weight_is_inverse_grad_norm_and_euler(oracle):
    pts, _ = oracle.sample_sz(2000, 8, psi=0.3, dwork=True)
    w, lo, ce = oracle.weights_logomega(pts, 0.3)
    f, g = oracle.quintic_eval(pts, 0.3)
    gc = g[:, 0::2] + 1j * g[:, 1::2]
    z = pts[:, 0::2] + 1j * pts[:, 1::2]
    np.testing.assert_allclose(w, 1.0 / np.sum(np.abs(gc) ** 2, 1), rtol=1e-14)
    # Euler: z . grad f = 5 f = 0 on the variety.
    assert np.max(np.abs(np.sum(z * gc, 1))) < 5e-2 * 1.0
    # log|Omega|^2 = ln|z_chart|^2 - ln|df_elim|^2 with the chart/elim argmax rules.
    chart = np.argmax(np.abs(z) ** 2, 1)
    assert np.array_equal(chart, ce[:, 0])
    np.testing.assert_allclose(lo, np.log(np.abs(z[np.arange(len(z)), ce[:, 0]]) ** 2)
                               - np.log(np.abs(gc[np.arange(len(z)), ce[:, 1]]) ** 2), rtol=1e-13, atol=1e-14)


# --------------------------------------------------------------------------------------------
# Sections and derivatives (quintic.rs:4021-4071, 1231-1296)
# --------------------------------------------------------------------------------------------
def test_sections_closed_form_and_complex_step_derivative(oracle):
    pts, _ = oracle.sample_sz(50, 4)
    mono = oracle.monomials(4)
    sv, sd = oracle.sections(pts, mono)
    z = pts[:, 0::2] + 1j * pts[:, 1::2]
    s = np.prod(z[:, None, :] ** mono[None, :, :].astype(int), axis=2)
    np.testing.assert_allclose(sv[:, 0::2] + 1j * sv[:, 1::2], s, rtol=1e-13, atol=1e-16)
    # Holomorphic derivative by a complex finite difference along each coordinate.
    h = 1e-6
    for i in range(5):
        zp = z.copy()
        zp[:, i] += h
        zm = z.copy()
        zm[:, i] -= h
        fd = (np.prod(zp[:, None, :] ** mono.astype(int), 2) - np.prod(zm[:, None, :] ** mono.astype(int), 2)) / (2 * h)
        np.testing.assert_allclose(sd[:, i, 0::2] + 1j * sd[:, i, 1::2], fd, rtol=1e-7, atol=1e-10)


# --------------------------------------------------------------------------------------------
# FS-Gram, T-step, inversion (quintic.rs:2957-3165, 1734-1894)
# --------------------------------------------------------------------------------------------
def test_fs_gram_orthonormalises(oracle):
    # quintic.rs:5721-5755: after FS-Gram the weighted Gram is the identity within 1e-6.
    P = oracle.Problem(3, 20000, 42)
    gre, gim = oracle.fs_gram(P.sv, P.w)
    assert np.max(np.abs(gre - np.eye(P.nb))) < 1e-10 and np.max(np.abs(gim)) < 1e-10


def test_lu_invert_is_inverse(oracle):
    rng = np.random.default_rng(0)
    a = rng.standard_normal((40, 40)) + 40 * np.eye(40)
    np.testing.assert_allclose(oracle.lu_invert(a) @ a, np.eye(40), atol=1e-13)
    with pytest.raises(ValueError):
        oracle.lu_invert(np.zeros((4, 4)))


def test_t_step_invariants_and_brute_force(oracle):
    P = oracle.Problem(2, 3000, 9)
    h1, r, (tre, tim, kinv) = oracle.donaldson_step(P.sv, P.w, P.h, want_t=True)
    S = P.sv[:, 0::2] + 1j * P.sv[:, 1::2]
    # Brute force: K = s^H G s with G = I, T = sum w/K conj(s) s^T / sum w, trace -> N.
    K = np.sum(np.abs(S) ** 2, 1)
    T = (S.conj() * (P.w / K)[:, None]).T @ S / P.w.sum()
    T *= P.nb / np.trace(T).real
    np.testing.assert_allclose(tre + 1j * tim, T, rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(kinv, 1 / K, rtol=1e-13)
    G = h1[0::2, 0::2] + 1j * h1[1::2, 0::2]
    np.testing.assert_allclose(G @ T, np.eye(P.nb), atol=1e-10)
    assert np.allclose(G, G.conj().T, atol=1e-14)
    assert abs(r - np.linalg.norm(h1 - P.h)) < 1e-13


def test_donaldson_residual_decreases(oracle):
    # quintic.rs:5536-5551 (Donaldson converges).
    P = oracle.Problem(2, 5000, 17)
    _, res = oracle.donaldson_solve(P.sv, P.w, P.h, 20, 1e-6)
    assert res[-1] < res[0]


def test_eta_chart_invariance(oracle):
    # quintic.rs:4894-4909: eta is the same in every (chart, elim) frame at a point. Rescaling a
    # point by a phase keeps the projective point; eta must not change.
    P = oracle.Problem(3, 400, 2)
    eta0 = oracle.eta(P.pts, P.sv, P.sd, P.log_om, P.h)
    ph = np.exp(1j * 0.7)
    z = (P.pts[:, 0::2] + 1j * P.pts[:, 1::2]) * ph
    pts2 = np.empty_like(P.pts)
    pts2[:, 0::2], pts2[:, 1::2] = z.real, z.imag
    sv2, sd2 = oracle.sections(pts2, P.mono)
    sv2, sd2, _ = oracle.orthonormalise(sv2, sd2, P.w)
    _, lo2, _ = oracle.weights_logomega(pts2)
    eta1 = oracle.eta(pts2, sv2, sd2, lo2, P.h)
    np.testing.assert_allclose(eta1, eta0, rtol=1e-9)


# --------------------------------------------------------------------------------------------
# End-to-end pins: the reference's published p5_9 per-seed values
# --------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def k3_seed42(oracle):
    return oracle.Problem(3, 200000, 42)


def test_published_sigma_fs_k3(k3_seed42):
    # output/p5_9_ako_apples_to_apples.json: k=3, 200k, seed 42, sigma_fs_identity.
    assert abs(k3_seed42.sigma() / 0.19735711129744954 - 1) < 1e-13


@pytest.mark.slow
def test_published_sigma_donaldson_k3(oracle, k3_seed42):
    # Same record: donaldson_solve(50, 1e-12) runs 28 iterations; sigma_donaldson.
    P = k3_seed42
    h, res = oracle.donaldson_solve(P.sv, P.w, P.h, 50, 1e-12)
    assert len(res) == 28
    assert abs(P.sigma(h) / 0.19332013761052813 - 1) < 1e-13


def test_published_sigma_fs_k4_seed100(oracle):
    # Same file, k=4 record, seed 100: sigma_fs_identity = 0.1405790096262153.
    P = oracle.Problem(4, 200000, 100)
    assert abs(P.sigma() / 0.1405790096262153 - 1) < 1e-12


# ------------------------------------------------------------------------------------------
# sigma_L2^2 analytic gradient and Adam (quintic.rs:2194-2860, 453-514)
# ------------------------------------------------------------------------------------------
def _herm_block(X, Y):
    nb = X.shape[0]
    G = X + 1j * Y
    H = np.zeros((2 * nb, 2 * nb))
    H[0::2, 0::2] = G.real
    H[0::2, 1::2] = -G.imag
    H[1::2, 0::2] = G.imag
    H[1::2, 1::2] = G.real
    return H


def test_sigma_sq_grad_matches_finite_differences(oracle):
    # The reference checks its analytic gradient against central differences
    # (quintic.rs:5646-5696); the same check pins the restatement: along random Hermitian
    # directions (X symmetric, Y antisymmetric) the derivative is
    # sum_{a<=b} g_re(a,b) X_ab + sum_{a<b} g_im(a,b) Y_ab.
    P = oracle.Problem(3, 2500, 7)
    nb = P.nb
    h, _ = oracle.donaldson_step(P.sv, P.w, P.h)
    s2, g, s1 = P.sigma_sq_grad(h)
    assert abs(s2 - P.sigma_l2sq(h)) < 1e-13 * s2
    assert abs(s1 - P.sigma(h)) < 1e-13 * s1
    gre, gim = g[: nb * nb].reshape(nb, nb), g[nb * nb:].reshape(nb, nb)
    assert np.allclose(np.tril(gim), 0.0)
    assert np.allclose(gre, gre.T)
    rng = np.random.default_rng(3)
    iu, iu1 = np.triu_indices(nb, 0), np.triu_indices(nb, 1)
    for _ in range(3):
        X = rng.standard_normal((nb, nb))
        X = 0.5 * (X + X.T)
        Y = rng.standard_normal((nb, nb))
        Y = 0.5 * (Y - Y.T)
        eps = 1e-6
        D = _herm_block(X, Y)
        fd = (P.sigma_l2sq(h + eps * D) - P.sigma_l2sq(h - eps * D)) / (2 * eps)
        an = np.sum(gre[iu] * X[iu]) + np.sum(gim[iu1] * Y[iu1])
        assert abs(fd - an) < 1e-6 * abs(an), (fd, an)


def test_adam_first_step_closed_form(oracle):
    # From zero moments the first Adam step is lr * g / (|g| + eps / sqrt(1 - beta2)) per
    # parameter (scale = lr sqrt(1-b2)/(1-b1), m = (1-b1) g, v = (1-b2) g^2); a symmetric real
    # gradient on one off-diagonal pair moves only Re G_01 = Re G_10, and the trace is unchanged.
    nb = 4
    h = np.eye(2 * nb)
    g = np.zeros(2 * nb * nb)
    g[0 * nb + 1] = g[1 * nb + 0] = 0.3
    m, v = np.zeros_like(g), np.zeros_like(g)
    lr = 1e-2
    h2, t = oracle.adam_step(h, g, m, v, 0, lr)
    assert t == 1
    step = lr * 0.3 / (0.3 + 1e-8 / np.sqrt(1 - 0.999))
    assert abs(h2[0, 2] - (-step)) < 1e-15 and abs(h2[2, 0] - (-step)) < 1e-15
    assert abs(h2[1, 3] - (-step)) < 1e-15
    d = h2 - h
    d[0, 2] = d[2, 0] = d[1, 3] = d[3, 1] = 0.0
    assert np.max(np.abs(d)) == 0.0


def test_legacy_weighted_step_is_the_uninverted_T(oracle):
    # donaldson_iteration_quintic_weighted (quintic.rs:4256-4331) builds the same trace-normalised
    # T as the fixed step (quintic.rs:1771-1812), without the inversion: two independent oracle
    # code paths must agree.
    import numpy as np
    P = oracle.Problem(2, 1500, 9)
    rng = np.random.default_rng(4)
    a = rng.standard_normal((P.nb, P.nb)) + 1j * rng.standard_normal((P.nb, P.nb))
    G = a @ a.conj().T / P.nb + np.eye(P.nb)
    h = np.zeros((2 * P.nb, 2 * P.nb))
    h[0::2, 0::2] = G.real
    h[0::2, 1::2] = -G.imag
    h[1::2, 0::2] = G.imag
    h[1::2, 1::2] = G.real
    hl = oracle.legacy_step(P.sv, h, P.w)
    _, _, (tre, tim, _) = oracle.donaldson_step(P.sv, P.w, h, want_t=True)
    assert np.abs(hl[0::2, 0::2] - tre).max() <= 1e-13 * np.abs(tre).max()
    assert np.abs(hl[1::2, 0::2] - tim).max() <= 1e-13 * np.abs(tre).max()
    # unweighted: the brute-force formula with uniform weights 1/n
    hu = oracle.legacy_step(P.sv, h, None)
    s = P.sv[:, 0::2] + 1j * P.sv[:, 1::2]
    K = np.einsum("pa,ab,pb->p", s.conj(), G, s).real
    T = np.einsum("p,pa,pb->ab", 1.0 / K, s.conj(), s) / P.n
    T *= P.nb / np.trace(T).real
    assert np.abs(hu[0::2, 0::2] - T.real).max() <= 1e-12 * np.abs(T).max()
    assert np.abs(hu[1::2, 0::2] - T.imag).max() <= 1e-12 * np.abs(T).max()


# --------------------------------------------------------------------------------------------
# Newton-projection sampler (quintic.rs:3334-3482, 3568-3595) and the legacy Monge-Ampere
# residual pipeline (quintic.rs:4224-4250, 4961-5049)
# --------------------------------------------------------------------------------------------
def test_newton_projects_random_to_quintic(oracle):
    # The reference's newton_projects_random_to_quintic (quintic.rs:5497-5519): the first LCG
    # draws of seed 42, projected with tol 1e-10, leave |f|^2 < 1e-15.
    pts, tried, _ = oracle.sample_newton(1, 42, tol=1e-10)
    assert tried == 1 and pts.shape == (1, 10)
    f, _ = oracle.quintic_eval(pts)
    assert f[0, 0] ** 2 + f[0, 1] ** 2 < 1e-15


def test_sampled_points_lie_on_quintic(oracle):
    # The reference's sampled_points_lie_on_quintic (quintic.rs:5521-5534).
    pts, _, _ = oracle.sample_newton(50, 7, tol=1e-10)
    assert pts.shape == (50, 10)
    f, _ = oracle.quintic_eval(pts)
    assert np.all(f[:, 0] ** 2 + f[:, 1] ** 2 < 1e-15)
    assert np.allclose(np.sum(pts * pts, axis=1), 1.0, atol=1e-14)


def test_newton_draw_order_and_normalisation(oracle):
    # With tol = 10 every start point is accepted before any step (|f| <= sum |z_j|^5 <= 1 on
    # the unit sphere), so the sampler returns try t's 10 normals (LCG draws 20t .. 20t + 19,
    # Box-Muller pairs) divided by their norm, bit for bit.
    n = 6
    pts, tried, acc = oracle.sample_newton(n, 99, tol=10.0)
    assert tried == n and acc.all()
    _, _, normals = oracle.lcg_stream(99, 10 * n)
    z = normals.reshape(n, 10)
    nrm = np.sqrt(np.array([sum(float(v) * float(v) for v in row) for row in z]))
    assert np.array_equal(pts, z / nrm[:, None])


def test_newton_projection_fixed_point_and_dwork(oracle):
    # A point on the variety is returned unchanged; the Dwork form at psi = 0 is the Fermat one.
    z = np.array([1, 0, 0, 0, 0, 0, 0, 0, -1, 0], float) / np.sqrt(2.0)
    ok, zp = oracle.newton_project(z, 0.0, 1e-12)
    assert ok and np.array_equal(zp, z)
    a, _, _ = oracle.sample_newton(40, 5, psi=0.0, dwork=False)
    b, _, _ = oracle.sample_newton(40, 5, psi=0.0, dwork=True)
    assert np.array_equal(a, b)
    # Dwork psi = 0.5: every point satisfies the deformed equation.
    c, _, _ = oracle.sample_newton(200, 5, psi=0.5)
    f, _ = oracle.quintic_eval(c, 0.5)
    assert np.all(np.hypot(f[:, 0], f[:, 1]) < 1e-11)


def test_monge_ampere_pipeline_in_published_bounds(oracle):
    # monge_ampere_residual_quintic_in_published_bounds_at_k2 (quintic.rs:5556-5580): Newton
    # points (1500, seed 23, tol 1e-10), raw k=2 basis, 30 un-inverted Donaldson iterations
    # (tol 1e-6), then sigma_L1 with the CY weights inside PublishedQuinticBounds k=2
    # ([0.05, 1.0], quintic.rs:5089-5098).
    pts, _, _ = oracle.sample_newton(1500, 23, tol=1e-10)
    mono = oracle.monomials(2)
    sv, sd = oracle.sections(pts, mono)
    w, log_om, _ = oracle.weights_logomega(pts)
    h, res = oracle.legacy_solve(sv, 30, 1e-6)
    assert res[-1] < res[0]
    s = oracle.sigma(pts, sv, sd, log_om, w, h)
    assert np.isfinite(s) and 0.05 <= s <= 1.0, s


def test_newton_polish_of_the_dk_roots(oracle):
    """The FP64 Newton polish after the Durand-Kerner sweeps (not in the reference, which stops
    after 50 sweeps). Converged roots are Newton fixed points, so after 50 sweeps the polish moves
    no point by more than rounding; after 12 sweeps it converges the roots DK left unconverged,
    which a wrong step (sign, derivative) would not do."""
    a = lambda f: np.abs(f[:, 0] + 1j * f[:, 1])  # noqa: E731
    for psi in (0.0, 0.5):
        p0, t0 = oracle.sample_sz(4000, 42, psi)
        p2, t2 = oracle.sample_sz(4000, 42, psi, polish=2)
        assert t0 == t2 and p0.shape == p2.shape
        assert np.max(np.abs(p2 - p0)) < 1e-12
        f2, _ = oracle.quintic_eval(p2, psi)
        assert np.max(a(f2)) < 1e-12
    p, _ = oracle.sample_sz(10000, 42, dk_iters=12)
    pp, _ = oracle.sample_sz(10000, 42, dk_iters=12, polish=4)
    f, _ = oracle.quintic_eval(p)
    fp, _ = oracle.quintic_eval(pp)
    assert np.sum(a(fp) > 1e-10) * 5 < np.sum(a(f) > 1e-10)
    assert np.quantile(a(fp), 0.9) < 1e-14 < np.quantile(a(f), 0.9)
