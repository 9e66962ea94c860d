"""Pins of the oracle's L-BFGS sigma refinement and random restarts (quintic.rs:545-582,
678-880, 1100-1223, 2003-2048) by mathematics that does not restate them: the two-loop
recursion against the explicit BFGS inverse-Hessian update, convergence to the known minimiser
of a quadratic and of the Rosenbrock function, the Hermitian projection's invariants, and the
restart loop's best-keeping on a small sigma problem. CPU only."""
import numpy as np
import pytest


def _spd(rng, n):
    a = rng.standard_normal((n, n))
    return a @ a.T + n * np.eye(n)


def test_two_loop_matches_explicit_bfgs(oracle):
    # H0 = gamma I with gamma = s.y / y.y of the newest pair, then for each pair (oldest first)
    # H <- (I - rho s y^T) H (I - rho y s^T) + rho s s^T; the recursion must give -H g.
    rng = np.random.default_rng(3)
    n = 40
    A = _spd(rng, n)
    pairs = []
    for _ in range(6):
        s = rng.standard_normal(n)
        y = A @ s
        pairs.append((s, y, 1.0 / float(s @ y)))
    g = rng.standard_normal(n)
    for m in range(0, 7):
        P = pairs[len(pairs) - m:] if m else []
        if P:
            s_l, y_l, _ = P[-1]
            H = (float(s_l @ y_l) / float(y_l @ y_l)) * np.eye(n)
        else:
            H = np.eye(n)
        for s, y, rho in P:
            V = np.eye(n) - rho * np.outer(y, s)
            H = V.T @ H @ V + rho * np.outer(s, s)
        d = oracle.lbfgs_direction(g, P)
        assert np.allclose(d, -H @ g, rtol=1e-10, atol=1e-12), m


def test_lbfgs_quadratic_minimiser(oracle):
    rng = np.random.default_rng(5)
    n = 30
    A = _spd(rng, n)
    b = rng.standard_normal(n)
    x_star = np.linalg.solve(A, b)

    def fgs(x):
        f = 0.5 * x @ A @ x - b @ x
        return f, A @ x - b, f

    hist, x_best, _ = oracle.lbfgs_minimize(np.zeros(n), fgs, lambda x, d, t: x + t * d, 200, 8)
    assert np.linalg.norm(x_best - x_star) < 1e-6 * np.linalg.norm(x_star)
    assert all(h2 <= h1 + 1e-12 for h1, h2 in zip(hist, hist[1:]))  # Armijo: f never rises


def test_lbfgs_rosenbrock(oracle):
    # A non-convex valley: the steepest-descent retries and history resets are exercised.
    def fgs(x):
        a, b = x
        f = (1 - a) ** 2 + 100 * (b - a * a) ** 2
        g = np.array([-2 * (1 - a) - 400 * a * (b - a * a), 200 * (b - a * a)])
        return f, g, f

    hist, x_best, _ = oracle.lbfgs_minimize(np.array([-1.2, 1.0]), fgs,
                                            lambda x, d, t: x + t * d, 500, 5)
    assert np.linalg.norm(x_best - np.array([1.0, 1.0])) < 1e-5
    assert hist[-1] < 1e-10


def test_apply_step_projection(oracle):
    rng = np.random.default_rng(7)
    nb = 9
    X = rng.standard_normal((nb, nb)) + 1j * rng.standard_normal((nb, nb))
    G = X @ X.conj().T + nb * np.eye(nb)
    G = 0.5 * (G + G.conj().T)  # exactly Hermitian, as the solver keeps G
    h = oracle._pack_h(G.real, G.imag)
    d = rng.standard_normal(2 * nb * nb)
    out = oracle.lbfgs_apply_step(h, d, 0.3)
    hre, him = oracle._unpack_h(out)
    assert np.array_equal(hre, hre.T)
    assert np.array_equal(him, -him.T)
    assert np.all(np.diag(him) == 0.0)
    assert abs(np.trace(hre) - nb) < 1e-12
    # For a direction already in the projected form the step is h + t D, renormalised.
    dre = d[: nb * nb].reshape(nb, nb)
    dre = 0.5 * (dre + dre.T)
    dim = np.triu(d[nb * nb:].reshape(nb, nb), 1)
    d2 = np.concatenate([dre.ravel(), dim.ravel()])
    out2 = oracle.lbfgs_apply_step(h, d2, 0.3)
    Gn = (G.real + 0.3 * dre) + 1j * (G.imag + 0.3 * (dim - dim.T))
    Gn *= nb / np.trace(Gn.real)
    r2, i2 = oracle._unpack_h(out2)
    assert np.allclose(r2 + 1j * i2, Gn, rtol=0, atol=1e-13)
    # A zero step only renormalises.
    r0, i0 = oracle._unpack_h(oracle.lbfgs_apply_step(h, d, 0.0))
    assert np.allclose(r0 + 1j * i0, G * (nb / np.trace(G.real)), rtol=0, atol=1e-13)


def test_perturbation_is_hermitian_and_bounded(oracle):
    nb, scale = 7, 0.05
    h = oracle.identity_block(nb)
    _, u, _ = oracle.lcg_stream(99, nb * nb)
    p = oracle.perturb_h_hermitian(h, scale, iter(u.tolist()))
    pre, pim = oracle._unpack_h(p)
    P = (pre - np.eye(nb)) + 1j * pim
    assert np.allclose(P, P.conj().T, atol=0)
    assert np.all(np.abs(P.real) <= scale) and np.all(np.abs(P.imag) <= scale)
    # The first draw goes to the (0, 0) diagonal, the next two to re and im of (0, 1).
    # (The diagonal is recovered as (1 + p) - 1, exact to an ulp of 1.)
    assert abs(P[0, 0].real - scale * (u[0] * 2 - 1)) < 4e-16
    assert P[0, 1].real == scale * (u[1] * 2 - 1) and P[0, 1].imag == scale * (u[2] * 2 - 1)
    # nb^2 draws per perturbation: diagonal nb, plus two per pair.
    assert abs(P[nb - 1, nb - 1].real - scale * (u[nb * nb - 1] * 2 - 1)) < 4e-16


def test_restarts_keep_the_best(oracle):
    P = oracle.Problem(2, 1500, 5)
    h, _ = oracle.donaldson_step(P.sv, P.w, P.h)

    def refine(hh):
        return oracle.sigma_refine_analytic(P, hh, 2, 1e-3)[1]

    best, history, h_best = oracle.refine_with_restarts(P, h, 3, refine, 0.01, 6789)
    assert len(history) == 4
    assert best == min(history)
    assert P.sigma(h_best) == best
    # The first entry is the plain run from h.
    assert history[0] == P.sigma(refine(h))
