"""The host loop of the 8-coordinate Donaldson solve (papers_amd/metric8.py) with the oracle's
route34 T-operator: the reference's adaptive-damping known answers (schoen_metric.rs tests),
iteration invariants (trace N, Hermitian, the damped blend), the Tikhonov shift against a
direct inverse, convergence, and the catastrophe guard's snapshot restore."""
import math

import numpy as np
import pytest

from papers_amd import metric8 as M8


def test_adaptive_damping_known_answers():
    # adaptive_damping_ramps_up_on_smooth_descent / drops_on_oscillation / clamps_to_min_and_max
    s = M8.AdaptiveDamping(0.3, 0.05, 1.0)
    s.update(1.0)
    assert abs(s.alpha() - 0.3) < 1e-12
    for r in (0.5, 0.25, 0.1):
        s.update(r)
    assert abs(s.alpha() - 0.45) < 1e-12
    s = M8.AdaptiveDamping(0.3, 0.05, 1.0)
    s.update(1.0)
    s.update(2.0)
    assert abs(s.alpha() - 0.21) < 1e-12
    s = M8.AdaptiveDamping(0.07, 0.05, 1.0)
    for _ in range(20):
        s.update(1.0)
        s.update(2.0)
    assert s.alpha() >= 0.05 - 1e-12
    s2 = M8.AdaptiveDamping(0.9, 0.05, 1.0)
    s2.update(1.0)
    r = 1.0
    for _ in range(20):
        r *= 0.5
        s2.update(r)
    assert s2.alpha() <= 1.0 + 1e-12


def _problem(seed=3, n=3000, nb=12):
    rng = np.random.default_rng(seed)
    sv = rng.standard_normal((n, 2 * nb))
    w = rng.uniform(0.1, 1.0, n)
    return sv, w, nb


def _top(oracle, sv, w):
    return lambda hr, hi: oracle.t_operator_raw_route34(sv, w, hr, hi)


def test_iteration_invariants_and_damping(oracle):
    sv, w, nb = _problem()
    top = _top(oracle, sv, w)
    h_re, h_im = np.eye(nb), np.zeros((nb, nb))
    a_re, a_im, r1 = M8.donaldson_iteration_route34(top, h_re, h_im, w)
    assert abs(np.trace(a_re) - nb) < 1e-12 and np.allclose(a_re, a_re.T) and \
        np.allclose(a_im, -a_im.T)
    assert abs(r1 - math.sqrt(np.sum((a_re - h_re) ** 2) + np.sum(a_im ** 2))) < 1e-14
    # alpha blends the full step with the old h; the residual is that of the full step
    b_re, b_im, r2 = M8.donaldson_iteration_route34(top, h_re, h_im, w, damping=0.25)
    assert np.allclose(b_re, 0.25 * a_re + 0.75 * h_re, rtol=0, atol=1e-15)
    assert np.allclose(b_im, 0.25 * a_im, rtol=0, atol=1e-15)
    assert r2 == r1


def test_tikhonov_is_a_shift_of_the_normalised_t(oracle):
    sv, w, nb = _problem(seed=8)
    top = _top(oracle, sv, w)
    rng = np.random.default_rng(1)
    a = rng.standard_normal((nb, nb)) + 1j * rng.standard_normal((nb, nb))
    h = a @ a.conj().T / nb + np.eye(nb)
    lam = 0.3
    g_re, g_im, _ = M8.donaldson_iteration_route34(top, h.real, h.imag, w, tikhonov=lam)
    pr, pi = top(h.real, h.imag)
    t = (pr + 1j * pi) / np.sum(w)
    t = 0.5 * (t + t.conj().T)
    t *= nb / np.trace(t).real
    g = np.linalg.inv(t + lam * np.eye(nb))
    g *= nb / np.trace(g).real
    assert np.allclose(g_re + 1j * g_im, g, rtol=0, atol=1e-12)


def test_solve_converges(oracle):
    sv, w, nb = _problem(seed=5, n=6000, nb=10)
    res = M8.solve_route34(_top(oracle, sv, w), None, np.eye(nb), np.zeros((nb, nb)), w,
                           max_iter=200, tol=1e-10)
    hist = res["donaldson_history"]
    assert res["final_donaldson_residual"] < 1e-10 and res["iterations_run"] == len(hist)
    assert hist[-1] < hist[0] * 1e-6
    # a fixed point: one more iteration moves h by less than the tolerance
    _, _, r = M8.donaldson_iteration_route34(_top(oracle, sv, w), res["h_re"], res["h_im"], w)
    assert r < 1e-9


def test_catastrophe_guard_restores_the_best_snapshot(oracle):
    sv, w, nb = _problem(seed=6)
    base = _top(oracle, sv, w)
    calls = {"n": 0}

    def flaky(hr, hi):
        calls["n"] += 1
        pr, pi = base(hr, hi)
        if calls["n"] >= 4:  # from the 4th call on, T is scaled by a huge factor on the imag part
            pi = pi * 1e6
        return pr, pi

    res = M8.solve_route34(flaky, None, np.eye(nb), np.zeros((nb, nb)), w, max_iter=40,
                           tol=0.0)
    good = M8.solve_route34(base, None, np.eye(nb), np.zeros((nb, nb)), w, max_iter=3, tol=0.0)
    hist = res["donaldson_history"]
    if res["iterations_run"] == 3:
        # restored to the snapshot of iteration 3 (the minimum before the blow-up)
        assert np.allclose(res["h_re"], good["h_re"]) and np.allclose(res["h_im"], good["h_im"])
        assert hist[-1] == min(hist[:-1])
    else:
        pytest.skip("the perturbation did not trigger the guard for this draw")


def test_nan_residual_restores_snapshot(oracle):
    sv, w, nb = _problem(seed=9)
    base = _top(oracle, sv, w)
    calls = {"n": 0}

    def dies(hr, hi):
        calls["n"] += 1
        pr, pi = base(hr, hi)
        if calls["n"] == 3:
            pr = pr * np.nan
        return pr, pi

    res = M8.solve_route34(dies, None, np.eye(nb), np.zeros((nb, nb)), w, max_iter=10, tol=0.0)
    good = M8.solve_route34(base, None, np.eye(nb), np.zeros((nb, nb)), w, max_iter=2, tol=0.0)
    assert res["iterations_run"] == 2
    assert np.array_equal(res["h_re"], good["h_re"])
    assert res["donaldson_history"][-1] == min(res["donaldson_history"][:2])
