"""The host loop of the 8-coordinate Donaldson solve (papers_amd/metric8.py) with the oracle's
route34 T-operator: the reference's adaptive-damping and Tikhonov-schedule known answers
(schoen_metric.rs, ty_metric.rs tests), iteration invariants (trace N, Hermitian, the damped
blend), the Tikhonov shift against a direct inverse, the schedule's lambda per iteration,
convergence, and the catastrophe guard's snapshot restore."""
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


def test_sigma_of_a_non_positive_definite_iterate_is_nan(oracle):
    """The sigma contract rejects an h that is not positive definite (ERR_NOT_PD); the loop
    records NaN for that iterate and runs on, with h and the residuals unchanged."""
    from papers_amd.quintic import ERR_NOT_PD, Cy3qError
    sv, w, nb = _problem(seed=4)
    base = _top(oracle, sv, w)
    calls = {"n": 0}

    def sigma(hr, hi):
        calls["n"] += 1
        if calls["n"] == 3:
            raise Cy3qError(ERR_NOT_PD, "h is not Hermitian positive definite")
        return float(calls["n"])

    res = M8.solve_route34(base, sigma, np.eye(nb), np.zeros((nb, nb)), w, max_iter=4, tol=0.0)
    ref = M8.solve_route34(base, None, np.eye(nb), np.zeros((nb, nb)), w, max_iter=4, tol=0.0)
    assert res["sigma_fs_identity"] == 1.0
    assert np.isnan(res["sigma_history"][1]) and res["sigma_history"][2] == 4.0
    assert res["donaldson_history"] == ref["donaldson_history"]
    assert np.array_equal(res["h_re"], ref["h_re"])

    def other(hr, hi):
        raise Cy3qError(-2, "device failure")

    with pytest.raises(Cy3qError):
        M8.solve_route34(base, other, np.eye(nb), np.zeros((nb, nb)), w, max_iter=2, tol=0.0)


def test_tikhonov_schedule_known_answers():
    # tikhonov_shift_resolution_rules (schoen_metric.rs and ty_metric.rs tests)
    assert M8.resolve_tikhonov(None) is None
    assert M8.auto_ty_tikhonov(2) is None and M8.auto_ty_tikhonov(3) is None
    assert M8.auto_schoen_tikhonov(3, 3, 1) is None and M8.auto_schoen_tikhonov(4, 4, 1) is None
    for t in (M8.auto_ty_tikhonov(4), M8.auto_schoen_tikhonov(4, 4, 2)):
        assert (t.lambda_max, t.lambda_min, t.schedule_exponent) == (1e-3, 1e-9, 1.0)
        assert not t.gating.stall_band
    g = M8.auto_ty_gated_tikhonov(4).gating
    assert g.stall_band and (g.residual_lo, g.residual_hi, g.min_stuck_iters) == (1e-7, 1e-3, 3)
    assert M8.auto_schoen_gated_tikhonov(3, 3, 1) is None
    custom = M8.TikhonovShift(2e-4, 1e-10, 1.5)
    assert M8.resolve_tikhonov(custom) is custom
    t = M8.TikhonovShift.k4_default()
    assert abs(t.lambda_at(1.0, 1.0) - 1e-3) < 1e-15        # geometric ceiling wins
    assert abs(t.lambda_at(1e-30, 1.0) - 1e-9) < 1e-15      # floor
    assert abs(t.lambda_at(math.nan, math.nan) - 1e-3) < 1e-15  # cold start
    # residual at 0.1 of its first value: geometric term 1e-4, under the 1% cap of 1e-3
    assert abs(t.lambda_at(0.1, 1.0) - 1e-4) < 1e-15
    # the 1% cap wins once the residual is small: 1e-5 -> 1e-7 (geometric term 1e-6)
    assert abs(t.lambda_at(1e-5, 1e-2) - 1e-7) < 1e-20
    # p = 2 halves twice: ratio 0.5 -> 0.25 lambda_max, under a 1% cap of 0.5
    assert abs(M8.TikhonovShift(1e-3, 1e-9, 2.0).lambda_at(0.5, 1.0) - 2.5e-4) < 1e-18


def test_gated_tikhonov_engages_in_stall_band():
    # gated_tikhonov_engages_in_stall_band (schoen_metric.rs)
    gating = M8.TikhonovGating.schoen_stall_band_default()
    state = M8.GatingState()
    traj = [1.0e-5, 1.01e-5, 0.99e-5, 1.005e-5, 0.998e-5, 1.002e-5, 1.001e-5, 0.999e-5]
    decisions, prev = [], math.nan
    for r in traj:
        decisions.append(state.is_open(prev, gating))
        state.update(r, 0.95, 1.05)
        prev = r
    assert decisions == [False, False, False, False, True, True, True, True]
    assert M8.TikhonovShift.k4_gated_default().lambda_at(1e-5, 1e-2) > 0.0
    # out of the band the gate stays closed whatever the streak
    assert not state.is_open(2e-3, gating) and not state.is_open(5e-8, gating)
    # a non-finite residual resets the streak
    state.update(math.nan, 0.95, 1.05)
    assert state.in_band_streak == 0 and not state.is_open(1e-5, gating)


def test_gated_tikhonov_closed_on_healthy_descent():
    # gated_tikhonov_does_not_engage_on_healthy_convergence (schoen_metric.rs)
    gating = M8.TikhonovGating.schoen_stall_band_default()
    state, r, prev = M8.GatingState(), 1.0, math.nan
    for _ in range(20):
        assert not state.is_open(prev, gating)
        state.update(r, 0.95, 1.05)
        prev, r = r, r * 0.5
    assert state.in_band_streak == 0


def _scripted(monkeypatch, residuals):
    seen = []

    def fake(top, h_re, h_im, w, damping=1.0, tikhonov=0.0, inverse=None):
        seen.append(tikhonov)
        return h_re, h_im, residuals[len(seen) - 1]

    monkeypatch.setattr(M8, "donaldson_iteration_route34", fake)
    return seen


def test_schedule_lambda_per_iteration(monkeypatch):
    # Always-on k4 schedule: cold start at lambda_max, then min(1e-3 r/r0, r/100) with r the
    # previous residual and r0 the first one.
    res = [1e-2, 5e-3, 1e-3, 1e-5, 1e-9]
    seen = _scripted(monkeypatch, res)
    nb = 2
    out = M8.solve_route34(None, None, np.eye(nb), np.zeros((nb, nb)), None, max_iter=5, tol=0.0,
                           tikhonov=M8.TikhonovShift.k4_default())
    want = [1e-3, 1e-4, 5e-5, 1e-5, 1e-7]
    assert len(seen) == 5 and out["tikhonov_history"] == seen
    assert all(abs(a - b) <= 1e-12 * b for a, b in zip(seen, want)), seen
    # Gated: closed until three in-band ratios have been seen, then lambda from the schedule
    res = [1e-4, 1.01e-4, 0.99e-4, 1.0e-4, 1.0e-4]
    seen = _scripted(monkeypatch, res)
    M8.solve_route34(None, None, np.eye(nb), np.zeros((nb, nb)), None, max_iter=5, tol=0.0,
                     tikhonov=M8.TikhonovShift.k4_gated_default())
    assert seen[:4] == [0.0, 0.0, 0.0, 0.0]
    assert abs(seen[4] - 1e-6) <= 1e-18  # ratio 1 -> 1e-3, capped at 1% of 1e-4
    # A constant shift and None pass straight through.
    seen = _scripted(monkeypatch, res)
    M8.solve_route34(None, None, np.eye(nb), np.zeros((nb, nb)), None, max_iter=3, tol=0.0,
                     tikhonov=0.25)
    assert seen == [0.25] * 3
    seen = _scripted(monkeypatch, res)
    M8.solve_route34(None, None, np.eye(nb), np.zeros((nb, nb)), None, max_iter=3, tol=0.0,
                     tikhonov=None)
    assert seen == [0.0] * 3


def test_gated_schedule_is_bitwise_inert_on_healthy_descent(oracle):
    # tikhonov_shift_no_regression: a gate that never opens leaves every inversion as without
    # the schedule, so the trajectory is the unshifted one bit for bit.
    sv, w, nb = _problem(seed=5, n=6000, nb=10)
    top = _top(oracle, sv, w)
    a = M8.solve_route34(top, None, np.eye(nb), np.zeros((nb, nb)), w, max_iter=40, tol=1e-12)
    b = M8.solve_route34(top, None, np.eye(nb), np.zeros((nb, nb)), w, max_iter=40, tol=1e-12,
                         tikhonov=M8.TikhonovShift.k4_gated_default())
    assert b["tikhonov_history"] == [0.0] * b["iterations_run"]
    assert np.array_equal(a["h_re"], b["h_re"]) and np.array_equal(a["h_im"], b["h_im"])
    assert a["donaldson_history"] == b["donaldson_history"]
    # the always-on schedule does shift the trajectory
    c = M8.solve_route34(top, None, np.eye(nb), np.zeros((nb, nb)), w, max_iter=3, tol=0.0,
                         tikhonov=M8.TikhonovShift.k4_default())
    assert c["tikhonov_history"][0] == 1e-3 and c["donaldson_history"][0] != \
        a["donaldson_history"][0]


def test_not_positive_definite_h_restores_snapshot(oracle):
    # The GPU T-operator raises CY3Q_ERR_NOT_PD for an h that is not positive definite; the
    # loop turns it into a NaN residual, so the minimum-residual snapshot comes back (the
    # reference keeps iterating on K clamped at 1e-30, cy3_donaldson_gpu.rs:84). Other errors
    # propagate.
    from papers_amd.quintic import ERR_HIP, ERR_NOT_PD, Cy3qError

    sv, w, nb = _problem(seed=9)
    base = _top(oracle, sv, w)
    calls = {"n": 0}

    def not_pd(hr, hi):
        calls["n"] += 1
        if calls["n"] == 3:
            raise Cy3qError(ERR_NOT_PD, "h is not Hermitian positive definite")
        return base(hr, hi)

    res = M8.solve_route34(not_pd, None, np.eye(nb), np.zeros((nb, nb)), w, max_iter=10, tol=0.0)
    good = M8.solve_route34(base, None, np.eye(nb), np.zeros((nb, nb)), w, max_iter=2, tol=0.0)
    assert res["iterations_run"] == 2
    assert np.array_equal(res["h_re"], good["h_re"])
    assert res["donaldson_history"][-1] == min(res["donaldson_history"][:2])

    def hip_fails(hr, hi):
        raise Cy3qError(ERR_HIP, "launch failed")

    with pytest.raises(Cy3qError):
        M8.solve_route34(hip_fails, None, np.eye(nb), np.zeros((nb, nb)), w, max_iter=3, tol=0.0)
