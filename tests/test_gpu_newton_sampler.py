"""The Newton-projection sampler (QuinticSolver::new / new_with_psi, quintic.rs:174-221,
3334-3482, 3568-3595) and the legacy Monge-Ampere residual (quintic.rs:4961-5049) on the GPU,
against the oracle and the reference's own tests. The GPU draws its normals with the device
libm (log, cos), so points agree with the oracle to rounding rather than bit for bit; the
tolerances are written in each test."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k,n,seed,psi", [(3, 3000, 2024, 0.0), (2, 2000, 7, 0.5)])
def test_newton_sampler_matches_oracle(cy3q, oracle, k, n, seed, psi):
    if psi == 0.0:
        s = cy3q.QuinticSolver.new(k, n, seed, 1e-12)
    else:
        s = cy3q.QuinticSolver.new_with_psi(k, n, seed, 1e-12, psi)
    ref, tried, _ = oracle.sample_newton(n, seed, psi=psi, tol=1e-12)
    pts = s.points
    print(f"k={k} psi={psi}: {pts.shape[0]} points, oracle {ref.shape[0]} in {tried} tries")
    assert pts.shape == ref.shape
    assert np.max(np.abs(pts - ref)) < 1e-12
    w, lo = s.weights, s.log_omega_sq
    w_ref, lo_ref, _ = oracle.weights_logomega(pts, psi)
    assert np.allclose(w, w_ref, rtol=1e-13, atol=0)
    assert np.allclose(lo, lo_ref, rtol=0, atol=1e-12)


def test_reference_newton_tests(cy3q, oracle):
    # newton_projects_random_to_quintic and sampled_points_lie_on_quintic
    # (quintic.rs:5497-5534), on the GPU points.
    s1 = cy3q.QuinticSolver.new(2, 1, 42, 1e-10)
    f, _ = oracle.quintic_eval(s1.points)
    assert s1.n_points == 1 and f[0, 0] ** 2 + f[0, 1] ** 2 < 1e-15
    s50 = cy3q.QuinticSolver.new(2, 50, 7, 1e-10)
    f, _ = oracle.quintic_eval(s50.points)
    assert s50.n_points == 50
    assert np.all(f[:, 0] ** 2 + f[:, 1] ** 2 < 1e-15)


def test_monge_ampere_residual_published_bounds(cy3q, oracle):
    # monge_ampere_residual_quintic_in_published_bounds_at_k2 (quintic.rs:5556-5580) end to end
    # on the GPU: Newton points, the raw k=2 basis, 30 un-inverted Donaldson iterations, the
    # residual in [0.05, 1.0]; and the residual against the oracle's sigma for the same h.
    s = cy3q.QuinticSolver.new(2, 1500, 23, 1e-10)
    res = s.donaldson_solve_legacy(30, 1e-6, weighted=False)
    assert res[-1] < res[0]
    h = s.h_block
    sig = cy3q.monge_ampere_residual_quintic(s.points, h, 2)
    assert np.isfinite(sig) and 0.05 <= sig <= 1.0, sig
    pts = s.points
    sv, sd = oracle.sections(pts, oracle.monomials(2))
    w, lo, _ = oracle.weights_logomega(pts)
    ref = oracle.sigma(pts, sv, sd, lo, w, h)
    assert abs(sig - ref) <= 1e-11 * ref
    assert abs(sig - s.sigma()) <= 1e-13 * sig
    # The weighted form with the caller's weights (uniform here).
    ones = np.ones(pts.shape[0])
    sig_u = cy3q.monge_ampere_residual_quintic(pts, h, 2, weights=ones)
    ref_u = oracle.sigma(pts, sv, sd, lo, ones, h)
    assert abs(sig_u - ref_u) <= 1e-11 * ref_u
