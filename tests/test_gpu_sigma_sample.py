"""GPU parity of sigma on a fresh sample (cy3q_solver_sigma_at_sample) against the oracle.

sigma_at_metric_closure (quintic.rs:1520-1616) with the solver's Bergman metric: the balanced G
of one sample evaluated on a second, independently seeded sample, as config 1 of the benchmark
protocol does (10 iterations at k = 3, sigma on a separate sample).
"""
import numpy as np
import pytest

from tests.test_oracle_sigma_sample import _fs_closure_eta, _multinom_diag

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_sigma_at_fresh_sample_matches_oracle(cy3q, oracle):
    print("case: k=3, 6000 points seed 42, 3 steps, sigma on 4000 points seed 4242", flush=True)
    P = oracle.Problem(3, 6000, 42)
    s = cy3q.QuinticSolver.new_with_sampler(3, 6000, 42)
    s.orthonormalise_basis_fs_gram()
    for _ in range(3):
        s.donaldson_step()
    h = s.h_block
    s1, kap, l2, n, eta, pts = s.sigma_at_sample(4000, 4242, want_eta=True)
    r1, rk, rl2, rn, reta, rpts = oracle.sigma_at_sample(P, h, 4000, 4242, want_eta=True)
    assert n == rn == 4000
    assert np.max(np.abs(pts - rpts)) < 1e-12
    ok = np.isfinite(reta)
    assert np.array_equal(ok, np.isfinite(eta))
    np.testing.assert_allclose(eta[ok], reta[ok], rtol=1e-10)
    assert abs(s1 / r1 - 1) < 1e-11
    assert abs(kap / rk - 1) < 1e-11
    assert abs(l2 / rl2 - 1) < 1e-10
    # The solver's own points and metric are untouched.
    assert np.array_equal(s.h_block, h)
    assert s.n_points == 6000


def test_sigma_at_fresh_sample_fubini_study(cy3q, oracle):
    print("case: k=3 raw basis, G = diag(multinomial), 3000 fresh points seed 9", flush=True)
    P = oracle.Problem(3, 800, 42, fs_gram=False)
    s = cy3q.QuinticSolver.from_points(3, P.pts)
    s.h_block = oracle._pack_h(np.diag(_multinom_diag(P.mono)), np.zeros((P.nb, P.nb)))
    s1, kap, l2, n, eta, pts = s.sigma_at_sample(3000, 9, want_eta=True)
    eta_fs = _fs_closure_eta(pts)
    np.testing.assert_allclose(eta, 27 * eta_fs, rtol=1e-9)


def test_sigma_at_own_sample_after_two_fs_grams(cy3q):
    # The fresh sample drawn with the solver's own seed and size is the solver's sample; after
    # two FS-Gram passes its sections go through the composed basis change in one contraction.
    print("case: k=3, 5000 points seed 3, two FS-Gram passes", flush=True)
    s = cy3q.QuinticSolver.new_with_sampler(3, 5000, 3)
    s.orthonormalise_basis_fs_gram()
    s.donaldson_step()
    s.orthonormalise_basis_fs_gram()
    s.donaldson_step()
    s.donaldson_step()
    own = s.sigma_ext()
    fresh = s.sigma_at_sample(5000, 3)
    assert abs(fresh[0] / own[0] - 1) < 1e-11
    assert abs(fresh[1] / own[1] - 1) < 1e-11


@pytest.mark.parametrize("sampler", ["newton", "sz"])
def test_sigma_at_fresh_sample_dwork(cy3q, oracle, sampler):
    print(f"case: Dwork psi=0.3 k=2, fresh {sampler} sample of 2000 points seed 5", flush=True)
    P = oracle.Problem(2, 3000, 42, psi=0.3)
    s = cy3q.QuinticSolver.new_with_psi_sz(2, 3000, 42, 0.3)
    s.orthonormalise_basis_fs_gram()
    s.donaldson_step()
    h = s.h_block
    code = cy3q.SAMPLER_NEWTON_DWORK if sampler == "newton" else cy3q.SAMPLER_SZ_DWORK
    got = s.sigma_at_sample(2000, 5, code)
    ref = oracle.sigma_at_sample(P, h, 2000, 5, sampler)
    assert got[3] == ref[3]
    assert abs(got[0] / ref[0] - 1) < 1e-9
    assert abs(got[1] / ref[1] - 1) < 1e-9


def test_sigma_at_fresh_sample_argument_errors(cy3q):
    s = cy3q.QuinticSolver.new_with_psi_sz(2, 500, 1, 0.2)
    with pytest.raises(cy3q.Cy3qError):
        s.sigma_at_sample(100, 1, cy3q.SAMPLER_SZ)  # Fermat sampler, Dwork metric
    with pytest.raises(cy3q.Cy3qError):
        s.sigma_at_sample(0, 1, cy3q.SAMPLER_SZ_DWORK)
