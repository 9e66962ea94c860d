"""L-BFGS sigma refinement and random restarts (quintic.rs:545-582, 678-1139) on the GPU against
the oracle, on the same points and starting metric. The line search makes discrete choices, so
the tolerances are set for trajectories that take the same choices; each case checks that the
oracle's trial points kept G positive definite (the GPU rejects any other point, where the
reference evaluates s^H G s directly). Tolerances are written in each test."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _setup(cy3q, oracle, k, n, seed, quotient=False):
    P = oracle.Problem(k, n, seed, quotient=quotient)
    s = cy3q.QuinticSolver.from_points(k, P.pts, quotient=quotient)
    s.orthonormalise_basis_fs_gram()
    h, _ = oracle.donaldson_step(P.sv, P.w, P.h)
    s.h_block = h
    return P, s, h


@pytest.mark.parametrize("k,n,seed,n_iter,m", [(3, 2500, 9, 6, 3), (2, 3000, 4, 8, 5)])
def test_lbfgs_matches_oracle(cy3q, oracle, k, n, seed, n_iter, m):
    P, s, h = _setup(cy3q, oracle, k, n, seed)
    hist_ref, h_ref, min_eig = oracle.sigma_refine_lbfgs(P, h, n_iter, m)
    assert min_eig > 0.0
    hist = s.sigma_refine_lbfgs_gpu(n_iter, m)
    print(f"k={k}: oracle {hist_ref}\n       gpu    {hist}")
    assert len(hist) == len(hist_ref)
    assert np.allclose(hist, hist_ref, rtol=1e-8, atol=0)
    assert _rel(s.h_block, h_ref) < 1e-7
    # The solver keeps the lowest-sigma_L1 point, and it improves on the start.
    assert abs(s.sigma() - min(hist)) <= 1e-12 * min(hist)
    assert min(hist) < hist[0]


def test_adam_restarts_match_oracle(cy3q, oracle):
    P, s, h = _setup(cy3q, oracle, 3, 2500, 9)
    n_restarts, n_iter, lr, scale, seed = 2, 3, 1e-3, 0.01, 6789

    def refine(hh):
        return oracle.sigma_refine_analytic(P, hh, n_iter, lr)[1]

    best_ref, hist_ref, h_ref = oracle.refine_with_restarts(P, h, n_restarts, refine, scale, seed)
    best, hist = s.sigma_refine_analytic_with_restarts(n_restarts, n_iter, lr, scale, seed)
    assert np.allclose(hist, hist_ref, rtol=1e-9, atol=0)
    assert abs(best - best_ref) <= 1e-9 * best_ref
    assert best == min(hist)
    assert _rel(s.h_block, h_ref) < 1e-8


def test_lbfgs_restarts_match_oracle(cy3q, oracle):
    P, s, h = _setup(cy3q, oracle, 2, 3000, 4)
    n_restarts, n_iter, m, scale, seed = 2, 3, 3, 0.005, 12345
    eigs = []

    def refine(hh):
        _, hb, me = oracle.sigma_refine_lbfgs(P, hh, n_iter, m)
        eigs.append(me)
        return hb

    best_ref, hist_ref, h_ref = oracle.refine_with_restarts(P, h, n_restarts, refine, scale, seed)
    assert min(eigs) > 0.0
    best, hist = s.sigma_refine_lbfgs_with_restarts_gpu(n_restarts, n_iter, m, scale, seed)
    assert np.allclose(hist, hist_ref, rtol=1e-8, atol=0)
    assert abs(best - best_ref) <= 1e-8 * best_ref
    assert _rel(s.h_block, h_ref) < 1e-7


def test_lbfgs_zero_iterations_keeps_h(cy3q, oracle):
    P, s, h = _setup(cy3q, oracle, 2, 2000, 3)
    h0 = s.h_block
    hist = s.sigma_refine_lbfgs_gpu(0, 4)
    assert len(hist) == 1
    assert abs(hist[0] - s.sigma()) <= 1e-14 * hist[0]
    assert _rel(s.h_block, h0) < 1e-14
