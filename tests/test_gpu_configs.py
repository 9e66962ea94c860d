"""The benchmark protocol's spectral/overlap and batching configurations on the GPU against the
oracle, at reduced point counts that the oracle finishes in seconds.

Spectral/overlap: degree k = 8 with the quotient basis (N = 460, padded width 928), Fermat and
Dwork psi = 0.5; FS-Gram and two FP64 Donaldson steps (donaldson_step_workspace,
quintic.rs:1734-1894), the intrinsic Laplacian on the section-degree-3 functions
s_a conj(s_b) / |z|^6 (B = 1225) of that metric, and the 20 random degree-5 modes with all 1540
triple overlaps (yukawa_overlap_real.rs:271-315 with the measure w |Omega|^2).

Batching: the 32 instance specs drawn by the LCG with seed 7 (psi in [0, 0.9], k in 2..6), with
point counts capped at 40,000, balanced for 15 iterations in one captured graph in split
precision with the refinement sweep, each instance against a stand-alone FP64 solve of its
own points; two instances' FP64 solves against the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _relmax(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


_K8 = {}


def _k8(cy3q, oracle, psi):
    """k = 8 quotient problem on 5000 points with two FP64 steps, GPU and oracle side by side."""
    if psi not in _K8:
        P = oracle.Problem(8, 5000, 42, psi=psi, quotient=True)
        s = cy3q.QuinticSolver.from_points(8, P.pts, psi=psi, quotient=True)
        s.orthonormalise_basis_fs_gram()
        h = P.h
        hs = []
        for _ in range(2):
            h, r_ref = oracle.donaldson_step(P.sv, P.w, h)
            r = s.donaldson_step()
            hs.append((h, r_ref, s.h_block, r))
        _K8[psi] = (P, s, hs)
    return _K8[psi]


@pytest.mark.parametrize("psi", [0.0, 0.5])
def test_k8_quotient_fs_gram_and_steps(cy3q, oracle, psi):
    print(f"case: k=8 quotient N=460, 5000 points, psi={psi}", flush=True)
    P, s, hs = _k8(cy3q, oracle, psi)
    assert P.nb == s.n_basis == 460
    sv, _ = s.sections(derivs=False)
    assert _rel(sv, P.sv) < 1e-11
    for it, (h_ref, r_ref, h, r) in enumerate(hs):
        assert _rel(h, h_ref) < 1e-11, it
        assert abs(r / r_ref - 1) < 1e-9, it
    e = s.sigma_ext(want_eta=True)
    assert abs(e[0] / P.sigma(hs[-1][0]) - 1) < 1e-10


def test_k8_intrinsic_laplacian_b1225(cy3q, oracle):
    import scipy.linalg as sl
    from papers_amd import galerkin as G
    print("case: intrinsic Laplacian B=1225 on the k=8 psi=0.5 metric, 5000 points", flush=True)
    P, s, hs = _k8(cy3q, oracle, 0.5)
    h = hs[-1][0]
    s.h_block = h
    low, rank, M, K = G.solver_laplacian_spectrum(s, 3, n_low=50, want_matrices=True)
    Mo, Ko, _ = oracle.lap_quintic_mk(P, h, 3)
    assert M.shape == (1225, 1225) and rank == 1225
    assert _relmax(M, Mo) <= 1e-10
    assert _relmax(K, Ko) <= 1e-10
    ref = sl.eigh(Ko, Mo, eigvals_only=True)
    assert np.abs(low - ref[:50]).max() <= 1e-8 * ref.max()
    assert abs(low[0]) <= 1e-9 * ref.max()


def test_k8_yukawa_twenty_degree5_modes(cy3q, oracle):
    from papers_amd import galerkin as G
    print("case: 20 degree-5 random modes, 1540 triples, k=8 psi=0.5 metric", flush=True)
    P, s, hs = _k8(cy3q, oracle, 0.5)
    y, coef = G.solver_yukawa_random_modes(s, degree=5, n_modes=20, coeff_seed=42)
    assert coef.shape == (126, 20)
    sv5, _ = oracle.sections(P.pts, oracle.monomials(5, False), derivs=False)
    S = sv5[:, 0::2] + 1j * sv5[:, 1::2]
    modes = (S @ coef).T
    ok = np.isfinite(P.w) & (P.w > 0) & np.isfinite(P.log_om)
    c = np.where(ok, P.w * np.exp(np.where(ok, P.log_om, 0.0)), np.nan)
    yo = oracle.yuk_triple(modes, np.ones(P.n, complex), c)
    t = G.triple_index(20)
    assert y.shape[0] == len(t) == 1540
    assert _relmax(y, yo[t[:, 0], t[:, 1], t[:, 2]]) <= 1e-11


def test_batch_32_drawn_instances(cy3q, oracle):
    from papers_amd.batch import Batch, draw_specs
    specs = [(k, min(n, 40000), seed, psi) for k, n, seed, psi in draw_specs(7, 32)]
    print(f"case: 32 drawn instances, {sum(x[1] for x in specs)} points", flush=True)
    b = Batch(specs)
    b.solve_graph(15)
    g_split = [x.h_block.copy() for x in b.solvers]
    sig_split = [x.sigma() for x in b.solvers]
    picked = []
    for i, (k, n, seed, psi) in enumerate(specs):
        bs = b.solvers[i]
        s = cy3q.QuinticSolver.from_points(k, bs.points, psi=psi, quotient=True)
        s.orthonormalise_basis_fs_gram()
        s.donaldson_solve(15, 0.0)
        g64 = s.h_block
        tol = 2e-6 * max(1.0, (40000.0 / bs.n_points) ** 0.5)
        err = _rel(g_split[i], g64)
        assert err < tol, (i, specs[i], err)
        assert abs(sig_split[i] / s.sigma() - 1) < 1e-6, i
        if k in (2, 6) and len([p for p in picked if p[0] == k]) == 0:
            picked.append((k, i, bs.points, psi, g64))
    assert len(picked) == 2
    for k, i, pts, psi, g64 in picked:
        mono = oracle.monomials(k, True)
        w, lo, _ = oracle.weights_logomega(pts, psi)
        sv, sd = oracle.sections(pts, mono, derivs=False)
        sv, _, _ = oracle.orthonormalise(sv, None, w)
        h = oracle.identity_block(mono.shape[0])
        for _ in range(15):
            h, _ = oracle.donaldson_step(sv, w, h)
        assert _rel(g64, h) < 1e-10, (k, i)
    b.close()
