"""Laplacian Galerkin assembly, the Hermitian eigensolver and the Yukawa triple overlaps on the
GPU, against the oracle restatements (oracle/cy3q_oracle.c) and library eigen routines.
Tolerances are written in each test. Both sides compute in FP64, and the GPU sums in a
different order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _relmax(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


# ------------------------------------------------------------------------------------------
# Hermitian eigensolver
# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 37, 128])
def test_hermitian_eigen_matches_eigh(cy3q, n):
    from papers_amd import galerkin as G
    rng = np.random.default_rng(n)
    a = _rand_c(rng, n, n)
    h = (a + a.conj().T) / 2
    ev, vec, sw = G.hermitian_eigen(h, vectors=True)
    ref = np.linalg.eigvalsh(h)
    scale = np.abs(ref).max()
    assert np.abs(ev - ref).max() <= 1e-12 * scale
    assert np.abs(h @ vec - vec * ev[None, :]).max() <= 1e-11 * scale
    assert np.abs(vec.conj().T @ vec - np.eye(n)).max() <= 1e-12
    assert sw <= 20


def test_hermitian_eigen_degenerate_and_diagonal(cy3q):
    from papers_amd import galerkin as G
    d = np.diag([3.0, 1.0, 1.0, 2.0, -4.0]).astype(complex)
    ev, vec, sw = G.hermitian_eigen(d, vectors=True)
    assert np.array_equal(ev, [-4.0, 1.0, 1.0, 2.0, 3.0]) and sw == 0
    rng = np.random.default_rng(3)
    q, _ = np.linalg.qr(_rand_c(rng, 6, 6))
    h = q @ np.diag([1.0, 1.0, 1.0, 5.0, 5.0, 9.0]) @ q.conj().T
    ev, _, _ = G.hermitian_eigen(h)
    assert np.abs(ev - [1, 1, 1, 5, 5, 9]).max() <= 1e-13 * 9


# ------------------------------------------------------------------------------------------
# Ambient Laplacian: compute_metric_laplacian_spectrum
# ------------------------------------------------------------------------------------------
def _ambient_inputs(rng, n):
    pts = _rand_c(rng, n, 8) * 0.5
    w = rng.uniform(0.2, 2.0, n)
    om = _rand_c(rng, n)
    w[3] = 0.0
    w[7] = np.nan
    om[11] = np.nan
    return pts, w, om


@pytest.mark.parametrize("deg,n", [(1, 500), (2, 3000), (3, 2500)])
def test_ambient_laplacian_matches_oracle(cy3q, oracle, deg, n):
    from papers_amd import galerkin as G
    rng = np.random.default_rng(deg)
    pts, w, om = _ambient_inputs(rng, n)
    cfg = G.LaplacianConfig.default(max_total_degree=deg, n_low_eigenvalues=10)
    res = G.compute_metric_laplacian_spectrum(pts, w, om, cfg, want_matrices=True)
    ex = oracle.lap_basis(deg)
    assert res.basis_dim == ex.shape[0]
    assert np.array_equal(G.build_test_basis(deg), ex)  # bit-exact basis table
    Mo, Ko = oracle.lap_ambient_mk(pts, w, om, ex)
    # assembly: FP64 MFMA vs the reference's serial loop, 1e-12 of the largest entry
    assert _relmax(res.mass, Mo) <= 1e-12
    assert _relmax(res.stiffness, Ko) <= 1e-12
    assert np.array_equal(res.mass, res.mass.conj().T)
    # spectrum of the same pencil: the reference path (Gauss-Jordan inverse + cyclic Jacobi) vs
    # the GPU path (eigen-inverse + parallel Jacobi), 1e-9 of the largest eigenvalue
    ref = oracle.lap_spectrum_from_mk(res.mass, res.stiffness, 128, 1e-12, 1e-10)
    assert np.abs(res.eigenvalues_full - ref).max() <= 1e-9 * ref.max()
    assert np.array_equal(res.eigenvalues, res.eigenvalues_full[:10])
    assert np.all(np.diff(res.eigenvalues_full) >= 0)


def test_ambient_laplacian_empty_inputs(cy3q):
    from papers_amd import galerkin as G
    res = G.compute_metric_laplacian_spectrum(np.zeros((0, 8), complex), np.zeros(0),
                                              np.zeros(0, complex))
    assert res.eigenvalues.size == 0 and res.basis_dim == 494


# ------------------------------------------------------------------------------------------
# Intrinsic Laplacian of the quintic metric
# ------------------------------------------------------------------------------------------
def _quintic(cy3q, oracle, k, n, seed, psi=0.0, steps=2):
    # From k = 5 the raw monomials are dependent on the quintic: the quotient basis is used.
    quotient = k >= 5
    P = oracle.Problem(k, n, seed, psi=psi, quotient=quotient)
    s = cy3q.QuinticSolver.from_points(k, P.pts, psi=psi, quotient=quotient)
    s.orthonormalise_basis_fs_gram()
    h = P.h
    for _ in range(steps):
        h, _ = oracle.donaldson_step(P.sv, P.w, h)
    s.h_block = h
    return P, s, h


@pytest.mark.parametrize("sdeg,k,n,psi", [(1, 2, 1500, 0.0), (2, 3, 1200, 0.0), (2, 2, 1000, 0.5)])
def test_quintic_laplacian_matches_oracle(cy3q, oracle, sdeg, k, n, psi):
    import scipy.linalg as sl
    from papers_amd import galerkin as G
    P, s, h = _quintic(cy3q, oracle, k, n, 42, psi)
    low, rank, M, K = G.solver_laplacian_spectrum(s, sdeg, n_low=12, want_matrices=True)
    Mo, Ko, _ = oracle.lap_quintic_mk(P, h, sdeg)
    assert _relmax(M, Mo) <= 1e-11
    assert _relmax(K, Ko) <= 1e-10
    B = M.shape[0]
    assert rank == B
    ref = sl.eigh(K, M, eigvals_only=True)
    assert np.abs(low - ref[:12]).max() <= 1e-9 * ref.max()
    # the constant function is in the span and in the kernel
    assert abs(low[0]) <= 1e-10 * ref.max()


@pytest.mark.parametrize("sdeg,k,n,psi", [(1, 2, 1500, 0.0), (3, 3, 20000, 0.5), (3, 6, 8000, 0.5),
                                          (3, 7, 6000, 0.0)])
def test_quintic_laplacian_split_assembly(cy3q, oracle, sdeg, k, n, psi):
    # The fp16 split assembly (the metric contraction y = W x and the rank-k update) against the
    # FP64 one on the same solver: M and K to 1e-6 of their largest entry (about 22-bit
    # products, FP32 partials over at most 4096 points or one column block of W), the
    # eigenvalues to 1e-5 relative, and bit-identical repeats. k = 6 and 7 (2N = 420 and 660
    # real columns) take two and three 256-column blocks of the split contraction.
    from papers_amd import galerkin as G
    from papers_amd.quintic import PREC_FP64, PREC_SPLIT16
    P, s, h = _quintic(cy3q, oracle, k, n, 42, psi)
    low64, r64, M64, K64 = G.solver_laplacian_spectrum(s, sdeg, n_low=20, want_matrices=True,
                                                        precision=PREC_FP64)
    low16, r16, M16, K16 = G.solver_laplacian_spectrum(s, sdeg, n_low=20, want_matrices=True,
                                                        precision=PREC_SPLIT16)
    print(f"sdeg={sdeg}: M {_relmax(M16, M64):.2e}, K {_relmax(K16, K64):.2e}, "
          f"eig {np.max(np.abs(low16[1:] - low64[1:]) / np.abs(low64[1:])):.2e}")
    assert r16 == r64 == M64.shape[0]
    assert _relmax(M16, M64) <= 1e-6
    assert _relmax(K16, K64) <= 1e-6
    assert np.all(np.abs(low16[1:] - low64[1:]) <= 1e-5 * np.abs(low64[1:]))
    assert abs(low16[0]) <= 1e-6 * low64.max()
    low16b, _, M16b, _ = G.solver_laplacian_spectrum(s, sdeg, n_low=20, want_matrices=True,
                                                     precision=PREC_SPLIT16)
    assert np.array_equal(M16b, M16) and np.array_equal(low16b, low16)


# ------------------------------------------------------------------------------------------
# Yukawa
# ------------------------------------------------------------------------------------------
def test_sz_weights_match_oracle(cy3q, oracle):
    from papers_amd import galerkin as G
    rng = np.random.default_rng(8)
    w = rng.uniform(0.1, 1.0, 5000)
    om = _rand_c(rng, 5000)
    w[5] = -1.0
    om[9] = 0.0
    om[10] = np.nan
    sz, sc = G.sz_weights(w, om)
    szo, sco = oracle.yuk_sz_weights(w, om)
    assert np.abs(sz - szo).max() <= 1e-15 * szo.max() * 10
    assert abs(sc - sco) <= 1e-12


@pytest.mark.parametrize("nm,n", [(1, 100), (7, 3000), (20, 4096), (40, 2500), (64, 1500)])
def test_triple_overlap_matches_oracle(cy3q, oracle, nm, n):
    # nm = 20 fits one part of 64 tiles of the pair-product GEMM; 40 and 64 modes split into
    # about 5 and 17 parts, and every n here leaves a ragged last stage of 16 points.
    print(f"case: triple overlap, {nm} modes, {n} points", flush=True)
    from papers_amd import galerkin as G
    rng = np.random.default_rng(nm)
    modes = _rand_c(rng, nm, n)
    om = _rand_c(rng, n)
    w = rng.uniform(0.1, 1.0, n)
    hd = rng.uniform(0.5, 2.0, nm)
    w[3] = np.nan
    om[8] = np.inf
    modes[nm - 1, 5] = np.nan
    y = G.triple_overlap(modes, om, w, hd)
    yo = oracle.yuk_triple(modes, om, w, hd)
    assert _relmax(y, yo) <= 1e-12


def test_yukawa_couplings_match_oracle(cy3q, oracle):
    from papers_amd import galerkin as G
    rng = np.random.default_rng(12)
    nm, n = 4, 1500
    modes = _rand_c(rng, nm, n)
    om = _rand_c(rng, n)
    w = rng.uniform(0.1, 1.0, n)
    r = G.compute_yukawa_couplings(modes, om, w, None, n_bootstrap=6, seed=99)
    yo, unco, sco, cvo = oracle.yuk_couplings(modes, om, w, None, 6, 99)
    assert _relmax(r.couplings, yo) <= 1e-12
    assert np.abs(r.couplings_uncertainty - unco).max() <= 1e-10 * np.abs(yo).max()
    assert abs(r.quadrature_uniformity_score - sco) <= 1e-12
    assert abs(r.convergence_ratio - cvo) <= 1e-10


def test_solver_random_modes_match_oracle(cy3q, oracle):
    from papers_amd import galerkin as G
    P, s, _ = _quintic(cy3q, oracle, 2, 3000, 5, steps=0)
    y, coef = G.solver_yukawa_random_modes(s, degree=5, n_modes=6, coeff_seed=42)
    # coefficients: Box-Muller normals of the lib.rs LCG, b-major, (g0 + i g1)/sqrt(2)
    nsec, nm = coef.shape
    _, u, _ = oracle.lcg_stream(42, 4 * nsec * nm)
    g = np.sqrt(-2.0 * np.log(np.maximum(u[0::2] + 1e-12, 1e-12))) * np.cos(2 * np.pi * u[1::2])
    cref = ((g[0::2] + 1j * g[1::2]) / np.sqrt(2)).reshape(nsec, nm)
    assert np.abs(coef - cref).max() <= 1e-14
    sv5, _ = oracle.sections(P.pts, oracle.monomials(5, False), derivs=False)
    S = sv5[:, 0::2] + 1j * sv5[:, 1::2]
    modes = (S @ coef).T
    ok = np.isfinite(P.w) & (P.w > 0) & np.isfinite(P.log_om)
    c = np.where(ok, P.w * np.exp(np.where(ok, P.log_om, 0.0)), np.nan)
    yo = oracle.yuk_triple(modes, np.ones(P.n, complex), c)
    t = G.triple_index(nm)
    assert y.shape[0] == len(t) == 56
    assert _relmax(y, yo[t[:, 0], t[:, 1], t[:, 2]]) <= 1e-11
