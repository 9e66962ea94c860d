"""GPU parity of the cy3q hot path against the CPU FP64 oracle (FP64 path).

Every test calls the HIP path through the C ABI (papers_amd.quintic over libcy3q.so) and
compares with oracle/ on the same seeded inputs. Tolerances are written in each test.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_sampler_fermat_matches_oracle(cy3q, oracle):
    # sample_quintic_points_sz (quintic.rs:3633-3785): same seed, same point order and count.
    s = cy3q.QuinticSolver.new_with_sampler(3, 5000, 42)
    ref, _ = oracle.sample_sz(5000, 42)
    got = s.points
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) < 1e-12


def test_sampler_dwork_matches_oracle(cy3q, oracle):
    # sample_deformed_quintic_points_sz (quintic.rs:3810-3983) at psi = 0.5.
    s = cy3q.QuinticSolver.new_with_psi_sz(2, 3000, 7, 0.5)
    ref, _ = oracle.sample_sz(3000, 7, psi=0.5, dwork=True)
    got = s.points
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) < 1e-11


@pytest.mark.parametrize("k,quotient", [(3, False), (6, True)])
def test_weights_logomega_sections(cy3q, oracle, k, quotient):
    pts, _ = oracle.sample_sz(700, 3)
    s = cy3q.QuinticSolver.from_points(k, pts, quotient=quotient)
    w_ref, lo_ref, _ = oracle.weights_logomega(pts)
    np.testing.assert_allclose(s.weights, w_ref, rtol=1e-14, atol=0)
    np.testing.assert_allclose(s.log_omega_sq, lo_ref, rtol=1e-13, atol=1e-14)
    mono = oracle.monomials(k, quotient)
    sv_ref, sd_ref = oracle.sections(pts, mono)
    sv, sd = s.sections(derivs=True)
    # The sections are bit-exact apart from libm-free products: same product order.
    assert np.array_equal(sv, sv_ref)
    assert np.array_equal(sd, sd_ref)


def test_fs_gram_matches_oracle(cy3q, oracle):
    pts, _ = oracle.sample_sz(20000, 11)
    s = cy3q.QuinticSolver.from_points(3, pts)
    s.orthonormalise_basis_fs_gram()
    P = oracle.Problem.__new__(oracle.Problem)
    mono = oracle.monomials(3)
    sv_raw, sd_raw = oracle.sections(pts, mono)
    w, _, _ = oracle.weights_logomega(pts)
    sv_ref, sd_ref, _ = oracle.orthonormalise(sv_raw, sd_raw, w)
    sv, sd = s.sections(derivs=True)
    assert _rel(sv, sv_ref) < 1e-12
    assert _rel(sd, sd_ref) < 1e-12


def _oracle_problem(oracle, k, n, seed, quotient=False):
    return oracle.Problem(k, n, seed, quotient=quotient)


def test_donaldson_steps_match_oracle(cy3q, oracle):
    P = _oracle_problem(oracle, 3, 20000, 5)
    s = cy3q.QuinticSolver.from_points(3, P.pts)
    s.orthonormalise_basis_fs_gram()
    h = P.h
    for it in range(3):
        h, r_ref = oracle.donaldson_step(P.sv, P.w, h)
        r = s.donaldson_step()
        # FP64 MFMA contractions + Cholesky inverse vs sequential sums + LU inverse.
        assert _rel(s.h_block, h) < 1e-12, it
        assert abs(r - r_ref) <= 1e-10 * max(r_ref, 1e-300) + 1e-14, it


def test_t_operator_matches_oracle(cy3q, oracle):
    P = _oracle_problem(oracle, 4, 8000, 9)
    s = cy3q.QuinticSolver.from_points(4, P.pts)
    s.orthonormalise_basis_fs_gram()
    _, _, (tre, tim, _) = oracle.donaldson_step(P.sv, P.w, P.h, want_t=True)
    s.donaldson_step()
    t = s.last_t()
    assert _rel(t, tre + 1j * tim) < 1e-13
    # Hermitian and trace-normalised to N (quintic.rs:1790-1812).
    assert np.allclose(t, t.conj().T, rtol=0, atol=1e-15 * np.abs(t).max())
    assert abs(np.trace(t).real - P.nb) < 1e-11


def test_t_operator_k6_quotient_matches_oracle(cy3q, oracle):
    # k=6 with the quotient basis (2N = 410, padded to M = 416) runs the FP64 T pass with the
    # compile-time wave programs (k_syrk13_f64), and the FS-Gram through the same kernel.
    P = _oracle_problem(oracle, 6, 4000, 13, quotient=True)
    s = cy3q.QuinticSolver.from_points(6, P.pts, quotient=True)
    s.orthonormalise_basis_fs_gram()
    sv, _ = s.sections()
    assert _rel(sv, P.sv) < 1e-11
    h1, r_ref, (tre, tim, _) = oracle.donaldson_step(P.sv, P.w, P.h, want_t=True)
    r = s.donaldson_step()
    assert _rel(s.last_t(), tre + 1j * tim) < 1e-12
    assert _rel(s.h_block, h1) < 1e-11
    assert abs(r - r_ref) <= 1e-9 * r_ref


def test_k6_quotient_fp64_steps_match_oracle(cy3q, oracle):
    """Three FP64 steps at M = 416 from the FS-Gram h: after the first step W is a general lower
    triangular factor, so every column tile's finished stages (whose B columns the FP64 D pass
    no longer copies) would show up here if a tile read them."""
    print("k=6 quotient FP64 steps, 4001 points", flush=True)
    P = _oracle_problem(oracle, 6, 4001, 21, quotient=True)
    s = cy3q.QuinticSolver.from_points(6, P.pts, quotient=True)
    s.orthonormalise_basis_fs_gram()
    h = P.h
    for it in range(3):
        h, r_ref = oracle.donaldson_step(P.sv, P.w, h)
        r = s.donaldson_step()
        assert _rel(s.h_block, h) < 1e-11, it
        assert abs(r - r_ref) <= 1e-9 * max(r_ref, 1e-300) + 1e-14, it


def test_sigma_matches_oracle(cy3q, oracle):
    P = _oracle_problem(oracle, 3, 20000, 5)
    s = cy3q.QuinticSolver.from_points(3, P.pts)
    s.orthonormalise_basis_fs_gram()
    eta_ref = oracle.eta(P.pts, P.sv, P.sd, P.log_om, P.h)
    sig_ref, kap_ref, l2_ref = oracle.sigma_reduce(eta_ref, P.w)
    sig, kap, l2, eta = s.sigma_ext(want_eta=True)
    ok = np.isfinite(eta_ref)
    assert np.array_equal(ok, np.isfinite(eta))
    np.testing.assert_allclose(eta[ok], eta_ref[ok], rtol=1e-11)
    assert abs(sig / sig_ref - 1) < 1e-12
    assert abs(kap / kap_ref - 1) < 1e-12
    assert abs(l2 / l2_ref - 1) < 1e-12
    h, _ = oracle.donaldson_step(P.sv, P.w, P.h)
    s.donaldson_step()
    assert abs(s.sigma() / P.sigma(h) - 1) < 1e-11


def _published_cases():
    import json
    import os

    path = os.path.join(os.path.dirname(__file__), "golden", "published_sigma_p5_9.json")
    with open(path) as f:
        d = json.load(f)
    return d, d["cases"]


_PUB, _PUB_CASES = _published_cases()


@pytest.mark.parametrize("case", _PUB_CASES, ids=lambda c: f"k{c['k']}-seed{c['seed']}")
def test_published_sigma(cy3q, case):
    # The reference's published p5_9 run (golden/published_sigma_p5_9.json): 200k SZ points,
    # sigma at the FS-Gram identity metric, then donaldson_solve(50, 1e-12) with its iteration
    # count and sigma. The FP64 GPU path reproduces all ten (k, seed) records.
    s = cy3q.QuinticSolver.new_with_sampler(case["k"], _PUB["n_pts"], case["seed"])
    assert s.n_points == _PUB["n_pts"] and s.n_basis == case["n_basis"]
    s.orthonormalise_basis_fs_gram()
    assert abs(s.sigma() / case["sigma_fs_identity"] - 1) < 1e-11
    iters = s.donaldson_solve(_PUB["donaldson_max_iter"], _PUB["donaldson_tol"])
    assert iters == case["donaldson_iters_run"]
    assert abs(s.sigma() / case["sigma_donaldson"] - 1) < 1e-10


def test_raw_basis_k6_gram_not_pd(cy3q):
    # The raw k=6 basis has the 5 relations z0^5 z_j = -(z1^5+..+z4^5) z_j on the variety, so
    # its FS Gram is singular: the reference's k=6 run failed here
    # (references/p3_10_kscan_results.md:42). The quotient basis (N=205) factors.
    s = cy3q.QuinticSolver.new_with_sampler(6, 20000, 42)
    with pytest.raises(ValueError):
        s.orthonormalise_basis_fs_gram()
    q = cy3q.QuinticSolver.new_with_sampler(6, 20000, 42, quotient=True)
    q.orthonormalise_basis_fs_gram()
    assert q.n_basis == 205


def test_degenerate_points_are_skipped(cy3q, oracle):
    # A point with z = 0 has grad f = 0: weight 0 and log|Omega|^2 = NaN, so the reference
    # skips it in T and sigma (quintic.rs:1322-1326, 1775-1777, 3520-3525).
    pts, _ = oracle.sample_sz(3000, 21)
    pts = np.concatenate([pts[:1500], np.zeros((1, 10)), pts[1500:]])
    s = cy3q.QuinticSolver.from_points(2, pts)
    w = s.weights
    assert w[1500] == 0.0 and np.isnan(s.log_omega_sq[1500])
    mono = oracle.monomials(2)
    sv, sd = oracle.sections(pts, mono)
    ww, lo, _ = oracle.weights_logomega(pts)
    h1, _ = oracle.donaldson_step(sv, ww, np.eye(2 * 15))
    s.donaldson_step()
    assert _rel(s.h_block, h1) < 1e-12
    eta_ref = oracle.eta(pts, sv, sd, lo, h1)
    assert abs(s.sigma() / oracle.sigma_reduce(eta_ref, ww)[0] - 1) < 1e-11


def test_run_twice_bit_identical(cy3q):
    # Fixed-order reductions everywhere: two runs give identical bits.
    out = []
    for _ in range(2):
        s = cy3q.QuinticSolver.new_with_sampler(4, 30000, 3)
        s.orthonormalise_basis_fs_gram()
        s.donaldson_solve(3, 0.0)
        out.append((s.h_block, s.sigma()))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


@pytest.mark.parametrize("weighted", [True, False])
def test_legacy_iteration_matches_oracle(cy3q, oracle, weighted):
    # donaldson_solve_quintic(_weighted) (quintic.rs:4224-4357): h and the residual history.
    P = oracle.Problem(3, 3000, 21)
    s = cy3q.QuinticSolver.from_points(3, P.pts)
    s.orthonormalise_basis_fs_gram()
    res = s.donaldson_solve_legacy(6, 0.0, weighted=weighted)
    h_ref, res_ref = oracle.legacy_solve(P.sv, 6, 0.0, P.w if weighted else None)
    assert len(res) == 6
    assert np.abs(np.array(res) - np.array(res_ref)).max() <= 1e-10 * max(res_ref)
    assert np.linalg.norm(s.h_block - h_ref) / np.linalg.norm(h_ref) < 1e-11


@pytest.mark.parametrize("psi", [0.0, 0.5])
def test_sampler_newton_polish_matches_oracle(cy3q, oracle, psi):
    # cy3q_options.newton_polish: FP64 Newton passes on each DK root (the north star's root
    # polish; the reference has none). After 50 DK sweeps two passes move no point beyond
    # rounding, so the accepted set and its order are the reference's.
    print(f"case: SZ sampler psi={psi}, 5000 points seed 42, newton_polish=2", flush=True)
    if psi == 0.0:
        s = cy3q.QuinticSolver.new_with_sampler(3, 5000, 42, newton_polish=2)
    else:
        s = cy3q.QuinticSolver.new_with_psi_sz(3, 5000, 42, psi, newton_polish=2)
    ref, _ = oracle.sample_sz(5000, 42, psi, polish=2)
    got = s.points
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) < 1e-12
    unpolished, _ = oracle.sample_sz(5000, 42, psi)
    assert np.max(np.abs(got - unpolished)) < 1e-12


def test_sampler_newton_polish_after_short_dk(cy3q, oracle):
    # With 12 DK sweeps, 4 polish passes converge most of the unconverged roots on the GPU as in
    # the oracle (the few near-collisions DK leaves stay as they are).
    print("case: SZ sampler, dk_iters=12, newton_polish=4, 10000 points", flush=True)
    a = lambda f: np.abs(f[:, 0] + 1j * f[:, 1])  # noqa: E731
    s = cy3q.QuinticSolver.new_with_sampler(3, 10000, 42, dk_iters=12, newton_polish=4)
    got = s.points
    ref, _ = oracle.sample_sz(10000, 42, dk_iters=12, polish=4)
    fg, _ = oracle.quintic_eval(got)
    fr, _ = oracle.quintic_eval(ref)
    bad_g, bad_r = np.sum(a(fg) > 1e-10), np.sum(a(fr) > 1e-10)
    assert abs(int(bad_g) - int(bad_r)) <= max(5, bad_r // 10)
    assert np.quantile(a(fg), 0.9) < 1e-14
