"""The route34 T-operator contract on the GPU (cy3q_top_*, Cy3DonaldsonGpu of
route34/cy3_donaldson_gpu.rs) against the oracle's restatement, on random section values with
an SPD h = A^H A + n I, as the reference's parity test builds them
(route34/tests/test_cy3_donaldson_gpu_parity.rs). Tolerance: 1e-12 relative (FP64 MFMA sums
in a fixed order against numpy's)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _case(seed, n, nb):
    rng = np.random.default_rng(seed)
    sv = rng.standard_normal((n, 2 * nb))
    w = rng.uniform(0.05, 2.0, n)
    a = rng.standard_normal((nb, nb)) + 1j * rng.standard_normal((nb, nb))
    h = a.conj().T @ a + nb * np.eye(nb)
    return sv, w, h


@pytest.mark.parametrize("n,nb", [(5000, 35), (3000, 87), (2000, 205)])
def test_t_operator_raw_matches_oracle(cy3q, oracle, n, nb):
    from papers_amd.donaldson_gpu import Cy3DonaldsonGpu
    print(f"t_operator_raw n={n} nb={nb}", flush=True)
    sv, w, h = _case(42 + nb, n, nb)
    g = Cy3DonaldsonGpu(n, nb)
    g.upload_static(sv, w)
    re, im = g.t_operator_raw(h.real, h.imag)
    rr, ri = oracle.t_operator_raw_route34(sv, w, h.real, h.imag)
    got, ref = re + 1j * im, (rr + 1j * ri).ravel()
    assert _rel(got, ref) < 1e-12
    # Repeated calls with new h reuse the static upload.
    re2, im2 = g.t_operator_raw(2.0 * h.real, 2.0 * h.imag)
    assert _rel(re2 + 1j * im2, ref / 2.0) < 1e-12
    g.close()


def test_t_operator_raw_skips_bad_weights(cy3q, oracle):
    from papers_amd.donaldson_gpu import Cy3DonaldsonGpu
    sv, w, h = _case(3, 1000, 20)
    w[::7] = np.nan
    w[3::11] = 0.0
    w[5::13] = -1.0
    g = Cy3DonaldsonGpu(1000, 20)
    g.upload_static(sv, w)
    re, im = g.t_operator_raw(h.real, h.imag)
    rr, ri = oracle.t_operator_raw_route34(sv, w, h.real, h.imag)
    assert np.all(np.isfinite(re)) and np.all(np.isfinite(im))
    assert _rel(re + 1j * im, (rr + 1j * ri).ravel()) < 1e-12


def test_t_operator_raw_errors(cy3q):
    from papers_amd.donaldson_gpu import Cy3DonaldsonGpu
    g = Cy3DonaldsonGpu(100, 4)
    with pytest.raises(cy3q.Cy3qError):
        g.t_operator_raw(np.eye(4), np.zeros((4, 4)))  # before upload_static
    rng = np.random.default_rng(1)
    g.upload_static(rng.standard_normal((100, 8)), np.ones(100))
    with pytest.raises(cy3q.Cy3qError):
        g.t_operator_raw(-np.eye(4), np.zeros((4, 4)))  # not positive definite
    with pytest.raises(ValueError):
        g.upload_static(np.zeros((10, 8)), np.ones(10))
    g.close()


def test_route34_convention_matches_solver_t(cy3q):
    # The solver's last T (quintic convention) at G equals the route34 h_pair at h = conj(G),
    # divided by sum w and trace-normalised.
    from papers_amd.donaldson_gpu import Cy3DonaldsonGpu
    s = cy3q.QuinticSolver.new_with_sampler(3, 6000, 17)
    s.orthonormalise_basis_fs_gram()
    s.donaldson_solve(2, 0.0)
    hb = s.h_block
    G = hb[0::2, 0::2] - 1j * hb[0::2, 1::2]
    sv, _ = s.sections()
    w = s.weights
    s.donaldson_step()
    t = s.last_t()
    g = Cy3DonaldsonGpu(s.n_points, s.n_basis)
    g.upload_static(sv[:, : 2 * s.n_basis], w)
    re, im = g.t_operator_raw(G.real, -G.imag)
    tr = (re + 1j * im).reshape(s.n_basis, s.n_basis) / np.sum(w[np.isfinite(w) & (w > 0)])
    tr *= s.n_basis / np.trace(tr).real
    assert _rel(tr, t) < 1e-12


def _sigma_case(seed, n, nb):
    rng = np.random.default_rng(seed)
    sv = rng.standard_normal((n, 2 * nb))
    sd = rng.standard_normal((n, 8, 2 * nb))
    frames = rng.standard_normal((n, 3, 8, 2))
    om = rng.uniform(0.2, 3.0, n)
    w = rng.uniform(0.05, 2.0, n)
    a = rng.standard_normal((nb, nb)) + 1j * rng.standard_normal((nb, nb))
    h = a.conj().T @ a / nb + np.eye(nb)
    return sv, sd, frames, om, w, h


@pytest.mark.parametrize("n,nb", [(3000, 20), (1500, 87)])
def test_sigma8_matches_oracle(cy3q, oracle, n, nb):
    from papers_amd.donaldson_gpu import Cy3SigmaGpu
    print(f"sigma8 n={n} nb={nb}", flush=True)
    sv, sd, frames, om, w, h = _sigma_case(7 + nb, n, nb)
    w[::17] = np.nan
    om[5::23] = 0.0
    g = Cy3SigmaGpu(n, nb)
    g.upload_static(sv, sd, frames, om, w)
    sig = g.compute_sigma(h.real, h.imag)
    ref, _ = oracle.sigma_route34(sv, sd, frames, om, w, h.real, h.imag)
    assert np.isfinite(ref) and abs(sig - ref) <= 1e-10 * abs(ref)
    g.close()


def test_sigma8_tiny_weights(cy3q, oracle):
    # Weights scaled by 1e-16 (a sum near 1e-13): the reference still returns sigma down to a
    # weight sum of 1e-30 (cy3_sigma_gpu.rs:514-519), and so does the contract. sigma is
    # invariant under a common weight scale.
    from papers_amd.donaldson_gpu import Cy3SigmaGpu
    print("sigma8 tiny weights n=900 nb=20", flush=True)
    sv, sd, frames, om, w, h = _sigma_case(41, 900, 20)
    g = Cy3SigmaGpu(900, 20)
    g.upload_static(sv, sd, frames, om, w * 1e-16)
    sig = g.compute_sigma(h.real, h.imag)
    ref, _ = oracle.sigma_route34(sv, sd, frames, om, w * 1e-16, h.real, h.imag)
    ref1, _ = oracle.sigma_route34(sv, sd, frames, om, w, h.real, h.imag)
    assert np.sum(w) * 1e-16 < 1e-12
    assert np.isfinite(sig) and abs(sig - ref) <= 1e-10 * abs(ref)
    assert abs(ref / ref1 - 1) < 1e-12
    g.upload_static(sv, sd, frames, om, w * 1e-40)
    assert np.isnan(g.compute_sigma(h.real, h.imag))
    g.close()


def test_sigma8_fubini_study_closed_form(cy3q):
    # Linear sections, H = I, frames orthonormal with T_a . z = 0: eta = 1 / (|z|^6 |Omega|^2).
    from papers_amd.donaldson_gpu import Cy3SigmaGpu
    import test_oracle_top as T
    rng = np.random.default_rng(9)
    sv, sd, frames, om, w, z = T._fs_case(rng, 500)
    g = Cy3SigmaGpu(500, 8)
    g.upload_static(sv, sd, frames, om, w)
    sig = g.compute_sigma(np.eye(8), np.zeros((8, 8)))
    eta = 1.0 / (np.sum(np.abs(z) ** 2, axis=1) ** 3 * om)
    kap = np.sum(w * eta) / np.sum(w)
    ref = np.sum(w * (eta / kap - 1) ** 2) / np.sum(w)
    assert abs(sig - ref) <= 1e-11 * ref


def test_route34_loop_on_gpu_matches_oracle_loop(cy3q, oracle):
    # The 8-coordinate Donaldson loop (papers_amd/metric8.py) driven by the GPU contracts
    # against the same loop driven by the oracle's T-operator and sigma: 6 iterations with
    # static damping 0.6 and a Tikhonov shift 1e-3.
    from papers_amd import metric8 as M8
    from papers_amd.donaldson_gpu import Cy3DonaldsonGpu, Cy3SigmaGpu
    n, nb = 2500, 16
    sv, sd, frames, om, w, _ = _sigma_case(31, n, nb)
    top = Cy3DonaldsonGpu(n, nb)
    top.upload_static(sv, w)
    sg = Cy3SigmaGpu(n, nb)
    sg.upload_static(sv, sd, frames, om, w)
    h0 = (np.eye(nb), np.zeros((nb, nb)))
    kw = dict(max_iter=6, tol=0.0, damping=0.6, tikhonov=1e-3)
    gpu = M8.solve_route34(top.t_operator_raw, sg.compute_sigma, *h0, w, **kw)
    ref = M8.solve_route34(lambda a, b: oracle.t_operator_raw_route34(sv, w, a, b),
                           lambda a, b: oracle.sigma_route34(sv, sd, frames, om, w, a, b)[0],
                           *h0, w, **kw)
    assert gpu["iterations_run"] == ref["iterations_run"]
    assert _rel(gpu["h_re"] + 1j * gpu["h_im"], ref["h_re"] + 1j * ref["h_im"]) < 1e-11
    # (damping < 1 scales the cap to 6 x 4 iterations; the late residuals reach the rounding
    # floor, so they compare to 1e-11 of the first one)
    r0 = ref["donaldson_history"][0]
    assert np.allclose(gpu["donaldson_history"], ref["donaldson_history"], rtol=1e-9,
                       atol=1e-11 * r0)
    assert np.allclose(gpu["sigma_history"], ref["sigma_history"], rtol=1e-9, atol=0)
    top.close()
    sg.close()
