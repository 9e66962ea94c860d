"""The AdamGradientKernel contract (cy3q_adamgrad_*, gpu_adam.rs:1563-2194) on the GPU against
the oracle: sigma_L2^2 and its gradient from caller-uploaded arrays, and the reference's
per-point buffer (cpu_compute_per_point_reference, gpu_adam.rs:2677-2739). Tolerances are
written in each test."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _problem(oracle, k, n, seed, quotient=False):
    P = oracle.Problem(k, n, seed, quotient=quotient)
    h, _ = oracle.donaldson_step(P.sv, P.w, P.h)
    return P, h


def _kernel(cy3q, P):
    g = cy3q.AdamGradientKernel(P.n, P.nb)
    g.upload_static_inputs(P.sv, P.sd, P.pts, P.w, P.log_om)
    return g


@pytest.mark.parametrize("k,n,seed,quotient", [(3, 3000, 7, False), (5, 1500, 3, True)])
def test_adamgrad_sigma_grad_matches_oracle(cy3q, oracle, k, n, seed, quotient):
    print(f"case: k={k} quotient={quotient} n={n} seed={seed}", flush=True)
    P, h = _problem(oracle, k, n, seed, quotient)
    g = _kernel(cy3q, P)
    s2, grad, s1 = g.compute_sigma_grad_h_only(h)
    s2_ref, g_ref, s1_ref = P.sigma_sq_grad(h)
    assert abs(s2 / s2_ref - 1) < 1e-11
    assert abs(s1 / s1_ref - 1) < 1e-11
    assert _rel(grad, g_ref) < 1e-9
    # A second h through the same uploaded inputs.
    h2, _ = oracle.donaldson_step(P.sv, P.w, h)
    s2b, gb, _ = g.compute_sigma_grad_h_only(h2)
    s2b_ref, gb_ref, _ = P.sigma_sq_grad(h2)
    assert abs(s2b / s2b_ref - 1) < 1e-11
    assert _rel(gb, gb_ref) < 1e-9
    assert g.total_device_bytes() > 8 * P.n * 12 * P.nb


def test_adamgrad_per_point_matches_oracle(cy3q, oracle):
    print("case: per-point buffer, k=3, 2000 points seed 5, invalid rows", flush=True)
    P, h = _problem(oracle, 3, 2000, 5)
    w = P.w.copy()
    lo = P.log_om.copy()
    w[10] = 0.0
    w[11] = -1.0
    lo[12] = np.nan
    g = cy3q.AdamGradientKernel(P.n, P.nb)
    buf = g.compute_per_point(P.sv, P.sd, h, P.pts, w, lo)
    ref = oracle.per_point(P.pts, P.sv, P.sd, lo, w, h)
    assert buf.shape == ref.shape
    assert np.array_equal(np.isnan(buf[:, 0]), np.isnan(ref[:, 0]))
    ok = np.isfinite(ref[:, 0])
    np.testing.assert_allclose(buf[ok, 0], ref[ok, 0], rtol=1e-11)
    assert np.array_equal(buf[:, 1], ref[:, 1])
    scale = np.max(np.abs(ref[:, 2:]), axis=1, keepdims=True) + 1e-300
    assert np.max(np.abs(buf[:, 2:] - ref[:, 2:]) / scale) < 1e-9
    for p in (10, 11, 12):
        assert np.isnan(buf[p, 0]) and buf[p, 1] == 0.0 and not np.any(buf[p, 2:])
    # h_only after the upload done by compute_per_point gives the same buffer.
    buf2 = g.compute_per_point_h_only(h)
    assert np.array_equal(buf, buf2, equal_nan=True)


def test_adamgrad_per_point_chunks_and_recombination(cy3q, oracle):
    # k = 5 quotient (N = 125): rows of 31252 values, several 1 GiB chunks at 9000 points; the
    # recombined buffer gives the gradient of compute_sigma_grad_h_only.
    print("case: per-point buffer, k=5 quotient, 9000 points", flush=True)
    P, h = _problem(oracle, 5, 9000, 17, quotient=True)
    g = _kernel(cy3q, P)
    buf = g.compute_per_point_h_only(h)
    eta, w, gl = buf[:, 0], buf[:, 1], buf[:, 2:]
    ok = np.isfinite(eta) & (w != 0)
    tw = np.sum(w[ok])
    kappa = np.sum(w[ok] * eta[ok]) / tw
    dk = (w[ok] * eta[ok]) @ gl[ok] / tw
    en = eta[ok] / kappa
    grad = (w[ok] * 2 * (en - 1) * en) @ (gl[ok] - dk / kappa) / tw
    s2, g_dev, _ = g.compute_sigma_grad_h_only(h)
    assert _rel(grad, g_dev) < 1e-9


def test_adamgrad_errors(cy3q, oracle):
    P, h = _problem(oracle, 2, 200, 1)
    g = cy3q.AdamGradientKernel(P.n, P.nb)
    with pytest.raises(cy3q.Cy3qError):
        g.compute_sigma_grad_h_only(h)  # before upload_static_inputs
    g.upload_static_inputs(P.sv, P.sd, P.pts, P.w, P.log_om)
    with pytest.raises(cy3q.Cy3qError):
        g.compute_sigma_grad_h_only(-np.eye(2 * P.nb))  # not positive definite
    with pytest.raises(ValueError):
        g.compute_sigma_grad_h_only(np.eye(3))
    with pytest.raises(cy3q.Cy3qError):
        cy3q.AdamGradientKernel(0, 10)
