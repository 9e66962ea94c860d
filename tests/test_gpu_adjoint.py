"""sigma_L2^2 gradient (adjoint kernels), Adam refinement and the torch wrapper on the GPU,
against the oracle and finite differences. Tolerances are written in each test."""
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


@pytest.mark.parametrize("k,n,seed,quotient", [(3, 3000, 7, False), (4, 2000, 11, False),
                                               (5, 1200, 3, True)])
def test_sigma_sq_grad_matches_oracle(cy3q, oracle, k, n, seed, quotient):
    # sigma_squared_and_gradient (quintic.rs:2714-2860): value to 1e-11, gradient to 1e-9 in
    # norm (FP64 on both sides; the GPU sums the contractions in a different order).
    P, s, h = _setup(cy3q, oracle, k, n, seed, quotient)
    s2_ref, g_ref, s1_ref = P.sigma_sq_grad(h)
    s2, g, s1 = s.sigma_squared_and_gradient()
    assert abs(s2 / s2_ref - 1) < 1e-11
    assert abs(s1 / s1_ref - 1) < 1e-11
    assert _rel(g, g_ref) < 1e-9
    nb = s.n_basis
    assert np.all(np.tril(g[nb * nb:].reshape(nb, nb)) == 0.0)


def test_sigma_sq_grad_finite_differences_gpu(cy3q, oracle):
    # The GPU gradient against central differences of the GPU sigma_L2^2 (step 1e-6).
    P, s, h = _setup(cy3q, oracle, 3, 4000, 5)
    nb = s.n_basis
    _, g, _ = s.sigma_squared_and_gradient()
    gre, gim = g[: nb * nb].reshape(nb, nb), g[nb * nb:].reshape(nb, nb)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((nb, nb))
    X = 0.5 * (X + X.T)
    Y = rng.standard_normal((nb, nb))
    Y = 0.5 * (Y - Y.T)
    D = np.zeros((2 * nb, 2 * nb))
    D[0::2, 0::2] = X
    D[0::2, 1::2] = -Y
    D[1::2, 0::2] = Y
    D[1::2, 1::2] = X
    eps = 1e-6
    s.h_block = h + eps * D
    fp = s.sigma_ext()[2]
    s.h_block = h - eps * D
    fm = s.sigma_ext()[2]
    fd = (fp - fm) / (2 * eps)
    iu, iu1 = np.triu_indices(nb, 0), np.triu_indices(nb, 1)
    an = np.sum(gre[iu] * X[iu]) + np.sum(gim[iu1] * Y[iu1])
    assert abs(fd - an) < 1e-6 * abs(an), (fd, an)


def test_sigma_refine_analytic_matches_oracle(cy3q, oracle):
    # sigma_refine_analytic (quintic.rs:453-514): Adam from zero moments, 4 steps at lr 1e-3.
    P, s, h = _setup(cy3q, oracle, 3, 2500, 9)
    hist_ref, h_ref = oracle.sigma_refine_analytic(P, h, 4, 1e-3)
    hist = s.sigma_refine_analytic(4, 1e-3)
    assert len(hist) == 4
    assert np.allclose(hist, hist_ref, rtol=1e-10, atol=0)
    assert _rel(s.h_block, h_ref) < 1e-9
    assert hist[-1] < hist[0]


def test_torch_autograd_sigma_squared(cy3q, oracle):
    # The autograd wrapper's gradient, pulled back through a Cholesky parametrisation
    # G = L L^H, against central differences of the forward along a random direction of L.
    torch = pytest.importorskip("torch")
    from papers_amd.torch_sigma import sigma_squared

    P, s, h = _setup(cy3q, oracle, 3, 3000, 13)
    nb = s.n_basis
    G = h[0::2, 0::2] + 1j * h[1::2, 0::2]
    L0 = np.linalg.cholesky(G)
    lr = torch.tensor(L0.real, dtype=torch.float64, requires_grad=True)
    li = torch.tensor(L0.imag, dtype=torch.float64, requires_grad=True)

    def f(lr_, li_):
        Lc = torch.complex(torch.tril(lr_), torch.tril(li_))
        Gc = Lc @ Lc.conj().T
        return sigma_squared(Gc.real, Gc.imag, s)

    val = f(lr, li)
    val.backward()
    rng = np.random.default_rng(4)
    dr = torch.tensor(np.tril(rng.standard_normal((nb, nb))), dtype=torch.float64)
    di = torch.tensor(np.tril(rng.standard_normal((nb, nb)), -1), dtype=torch.float64)
    an = float((lr.grad * dr).sum() + (li.grad * di).sum())
    eps = 1e-6
    with torch.no_grad():
        fp = float(f(lr + eps * dr, li + eps * di))
        fm = float(f(lr - eps * dr, li - eps * di))
    fd = (fp - fm) / (2 * eps)
    assert abs(fd - an) < 1e-6 * abs(an), (fd, an)
