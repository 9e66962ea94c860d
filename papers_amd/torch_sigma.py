"""sigma_L2^2 as a torch.autograd.Function over the Hermitian metric G = G_re + i G_im.

forward runs the solver's GPU sigma pass; backward returns the analytic gradient of
sigma_squared_and_gradient (cy3/src/quintic.rs:2714-2860), computed by the adjoint kernels in
the same call. The solver depends only on the Hermitian part of G, so for full-matrix inputs
the gradient is spread over (a, b) and (b, a):
    dL/dG_re[a, b] = 1/2 g_re(a, b) (a != b),   dL/dG_re[a, a] = g_re(a, a),
    dL/dG_im[a, b] = 1/2 g_im(a, b),            dL/dG_im[b, a] = -1/2 g_im(a, b)  (a < b),
where g_re, g_im are the reference's per-parameter derivatives (symmetric pairs, diagonal
halved). A caller can therefore parametrise G freely (for example by a Cholesky factor,
G = L L^H) and let autograd apply the chain rule.
"""
from __future__ import annotations

import numpy as np
import torch


def _block_from(g_re: np.ndarray, g_im: np.ndarray) -> np.ndarray:
    nb = g_re.shape[0]
    h = np.zeros((2 * nb, 2 * nb))
    h[0::2, 0::2] = g_re
    h[0::2, 1::2] = -g_im
    h[1::2, 0::2] = g_im
    h[1::2, 1::2] = g_re
    return h


class SigmaSquared(torch.autograd.Function):
    """sigma_L2^2(G) for a QuinticSolver; inputs G_re, G_im ([N, N] float64 tensors)."""

    @staticmethod
    def forward(ctx, g_re: torch.Tensor, g_im: torch.Tensor, solver) -> torch.Tensor:
        gr = g_re.detach().to("cpu", torch.float64).numpy()
        gi = g_im.detach().to("cpu", torch.float64).numpy()
        solver.h_block = _block_from(gr, gi)
        s2, grad, _ = solver.sigma_squared_and_gradient()
        nb = gr.shape[0]
        dre = grad[: nb * nb].reshape(nb, nb)
        dim = grad[nb * nb:].reshape(nb, nb)
        off = ~np.eye(nb, dtype=bool)
        d_re = np.where(off, 0.5 * dre, dre)
        d_im = 0.5 * (np.triu(dim, 1) - np.triu(dim, 1).T)
        ctx.save_for_backward(torch.from_numpy(d_re).to(g_re.device, g_re.dtype),
                              torch.from_numpy(d_im).to(g_im.device, g_im.dtype))
        return torch.tensor(s2, dtype=g_re.dtype, device=g_re.device)

    @staticmethod
    def backward(ctx, grad_out):
        d_re, d_im = ctx.saved_tensors
        return grad_out * d_re, grad_out * d_im, None


def sigma_squared(g_re: torch.Tensor, g_im: torch.Tensor, solver) -> torch.Tensor:
    return SigmaSquared.apply(g_re, g_im, solver)
