"""Split-precision (fp16 hi/lo MFMA) T-operator against the FP64 path and the oracle.

Tolerances follow the stated precision target: after the iterations, G and sigma within
1e-6 relative of the FP64 values on the same points.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("k,quotient,n", [(3, False, 50000), (5, True, 60000), (6, True, 40000)])
def test_split16_step_matches_fp64(cy3q, k, quotient, n):
    a = cy3q.QuinticSolver.new_with_sampler(k, n, 42, quotient=quotient)
    b = cy3q.QuinticSolver.new_with_sampler(k, n, 42, quotient=quotient, precision=cy3q.PREC_SPLIT16)
    for s in (a, b):
        s.orthonormalise_basis_fs_gram()
    for it in range(3):
        ra = a.donaldson_step()
        rb = b.donaldson_step()
        assert _rel(b.last_t(), a.last_t()) < 2e-6, it
        assert _rel(b.h_block, a.h_block) < 2e-6, it
        assert abs(rb - ra) <= 1e-5 * ra + 1e-12
    sa, sb = a.sigma(), b.sigma()
    assert abs(sb / sa - 1) < 1e-6


def test_split16_matches_oracle_one_step(cy3q, oracle):
    P = oracle.Problem(4, 30000, 13)
    s = cy3q.QuinticSolver.from_points(4, P.pts, precision=cy3q.PREC_SPLIT16)
    s.orthonormalise_basis_fs_gram()
    h, _ = oracle.donaldson_step(P.sv, P.w, P.h)
    h2, _ = oracle.donaldson_step(P.sv, P.w, h)
    s.donaldson_step()
    assert _rel(s.h_block, h) < 2e-6
    s.donaldson_step()
    assert _rel(s.h_block, h2) < 2e-6


def test_split16_switch_precision_midway(cy3q):
    # The FP64 and split paths share the same h state: switching between them mid-solve is valid.
    a = cy3q.QuinticSolver.new_with_sampler(3, 30000, 9)
    a.orthonormalise_basis_fs_gram()
    a.donaldson_solve(3, 0.0)
    a.set_precision(cy3q.PREC_SPLIT16)
    a.donaldson_solve(3, 0.0)
    a.set_precision(cy3q.PREC_FP64)
    a.donaldson_solve(2, 0.0)
    b = cy3q.QuinticSolver.new_with_sampler(3, 30000, 9)
    b.orthonormalise_basis_fs_gram()
    b.donaldson_solve(8, 0.0)
    assert _rel(a.h_block, b.h_block) < 1e-6
