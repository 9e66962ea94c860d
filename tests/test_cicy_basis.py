"""The product's Groebner-reduced section bases (cy3q_cicy_basis, host code of libcy3q.so, no
GPU) against the oracle's (Buchberger over exact rationals in oracle/cicy_oracle.py), bit for
bit: the same monomials in the reference's enumeration order (ty_metric.rs:792-895,
schoen_metric.rs:1185-1212). The TY k = 3 and Schoen (3, 3, 1) sizes, 87 and 27, are the
reference's own (determinism golden file)."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def co(oracle):
    from oracle import cicy_oracle as CO
    return CO


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6])
def test_ty_basis_matches_oracle(cy3q, co, k):
    from papers_amd import cy3_metric as CM
    got = CM.invariant_reduced_basis(CM.CICY_TY, [k])
    ref = co.ty_basis(k)
    assert got.dtype == np.uint32 and np.array_equal(got, ref)
    if k == 3:
        assert len(got) == 87


@pytest.mark.parametrize("d", [(3, 3, 1), (2, 2, 1), (3, 3, 2), (4, 4, 1), (6, 6, 3), (6, 3, 3)])
def test_schoen_basis_matches_oracle(cy3q, co, d):
    from papers_amd import cy3_metric as CM
    got = CM.invariant_reduced_basis(CM.CICY_SCHOEN, d)
    assert np.array_equal(got, co.schoen_basis(*d))
    if d == (3, 3, 1):
        assert len(got) == 27


def test_bad_arguments_fail_before_the_gpu(cy3q):
    from papers_amd import cy3_metric as CM
    from papers_amd.quintic import Cy3qError
    with pytest.raises(Cy3qError):
        CM.invariant_reduced_basis(5, [3])
    with pytest.raises(Cy3qError):
        CM.sample_points(CM.CICY_SCHOEN, 1, 10, mode="lanes")
    with pytest.raises(Cy3qError):
        CM.sample_points(CM.CICY_TY, 1, 0)
    with pytest.raises(Cy3qError):
        CM.affine_chart_frames(CM.CICY_TY, np.zeros((2, 16)), np.array([[0, 5, -1], [0, 0, -1]]))
