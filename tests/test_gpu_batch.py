"""Batched instances: one captured hipGraph launch against the sequential solves and against
stand-alone solvers (bit for bit: the same kernels in the same order per instance), graph
replays, and the autotuner with its cache."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPECS = [(2, 3000, 1, 0.0), (3, 5000, 2, 0.3), (4, 4000, 3, 0.7), (6, 6000, 4, 0.1)]


def _hs(b):
    return [s.h_block.copy() for s in b.solvers]


def test_graph_equals_sequential_and_replays(cy3q):
    from papers_amd.batch import Batch
    b = Batch(SPECS)
    b.solve_sequential(4)
    hs = _hs(b)
    b.solve_graph(4)
    hg = _hs(b)
    for x, y in zip(hs, hg):
        assert np.array_equal(x, y)
    b.solve_graph(4)  # replay of the captured graph starts again from h = I
    for x, y in zip(hs, _hs(b)):
        assert np.array_equal(x, y)
    b.solve_graph(3)  # a new capture for a new iteration count
    h3 = _hs(b)
    b.solve_sequential(3)
    for x, y in zip(h3, _hs(b)):
        assert np.array_equal(x, y)
    b.close()


def test_batch_instance_matches_standalone_solver(cy3q):
    from papers_amd.batch import Batch
    b = Batch(SPECS[:2])
    b.solve_graph(5)
    k, n, seed, psi = SPECS[1]
    s = cy3q.QuinticSolver.new_with_psi_sz(k, n, seed, psi, quotient=True)
    s.orthonormalise_basis_fs_gram()
    s.set_precision(cy3q.PREC_SPLIT16)
    s.donaldson_solve(5, 0.0)
    assert np.array_equal(s.h_block, b.solvers[1].h_block)
    assert np.array_equal(s.points, b.solvers[1].points)
    b.close()


def test_autotune_cache_and_results(cy3q, tmp_path):
    from papers_amd.batch import Batch
    cache = str(tmp_path / "tune.txt")
    b = Batch(SPECS)
    n1 = b.autotune(cache)
    assert n1 == len(SPECS)
    lines = open(cache).read().split("\n")
    assert len([x for x in lines if x.strip()]) == len(SPECS)
    tun = b.tuning()
    assert all(t[0] in (1, 2, 4, 8) and t[1] in (256, 512, 1024, 2048) for t in tun)
    b2 = Batch(SPECS)
    assert b2.autotune(cache) == 0  # every shape found in the cache
    assert b2.tuning() == tun
    b2.solve_graph(4)
    hg = _hs(b2)
    b2.solve_sequential(4)
    for x, y in zip(hg, _hs(b2)):
        assert np.array_equal(x, y)
    # tuning changes only the summation order: sigma agrees with the untuned batch
    b.close()
    b3 = Batch(SPECS)
    b3.solve_graph(4)
    for s2, s3 in zip(b2.solvers, b3.solvers):
        assert abs(s2.sigma() / s3.sigma() - 1) < 1e-6
    b2.close()
    b3.close()
    assert os.path.exists(cache)
