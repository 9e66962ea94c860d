"""The C-ABI library loads and exports every symbol include/cy3q.h declares (no GPU calls)."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(cy3q_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_header_symbols(cy3q):
    lib = ctypes.CDLL(cy3q.lib_path())
    names = _declared()
    assert len(names) > 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_options(cy3q):
    from papers_amd import quintic as q
    L = q._load()
    assert b"gfx950" in L.cy3q_version()
    o = q._options()
    assert (o.quotient, o.precision, o.device, o.dk_iters, o.newton_polish) == (0, 0, 0, 50, 0)


def test_bad_arguments_fail_loudly(cy3q):
    # Argument checks run before any device call.
    from papers_amd import quintic as q
    import pytest
    with pytest.raises(q.Cy3qError):
        q.QuinticSolver.new_with_sampler(3, 0, 1)
    with pytest.raises(q.Cy3qError):
        q.QuinticSolver._new(3, 100, 1, 0.5, q.SAMPLER_SZ)
    # The refinement drivers check their handle before touching the device.
    L = q._load()
    import ctypes as C
    best = C.c_double()
    assert L.cy3q_solver_sigma_refine_lbfgs(None, 5, 4, None, None) == -1
    assert L.cy3q_solver_sigma_refine_analytic_restarts(None, 1, 1, 1e-3, 0.01, 1,
                                                        C.byref(best), None) == -1
    assert L.cy3q_solver_sigma_refine_lbfgs_restarts(None, 1, 1, 4, 0.01, 1, C.byref(best),
                                                     None) == -1
    assert L.cy3q_solver_set_refine_kind(None, cy3q.REFINE_T_FP64) == -1
    from papers_amd.donaldson_gpu import _lib as top_lib
    L = top_lib()  # binds the cy3q_top_* signatures
    out = C.c_void_p()
    assert L.cy3q_top_new(0, 4, 0, C.byref(out)) == -1
    assert L.cy3q_top_new(10, 0, 0, C.byref(out)) == -1
    assert L.cy3q_top_upload_static(None, None, None) == -1
    assert L.cy3q_top_t_operator_raw(None, None, None, None, None) == -1
    assert L.cy3q_sig8_new(0, 4, 0, C.byref(out)) == -1
    assert L.cy3q_sig8_upload_static(None, None, None, None, None, None) == -1
    sig = C.c_double()
    assert L.cy3q_sig8_compute_sigma(None, None, None, C.byref(sig)) == -1


def test_laplacian_basis_table_bit_exact(cy3q, oracle):
    # build_test_basis (metric_laplacian.rs:87-115) is host-side table code: bit-exact vs the oracle.
    from papers_amd import galerkin as G
    for deg in range(0, 6):
        assert np.array_equal(G.build_test_basis(deg), oracle.lap_basis(deg))


def test_laplacian_config_defaults(cy3q):
    # MetricLaplacianConfig::default (metric_laplacian.rs:277-290)
    from papers_amd import galerkin as G
    c = G.LaplacianConfig.default()
    assert (c.max_total_degree, c.n_low_eigenvalues, c.jacobi_max_sweeps) == (4, 15, 128)
    assert (c.jacobi_tol, c.mass_regularisation) == (1e-12, 1e-10)


def test_galerkin_bad_arguments_fail_loudly(cy3q):
    # Argument checks run before any device call.
    import pytest
    from papers_amd import galerkin as G
    from papers_amd import quintic as q
    with pytest.raises(q.Cy3qError):
        G.triple_overlap(np.zeros((65, 4), complex), np.ones(4, complex), np.ones(4))
    with pytest.raises(q.Cy3qError):
        G.compute_metric_laplacian_spectrum(np.zeros((4, 8), complex), np.ones(4),
                                            np.ones(4, complex),
                                            G.LaplacianConfig.default(max_total_degree=7))


def test_batch_specs_from_lcg(cy3q, oracle):
    # psi = 0.9 u1, k = 2 + floor(5 u2), N_p = round(exp(ln 2e4 + u3 ln 25)) from the lib.rs LCG.
    import math
    from papers_amd.batch import draw_specs
    specs = draw_specs(7, 32)
    _, u, _ = oracle.lcg_stream(7, 96)
    for i, (k, n, seed, psi) in enumerate(specs):
        assert psi == 0.9 * u[3 * i]
        assert k == 2 + min(4, math.floor(5 * u[3 * i + 1]))
        assert n == round(math.exp(math.log(2e4) + u[3 * i + 2] * (math.log(5e5) - math.log(2e4))))
        assert seed == 42 + i and 2 <= k <= 6 and 20000 <= n <= 500000
