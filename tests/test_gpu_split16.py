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



@pytest.mark.parametrize("n", [3001, 70001])
def test_split16_k6_block_counts(cy3q, n):
    # M = 416 runs the balanced D pass (k-steps j and 12 - j per barrier). 3001 points give 24
    # ragged blocks (one per workgroup); 70001 give 547, two or three per workgroup, so the
    # prefetch across blocks and the register-set swap of an odd block count both run.
    # The split error of T is a sum of independent per-point rounding errors, so its relative
    # size grows as 1/sqrt(n): the 2e-6 bound at 40k points becomes 2e-6 sqrt(40000 / n) below
    # that (7.3e-6 at 3001). A (k-step, column group) unit dealt twice or not at all would move
    # D_p by a share of one of 26 groups, percents, and T by orders of magnitude more.
    tol = 2e-6 * max(1.0, (40000.0 / n) ** 0.5)
    print(f"k=6 quotient split16 vs fp64, n={n}, tol={tol:.2e}", flush=True)
    a = cy3q.QuinticSolver.new_with_sampler(6, n, 11, quotient=True)
    b = cy3q.QuinticSolver.new_with_sampler(6, n, 11, quotient=True, precision=cy3q.PREC_SPLIT16)
    for s in (a, b):
        s.orthonormalise_basis_fs_gram()
    for it in range(2):
        a.donaldson_step()
        b.donaldson_step()
        assert _rel(b.last_t(), a.last_t()) < tol, it
        assert _rel(b.h_block, a.h_block) < tol, it

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
    a.set_refine(0)
    a.donaldson_solve(3, 0.0)
    a.set_precision(cy3q.PREC_FP64)
    a.donaldson_solve(2, 0.0)
    b = cy3q.QuinticSolver.new_with_sampler(3, 30000, 9)
    b.orthonormalise_basis_fs_gram()
    b.donaldson_solve(8, 0.0)
    assert _rel(a.h_block, b.h_block) < 1e-6


def test_refinement_sweep_is_one_fp64_step(cy3q):
    # A split-precision solve of K iterations with one whole-step refinement sweep equals K - 1
    # split iterations followed by one FP64 T-operator step, bit for bit (deterministic
    # reductions).
    mk = lambda: cy3q.QuinticSolver.new_with_sampler(4, 30000, 5, precision=cy3q.PREC_SPLIT16)
    a, b = mk(), mk()
    for s in (a, b):
        s.orthonormalise_basis_fs_gram()
    a.set_refine(1)
    a.set_refine_kind(cy3q.REFINE_STEP_FP64)
    assert a.donaldson_solve(4, 0.0) == 4
    b.set_refine(0)
    b.donaldson_solve(3, 0.0)
    b.set_precision(cy3q.PREC_FP64)
    r = b.donaldson_step()
    assert np.array_equal(a.h_block, b.h_block)
    assert a.donaldson_residuals[-1] == r
    # With an early-stop tolerance the sweep still ends the solve.
    c = mk()
    c.orthonormalise_basis_fs_gram()
    c.set_refine(1)
    it = c.donaldson_solve(30, 1e-4)
    assert 2 <= it < 30
    assert c.donaldson_residuals[-2] < 1e-4


def test_refinement_sweep_improves_fixed_point(cy3q):
    # After the split iterations have settled, two FP64 sweeps bring G closer to the FP64
    # iterate of the same index than two more split iterations do.
    a = cy3q.QuinticSolver.new_with_sampler(5, 60000, 11, quotient=True)
    a.orthonormalise_basis_fs_gram()
    a.donaldson_solve(42, 0.0)
    g42 = a.h_block
    b = cy3q.QuinticSolver.new_with_sampler(5, 60000, 11, quotient=True, precision=cy3q.PREC_SPLIT16)
    b.orthonormalise_basis_fs_gram()
    b.set_refine(0)
    b.donaldson_solve(40, 0.0)
    h40 = b.h_block
    b.donaldson_solve(2, 0.0)
    e_split = _rel(b.h_block, g42)
    b.h_block = h40
    b.set_refine(2)
    b.donaldson_solve(2, 0.0)
    e_ref = _rel(b.h_block, g42)
    assert e_ref < e_split


def test_refinement_kinds(cy3q):
    # Both sweep kinds move the settled split iterate towards the FP64 iterate of the same
    # index; the default (split D pass, FP64 T pass) is deterministic run to run and lands
    # within 1e-6 of the whole-step sweep.
    a = cy3q.QuinticSolver.new_with_sampler(5, 60000, 11, quotient=True)
    a.orthonormalise_basis_fs_gram()
    a.donaldson_solve(42, 0.0)
    g42 = a.h_block
    b = cy3q.QuinticSolver.new_with_sampler(5, 60000, 11, quotient=True, precision=cy3q.PREC_SPLIT16)
    b.orthonormalise_basis_fs_gram()
    b.set_refine(0)
    b.donaldson_solve(40, 0.0)
    h40 = b.h_block
    b.donaldson_solve(2, 0.0)
    e_split = _rel(b.h_block, g42)
    out = {}
    for kind in (cy3q.REFINE_T_FP64, cy3q.REFINE_T_FP64, cy3q.REFINE_STEP_FP64):
        b.h_block = h40
        b.set_refine(2)
        b.set_refine_kind(kind)
        b.donaldson_solve(2, 0.0)
        out.setdefault(kind, []).append(b.h_block)
    t64, step = out[cy3q.REFINE_T_FP64], out[cy3q.REFINE_STEP_FP64][0]
    assert np.array_equal(t64[0], t64[1])
    assert _rel(t64[0], g42) < e_split
    assert _rel(step, g42) < e_split
    assert _rel(t64[0], step) < 1e-6
    with pytest.raises(cy3q.Cy3qError):
        b.set_refine_kind(7)



@pytest.mark.parametrize("k,quotient,n", [(7, True, 40000), (8, True, 40000), (8, True, 3001)])
def test_split16_past_416_columns(cy3q, k, quotient, n):
    # 2N = 630 (M = 640) and 920 (M = 928): the D pass in column ranges whose partial sums are
    # added, the T pass in parts (the lower triangle of each 192-column segment and the
    # rectangle of each segment pair); 3001 points leave a ragged last block and chunk. A tile
    # or range dealt twice or not at all would move T by whole column blocks, far above the
    # split rounding (about 1e-6 relative at 40k points). (The raw k = 8 basis, M = 992, has
    # linearly dependent sections on the quintic, so its FS-Gram matrix is singular.)
    tol = 3e-6 * max(1.0, (40000.0 / n) ** 0.5)
    print(f"k={k} quotient={quotient} split16 vs fp64, n={n}, tol={tol:.2e}", flush=True)
    a = cy3q.QuinticSolver.new_with_sampler(k, n, 5, quotient=quotient)
    b = cy3q.QuinticSolver.new_with_sampler(k, n, 5, quotient=quotient, precision=cy3q.PREC_SPLIT16)
    for s in (a, b):
        s.orthonormalise_basis_fs_gram()
    for it in range(2):
        ra = a.donaldson_step()
        rb = b.donaldson_step()
        et, eh = _rel(b.last_t(), a.last_t()), _rel(b.h_block, a.h_block)
        print(f"  step {it}: T rel {et:.2e}, h rel {eh:.2e}, residual {ra:.6e} / {rb:.6e}",
              flush=True)
        assert et < tol, it
        assert eh < tol, it
        assert abs(rb - ra) <= 1e-5 * ra + 1e-12


def test_split16_k8_solve_with_refinement_sigma(cy3q):
    # The config-5 metric (Dwork psi = 0.5, k = 8 quotient) balanced in split precision with the
    # FP64 refinement sweep, against the FP64 solve: sigma within the 1e-6 target.
    n = 60000
    a = cy3q.QuinticSolver.new_with_psi_sz(8, n, 42, 0.5, quotient=True)
    b = cy3q.QuinticSolver.new_with_psi_sz(8, n, 42, 0.5, quotient=True,
                                           precision=cy3q.PREC_SPLIT16)
    for s in (a, b):
        s.orthonormalise_basis_fs_gram()
        s.donaldson_solve(6, 0.0)
    sa, sb = a.sigma(), b.sigma()
    print(f"k=8 psi=0.5: sigma fp64 {sa:.12e}, split+refine {sb:.12e}, rel {abs(sb / sa - 1):.2e}",
          flush=True)
    assert abs(sb / sa - 1) < 1e-6
    assert _rel(b.h_block, a.h_block) < 1e-5
