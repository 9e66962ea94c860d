"""The metric cache and the replay log around real GPU solves: a cached metric is the solver's
metric bit for bit, sigma re-measured from the cached H reproduces the stored value to rounding,
and a replay log of per-seed solves verifies."""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _direct(cy3q, cfg):
    s = cy3q.QuinticSolver.new_with_sampler(cfg.k_degree, cfg.n_sample, cfg.seed,
                                            sampler=cy3q.SAMPLER_SZ, quotient=cfg.apply_quotient)
    if cfg.precision == "split16":
        s.set_precision(cy3q.PREC_SPLIT16)
        s.set_refine(cfg.refine)
    it = s.donaldson_solve(cfg.max_iter, cfg.donaldson_tol)
    return s, it


@pytest.mark.parametrize("precision", ["fp64", "split16"])
def test_cached_metric_is_the_solver_metric(cy3q, tmp_path, precision):
    from papers_amd import metric_cache as mc
    cfg = mc.QuinticMetricConfig(k_degree=3, n_sample=4000, max_iter=8, donaldson_tol=1e-14,
                                 seed=31, precision=precision)
    print(f"metric cache solve: {cfg}", flush=True)
    # k=3 on 4000 points does not reach publication grade: strict mode refuses to cache it.
    with pytest.raises(mc.StaleQuality):
        mc.load_or_solve(str(tmp_path), cfg)
    r, out, path, _ = mc.load_or_solve(str(tmp_path), cfg, strict_quality=False)
    assert out == "miss" and os.path.exists(path)
    assert r.iterations_run == 8 and len(r.donaldson_history) == 8
    assert r.n_basis == 35 and len(r.basis_monomials) == 35
    assert r.sigma_history == [r.sigma_fs_identity, r.sigma_after_donaldson]
    assert r.final_sigma_residual == r.sigma_after_donaldson
    assert 0.0 < r.sigma_after_donaldson and 0.0 < r.sigma_fs_identity

    s, it = _direct(cy3q, cfg)
    assert it == r.iterations_run
    assert np.array_equal(s.h_block, r.h_block())
    assert np.array_equal(np.asarray(s.donaldson_residuals), np.asarray(r.donaldson_history))

    r2, out2, _, _ = mc.load_or_solve(str(tmp_path), cfg, strict_quality=False)
    assert out2 == "hit"
    assert r2.h_complex().tobytes() == r.h_complex().tobytes()
    # sigma of the cached H on the same points is the stored sigma. A fresh solver factors the
    # loaded H itself, where the solved one kept the factor of its last step, so the two agree
    # to rounding (tolerance 1e-13 relative), not bit for bit.
    t = cy3q.QuinticSolver.new_with_sampler(cfg.k_degree, cfg.n_sample, cfg.seed,
                                            sampler=cy3q.SAMPLER_SZ, quotient=True)
    t.h_block = r2.h_block()
    assert abs(t.sigma() - r2.sigma_after_donaldson) <= 1e-13 * r2.sigma_after_donaldson


def test_replog_of_per_seed_solves(cy3q, tmp_path):
    from papers_amd import metric_cache as mc
    from papers_amd import repro
    w = repro.ReplogWriter()
    config = {"k": 3, "n_pts": 3000, "donaldson_iters": 6, "seeds": [5, 6]}
    w.push(repro.run_start_event("test_gpu_formats", repro.ReproManifest.collect(), config))
    for seed in config["seeds"]:
        cfg = mc.QuinticMetricConfig(k_degree=3, n_sample=3000, max_iter=6,
                                     donaldson_tol=1e-14, seed=seed)
        r = mc.solve_quintic_metric(cfg)
        w.push(repro.per_seed_event(repro.PerSeedEvent(
            seed=seed, candidate="fermat_quintic", k=3, iters_run=r.iterations_run,
            final_residual=r.final_donaldson_residual, sigma_fs_identity=r.sigma_fs_identity,
            sigma_final=r.final_sigma_residual, n_basis=r.n_basis,
            elapsed_ms=1e3 * r.run_metadata["wall_clock_seconds"])))
        assert math.isfinite(r.final_sigma_residual)
    w.push(repro.run_end_event({"n_seeds": 2}, 1.0))
    p = tmp_path / "run.replog"
    w.write_to_path(str(p))
    assert repro.verify_replog(str(p)) == w.final_chain_hex()
    recs, trailer = repro.read_replog(str(p))
    assert trailer["n_events"] == 4 and recs[0][3]["manifest"]["rust_toolchain"].startswith("cy3q")
