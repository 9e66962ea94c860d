"""Content-addressed on-disk cache of balanced quintic metrics.

Restates the cache of the reference's ``route34/metric_cache.rs`` for this library's metric
solve (the Fermat/Dwork quintic pipeline of :mod:`papers_amd.quintic`):

* the key is a fixed, versioned record of every input that changes the metric (manifold and
  psi, basis degree, sample count, sampler and seed, iteration cap, tolerance, precision and
  refinement), written as compact JSON; floats enter by their IEEE-754 bit patterns, so ``1e-6``
  and the nearest double above it are different keys (``metric_cache.rs:131-160``);
* the file name is ``quintic_k{k}_n{n_sample}_i{max_iter}_seed{seed}_{h}.json`` where ``h`` is the
  first 16 hex digits of the SHA-256 of the key JSON (``metric_cache.rs:162-193``): the readable
  prefix identifies a file by eye, the hash makes every parameter change a miss;
* :func:`load_or_solve` returns a cached result on a hit, otherwise solves, writes the result to
  ``<name>.json.tmp`` and renames it into place, so a crash never leaves a partial file under the
  final name (``metric_cache.rs:222-262``);
* the quality guard rejects, in strict mode, a result whose final sigma residual is not finite
  or is at or above :data:`SIGMA_QUALITY_THRESHOLD` = 1e-3 (Donaldson 2009, section 5); in
  non-strict mode it warns on stderr and goes on (``metric_cache.rs:297-322``).

The stored result follows the reference's ``TyMetricResult`` layout (``ty_metric.rs:352-395``):
``balanced_h`` is the ``n_basis x n_basis`` Hermitian ``H`` row-major with real and imaginary
parts interleaved, ``balanced_h_re``/``balanced_h_im`` the same data split, then the residuals,
histories, basis monomials and run metadata. Floats are written in their shortest round-trip
form, so ``H`` reads back bit for bit.

The same module also computes the reference's own cache paths for the Tian-Yau and Schoen
configurations (:func:`ty_cache_path`, :func:`schoen_cache_path`) so that a cache directory
shared with the reference can be indexed.
"""
from __future__ import annotations

import json
import math
import os
import struct
import sys
import time
from dataclasses import dataclass, field, fields

import numpy as np

from .repro import sorted_value, to_json_bytes, sha256

__all__ = [
    "SIGMA_QUALITY_THRESHOLD", "SCHEMA_VERSION", "MetricCacheError", "StaleQuality",
    "QuinticMetricConfig", "QuinticMetricResult", "key_material", "hash_prefix",
    "quintic_cache_path", "ty_cache_path", "schoen_cache_path", "check_sigma_quality",
    "solve_quintic_metric", "load_or_solve", "save_result", "load_result", "f64_bits",
]

SIGMA_QUALITY_THRESHOLD = 1.0e-3
SCHEMA_VERSION = 1


class MetricCacheError(RuntimeError):
    pass


class StaleQuality(MetricCacheError):
    """MetricCacheError::StaleQuality (metric_cache.rs:78-82)."""

    def __init__(self, path: str, sigma: float, threshold: float):
        self.path, self.sigma, self.threshold = path, sigma, threshold
        super().__init__(
            f"metric_cache: cache file {path} has sigma = {sigma:.3e} >= threshold "
            f"{threshold:.3e}; publication-grade quality not satisfied; recompute fresh")


def f64_bits(x: float) -> int:
    """``f64::to_bits``: the IEEE-754 bit pattern of ``x`` as an unsigned 64-bit integer."""
    return struct.unpack("<Q", struct.pack("<d", float(x)))[0]


def hash_prefix(key: dict) -> str:
    """First 16 lower-case hex digits of the SHA-256 of the key's compact JSON."""
    return sha256(to_json_bytes(key)).hex()[:16]


# ---- this library's quintic cache -------------------------------------------------------------

_SAMPLERS = ("sz", "sz_dwork", "newton", "newton_dwork")
_PRECISIONS = ("fp64", "split16")


@dataclass
class QuinticMetricConfig:
    """Inputs of one quintic metric solve. ``sampler`` is one of "sz", "sz_dwork", "newton",
    "newton_dwork"; ``precision`` "fp64" or "split16" (with ``refine`` FP64 sweeps at the end)."""
    k_degree: int
    n_sample: int
    max_iter: int
    donaldson_tol: float
    seed: int
    psi: float = 0.0
    sampler: str = "sz"
    apply_quotient: bool = True
    precision: str = "fp64"
    refine: int = 1
    newton_tol: float = 1e-12
    orthonormalise_fs_gram: bool = False
    adam_iters: int = 0
    adam_lr: float = 1e-3
    store_points: bool = False
    device: int = 0

    def validate(self):
        if self.sampler not in _SAMPLERS:
            raise ValueError(f"sampler must be one of {_SAMPLERS}")
        if self.precision not in _PRECISIONS:
            raise ValueError(f"precision must be one of {_PRECISIONS}")
        if self.sampler in ("sz", "newton") and self.psi != 0.0:
            raise ValueError("the Fermat samplers need psi = 0; use sz_dwork / newton_dwork")
        if self.k_degree < 1 or self.n_sample < 1 or self.max_iter < 0:
            raise ValueError("k_degree and n_sample must be positive, max_iter non-negative")


def key_material(cfg: QuinticMetricConfig) -> dict:
    """The hashed key: every field that changes the balanced metric, in a fixed order. The
    device, ``store_points`` and the quality mode do not enter (they change neither H nor the
    residuals)."""
    manifold = "fermat_quintic" if cfg.psi == 0.0 and cfg.sampler in ("sz", "newton") \
        else "dwork_quintic"
    return {
        "schema_version": SCHEMA_VERSION,
        "manifold": manifold,
        "psi_bits": f64_bits(cfg.psi),
        "k_degree": int(cfg.k_degree),
        "n_sample": int(cfg.n_sample),
        "max_iter": int(cfg.max_iter),
        "donaldson_tol_bits": f64_bits(cfg.donaldson_tol),
        "seed": int(cfg.seed),
        "sampler": cfg.sampler,
        "newton_tol_bits": f64_bits(cfg.newton_tol) if cfg.sampler.startswith("newton") else 0,
        "apply_quotient": bool(cfg.apply_quotient),
        "precision": cfg.precision,
        "refine": int(cfg.refine) if cfg.precision != "fp64" else 0,
        "orthonormalise_fs_gram": bool(cfg.orthonormalise_fs_gram),
        "adam_iters": int(cfg.adam_iters),
        "adam_lr_bits": f64_bits(cfg.adam_lr) if cfg.adam_iters > 0 else 0,
    }


def quintic_cache_path(cache_root: str, cfg: QuinticMetricConfig) -> str:
    cfg.validate()
    name = (f"quintic_k{cfg.k_degree}_n{cfg.n_sample}_i{cfg.max_iter}_seed{cfg.seed}_"
            f"{hash_prefix(key_material(cfg))}.json")
    return os.path.join(cache_root, name)


# ---- the reference's TY / Schoen cache names (metric_cache.rs:131-215) ------------------------

def ty_cache_path(cache_root: str, k_degree: int, n_sample: int, max_iter: int,
                  donaldson_tol: float, seed: int, apply_z3_quotient: bool) -> str:
    key = {
        "schema_version": 1, "manifold": "tian_yau_z3", "k_degree": int(k_degree),
        "n_sample": int(n_sample), "max_iter": int(max_iter),
        "donaldson_tol_bits": f64_bits(donaldson_tol), "seed": int(seed),
        "apply_z3_quotient": bool(apply_z3_quotient),
    }
    return os.path.join(cache_root, f"ty_k{k_degree}_n{n_sample}_i{max_iter}_seed{seed}_"
                                    f"{hash_prefix(key)}.json")


def schoen_cache_path(cache_root: str, d_x: int, d_y: int, d_t: int, n_sample: int,
                      max_iter: int, donaldson_tol: float, seed: int,
                      apply_z3xz3_quotient: bool) -> str:
    key = {
        "schema_version": 1, "manifold": "schoen_z3xz3", "d_x": int(d_x), "d_y": int(d_y),
        "d_t": int(d_t), "n_sample": int(n_sample), "max_iter": int(max_iter),
        "donaldson_tol_bits": f64_bits(donaldson_tol), "seed": int(seed),
        "apply_z3xz3_quotient": bool(apply_z3xz3_quotient), "schoen_bundle_canonical": False,
    }
    return os.path.join(cache_root, f"schoen_dx{d_x}_dy{d_y}_dt{d_t}_n{n_sample}_i{max_iter}_"
                                    f"seed{seed}_{hash_prefix(key)}.json")


# ---- result record ----------------------------------------------------------------------------

@dataclass
class QuinticMetricResult:
    balanced_h: list
    balanced_h_re: list
    balanced_h_im: list
    n_basis: int
    sample_points: list
    final_sigma_residual: float
    sigma_fs_identity: float
    final_donaldson_residual: float
    iterations_run: int
    basis_monomials: list
    sigma_history: list
    donaldson_history: list
    run_metadata: dict
    sigma_after_donaldson: float
    sigma_after_adam: float = float("nan")
    adam_iters_run: int = 0
    adam_sigma_history: list = field(default_factory=list)

    def h_complex(self) -> np.ndarray:
        n = self.n_basis
        return (np.asarray(self.balanced_h_re, np.float64).reshape(n, n)
                + 1j * np.asarray(self.balanced_h_im, np.float64).reshape(n, n))

    def h_block(self) -> np.ndarray:
        """H in the solver's real 2N x 2N block form (``QuinticSolver.h_block``)."""
        n = self.n_basis
        hre = np.asarray(self.balanced_h_re, np.float64).reshape(n, n)
        him = np.asarray(self.balanced_h_im, np.float64).reshape(n, n)
        h = np.zeros((2 * n, 2 * n))
        h[0::2, 0::2] = hre
        h[0::2, 1::2] = -him
        h[1::2, 0::2] = him
        h[1::2, 1::2] = hre
        return h

    def to_dict(self) -> dict:
        d = {f.name: getattr(self, f.name) for f in fields(self)}
        d["run_metadata"] = sorted_value(d["run_metadata"])
        return d

    @classmethod
    def from_dict(cls, d: dict) -> "QuinticMetricResult":
        known = {f.name for f in fields(cls)}
        kw = {k: v for k, v in d.items() if k in known}
        for name in ("sigma_after_adam", "final_sigma_residual", "sigma_fs_identity",
                     "final_donaldson_residual", "sigma_after_donaldson"):
            if name in kw and kw[name] is None:  # non-finite floats are written as null
                kw[name] = float("nan")
        return cls(**kw)


def check_sigma_quality(path: str, sigma: float, strict: bool):
    """check_sigma_quality (metric_cache.rs:297-322)."""
    sigma = float("nan") if sigma is None else float(sigma)
    if math.isfinite(sigma) and sigma < SIGMA_QUALITY_THRESHOLD:
        return
    if strict:
        raise StaleQuality(path, sigma, SIGMA_QUALITY_THRESHOLD)
    print(f"WARNING: metric_cache: {path} has sigma = {sigma:.3e} >= Donaldson-2009 publication "
          f"threshold {SIGMA_QUALITY_THRESHOLD:.3e}. Cached / loaded anyway because "
          f"strict_quality=false.", file=sys.stderr)


def save_result(path: str, result: QuinticMetricResult):
    """Writes ``result`` as compact JSON to ``path + ".tmp"`` and renames it to ``path``."""
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(to_json_bytes(result.to_dict()))
    os.replace(tmp, path)


def load_result(path: str) -> QuinticMetricResult:
    with open(path, "rb") as f:
        return QuinticMetricResult.from_dict(json.loads(f.read()))


# ---- solve ------------------------------------------------------------------------------------

def solve_quintic_metric(cfg: QuinticMetricConfig) -> QuinticMetricResult:
    """One metric solve on the GPU: sample, (optionally) FS-Gram orthonormalise, Donaldson
    balance to ``donaldson_tol`` or ``max_iter`` iterations, (optionally) Adam sigma refinement.
    sigma is measured at the starting metric, after the Donaldson loop and after refinement."""
    from . import quintic as q

    cfg.validate()
    t0 = time.perf_counter()
    opts = dict(quotient=cfg.apply_quotient, device=cfg.device)
    if cfg.sampler == "sz":
        s = q.QuinticSolver.new_with_sampler(cfg.k_degree, cfg.n_sample, cfg.seed,
                                             sampler=q.SAMPLER_SZ, **opts)
    elif cfg.sampler == "sz_dwork":
        s = q.QuinticSolver.new_with_psi_sz(cfg.k_degree, cfg.n_sample, cfg.seed, cfg.psi, **opts)
    elif cfg.sampler == "newton":
        s = q.QuinticSolver.new(cfg.k_degree, cfg.n_sample, cfg.seed, cfg.newton_tol, **opts)
    else:
        s = q.QuinticSolver.new_with_psi(cfg.k_degree, cfg.n_sample, cfg.seed, cfg.newton_tol,
                                         cfg.psi, **opts)
    if s is None:
        raise MetricCacheError("the sampler produced no points")
    try:
        if cfg.orthonormalise_fs_gram:
            s.orthonormalise_basis_fs_gram()
        if cfg.precision == "split16":
            s.set_precision(q.PREC_SPLIT16)
            s.set_refine(cfg.refine)
        sigma0 = s.sigma()
        iters = s.donaldson_solve(cfg.max_iter, cfg.donaldson_tol)
        dhist = [float(r) for r in s.donaldson_residuals]
        sigma_d = s.sigma()
        adam_hist, sigma_a = [], float("nan")
        if cfg.adam_iters > 0:
            adam_hist = [float(x) for x in s.sigma_refine_analytic(cfg.adam_iters, cfg.adam_lr)]
            sigma_a = s.sigma()
        hb = s.h_block
        hre, him = hb[0::2, 0::2], -hb[0::2, 1::2]
        n = s.n_basis
        inter = np.empty((n, n, 2))
        inter[..., 0], inter[..., 1] = hre, him
        mons = q.build_degree_k_quintic_monomials(cfg.k_degree, cfg.apply_quotient)
        pts = s.points.tolist() if cfg.store_points else []
        final = sigma_a if cfg.adam_iters > 0 else sigma_d
        meta = {
            "seed": int(cfg.seed), "n_points": int(s.n_points), "psi": float(cfg.psi),
            "sampler": cfg.sampler, "precision": cfg.precision, "k_degree": int(cfg.k_degree),
            "max_iter": int(cfg.max_iter), "donaldson_tol": float(cfg.donaldson_tol),
            "wall_clock_seconds": time.perf_counter() - t0, "library": q.library_version(),
            "key": key_material(cfg),
        }
        return QuinticMetricResult(
            balanced_h=inter.ravel().tolist(), balanced_h_re=hre.ravel().tolist(),
            balanced_h_im=him.ravel().tolist(), n_basis=int(n), sample_points=pts,
            final_sigma_residual=float(final), sigma_fs_identity=float(sigma0),
            final_donaldson_residual=dhist[-1] if dhist else float("nan"),
            iterations_run=int(iters), basis_monomials=np.asarray(mons).tolist(),
            sigma_history=[float(sigma0), float(sigma_d)], donaldson_history=dhist,
            run_metadata=meta, sigma_after_donaldson=float(sigma_d),
            sigma_after_adam=float(sigma_a), adam_iters_run=len(adam_hist),
            adam_sigma_history=adam_hist)
    finally:
        s.close()


def load_or_solve(cache_root: str, cfg: QuinticMetricConfig, force_recompute: bool = False,
                  strict_quality: bool = True, solver=None):
    """load_or_solve_ty (metric_cache.rs:222-262) for the quintic: returns
    ``(result, outcome, path, elapsed_s)`` with ``outcome`` in {"hit", "miss",
    "forced_recompute"}. ``solver`` replaces :func:`solve_quintic_metric` (same signature)."""
    os.makedirs(cache_root, exist_ok=True)
    path = quintic_cache_path(cache_root, cfg)
    t0 = time.perf_counter()
    if not force_recompute and os.path.exists(path):
        result = load_result(path)
        check_sigma_quality(path, result.final_sigma_residual, strict_quality)
        return result, "hit", path, time.perf_counter() - t0
    outcome = "forced_recompute" if force_recompute else "miss"
    result = (solver or solve_quintic_metric)(cfg)
    check_sigma_quality(path, result.final_sigma_residual, strict_quality)
    save_result(path, result)
    return result, outcome, path, time.perf_counter() - t0
