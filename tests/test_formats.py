"""On-disk formats: the chained-SHA replay log, the run manifest and the metric cache.

The replay-log fixtures under tests/golden/replog/ are files the reference wrote
(cy3_solver/output/*.replog and their .kernel.replog sidecars, copied as data). They pin the
chain rule, the canonical JSON bytes (field order, sorted free-form maps, float layout) and
the sidecar layout: rebuilding each run's events through this library's writer must reproduce
the reference's files byte for byte.
"""
import calendar
import glob
import hashlib
import json
import math
import os
import random
import struct

import numpy as np
import pytest

from papers_amd import metric_cache as mc
from papers_amd import repro

GOLD = os.path.join(os.path.dirname(__file__), "golden", "replog")
GOOD = ["p5_10_legacy_lcg_reduced", "p5_10_smoke", "p_ext1_compute"]
STALE_SIDECAR = "p_complete_prediction_set"


def _rebuild(path):
    """Re-create a replay log from its parsed events through the writer's event constructors."""
    recs, _ = repro.read_replog(path)
    w = repro.ReplogWriter()
    for _, _, _, ev in recs:
        kind = ev["kind"]
        if kind == "run_start":
            e = repro.run_start_event(ev["binary"], repro.ReproManifest.from_dict(ev["manifest"]),
                                      ev["config_json"])
        elif kind == "per_seed":
            e = repro.per_seed_event(repro.PerSeedEvent(
                **{k: v for k, v in ev.items() if k != "kind"}))
        else:
            e = repro.run_end_event(ev["summary"], ev["total_elapsed_s"])
        w.push(e)
    return recs, w


# ---- float layout -----------------------------------------------------------------------------

@pytest.mark.parametrize("x,text", [
    (1e-6, "1e-6"), (6.13866887128792e-05, "0.0000613866887128792"), (1234.5, "1234.5"),
    (1.0, "1.0"), (0.0, "0.0"), (-0.0, "-0.0"), (1e16, "1e16"), (1e15, "1000000000000000.0"),
    (1.5e300, "1.5e300"), (5e-324, "5e-324"), (0.1 + 0.2, "0.30000000000000004"),
    (1e-5, "0.00001"), (1e-4, "0.0001"), (123456789012345680.0, "1.2345678901234568e17"),
    (-2.5e-7, "-2.5e-7"), (float("nan"), "null"), (float("inf"), "null"), (100.0, "100.0"),
    (0.029800000000000004, "0.029800000000000004"),
])
def test_format_f64_layout(x, text):
    assert repro.format_f64(x) == text


def test_format_f64_round_trips_random_doubles():
    rng = random.Random(7)
    for _ in range(20000):
        bits = rng.getrandbits(64)
        x = struct.unpack("<d", struct.pack("<Q", bits))[0]
        t = repro.format_f64(x)
        if not math.isfinite(x):
            assert t == "null"
            continue
        y = float(t)
        assert struct.pack("<d", y) == struct.pack("<d", x), (x, t)
        # shortest: the digit string is the one repr() finds
        assert len(t.replace("-", "").replace(".", "").split("e")[0].strip("0")) <= 17


# ---- replay log fixtures ----------------------------------------------------------------------

@pytest.mark.parametrize("name", GOOD)
def test_reference_replog_verifies(name):
    path = os.path.join(GOLD, name + ".replog")
    final = repro.verify_replog(path)
    recs, trailer = repro.read_replog(path)
    assert final == trailer["final_chain_sha256_hex"] == recs[-1][1]
    # the chain rule, recomputed from the raw lines with hashlib alone
    h = hashlib.sha256(b"").digest()
    for _, hexd, eb, _ in recs:
        h = hashlib.sha256(h + eb).digest()
        assert h.hex() == hexd


@pytest.mark.parametrize("name", GOOD + [STALE_SIDECAR])
def test_event_reencoding_is_byte_identical(name):
    recs, _ = repro.read_replog(os.path.join(GOLD, name + ".replog"))
    for _, _, eb, ev in recs:
        assert repro.to_json_bytes(ev) == eb


@pytest.mark.parametrize("name", GOOD)
def test_writer_reproduces_reference_files(name, tmp_path):
    path = os.path.join(GOLD, name + ".replog")
    recs, w = _rebuild(path)
    assert w.final_chain_hex() == recs[-1][1]
    with open(path, "rb") as f:
        assert w.to_lines() == f.read()
    with open(os.path.join(GOLD, name + ".kernel.replog"), "rb") as f:
        assert w.kernel_sidecar() == f.read()
    out = tmp_path / "copy.replog"
    side = w.write_to_path(str(out))
    assert side == str(tmp_path / "copy.kernel.replog")
    assert repro.verify_replog(str(out)) == recs[-1][1]
    assert not glob.glob(str(tmp_path / "*.tmp"))


def test_stale_sidecar_is_rejected():
    # The reference's p_complete_prediction_set.kernel.replog was left from another run: its
    # entries do not hash the events of the .replog beside it.
    with pytest.raises(repro.ReplogError, match="sidecar"):
        repro.verify_replog(os.path.join(GOLD, STALE_SIDECAR + ".replog"))


def test_sidecar_layout_fields():
    with open(os.path.join(GOLD, "p5_10_legacy_lcg_reduced.kernel.replog"), "rb") as f:
        data = f.read()
    ents = repro.parse_kernel_replog(data)
    recs, _ = repro.read_replog(os.path.join(GOLD, "p5_10_legacy_lcg_reduced.replog"))
    assert len(data) == 16 + len(ents) * (4 + 19 + 8 + 32 + 32 + 4 + 4 + 8)
    for e, (seq, hexd, eb, _) in zip(ents, recs):
        assert e["name"] == "discriminator_event" and e["seed"] == seq
        assert e["n_out"] == 32 and e["t_ns"] == 0 and e["n_in"] == len(eb)


def _tampered(tmp_path, edit):
    src = os.path.join(GOLD, "p5_10_legacy_lcg_reduced.replog")
    with open(src, "rb") as f:
        lines = f.read().split(b"\n")
    lines = edit(lines)
    p = tmp_path / "t.replog"
    p.write_bytes(b"\n".join(lines))
    return str(p)


def test_tampering_is_detected(tmp_path):
    def change_digit(lines):
        lines[2] = lines[2].replace(b'"iters_run":18', b'"iters_run":19')
        return lines

    def drop_record(lines):
        return lines[:3] + lines[4:]

    def swap_records(lines):
        lines[1], lines[2] = lines[2], lines[1]
        return lines

    def bad_trailer(lines):
        lines[-2] = lines[-2].replace(b'"n_events":8', b'"n_events":7')
        return lines

    for edit, msg in [(change_digit, "chain mismatch"), (drop_record, "seq"),
                      (swap_records, "seq"), (bad_trailer, "trailer")]:
        with pytest.raises(repro.ReplogError, match=msg):
            repro.verify_replog(_tampered(tmp_path, edit))


def test_chain_is_deterministic_and_order_sensitive():
    # replog_chain_is_deterministic (repro.rs:378-402)
    def mk(seed):
        return repro.per_seed_event(repro.PerSeedEvent(
            seed=seed, candidate="TY", k=3, iters_run=17, final_residual=1e-7,
            sigma_fs_identity=0.123, sigma_final=0.456, n_basis=27, elapsed_ms=1234.5))

    a, b, c = repro.ReplogWriter(), repro.ReplogWriter(), repro.ReplogWriter()
    for w, seeds in ((a, (42, 99)), (b, (42, 99)), (c, (99, 42))):
        for s in seeds:
            w.push(mk(s))
    assert a.final_chain_hex() == b.final_chain_hex() != c.final_chain_hex()
    assert repro.to_json_bytes(mk(42)) == (
        b'{"kind":"per_seed","seed":42,"candidate":"TY","k":3,"iters_run":17,'
        b'"final_residual":1e-7,"sigma_fs_identity":0.123,"sigma_final":0.456,"n_basis":27,'
        b'"elapsed_ms":1234.5}')
    empty = repro.ReplogWriter()
    assert empty.final_chain_hex() == (
        "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855")


def test_free_form_maps_are_key_sorted():
    ev = repro.run_end_event({"b": 1, "a": {"z": 1.0, "c": [2, {"y": None, "x": True}]}}, 2.0)
    assert repro.to_json_bytes(ev) == (
        b'{"kind":"run_end","summary":{"a":{"c":[2,{"x":true,"y":null}],"z":1.0},"b":1},'
        b'"total_elapsed_s":2.0}')


# ---- manifest ---------------------------------------------------------------------------------

def test_manifest_collects():
    # manifest_collects_without_panic (repro.rs:362-375)
    m = repro.ReproManifest.collect()
    assert m.rust_toolchain and m.target_triple and m.n_threads_rayon >= 1
    assert len(m.timestamp_utc) == 20 and m.timestamp_utc.endswith("Z")
    d = json.loads(repro.to_json_bytes(repro.run_start_event("bench", m, {"k": 6})))
    assert list(d["manifest"])[:8] == [
        "git_revision", "rust_toolchain", "target_triple", "cpu_features", "hostname",
        "timestamp_utc", "command_line", "n_threads_rayon"]


@pytest.mark.parametrize("stamp", ["1970-01-01T00:00:00Z", "2000-02-29T23:59:59Z",
                                   "2001-09-09T01:46:40Z", "2026-05-01T14:48:41Z",
                                   "2100-03-01T00:00:00Z"])
def test_rfc3339_civil_dates(stamp):
    t = calendar.timegm((int(stamp[:4]), int(stamp[5:7]), int(stamp[8:10]), int(stamp[11:13]),
                         int(stamp[14:16]), int(stamp[17:19]), 0, 0, 0))
    assert repro.rfc3339_utc_now(t) == stamp


# ---- metric cache -----------------------------------------------------------------------------

def test_reference_cache_key_bytes():
    key = (b'{"schema_version":1,"manifold":"tian_yau_z3","k_degree":4,"n_sample":1000,'
           b'"max_iter":20,"donaldson_tol_bits":' + str(
               struct.unpack("<Q", struct.pack("<d", 1e-3))[0]).encode()
           + b',"seed":42,"apply_z3_quotient":true}')
    h = hashlib.sha256(key).hexdigest()[:16]
    p = mc.ty_cache_path("root", 4, 1000, 20, 1e-3, 42, True)
    assert p == os.path.join("root", f"ty_k4_n1000_i20_seed42_{h}.json")
    assert mc.f64_bits(1e-3) == 0x3F50624DD2F1A9FC


def test_reference_cache_paths_include_all_key_fields():
    # ty_path_includes_all_key_fields / schoen_path_includes_all_key_fields (metric_cache.rs)
    a = mc.ty_cache_path("r", 4, 1000, 20, 1e-3, 42, True)
    assert a != mc.ty_cache_path("r", 4, 1000, 20, 1e-3, 43, True)
    assert a != mc.ty_cache_path("r", 5, 1000, 20, 1e-3, 42, True)
    assert a != mc.ty_cache_path("r", 4, 1000, 20, 1e-4, 42, True)
    assert a != mc.ty_cache_path("r", 4, 1000, 20, 1e-3, 42, False)
    s = mc.schoen_cache_path("r", 4, 4, 2, 1000, 20, 1e-3, 42, True)
    assert s != mc.schoen_cache_path("r", 5, 4, 2, 1000, 20, 1e-3, 42, True)
    assert s != mc.schoen_cache_path("r", 4, 4, 2, 1000, 20, 1e-3, 42, False)
    assert os.path.basename(s).startswith("schoen_dx4_dy4_dt2_n1000_i20_seed42_")


def _cfg(**kw):
    base = dict(k_degree=4, n_sample=1000, max_iter=20, donaldson_tol=1e-3, seed=42)
    base.update(kw)
    return mc.QuinticMetricConfig(**base)


def test_quintic_cache_path_includes_all_key_fields():
    a = mc.quintic_cache_path("r", _cfg())
    assert os.path.basename(a).startswith("quintic_k4_n1000_i20_seed42_")
    changes = [dict(seed=43), dict(k_degree=5), dict(n_sample=1001), dict(max_iter=21),
               dict(donaldson_tol=np.nextafter(1e-3, 1.0)), dict(apply_quotient=False),
               dict(sampler="newton"), dict(sampler="sz_dwork", psi=0.5),
               dict(precision="split16"), dict(orthonormalise_fs_gram=True),
               dict(adam_iters=10)]
    paths = {a}
    for ch in changes:
        p = mc.quintic_cache_path("r", _cfg(**ch))
        assert p not in paths, ch
        paths.add(p)
    assert mc.quintic_cache_path("r", _cfg(sampler="sz_dwork", psi=0.5)) != \
        mc.quintic_cache_path("r", _cfg(sampler="sz_dwork", psi=np.nextafter(0.5, 1.0)))
    assert mc.quintic_cache_path("r", _cfg(precision="split16", refine=1)) != \
        mc.quintic_cache_path("r", _cfg(precision="split16", refine=2))
    # inputs that change neither H nor the residuals do not change the key
    assert mc.quintic_cache_path("r", _cfg(device=1, store_points=True)) == a
    assert mc.quintic_cache_path("r", _cfg(refine=3)) == a  # refine only counts in split16
    with pytest.raises(ValueError):
        mc.quintic_cache_path("r", _cfg(sampler="sz", psi=0.5))


def test_quality_threshold_and_guard(capsys):
    assert mc.SIGMA_QUALITY_THRESHOLD == 1.0e-3
    mc.check_sigma_quality("p", 9.99e-4, strict=True)
    for bad in (1e-3, 0.2, float("nan"), float("inf"), None):
        with pytest.raises(mc.StaleQuality):
            mc.check_sigma_quality("p", bad, strict=True)
    mc.check_sigma_quality("p", 0.2, strict=False)
    assert "WARNING" in capsys.readouterr().err


def _fake_result(cfg, sigma):
    rng = np.random.default_rng(cfg.seed)
    n = 6
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    h = a @ a.conj().T + n * np.eye(n)
    inter = np.stack([h.real, h.imag], -1)
    return mc.QuinticMetricResult(
        balanced_h=inter.ravel().tolist(), balanced_h_re=h.real.ravel().tolist(),
        balanced_h_im=h.imag.ravel().tolist(), n_basis=n, sample_points=[],
        final_sigma_residual=sigma, sigma_fs_identity=0.5, final_donaldson_residual=1e-9,
        iterations_run=7, basis_monomials=[[4, 0, 0, 0, 0]], sigma_history=[0.5, sigma],
        donaldson_history=[1e-1, 1e-9], run_metadata={"z": 1, "a": float("nan")},
        sigma_after_donaldson=sigma)


def test_load_or_solve_cycle(tmp_path):
    calls = []

    def solver(cfg):
        calls.append(cfg.seed)
        return _fake_result(cfg, 5e-4)

    cfg = _cfg()
    r1, out1, p1, _ = mc.load_or_solve(str(tmp_path), cfg, solver=solver)
    assert out1 == "miss" and calls == [42] and os.path.exists(p1)
    assert not glob.glob(str(tmp_path / "*.tmp"))
    r2, out2, p2, _ = mc.load_or_solve(str(tmp_path), cfg, solver=solver)
    assert out2 == "hit" and p2 == p1 and calls == [42]
    # H reads back bit for bit; NaN (written as null) reads back as NaN
    assert np.array_equal(np.asarray(r2.balanced_h), np.asarray(r1.balanced_h))
    assert r2.h_complex().tobytes() == r1.h_complex().tobytes()
    assert math.isnan(r2.sigma_after_adam) and r2.run_metadata["a"] is None
    hb = r2.h_block()
    assert np.array_equal(hb[0::2, 0::2], r2.h_complex().real)
    assert np.array_equal(-hb[0::2, 1::2], r2.h_complex().imag)
    _, out3, _, _ = mc.load_or_solve(str(tmp_path), cfg, force_recompute=True, solver=solver)
    assert out3 == "forced_recompute" and calls == [42, 42]


def test_load_or_solve_quality_guard(tmp_path):
    cfg = _cfg(seed=7)
    with pytest.raises(mc.StaleQuality):
        mc.load_or_solve(str(tmp_path), cfg, solver=lambda c: _fake_result(c, 2e-3))
    assert not os.path.exists(mc.quintic_cache_path(str(tmp_path), cfg))
    # non-strict: cached anyway; a strict load of that file is then refused
    _, out, p, _ = mc.load_or_solve(str(tmp_path), cfg, strict_quality=False,
                                    solver=lambda c: _fake_result(c, 2e-3))
    assert out == "miss" and os.path.exists(p)
    with pytest.raises(mc.StaleQuality):
        mc.load_or_solve(str(tmp_path), cfg, solver=lambda c: _fake_result(c, 2e-3))
