"""CPU oracle for the 8-coordinate (Tian-Yau, Schoen) callers of the route34 GPU contracts.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker; the product never imports it.

* :func:`sample` wraps ``orc_cicy_sample`` (oracle/cy3q_oracle_cicy.c): ``CicySampler::sample_batch``
  (cicy_sampler.rs:591-617) and ``SchoenSampler::sample_points`` (route34/schoen_sampler.rs:482-498)
  with the reference's ChaCha8 + ziggurat stream.
* :func:`groebner_lms` is Buchberger's algorithm over the rationals (exact ``Fraction``
  arithmetic) under degree-lex with x0 > x1 > ... > x7, as ``reduced_groebner``
  (route34/groebner.rs:470-660) computes it; only the leading monomials of the reduced basis are
  returned, which is all the section-basis filter uses (``monomial_in_lm_ideal``, :769-787).
* :func:`ty_basis` and :func:`schoen_basis` are ``build_ty_invariant_reduced_basis``
  (route34/ty_metric.rs:792-895) and ``build_schoen_invariant_reduced_basis``
  (route34/schoen_metric.rs:1185-1212).
* :func:`solve` runs the whole ``solve_ty_metric`` / ``solve_schoen_metric`` protocol for static
  damping without a Tikhonov shift (the k = 3 and (3,3,1) defaults): sigma at h = I, then the
  Donaldson iterations of ``donaldson_iteration_impl`` (ty_metric.rs:1063-1343) with sigma after
  each, the minimum-residual snapshot and the catastrophe guard (:573-708).
"""
from __future__ import annotations

import ctypes as C
import itertools
from fractions import Fraction

import numpy as np
import scipy.linalg as sla

from . import oracle as O

TY, SCHOEN = 0, 1

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


def _lib():
    L = O.lib()
    if not getattr(L, "_cicy_ready", False):
        i64, u64 = C.c_int64, C.c_uint64
        L.orc_cicy_sample.argtypes = [C.c_int, u64, i64, _dp, _dp, _dp, _i32p, C.POINTER(i64),
                                      C.c_void_p, i64]
        L.orc_cicy_sample.restype = i64
        L.orc_cicy_frames.argtypes = [C.c_int, i64, _dp, _i32p, _dp]
        L.orc_cicy_sections.argtypes = [i64, _dp, i64, _u32p, _dp, _dp]
        L.orc_chacha_block.argtypes = [_u32p, u64, u64, C.c_int, _u32p]
        L.orc_chacha8_u64_stream.argtypes = [u64, i64, _u64p]
        L.orc_std_normal_stream.argtypes = [u64, i64, _dp]
        L.orc_zig_tables.argtypes = [_dp, _dp]
        L.orc_ty_sample_lanes.argtypes = [u64, i64, _dp, _dp, _dp, _i32p, C.c_void_p,
                                          C.POINTER(i64)]
        L.orc_ty_sample_lanes.restype = i64
        L._cicy_ready = True
    return L


def chacha_block(key, counter: int, stream: int, rounds: int) -> np.ndarray:
    out = np.zeros(16, np.uint32)
    _lib().orc_chacha_block(np.ascontiguousarray(key, np.uint32), counter, stream, rounds, out)
    return out


def chacha8_u64(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint64)
    _lib().orc_chacha8_u64_stream(seed, n, out)
    return out


def std_normals(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n)
    _lib().orc_std_normal_stream(seed, n, out)
    return out


def zig_tables():
    x, f = np.zeros(257), np.zeros(257)
    _lib().orc_zig_tables(x, f)
    return x, f


def sample(which: int, seed: int, n_target: int, line_log: bool = False):
    """Returns (coords [n][8] complex, omega_sq, weight, idx [n][3], n_lines[, log])."""
    cap = max(n_target * 64, 1024)
    coords = np.zeros((n_target, 16))
    om = np.zeros(n_target)
    w = np.zeros(n_target)
    idx = np.zeros((n_target, 3), np.int32)
    nl = C.c_int64(0)
    log = np.zeros((cap, 3), np.int32) if line_log else None
    n = _lib().orc_cicy_sample(which, seed, n_target, coords, om, w, idx, C.byref(nl),
                               log.ctypes.data if line_log else None, cap if line_log else 0)
    c = coords[:n, 0::2] + 1j * coords[:n, 1::2]
    out = (c, om[:n].copy(), w[:n].copy(), idx[:n].copy(), int(nl.value))
    if line_log:
        out = out + (log[: nl.value].copy(),)
    return out


def sample_lanes(seed: int, n_target: int):
    """The reference GPU build's TY batch sampler: (coords, omega_sq, weight, idx, n_lanes,
    lane_ok)."""
    nl = max(4 * n_target, 64)
    coords = np.zeros((n_target, 16))
    om = np.zeros(n_target)
    w = np.zeros(n_target)
    idx = np.zeros((n_target, 3), np.int32)
    ok = np.zeros(nl, np.int32)
    nlo = C.c_int64(0)
    n = _lib().orc_ty_sample_lanes(seed, n_target, coords, om, w, idx, ok.ctypes.data,
                                   C.byref(nlo))
    c = coords[:n, 0::2] + 1j * coords[:n, 1::2]
    return c, om[:n].copy(), w[:n].copy(), idx[:n].copy(), int(nlo.value), ok


def frames(which: int, coords: np.ndarray, idx: np.ndarray) -> np.ndarray:
    """[n][3][8][2] tangent frames (NaN for a degenerate chart)."""
    c = np.ascontiguousarray(np.stack([coords.real, coords.imag], -1).reshape(len(coords), 16))
    out = np.zeros((len(coords), 48))
    _lib().orc_cicy_frames(which, len(coords), c, np.ascontiguousarray(idx, np.int32), out)
    return out.reshape(-1, 3, 8, 2)


def sections(coords: np.ndarray, exps: np.ndarray):
    """(sv [n][2 nb], sd [n][8][2 nb]) of the monomial basis ``exps`` [nb][8]."""
    n, nb = len(coords), len(exps)
    c = np.ascontiguousarray(np.stack([coords.real, coords.imag], -1).reshape(n, 16))
    sv = np.zeros((n, 2 * nb))
    sd = np.zeros((n, 8, 2 * nb))
    _lib().orc_cicy_sections(n, c, nb, np.ascontiguousarray(exps, np.uint32), sv, sd)
    return sv, sd


# ---------------------------------------------------------------------------------------------
# Buchberger over Q (Cox, Little, O'Shea, Ideals, Varieties and Algorithms, ch. 2 sections 6-7)

def _key(e):
    """Degree-lex sort key: total degree, then lex with x0 largest (groebner.rs:125-157)."""
    return (sum(e), tuple(e))


def _lm(p):
    return max(p, key=_key)


def _sub(p, q, c, shift):
    """p - c * x^shift * q (dicts exponent -> Fraction)."""
    r = dict(p)
    for e, v in q.items():
        e2 = tuple(a + b for a, b in zip(e, shift))
        r[e2] = r.get(e2, Fraction(0)) - c * v
        if r[e2] == 0:
            del r[e2]
    return r


def _divides(a, b):
    return all(x <= y for x, y in zip(a, b))


def _reduce(f, basis):
    """Full normal form of f modulo basis."""
    r, p = {}, dict(f)
    while p:
        lm = _lm(p)
        for g in basis:
            lg = _lm(g)
            if _divides(lg, lm):
                shift = tuple(x - y for x, y in zip(lm, lg))
                p = _sub(p, g, p[lm] / g[lg], shift)
                break
        else:
            r[lm] = p.pop(lm)
    return r


def _spoly(f, g):
    lf, lg = _lm(f), _lm(g)
    l = tuple(max(a, b) for a, b in zip(lf, lg))
    sf = tuple(a - b for a, b in zip(l, lf))
    sg = tuple(a - b for a, b in zip(l, lg))
    a = {tuple(x + y for x, y in zip(e, sf)): v / f[lf] for e, v in f.items()}
    return _sub(a, g, 1 / g[lg], sg)


def groebner_lms(gens):
    """Leading monomials of the reduced Groebner basis of the ideal of ``gens`` (lists of
    (coefficient, exponent tuple)). S-pairs are taken in order of increasing lcm degree (the
    normal selection strategy); pairs with coprime leading monomials are skipped (Buchberger's
    first criterion). The result depends only on the ideal and the order."""
    G = [{tuple(e): Fraction(c) for c, e in g} for g in gens]

    def lcm_deg(i, j):
        return sum(max(a, b) for a, b in zip(_lm(G[i]), _lm(G[j])))

    pairs = set(itertools.combinations(range(len(G)), 2))
    while pairs:
        i, j = min(pairs, key=lambda q: (lcm_deg(*q), q))
        pairs.discard((i, j))
        li, lj = _lm(G[i]), _lm(G[j])
        if all(a == 0 or b == 0 for a, b in zip(li, lj)):
            continue
        r = _reduce(_spoly(G[i], G[j]), G)
        if r:
            G.append(r)
            pairs.update((k, len(G) - 1) for k in range(len(G) - 1))
    lms = [_lm(g) for g in G]
    keep = []
    for a, la in enumerate(lms):
        if any(b != a and _divides(lb, la) and (lb != la or b < a) for b, lb in enumerate(lms)):
            continue
        keep.append(la)
    return sorted(set(keep), key=_key)


def ty_generators():
    """ty_generators (route34/groebner.rs:665-690)."""
    f1 = [(1, tuple(3 if j == k else 0 for j in range(8))) for k in range(4)]
    f2 = [(1, tuple(3 if j == 4 + k else 0 for j in range(8))) for k in range(4)]
    f3 = [(1, tuple(1 if j in (k, 4 + k) else 0 for j in range(8))) for k in range(4)]
    return [f1, f2, f3]


def schoen_generators():
    """schoen_generators (route34/groebner.rs:705-744)."""
    f1, f2 = [], []
    for k in range(3):
        e = [0] * 8; e[k] = 3; e[6] = 1
        f1.append((1, tuple(e)))
        e2 = [0] * 8; e2[k] = 3; e2[7] = 1
        f1.append((1 if k == 0 else -1, tuple(e2)))
        e = [0] * 8; e[3 + k] = 3; e[6] = 1
        f2.append((1, tuple(e)))
        e2 = [0] * 8; e2[3 + k] = 3; e2[7] = 1
        f2.append((1 if k == 0 else -1, tuple(e2)))
    return [f1, f2]


def in_lm_ideal(m, lms) -> bool:
    return any(_divides(l, m) for l in lms)


def ty_basis(k: int, lms=None) -> np.ndarray:
    """build_ty_invariant_reduced_basis: the (k, k) monomials in z0..z3 x w0..w3, z-part major
    (ty_metric.rs:792-820), with z3_character = (m1 + 2 m2 + m5 + 2 m6) mod 3 == 0
    (quotient.rs:55-57) and outside the leading-monomial ideal."""
    lms = groebner_lms(ty_generators()) if lms is None else lms
    z = [(a0, a1, a2, k - a0 - a1 - a2) for a0 in range(k + 1) for a1 in range(k - a0 + 1)
         for a2 in range(k - a0 - a1 + 1)]
    out = []
    for zz in z:
        for ww in z:
            m = zz + ww
            if (m[1] + 2 * m[2] + m[5] + 2 * m[6]) % 3 != 0:
                continue
            if in_lm_ideal(m, lms):
                continue
            out.append(m)
    return np.array(out, np.uint32).reshape(-1, 8)


def schoen_basis(d_x: int, d_y: int, d_t: int, lms=None) -> np.ndarray:
    """build_schoen_invariant_reduced_basis: enumerate_bidegree_monomials order
    (route34/z3xz3_projector.rs:298-315), alpha character (m1 + 2 m2 + m4 + 2 m5) mod 3 and beta
    character m7 mod 3 both zero (:105-113), outside the leading-monomial ideal."""
    lms = groebner_lms(schoen_generators()) if lms is None else lms
    out = []
    for a0 in range(d_x + 1):
        for a1 in range(d_x - a0 + 1):
            a2 = d_x - a0 - a1
            for b0 in range(d_y + 1):
                for b1 in range(d_y - b0 + 1):
                    b2 = d_y - b0 - b1
                    for c0 in range(d_t + 1):
                        m = (a0, a1, a2, b0, b1, b2, c0, d_t - c0)
                        if (m[1] + 2 * m[2] + m[4] + 2 * m[5]) % 3 or m[7] % 3:
                            continue
                        if in_lm_ideal(m, lms):
                            continue
                        out.append(m)
    return np.array(out, np.uint32).reshape(-1, 8)


# ---------------------------------------------------------------------------------------------
# The solve protocol

def donaldson_iteration(sv, w, h_re, h_im):
    """donaldson_iteration_impl (ty_metric.rs:1063-1343) with alpha = 1 and no Tikhonov shift.
    Returns (h_re, h_im, residual)."""
    nb = h_re.shape[0]
    hp_re, hp_im = O.t_operator_raw_route34(sv, w, h_re, h_im)
    ws = float(np.sum(w))
    if not ws > 1e-30:
        return h_re, h_im, float("nan")
    a = hp_re * (1.0 / ws)
    b = hp_im * (1.0 / ws)
    a = 0.5 * (a + a.T)
    b = 0.5 * (b - b.T)
    np.fill_diagonal(b, 0.0)
    tr = np.trace(a)
    if abs(tr) > 1e-30:
        a, b = a * (nb / tr), b * (nb / tr)
    blk = np.zeros((2 * nb, 2 * nb))
    blk[0::2, 0::2] = a
    blk[0::2, 1::2] = -b
    blk[1::2, 0::2] = b
    blk[1::2, 1::2] = a
    lu, piv = sla.lu_factor(blk)
    d = np.abs(np.diag(lu))
    if not np.all(np.isfinite(d)) or np.any(d < np.finfo(float).tiny):
        return h_re, h_im, float("nan")
    inv = sla.lu_solve((lu, piv), np.eye(2 * nb))
    if not np.all(np.isfinite(inv)):
        return h_re, h_im, float("nan")
    na = 0.5 * (inv[0::2, 0::2] + inv[1::2, 1::2])
    nbm = 0.5 * (inv[1::2, 0::2] - inv[0::2, 1::2])
    tr = np.trace(na)
    if abs(tr) > 1e-30:
        na, nbm = na * (nb / tr), nbm * (nb / tr)
    r = float(np.sqrt(np.sum((na - h_re) ** 2 + (nbm - h_im) ** 2)))
    return na, nbm, r


def setup(which: int, seed: int, n_sample: int, degrees, ty_mode: str = "lanes"):
    """Points (with the quotient factor), basis, sections, derivatives and frames. TY points come
    from the batch sampler of the reference's GPU build (``ty_mode="lanes"``, the build its
    determinism golden file was written with) or the sequential sampler (``"restarts"``)."""
    if which == TY and ty_mode == "lanes":
        c, om, w, idx = sample_lanes(seed, n_sample)[:4]
    else:
        c, om, w, idx, _ = sample(which, seed, n_sample)
    if which == TY:
        w = w * (1.0 / 3.0)
        exps = ty_basis(*degrees)
    else:
        w = w * (1.0 / 9.0)
        exps = schoen_basis(*degrees)
    sv, sd = sections(c, exps)
    fr = frames(which, c, idx)
    return dict(coords=c, omega_sq=om, w=w, idx=idx, exps=exps, sv=sv, sd=sd, frames=fr)


def solve(which: int, seed: int, n_sample: int, degrees, max_iter: int, tol: float,
          ty_mode: str = "lanes"):
    """Returns dict(n_basis, sigma_fs_identity, sigma_history, residual_history,
    final_sigma, final_residual, iterations)."""
    S = setup(which, seed, n_sample, degrees, ty_mode)
    sv, sd, fr, om, w = S["sv"], S["sd"], S["frames"], S["omega_sq"], S["w"]
    nb = len(S["exps"])

    def sig(hr, hi):
        return O.sigma_route34(sv, sd, fr, om, w, hr, hi)[0]

    h_re, h_im = np.eye(nb), np.zeros((nb, nb))
    s0 = sig(h_re, h_im)
    res_h, sig_h = [], []
    min_r, min_h, min_it, bad = float("inf"), (h_re, h_im), 0, 0
    final_r, iters = float("inf"), 0
    for it in range(max_iter):
        h_re2, h_im2, r = donaldson_iteration(sv, w, h_re, h_im)
        if not np.isfinite(r):
            if np.isfinite(min_r):
                h_re, h_im = min_h
                res_h.append(min_r)
                sig_h.append(sig(h_re, h_im))
                final_r, iters = min_r, min_it
                break
            raise RuntimeError("non-finite residual before any snapshot")
        h_re, h_im = h_re2, h_im2
        res_h.append(r)
        sig_h.append(sig(h_re, h_im))
        final_r, iters = r, it + 1
        if r < min_r:
            min_r, min_h, min_it, bad = r, (h_re, h_im), iters, 0
        elif min_r < 1e-2 and r > 100.0 * min_r:
            bad += 1
            if bad >= 5:
                h_re, h_im = min_h
                res_h.append(min_r)
                sig_h.append(sig(h_re, h_im))
                final_r, iters = min_r, min_it
                break
        else:
            bad = 0
        if r < tol:
            break
    return dict(n_basis=nb, sigma_fs_identity=s0, sigma_history=sig_h, residual_history=res_h,
                final_sigma=sig_h[-1] if sig_h else float("nan"), final_residual=final_r,
                iterations=iters, h_re=h_re, h_im=h_im, setup=S)
