"""Batched multi-instance balancing over the cy3q C ABI (include/cy3q.h, cy3q_batch_*).

Many (psi, k, N_p) quintic instances share one GPU. Each keeps its own QuinticSolver and HIP
stream. The balancing iterations of all of them run either one instance after another
(`solve_sequential`) or as one captured hipGraph launch (`solve_graph`). `autotune` picks the
T-pass chunking and the FP64 rank-k workgroup count per instance shape and caches the choice.
The reference has no batch mode (its instances run one at a time).
"""
from __future__ import annotations

import ctypes as C

from .quintic import PREC_SPLIT16, QuinticSolver, _check, _load, _options

_bound = False


class InstanceSpec(C.Structure):
    _fields_ = [("k", C.c_uint32), ("n_points", C.c_int64), ("seed", C.c_uint64),
                ("psi", C.c_double)]

    def as_tuple(self):
        return (int(self.k), int(self.n_points), int(self.seed), float(self.psi))


def _lib():
    global _bound
    L = _load()
    if not _bound:
        i64, u64, vp, ip = C.c_int64, C.c_uint64, C.c_void_p, C.c_int
        L.cy3q_batch_draw_specs.argtypes = [u64, ip, vp]
        L.cy3q_batch_new.argtypes = [vp, ip, vp, C.POINTER(vp)]
        L.cy3q_batch_free.argtypes = [vp]
        L.cy3q_batch_free.restype = None
        L.cy3q_batch_size.argtypes = [vp, C.POINTER(ip)]
        L.cy3q_batch_solver.argtypes = [vp, ip, C.POINTER(vp)]
        L.cy3q_batch_solve_sequential.argtypes = [vp, i64]
        L.cy3q_batch_solve_graph.argtypes = [vp, i64]
        L.cy3q_batch_autotune.argtypes = [vp, C.c_char_p, C.POINTER(i64)]
        L.cy3q_solver_set_tuning.argtypes = [vp, i64, i64]
        L.cy3q_solver_get_tuning.argtypes = [vp, C.POINTER(i64), C.POINTER(i64)]
        _bound = True
    return L


def draw_specs(seed: int, n: int) -> list:
    """n instance specs (psi, k, N_p) from the lib.rs LCG (see cy3q_batch_draw_specs)."""
    arr = (InstanceSpec * n)()
    _check(_lib().cy3q_batch_draw_specs(seed, n, arr))
    return [arr[i].as_tuple() for i in range(n)]


class _Borrowed(QuinticSolver):
    """A batch-owned solver handle: never freed from Python."""

    def close(self):
        self._h = None

    def __del__(self):
        pass


class Batch:
    def __init__(self, specs, quotient: bool = True, precision: int = PREC_SPLIT16, device: int = 0):
        L = _lib()
        n = len(specs)
        arr = (InstanceSpec * n)()
        for i, (k, npts, seed, psi) in enumerate(specs):
            arr[i].k, arr[i].n_points, arr[i].seed, arr[i].psi = k, npts, seed, psi
        o = _options(quotient=quotient, precision=precision, device=device)
        h = C.c_void_p()
        _check(L.cy3q_batch_new(arr, n, C.byref(o), C.byref(h)))
        self._h = h
        self.specs = [tuple(s) for s in specs]
        self.solvers = []
        for i in range(n):
            sh = C.c_void_p()
            _check(L.cy3q_batch_solver(self._h, i, C.byref(sh)))
            self.solvers.append(_Borrowed(sh.value))

    def __len__(self):
        return len(self.specs)

    def close(self):
        if getattr(self, "_h", None):
            for s in self.solvers:
                s.close()
            _lib().cy3q_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve_sequential(self, iters: int):
        _check(_lib().cy3q_batch_solve_sequential(self._h, iters))

    def solve_graph(self, iters: int):
        _check(_lib().cy3q_batch_solve_graph(self._h, iters))

    def autotune(self, cache_path: str | None = None) -> int:
        n = C.c_int64()
        _check(_lib().cy3q_batch_autotune(self._h, None if cache_path is None
                                          else cache_path.encode(), C.byref(n)))
        return n.value

    def tuning(self) -> list:
        out = []
        for s in self.solvers:
            a, b = C.c_int64(), C.c_int64()
            _check(_lib().cy3q_solver_get_tuning(s._h, C.byref(a), C.byref(b)))
            out.append((a.value, b.value))
        return out
