"""cy3q_hermitian_eigen_noalloc, the allocation-free form of cy3q_hermitian_eigen, through the C
ABI: device inputs and outputs (the caller's memory, from the HIP runtime the library itself
links), calls in a row on the shared scratch array, each smaller than the last, each equal to the
allocating entry point bit for bit (the same kernels in the same order), and the size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib(cy3q):
    from papers_amd import galerkin as G
    L = G._lib()
    vp, i64, ip, dbl = C.c_void_p, C.c_int64, C.c_int, C.c_double
    L.cy3q_hermitian_eigen_noalloc.argtypes = [vp, i64, i64, dbl, ip, vp, vp, C.POINTER(i64)]
    return L


def _herm(rng, n):
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return (a + a.conj().T) / 2


def _hip():
    # The runtime libcy3q.so is linked against (already loaded under this soname).
    H = C.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipFree.argtypes = [C.c_void_p]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipDeviceSynchronize.argtypes = []
    return H


_H2D, _D2H = 1, 2


def _dev(H, nbytes):
    p = C.c_void_p()
    assert H.hipMalloc(C.byref(p), nbytes) == 0
    return p


def _noalloc(L, a, vectors, max_sweeps=64, tol=1e-14):
    H = _hip()
    n = a.shape[0]
    a = np.ascontiguousarray(a, np.complex128)
    da, de = _dev(H, a.nbytes), _dev(H, 8 * n)
    dv = _dev(H, 16 * n * n) if vectors else None
    ev = np.full(n, np.nan)
    vec = np.full((n, n), np.nan + 0j) if vectors else None
    try:
        assert H.hipMemcpy(da, a.ctypes.data, a.nbytes, _H2D) == 0
        assert H.hipMemcpy(de, ev.ctypes.data, ev.nbytes, _H2D) == 0
        sw = C.c_int64()
        rc = L.cy3q_hermitian_eigen_noalloc(da, n, max_sweeps, tol, 0, de, dv, C.byref(sw))
        assert H.hipDeviceSynchronize() == 0
        assert H.hipMemcpy(ev.ctypes.data, de, ev.nbytes, _D2H) == 0
        if vectors:
            assert H.hipMemcpy(vec.ctypes.data, dv, vec.nbytes, _D2H) == 0
    finally:
        for p in (da, de, dv):
            if p is not None:
                H.hipFree(p)
    return rc, ev, vec, sw.value


def test_noalloc_eigen_matches_allocating_entry_point_twice(cy3q):
    from papers_amd import galerkin as G
    L = _lib(cy3q)
    rng = np.random.default_rng(31)
    for n, vectors in ((301, True), (64, True), (17, False)):
        print(f"case: n={n} vectors={vectors}", flush=True)
        a = _herm(rng, n)
        rc, ev, vec, sw = _noalloc(L, a, vectors)
        assert rc == 0, L.cy3q_last_error()
        ev_ref, vec_ref, sw_ref = G.hermitian_eigen(a, max_sweeps=64, tol=1e-14, vectors=vectors)
        assert sw == sw_ref
        assert np.array_equal(ev, ev_ref)
        if vectors:
            assert np.array_equal(vec, vec_ref)
        np.testing.assert_allclose(ev, np.linalg.eigvalsh(a), atol=1e-10 * np.abs(ev).max())


def test_noalloc_eigen_size_limit(cy3q):
    L = _lib(cy3q)
    a = np.eye(513, dtype=np.complex128)
    rc, _, _, _ = _noalloc(L, a, False)
    assert rc == -1
    assert b"513" in L.cy3q_last_error()
