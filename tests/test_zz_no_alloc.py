"""The no-allocation PDF walk (EZmock_ext_pdf_walk_noalloc) through the C
ABI, against the existing EZmock_ext_pdf_walk on the same densities: tracers
per cell, cells per PDF bin, ncell and ntot bit-exact, on two calls in a row
(the second with a smaller input, so stale contents of the shared static
buffer would show), and against the oracle's literal walk
(pop_tracer.c:359-514).  Device arrays come from hipMalloc of the HIP
runtime the engine itself links."""
import ctypes as C

import numpy as np
import pytest

from ezmock_amd import _lib, pdf_walk
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    _lib.load("f32")                    # the engine's runtime is loaded first
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


class Dev:
    """A device array filled from / read back to numpy."""

    def __init__(self, h, nbytes):
        self.h, self.p = h, C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.h.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, dtype, n):
        out = np.empty(n, dtype=dtype)
        assert self.h.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.p)


def densities(n, seed, prec, nexp):
    rs = np.random.default_rng(seed)
    rhot = rs.exponential(2.0, n)
    rhot[rs.random(n) < 0.4] = 0
    rhot = rhot.astype(po.rtype(prec))
    e, info = po.eval_pdf(nexp, 0.3 + 0.05 * (seed % 4), n, prec)
    assert e == 0
    order = np.argsort(-rhot.astype(np.float64), kind="stable")
    for c in np.cumsum(info["pdf"])[:-1]:           # ties at the bin boundaries
        rhot[order[c:c + 3]] = rhot[order[c - 1]]
    return rhot, info


def walk_noalloc(L, h, rhot, info, prec):
    n = len(rhot)
    nmax = len(info["pdf"])
    d_r, d_nt = Dev(h, rhot.nbytes), Dev(h, 2 * n)
    try:
        d_r.put(rhot)
        pdf = np.ascontiguousarray(info["pdf"], dtype=np.uint64)
        pdf_out = np.zeros(nmax, dtype=np.uint64)
        nc, nt_ = C.c_size_t(), C.c_size_t()
        err = C.c_int(0)
        r = L.EZmock_ext_pdf_walk_noalloc(d_r.p, n, info["ncell"], pdf.ctypes.data, nmax, d_nt.p,
                                         pdf_out.ctypes.data, C.byref(nc), C.byref(nt_), None,
                                         C.byref(err))
        nt = d_nt.get(np.int16, n) if r == 0 else None
    finally:
        d_r.free()
        d_nt.free()
    return r, nt, pdf_out.astype(np.int64), nc.value, nt_.value


def lib(prec):
    L = _lib.load(prec)
    L.EZmock_ext_pdf_walk_noalloc.restype = C.c_int
    L.EZmock_ext_pdf_walk_noalloc.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p,
                                             C.c_int, C.c_void_p, C.c_void_p,
                                             C.POINTER(C.c_size_t), C.POINTER(C.c_size_t),
                                             C.c_void_p, C.POINTER(C.c_int)]
    L.EZmock_ext_noalloc_scratch_bytes.restype = C.c_size_t
    return L


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_noalloc_walk_matches_allocating_walk(gpu_available, prec):
    L, h = lib(prec), hip()
    assert 0 < L.EZmock_ext_noalloc_scratch_bytes() < 80 * 2 ** 20
    for n, seed, nexp in ((400000, 1, 30000), (150000, 2, 9000)):    # second call smaller
        rhot, info = densities(n, seed, prec, nexp)
        r, nt, p, nc, ntot = walk_noalloc(L, h, rhot, info, prec)
        assert r == 0
        e2, nt2, p2, nc2, ntot2 = pdf_walk(rhot, info["ncell"], info["pdf"], prec)
        assert e2 == 0
        assert np.array_equal(nt, nt2)
        assert np.array_equal(p, p2) and (nc, ntot) == (nc2, ntot2)
        e3, nt3, p3, nc3, ntot3 = po.pdf_walk(rhot, info["ncell"], info["pdf"], prec)
        assert e3 == 0 and np.array_equal(nt, nt3) and (nc, ntot) == (nc3, ntot3)
        assert nc > info["ncell"]                   # the planted ties were absorbed


def test_noalloc_walk_rejects_too_many_candidates(gpu_available):
    """More candidate cells than the static buffer holds: a clear error
    (EZMOCK_ERR_ARG_HIEXP) before any candidate is written."""
    L, h = lib("f32"), hip()
    n = (1 << 22) + 100000
    rhot = np.ones(n, dtype=np.float32)
    info = {"ncell": n - 1000, "pdf": np.array([n - 1000], dtype=np.int64)}
    r, *_ = walk_noalloc(L, h, rhot, info, "f32")
    assert r == 16                                  # EZMOCK_ERR_ARG_HIEXP
    # the next call on the same GPU is unaffected
    rhot, info = densities(100000, 3, "f32", 6000)
    r, nt, p, nc, ntot = walk_noalloc(L, h, rhot, info, "f32")
    e2, nt2, p2, nc2, ntot2 = pdf_walk(rhot, info["ncell"], info["pdf"], "f32")
    assert r == 0 and np.array_equal(nt, nt2) and (nc, ntot) == (nc2, ntot2)
