"""Request pool: continuous batching of mocks on one GPU (EZmock_ext_pool_*).

``MockPool`` owns ``nslot`` worker threads of the engine library, each with its
own instance and HIP stream.  ``submit`` queues a mock request and returns at
once; ``wait`` returns its result.  Requests are admitted as workers and the
device-memory budget free up, so mocks of different grid sizes run
concurrently on the device; a request is computed by the same calls as a
stand-alone mock (``EZmock``) with the pool's cosmology and linear P(k).
"""
import ctypes as C

import numpy as np

from . import _lib


def _structs(real):
    class Request(C.Structure):
        _fields_ = [("Ngrid", C.c_int), ("Lbox", C.c_double), ("seed", C.c_uint64),
                    ("params", real * 4), ("nexp", C.c_size_t), ("rsd_fac", C.c_double),
                    ("pk_dk", C.c_double), ("pk_nbin", C.c_int), ("want_catalog", C.c_int)]

    class Result(C.Structure):
        _fields_ = [("id", C.c_long), ("err", C.c_int), ("ntracer", C.c_size_t),
                    ("t_submit", C.c_double), ("t_start", C.c_double), ("t_end", C.c_double),
                    ("k", C.POINTER(C.c_double)), ("P0", C.POINTER(C.c_double)),
                    ("P2", C.POINTER(C.c_double)), ("nmodes", C.POINTER(C.c_double)),
                    ("cat", C.POINTER(real) * 6)]
    return Request, Result


class MockPool:
    def __init__(self, k, Pk, z_out, Omega_m, z_pk=0.0, Omega_nu=0.0, w=-1.0,
                 nslot=4, mem_budget=200e9, precision="f32", options=None):
        self._L = _lib.load(precision)
        self._rt = np.float32 if precision == "f32" else np.float64
        self._Req, self._Res = _structs(self._L._real)
        k = np.ascontiguousarray(k, dtype=np.float64)
        Pk = np.ascontiguousarray(Pk, dtype=np.float64)
        err = C.c_int(0)
        self._p = self._L.EZmock_ext_pool_init(int(nslot), int(mem_budget), k.ctypes.data,
                                               len(k), Pk.ctypes.data, float(z_out),
                                               float(z_pk), float(Omega_m), float(Omega_nu),
                                               float(w), C.byref(err))
        if not self._p:
            raise RuntimeError(self._L.EZmock_errmsg(err.value).decode())
        for o, v in (options or {}).items():        # EZmock.set_option of every instance
            if self._L.EZmock_ext_pool_option(self._p, int(o), int(v)):
                raise ValueError(f"pool option {o} = {v} refused")
        self._nbin = {}

    def footprint(self, Ngrid, with_powspec=False):
        return self._L.EZmock_ext_footprint(int(Ngrid), int(bool(with_powspec)))

    def submit(self, Ngrid, Lbox, seed, params, nexp, rsd_fac=0.0, pk_dk=0.0,
               pk_nbin=0, want_catalog=False):
        r = self._Req()
        r.Ngrid, r.Lbox, r.seed = int(Ngrid), float(Lbox), int(seed)
        for i, v in enumerate(np.asarray(params, dtype=self._rt)):
            r.params[i] = v
        r.nexp, r.rsd_fac, r.pk_dk, r.pk_nbin = int(nexp), float(rsd_fac), float(pk_dk), int(pk_nbin)
        r.want_catalog = int(bool(want_catalog))
        err = C.c_int(0)
        rid = self._L.EZmock_ext_pool_submit(self._p, C.byref(r), C.byref(err))
        if rid < 0:
            raise RuntimeError(self._L.EZmock_errmsg(err.value).decode())
        self._nbin[rid] = int(pk_nbin)
        return rid

    def wait(self, rid):
        """Result dict: id, err (and errmsg), ntracer, t_submit / t_start /
        t_end (steady-clock seconds), latency (s), and, when requested,
        k / P0 / P2 / nmodes and the catalogue ``cat`` = (x, y, z, vx, vy, vz)."""
        res = self._Res()
        self._L.EZmock_ext_pool_wait(self._p, int(rid), C.byref(res))
        nbin = self._nbin.pop(int(rid), 0)
        out = {"id": res.id, "err": res.err, "ntracer": res.ntracer,
               "t_submit": res.t_submit, "t_start": res.t_start, "t_end": res.t_end,
               "latency": res.t_end - res.t_submit}
        if res.err:
            out["errmsg"] = self._L.EZmock_errmsg(res.err).decode()
        if res.k and nbin:
            for name in ("k", "P0", "P2", "nmodes"):
                out[name] = np.ctypeslib.as_array(getattr(res, name), shape=(nbin,)).copy()
        if res.cat[0] and res.ntracer:
            out["cat"] = tuple(np.ctypeslib.as_array(res.cat[i], shape=(res.ntracer,)).copy()
                               for i in range(6))
        self._L.EZmock_ext_result_free(C.byref(res))
        return out

    def close(self):
        if getattr(self, "_p", None):
            self._L.EZmock_ext_pool_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
