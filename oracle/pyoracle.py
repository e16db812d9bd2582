"""ctypes front-end of the CPU oracle (test infrastructure only).

Loads ``oracle/liboracle_f32.so`` / ``liboracle_f64.so`` (built by
``oracle/Makefile``) and mirrors the reference's Python wrapper methods
(python/EZmock/EZmock.pyx:36-485) closely enough that parity tests read the
same on both sides.  Only tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


def build(quiet=True):
    import subprocess
    out = subprocess.run(["make", "-B", "-C", _HERE], capture_output=True,
                         text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def lib(prec="f32"):
    if prec in _LIBS:
        return _LIBS[prec]
    path = os.path.join(_HERE, f"liboracle_{prec}.so")
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    vp, i32, u64, dbl, sz = C.c_void_p, C.c_int, C.c_uint64, C.c_double, C.c_size_t
    pint = C.POINTER(C.c_int)
    L.orc_init.restype = vp
    L.orc_init.argtypes = [dbl, i32, i32, u64, i32, pint]
    L.orc_set_cosmology.argtypes = [vp, dbl, dbl, C.c_bool, dbl, dbl, dbl, dbl, dbl, pint]
    L.orc_setup_linear_pk.argtypes = [vp, vp, i32, vp, vp, dbl, C.c_bool, pint]
    L.orc_create_dens_field.argtypes = [vp, vp, C.c_bool, vp, C.c_bool, C.c_bool, pint]
    L.orc_populate_tracer.argtypes = [vp, vp, sz, C.c_bool, C.POINTER(sz)] + [C.POINTER(vp)] * 6 + [pint]
    L.orc_destroy.argtypes = [vp]
    L.orc_errmsg.restype = C.c_char_p
    L.orc_errmsg.argtypes = [i32]
    L.orc_free.argtypes = [vp]
    L.orc_rsd_z.restype = dbl
    L.orc_get_info.argtypes = [vp, vp]
    L.orc_copy_field.argtypes = [vp, i32, vp]
    L.orc_pk_nbin.restype = sz
    L.orc_pk_nbin.argtypes = [vp]
    L.orc_pk_table.argtypes = [vp, vp, vp, vp]
    L.orc_pk_interp.restype = dbl
    L.orc_pk_interp.argtypes = [vp, dbl]
    L.orc_pdf_info.argtypes = [vp, vp, sz]
    L.orc_eval_pdf.argtypes = [sz, C.c_float if prec == "f32" else C.c_double, sz,
                               C.POINTER(sz), C.POINTER(sz), C.POINTER(sz), vp, sz]
    L.orc_stage_time.restype = dbl
    L.orc_stage_time.argtypes = [vp, i32]
    L.orc_set_keep_spec.argtypes = [vp, i32]
    L.orc_pdf_walk.argtypes = [vp, sz, sz, vp, i32, vp, vp, C.POINTER(sz), C.POINTER(sz)]
    L.orc_rng_reset.argtypes = [vp, i32, u64, u64]
    L.orc_rng_double.restype = dbl
    L.orc_rng_double.argtypes = [vp]
    L.orc_rng_double_pos.restype = dbl
    L.orc_rng_double_pos.argtypes = [vp]
    L.orc_mrg_set_state.argtypes = [vp, vp]
    L.orc_mrg_get_state.argtypes = [vp, vp]
    L.orc_mrg_jump.argtypes = [vp, u64]
    L.orc_mrg_matpow.argtypes = [i32, i32, vp]
    L.orc_mt_next.restype = C.c_uint32
    L.orc_mt_next.argtypes = [vp]
    L.orc_hyp2f1.argtypes = [dbl, dbl, dbl, dbl, dbl, i32, C.POINTER(dbl)]
    L.orc_cspline_ypp.argtypes = [vp, vp, sz, vp]
    L.orc_cspline_eval.restype = dbl
    L.orc_cspline_eval.argtypes = [vp, vp, vp, dbl, sz]
    L.orc_fft_c2r_3d.argtypes = [i32, vp, vp, i32]
    L.orc_fft_r2c_3d.argtypes = [i32, vp, vp, i32]
    L.orc_fft_1d.argtypes = [i32, vp, i32]
    _LIBS[prec] = L
    return L


def rtype(prec):
    return np.float32 if prec == "f32" else np.float64


class Rng:
    """One oracle random stream: reset(seed, step), then draws."""

    SIZE = 6 * 8 + 624 * 4 + 16

    def __init__(self, kind, seed, step=0, prec="f64"):
        self.L = lib(prec)
        self.buf = C.create_string_buffer(self.SIZE + 64)
        if self.L.orc_rng_reset(self.buf, kind, seed, step):
            raise ValueError("bad rng type")

    def double(self, n=1):
        return np.array([self.L.orc_rng_double(self.buf) for _ in range(n)])

    def double_pos(self, n=1):
        return np.array([self.L.orc_rng_double_pos(self.buf) for _ in range(n)])


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class EZmockOracle:
    """Same methods and argument meaning as the reference's Python class."""

    def __init__(self, Lbox, Ngrid, rng=1, seed=1, nthread=0, prec="f32",
                 keep_spec=False):
        self.prec = prec
        self.L = lib(prec)
        self.rt = rtype(prec)
        self.Ng = int(Ngrid)
        self.Lbox = float(Lbox)
        self._err = C.c_int(0)
        self._ez = self.L.orc_init(self.Lbox, self.Ng, int(rng), int(seed),
                                   int(nthread), C.byref(self._err))
        if not self._ez:
            raise OracleError(self._err.value, self.errmsg(self._err.value))
        self.L.orc_set_keep_spec(self._ez, int(keep_spec))

    def errmsg(self, code):
        return self.L.orc_errmsg(code).decode()

    def _check(self):
        if self._err.value != 0:
            raise OracleError(self._err.value, self.errmsg(self._err.value))

    def close(self):
        if self._ez:
            self.L.orc_destroy(self._ez)
            self._ez = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_growth_params(self, growth_pk, vel_fac):
        r = self.L.orc_set_cosmology(self._ez, growth_pk, vel_fac, False, 0, 0, 0, 0, 0,
                                     C.byref(self._err))
        self._err.value = r
        self._check()

    def eval_growth_params(self, z_out, Omega_m, z_pk=0, Omega_nu=0, w=-1):
        r = self.L.orc_set_cosmology(self._ez, 0.0, 0.0, True, z_out, z_pk, Omega_m,
                                     Omega_nu, w, C.byref(self._err))
        self._err.value = r
        self._check()

    def setup_linear_pk(self, k, Plin, Pnw=None, BAO_enhance=0, interp_log=False):
        k = np.ascontiguousarray(k, dtype=np.float64)
        Plin = np.ascontiguousarray(Plin, dtype=np.float64)
        pnw = None
        if BAO_enhance != 0:
            pnw = np.ascontiguousarray(Pnw, dtype=np.float64)
        self.L.orc_setup_linear_pk(self._ez, k.ctypes.data, len(k), Plin.ctypes.data,
                                   pnw.ctypes.data if pnw is not None else None,
                                   float(BAO_enhance), bool(interp_log), C.byref(self._err))
        self._check()

    def create_dens_field_from_ic(self, fixamp=False, iphase=False):
        self.L.orc_create_dens_field(self._ez, None, False, None, bool(fixamp),
                                     bool(iphase), C.byref(self._err))
        self._check()

    def create_dens_field_from_wn(self, delta):
        delta = np.ascontiguousarray(delta, dtype=self.rt)
        self.L.orc_create_dens_field(self._ez, None, False, delta.ctypes.data, False,
                                     False, C.byref(self._err))
        self._check()

    def create_dens_field_from_disp(self, dx, dy, dz, deepcopy=True):
        self._disp = [np.ascontiguousarray(a, dtype=self.rt) for a in (dx, dy, dz)]
        arr = (C.c_void_p * 3)(*[a.ctypes.data for a in self._disp])
        self.L.orc_create_dens_field(self._ez, arr, bool(deepcopy), None, False, False,
                                     C.byref(self._err))
        self._check()

    def populate_tracer(self, rho_c, rho_exp, pdf_base, sigma_v, ntracer, att_part=False):
        params = np.array([rho_c, rho_exp, pdf_base, sigma_v], dtype=self.rt)
        n = C.c_size_t(0)
        ptrs = [C.c_void_p() for _ in range(6)]
        self.L.orc_populate_tracer(self._ez, params.ctypes.data, int(ntracer),
                                   bool(att_part), C.byref(n),
                                   *[C.byref(p) for p in ptrs], C.byref(self._err))
        self._check()
        out = []
        ct = C.c_float if self.prec == "f32" else C.c_double
        for p in ptrs:
            a = np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), shape=(n.value,)).copy()
            self.L.orc_free(p)
            out.append(a)
        return tuple(out)

    # ---- test accessors
    def field(self, which):
        N = self.Ng ** 3
        if which in ("psi_x", "psi_y", "psi_z", "rho", "rhot"):
            code = {"psi_x": 0, "psi_y": 1, "psi_z": 2, "rho": 3, "rhot": 4}[which]
            a = np.empty(N, dtype=self.rt)
        elif which == "nt":
            code, a = 5, np.empty(N, dtype=np.int16)
        elif which == "spec":
            code, a = 6, np.empty(self.Ng * self.Ng * (self.Ng // 2 + 1) * 2, dtype=self.rt)
        elif which == "hist":
            code, a = 7, np.empty(64, dtype=np.uint64)
        else:
            raise KeyError(which)
        if self.L.orc_copy_field(self._ez, code, a.ctypes.data):
            raise RuntimeError(f"field {which} not available")
        return a

    def pdf_info(self):
        out = np.zeros(5 + 1025, dtype=np.uint64)
        self.L.orc_pdf_info(self._ez, out.ctypes.data, len(out))
        nmax = int(out[0])
        return dict(nmax=nmax, ncell=int(out[1]), ntot=int(out[2]), ncell0=int(out[3]),
                    ntot0=int(out[4]), pdf=out[5:5 + nmax].astype(np.int64))

    def info(self):
        o = np.zeros(8)
        self.L.orc_get_info(self._ez, o.ctypes.data)
        return dict(Ng=int(o[0]), Lbox=o[1], growth2=o[2], vfac=o[3], nthread=int(o[4]),
                    nvalid=int(o[5]), hist_imax=int(o[6]), iv=int(o[7]))

    def pk_interp(self, k):
        return self.L.orc_pk_interp(self._ez, float(k))

    def stage_times(self):
        names = ["noise", "fft", "cic", "bias", "walk", "pos", "vel", "sigv"]
        return {n: self.L.orc_stage_time(self._ez, i) for i, n in enumerate(names)}

    def rsd_z(self, z, vz, rsd_fac):
        L = self.L
        ct = C.c_float if self.prec == "f32" else C.c_double
        L.orc_rsd_z.argtypes = [C.c_void_p, ct, ct, C.c_double]
        return np.array([L.orc_rsd_z(self._ez, float(a), float(b), rsd_fac)
                         for a, b in zip(z, vz)])


def eval_pdf(nexp, pbase, ncells, prec="f32"):
    L = lib(prec)
    nmax, ncell, ntot = C.c_size_t(), C.c_size_t(), C.c_size_t()
    num = np.zeros(1100, dtype=np.uint64)
    e = L.orc_eval_pdf(int(nexp), float(pbase), int(ncells), C.byref(nmax), C.byref(ncell),
                       C.byref(ntot), num.ctypes.data, len(num))
    if e:
        return e, None
    return 0, dict(nmax=nmax.value, ncell=ncell.value, ntot=ntot.value,
                   pdf=num[:nmax.value].astype(np.int64))


def pdf_walk(rhot, ncell, pdf, prec="f32"):
    """Literal reference walk on an arbitrary tracer-density array."""
    L = lib(prec)
    rhot = np.ascontiguousarray(rhot, dtype=rtype(prec))
    pdf_in = np.ascontiguousarray(pdf, dtype=np.uint64)
    nmax = len(pdf_in)
    nt = np.zeros(len(rhot), dtype=np.int16)
    pdf_out = np.zeros(nmax, dtype=np.uint64)
    nc, ntot = C.c_size_t(), C.c_size_t()
    e = L.orc_pdf_walk(rhot.ctypes.data, len(rhot), int(ncell), pdf_in.ctypes.data, nmax,
                       nt.ctypes.data, pdf_out.ctypes.data, C.byref(nc), C.byref(ntot))
    return e, nt, pdf_out.astype(np.int64), nc.value, ntot.value


def fft_c2r_3d(H, prec="f64", nthread=1):
    """H: complex array (Ng, Ng, Ng//2+1) -> real (Ng, Ng, Ng)."""
    L = lib(prec)
    Ng = H.shape[0]
    rt = rtype(prec)
    inp = np.ascontiguousarray(np.stack([H.real, H.imag], axis=-1).astype(rt))
    out = np.empty((Ng, Ng, Ng), dtype=rt)
    if L.orc_fft_c2r_3d(Ng, inp.ctypes.data, out.ctypes.data, nthread):
        raise RuntimeError("fft failed")
    return out


def fft_r2c_3d(x, prec="f64", nthread=1):
    L = lib(prec)
    Ng = x.shape[0]
    rt = rtype(prec)
    inp = np.ascontiguousarray(x.astype(rt))
    out = np.empty((Ng, Ng, Ng // 2 + 1, 2), dtype=rt)
    if L.orc_fft_r2c_3d(Ng, inp.ctypes.data, out.ctypes.data, nthread):
        raise RuntimeError("fft failed")
    return out[..., 0] + 1j * out[..., 1]
