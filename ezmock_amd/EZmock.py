"""Python front-end of the MI355X EZmock engine.

``class EZmock`` keeps the constructor keywords, method names, argument
meaning, dtype casts and ``RuntimeError(errmsg)`` mapping of the reference's
Cython class (python/EZmock/EZmock.pyx:36-485), so that code written for the
reference runs unchanged:

    from ezmock_amd import EZmock
    ez = EZmock(Lbox=1000, Ngrid=256, rng=0, seed=1)
    ez.eval_growth_params(0.5, 0.31)
    ez.setup_linear_pk(k, P)
    ez.create_dens_field_from_ic()
    x, y, z, vx, vy, vz = ez.populate_tracer(0.5, 10, 0.3, 150, 300000)

It calls the C ABI of ``ezmock_amd/lib/libEZmock_{f32,f64}.so`` through
ctypes; ``precision='f32'`` (default, as the reference's wrapper is built
with SINGLE_PREC, setup.py:56) or ``'f64'``.  The hot path runs on the GPU;
the extra methods prefixed ``device_`` / ``field`` expose the
device-resident results.
"""
import ctypes as C

import numpy as np

from . import _lib


class EZmock:
    def __init__(self, *args, **kwargs):
        # EZmock.pyx:42-84
        if 'Lbox' in kwargs:
            Lbox = float(kwargs['Lbox'])
        else:
            raise KeyError('`Lbox` is a required argument')
        if 'Ngrid' in kwargs:
            Ngrid = int(kwargs['Ngrid'])
        else:
            raise KeyError('`Ngrid` is a required argument')
        rng = int(kwargs.get('rng', 1))
        seed = int(kwargs.get('seed', 1))
        nthread = int(kwargs.get('nthread', 0))
        self._prec = kwargs.get('precision', 'f32')
        self._L = _lib.load(self._prec)
        self._rt = np.float32 if self._prec == 'f32' else np.float64
        self._ct = C.c_float if self._prec == 'f32' else C.c_double
        self._err = C.c_int(0)
        self._Ng = Ngrid
        self._Lbox = Lbox
        self._ez = self._L.EZmock_init(Lbox, Ngrid, rng, seed, nthread,
                                       C.byref(self._err))
        if not self._ez:
            raise RuntimeError(self._errmsg(self._err.value))

    def __del__(self):
        ez = getattr(self, '_ez', None)
        if ez:
            self._L.EZmock_destroy(ez)
            self._ez = None

    def _errmsg(self, code):
        return self._L.EZmock_errmsg(code).decode('utf-8')

    def _assert_return_value(self):
        if self._err.value != 0:
            raise RuntimeError(self._errmsg(self._err.value))

    # EZmock.pyx:103-122
    def set_growth_params(self, growth_pk, vel_fac):
        self._err.value = self._L.EZmock_set_cosmology(
            self._ez, float(growth_pk), float(vel_fac), False, 0., 0., 0., 0., 0.,
            C.byref(self._err))
        self._assert_return_value()

    # EZmock.pyx:125-148
    def eval_growth_params(self, z_out, Omega_m, z_pk=0, Omega_nu=0, w=-1):
        self._err.value = self._L.EZmock_set_cosmology(
            self._ez, 0., 0., True, float(z_out), float(z_pk), float(Omega_m),
            float(Omega_nu), float(w), C.byref(self._err))
        self._assert_return_value()

    # EZmock.pyx:151-206
    def setup_linear_pk(self, k, Plin, Pnw=None, BAO_enhance=0, interp_log=False):
        try:
            k = np.asarray(k).astype(np.float64)
            Plin = np.asarray(Plin).astype(np.float64)
        except Exception:
            raise TypeError('`k` and `Plin` are expected to be float numpy arrays')
        assert k.ndim == 1 and Plin.ndim == 1 and k.shape[0] == Plin.shape[0], \
            '`k` and `Plin` must be 1d arrays with the same length'
        k = np.ascontiguousarray(k)
        Plin = np.ascontiguousarray(Plin)
        pnw = None
        if BAO_enhance != 0:
            try:
                Pnw = np.asarray(Pnw).astype(np.float64)
            except Exception:
                raise TypeError('`Pnw` is expected to be a float numpy array')
            assert Pnw.ndim == 1 and k.shape[0] == Pnw.shape[0], \
                '`Pnw` must be an 1d array with the same length as k'
            pnw = np.ascontiguousarray(Pnw)
        self._err.value = self._L.EZmock_setup_linear_pk(
            self._ez, k.ctypes.data, k.shape[0], Plin.ctypes.data,
            pnw.ctypes.data if pnw is not None else None, float(BAO_enhance),
            bool(interp_log), C.byref(self._err))
        self._assert_return_value()

    # EZmock.pyx:209-229
    def create_dens_field_from_ic(self, fixamp=False, iphase=False):
        self._err.value = self._L.EZmock_create_dens_field(
            self._ez, None, False, None, bool(fixamp), bool(iphase), C.byref(self._err))
        self._assert_return_value()

    # EZmock.pyx:232-264
    def create_dens_field_from_wn(self, delta):
        try:
            delta = np.asarray(delta).astype(self._rt)
        except Exception:
            raise TypeError('`delta` is expected to be a float numpy array')
        assert delta.ndim == 1 and delta.shape[0] == self._Ng ** 3, \
            f'`delta` must be a 1d array with {self._Ng:d}**3 elements'
        delta = np.ascontiguousarray(delta)
        self._err.value = self._L.EZmock_create_dens_field(
            self._ez, None, False, delta.ctypes.data, False, False, C.byref(self._err))
        self._assert_return_value()

    # EZmock.pyx:267-320
    def create_dens_field_from_disp(self, dx, dy, dz, deepcopy=False):
        try:
            d = [np.asarray(a).astype(self._rt) for a in (dx, dy, dz)]
        except Exception:
            raise TypeError('`dx`, `dy`, and `dz` must be float numpy arrays')
        Ntot = self._Ng ** 3
        assert all(a.ndim == 1 and a.shape[0] == Ntot for a in d), \
            f'`dx`, `dy`, `dz` must be 1d arrays with {self._Ng:d}**3 elements'
        d = [np.ascontiguousarray(a) for a in d]
        if not deepcopy:
            self._disp_ref = d          # the library keeps only the pointers
        arr = (C.c_void_p * 3)(*[a.ctypes.data for a in d])
        self._err.value = self._L.EZmock_create_dens_field(
            self._ez, arr, bool(deepcopy), None, False, False, C.byref(self._err))
        self._assert_return_value()

    def _wrap(self, ptr, n):
        """numpy array owning a malloc'd buffer (freed with free())."""
        libc = C.CDLL(None)
        libc.free.argtypes = [C.c_void_p]
        if n == 0:
            libc.free(ptr)
            return np.empty(0, dtype=self._rt)
        buf = (self._ct * n).from_address(ptr)
        a = np.frombuffer(buf, dtype=self._rt)
        # release with free() when the array (and its buffer) dies
        import weakref
        weakref.finalize(buf, libc.free, ptr)
        return a

    # EZmock.pyx:333-407
    def populate_tracer(self, rho_c, rho_exp, pdf_base, sigma_v, ntracer, att_part=False):
        params = np.array([rho_c, rho_exp, pdf_base, sigma_v], dtype=self._rt)
        n = C.c_size_t(0)
        ptrs = [C.c_void_p() for _ in range(6)]
        self._err.value = self._L.EZmock_populate_tracer(
            self._ez, params.ctypes.data, int(ntracer), bool(att_part), C.byref(n),
            *[C.byref(p) for p in ptrs], C.byref(self._err))
        self._assert_return_value()
        return tuple(self._wrap(p.value, n.value) for p in ptrs)

    # EZmock.pyx:410-485
    def populate_tracer_to_file(self, rho_c, rho_exp, pdf_base, sigma_v, ntracer, fname,
                                att_part=False, rsd_fac=0, header=True):
        params = np.array([rho_c, rho_exp, pdf_base, sigma_v], dtype=self._rt)
        n = C.c_size_t(0)
        ptrs = [C.c_void_p() for _ in range(6)]
        self._err.value = self._L.EZmock_populate_tracer(
            self._ez, params.ctypes.data, int(ntracer), bool(att_part), C.byref(n),
            *[C.byref(p) for p in ptrs], C.byref(self._err))
        self._assert_return_value()
        libc = C.CDLL(None)
        libc.free.argtypes = [C.c_void_p]
        try:
            self._err.value = self._L.EZmock_write_ascii(
                self._ez, *[p.value for p in ptrs], n.value, float(rsd_fac), bool(header),
                str(fname).encode('utf-8'), C.byref(self._err))
            self._assert_return_value()
        finally:
            for p in ptrs:
                libc.free(p.value)

    # ------------------------------------------------------ extensions
    def populate_tracer_device(self, rho_c, rho_exp, pdf_base, sigma_v, ntracer,
                               att_part=False):
        """Populate and keep the catalogue on the GPU; returns its size."""
        params = np.array([rho_c, rho_exp, pdf_base, sigma_v], dtype=self._rt)
        n = C.c_size_t(0)
        self._err.value = self._L.EZmock_ext_populate_device(
            self._ez, params.ctypes.data, int(ntracer), bool(att_part), C.byref(n),
            C.byref(self._err))
        self._assert_return_value()
        return n.value

    def mock(self, rho_c, rho_exp, pdf_base, sigma_v, ntracer, fixamp=False, iphase=False):
        """create_dens_field_from_ic + populate_tracer_device in one call,
        replayed from captured HIP graphs after the first; returns the
        catalogue size (catalogue on the GPU)."""
        params = np.array([rho_c, rho_exp, pdf_base, sigma_v], dtype=self._rt)
        n = C.c_size_t(0)
        self._err.value = self._L.EZmock_ext_mock(
            self._ez, params.ctypes.data, int(ntracer), bool(fixamp), bool(iphase),
            C.byref(n), C.byref(self._err))
        self._assert_return_value()
        return n.value

    def path_counts(self):
        """How this instance's calls ran: device-selection populates,
        host-driven populates, capacity redos, graph replays, captures, and
        the candidate cells of the last device selection."""
        o = np.zeros(6, dtype=np.uint64)
        self._L.EZmock_ext_path_counts(self._ez, o.ctypes.data)
        return dict(devsel=int(o[0]), host=int(o[1]), retry=int(o[2]), replay=int(o[3]),
                    capture=int(o[4]), candidates=int(o[5]))

    def device_catalog(self):
        """Copy the device-resident catalogue to host numpy arrays."""
        ptrs = (C.c_void_p * 6)()
        n = C.c_size_t(0)
        self._L.EZmock_ext_device_tracers(self._ez, ptrs, C.byref(n), C.byref(self._err))
        self._assert_return_value()
        out = [np.empty(n.value, dtype=self._rt) for _ in range(6)]
        self._err.value = self._L.EZmock_ext_copy_tracers(
            self._ez, *[a.ctypes.data for a in out], n.value, C.byref(self._err))
        self._assert_return_value()
        return tuple(out)

    def device_pointers(self):
        ptrs = (C.c_void_p * 6)()
        n = C.c_size_t(0)
        self._L.EZmock_ext_device_tracers(self._ez, ptrs, C.byref(n), C.byref(self._err))
        self._assert_return_value()
        return [p for p in ptrs], n.value

    def rsd_z(self, rsd_fac):
        """Redshift-space z of the device catalogue (float64, on the GPU)."""
        _, n = self.device_pointers()
        out = np.empty(n, dtype=np.float64)
        self._err.value = self._L.EZmock_ext_rsd_z(self._ez, float(rsd_fac),
                                                   out.ctypes.data, C.byref(self._err))
        self._assert_return_value()
        return out

    def powspec(self, rsd_fac=0.0, interlace=True, dk=None, nbin=None):
        """P0 and P2 of the device catalogue (CIC, optional interlacing,
        line of sight z; rsd_fac > 0 shifts z to redshift space).  Default
        bins: width k_f = 2 pi / Lbox up to the Nyquist frequency.
        Returns (k_mean, P0, P2, nmodes) as float64 arrays."""
        kf = 2 * np.pi / self._Lbox
        if dk is None:
            dk = kf
        if nbin is None:
            nbin = min(1024, int(np.pi * self._Ng / self._Lbox / dk))
        out = [np.zeros(int(nbin), dtype=np.float64) for _ in range(4)]
        self._err.value = self._L.EZmock_ext_powspec(
            self._ez, float(rsd_fac), bool(interlace), float(dk), int(nbin),
            *[a.ctypes.data for a in out], C.byref(self._err))
        self._assert_return_value()
        return tuple(out)

    def field(self, which):
        N = self._Ng ** 3
        codes = {"psi_x": 0, "psi_y": 1, "psi_z": 2, "rho": 3, "rhot": 4}
        if which in codes:
            code, a = codes[which], np.empty(N, dtype=self._rt)
        elif which == "nt":
            code, a = 5, np.empty(N, dtype=np.int16)
        elif which == "spec":
            code, a = 6, np.empty(self._Ng * self._Ng * (self._Ng // 2 + 1) * 2, dtype=self._rt)
        elif which == "hist":
            code, a = 7, np.empty(64, dtype=np.uint64)
        else:
            raise KeyError(which)
        self._err.value = self._L.EZmock_ext_copy_field(self._ez, code, a.ctypes.data,
                                                        C.byref(self._err))
        self._assert_return_value()
        return a

    def pdf_info(self):
        out = np.zeros(5 + 1025, dtype=np.uint64)
        self._L.EZmock_ext_pdf_info(self._ez, out.ctypes.data, len(out))
        nmax = int(out[0])
        return dict(nmax=nmax, ncell=int(out[1]), ntot=int(out[2]), ncell0=int(out[3]),
                    ntot0=int(out[4]), pdf=out[5:5 + nmax].astype(np.int64))

    def info(self):
        o = np.zeros(8)
        self._L.EZmock_ext_info(self._ez, o.ctypes.data)
        return dict(Ng=int(o[0]), Lbox=o[1], growth2=o[2], vfac=o[3], pk_n=int(o[4]),
                    kmin=o[5], kmax=o[6], pk_log=bool(o[7]))

    def pk_table(self):
        n = self.info()["pk_n"]
        k, P, P2 = np.zeros(n), np.zeros(n), np.zeros(n)
        self._L.EZmock_ext_pk_table(self._ez, k.ctypes.data, P.ctypes.data, P2.ctypes.data, n)
        return k, P, P2

    def set_seed(self, seed):
        """Seed of the random streams of the next field and catalogues (the
        instance is kept: its device buffers, FFT schedule and, for MT19937,
        jump windows of an unchanged seed are reused)."""
        self._L.EZmock_ext_set_seed(self._ez, int(seed), C.byref(self._err))
        self._assert_return_value()

    def set_option(self, option, value):
        return self._L.EZmock_ext_set_option(self._ez, int(option), int(value))

    def keep_spectrum(self, on=True):
        self.set_option(1, on)

    def profile(self, on=True):
        self.set_option(2, on)

    def cic_method(self, method):
        """0: particle binning into fixed list ranges (default; exact ranges
        after an overflow); 1: particle binning with exact ranges; 2: tile
        scan of the displaced lattice wherever the displacement range admits
        it, else binning."""
        if self.set_option(3, int(method)):
            raise ValueError(f"unknown CIC method {method}")

    def kernel_times(self):
        n = self._L.EZmock_ext_kernel_times(self._ez, None, None, 0, None, 0)
        ms = np.zeros(max(n, 1))
        cnt = np.zeros(max(n, 1), dtype=np.int64)
        names = C.create_string_buffer(8192)
        self._L.EZmock_ext_kernel_times(self._ez, ms.ctypes.data, cnt.ctypes.data, n,
                                        names, 8192)
        keys = names.value.decode().split("\n")[:n]
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(keys)}

    def reset_kernel_times(self):
        self._L.EZmock_ext_reset_kernel_times(self._ez)

    def sync(self):
        self._L.EZmock_ext_sync(self._ez, C.byref(self._err))
        self._assert_return_value()


def device_count(precision="f32"):
    return _lib.load(precision).EZmock_ext_device_count()


def uniform(seed, start, n, rng=0, precision="f32", pos=False):
    """GPU draws of stream `rng` (reset(seed, start), then n get_double, or
    get_double_pos with pos=True)."""
    L = _lib.load(precision)
    out = np.empty(int(n), dtype=np.float64)
    err = C.c_int(0)
    L.EZmock_ext_uniform(int(rng), int(seed), int(start), int(n), bool(pos), out.ctypes.data,
                         C.byref(err))
    if err.value:
        raise RuntimeError(L.EZmock_errmsg(err.value).decode())
    return out


def fft_c2r(H, precision="f32"):
    """GPU 3-D c2r of a (Ng, Ng, Ng//2+1) complex array (unnormalised)."""
    L = _lib.load(precision)
    rt = np.float32 if precision == "f32" else np.float64
    Ng = H.shape[0]
    inp = np.ascontiguousarray(np.stack([H.real, H.imag], axis=-1).astype(rt))
    out = np.empty((Ng, Ng, Ng), dtype=rt)
    err = C.c_int(0)
    L.EZmock_ext_fft_c2r(Ng, inp.ctypes.data, out.ctypes.data, C.byref(err))
    if err.value:
        raise RuntimeError(L.EZmock_errmsg(err.value).decode())
    return out


def fft_r2c(x, precision="f32"):
    L = _lib.load(precision)
    rt = np.float32 if precision == "f32" else np.float64
    Ng = x.shape[0]
    inp = np.ascontiguousarray(x.astype(rt))
    out = np.empty((Ng, Ng, Ng // 2 + 1, 2), dtype=rt)
    err = C.c_int(0)
    L.EZmock_ext_fft_r2c(Ng, inp.ctypes.data, out.ctypes.data, C.byref(err))
    if err.value:
        raise RuntimeError(L.EZmock_errmsg(err.value).decode())
    return out[..., 0] + 1j * out[..., 1]


def fft_schedule(Ng, precision="f32"):
    """FFT schedule of grid size Ng in this process (EZmock_ext_fft_schedule):
    dict(tk_j, tk_i, split, tuned, tk_g)."""
    L = _lib.load(precision)
    out = np.zeros(5, dtype=np.int32)
    if L.EZmock_ext_fft_schedule(int(Ng), out.ctypes.data):
        raise RuntimeError(f"no FFT schedule for grid size {Ng}")
    return dict(tk_j=int(out[0]), tk_i=int(out[1]), split=int(out[2]), tuned=bool(out[3]),
                tk_g=int(out[4]))


def fft_set_schedule(Ng, tk_j, tk_i, split, precision="f32"):
    L = _lib.load(precision)
    if L.EZmock_ext_fft_set_schedule(int(Ng), int(tk_j), int(tk_i), int(split)):
        raise ValueError("invalid FFT schedule")


def fft_check(Ng, path, modes, amps, precision="f32"):
    """Plane waves through the GPU c2r passes, checked on the device
    (EZmock_ext_fft_check): path 0 = generic passes (one output), path 1 =
    production first pass of the three displacement components.  modes:
    [(i, j, kk), ...] with 0 < kk < Ng/2; amps: complex amplitudes.
    Returns (max abs error, max abs closed form) per output."""
    L = _lib.load(precision)
    nout = 3 if path else 1
    ijk = np.ascontiguousarray(np.asarray(modes, dtype=np.int32).reshape(-1))
    a = np.ascontiguousarray(np.stack([np.real(amps), np.imag(amps)], -1).astype(np.float64))
    err_, ref = np.zeros(nout), np.zeros(nout)
    err = C.c_int(0)
    L.EZmock_ext_fft_check(int(Ng), int(path), len(modes), ijk.ctypes.data, a.ctypes.data,
                           err_.ctypes.data, ref.ctypes.data, C.byref(err))
    if err.value:
        raise RuntimeError(L.EZmock_errmsg(err.value).decode())
    return err_, ref


def pdf_walk(rhot, ncell, pdf, precision="f32"):
    """GPU PDF walk on an arbitrary tracer-density array."""
    L = _lib.load(precision)
    rt = np.float32 if precision == "f32" else np.float64
    rhot = np.ascontiguousarray(rhot, dtype=rt)
    pdf_in = np.ascontiguousarray(pdf, dtype=np.uint64)
    nmax = len(pdf_in)
    nt = np.zeros(len(rhot), dtype=np.int16)
    pdf_out = np.zeros(nmax, dtype=np.uint64)
    nc, ntot = C.c_size_t(), C.c_size_t()
    err = C.c_int(0)
    L.EZmock_ext_pdf_walk(rhot.ctypes.data, len(rhot), int(ncell), pdf_in.ctypes.data, nmax,
                          nt.ctypes.data, pdf_out.ctypes.data, C.byref(nc), C.byref(ntot),
                          C.byref(err))
    return err.value, nt, pdf_out.astype(np.int64), nc.value, ntot.value
