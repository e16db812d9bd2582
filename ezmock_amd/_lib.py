"""Loader of the engine's C-ABI libraries (``ezmock_amd/lib``).

``libEZmock_f32.so`` (real = float, the reference default) and
``libEZmock_f64.so`` (real = double) export the eight functions of
``include/EZmock.h`` plus the extensions of ``include/EZmock_ext.h``.
There is no fallback: if the library is missing, loading fails with an error
that says how to build it.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(_HERE, "lib")
_CACHE = {}

# Every symbol the public headers declare (checked by the CPU test suite).
API_SYMBOLS = [
    "EZmock_init", "EZmock_set_cosmology", "EZmock_setup_linear_pk",
    "EZmock_create_dens_field", "EZmock_populate_tracer", "EZmock_write_ascii",
    "EZmock_destroy", "EZmock_errmsg",
]
EXT_SYMBOLS = [
    "EZmock_ext_real_size", "EZmock_ext_device_count", "EZmock_ext_populate_device",
    "EZmock_ext_mock", "EZmock_ext_path_counts",
    "EZmock_ext_copy_tracers", "EZmock_ext_device_tracers", "EZmock_ext_rsd_z",
    "EZmock_ext_copy_field", "EZmock_ext_pdf_info", "EZmock_ext_set_option",
    "EZmock_ext_kernel_times", "EZmock_ext_reset_kernel_times", "EZmock_ext_sync",
    "EZmock_ext_uniform", "EZmock_ext_fft_c2r", "EZmock_ext_fft_r2c",
    "EZmock_ext_pdf_walk", "EZmock_ext_info", "EZmock_ext_pk_table",
    "EZmock_ext_powspec", "EZmock_ext_set_seed", "EZmock_ext_footprint",
    "EZmock_ext_pool_init", "EZmock_ext_pool_submit", "EZmock_ext_pool_wait",
    "EZmock_ext_pool_option",
    "EZmock_ext_result_free", "EZmock_ext_pool_destroy", "EZmock_ext_fft_check",
    "EZmock_ext_fft_schedule", "EZmock_ext_fft_set_schedule",
    "EZmock_ext_noalloc_scratch_bytes", "EZmock_ext_pdf_walk_noalloc",
]


def lib_path(precision="f32"):
    return os.path.join(LIBDIR, f"libEZmock_{precision}.so")


def load(precision="f32"):
    if precision in _CACHE:
        return _CACHE[precision]
    if precision not in ("f32", "f64"):
        raise ValueError("precision must be 'f32' or 'f64'")
    path = lib_path(precision)
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; "
            f"g.build()'` (or `make -C ezmock_amd/csrc`)")
    L = C.CDLL(path)
    rt = C.c_float if precision == "f32" else C.c_double
    vp, i32, u64, dbl, sz, b = C.c_void_p, C.c_int, C.c_uint64, C.c_double, C.c_size_t, C.c_bool
    pint = C.POINTER(C.c_int)
    L.EZmock_init.restype = vp
    L.EZmock_init.argtypes = [dbl, i32, i32, u64, i32, pint]
    L.EZmock_set_cosmology.restype = i32
    L.EZmock_set_cosmology.argtypes = [vp, dbl, dbl, b, dbl, dbl, dbl, dbl, dbl, pint]
    L.EZmock_setup_linear_pk.restype = i32
    L.EZmock_setup_linear_pk.argtypes = [vp, vp, i32, vp, vp, dbl, b, pint]
    L.EZmock_create_dens_field.restype = i32
    L.EZmock_create_dens_field.argtypes = [vp, vp, b, vp, b, b, pint]
    L.EZmock_populate_tracer.restype = i32
    L.EZmock_populate_tracer.argtypes = [vp, vp, sz, b, C.POINTER(sz)] + [C.POINTER(vp)] * 6 + [pint]
    L.EZmock_write_ascii.restype = i32
    L.EZmock_write_ascii.argtypes = [vp] + [vp] * 6 + [sz, dbl, b, C.c_char_p, pint]
    L.EZmock_destroy.restype = None
    L.EZmock_destroy.argtypes = [vp]
    L.EZmock_errmsg.restype = C.c_char_p
    L.EZmock_errmsg.argtypes = [i32]
    L.EZmock_ext_real_size.restype = i32
    L.EZmock_ext_device_count.restype = i32
    L.EZmock_ext_populate_device.restype = i32
    L.EZmock_ext_populate_device.argtypes = [vp, vp, sz, b, C.POINTER(sz), pint]
    L.EZmock_ext_path_counts.restype = i32
    L.EZmock_ext_path_counts.argtypes = [vp, vp]
    L.EZmock_ext_mock.restype = i32
    L.EZmock_ext_mock.argtypes = [vp, vp, sz, b, b, C.POINTER(sz), pint]
    L.EZmock_ext_copy_tracers.restype = i32
    L.EZmock_ext_copy_tracers.argtypes = [vp] + [vp] * 6 + [sz, pint]
    L.EZmock_ext_device_tracers.restype = i32
    L.EZmock_ext_device_tracers.argtypes = [vp, vp, C.POINTER(sz), pint]
    L.EZmock_ext_rsd_z.restype = i32
    L.EZmock_ext_rsd_z.argtypes = [vp, dbl, vp, pint]
    L.EZmock_ext_copy_field.restype = i32
    L.EZmock_ext_copy_field.argtypes = [vp, i32, vp, pint]
    L.EZmock_ext_pdf_info.restype = i32
    L.EZmock_ext_pdf_info.argtypes = [vp, vp, sz]
    L.EZmock_ext_set_option.restype = i32
    L.EZmock_ext_set_option.argtypes = [vp, i32, C.c_long]
    L.EZmock_ext_kernel_times.restype = i32
    L.EZmock_ext_kernel_times.argtypes = [vp, vp, vp, i32, C.c_char_p, sz]
    L.EZmock_ext_reset_kernel_times.restype = None
    L.EZmock_ext_reset_kernel_times.argtypes = [vp]
    L.EZmock_ext_sync.restype = i32
    L.EZmock_ext_sync.argtypes = [vp, pint]
    L.EZmock_ext_uniform.restype = i32
    L.EZmock_ext_uniform.argtypes = [i32, u64, u64, sz, b, vp, pint]
    L.EZmock_ext_fft_c2r.restype = i32
    L.EZmock_ext_fft_c2r.argtypes = [i32, vp, vp, pint]
    L.EZmock_ext_fft_schedule.restype = i32
    L.EZmock_ext_fft_schedule.argtypes = [i32, vp]
    L.EZmock_ext_fft_set_schedule.restype = i32
    L.EZmock_ext_fft_set_schedule.argtypes = [i32, i32, i32, i32]
    L.EZmock_ext_fft_check.restype = i32
    L.EZmock_ext_fft_check.argtypes = [i32, i32, i32, vp, vp, vp, vp, pint]
    L.EZmock_ext_fft_r2c.restype = i32
    L.EZmock_ext_fft_r2c.argtypes = [i32, vp, vp, pint]
    L.EZmock_ext_pdf_walk.restype = i32
    L.EZmock_ext_pdf_walk.argtypes = [vp, sz, sz, vp, i32, vp, vp, C.POINTER(sz),
                                      C.POINTER(sz), pint]
    L.EZmock_ext_info.restype = i32
    L.EZmock_ext_info.argtypes = [vp, vp]
    L.EZmock_ext_pk_table.restype = i32
    L.EZmock_ext_pk_table.argtypes = [vp, vp, vp, vp, i32]
    L.EZmock_ext_powspec.restype = i32
    L.EZmock_ext_powspec.argtypes = [vp, dbl, b, dbl, i32, vp, vp, vp, vp, pint]
    L.EZmock_ext_set_seed.restype = i32
    L.EZmock_ext_set_seed.argtypes = [vp, u64, pint]
    L.EZmock_ext_footprint.restype = sz
    L.EZmock_ext_footprint.argtypes = [i32, i32]
    L.EZmock_ext_pool_init.restype = vp
    L.EZmock_ext_pool_init.argtypes = [i32, sz, vp, i32, vp, dbl, dbl, dbl, dbl, dbl, pint]
    L.EZmock_ext_pool_option.restype = i32
    L.EZmock_ext_pool_option.argtypes = [vp, i32, C.c_long]
    L.EZmock_ext_pool_submit.restype = C.c_long
    L.EZmock_ext_pool_submit.argtypes = [vp, vp, pint]
    L.EZmock_ext_pool_wait.restype = i32
    L.EZmock_ext_pool_wait.argtypes = [vp, C.c_long, vp]
    L.EZmock_ext_result_free.restype = None
    L.EZmock_ext_result_free.argtypes = [vp]
    L.EZmock_ext_pool_destroy.restype = None
    L.EZmock_ext_pool_destroy.argtypes = [vp]
    L._real = rt
    _CACHE[precision] = L
    return L
