"""MI355X-native EZmock engine: the libEZmock C API on hand-written HIP
kernels for gfx950, with a Python class mirroring the reference wrapper."""
from .EZmock import EZmock, device_count, uniform, fft_c2r, fft_r2c, pdf_walk  # noqa: F401
