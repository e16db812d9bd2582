"""Writes tests/golden/fft_golden.npz: small 3-D real FFT vectors from
numpy.fft (pocketfft), the published definition the oracle's and the GPU
engine's transforms are checked against (FFTW's c2r/r2c compute the same
unnormalised sums).  Run: python tests/golden/make_fft_golden.py"""
import os

import numpy as np

out = {}
rs = np.random.default_rng(20240611)
for Ng in (4, 6, 8, 12, 16):
    H = rs.standard_normal((Ng, Ng, Ng // 2 + 1)) + 1j * rs.standard_normal((Ng, Ng, Ng // 2 + 1))
    out[f"c2r_in_{Ng}"] = H
    out[f"c2r_out_{Ng}"] = np.fft.irfftn(H, s=(Ng, Ng, Ng), axes=(0, 1, 2)) * Ng ** 3
    x = rs.standard_normal((Ng, Ng, Ng))
    out[f"r2c_in_{Ng}"] = x
    out[f"r2c_out_{Ng}"] = np.fft.rfftn(x, axes=(0, 1, 2))
np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.npz"), **out)
print("wrote", len(out), "arrays")
