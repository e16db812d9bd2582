"""Pins the oracle's 3-D real FFTs (oracle/orc_fft.c) against numpy.fft
golden vectors (tests/golden/fft_golden.npz, made by make_fft_golden.py), a
brute-force DFT and closed forms (delta, plane wave, Parseval)."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz"))
TOL = {"f64": 1e-12, "f32": 2e-6}


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("Ng", [4, 6, 8, 12, 16])
def test_c2r_golden(Ng, prec):
    H, ref = GOLD[f"c2r_in_{Ng}"], GOLD[f"c2r_out_{Ng}"]
    out = po.fft_c2r_3d(H, prec)
    assert np.abs(out - ref).max() <= TOL[prec] * np.abs(ref).max()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("Ng", [4, 6, 8, 12, 16])
def test_r2c_golden(Ng, prec):
    x, ref = GOLD[f"r2c_in_{Ng}"], GOLD[f"r2c_out_{Ng}"]
    out = po.fft_r2c_3d(x, prec)
    assert np.abs(out - ref).max() <= TOL[prec] * np.abs(ref).max()


def test_c2r_brute_force_drops_dc_nyquist_imaginary_parts():
    """Direct sum over the Hermitian-extended spectrum; the kz = 0 and
    kz = Ng/2 terms enter through their real parts only."""
    Ng = 4
    rs = np.random.default_rng(5)
    H = rs.standard_normal((Ng, Ng, 3)) + 1j * rs.standard_normal((Ng, Ng, 3))
    out = po.fft_c2r_3d(H, "f64")
    x = np.arange(Ng)
    ref = np.zeros((Ng, Ng, Ng))
    for i in range(Ng):
        for j in range(Ng):
            for k in range(3):
                ph = np.exp(2j * np.pi * (i * x[:, None, None] + j * x[None, :, None]
                                          + k * x[None, None, :]) / Ng)
                if k in (0, 2):
                    ref += (H[i, j, k] * ph).real
                else:
                    ref += 2 * (H[i, j, k] * ph).real
    assert np.allclose(out, ref, atol=1e-11)


def test_c2r_plane_wave_and_delta():
    Ng = 16
    H = np.zeros((Ng, Ng, Ng // 2 + 1), complex)
    H[0, 0, 0] = 1.0
    assert np.allclose(po.fft_c2r_3d(H, "f64"), 1.0)
    H[:] = 0
    H[3, 0, 2] = 0.5
    xs = np.arange(Ng)
    ref = np.cos(2 * np.pi * (3 * xs[:, None, None] + 2 * xs[None, None, :]) / Ng) * np.ones((1, Ng, 1))
    assert np.allclose(po.fft_c2r_3d(H, "f64"), ref, atol=1e-12)


def test_r2c_parseval_and_roundtrip():
    Ng = 12
    rs = np.random.default_rng(9)
    x = rs.standard_normal((Ng, Ng, Ng))
    X = po.fft_r2c_3d(x, "f64")
    w = np.full(Ng // 2 + 1, 2.0)
    w[0] = 1
    w[-1] = 1
    assert np.sum(w * np.abs(X) ** 2) / Ng ** 3 == pytest.approx(np.sum(x ** 2), rel=1e-11)
    back = po.fft_c2r_3d(X, "f64") / Ng ** 3
    assert np.allclose(back, x, atol=1e-12)


@pytest.mark.parametrize("n", [2, 3, 5, 7, 12, 30, 64, 96, 100])
def test_1d_against_numpy(n):
    rs = np.random.default_rng(n)
    z = rs.standard_normal(n) + 1j * rs.standard_normal(n)
    d = np.stack([z.real, z.imag], -1).ravel().astype(np.float64)
    assert po.lib("f64").orc_fft_1d(n, d.ctypes.data, -1) == 0
    out = d[0::2] + 1j * d[1::2]
    assert np.allclose(out, np.fft.fft(z), atol=1e-12)
