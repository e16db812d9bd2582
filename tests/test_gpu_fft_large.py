"""The FFT passes at the large grid sizes (their kernels are separate
instantiations per size: 768, 1024, 1536, 2048, plus 512 and 96 for
reference), through the generic pass sequence and through the production
first pass of the three displacement components (the c2r's of
perturb.c:361,397-398), against plane-wave closed forms evaluated on the
device: out(x) = sum_m 2 w_m Re(A_m exp(2 pi i n_m.x / Ng)), w_m = 1
(generic) or the mode's n_x, n_y, n_z (production).  Modes include
negative i, j wavenumbers and the i = Ng/2 Nyquist row.  Tolerance: the
largest error within 2e-6 (FP32) / 1e-13 (FP64) of the largest value, times
log2(Ng)."""
import math

import numpy as np
import pytest

from ezmock_amd import EZmock
from ezmock_amd.EZmock import fft_check

pytestmark = pytest.mark.gpu


def modes_for(Ng):
    h = Ng // 2
    return ([(3, 5, 7), (Ng - 2, 1, h - 1), (h, Ng - 7, 1), (1, h, 2)],
            [1.0 + 0.5j, -0.75 + 0.25j, 0.5 - 1.0j, 0.3 + 0.2j])


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("Ng", [96, 512, 768, 1024, 1536, 2048])
def test_plane_waves_large_grids(gpu_available, Ng, prec):
    if prec == "f64" and Ng > 1536:
        pytest.skip("2048^3 FP64 needs 275 GB for the production path")
    modes, amps = modes_for(Ng)
    tol = (2e-6 if prec == "f32" else 1e-13) * math.log2(Ng)
    for path in (0, 1):
        err, ref = fft_check(Ng, path, modes, amps, prec)
        print(f"Ng {Ng} {prec} path {path}: max err / max value {err / ref}")
        assert np.all(ref > 1.0)
        assert np.all(err <= tol * ref), (path, err / ref)


def test_fft_check_rejects_bad_modes(gpu_available):
    with pytest.raises(RuntimeError):
        fft_check(64, 0, [(1, 1, 0)], [1.0])        # kk = 0 is not a free mode
    with pytest.raises(RuntimeError):
        fft_check(64, 0, [(1, 1, 32)], [1.0])       # kk = Nyquist
    with pytest.raises(RuntimeError):
        fft_check(40, 0, [(1, 1, 2)], [1.0])        # unsupported size
