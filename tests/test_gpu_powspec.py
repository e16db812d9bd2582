"""Power-spectrum multipoles of the device catalogue (EZmock_ext_powspec).

1. The GPU estimator against a float64 numpy implementation of the same
   definition (CIC assignment with np.add.at, rfftn, half-cell interlacing,
   CIC window compensation, Legendre weights), on the same catalogue:
   FP64 engine within 1e-9 relative per bin, FP32 engine within 2e-3 of
   (|P| + shot noise) per bin.
2. FP32 against FP64: the density field's P(k) within 1e-4 (256^3); the
   tracer multipoles of the 512^3 single mock within the statistical scatter
   of two realisations of the stochastic bias (see the test's docstring).
"""
import numpy as np
import pytest

from ezmock_amd import EZmock, inputs

pytestmark = pytest.mark.gpu


def pk_numpy(cat, Ng, L, rsd_fac, interlace, dk, nbin):
    x, y, z, vx, vy, vz = [np.asarray(a, dtype=np.float64) for a in cat]
    ntr = len(x)
    if rsd_fac:
        z = z + rsd_fac * vz
        z[z >= L] -= L
        z[z < 0] += L
    pos = [x, y, z]
    spectra = []
    for shift in ([0.0, 0.5] if interlace else [0.0]):
        grid = np.zeros((Ng, Ng, Ng))
        idx, wts = [], []
        for a in range(3):
            u = pos[a] * (Ng / L) + shift
            f = np.floor(u)
            d = u - f
            i0 = f.astype(np.int64) % Ng
            idx.append((i0, (i0 + 1) % Ng))
            wts.append((1 - d, d))
        for a in range(2):
            for b in range(2):
                for c in range(2):
                    np.add.at(grid, (idx[0][a], idx[1][b], idx[2][c]),
                              wts[0][a] * wts[1][b] * wts[2][c])
        spectra.append(np.fft.rfftn(grid))
    n = np.fft.fftfreq(Ng, 1.0 / Ng).astype(np.int64)
    nk = np.arange(Ng // 2 + 1)
    NI, NJ, NK = np.meshgrid(n, n, nk, indexing="ij")
    X = spectra[0]
    if interlace:
        X = 0.5 * (X + spectra[1] * np.exp(1j * np.pi * (NI + NJ + NK) / Ng))
    W = (np.sinc(NI / Ng) * np.sinc(NJ / Ng) * np.sinc(NK / Ng)) ** 2
    P = np.abs(X) ** 2 * L ** 3 / ntr ** 2 / W ** 2
    kf = 2 * np.pi / L
    k = kf * np.sqrt(NI ** 2 + NJ ** 2 + NK ** 2)
    w = np.where((NK == 0) | (NK == Ng // 2), 1.0, 2.0)
    mu = np.divide(kf * NK, k, out=np.zeros_like(k), where=k > 0)
    L2 = 1.5 * mu ** 2 - 0.5
    ok = (k > 0) & (k < dk * nbin)
    b = np.floor(k / dk).astype(np.int64)
    ok &= b < nbin
    b, w, P, L2, kk = b[ok], w[ok], P[ok], L2[ok], k[ok]
    nm = np.bincount(b, w, nbin)
    with np.errstate(invalid="ignore", divide="ignore"):
        kmean = np.bincount(b, w * kk, nbin) / nm
        P0 = np.bincount(b, w * P, nbin) / nm - L ** 3 / ntr
        P2 = 5 * np.bincount(b, w * P * L2, nbin) / nm
    return kmean, P0, P2, nm


def _mock(prec, Ng, L, seed, params, nexp, pk):
    k, P = pk
    ez = EZmock(Lbox=L, Ngrid=Ng, rng=0, seed=seed, precision=prec)
    ez.eval_growth_params(0.5, 0.31)
    ez.setup_linear_pk(k, P)
    ez.create_dens_field_from_ic()
    ez.populate_tracer_device(*params, nexp)
    return ez


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("interlace", [False, True])
@pytest.mark.parametrize("rsd", [0.0, 0.011384])
def test_powspec_matches_numpy(gpu_available, pk_table, prec, interlace, rsd):
    Ng, L = 64, 250.0
    ez = _mock(prec, Ng, L, 5, (0.5, 10.0, 0.3, 150.0), 40000, pk_table)
    cat = ez.device_catalog()
    dk, nbin = 0.02, int(np.pi * Ng / L / 0.02)
    kg, P0g, P2g, nmg = ez.powspec(rsd_fac=rsd, interlace=interlace, dk=dk, nbin=nbin)
    kn, P0n, P2n, nmn = pk_numpy(cat, Ng, L, rsd, interlace, dk, nbin)
    assert np.array_equal(nmg, nmn)
    have = nmn > 0
    shot = L ** 3 / len(cat[0])
    if prec == "f64":
        assert np.allclose(kg[have], kn[have], rtol=1e-12, atol=0)
        for g, n in ((P0g, P0n), (P2g, P2n)):
            assert np.all(np.abs(g[have] - n[have]) <= 1e-9 * (np.abs(n[have]) + shot))
    else:
        assert np.allclose(kg[have], kn[have], rtol=1e-12, atol=0)
        for g, n in ((P0g, P0n), (P2g, P2n)):
            assert np.all(np.abs(g[have] - n[have]) <= 2e-3 * (np.abs(n[have]) + shot))


def _field_pk(rho, L, dk, nbin):
    """Monopole of a density field (CIC window not removed: both sides
    carry the same one)."""
    Ng = rho.shape[0]
    d = rho / rho.mean() - 1
    X = np.fft.rfftn(d)
    n = np.fft.fftfreq(Ng, 1.0 / Ng)
    nk = np.arange(Ng // 2 + 1)
    kf = 2 * np.pi / L
    k = kf * np.sqrt(n[:, None, None] ** 2 + n[None, :, None] ** 2 + nk[None, None, :] ** 2)
    w = np.where((nk == 0) | (nk == Ng // 2), 1.0, 2.0)[None, None, :] * np.ones_like(k)
    b = np.floor(k / dk).astype(np.int64)
    ok = (k > 0) & (b < nbin)
    P = np.abs(X) ** 2 * L ** 3 / Ng ** 6
    nm = np.bincount(b[ok], w[ok], nbin)
    return np.bincount(b[ok], (w * P)[ok], nbin) / np.maximum(nm, 1), nm


def test_density_field_pk_fp32_against_fp64(gpu_available, pk_table):
    """The FP32 density field has the FP64 one's power spectrum: P(k) of the
    CIC density of the 256^3 single mock (Lbox 1000, seed 1) within 1e-4
    relative for k < 0.4 h/Mpc."""
    res = {}
    for prec in ("f64", "f32"):
        k, P = pk_table
        ez = EZmock(Lbox=1000.0, Ngrid=256, rng=0, seed=1, precision=prec)
        ez.eval_growth_params(0.5, 0.31)
        ez.setup_linear_pk(k, P)
        ez.create_dens_field_from_ic()
        res[prec] = _field_pk(ez.field("rho").astype(np.float64).reshape((256,) * 3),
                              1000.0, 0.01, 40)
        del ez
    (Pd, nm), (Ps, _) = res["f64"], res["f32"]
    sel = nm > 0
    dev = np.abs(Ps[sel] - Pd[sel]) / Pd[sel]
    print(f"density field, FP32 vs FP64, k < 0.4: max |dP|/P {dev.max():.2e}")
    assert dev.max() <= 1e-4


def test_powspec_fp32_against_fp64_single_mock(gpu_available, pk_table):
    """Tracer multipoles of the 512^3 single mock, FP32 vs FP64 engine.
    The two catalogues are different realisations of the stochastic bias:
    a cell whose density rounds to the other side of rho_c shifts the rank of
    every later valid cell, hence its Box-Muller pair (pop_tracer.c:236,256),
    exactly as between the reference's own FP32 and FP64 builds.  Bins of
    0.02 h/Mpc from 0.02 to 0.4: each difference within 3 sigma of the
    difference of two independent realisations, sigma = (P0 + P_shot)
    * 2 / sqrt(N_modes) (P2: 5 times that), and the monopoles within 2%."""
    Ng, L, seed = 512, 2500.0, 42
    params = (0.8, 8.0, 0.25, 200.0)
    rsd = 0.011384
    dk, nbin = 0.02, 20
    res = {}
    for prec in ("f64", "f32"):
        ez = _mock(prec, Ng, L, seed, params, 3125000, pk_table)
        res[prec] = ez.powspec(rsd_fac=rsd, interlace=True, dk=dk, nbin=nbin)
        del ez
    k, P0d, P2d, nm = res["f64"]
    _, P0s, P2s, _ = res["f32"]
    sel = (nm > 0) & (k >= 0.02) & (k < 0.4)
    shot = L ** 3 / 3125000
    sig = (P0d[sel] + shot) * 2 / np.sqrt(nm[sel])
    d0 = np.abs(P0s[sel] - P0d[sel])
    d2 = np.abs(P2s[sel] - P2d[sel])
    print(f"tracers, FP32 vs FP64, 0.02 <= k < 0.4: max |dP0|/P0 "
          f"{(d0 / P0d[sel]).max():.2e}, max |dP0|/sigma {(d0 / sig).max():.2f}, "
          f"max |dP2|/(5 sigma) {(d2 / (5 * sig)).max():.2f}")
    assert np.all(d0 <= 3 * sig)
    assert np.all(d2 <= 3 * 5 * sig)
    assert np.all(d0 / P0d[sel] <= 0.02)
    assert np.all(P2d[sel][:5] > 0)          # Kaiser quadrupole of a biased sample
