"""Pins the oracle's initial field and CIC density (oracle/orc_field.c).

Closed forms per mode: with fixed amplitudes every mode has
|delta|^2 = P(|k|) / L^3 exactly and phase 2*pi*u where u is the draw of
the cell's stream position (perturb.c:112-118: draw idx-1 of the
half-complex cell idx); Hermitian and self-conjugate invariants of the
kz = 0 and Nyquist planes; psi is curl-free with div psi = -delta.
CIC: mass conservation and an independent numpy np.add.at CIC."""
import numpy as np
import pytest

from oracle import pyoracle as po


def make(Ng=16, L=200., prec="f64", seed=3, nthread=1, **kw):
    from ezmock_amd import inputs
    k, P = inputs.eh_nowiggle_pk()
    ez = po.EZmockOracle(Lbox=L, Ngrid=Ng, rng=0, seed=seed, nthread=nthread, prec=prec,
                         keep_spec=True)
    ez.eval_growth_params(0.5, 0.31)
    ez.setup_linear_pk(k, P, **kw)
    return ez


def spec3(ez):
    Ng = ez.Ng
    s = ez.field("spec").astype(np.float64)
    return (s[0::2] + 1j * s[1::2]).reshape(Ng, Ng, Ng // 2 + 1)


def kgrid(Ng):
    n = np.fft.fftfreq(Ng, 1.0 / Ng)
    n[Ng // 2] = Ng // 2
    return n


def test_fixamp_modes_have_exact_power_and_stream_phases():
    Ng, L = 16, 200.
    ez = make(Ng, L)
    ez.create_dens_field_from_ic(fixamp=True)
    H = spec3(ez)
    n = kgrid(Ng)
    kz = np.arange(Ng // 2 + 1)
    nsq = n[:, None, None] ** 2 + n[None, :, None] ** 2 + kz[None, None, :] ** 2
    kf = 2 * np.pi / L
    draws = po.Rng(0, 3, 0).double(Ng * Ng * (Ng // 2 + 1))
    for (i, j, k) in [(1, 0, 0), (0, 1, 3), (2, 5, 4), (7, 9, 1), (3, 3, 8), (8, 2, 5)]:
        idx = (i * Ng + j) * (Ng // 2 + 1) + k
        h = H[i, j, k]
        delta_amp = abs(h) * kf * nsq[i, j, k]
        P = ez.pk_interp(kf * np.sqrt(nsq[i, j, k]))
        assert delta_amp == pytest.approx(np.sqrt(P) * L ** -1.5, rel=1e-12)
        # rhok = i*|.|*exp(i*phase), phase = 2 pi u_(idx-1)
        ph = np.angle(h / 1j) % (2 * np.pi)
        assert ph == pytest.approx((2 * np.pi * draws[idx - 1]) % (2 * np.pi), abs=1e-9)


def test_random_amplitudes_follow_stream_positions():
    Ng, L = 16, 200.
    ez = make(Ng, L)
    ez.create_dens_field_from_ic()
    H = spec3(ez)
    n = kgrid(Ng)
    kf = 2 * np.pi / L
    d = po.Rng(0, 3, 0).double(2 * Ng * Ng * (Ng // 2 + 1))
    for (i, j, k) in [(1, 2, 3), (5, 0, 7), (9, 9, 2), (2, 14, 6)]:
        idx = (i * Ng + j) * (Ng // 2 + 1) + k
        nsq = n[i] ** 2 + n[j] ** 2 + k * k
        P = ez.pk_interp(kf * np.sqrt(nsq))
        amp2 = (abs(H[i, j, k]) * kf * nsq) ** 2 * L ** 3 / P
        assert amp2 == pytest.approx(-np.log(d[2 * (idx - 1)]), rel=1e-10)
        ph = np.angle(H[i, j, k] / 1j) % (2 * np.pi)
        assert ph == pytest.approx(2 * np.pi * d[2 * (idx - 1) + 1] % (2 * np.pi), abs=1e-9)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_hermitian_planes_and_self_conjugate_modes(prec):
    Ng = 16
    ez = make(Ng, prec=prec)
    ez.create_dens_field_from_ic()
    H = spec3(ez)
    Ngh = Ng // 2
    for kk in (0, Ngh):
        for i in range(Ng):
            for j in range(Ng):
                a, b = H[i, j, kk], H[-i % Ng, -j % Ng, kk]
                if (i, j) == (-i % Ng, -j % Ng):
                    assert a.real == 0
                else:
                    assert a == -np.conj(b)      # rhok is i * (Hermitian)
    assert H[0, 0, 0] == 0


def test_thread_count_independent_field():
    a = make(16, nthread=1); a.create_dens_field_from_ic()
    b = make(16, nthread=5); b.create_dens_field_from_ic()
    assert np.array_equal(a.field("spec"), b.field("spec"))
    assert np.array_equal(a.field("psi_x"), b.field("psi_x"))


def test_psi_is_curl_free_gradient_of_delta():
    Ng, L = 16, 200.
    ez = make(Ng, L)
    ez.create_dens_field_from_ic()
    H = spec3(ez)
    psi = [ez.field(f).reshape(Ng, Ng, Ng) for f in ("psi_x", "psi_y", "psi_z")]
    F = [np.fft.rfftn(p) for p in psi]
    n = kgrid(Ng)
    nx, ny = n[:, None, None], n[None, :, None]
    nz = np.arange(Ng // 2 + 1)[None, None, :]
    # interior modes (no Nyquist index): the unnormalised c2r of
    # psi_a(k) = rhok(k) * n_a transforms forward to Ng^3 psi_a(k)
    sel = (np.abs(nx) < Ng // 2) & (np.abs(ny) < Ng // 2) & (nz > 0) & (nz < Ng // 2)
    for Fa, na in zip(F, (nx, ny, nz)):
        ref = Ng ** 3 * np.broadcast_to(H * na, H.shape)[sel]
        assert np.allclose(Fa[sel], ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())
    # curl: n_x psi_y - n_y psi_x = 0
    c = (nx * F[1] - ny * F[0])[sel]
    assert np.abs(c).max() < 1e-9 * np.abs(F[0][sel]).max() * Ng
    assert abs(psi[0].mean()) < 1e-12


def cic_numpy(psi, Ng, L):
    """Independent CIC: vectorised numpy with np.add.at."""
    igs = Ng / L
    i, j, k = np.meshgrid(np.arange(Ng), np.arange(Ng), np.arange(Ng), indexing="ij")
    pos = [psi[0] * igs + i, psi[1] * igs + j, psi[2] * igs + k]
    f = [np.floor(p) for p in pos]
    w1 = [p - q for p, q in zip(pos, f)]
    b = [q.astype(int) % Ng for q in f]
    rho = np.zeros((Ng, Ng, Ng))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = ((w1[0] if dx else 1 - w1[0]) * (w1[1] if dy else 1 - w1[1])
                     * (w1[2] if dz else 1 - w1[2]))
                np.add.at(rho, ((b[0] + dx) % Ng, (b[1] + dy) % Ng, (b[2] + dz) % Ng), w)
    return rho


def test_cic_mass_and_independent_cic():
    Ng, L = 16, 200.
    ez = make(Ng, L)
    ez.create_dens_field_from_ic()
    rho = ez.field("rho").reshape(Ng, Ng, Ng)
    assert rho.sum() == pytest.approx(Ng ** 3, rel=1e-13)
    psi = [ez.field(f).reshape(Ng, Ng, Ng) for f in ("psi_x", "psi_y", "psi_z")]
    assert np.allclose(rho, cic_numpy(psi, Ng, L), atol=1e-12)


def test_cic_simple_displacements():
    Ng, L = 8, 80.
    ez = po.EZmockOracle(Lbox=L, Ngrid=Ng, rng=0, seed=1, prec="f64")
    zero = np.zeros(Ng ** 3)
    ez.create_dens_field_from_disp(zero, zero, zero, deepcopy=True)
    assert np.array_equal(ez.field("rho"), np.ones(Ng ** 3))
    half = np.full(Ng ** 3, 0.5 * L / Ng)
    ez2 = po.EZmockOracle(Lbox=L, Ngrid=Ng, rng=0, seed=1, prec="f64")
    ez2.create_dens_field_from_disp(half, zero, -half, deepcopy=True)
    assert np.allclose(ez2.field("rho"), 1.0)
    one = np.full(Ng ** 3, 1.0 * L / Ng)
    ez3 = po.EZmockOracle(Lbox=L, Ngrid=Ng, rng=0, seed=1, prec="f64")
    ez3.create_dens_field_from_disp(-one, zero, zero, deepcopy=True)   # wraps
    assert np.array_equal(ez3.field("rho"), np.ones(Ng ** 3))


def test_white_noise_path_matches_rng_path_statistics():
    Ng, L = 16, 200.
    ez = make(Ng, L)
    rs = np.random.default_rng(1)
    delta = rs.standard_normal(Ng ** 3)
    ez.create_dens_field_from_wn(delta)
    H = spec3(ez)
    D = np.fft.rfftn(delta.reshape(Ng, Ng, Ng))
    n = kgrid(Ng)
    kf = 2 * np.pi / L
    i, j, k = 3, 5, 2
    nsq = n[i] ** 2 + n[j] ** 2 + k * k
    P = ez.pk_interp(kf * np.sqrt(nsq))
    ref = 1j * np.sqrt(P) * L ** -1.5 / (nsq * kf) * D[i, j, k]
    assert H[i, j, k] == pytest.approx(ref, rel=1e-10)
