"""Synthetic inputs for the EZmock workloads.

The reference takes a tabulated linear power spectrum from the caller
(``EZmock_setup_linear_pk``, linear_pk.c:75) and has no P(k) code of its own.
The benchmark workloads need one, so this module tabulates the
Eisenstein & Hu (1998, ApJ 496, 605; eqs. 26-31) no-wiggle transfer function
with a power-law primordial spectrum, normalised to sigma_8, in float64.

Workload parameters: Omega_m = 0.31, h = 0.6766, n_s = 0.97,
sigma_8 = 0.81, Omega_b h^2 = 0.02242, T_CMB = 2.7255 K, 1024 log-spaced k in
[1e-4, 10] h/Mpc.
"""
import numpy as np

PLANCK_LIKE = dict(Omega_m=0.31, h=0.6766, n_s=0.97, sigma8=0.81,
                   ombh2=0.02242, Tcmb=2.7255)


def eh_nowiggle_transfer(k, Omega_m, h, ombh2, Tcmb):
    """No-wiggle transfer function T(k), k in h/Mpc (EH98 eqs. 26-31)."""
    k = np.asarray(k, dtype=np.float64)
    omh2 = Omega_m * h * h
    fb = ombh2 / omh2
    theta = Tcmb / 2.7
    s = 44.5 * np.log(9.83 / omh2) / np.sqrt(1.0 + 10.0 * ombh2 ** 0.75)
    alpha = 1.0 - 0.328 * np.log(431.0 * omh2) * fb + 0.38 * np.log(22.3 * omh2) * fb * fb
    kmpc = k * h
    gamma_eff = Omega_m * h * (alpha + (1.0 - alpha) / (1.0 + (0.43 * kmpc * s) ** 4))
    q = k * theta * theta / gamma_eff
    L0 = np.log(2.0 * np.e + 1.8 * q)
    C0 = 14.2 + 731.0 / (1.0 + 62.5 * q)
    return L0 / (L0 + C0 * q * q)


def sigma_R(k, P, R=8.0):
    """RMS of the top-hat filtered field (trapezoid in ln k)."""
    x = k * R
    W = 3.0 * (np.sin(x) - x * np.cos(x)) / x ** 3
    integrand = P * W * W * k ** 3 / (2.0 * np.pi ** 2)
    lk = np.log(k)
    return np.sqrt(np.sum(0.5 * (integrand[1:] + integrand[:-1]) * np.diff(lk)))


def eh_nowiggle_pk(nk=1024, kmin=1e-4, kmax=10.0, **cosmo):
    """(k, P) at z = 0: 1024 log-spaced k in [1e-4, 10] h/Mpc by default."""
    c = dict(PLANCK_LIKE)
    c.update(cosmo)
    k = np.logspace(np.log10(kmin), np.log10(kmax), nk)
    T = eh_nowiggle_transfer(k, c["Omega_m"], c["h"], c["ombh2"], c["Tcmb"])
    P = k ** c["n_s"] * T * T
    # normalise on a dense grid so the sigma_8 integral is converged
    kd = np.logspace(-5, 2, 20000)
    Td = eh_nowiggle_transfer(kd, c["Omega_m"], c["h"], c["ombh2"], c["Tcmb"])
    Pd = kd ** c["n_s"] * Td * Td
    A = (c["sigma8"] / sigma_R(kd, Pd)) ** 2
    return k, A * P


def rsd_factor(z, Omega_m, w=-1.0):
    """(1+z) / (100 E(z)) for flat wCDM: the factor that turns a peculiar
    velocity in km/s into a line-of-sight shift in Mpc/h."""
    a = 1.0 / (1.0 + z)
    E = np.sqrt(Omega_m * a ** -3 + (1.0 - Omega_m) * a ** (-3.0 * (1.0 + w)))
    return (1.0 + z) / (100.0 * E)


# The single-mock workloads (grid, box, seed, bias parameters, tracer density).
WORKLOADS = {
    "c1_256_f64": dict(Ng=256, Lbox=1000.0, z=0.5, seed=1, rng=0,
                       params=(0.5, 10.0, 0.3, 150.0), density=3e-4),
    "c2_512": dict(Ng=512, Lbox=2500.0, z=0.5, seed=42, rng=0,
                   params=(0.8, 8.0, 0.25, 200.0), density=2e-4),
    "c_1024": dict(Ng=1024, Lbox=2500.0, z=0.5, seed=42, rng=0,
                   params=(0.8, 8.0, 0.25, 200.0), density=2e-4),
    "c4_1536": dict(Ng=1536, Lbox=3000.0, z=1.0, seed=7, rng=0,
                    params=(0.8, 8.0, 0.3, 200.0), density=5e-4),
}


def nexp_of(workload):
    w = WORKLOADS[workload] if isinstance(workload, str) else workload
    return int(round(w["density"] * w["Lbox"] ** 3))
