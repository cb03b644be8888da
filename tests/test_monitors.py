"""The monitors' checker (oracle/monitors.py): the energy spectrum of EnergySpectrum.cpp:62-154 and
the statistics of ana_isotropic.cpp:226-330, pinned by closed forms
(single Fourier modes, Parseval), by a brute-force DFT on a tiny grid and by
the shell rule (floor of |k|)."""
import itertools
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import monitors  # noqa: E402  (the monitors' checker, oracle/monitors.py)


def grid(n):
    x = np.arange(n) * 2 * np.pi / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    return X, Y, Z


def test_single_mode_spectrum():
    n = 16
    X, Y, Z = grid(n)
    k, ek = monitors.energy_spectrum(np.sin(X), 0 * X, 0 * X)
    assert len(k) == 2 * n + 1
    assert abs(ek[1] - 0.25) < 1e-15
    assert np.sum(np.abs(np.delete(ek, 1))) < 1e-28


def test_shell_is_floor_of_wavenumber_magnitude():
    n = 16
    X, Y, Z = grid(n)
    # |k| = sqrt(2) -> shell 1; |k| = 3 (2,2,1) -> shell 3; |k| = sqrt(13) -> 3
    for (a, b, c), shell in (((1, 1, 0), 1), ((2, 2, 1), 3), ((3, 2, 0), 3), ((0, 0, 4), 4)):
        u = np.cos(a * X + b * Y + c * Z)
        k, ek = monitors.energy_spectrum(u, 0 * u, 0 * u)
        assert abs(ek[shell] - 0.25) < 1e-14, (a, b, c)
        assert abs(np.sum(ek) - 0.25) < 1e-14


def test_nyquist_mode_is_labelled_positive():
    n = 8
    X, Y, Z = grid(n)
    u = np.cos(4 * X)            # the Nyquist mode, i = N/2 -> k1 = +4
    k, ek = monitors.energy_spectrum(u, 0 * u, 0 * u)
    assert abs(ek[4] - 0.5) < 1e-14


def test_parseval():
    rng = np.random.default_rng(0)
    u, v, w = (rng.normal(size=(12, 12, 12)) for _ in range(3))
    k, ek = monitors.energy_spectrum(u, v, w)
    assert abs(np.sum(ek) - 0.5 * np.mean(u * u + v * v + w * w)) < 1e-13


def test_brute_force_dft_on_tiny_grid():
    n = 4
    rng = np.random.default_rng(1)
    vel = [rng.normal(size=(n, n, n)) for _ in range(3)]
    kmax = 2 * n
    ref = np.zeros(kmax + 1)
    idx = list(itertools.product(range(n), repeat=3))
    for (k, j, i) in idx:       # mode (i, j, k) in the reference's loop order
        k1 = i if i <= n // 2 else i - n
        k2 = j if j <= n // 2 else j - n
        k3 = k if k <= n // 2 else k - n
        shell = int(int(np.sqrt(1.0 * (k1 * k1 + k2 * k2 + k3 * k3))) + 0.5)
        e = 0.0
        for a in vel:
            s = 0j
            for (z, y, x) in idx:
                s += a[z, y, x] * np.exp(-2j * np.pi * (i * x + j * y + k * z) / n)
            s /= n ** 3
            e += abs(s) ** 2
        ref[shell] += 0.5 * e
    _, ek = monitors.energy_spectrum(*vel)
    assert np.allclose(ek, ref, rtol=1e-12, atol=1e-15)


class _P:
    gamma = 1.4
    Rgas = 1.0 / (1.4 * 0.1 ** 2)
    mu = 1.0 / 100.0
    mu_law = 1          # constant
    mu_power = 0.76
    sutherland_s = 110.4 / 273.15


def test_turbulence_statistics_closed_forms():
    n = 16
    X, Y, Z = grid(n)
    u = np.sin(X) * np.cos(Y)
    v = -np.cos(X) * np.sin(Y)
    w = 0 * X
    rho = np.ones_like(u)
    T = np.ones_like(u)
    du_dx = np.cos(X) * np.cos(Y)
    st = monitors.turbulence_statistics(_P, rho, u, v, w, T, du_dx)
    ke = 0.5 * np.mean(u * u + v * v)
    assert abs(st["kinetic_energy"] - 0.25) < 1e-14 and abs(ke - 0.25) < 1e-14
    assert abs(st["u_rms"] - np.sqrt(0.5 / 3.0)) < 1e-14
    assert abs(st["tke"] - 0.25) < 1e-14          # zero-mean field
    nu = _P.mu
    # every mode of the field has |k|^2 = 2 -> shell 1, E(1) = KE
    assert abs(st["dissipation"] - 2.0 * nu * 1.0 * 0.25) < 1e-15
    c = np.sqrt(_P.gamma * _P.Rgas)
    assert abs(st["Mach_t"] - np.sqrt(0.5) / c) < 1e-14
    assert abs(st["taylor_Li"] - np.sqrt(0.5 / 3.0) / 0.5) < 1e-14
