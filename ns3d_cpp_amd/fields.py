"""Synthetic initial fields of the workloads (interior arrays, (nz, ny, nx)).

The reference reads its initial field from a Tecplot file that it does not
ship (main.cpp:67), so the workloads define their own inputs:

* Taylor-Green vortex on [0, 2pi)^3 at grid points x_i = 2*pi*i/n.
* Decaying isotropic turbulence: a seeded, solenoidal, random-phase velocity
  field with the spectrum E(k) = 16 sqrt(2/pi) (u0^2/k0) (k/k0)^4
  exp(-2 (k/k0)^2), rescaled to u_rms = u0 per component, rho = 1, T = 1.
  With the reference's Mach_t = sqrt(<|u|^2>) / <c> (ana_isotropic.cpp:289),
  Mt = sqrt(3) * Ma when u_rms = 1 per component.
"""
import numpy as np


def tgv(n, Ma=0.1, gamma=1.4):
    """Taylor-Green vortex, u = sin x cos y cos z, v = -cos x sin y cos z,
    w = 0, rho = 1, p = 1/(gamma Ma^2) + (cos 2x + cos 2y)(cos 2z + 2)/16."""
    x = 2.0 * np.pi * np.arange(n) / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    u = np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -np.cos(X) * np.sin(Y) * np.cos(Z)
    w = np.zeros_like(u)
    rho = np.ones_like(u)
    p = 1.0 / (gamma * Ma * Ma) + (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0) / 16.0
    return rho, u, v, w, p


def hit_spectrum(k, k0=4.0, u0=1.0):
    return 16.0 * np.sqrt(2.0 / np.pi) * (u0 ** 2 / k0) * (k / k0) ** 4 * np.exp(-2.0 * (k / k0) ** 2)


def hit_velocity(n, k0=4.0, u0=1.0, seed=1234):
    """Solenoidal random-phase velocity with the spectrum hit_spectrum.

    Each wavevector gets a Gaussian complex vector (numpy PCG64 stream of
    `seed`) projected onto the plane normal to k and scaled to the shell
    density sqrt(E(|k|) / (4 pi |k|^2)); the velocity is the real part of the
    inverse FFT (still divergence-free, spectrally), rescaled so that
    sqrt(<u^2 + v^2 + w^2> / 3) = u0.  Modes with |k| >= n/2 are dropped.
    """
    rng = np.random.default_rng(seed)
    kf = np.fft.fftfreq(n, d=1.0 / n)
    # broadcast (z, y, x) wavenumber axes: the same elementwise arithmetic as
    # full meshgrids, without their memory
    KZ, KY, KX = kf[:, None, None], kf[None, :, None], kf[None, None, :]
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    amp = np.zeros_like(kk)
    keep = (kk > 0) & (kk < n / 2)
    amp[keep] = np.sqrt(hit_spectrum(kk[keep], k0, u0) / (4.0 * np.pi * kk[keep] ** 2))
    del keep
    a = []
    for _ in range(3):
        c = rng.standard_normal(kk.shape).astype(np.complex128)
        c += 1j * rng.standard_normal(kk.shape)
        a.append(c)
    k2 = np.where(kk > 0, kk ** 2, 1.0)
    del kk
    kdota = KX * a[0]
    kdota += KY * a[1]
    kdota += KZ * a[2]
    kdota /= k2
    del k2
    a[0] -= KX * kdota
    a[1] -= KY * kdota
    a[2] -= KZ * kdota
    del kdota
    norm = np.abs(a[0]) ** 2
    norm += np.abs(a[1]) ** 2
    norm += np.abs(a[2]) ** 2
    np.sqrt(norm, out=norm)
    norm[norm == 0] = 1.0
    vel = []
    for c in range(3):
        uh = amp * a[c]
        uh /= norm
        a[c] = None
        v = np.real(np.fft.ifftn(uh))
        del uh
        v *= n ** 3
        vel.append(v)
    u, v, w = vel
    scale = u0 / np.sqrt(np.mean(u * u + v * v + w * w) / 3.0)
    return u * scale, v * scale, w * scale


def hit(n, k0=4.0, Mt=0.3, u0=1.0, seed=1234, gamma=1.4):
    """Isotropic-turbulence state: returns ((rho,u,v,w,p), Ma) with rho = 1,
    T = 1 and p = rho T / (gamma Ma^2), Ma = Mt / sqrt(3) (u_rms = u0 = 1)."""
    u, v, w = hit_velocity(n, k0, u0, seed)
    Ma = Mt / np.sqrt(3.0) * u0
    rho = np.ones_like(u)
    p = np.full_like(u, 1.0 / (gamma * Ma * Ma))
    return (rho, u, v, w, p), Ma
