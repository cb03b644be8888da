"""Checker of the device monitors: the reference's post-processing restated
in numpy on fields downloaded off the device.  Test infrastructure, like the
C oracle beside it: imported only by tests/ as the checker of the device
spectrum, statistics and dissipation; the product computes its monitors on
the device (ns3d_monitors.hip).

* ``energy_spectrum`` restates ``compute_energy_spectrum_rank0``
  (EnergySpectrum.cpp:62-154) with numpy's FFT in place of FFTW: forward DFT
  normalised by 1/N, E = (|U|^2 + |V|^2 + |W|^2) / 2 per mode, summed over
  shells of integer |k| = floor(sqrt(k1^2 + k2^2 + k3^2)) (the reference's
  ``(int) sqrt(...) + 0.5`` truncates twice, which is floor), wave numbers
  ``i <= N/2 ? i : i - N``, bins 0 .. Kmax = 2 NX.
* ``turbulence_statistics`` restates ``compute_turbulence_statistics``
  (ana_isotropic.cpp:226-330): volume means, u_rms, TKE, Taylor scale from
  <(du/dx)^2>, Mach_t = sqrt(2 KE) / <c>, nu = <mu(T)> / <rho>, spectral
  dissipation 2 nu sum k^2 E(k) and the Kolmogorov scales.
* ``dissipation_history`` / ``kinetic_energy`` use the device reductions.
"""
import numpy as np


def energy_spectrum(u, v, w):
    """(k, E(k)) for interior velocity arrays of shape (nz, ny, nx)."""
    u, v, w = (np.asarray(a, dtype=np.float64) for a in (u, v, w))
    nz, ny, nx = u.shape
    n = u.size
    e = np.zeros(u.shape)
    for a in (u, v, w):
        f = np.fft.fftn(a) / n
        e += f.real ** 2 + f.imag ** 2
    e *= 0.5

    def wav(m):
        i = np.arange(m)
        return np.where(i <= m // 2, i, i - m)

    k3, k2, k1 = np.meshgrid(wav(nz), wav(ny), wav(nx), indexing="ij")
    kk = np.floor(np.sqrt((k1 * k1 + k2 * k2 + k3 * k3).astype(np.float64))).astype(np.int64)
    kmax = 2 * nx
    ek = np.bincount(kk.ravel(), weights=e.ravel(), minlength=kmax + 1)[: kmax + 1]
    return np.arange(kmax + 1), ek


def mu_of_T(params, T):
    """SolverParams::get_mu (field_structures.h:94-105) and the extra laws."""
    law = int(params.mu_law)
    if law == 1:
        return np.full_like(T, params.mu)
    if law == 2:
        return params.mu * T ** params.mu_power
    S = params.sutherland_s
    return params.mu * T ** 1.5 * (1.0 + S) / (T + S)


def turbulence_statistics(params, rho, u, v, w, T, du_dx, spectrum=None):
    """The reference's isotropic-turbulence statistics from interior arrays."""
    n = u.size
    q2 = u * u + v * v + w * w
    ke = 0.5 * float(np.sum(q2)) / n
    u_rms = np.sqrt(float(np.sum(q2)) / (3 * n))
    mu_mean = float(np.sum(mu_of_T(params, T))) / n
    rho_mean = float(np.sum(rho)) / n
    nu = mu_mean / rho_mean
    var = [float(np.sum(a * a)) / n - (float(np.sum(a)) / n) ** 2 for a in (u, v, w)]
    tke = 0.5 * sum(var)
    taylor_li = u_rms / np.sqrt(float(np.sum(du_dx * du_dx)) / n)
    c_mean = float(np.sum(np.sqrt(params.gamma * params.Rgas * T))) / n
    mach_t = np.sqrt(2.0 * ke) / c_mean
    if spectrum is None:
        spectrum = energy_spectrum(u, v, w)
    k, ek = spectrum
    eps = float(np.sum(2.0 * nu * k * k * ek))
    return {
        "kinetic_energy": ke, "tke": tke, "u_rms": u_rms, "dissipation": eps,
        "taylor": np.sqrt(15.0 * nu * u_rms * u_rms / eps) if eps > 0 else float("inf"),
        "eta": (nu ** 3 / eps) ** 0.25 if eps > 0 else float("inf"),
        "u_eta": (nu * eps) ** 0.25, "tau_eta": (nu / eps) ** 0.5 if eps > 0 else float("inf"),
        "taylor_Li": taylor_li, "Re_lambda": u_rms * taylor_li / nu, "Mach_t": mach_t,
        "mean_mu": mu_mean, "mean_rho": rho_mean,
    }


def solver_statistics(solver):
    """turbulence_statistics of a Solver's current state (one download)."""
    pr = solver.primitive()
    g = solver.compute_gradients() if not _is_mixed(solver) else None
    it = solver.interior
    if g is None:
        # the mixed path has no separate gradient entry point: x-derivative
        # with the same central scheme on the downloaded primitives
        du_dx = _diff_x(it(pr["u"]), solver)
    else:
        du_dx = it(g[0])
    return turbulence_statistics(solver.params, it(pr["rho"]), it(pr["u"]), it(pr["v"]), it(pr["w"]),
                                 it(pr["T"]), du_dx)


def _is_mixed(solver):
    return int(solver.params.precision) == 1


# central first-derivative coefficients of ViscousFlux.cpp:20-40 (periodic)
_CENTRAL = {
    0: [0.8, -0.2, 0.0380952380952381, -0.00357142857142857],
    1: [45.0 / 60.0, -9.0 / 60.0, 1.0 / 60.0],
    2: [8.0 / 12.0, -1.0 / 12.0],
}


def _diff_x(a, solver):
    c = _CENTRAL[int(solver.params.vis_scheme)]
    d = np.zeros_like(a)
    for m, cm in enumerate(c, start=1):
        d += cm * (np.roll(a, -m, axis=2) - np.roll(a, m, axis=2))
    return d / solver.grid.dx
