"""ns3d_cpp_amd -- MI355X RHS + RK3 engine for the 3D compressible
Navier-Stokes solver of wylx-2/ns3d_cpp.

The product is the C ABI in include/ns3d_gpu.h, implemented by
libns3d_gpu.so (HIP kernels for gfx950).  This module is a thin ctypes
host wrapper over that ABI; it never computes anything on the CPU, and it
raises if the library is missing.
"""
import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import BC, FVS, MU, PREC, RECON, VIS, Grid, Params

__all__ = ["NS3DError", "Solver", "make_params", "Params", "Grid", "lib"]


class NS3DError(RuntimeError):
    def __init__(self, status, msg=""):
        self.status = status
        super().__init__(f"ns3d status {status}: {msg}")


def lib():
    return _lib.load()


def _enum(table, v):
    return table[v] if isinstance(v, str) else int(v)


def make_params(n, L=2.0 * np.pi, **kw):
    """Params/Grid for an n^3 (or (nx,ny,nz)) periodic box of edge L.

    Keyword arguments override SolverParams members by their solver.in
    meaning; enum members accept the solver.in spelling ("WENO5",
    "StagerWarming", "C8th", ...).  Returns finalized (Params, Grid).
    """
    L_ = lib()
    p, g = Params(), Grid()
    L_.ns3d_params_default(C.byref(p), C.byref(g))
    if isinstance(n, int):
        n = (n, n, n)
    g.nx, g.ny, g.nz = n
    g.Lx = g.Ly = g.Lz = float(L)
    for k, v in kw.items():
        if k == "fvs_type":
            p.fvs_type = _enum(FVS, v)
        elif k == "recon":
            p.recon = _enum(RECON, v)
        elif k == "vis_scheme":
            p.vis_scheme = _enum(VIS, v)
        elif k == "mu_law":
            p.mu_law = _enum(MU, v)
        elif k == "precision":
            p.precision = _enum(PREC, v)
        elif k == "bc":
            for f in range(6):
                p.bc[f] = _enum(BC, v[f])
        elif hasattr(p, k):
            setattr(p, k, v)
        elif hasattr(g, k):
            setattr(g, k, v)
        else:
            raise KeyError(k)
    st = L_.ns3d_params_finalize(C.byref(p), C.byref(g))
    if st:
        raise NS3DError(st, "ns3d_params_finalize")
    return p, g


def split_batch(params, W, axis):
    """split_StagerWarming / split_LaxFriedrichs of n states on the device:
    W = (rho, u, v, w, p) arrays of n; returns (Fp, Fm), 5 arrays each."""
    W = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in W]
    n = W[0].size
    Fp = [np.empty(n) for _ in range(5)]
    Fm = [np.empty(n) for _ in range(5)]
    arr = C.c_void_p * 5
    nvec = [1 if axis == d else 0 for d in range(3)]
    st = lib().ns3d_split_batch(C.byref(params), n, arr(*[a.ctypes.data for a in W]), *nvec,
                                arr(*[a.ctypes.data for a in Fp]), arr(*[a.ctypes.data for a in Fm]))
    if st:
        raise NS3DError(st, "ns3d_split_batch")
    return Fp, Fm


def reconstruct_batch(params, stencils, flag):
    """reconstruct_select of n stencils (an (n, width) array) on the device."""
    st_ = np.ascontiguousarray(stencils, dtype=np.float64)
    n, width = st_.shape
    out = np.empty(n)
    st = lib().ns3d_reconstruct_batch(C.byref(params), n, width, st_.ctypes.data, float(flag), out.ctypes.data)
    if st:
        raise NS3DError(st, "ns3d_reconstruct_batch")
    return out


def read_params(path):
    """(Params, Grid) of a solver.in file (read_solver_params_from_file,
    initialization.cpp:199-380), finalized."""
    p, g = Params(), Grid()
    st = lib().ns3d_params_from_file(os.fsencode(path), C.byref(p), C.byref(g))
    if st:
        raise NS3DError(st, f"ns3d_params_from_file({path})")
    return p, g


def _ptr(a):
    return a.ctypes.data_as(_lib.DP)


class Solver:
    """One device-resident periodic box (one ns3d_ctx)."""

    def __init__(self, params, grid):
        self._L = lib()
        ctx = C.c_void_p()
        st = self._L.ns3d_create(C.byref(params), C.byref(grid), C.byref(ctx))
        if st:
            raise NS3DError(st, self._L.ns3d_status_string(st).decode())
        self.ctx = ctx
        self.params = params
        self.grid = grid
        sx, sy, sz, ng = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self._L.ns3d_sizes(ctx, C.byref(sx), C.byref(sy), C.byref(sz), C.byref(ng))
        self.shape = (sz.value, sy.value, sx.value)
        self.ng = ng.value

    # -- helpers -----------------------------------------------------------
    def _check(self, st, what):
        if st:
            err = self._L.ns3d_last_error(self.ctx).decode() if self.ctx else ""
            raise NS3DError(st, f"{what}: {self._L.ns3d_status_string(st).decode()} {err}")

    def _full(self, a):
        """Interior (nz,ny,nx) or full (sz,sy,sx) array -> contiguous full array."""
        a = np.asarray(a, dtype=np.float64)
        if a.shape == self.shape:
            return np.ascontiguousarray(a)
        out = np.zeros(self.shape)
        ng = self.ng
        out[ng:-ng, ng:-ng, ng:-ng] = a
        return out

    def interior(self, a):
        ng = self.ng
        return a[ng:-ng, ng:-ng, ng:-ng]

    def close(self):
        if self.ctx:
            self._L.ns3d_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state -------------------------------------------------------------
    def upload_primitive(self, rho, u, v, w, p):
        arrs = [self._full(x) for x in (rho, u, v, w, p)]
        self._check(self._L.ns3d_upload_primitive(self.ctx, *[_ptr(a) for a in arrs]), "upload_primitive")

    def upload_conserved(self, rho, ru, rv, rw, E):
        arrs = [self._full(x) for x in (rho, ru, rv, rw, E)]
        self._check(self._L.ns3d_upload_conserved(self.ctx, *[_ptr(a) for a in arrs]), "upload_conserved")

    def primitive(self):
        out = [np.empty(self.shape) for _ in range(6)]
        self._check(self._L.ns3d_download_primitive(self.ctx, *[_ptr(a) for a in out]), "download_primitive")
        return dict(zip(("rho", "u", "v", "w", "p", "T"), out))

    def conserved(self):
        out = [np.empty(self.shape) for _ in range(5)]
        self._check(self._L.ns3d_download_conserved(self.ctx, *[_ptr(a) for a in out]), "download_conserved")
        return out

    # -- hot path ----------------------------------------------------------
    def compute_rhs(self):
        self._check(self._L.ns3d_compute_rhs(self.ctx), "compute_rhs")
        out = [np.empty(self.shape) for _ in range(5)]
        self._check(self._L.ns3d_download_rhs(self.ctx, *[_ptr(a) for a in out]), "download_rhs")
        return out

    def du_inviscous(self):
        self._check(self._L.ns3d_du_inviscous(self.ctx), "du_inviscous")

    def du_inviscous_character(self):
        self._check(self._L.ns3d_du_inviscous_character(self.ctx), "du_inviscous_character")

    def face_flux(self, d):
        sz, sy, sx = self.shape
        shp = [(sz, sy, sx - 1), (sz, sy - 1, sx), (sz - 1, sy, sx)][d]
        out = [np.empty(shp) for _ in range(5)]
        self._check(self._L.ns3d_download_face_flux(self.ctx, d, *[_ptr(a) for a in out]), "face_flux")
        return out

    def compute_gradients(self):
        self._check(self._L.ns3d_compute_gradients(self.ctx), "compute_gradients")
        out = [np.empty(self.shape) for _ in range(12)]
        arr = (_lib.DP * 12)(*[_ptr(a) for a in out])
        self._check(self._L.ns3d_download_gradients(self.ctx, arr), "download_gradients")
        return out

    def compute_viscous_flux(self):
        self._check(self._L.ns3d_compute_viscous_flux(self.ctx), "compute_viscous_flux")
        out = [np.empty(self.shape) for _ in range(12)]
        arr = (_lib.DP * 12)(*[_ptr(a) for a in out])
        self._check(self._L.ns3d_download_viscous_flux(self.ctx, arr), "download_viscous_flux")
        return out

    def compute_vis_flux(self):
        """Viscous divergence of the stored fluxes on rhs = 0 (5 arrays)."""
        self._check(self._L.ns3d_compute_vis_flux(self.ctx), "compute_vis_flux")
        out = [np.empty(self.shape) for _ in range(5)]
        self._check(self._L.ns3d_download_rhs(self.ctx, *[_ptr(a) for a in out]), "download_rhs")
        return out

    def set_precision(self, precision):
        """Switch between "fp64" and "mixed" arithmetic, keeping the state."""
        p = _enum(_lib.PREC, precision)
        self._check(self._L.ns3d_set_precision(self.ctx, p), "set_precision")
        self.params.precision = p

    def rk3_step(self, dt):
        st = self._L.ns3d_rk3_step(self.ctx, float(dt))
        self._check(st, "rk3_step")

    def advance(self, nsteps, dt):
        self._check(self._L.ns3d_advance(self.ctx, int(nsteps), float(dt)), "advance")

    def advance_cfl(self, nsteps, time=0.0):
        """nsteps RK3 steps, each with the CFL time step of its initial state
        computed on the device; returns the time after them."""
        t = C.c_double(time)
        self._check(self._L.ns3d_advance_cfl(self.ctx, int(nsteps), C.byref(t)), "advance_cfl")
        return t.value

    def nonphysical_cell(self):
        i, j, k = C.c_int(), C.c_int(), C.c_int()
        self._L.ns3d_nonphysical_cell(self.ctx, C.byref(i), C.byref(j), C.byref(k))
        return i.value, j.value, k.value

    def timestep(self):
        dt = C.c_double()
        self._check(self._L.ns3d_compute_timestep_local(self.ctx, C.byref(dt)), "compute_timestep")
        return dt.value

    def total_energy(self):
        e = C.c_double()
        self._check(self._L.ns3d_total_energy(self.ctx, C.byref(e)), "total_energy")
        return e.value

    def sync(self):
        self._check(self._L.ns3d_sync(self.ctx), "sync")

    def stream(self):
        return self._L.ns3d_stream(self.ctx)

    def copy_probe(self, reps=10):
        """Achieved GB/s of a plain streaming copy over 12 cell arrays."""
        g = C.c_double()
        self._check(self._L.ns3d_copy_probe(self.ctx, int(reps), C.byref(g)), "copy_probe")
        return g.value

    def profile(self, on=True):
        """Start (or stop) per-kernel event timing; resets the totals."""
        self._check(self._L.ns3d_profile_enable(self.ctx, 1 if on else 0), "profile_enable")

    def profile_report(self):
        """{kernel name: (total ms, launches)} since profile() was called."""
        n = self._L.ns3d_profile_count(self.ctx)
        if n < 0:
            self._check(-n, "profile_count")
        out = {}
        for k in range(n):
            buf = C.create_string_buffer(64)
            ms = C.c_double()
            cnt = C.c_longlong()
            self._check(self._L.ns3d_profile_entry(self.ctx, k, buf, 64, C.byref(ms), C.byref(cnt)), "profile_entry")
            out[buf.value.decode()] = (ms.value, cnt.value)
        return out

    # -- differentiation -----------------------------------------------------
    def _set5(self, arrs):
        full = [self._full(a) for a in arrs]
        return full, (_lib.DP * 5)(*[_ptr(a) for a in full])

    def _out5(self):
        out = [np.zeros(self.shape) for _ in range(5)]
        return out, (_lib.DP * 5)(*[_ptr(a) for a in out])

    def tlm_rhs(self, dU):
        """dR/dU * dU at the current state (5 conserved arrays)."""
        keep, src = self._set5(dU)
        out, dst = self._out5()
        self._check(self._L.ns3d_tlm_rhs(self.ctx, src, dst), "tlm_rhs")
        return out

    def adj_rhs(self, lam, w_obj=0.0):
        """((dR/dU)^T lam + w_obj d(sum Phi)/dU, d/d(1/Re) of lam.R + w_obj sum Phi)."""
        keep, src = self._set5(lam)
        out, dst = self._out5()
        th = C.c_double()
        self._check(self._L.ns3d_adj_rhs(self.ctx, src, float(w_obj), dst, C.byref(th)), "adj_rhs")
        return out, th.value

    def dissipation(self):
        """Volume-mean viscous dissipation of the current state."""
        e = C.c_double()
        self._check(self._L.ns3d_dissipation(self.ctx, C.byref(e)), "dissipation")
        return e.value

    def adj_forward(self, nsteps, dt, ckpt_every=None):
        """J = sum_{n<nsteps} dt * dissipation(U^n); stores checkpoints."""
        J = C.c_double()
        ck = int(ckpt_every) if ckpt_every else int(nsteps)
        self._check(self._L.ns3d_adj_forward(self.ctx, int(nsteps), float(dt), ck, C.byref(J)), "adj_forward")
        return J.value

    def adj_backward(self, lamN=None):
        """(dJ/dU^0 as 5 conserved arrays, dJ/d(1/Re)) of the last adj_forward."""
        keep, src = self._set5(lamN) if lamN is not None else (None, None)
        out, dst = self._out5()
        g = C.c_double()
        srcp = C.cast(src, C.c_void_p) if src is not None else None
        self._check(self._L.ns3d_adj_backward(self.ctx, srcp, dst, C.byref(g)), "adj_backward")
        return out, g.value

    def adj_memory(self):
        b = C.c_longlong()
        self._check(self._L.ns3d_adj_memory(self.ctx, C.byref(b)), "adj_memory")
        return b.value

    # -- monitors ------------------------------------------------------------
    STAT_NAMES = ("kinetic_energy", "tke", "u_rms", "dissipation", "taylor", "eta", "u_eta", "tau_eta",
                  "taylor_Li", "Re_lambda", "Mach_t", "mean_mu", "mean_rho")

    def energy_spectrum(self, nbins=None):
        """(k, E(k)) of the current velocity, computed on the device."""
        nb = int(nbins) if nbins else 2 * self.grid.nx + 1
        ek = np.zeros(nb)
        self._check(self._L.ns3d_energy_spectrum(self.ctx, _ptr(ek), nb), "energy_spectrum")
        return np.arange(nb), ek

    def turbulence_statistics(self):
        """The reference's isotropic-turbulence statistics, on the device."""
        out = np.zeros(len(self.STAT_NAMES))
        self._check(self._L.ns3d_turbulence_statistics(self.ctx, _ptr(out), len(out)), "turbulence_statistics")
        return dict(zip(self.STAT_NAMES, out.tolist()))

    def residuals(self):
        """Normalised L2 change of each conserved variable over the last step."""
        out = np.zeros(5)
        self._check(self._L.ns3d_residuals(self.ctx, _ptr(out)), "residuals")
        return out
