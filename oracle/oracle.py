"""ctypes wrapper of the CPU FP64 oracle (liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker, never by the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

# array ids (ns3d_oracle.h)
RHO, RHOU, RHOV, RHOW, E = 0, 1, 2, 3, 4
RHO0 = 5
U, V, W, P, T, CS = 10, 11, 12, 13, 14, 15
RHS0 = 16
GRAD0 = 21
VF0 = 33
FX0, FY0, FZ0 = 45, 50, 55

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    from ns3d_cpp_amd._lib import Grid, Params
    lib = C.CDLL(LIB_PATH)
    PP, GP, DP = C.POINTER(Params), C.POINTER(Grid), C.POINTER(C.c_double)
    sig = {
        "orc_params_finalize": (None, [PP, GP]),
        "orc_create": (C.c_void_p, [PP, GP]),
        "orc_destroy": (None, [C.c_void_p]),
        "orc_array": (DP, [C.c_void_p, C.c_int]),
        "orc_array_len": (C.c_longlong, [C.c_void_p, C.c_int]),
        "orc_sizes": (None, [C.c_void_p] + [C.POINTER(C.c_int)] * 4),
        "orc_c2p": (None, [C.c_void_p]),
        "orc_p2c": (None, [C.c_void_p]),
        "orc_record_u0": (None, [C.c_void_p]),
        "orc_apply_boundary": (None, [C.c_void_p]),
        "orc_exchange_vflux": (None, [C.c_void_p]),
        "orc_split_state": (None, [C.c_int, C.c_double, C.c_double] + [C.c_double] * 7
                            + [C.c_int] * 3 + [DP, DP]),
        "orc_reconstruct": (C.c_double, [DP, C.c_int, C.c_double, PP, C.POINTER(C.c_int)]),
        "orc_diff": (C.c_double, [DP, C.c_int, C.c_double]),
        "orc_char_matrices": (None, [DP, DP, C.c_int, C.c_int, C.c_int, C.c_double, DP, DP]),
        "orc_choose_scheme": (C.c_int, [C.c_int] * 5),
        "orc_get_mu": (C.c_double, [PP, C.c_double]),
        "orc_du_inviscous": (None, [C.c_void_p]),
        "orc_du_inviscous_character": (None, [C.c_void_p]),
        "orc_compute_gradients": (None, [C.c_void_p]),
        "orc_compute_viscous_flux": (None, [C.c_void_p]),
        "orc_compute_vis_flux": (None, [C.c_void_p]),
        "orc_compute_rhs": (None, [C.c_void_p]),
        "orc_rk3_step": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(C.c_longlong)]),
        "orc_step": (C.c_int, [C.c_void_p, C.c_double, DP, C.POINTER(C.c_longlong)]),
        "orc_compute_timestep": (C.c_double, [C.c_void_p]),
        "orc_total_energy": (C.c_double, [C.c_void_p]),
        "orc_num_threads": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class Oracle:
    """One oracle field (the reference's Field3D on one periodic rank)."""

    def __init__(self, params, grid):
        self.L = load()
        self.params = params
        self.grid = grid
        self.h = self.L.orc_create(C.byref(params), C.byref(grid))
        if not self.h:
            raise MemoryError("orc_create failed")
        sx, sy, sz, ng = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self.L.orc_sizes(self.h, C.byref(sx), C.byref(sy), C.byref(sz), C.byref(ng))
        self.shape = (sz.value, sy.value, sx.value)
        self.ng = ng.value

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def arr(self, aid):
        """numpy view of an oracle array: cell arrays (sz,sy,sx), face arrays
        shaped like idx_fx / idx_fy / idx_fz."""
        n = self.L.orc_array_len(self.h, aid)
        ptr = self.L.orc_array(self.h, aid)
        flat = np.ctypeslib.as_array(ptr, shape=(n,))
        sz, sy, sx = self.shape
        if aid < FX0:
            return flat.reshape(sz, sy, sx)
        if aid < FY0:
            return flat.reshape(sz, sy, sx - 1)
        if aid < FZ0:
            return flat.reshape(sz, sy - 1, sx)
        return flat.reshape(sz - 1, sy, sx)

    def interior(self, a):
        ng = self.ng
        return a[ng:-ng, ng:-ng, ng:-ng]

    def set_primitive(self, rho, u, v, w, p):
        """Interior primitives in, then main.cpp:70-71 (apply_boundary, P2C)."""
        for aid, val in zip((RHO, U, V, W, P), (rho, u, v, w, p)):
            a = self.arr(aid)
            if np.shape(val) == self.shape:
                a[...] = val
            else:
                self.interior(a)[...] = val
        self.L.orc_apply_boundary(self.h)
        self.L.orc_p2c(self.h)

    def set_conserved(self, rho, ru, rv, rw, En):
        for aid, val in zip((RHO, RHOU, RHOV, RHOW, E), (rho, ru, rv, rw, En)):
            a = self.arr(aid)
            if np.shape(val) == self.shape:
                a[...] = val
            else:
                self.interior(a)[...] = val
        self.L.orc_c2p(self.h)
        self.L.orc_apply_boundary(self.h)
        self.L.orc_p2c(self.h)

    def primitive(self):
        return {k: self.arr(a).copy() for k, a in
                (("rho", RHO), ("u", U), ("v", V), ("w", W), ("p", P), ("T", T))}

    def conserved(self):
        return [self.arr(a).copy() for a in (RHO, RHOU, RHOV, RHOW, E)]

    def compute_rhs(self):
        self.L.orc_compute_rhs(self.h)
        return [self.arr(RHS0 + m).copy() for m in range(5)]

    def du_inviscous(self):
        if self.params.char_recon:
            self.L.orc_du_inviscous_character(self.h)
        else:
            self.L.orc_du_inviscous(self.h)

    def face_flux(self, d):
        base = (FX0, FY0, FZ0)[d]
        return [self.arr(base + m).copy() for m in range(5)]

    def compute_gradients(self):
        self.L.orc_compute_gradients(self.h)
        return [self.arr(GRAD0 + m).copy() for m in range(12)]

    def compute_viscous_flux(self):
        self.L.orc_compute_viscous_flux(self.h)
        return [self.arr(VF0 + m).copy() for m in range(12)]

    def step(self, dt):
        """One time_advance step (record U0, RK3); returns (status, dt_used)."""
        dtu = C.c_double()
        bad = C.c_longlong()
        st = self.L.orc_step(self.h, float(dt), C.byref(dtu), C.byref(bad))
        self.bad_id = bad.value
        return st, dtu.value

    def timestep(self):
        return self.L.orc_compute_timestep(self.h)

    def total_energy(self):
        return self.L.orc_total_energy(self.h)
