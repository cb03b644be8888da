"""Exact pins of the oracle's WENO5 constants and of the characteristic
path's tangential basis.

WENO5 (reconstruction.cpp:93-124).  Everything the scheme's constants encode
is derived here from its published definition (Jiang & Shu, J. Comput.
Phys. 126, 1996) in exact rational arithmetic, without the constants:
  - p_k: the quadratic whose cell averages over the sub-stencil k are the
    data, evaluated at the right face of the centre cell;
  - beta_k = sum_{l=1,2} int_cell (d^l p_k / dx^l)^2 dx (unit spacing);
  - d_k: the weights that turn the three p_k into the 5-cell (quartic)
    reconstruction;
  - alpha_k = d_k / (eps + beta_k)^2 with eps = 1e-6, omega = alpha / sum.
The oracle must agree with the exact value on stencils of exactly
representable data, to a rounding bound.  The stencils are chosen so that
eps, 13/12, 1/4, each d_k and each p_k change the result by far more than
that bound; test_weno5_pins_catch_wrong_constants shows it for a list of
plausible slips.

Characteristic path (InviscousFlux_new.cpp:165-437, 532-659).  L R = I and
the diagonalisation of the flux Jacobian leave the basis of the repeated
eigenvalue u_n open.  The reference takes the entropy wave plus the two
coordinate axes other than the normal as shear waves (the permuted
(nz, nx, ny) / (ny, nz, nx) vectors).  With a nonlinear reconstruction that
choice matters: a basis rotated in the tangent plane mixes a tangential
velocity that is zero everywhere with the other one.  Pinned here:
  - the structure: for an axis normal, columns 2 and 3 of R (rows 2 and 3 of
    L) are the two other unit axes, with no mass or entropy part;
  - the consequence: a tangential velocity that is zero everywhere has a zero
    momentum face flux, exactly, with WENO5 and discontinuous data;
  - the composition: projecting each face's stencil with L, reconstructing
    each characteristic with the oracle's WENO5 and projecting back with R,
    done here in Python from the oracle's split and matrices, reproduces the
    oracle's characteristic faces bit for bit;
  - sensitivity: the same composition with a rotated tangential pair breaks
    the exact zero, and with the pair swapped changes nothing beyond
    rounding (the order of the two is immaterial).
"""
import ctypes as C
import math
import os
import sys
from fractions import Fraction as Fr

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import ns3d_cpp_amd as ns  # noqa: E402
from oracle import CS, Oracle, load  # noqa: E402

L = load()
DP = C.POINTER(C.c_double)


# ------------------------------------------------------- exact WENO5 algebra

def _solve(A, b):
    """Gauss-Jordan elimination in exact arithmetic; A is m x n with m >= n
    and a consistent system (extra rows are checked)."""
    m, n = len(A), len(A[0])
    M = [list(map(Fr, row)) + [Fr(v)] for row, v in zip(A, b)]
    r = 0
    for c in range(n):
        p = next((i for i in range(r, m) if M[i][c] != 0), None)
        assert p is not None, "singular system"
        M[r], M[p] = M[p], M[r]
        M[r] = [x / M[r][c] for x in M[r]]
        for i in range(m):
            if i != r and M[i][c] != 0:
                f = M[i][c]
                M[i] = [x - f * y for x, y in zip(M[i], M[r])]
        r += 1
    for i in range(r, m):
        assert M[i][n] == 0, "inconsistent system"
    return [M[i][n] for i in range(n)]


def _cell_average_row(j, deg):
    """Averages of x^q over the unit cell [j - 1/2, j + 1/2], q = 0..deg."""
    a, b = Fr(2 * j - 1, 2), Fr(2 * j + 1, 2)
    return [(b ** (q + 1) - a ** (q + 1)) / (q + 1) for q in range(deg + 1)]


def _interp_coeffs(cells):
    """Polynomial of degree len(cells)-1 with given cell averages: returns
    C such that coefficient q of the polynomial is sum_j C[q][j] f_j."""
    deg = len(cells) - 1
    A = [_cell_average_row(j, deg) for j in cells]
    cols = []
    for jj in range(len(cells)):
        e = [Fr(int(jj == t)) for t in range(len(cells))]
        cols.append(_solve(A, e))
    return [[cols[j][q] for j in range(len(cells))] for q in range(deg + 1)]


def _beta_quadratic(coeffs):
    """sum_{l=1,2} int_{-1/2}^{1/2} (d^l p/dx^l)^2 dx for p = sum c_q x^q,
    integrated term by term."""
    def deriv(c):
        return [q * c[q] for q in range(1, len(c))]

    def integral_of_square(c):
        tot = Fr(0)
        for a, ca in enumerate(c):
            for b, cb in enumerate(c):
                e = a + b + 1
                tot += ca * cb * (Fr(1, 2) ** e - Fr(-1, 2) ** e) / e
        return tot
    d1 = deriv(coeffs)
    d2 = deriv(d1)
    return integral_of_square(d1) + integral_of_square(d2)


# centre cell 0, face x = +1/2; the stencil f0..f4 holds cells -2..2
_SUB = [(-2, -1, 0), (-1, 0, 1), (0, 1, 2)]
_SUB_C = [_interp_coeffs(list(s)) for s in _SUB]
_FULL_C = _interp_coeffs([-2, -1, 0, 1, 2])


def _sub_face_functional(k):
    """p_k(1/2) as coefficients on f0..f4."""
    out = [Fr(0)] * 5
    for jj, j in enumerate(_SUB[k]):
        out[j + 2] = sum(_SUB_C[k][q][jj] * Fr(1, 2) ** q for q in range(3))
    return out


P_FACE = [_sub_face_functional(k) for k in range(3)]
FULL_FACE = [sum(_FULL_C[q][j] * Fr(1, 2) ** q for q in range(5)) for j in range(5)]
# the linear weights: sum_k d_k p_k(1/2) = the quartic's face value, for all data
D_LIN = _solve([[P_FACE[k][j] for k in range(3)] for j in range(5)], FULL_FACE)
EPS = Fr(1, 10 ** 6)


def _sub_coeffs(f, k):
    vals = [f[j + 2] for j in _SUB[k]]
    return [sum(_SUB_C[k][q][jj] * vals[jj] for jj in range(3)) for q in range(3)]


def weno5_exact(f, eps=EPS, d=None, AB=None, pface=None):
    """Exact WENO5-JS face value of f0..f4 (Fractions), its smoothness
    indicators and the three weighted terms.  beta_k is the integral of the
    definition; AB = (A, B) replaces it by A (2c)^2 + B (2b)^2 for
    p_k = a + b x + c x^2, and d / pface replace the derived weights and
    faces (for the mutation test below)."""
    f = [Fr(x) for x in f[:5]]
    d = D_LIN if d is None else d
    pface = P_FACE if pface is None else pface
    betas = []
    for k in range(3):
        c = _sub_coeffs(f, k)
        if AB is None:
            betas.append(_beta_quadratic(c))
        else:
            betas.append(Fr(AB[0]) * (2 * c[2]) ** 2 + Fr(AB[1]) * (2 * c[1]) ** 2)
    alpha = [d[k] / (eps + betas[k]) ** 2 for k in range(3)]
    s = sum(alpha)
    p = [sum(pface[k][j] * f[j] for j in range(5)) for k in range(3)]
    terms = [alpha[k] / s * p[k] for k in range(3)]
    return sum(terms), betas, terms


def oracle_weno5(st, flag):
    p, _ = ns.make_params(8, recon="WENO5")
    a = np.ascontiguousarray(st, dtype=np.float64)
    err = C.c_int(0)
    v = L.orc_reconstruct(a.ctypes.data_as(DP), len(a), float(flag), C.byref(p), C.byref(err))
    assert err.value == 0
    return v


def _stencils():
    """Stencils of exactly representable data: a unit step both ways, an
    odd-even mode, random non-smooth dyadic data with O(1) smoothness
    indicators, and small-amplitude data whose indicators are ~eps."""
    out = [np.array([0.0, 0.0, 0.0, 1.0, 1.0, 1.0]), np.array([1.0, 1.0, 1.0, 0.0, 0.0, 0.0]),
           np.array([0.0, 1.0, 0.0, 1.0, 0.0, 1.0]), np.array([3.0, -1.0, 2.0, 0.5, -4.0, 0.0])]
    rng = np.random.default_rng(1234)
    for _ in range(12):
        out.append(rng.integers(-16, 17, size=6) / 8.0)
    for _ in range(12):
        out.append(rng.integers(-16, 17, size=6) * 2.0 ** -11)
    return out


def _tolerance(terms):
    # a rounding bound: a few dozen roundings of the largest term
    return 64 * 2.0 ** -52 * float(sum(abs(t) for t in terms)) + 1e-300


def test_weno5_derived_constants():
    """The derivation alone gives the published constants: linear weights
    (1/10, 6/10, 3/10), the sub-stencil faces (2,-7,11)/6, (-1,5,2)/6,
    (2,5,-1)/6 and beta = 13/12 (second difference)^2 + 1/4 (first)^2."""
    assert D_LIN == [Fr(1, 10), Fr(6, 10), Fr(3, 10)]
    assert P_FACE[0] == [Fr(2, 6), Fr(-7, 6), Fr(11, 6), 0, 0]
    assert P_FACE[1] == [0, Fr(-1, 6), Fr(5, 6), Fr(2, 6), 0]
    assert P_FACE[2] == [0, 0, Fr(2, 6), Fr(5, 6), Fr(-1, 6)]
    _, betas, _ = weno5_exact([0, 0, 0, 1, 1])
    assert betas == [0, Fr(4, 3), Fr(10, 3)]
    rng = np.random.default_rng(5)
    for _ in range(20):
        f = [Fr(int(x)) for x in rng.integers(-9, 10, size=5)]
        _, betas, _ = weno5_exact(f)
        # the integral equals 13/12 (second difference)^2 + 1/4 (slope)^2
        assert betas == weno5_exact(f, AB=(Fr(13, 12), Fr(1, 4)))[1]
        f0, f1, f2, f3, f4 = f
        assert betas[0] == Fr(13, 12) * (f0 - 2 * f1 + f2) ** 2 + Fr(1, 4) * (f0 - 4 * f1 + 3 * f2) ** 2
        assert betas[1] == Fr(13, 12) * (f1 - 2 * f2 + f3) ** 2 + Fr(1, 4) * (f1 - f3) ** 2
        assert betas[2] == Fr(13, 12) * (f2 - 2 * f3 + f4) ** 2 + Fr(1, 4) * (3 * f2 - 4 * f3 + f4) ** 2


@pytest.mark.parametrize("flag", [1.0, -1.0])
def test_weno5_oracle_equals_exact(flag):
    """orc_reconstruct (WENO5) against the exact derivation; the minus side
    reconstructs the reversed stencil (reconstruction.cpp:29)."""
    for st in _stencils():
        f = st if flag > 0 else st[::-1]
        exact, _, terms = weno5_exact([float(x) for x in f])
        got = oracle_weno5(st, flag)
        assert abs(got - float(exact)) <= _tolerance(terms), (st, flag, got, float(exact))


MUTANTS = {
    "eps 1e-40": dict(eps=Fr(10) ** -40),
    "eps 1e-7": dict(eps=Fr(1, 10 ** 7)),
    "eps 1e-5": dict(eps=Fr(1, 10 ** 5)),
    "13/12 -> 1": dict(AB=(1, Fr(1, 4))),
    "13/12 -> 12/13": dict(AB=(Fr(12, 13), Fr(1, 4))),
    "1/4 -> 1/2": dict(AB=(Fr(13, 12), Fr(1, 2))),
    "1/4 -> 1": dict(AB=(Fr(13, 12), 1)),
    "d reversed": dict(d=[Fr(3, 10), Fr(6, 10), Fr(1, 10)]),
    "d (1/3,1/3,1/3)": dict(d=[Fr(1, 3)] * 3),
    "p1 = (-1,5,2)/6 -> (1,5,0)/6": dict(pface=[P_FACE[0], [0, Fr(1, 6), Fr(5, 6), 0, 0], P_FACE[2]]),
}


@pytest.mark.parametrize("name", list(MUTANTS))
def test_weno5_pins_catch_wrong_constants(name):
    """Each plausible slip in a WENO5 constant moves the face value of at
    least one stencil above by more than the rounding bound the oracle is
    held to, so it cannot pass test_weno5_oracle_equals_exact."""
    kw = MUTANTS[name]
    caught = False
    for st in _stencils():
        for flag in (1.0, -1.0):
            f = [float(x) for x in (st if flag > 0 else st[::-1])]
            exact, _, terms = weno5_exact(f)
            mutant, _, _ = weno5_exact(f, **kw)
            if abs(float(mutant) - float(exact)) > 4 * _tolerance(terms):
                caught = True
    assert caught, name


# ------------------------------------------------- characteristic basis

def _char_matrices(ul, ur, axis, gamma=1.4):
    Lm, Rm = np.zeros(25), np.zeros(25)
    n = [int(d == axis) for d in range(3)]
    L.orc_char_matrices(np.ascontiguousarray(ul, dtype=np.float64).ctypes.data_as(DP),
                        np.ascontiguousarray(ur, dtype=np.float64).ctypes.data_as(DP), *n, gamma,
                        Lm.ctypes.data_as(DP), Rm.ctypes.data_as(DP))
    return Lm.reshape(5, 5), Rm.reshape(5, 5)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_char_shear_vectors_are_the_other_axes(axis):
    """Columns 2, 3 of R and rows 2, 3 of L: unit momentum vectors along the
    two axes other than the normal, no mass part; the entropy and acoustic
    waves carry no tangential mixing."""
    rng = np.random.default_rng(axis)
    others = {d for d in range(3) if d != axis}
    for _ in range(4):
        ul = np.array([rng.uniform(0.5, 2), *rng.normal(size=3), rng.uniform(0.5, 3)])
        ur = np.array([rng.uniform(0.5, 2), *rng.normal(size=3), rng.uniform(0.5, 3)])
        Lm, Rm = _char_matrices(ul, ur, axis)
        seen = set()
        for col in (2, 3):
            assert Rm[0, col] == 0.0 and Lm[col, 4] == 0.0
            mom = Rm[1:4, col]
            assert sorted(np.abs(mom)) == [0.0, 0.0, 1.0]
            d = int(np.argmax(np.abs(mom)))
            assert d in others
            assert list(Lm[col, 1:4]) == list(mom)
            seen.add(d)
        assert seen == others
        assert np.allclose(Lm @ Rm, np.eye(5), atol=1e-13)


def _line_state(axis, zero_t, n=12, seed=0):
    """A state varying only along `axis`, with jumps: rho, the normal
    velocity, p and the other tangential velocity vary; the tangential
    velocity `zero_t` is zero everywhere."""
    rng = np.random.default_rng(seed)
    x = np.arange(n)
    step = (x >= n // 2).astype(float)
    prof = {"rho": 1.0 + 0.3 * step + 0.05 * np.sin(2 * np.pi * x / n),
            "p": 2.0 + 0.8 * step + 0.1 * np.cos(2 * np.pi * x / n),
            "vn": 0.4 * np.sin(2 * np.pi * x / n) - 0.2 * step,
            "vo": 0.7 * step - 0.35 + 0.1 * rng.standard_normal(n)}
    other = ({0, 1, 2} - {axis, zero_t}).pop()
    shape = [1, 1, 1]
    shape[2 - axis] = n          # arrays are (z, y, x)
    full = (n, n, n)
    vel = [None, None, None]
    vel[axis] = prof["vn"]
    vel[other] = prof["vo"]
    vel[zero_t] = np.zeros(n)

    def field(a):
        return np.broadcast_to(a.reshape(shape), full).copy()
    return [field(prof["rho"]), field(vel[0]), field(vel[1]), field(vel[2]), field(prof["p"])]


def _char_faces(o, axis, Lrot=None, swap=False):
    """Characteristic faces of one pencil along `axis`, composed in Python
    from the oracle's split (orc_split_state), matrices (orc_char_matrices)
    and WENO5 (orc_reconstruct), in the order of pencil_faces_char."""
    p = o.params
    prim = o.primitive()
    E = o.conserved()[4]
    cs = o.arr(CS).copy()
    ng = o.ng
    S, mid = 6, 2
    npts = o.shape[2 - axis]

    def at(a, q):
        idx = [ng, ng, ng]
        idx[2 - axis] = q
        return a[tuple(idx)]
    Fp, Fm = [], []
    nvec = [int(d == axis) for d in range(3)]
    for q in range(npts):
        fp, fm = (C.c_double * 5)(), (C.c_double * 5)()
        L.orc_split_state(p.fvs_type, p.gamma, p.sw_eps, at(prim["rho"], q), at(prim["u"], q), at(prim["v"], q),
                          at(prim["w"], q), at(E, q), at(prim["p"], q), at(cs, q), *nvec, fp, fm)
        Fp.append(list(fp))
        Fm.append(list(fm))
    faces = []
    n_int = npts - 2 * ng
    for q in range(ng - 1, ng + n_int):
        ul = [at(prim[k], q) for k in ("rho", "u", "v", "w", "p")]
        ur = [at(prim[k], q + 1) for k in ("rho", "u", "v", "w", "p")]
        Lm, Rm = _char_matrices(ul, ur, axis)
        if Lrot is not None:
            Lm, Rm = Lrot @ Lm, Rm @ Lrot.T
        if swap:
            P = np.eye(5)[[0, 1, 3, 2, 4]]
            Lm, Rm = P @ Lm, Rm @ P.T
        wchar = []
        for nn in range(5):
            lp, lm = [], []
            for m in range(S):
                ii = q + (m - mid)
                sp = sm = 0.0
                for r in range(5):
                    sp += Lm[nn, r] * Fp[ii][r]
                    sm += Lm[nn, r] * Fm[ii][r]
                lp.append(sp)
                lm.append(sm)
            wchar.append(oracle_weno5(np.array(lp), 1.0) + oracle_weno5(np.array(lm), -1.0))
        face = []
        for nn in range(5):
            s = 0.0
            for r in range(5):
                s += Rm[nn, r] * wchar[r]
            face.append(s)
        faces.append(face)
    return np.array(faces)


def _oracle_char_pencil(o, axis):
    o.du_inviscous()
    F = o.face_flux(axis)
    ng = o.ng
    out = []
    for m in range(5):
        a = F[m]
        idx = [ng, ng, ng]
        idx[2 - axis] = slice(ng - 1, ng + o.shape[2 - axis] - 2 * ng)
        out.append(a[tuple(idx)])
    return np.array(out).T


CASES = [(a, t) for a in range(3) for t in range(3) if t != a]


@pytest.mark.parametrize("axis,zero_t", CASES)
@pytest.mark.parametrize("fvs", ["StagerWarming", "LaxFriedrichs"])
def test_char_zero_tangential_velocity_has_zero_face_flux(axis, zero_t, fvs):
    """With the coordinate-axis shear basis the momentum face flux of a
    tangential velocity that is zero everywhere is exactly zero, while the
    other tangential velocity jumps (WENO5, characteristic path)."""
    n = 12
    p, g = ns.make_params(n, Ma=0.5, recon="WENO5", fvs_type=fvs, vis_scheme="C8th", char_recon=1)
    o = Oracle(p, g)
    o.set_primitive(*_line_state(axis, zero_t, n))
    o.du_inviscous()
    F = o.face_flux(axis)
    ng = o.ng
    core = F[1 + zero_t][ng:-ng, ng:-ng, ng:-ng]
    assert np.all(core == 0.0), np.abs(core).max()
    other = ({0, 1, 2} - {axis, zero_t}).pop()
    assert np.abs(F[1 + other][ng:-ng, ng:-ng, ng:-ng]).max() > 1e-3


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_char_composition_and_basis_sensitivity(axis):
    """The Python composition reproduces the oracle's characteristic faces
    bit for bit; swapping the shear pair changes them only by rounding; a
    pair rotated by 0.3 rad in the tangent plane breaks the exact zero of the
    test above (so that test detects such a basis)."""
    n = 12
    zero_t = (axis + 1) % 3
    p, g = ns.make_params(n, Ma=0.5, recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th", char_recon=1)
    o = Oracle(p, g)
    o.set_primitive(*_line_state(axis, zero_t, n, seed=axis))
    ref = _oracle_char_pencil(o, axis)
    mine = _char_faces(o, axis)
    assert np.array_equal(mine, ref)
    swapped = _char_faces(o, axis, swap=True)
    assert np.allclose(swapped, ref, rtol=0, atol=1e-14 * np.abs(ref).max())
    th = 0.3
    Q = np.eye(5)
    Q[2, 2] = Q[3, 3] = math.cos(th)
    Q[2, 3], Q[3, 2] = -math.sin(th), math.sin(th)
    rot = _char_faces(o, axis, Lrot=Q)
    assert np.abs(rot[:, 1 + zero_t]).max() > 1e-8
