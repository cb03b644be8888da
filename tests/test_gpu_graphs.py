"""RK3 steps replayed from captured HIP graphs (NS3D_GRAPHS=1: ns3d_advance
and ns3d_advance_cfl after their first step) against eager steps: the same
states bit for bit, replay after replay, with eager work and host
synchronisation between the replays, across a change of dt (a new capture)
and for the device-side CFL time step (with the mixed third stage reducing
the next step's CFL minimum where its workgroups tile the grid).  The
non-physical-cell report keeps the earliest step's first cell when later
steps run from the graph."""
import numpy as np
import pytest

import ns3d_cpp_amd as ns
from ns3d_cpp_amd import fields

pytestmark = pytest.mark.gpu

CASES = {
    "fp64-periodic": dict(precision="fp64", recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th"),
    "mixed-periodic": dict(precision="mixed", recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th"),
    "fp64-walls": dict(precision="fp64", recon="MDCD", fvs_type="LaxFriedrichs", vis_scheme="C6th",
                       bc=["Wall", "Outflow", "Periodic", "Periodic", "Symmetry", "Wall"]),
    "mixed-walls": dict(precision="mixed", recon="UPWIND_7TH", fvs_type="StagerWarming", vis_scheme="C4th",
                        bc=["Periodic", "Periodic", "Wall", "Symmetry", "Periodic", "Periodic"]),
    "mixed-char": dict(precision="mixed", recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th",
                       char_recon=1),
    "fp64-char": dict(precision="fp64", recon="WENO5", fvs_type="LaxFriedrichs", vis_scheme="C8th",
                      char_recon=1),
}


@pytest.fixture(autouse=True)
def graphs_on(monkeypatch):
    """Graph replay is opt-in; the contexts of these tests take it."""
    monkeypatch.setenv("NS3D_GRAPHS", "1")


def pair(case, n=16, seed=3):
    st, Ma = fields.hit(n, k0=2, Mt=0.3, seed=seed)
    p, g = ns.make_params(n, Re=200.0, Ma=Ma, **CASES[case])
    a, b = ns.Solver(p, g), ns.Solver(p, g)
    a.upload_primitive(*st)
    b.upload_primitive(*st)
    return a, b


def same(a, b):
    for x, y in zip(a.conserved(), b.conserved()):
        if not np.array_equal(x, y):
            return False
    return True


@pytest.mark.parametrize("case", list(CASES))
def test_graph_replays_match_eager_steps(gpu, case):
    a, b = pair(case)
    dt = 2e-3
    a.advance(4, dt)                 # 1 eager step, capture, 3 replays
    for _ in range(4):
        b.rk3_step(dt)
    assert same(a, b)
    # eager work and synchronisation between replays
    a.rk3_step(dt)
    b.rk3_step(dt)
    assert a.total_energy() == b.total_energy()
    assert same(a, b)
    for _ in range(2):
        a.advance(3, dt)             # the same graph, replayed again
        for _ in range(3):
            b.rk3_step(dt)
        assert same(a, b)
    a.advance(3, 1.5e-3)             # another dt: captured anew
    for _ in range(3):
        b.rk3_step(1.5e-3)
    assert same(a, b)
    a.advance(2, dt)                 # and back
    for _ in range(2):
        b.rk3_step(dt)
    assert same(a, b)


@pytest.mark.parametrize("case", ["fp64-periodic", "mixed-periodic", "fp64-walls", "mixed-char"])
def test_advance_cfl_matches_host_time_loop(gpu, case):
    """maincycle's loop (dt = compute_timestep; runge_kutta_3; t += dt) with
    dt on the device equals the host loop in dt, time and state."""
    a, b = pair(case, seed=5)
    t_a = a.advance_cfl(5, 0.25)
    t_b = 0.25
    for _ in range(5):
        dt = b.timestep()
        b.rk3_step(dt)
        t_b += dt
    assert t_a == t_b
    assert same(a, b)
    # replays continue from the device state; eager steps in between
    a.rk3_step(1e-3)
    b.rk3_step(1e-3)
    t_a = a.advance_cfl(4, t_a)
    for _ in range(4):
        dt = b.timestep()
        b.rk3_step(dt)
        t_b += dt
    assert t_a == t_b
    assert same(a, b)


@pytest.mark.parametrize("graphs", ["1", "0"])
@pytest.mark.parametrize("case", ["mixed-periodic", "mixed-walls"])
def test_advance_cfl_fused_reduction_matches_host_time_loop(gpu, monkeypatch, case, graphs):
    """Mixed mode on a grid the stage kernel's 64 x 4 workgroups tile
    exactly: the third stage reduces the next step's CFL minimum, and only
    the first step of a call runs the CFL pass.  dt, time and state equal
    the host loop's bit for bit."""
    monkeypatch.setenv("NS3D_GRAPHS", graphs)
    n = 64
    st, Ma = fields.hit(n, k0=4, Mt=0.3, seed=7)
    p, g = ns.make_params(n, Re=200.0, Ma=Ma, **CASES[case])
    a, b = ns.Solver(p, g), ns.Solver(p, g)
    a.upload_primitive(*st)
    b.upload_primitive(*st)
    a.profile(True)
    t_a = a.advance_cfl(4, 0.0)
    rep = a.profile_report()
    a.profile(False)
    # the stage kernel (on the periodic box with the z sweep folded in)
    stages = rep.get("vdiv_stage", (0, 0))[1] + rep.get("vdiv_zs_stage", (0, 0))[1]
    assert rep["cfl"][1] == 1 and stages == 12
    t_b = 0.0
    for _ in range(4):
        dt = b.timestep()
        b.rk3_step(dt)
        t_b += dt
    assert t_a == t_b
    assert same(a, b)
    # a host step in between: the next call starts from a fresh CFL pass
    a.rk3_step(1e-3)
    b.rk3_step(1e-3)
    t_a = a.advance_cfl(5, t_a)
    for _ in range(5):
        dt = b.timestep()
        b.rk3_step(dt)
        t_b += dt
    assert t_a == t_b
    assert same(a, b)
    assert a.timestep() == b.timestep()


def test_graphs_off_and_profiling_give_the_same_states(gpu, monkeypatch):
    """With per-kernel profiling on, the steps run eagerly; the states do not
    depend on the path (nor on graphs being off)."""
    a, b = pair("mixed-periodic")
    a.profile(True)
    a.advance(4, 2e-3)
    rep = a.profile_report()
    a.profile(False)
    stages = rep.get("vdiv_stage", (0, 0))[1] + rep.get("vdiv_zs_stage", (0, 0))[1]
    assert stages == 12                       # every launch seen: no graph
    b.advance(4, 2e-3)
    assert same(a, b)
    monkeypatch.setenv("NS3D_GRAPHS", "0")
    c, _ = pair("mixed-periodic")
    c.advance(4, 2e-3)
    assert same(a, c)


def test_nonphysical_report_through_replays(gpu):
    """A cell made non-physical before the first step: the report after
    several (replayed) steps names the first step's first bad cell, as a
    single eager step does."""
    n = 12
    st = [np.ones((n, n, n)), np.zeros((n, n, n)), np.zeros((n, n, n)), np.zeros((n, n, n)),
          np.full((n, n, n), 10.0)]
    p, g = ns.make_params(n, Re=100.0, Ma=0.3, recon="WENO5", vis_scheme="C8th")
    ref = ns.Solver(p, g)
    ref.upload_primitive(*st)
    U = [ref.interior(x).copy() for x in ref.conserved()]
    U[1][5, 6, 7] = 3.0
    U[4][5, 6, 7] = 1.0          # E below the kinetic energy: p < 0
    ref.upload_conserved(*U)
    with pytest.raises(ns.NS3DError):
        ref.rk3_step(1e-3)
    want = ref.nonphysical_cell()
    s = ns.Solver(p, g)
    s.upload_conserved(*U)
    with pytest.raises(ns.NS3DError) as e:
        s.advance(4, 1e-3)
    assert e.value.status == ns._lib.NS3D_ERR_NONPHYSICAL
    assert s.nonphysical_cell() == want
