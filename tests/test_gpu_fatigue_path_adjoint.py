"""Sensitivity of a displacement objective to the HCF law's parameters (A, B, d per particle)
through a whole fatigue load path of several cycles, on the GPU (lpm_fatigue_record +
lpm_adjoint_fatigue_path), against central differences of the FP64 oracle running the same path.

The path: a loading fatigue step, then three cycles of (unloading step, state update, fatigue
damage update with a fixed cycle step, reloading step, state update), the objective read at the
last reloaded equilibrium. Each update's rate depends on the cycle's energy extremes, which depend
on the bond damage the earlier updates raised, so the path's gradient differs from the one-update
sensitivity (lpm_adjoint_fatigue_params), which holds the earlier history fixed."""
import sys

import numpy as np
import pytest

import oracle as O
from scenarios import G, faces, hex_plate

pytestmark = pytest.mark.gpu

A0, B0, D0 = 5.206e-5, 2.019, 0.5
DT = 1e8          # cycles per update: damage grows by a few percent per update, none clamped
FMAX = 400.0
CYCLES = 3


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _setup():
    xyz, box = hex_plate(10, 16)
    left = np.nonzero(xyz[:, 0] < 0.5)[0]
    bottom, top = faces(xyz, 1)
    t = np.zeros(len(xyz), dtype=np.int32)
    t[top] = 4
    t[bottom] = 4
    return xyz, box, left, bottom, top, t


def _path(solve, update_state, update_damage, mk):
    assert solve(mk(FMAX))
    update_state()
    for _ in range(CYCLES):
        assert solve(mk(-0.9 * FMAX))
        update_state()
        update_damage()
        assert solve(mk(0.9 * FMAX))
        update_state()


def _oracle(xyz, box, left, bottom, top, t, A, dB=0.0, dd=0.0):
    """Top-edge y displacement sum at the end of the path, per-particle A (array)."""
    s = O.System(O.HEXAGON2D, 0.5, O.HCF, xyz, box)
    s.mode(elmap_literal=0, cg_tol=1e-14, tol_NR=1e-10, max_NR_iter=60, cg_transpose=0, diag10_literal=0)
    s.set_hcf(4.0, True, 110e3, 0.34, A0, B0 + dB, D0 + dd)
    for j in np.nonzero(A != A0)[0]:
        s.set_hcf(4.0, True, 110e3, 0.34, float(A[j]), B0 + dB, D0 + dd, pid=int(j))
    s.seti("type", t)
    s.search_nonlocal(1.5)
    s.update_geometry()
    s.update_force_state()

    def mk(f):
        return O.LoadStep().add_disp(left, "x", 0.0).add_disp(bottom, "y", 0.0).add_force(top, 0.0, f, 0.0)

    _path(s.solve_step_fatigue, s.update_state_var, lambda: s.update_fatigue_damage(DT, 4), mk)
    return (s.getd("xyz").reshape(-1, 3) - xyz)[top, 1].sum(), s.getd("damage")


def test_fatigue_path_sensitivity_matches_central_differences():
    say("[gpu] HCF parameter sensitivity through a recorded fatigue path")
    xyz, box, left, bottom, top, t = _setup()
    n = len(xyz)
    g = G.Context(G.HEXAGON2D, 0.5, G.HCF, xyz, bonds="device")
    g.set_hcf(4.0, True, 110e3, 0.34, A0, B0, D0, elmap_literal=0)
    g.set("type", t)
    g.search_nonlocal(1.5)
    g.mode(cg_tol=1e-13, tol_NR=1e-10, max_NR_iter=60, cg_maxit=20000, precision=1, cg_transpose=0, diag10_literal=0)
    g.update_geometry()
    g.update_force_state()

    def mk(f):
        return g.load_step(disp=[(left, "x", 0.0), (bottom, "y", 0.0)], force=[(top, 0.0, f, 0.0)])

    g.fatigue_record(True)
    _path(g.solve_step_fatigue, g.update_state_var, lambda: g.update_fatigue_damage(DT, 4), mk)
    Dg = g.get("damage")
    x_end = g.get("xyz").copy()
    w = np.zeros(2 * n)
    w[2 * top + 1] = 1.0
    Gv, gA, gB, gd, nsolve, rres = g.adjoint_fatigue_path(w_x=w)
    say(f"   G {Gv:.12e} transposed solves {nsolve} largest rel residual {rres:.1e}")
    assert rres < 1e-9
    assert nsolve >= 1 + 2 * CYCLES  # the current state and the two states of every cycle
    np.testing.assert_array_equal(g.get("xyz"), x_end)  # the current state is restored
    # the one-update sensitivity at the same state, for comparison
    _, gA1, _, _, _, _ = g.adjoint_fatigue_params(w_x=w)
    g.fatigue_record(False)
    A = np.full(n, A0)
    G0, Do = _oracle(xyz, box, left, bottom, top, t, A)
    say("   damage max", Do.max(), "GPU vs oracle", np.abs(Dg - Do).max())
    np.testing.assert_allclose(Dg, Do, rtol=1e-7, atol=1e-15)
    assert 0.05 < Do.max() < 0.6
    assert Gv - xyz[top, 1].sum() == pytest.approx(G0, rel=1e-8)
    assert np.all(gA[t == 4] == 0)
    rng = np.random.default_rng(5)
    r = rng.uniform(-1, 1, n)
    # A step of 1e-3 along the random direction moves some particles across a kink of the path
    # (a later cycle's energy maximum overtaking s0, or one end of a bond overtaking the other in
    # max(D_i, D_j)): the central difference there is -4.8152e-4 at 1e-3 and -4.81857425e-4 at
    # 3e-4 and 1e-4 (oracle runs alone). The steps below stay inside one branch.
    h, hr, hb = 1e-3, 1e-4, 1e-4
    cases = {
        "A scaled": (np.sum(gA * A), dict(A=A * (1 + h)), dict(A=A * (1 - h)), h),
        "A along a random direction": (np.sum(gA * A * r), dict(A=A * (1 + hr * r)), dict(A=A * (1 - hr * r)), hr),
        "B shifted": (np.sum(gB), dict(A=A, dB=hb), dict(A=A, dB=-hb), hb),
        "d shifted": (np.sum(gd), dict(A=A, dd=hb), dict(A=A, dd=-hb), hb),
    }
    fd_A = None
    for name, (adj, up, dn, hh) in cases.items():
        fd = (_oracle(xyz, box, left, bottom, top, t, **up)[0] - _oracle(xyz, box, left, bottom, top, t, **dn)[0]) / (2 * hh)
        err = abs(adj - fd) / abs(fd)
        say(f"   {name}: adjoint {adj:.10e} fd {fd:.10e} rel {err:.2e}")
        assert err < 1e-5
        if fd_A is None:
            fd_A = fd
    # the earlier cycles carry a large share of the path's gradient
    one = np.sum(gA1 * A)
    say(f"   one-update sensitivity (A scaled) {one:.10e}: rel {abs(one - fd_A) / abs(fd_A):.2e} from the path's")
    assert abs(one - fd_A) > 0.2 * abs(fd_A)


def _driver_oracle(xyz, box, left, bottom, top, t, A, dB=0.0, dd=0.0):
    s = O.System(O.HEXAGON2D, 0.5, O.HCF, xyz, box)
    s.mode(elmap_literal=0, cg_tol=1e-14, tol_NR=1e-10, max_NR_iter=60, cg_transpose=0, diag10_literal=0)
    s.set_hcf(4.0, True, 110e3, 0.34, A_DRV, B0 + dB, D0 + dd)
    for j in np.nonzero(A != A_DRV)[0]:
        s.set_hcf(4.0, True, 110e3, 0.34, float(A[j]), B0 + dB, D0 + dd, pid=int(j))
    s.seti("type", t)
    s.search_nonlocal(1.5)
    s.update_geometry()
    s.update_force_state()
    pre = O.LoadStep().add_disp(left, "x", 0.0).add_disp(bottom, "y", 0.0).add_force(top, 0.0, SPEC[0], 0.0)
    s.solve_problem_fatigue_static([pre])
    N = s.solve_problem_cyclic(1, 1, SPEC, [left, top, bottom], 0.5, 4, tmode=0)
    hold = O.LoadStep().add_disp(left, "x", 0.0).add_disp(bottom, "y", 0.0).add_force(top, 0.0, 0.0, 0.0)
    assert s.solve_step_fatigue(hold)
    return (s.getd("xyz").reshape(-1, 3) - xyz)[top, 1].sum(), s.getd("damage"), N


A_DRV = 1000.0
SPEC = [40.0] + [400.0, 40.0] * 8


def test_fatigue_path_sensitivity_through_the_cycle_driver():
    """The same sensitivity recorded through SolverFatigue::solveProblemStatic and
    solveProblemCyclic (solver_fatigue.h:128-218, LoadUniaxialForce, n_interval 1, linear time map
    with tau 0.5: four cycles jumping 2 spectrum cycles each), then a step holding the last load;
    against the oracle's central differences of the whole sequence."""
    say("[gpu] HCF parameter sensitivity through the cycle driver")
    xyz, box, left, bottom, top, t = _setup()
    n = len(xyz)
    g = G.Context(G.HEXAGON2D, 0.5, G.HCF, xyz, bonds="device")
    g.set_hcf(4.0, True, 110e3, 0.34, A_DRV, B0, D0, elmap_literal=0)
    g.set("type", t)
    g.search_nonlocal(1.5)
    g.mode(cg_tol=1e-13, tol_NR=1e-10, max_NR_iter=60, cg_maxit=20000, precision=1, cg_transpose=0, diag10_literal=0)
    g.update_geometry()
    g.update_force_state()
    g.fatigue_record(True)
    g.solve_problem_fatigue_static([g.load_step(disp=[(left, "x", 0.0), (bottom, "y", 0.0)],
                                                force=[(top, 0.0, SPEC[0], 0.0)])])
    N = g.solve_problem_cyclic(1, 1, SPEC, [left, top, bottom], 0.5, 4, tmode=0)
    assert g.solve_step_fatigue(g.load_step(disp=[(left, "x", 0.0), (bottom, "y", 0.0)], force=[(top, 0.0, 0.0, 0.0)]))
    w = np.zeros(2 * n)
    w[2 * top + 1] = 1.0
    Gv, gA, gB, gd, nsolve, rres = g.adjoint_fatigue_path(w_x=w)
    g.fatigue_record(False)
    A = np.full(n, A_DRV)
    G0, Do, No = _driver_oracle(xyz, box, left, bottom, top, t, A)
    say(f"   N {N} (oracle {No}), damage max {Do.max():.3e}, transposed solves {nsolve}, rel residual {rres:.1e}")
    assert N == No == 8
    assert rres < 1e-9
    assert 0.02 < Do.max() < 0.6
    assert Gv - xyz[top, 1].sum() == pytest.approx(G0, rel=1e-8)
    rng = np.random.default_rng(7)
    r = rng.uniform(-1, 1, n)
    h = 1e-4
    cases = {
        "A scaled": (np.sum(gA * A), dict(A=A * (1 + h)), dict(A=A * (1 - h))),
        "A along a random direction": (np.sum(gA * A * r), dict(A=A * (1 + h * r)), dict(A=A * (1 - h * r))),
        "B shifted": (np.sum(gB), dict(A=A, dB=h), dict(A=A, dB=-h)),
        "d shifted": (np.sum(gd), dict(A=A, dd=h), dict(A=A, dd=-h)),
    }
    for name, (adj, up, dn) in cases.items():
        fd = (_driver_oracle(xyz, box, left, bottom, top, t, **up)[0] -
              _driver_oracle(xyz, box, left, bottom, top, t, **dn)[0]) / (2 * h)
        err = abs(adj - fd) / abs(fd)
        say(f"   {name}: adjoint {adj:.10e} fd {fd:.10e} rel {err:.2e}")
        assert err < 1e-5
