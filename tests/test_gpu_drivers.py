"""Solver drivers on the GPU against the oracle: Absolute-mode boundary conditions, the static
step-halving sequence, and the fatigue cycle driver with every load template and with cycle
jumps (linear tau < 1 and the exponential time map). FP64 on both sides; equilibria within 1e-8."""
import math
import sys

import numpy as np
import pytest

import oracle as O
from scenarios import Pair, faces, hex_plate, rel
from test_oracle_laws import _halving_setup, _ratio_after_one_iteration

pytestmark = pytest.mark.gpu

FP64 = dict(precision=1, cg_tol=1e-13, cg_maxit=20000, cg_transpose=0, diag10_literal=0)
ORA = dict(cg_tol=1e-14, cg_transpose=0, diag10_literal=0)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# ------------------------------------------------------------------ Absolute boundary conditions
def test_load_step_bcs_apply_in_order():
    """Solver::updateForceBC then updateDisplacementBC (solver.h:130-190), one BC after the other:
    overlapping groups, Absolute after Relative and back on the same particles, a group with a
    repeated id (applied once per occurrence), and more BCs than one launch of the batched BC kernel
    holds (64), against the oracle bit for bit in positions, Pex and constraint flags."""
    say("[gpu] load-step BCs in order")
    xyz, box = hex_plate(12, 12)
    p = Pair(O.HEXAGON2D, 0.5, O.ELASTIC, xyz, box)
    p.set_elastic(0.0, True, 69e3, 0.3, 1.0)
    n = len(xyz)
    a, b = np.arange(0, 40), np.arange(25, 70)
    dup = np.array([3, 3, 4, 90, 3])
    disp = [(a, "x", 1e-3), (b, "x", 2.5e-3, O.ABSOLUTE), (a, "y", -4e-4), (dup, "y", 1e-4), (b, "z", 0.0),
            (dup, "x", 7e-4, O.ABSOLUTE), (a, "x", 3e-4)]
    force = [(a, 1.0, 2.0, 0.0), (b, -3.0, 0.5, 0.0, O.ABSOLUTE), (dup, 0.25, 0.0, 0.0), (a, 0.0, 1.0, 0.0)]
    rng = np.random.default_rng(11)
    for k in range(70):  # one BC per particle group of 3, past the 64 of one batch
        ids = rng.choice(n, 3, replace=False)
        disp.append((ids, "xyz"[k % 2], float(rng.uniform(-1e-3, 1e-3)), O.ABSOLUTE if k % 5 == 0 else O.RELATIVE))
    ls_o, ls_g = p.load(disp=disp, force=force)
    for _ in range(2):
        p.o.apply_bc(ls_o)
        p.g.apply_bc(ls_g)
        np.testing.assert_array_equal(p.g.get("xyz"), p.o.getd("xyz"))
        np.testing.assert_array_equal(p.g.get("Pex"), p.o.getd("Pex"))
        np.testing.assert_array_equal(p.g.get("disp_constraint"), p.o.geti("cons"))


def test_absolute_bcs_match_oracle():
    """Solver::updateDisplacementBC / updateForceBC in LoadMode::Absolute (solver.h:140-161,
    176-183): a displacement BC moves each particle of the group *to* (step, 0, 0) (or (0, step, 0),
    (0, 0, step)), and a force BC sets Pex = f / |group| instead of adding it (Q18). A corner
    particle at the origin is pinned by an Absolute x BC; the top row carries an Absolute force."""
    say("[gpu] absolute BCs")
    xyz, box = hex_plate(10, 10)
    k0 = int(np.lexsort((xyz[:, 0], xyz[:, 1]))[0])  # lowest row, leftmost
    xyz = xyz - xyz[k0]
    p = Pair(O.HEXAGON2D, 0.5, O.ELASTIC, xyz, box)
    p.set_elastic(0.0, True, 69e3, 0.3, 1.0)
    bottom, top = faces(xyz, 1)
    p.o.mode(dirichlet_static=1, tol_NR=1e-10, max_NR_iter=50, **ORA)
    p.g.mode(dirichlet_static=1, tol_NR=1e-10, max_NR_iter=50, **FP64)
    for s in (p.o, p.g):
        s.update_geometry()
        s.update_force_state()
    corner = np.array([k0])
    for step, (sx, fy) in enumerate([(0.0, 30.0), (2e-3, 45.0), (-1e-3, 10.0)]):
        ls_o, ls_g = p.load(disp=[(bottom, "y", 0.0), (corner, "x", sx, O.ABSOLUTE)],
                            force=[(top, 0.0, fy, 0.0, O.ABSOLUTE)])
        p.o.apply_bc(ls_o)
        p.g.apply_bc(ls_g)
        xo, xg = p.o.getd("xyz").reshape(-1, 3), p.g.get("xyz").reshape(-1, 3)
        if step == 0:
            np.testing.assert_array_equal(xg, xo)
        else:  # the previous solves agree to rounding, the BC's target exactly
            assert rel(xg - xyz, xo - xyz) < 1e-12
        np.testing.assert_array_equal(xg[k0], [sx, 0.0, 0.0])
        assert xo[k0, 0] == sx and xo[k0, 1] == 0.0 and xo[k0, 2] == 0.0    # moved *to* (sx, 0, 0)
        pex = p.g.get("Pex").reshape(-1, 3)
        np.testing.assert_array_equal(pex, p.o.getd("Pex").reshape(-1, 3))  # set from the BC alone
        assert pex[top, 1] == pytest.approx(fy / len(top), rel=1e-15)       # set, not accumulated
        assert p.o.solve_step_static(ls_o, -1)
        assert p.g.solve_step_static(ls_g, -1)
        err = rel(p.g.get("xyz") - xyz.ravel(), p.o.getd("xyz") - xyz.ravel())
        say(f"   step {step}: displacement rel err {err:.2e}")
        assert err < 1e-8
        assert rel(p.g.get("reaction"), p.o.getd("reaction")) < 1e-8


# ------------------------------------------------------------------ static halving sequence
def test_static_halving_sequence_matches_oracle():
    """SolverStatic::solveProblem (solver_static.h:64-95) with a forced failure: max_NR_iter = 2
    and a Newton tolerance between the one-iteration ratio of the full step and that of its halves
    (measured on the oracle), so [D] runs as D (fails), D/2, D/2 on both sides: the same number
    of steps, positions within 1e-8, and the failed attempt's force left in Pex (Q7) exactly."""
    say("[gpu] halving sequence")
    D, F = 0.2, 1e-6
    s, xyz, bottom, top, mid = _halving_setup()

    def step_o(scale):
        ls = O.LoadStep()
        for ax in "xyz":
            ls.add_disp(bottom, ax, 0.0)
        ls.add_disp(top, "z", D * scale)
        ls.add_force(mid, 0.0, 0.0, F * scale)
        return ls

    pex0 = s.getd("Pex").copy()
    rD = _ratio_after_one_iteration(s, step_o(1.0))
    s.reset(True)
    s.setd("Pex", pex0)
    r1 = _ratio_after_one_iteration(s, step_o(0.5))
    r2 = _ratio_after_one_iteration(s, step_o(0.5))
    assert max(r1, r2) < 0.5 * rD
    tol = math.sqrt(max(r1, r2) * rD)
    # the same bar on a fresh pair: pre-load, then the step under test
    p = Pair(O.SC3D, 0.5, O.ELASTIC, xyz, None)
    p.set_elastic(0.0, False, 200e3, 0.3, 1.0)
    p.o.mode(dirichlet_static=1, tol_NR=1e-11, max_NR_iter=50, **ORA)
    p.g.mode(dirichlet_static=1, tol_NR=1e-11, max_NR_iter=50, **FP64)
    for q in (p.o, p.g):
        q.update_geometry()
        q.update_force_state()
    pre_o, pre_g = p.load(disp=[(bottom, "x", 0.0), (bottom, "y", 0.0), (bottom, "z", 0.0), (top, "z", 0.02)])
    assert p.o.solve_step_static(pre_o, -1) and p.g.solve_step_static(pre_g, -1)
    x_conv = p.o.getd("xyz").reshape(-1, 3).copy()
    steps_o, steps_g = p.load(disp=[(bottom, "x", 0.0), (bottom, "y", 0.0), (bottom, "z", 0.0), (top, "z", D)],
                              force=[(mid, 0.0, 0.0, F)])
    for q in (p.o, p.g):
        q.mode(max_NR_iter=2, tol_NR=tol)
    p.o.solve_problem_static([steps_o], -1)
    n_g = p.g.solve_problem_static([steps_g], -1)
    assert p.o.info("n_steps") == 2 and n_g == 2
    assert list(p.o.geti("newton_log")[-3:]) == [2, 1, 1]
    xo, xg = p.o.getd("xyz").reshape(-1, 3), p.g.get("xyz").reshape(-1, 3)
    np.testing.assert_allclose(xg[top, 2] - x_conv[top, 2], D, rtol=1e-13)
    assert rel(xg - xyz, xo - xyz) < 1e-8
    pex_o, pex_g = p.o.getd("Pex").reshape(-1, 3), p.g.get("Pex").reshape(-1, 3)
    np.testing.assert_array_equal(pex_g, pex_o)
    assert pex_g[mid, 2].sum() == pytest.approx(2 * F, rel=1e-12)   # F (failed) + F/2 + F/2


# ------------------------------------------------------------------ cycle driver, all templates
def _hcf_plate():
    xyz, box = hex_plate(8, 12)
    p = Pair(O.HEXAGON2D, 0.5, O.HCF, xyz, box)
    p.set_hcf(3.0, True, 110e3, 0.34, 10.0, 2.019, 0.5)
    x, y = xyz[:, 0], xyz[:, 1]
    left = np.nonzero(x < x.min() + 0.5)[0]
    bottom, top = faces(xyz, 1)
    mid = np.nonzero(np.abs(y - np.median(y)) < 0.45)[0]
    t = np.zeros(len(xyz), dtype=np.int32)
    t[top] = 4
    t[bottom] = 4
    p.set_type(t)
    p.o.search_nonlocal(1.5)
    p.g.search_nonlocal(1.5)
    p.o.mode(dirichlet_static=1, tol_NR=1e-10, max_NR_iter=60, **ORA)
    p.g.mode(dirichlet_static=1, tol_NR=1e-10, max_NR_iter=60, **FP64)
    for s in (p.o, p.g):
        s.update_geometry()
        s.update_force_state()
    return p, xyz, left, bottom, top, mid


@pytest.mark.parametrize("ltype,tau,tmode", [(0, 0.5, 0), (2, 0.3, 0), (3, 2.0, 1), (1, 2.0, 1)])
def test_cycle_driver_templates_and_jumps_match_oracle(ltype, tau, tmode):
    """SolverFatigue::solveProblemCyclic (solver_fatigue.h:182-218) with LoadUniaxialDisp,
    LoadDogBoneDisp, LoadCTForce and LoadUniaxialForce (load_step_fatigue.h:22-85), n_interval 2,
    over a spectrum of 8 cycles with dN = int(1/tau) (2, 3) or int(1/tau exp(t/tau)) (tau 2:
    0, 0, 1, 2, 3, 6): the final cycle count, loads, constraint flags, positions and fatigue damage
    against the oracle."""
    say("[gpu] cycle driver", ltype, tau, tmode)
    p, xyz, left, bottom, top, mid = _hcf_plate()
    if ltype in (0, 2):
        spec = [0.0] + sum(([4e-3 + 1e-4 * k, 1e-3 + 2e-4 * k] for k in range(8)), [])
    else:
        spec = [40.0] + sum(([400.0 - 5.0 * k, 40.0 + 2.0 * k] for k in range(8)), [])
    groups = [top, bottom] if ltype == 2 else ([mid, top, bottom] if ltype == 3 else [left, top, bottom])
    No = p.o.solve_problem_cyclic(ltype, 2, spec, groups, tau, 4, tmode=tmode)
    Ng = p.g.solve_problem_cyclic(ltype, 2, spec, groups, tau, 4, tmode=tmode)
    assert No == Ng and No >= 8
    np.testing.assert_array_equal(p.g.get("disp_constraint"), p.o.geti("cons"))
    np.testing.assert_allclose(p.g.get("Pex"), p.o.getd("Pex"), rtol=1e-13, atol=1e-12)
    err = rel(p.g.get("xyz") - xyz.ravel(), p.o.getd("xyz") - xyz.ravel())
    Do, Dg = p.o.getd("damage"), p.g.get("damage")
    say(f"   N {No}: displacement rel err {err:.2e}, damage max {Do.max():.3e} rel err {rel(Dg, Do):.2e}")
    assert err < 1e-8
    assert Do.max() > 1e-4
    assert rel(Dg, Do) < 1e-8
    svo, svg = p.o.getd("state_var").reshape(-1, 7)[:, :3], p.g.get("state_var").reshape(-1, 7)[:, :3]
    np.testing.assert_allclose(svg, svo, rtol=1e-8, atol=1e-12)
