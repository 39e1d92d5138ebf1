"""Reuse of the two-level coarse space across assemblies (mode coarse_reuse): the preconditioner
changes, the equilibrium does not. Against the same steps with the coarse space rebuilt at every
assembly: an elastic SC bar, and the cracked 40 x 40 plate whose bonds break (flags bit-exact)."""
import sys

import numpy as np
import pytest

from scenarios import G, faces, rel, sc_cube
from test_gpu_configs import INC, _cracked_plate

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _sc(reuse, steps=6):
    xyz, _ = sc_cube(12)
    g = G.Context(G.SC3D, 0.5, G.ELASTIC, xyz, bonds="device")
    g.set_elastic(0.0, False, 200e3, 0.3, 1.0, elmap_literal=0)
    g.mode(dirichlet_static=1, cg_transpose=0, diag10_literal=0, precision=0, precond=2, cg_tol=1e-3, cg_forcing=1,
           tol_NR=1e-10, max_NR_iter=40, cg_maxit=20000, predictor=1, coarse_reuse=reuse)
    lo, hi = faces(xyz, 2)
    ls = g.load_step(disp=[(lo, "x", 0.0), (lo, "y", 0.0), (lo, "z", 0.0), (hi, "z", 0.004)])
    g.update_geometry()
    g.update_force_state()
    u = []
    for _ in range(steps):
        assert g.solve_step_static(ls, -1)
        u.append(g.get("xyz").reshape(-1, 3) - xyz)
    return u, g.stat("coarse_builds"), g.stat("cg_iters")


def test_coarse_reuse_same_equilibrium_sc():
    say("[gpu] coarse reuse, SC 12^3, 6 steps")
    u0, b0, c0 = _sc(0)
    u1, b1, c1 = _sc(1)
    say("   builds", b0, "->", b1, "CG iterations", c0, "->", c1)
    assert b1 < b0
    for a, b in zip(u1, u0):
        assert rel(a, b) < 1e-8


def _crack(reuse, steps=5):
    xyz, box, bl, csr = _cracked_plate(40)
    g = G.Context(G.HEXAGON2D, 0.5, G.ELASTIC, xyz, bonds=csr)
    g.set_elastic(0.0, True, 69e3, 0.3, 1e-3, elmap_literal=0)
    g.mode(dirichlet_static=1, cg_transpose=0, diag10_literal=0, precision=1, precond=2, cg_tol=1e-12, tol_NR=1e-10,
           max_NR_iter=60, cg_maxit=50000, coarse_reuse=reuse)
    y = xyz[:, 1]
    lo, hi = np.nonzero(y < y.min() + 1e-6)[0], np.nonzero(y > y.max() - 1e-6)[0]
    ls = g.load_step(disp=[(lo, "x", 0.0), (lo, "y", -INC), (hi, "x", 0.0), (hi, "y", INC)])
    g.update_geometry()
    g.update_force_state()
    for _ in range(steps):
        assert g.solve_step_static(ls, -1)
    return g.get("xyz").reshape(-1, 3) - xyz, g.get("bdamage"), g.stat("coarse_builds")


def test_coarse_reuse_cracked_plate_breaking():
    say("[gpu] coarse reuse, cracked 40 x 40 plate with bond breaking")
    u0, bd0, b0 = _crack(0)
    u1, bd1, b1 = _crack(1)
    say("   builds", b0, "->", b1, "broken", int((bd0 > 0.5).sum()))
    assert int((bd0 > 0.5).sum()) > 0
    np.testing.assert_array_equal(bd1, bd0)
    assert rel(u1, u0) < 1e-8
