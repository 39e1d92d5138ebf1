"""CG batch schedules (modes cg_batch, cg_batch_auto, cg_batch_next) change only how many
iterations are queued between reads of the device's stop flag: iterations queued past the stop
are no-ops. Every schedule must therefore give the same bits, the same iteration counts and the
same Newton counts, on the single-specimen Newton loop and on the specimen-ensemble one. The
SpMV launch shape is fixed (spmv_tune=0), since the tuner's pick sets the order of its sums."""
import sys

import numpy as np
import pytest

from scenarios import G, faces, hex_plate, sc_cube
from test_gpu_ensemble import _union

pytestmark = pytest.mark.gpu

# (cg_batch, cg_batch_auto, cg_batch_next): fixed batches of 16, batches of 2 (a read after every
# other iteration), the default sized first batch with batches of 6, and a sized first batch with
# batches of 2 after it
SCHEDULES = [(16, 0, 0), (2, 0, 0), (16, 1, 6), (16, 1, 2)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _sc_run(schedule, steps=6):
    xyz, _ = sc_cube(12)
    g = G.Context(G.SC3D, 0.5, G.ELASTIC, xyz, bonds="device")
    g.set_elastic(0.0, False, 200e3, 0.3, 1.0, elmap_literal=0)
    g.mode(dirichlet_static=1, cg_transpose=0, diag10_literal=0, precision=0, precond=2, cg_tol=1e-3, cg_forcing=1,
           tol_NR=1e-10, max_NR_iter=40, cg_maxit=20000, predictor=2, coarse_reuse=1)
    b, auto, nxt = schedule
    # one fixed SpMV launch shape: the tuner's pick sets the SpMV's summation order
    g.mode(cg_batch=b, cg_batch_auto=auto, cg_batch_next=nxt, spmv_tune=0)
    lo, hi = faces(xyz, 2)
    ls = g.load_step(disp=[(lo, "x", 0.0), (lo, "y", 0.0), (lo, "z", 0.0), (hi, "z", 0.004)])
    g.update_geometry()
    g.update_force_state()
    counts = []
    for _ in range(steps):
        assert g.solve_step_static(ls, -1)
        counts.append((g.stat("newton_iters"), g.stat("cg_iters"), g.stat("coarse_builds")))
    return g.get("xyz"), counts


def test_cg_batch_schedules_give_the_same_bits_sc():
    say("[gpu] CG batch schedules, SC 12^3, 6 load steps, predictor 2, two-level CG")
    ref_x, ref_c = _sc_run(SCHEDULES[0])
    say("   counts (Newton, CG, coarse builds) per step", ref_c)
    for sch in SCHEDULES[1:]:
        x, c = _sc_run(sch)
        say("   schedule", sch, "identical" if np.array_equal(x, ref_x) else "DIFFERENT")
        assert c == ref_c, (sch, c)
        np.testing.assert_array_equal(x, ref_x)


def _ens_run(schedule, steps=4):
    sizes, strains = [(30, 34), (50, 60), (24, 20)], [2e-4, 1e-4, 3e-4]
    parts, cases = [], []
    for (nx, ny), e in zip(sizes, strains):
        x, box = hex_plate(nx, ny)
        parts.append((x, G.build_bonds(G.HEXAGON2D, 0.5, x)))
        cases.append((x, e * (x[:, 1].max() - x[:, 1].min())))
    xyz, bonds, first = _union(parts)
    g = G.Context(G.HEXAGON2D, 0.5, G.ELASTIC, xyz, bonds=bonds)
    g.set_specimens(first)
    g.set_elastic(0.0, True, 110e3, 0.34, 1.0, elmap_literal=0)
    g.mode(dirichlet_static=1, cg_transpose=0, diag10_literal=0, precision=0, precond=2, cg_tol=1e-3, cg_forcing=1,
           tol_NR=1e-10, max_NR_iter=30, cg_maxit=50000, predictor=2)
    b, auto, nxt = schedule
    # one fixed SpMV launch shape: the tuner's pick sets the SpMV's summation order
    g.mode(cg_batch=b, cg_batch_auto=auto, cg_batch_next=nxt, spmv_tune=0)
    disp = []
    for s, (x, pull) in enumerate(cases):
        ids = np.arange(first[s], first[s + 1])
        y = x[:, 1]
        lo, hi = ids[y < y.min() + 1e-6], ids[y > y.max() - 1e-6]
        disp += [(lo, "x", 0.0), (lo, "y", 0.0), (hi, "y", pull)]
    ls = g.load_step(disp=disp)
    g.update_geometry()
    g.update_force_state()
    counts = []
    for _ in range(steps):
        assert g.solve_step_static(ls, -1)
        counts.append((g.stat("cg_iters"), tuple(g.specimen_stat("newton_iters").tolist())))
    return g.get("xyz"), counts


def test_cg_batch_schedules_give_the_same_bits_ensemble():
    say("[gpu] CG batch schedules, 3 hex specimens, 4 load steps, predictor 2, two-level CG")
    ref_x, ref_c = _ens_run(SCHEDULES[0])
    say("   counts (CG, Newton per specimen) per step", ref_c)
    for sch in SCHEDULES[1:]:
        x, c = _ens_run(sch)
        say("   schedule", sch, "identical" if np.array_equal(x, ref_x) else "DIFFERENT")
        assert c == ref_c, (sch, c)
        np.testing.assert_array_equal(x, ref_x)
