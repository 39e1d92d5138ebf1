"""CG batches replayed from a captured HIP graph (mode cg_graph): every replay, with the
program's own eager work (assembly, residuals, coarse inversion, damage) and synchronisations
between them, gives the same bits as the same launches issued directly, for a single system
(Jacobi and two-level, FP32 and FP64, with a precision switch that forces a re-capture) and for
a specimen ensemble."""
import sys

import numpy as np
import pytest

from scenarios import G, faces, hex_plate, sc_cube
from test_gpu_ensemble import _union

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _sc_run(graph, precond, precisions, timing=0):
    xyz, box = sc_cube(10)
    g = G.Context(G.SC3D, 0.5, G.ELASTIC, xyz, bonds="device")
    g.set_timing(timing)
    g.set_elastic(0.0, False, 200e3, 0.3, 1.0, elmap_literal=0)
    lo, hi = faces(xyz, 2)
    ls = g.load_step(disp=[(lo, "x", 0.0), (lo, "y", 0.0), (lo, "z", 0.0), (hi, "z", 0.005)])
    g.update_geometry()
    g.update_force_state()
    out = []
    for prec in precisions:
        g.mode(dirichlet_static=1, cg_transpose=0, diag10_literal=0, precision=prec, precond=precond, cg_forcing=1,
               cg_tol=1e-3 if prec == 0 else 1e-10, tol_NR=1e-10, max_NR_iter=40, cg_maxit=20000, cg_graph=graph,
               # a fixed SpMV launch (the tuner's choice depends on timings): cooperative slices
               # with the coarse space, one wavefront per slice without
               spmv_tune=0, spmv_unr=2, spmv_wpb=4, spmv_coop=16 if precond == 2 else 0, spmv_spw=2)
        assert g.solve_step_static(ls, -1)
        out.append(g.get("xyz"))
    return out, (g.stat("cg_iters"), g.stat("newton_iters"), g.stat("spmv_launches"))


@pytest.mark.parametrize("precond", [0, 2])
def test_cg_graph_replays_match_direct_launches(precond):
    precisions = [0, 0, 0, 1, 0]  # FP32, a switch to FP64 (re-capture), back to FP32
    say("[gpu] cg graph vs direct, SC 10^3, precond", precond)
    direct, cd = _sc_run(0, precond, precisions)
    graph, cg = _sc_run(1, precond, precisions)
    say("   (cg iterations, newton iterations, SpMVs that did work) direct", cd, "graph", cg)
    assert cd == cg
    for a, b in zip(direct, graph):
        np.testing.assert_array_equal(a, b)


def test_cg_graph_ensemble_matches_direct_launches():
    say("[gpu] cg graph vs direct, 3-specimen hex ensemble, two-level FP32")
    sizes, pulls = [(12, 10), (30, 20), (20, 36)], [0.001, 0.002, 0.0015]
    parts, cases = [], []
    for (nx, ny), pull in zip(sizes, pulls):
        x, box = hex_plate(nx, ny)
        parts.append((x, G.build_bonds(G.HEXAGON2D, 0.5, x)))
        cases.append((x, pull))
    xyz, bonds, first = _union(parts)
    results = []
    for graph in (0, 1):
        g = G.Context(G.HEXAGON2D, 0.5, G.ELASTIC, xyz, bonds=bonds)
        g.set_specimens(first)
        g.set_elastic(0.0, True, 69e3, 0.3, 1.0, elmap_literal=0)
        g.mode(dirichlet_static=1, cg_transpose=0, diag10_literal=0, precision=0, precond=2, coarse_cells=4, cg_tol=1e-3,
               cg_forcing=1, tol_NR=1e-10, max_NR_iter=50, cg_maxit=50000, cg_graph=graph, spmv_tune=0, spmv_unr=2,
               spmv_wpb=4, spmv_coop=0)
        disp = []
        for s, (x, pull) in enumerate(cases):
            ids = np.arange(first[s], first[s + 1])
            y = x[:, 1]
            lo, hi = ids[y < y.min() + 1e-6], ids[y > y.max() - 1e-6]
            disp += [(lo, "x", 0.0), (lo, "y", 0.0), (hi, "y", pull)]
        ls = g.load_step(disp=disp)
        g.update_geometry()
        g.update_force_state()
        xs = []
        for _ in range(3):
            assert g.solve_step_static(ls, -1)
            xs.append(g.get("xyz"))
        results.append((xs, g.specimen_stat("newton_iters").tolist(), g.stat("cg_iters")))
    say("   newton per specimen", results[0][1], results[1][1], "cg", results[0][2], results[1][2])
    assert results[0][1:] == results[1][1:]
    for a, b in zip(results[0][0], results[1][0]):
        np.testing.assert_array_equal(a, b)
