"""Asynchronous dump writer (lpm_dump_async): the frames it appends must be byte-identical to
the reference's Assembly::writeDump (assembly.h:309-335) for the state at the call, even though
the context keeps stepping while the writer thread formats."""
import sys

import numpy as np
import pytest

from scenarios import G, faces, sc_cube

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _frame(g, step, box):
    """writeDump's text, formatted here with the reference's format strings."""
    n = g.n
    x = g.get("xyz").reshape(-1, 3)
    x0 = g.get("xyz_initial").reshape(-1, 3)
    st = g.get("stress").reshape(-1, 6)
    dm, dv = g.get("damage"), g.get("damage_visual")
    sv = g.get("state_var").reshape(-1, 7)
    t = g.get("type")
    out = ["ITEM: TIMESTEP\n%d\n" % step, "ITEM: NUMBER OF ATOMS\n%d\n" % n, "ITEM: BOX BOUNDS pp pp pp\n"]
    out += ["%8.8f %8.8f\n" % (box[2 * k], box[2 * k + 1]) for k in range(3)]
    out.append("ITEM: ATOMS id type x y z dx dy dz s11 s22 s33 s23 s13 s12 damage damage_visual state_var[0] \n")
    fmt = "%d %d " + " ".join(["%.4e"] * 15) + "\n"
    for i in range(n):
        d = x[i] - x0[i]
        out.append(fmt % (i, t[i], x[i, 0], x[i, 1], x[i, 2], d[0], d[1], d[2], *st[i], dm[i], dv[i], sv[i, 0]))
    return "".join(out)


def test_async_dump_matches_reference_format(tmp_path):
    say("[gpu] async dump")
    xyz, box = sc_cube(8)
    g = G.Context(G.SC3D, 0.5, G.ELASTIC, xyz)
    g.set_elastic(0.0, False, 200e3, 0.3, 1.0, elmap_literal=0)
    g.mode(dirichlet_static=1, cg_transpose=0, diag10_literal=0, cg_tol=1e-8, tol_NR=1e-10, max_NR_iter=30,
           cg_maxit=5000)
    t = np.zeros(len(xyz), dtype=np.int32)
    t[::7] = 3
    g.set("type", t)
    bottom, top = faces(xyz, 2)
    ls = g.load_step(disp=[(bottom, "x", 0.0), (bottom, "y", 0.0), (bottom, "z", 0.0), (top, "z", 0.02)])
    g.update_geometry()
    g.update_force_state()
    path = tmp_path / "run.dump"
    expected = []
    for step in range(1, 4):
        assert g.solve_step_static(ls, -1)
        g.dump_async(path, step, box)
        expected.append(_frame(g, step, box))  # the state the snapshot holds (stream order)
    g.dump_wait()
    text = path.read_text()
    say("   frames", text.count("ITEM: TIMESTEP"), "bytes", len(text))
    assert text == "".join(expected)
