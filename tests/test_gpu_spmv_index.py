"""Column indices of the CG SpMV: int16 offsets j - i where every offset of the pattern fits,
int32 ids otherwise. Both give the same bits (the products and their order do not change), and a
pattern whose 2-hop columns lie more than 32767 rows away falls back to int32."""
import sys

import numpy as np
import pytest

from scenarios import G, faces, hex_plate, sc_cube

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(xyz, lattice, axis, idx16, steps=3):
    g = G.Context(lattice, 0.5, G.ELASTIC, xyz, bonds="device")
    if lattice == G.SC3D:
        g.set_elastic(0.0, False, 200e3, 0.3, 1.0, elmap_literal=0)
        axes = "xyz"
    else:
        g.set_elastic(0.0, True, 69e3, 0.3, 1.0, elmap_literal=0)
        axes = "xy"
    lo, hi = faces(xyz, axis)
    ls = g.load_step(disp=[(lo, a, 0.0) for a in axes] + [(hi, axes[axis], 0.002)])
    g.mode(dirichlet_static=1, cg_transpose=0, diag10_literal=0, precision=0, precond=0, cg_tol=1e-5, tol_NR=1e-9,
           max_NR_iter=40, cg_maxit=200000, spmv_tune=0, spmv_unr=2, spmv_wpb=4, spmv_coop=0, spmv_idx16=idx16)
    g.update_geometry()
    g.update_force_state()
    for _ in range(steps):
        assert g.solve_step_static(ls, -1)
    return g.get("xyz"), int(g.get_mode("spmv_index_bytes"))


def test_int16_offsets_match_int32_ids_bitwise():
    say("[gpu] SpMV int16 column offsets vs int32 ids, SC 10^3")
    xyz, _ = sc_cube(10)
    x16, b16 = _run(xyz, G.SC3D, 2, 1)
    x32, b32 = _run(xyz, G.SC3D, 2, 0)
    assert (b16, b32) == (2, 4)
    np.testing.assert_array_equal(x16, x32)


def test_far_columns_fall_back_to_int32():
    say("[gpu] SpMV column offsets beyond int16: hex strip 16400 x 3")
    xyz, _ = hex_plate(16400, 3)
    x, nbytes = _run(xyz, G.HEXAGON2D, 1, 1, steps=1)  # pulled across its three rows (well conditioned)
    say("   particles", len(xyz), "index bytes", nbytes)
    assert nbytes == 4
    assert np.all(np.isfinite(x))
