"""The cell-grid nonlocal neighbour search (mode nl_cells, the default) against the all-pairs search
(nl_cells=0) and the FP64 oracle's Assembly::searchNonlocalNeighbors restatement
(assembly.h:117-130): every list |x_i - x_j| < ratio * L_i, self included, in id order, bit for bit.

Per-particle nonlocal lengths differ (a random fifth of the particles has a 1.7x longer one, so the
grid is sized by the largest cutoff and the lists are not symmetric), positions are jittered so
that distances near the cutoff vary, and an ensemble of specimens that overlap in space checks that
no list crosses a specimen boundary. A cube whose interior lists exceed 4096 entries exercises the
fallback to the all-pairs search."""
import sys

import numpy as np
import pytest

import oracle as O
from scenarios import G, hex_plate, sc_cube

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _lattice(case):
    rng = np.random.default_rng(3)
    if case == "hex":
        xyz, box = hex_plate(34, 30)
        lat, L0 = O.HEXAGON2D, 2.5
    elif case == "sc":
        xyz, box = sc_cube(12)
        lat, L0 = O.SC3D, 1.6
    else:
        box = [0.0, 5.0, 0.0, 5.0, 0.0, 5.0]
        xyz = O.generate(O.FCC3D, box, 0.5)
        lat, L0 = O.FCC3D, 1.3
    xyz = xyz + rng.uniform(-0.01, 0.01, xyz.shape) * (1.0 if lat != O.HEXAGON2D else np.array([1.0, 1.0, 0.0]))
    longer = np.nonzero(rng.uniform(size=len(xyz)) < 0.2)[0]
    return lat, xyz, box, L0, longer


def _gpu_lists(g, ratio, cells):
    g.mode(nl_cells=cells)
    g.search_nonlocal(ratio)
    return g.get("nl_ptr").copy(), g.get("nl_idx").copy()


@pytest.mark.parametrize("case", ["hex", "sc", "fcc"])
def test_cell_search_matches_all_pairs_and_oracle(case):
    lat, xyz, box, L0, longer = _lattice(case)
    say(f"[gpu] nonlocal search {case}: {len(xyz)} particles, {len(longer)} with a longer length")
    g = G.Context(lat, 0.5, G.HCF, xyz, bonds="device")
    o = O.System(lat, 0.5, O.HCF, xyz, box)
    plane = lat == O.HEXAGON2D
    g.set_hcf(L0, plane, 110e3, 0.34, 5e-5, 2.019, 0.5)
    o.set_hcf(L0, plane, 110e3, 0.34, 5e-5, 2.019, 0.5)
    for j in longer:
        g.set_hcf(1.7 * L0, plane, 110e3, 0.34, 5e-5, 2.019, 0.5, pid=int(j))
        o.set_hcf(1.7 * L0, plane, 110e3, 0.34, 5e-5, 2.019, 0.5, pid=int(j))
    p1, i1 = _gpu_lists(g, 1.5, 1)
    p0, i0 = _gpu_lists(g, 1.5, 0)
    o.search_nonlocal(1.5)
    po, io = o.geti("nlptr"), o.geti("nlidx")
    say(f"   entries {len(i1)}, longest list {np.diff(p1).max()}")
    np.testing.assert_array_equal(p1, po)
    np.testing.assert_array_equal(i1, io)
    np.testing.assert_array_equal(p0, po)
    np.testing.assert_array_equal(i0, io)
    # a second search on the same context replaces the lists
    p2, i2 = _gpu_lists(g, 1.1, 1)
    o.search_nonlocal(1.1)
    np.testing.assert_array_equal(p2, o.geti("nlptr"))
    np.testing.assert_array_equal(i2, o.geti("nlidx"))
    g.close()


def test_cell_search_keeps_overlapping_specimens_apart():
    say("[gpu] nonlocal search over an ensemble of overlapping specimens")
    parts = [hex_plate(20, 18)[0], hex_plate(14, 24)[0], hex_plate(26, 10)[0]]
    first = np.cumsum([0] + [len(x) for x in parts])
    union = []
    for x in parts:
        b_ptr, b_split, b_p2 = (np.asarray(a, dtype=np.int64) for a in G.build_bonds(G.HEXAGON2D, 0.5, x))
        union.append((b_ptr, b_split, b_p2))
    bp = [np.zeros(1, np.int64)]
    bs, p2 = [], []
    for (b_ptr, b_split, b_p2), f in zip(union, first[:-1]):
        off = int(bp[-1][-1])
        bp.append(b_ptr[1:] + off)
        bs.append(b_split + off)
        p2.append(b_p2 + f)
    bonds = tuple(np.concatenate(a).astype(np.int32) for a in (bp, bs, p2))
    g = G.Context(G.HEXAGON2D, 0.5, G.HCF, np.vstack(parts), bonds=bonds)
    g.set_specimens(first)
    g.set_hcf(3.0, True, 110e3, 0.34, 5e-5, 2.019, 0.5)
    p1, i1 = _gpu_lists(g, 1.5, 1)
    p0, i0 = _gpu_lists(g, 1.5, 0)
    np.testing.assert_array_equal(p1, p0)
    np.testing.assert_array_equal(i1, i0)
    for s, x in enumerate(parts):
        box = [x[:, 0].min() - 1, x[:, 0].max() + 1, x[:, 1].min() - 1, x[:, 1].max() + 1, -1.0, 1.0]
        o = O.System(O.HEXAGON2D, 0.5, O.HCF, x, box)
        o.set_hcf(3.0, True, 110e3, 0.34, 5e-5, 2.019, 0.5)
        o.search_nonlocal(1.5)
        po, io = o.geti("nlptr"), o.geti("nlidx")
        a, b = first[s], first[s + 1]
        np.testing.assert_array_equal(p1[a:b + 1] - p1[a], po)
        np.testing.assert_array_equal(i1[p1[a]:p1[b]], io + a)
    say(f"   {len(parts)} specimens, {len(i1)} entries, all within their specimen")
    g.close()


def test_long_lists_fall_back_to_all_pairs():
    xyz, box = sc_cube(22)
    say(f"[gpu] nonlocal search with lists above 4096 entries: {len(xyz)} particles")
    g = G.Context(G.SC3D, 0.5, G.HCF, xyz, bonds="device")
    o = O.System(O.SC3D, 0.5, O.HCF, xyz, box)
    g.set_hcf(6.8, False, 110e3, 0.34, 5e-5, 2.019, 0.5)
    o.set_hcf(6.8, False, 110e3, 0.34, 5e-5, 2.019, 0.5)
    p1, i1 = _gpu_lists(g, 1.5, 1)
    o.search_nonlocal(1.5)
    assert np.diff(p1).max() > 4096
    np.testing.assert_array_equal(p1, o.geti("nlptr"))
    np.testing.assert_array_equal(i1, o.geti("nlidx"))
    g.close()
