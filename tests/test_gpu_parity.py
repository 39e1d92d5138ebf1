"""GPU parity of every phase of the load-step path against the FP64 oracle, through the
C ABI (lpmgpu -> liblpm_mi355x.so). Integer/topology outputs and FP64 phases that run the
reference's operation order are compared bit for bit; mixed-precision solves are compared
at converged equilibrium with tolerances written in each test."""
import sys

import numpy as np
import pytest

import oracle as O
from scenarios import G, Pair, faces, hex_plate, rel, sc_cube

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _banner():
    say("[gpu] parity tests start")
    yield


def deformed(xyz, seed, amp=2e-3, dim=3):
    rng = np.random.default_rng(seed)
    d = amp * rng.standard_normal(xyz.shape)
    if dim == 2:
        d[:, 2] = 0
    return xyz + d


# ----------------------------------------------------------------- topology
@pytest.mark.parametrize("case", ["sc", "hex", "fcc"])
def test_topology_bitwise(case):
    say("[gpu] topology", case)
    if case == "sc":
        xyz, box = sc_cube(6)
        lat, r = O.SC3D, 0.5
    elif case == "hex":
        xyz, box = hex_plate(14, 14)
        lat, r = O.HEXAGON2D, 0.5
    else:
        box = [0, 4, 0, 4, 0, 4]
        lat, r = O.FCC3D, 0.3
        xyz = O.generate(lat, box, r, 0, (0.3, 0.7, 1.1))
    p = Pair(lat, r, O.ELASTIC, xyz, box)
    for f_o, f_g in (("bptr", "bond_ptr"), ("bsplit", "bond_split"), ("p2", "bond_p2"), ("opp", "opp"),
                     ("cptr", "conn_ptr"), ("cidx", "conn_idx"), ("nb", "nb")):
        np.testing.assert_array_equal(p.g.get(f_g), p.o.geti(f_o), err_msg=f_o)
    np.testing.assert_array_equal(p.g.get("dis_initial"), p.o.getd("dis0"))
    np.testing.assert_array_equal(p.g.get("cs"), p.o.getd("cs"))


@pytest.mark.parametrize("case", ["sc", "hex", "square", "fcc_rotated", "jittered", "dogbone"])
def test_device_bond_search_bitwise(case):
    """Bonds searched on the device inside lpm_create (cell grid, the reference's distance test,
    lists sorted by partner id) equal the host search, itself pinned to the all-pairs loop: bond
    lists, layer split, reverse bonds, bond counts, the 2-hop pattern and the initial lengths and
    direction cosines, bit for bit."""
    say("[gpu] device bond search", case)
    rng = np.random.default_rng(3)
    if case == "sc":
        xyz, _ = sc_cube(20)
        lat, r = G.SC3D, 0.5
    elif case == "hex":
        xyz, _ = hex_plate(60, 50)
        lat, r = G.HEXAGON2D, 0.5
    elif case == "square":
        xyz = G.generate_lattice(G.SQUARE2D, [0, 40, 0, 30, 0, 1], 0.5)
        lat, r = G.SQUARE2D, 0.5
    elif case == "fcc_rotated":
        xyz = G.generate_lattice(G.FCC3D, [0, 8, 0, 8, 0, 8], 0.3, 0, (0.3, 0.7, 1.1))
        lat, r = G.FCC3D, 0.3
    elif case == "jittered":
        xyz, _ = sc_cube(16)
        xyz = xyz + 0.05 * rng.standard_normal(xyz.shape)
        lat, r = G.SC3D, 0.5
    else:
        x, _ = hex_plate(80, 160)
        y = (x[:, 1] - x[:, 1].min()) / np.ptp(x[:, 1])
        w = np.where(np.abs(y - 0.5) <= 0.15, 0.5, np.where(np.abs(y - 0.5) >= 0.25, 1.0,
                                                             0.5 + 0.5 * (np.abs(y - 0.5) - 0.15) / 0.10))
        xyz = np.ascontiguousarray(x[np.abs(x[:, 0] - 0.5 * np.ptp(x[:, 0])) <= 0.5 * w * np.ptp(x[:, 0]) + 1e-9])
        lat, r = G.HEXAGON2D, 0.5
    host = G.Context(lat, r, G.ELASTIC, xyz)
    dev = G.Context(lat, r, G.ELASTIC, xyz, bonds="device")
    say("   particles", len(xyz), "bonds", host.size("nbonds"), dev.size("nbonds"))
    for f in ("bond_ptr", "bond_split", "bond_p2", "opp", "nb", "conn_ptr", "conn_idx", "dis_initial", "cs"):
        np.testing.assert_array_equal(dev.get(f), host.get(f), err_msg=f)
    assert host.size("nbonds") > 0


@pytest.mark.parametrize("case", ["ct_square", "hex_notch", "image"])
def test_device_cut_search_matches_geometry_tool(case):
    """lpm_create_cut (device search that drops the bonds the reference's geometry tools drop) on
    specimens with holes and notches, after the tools' particle filter: bond lists, layer split
    and order equal the oracle's all-pairs searchNeighbor (geometry/geo1_CT_configuration.cpp)."""
    say("[gpu] device cut search", case)
    import os

    if case == "ct_square":
        xyz = G.generate_lattice(G.SQUARE2D, [0, 40, 0, 40, 0, 1], 0.5)
        lat, r = G.SQUARE2D, 0.5
        cuts = [(G.CUT_CIRCLE, 30.0, 10.2, 4.0), (G.CUT_CIRCLE, 30.0, 30.2, 4.0), (G.CUT_NOTCH, 18.3, 20.5, 0.3)]
    elif case == "hex_notch":
        xyz, _ = hex_plate(50, 40)
        lat, r = G.HEXAGON2D, 0.5
        cuts = [(G.CUT_CIRCLE, 12.0, 17.0, 3.5), (G.CUT_REVERSE_NOTCH, 35.2, 17.75, 0.35)]
    else:
        img = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "G6a_XZ_128.npy"))
        pixel = 0.0061
        xyz = G.image_particles(img, pixel)
        lat, r = G.SQUARE2D, pixel / 2
        cuts = [(G.CUT_CIRCLE, 64 * pixel, 30 * pixel, 12 * pixel), (G.CUT_NOTCH, 40 * pixel, 64 * pixel, 0.3 * pixel)]
    keep = G.cut_particles(xyz, cuts)
    np.testing.assert_array_equal(keep, O.cut_particles(xyz, cuts))
    xyz = np.ascontiguousarray(xyz[keep == 1])
    g = G.Context(lat, r, G.ELASTIC, xyz, bonds="device", cuts=cuts)
    lay, p1, p2 = O.search_neighbor_cut(lat, r, xyz, cuts)
    full = O.search_neighbor_cut(lat, r, xyz, [])
    say("   particles", len(xyz), "bonds", len(lay), "cut", len(full[0]) - len(lay))
    assert len(lay) < len(full[0])
    n = len(xyz)
    bp = np.zeros(n + 1, dtype=np.int64)
    np.add.at(bp, p1 + 1, 1)
    bp = np.cumsum(bp)
    order = np.lexsort((np.arange(len(p1)), lay, p1))  # per particle: layer 0 then 1, list order kept
    bs = bp[:-1] + np.bincount(p1[lay == 0], minlength=n)
    np.testing.assert_array_equal(g.get("bond_ptr"), bp)
    np.testing.assert_array_equal(g.get("bond_split"), bs)
    np.testing.assert_array_equal(g.get("bond_p2"), p2[order])


# ----------------------------------------------------------------- geometry, force, residual (FP64, bitwise)
@pytest.mark.parametrize("model", [O.ELASTIC, O.ELASTIC_DAMAGE])
def test_force_state_bitwise(model):
    say("[gpu] force state model", model)
    xyz, box = sc_cube(6)
    p = Pair(O.SC3D, 0.5, model, xyz, box)
    if model == O.ELASTIC:
        p.set_elastic(0.0, False, 200e3, 0.3, 1.0)
    else:
        p.set_elastic_damage(1.0, False, 200e3, 0.3, 1e-3, 1.0, 10.0, 0.99)
        dmg = np.linspace(0, 0.5, len(xyz))
        p.o.setd("damage", dmg)
        p.g.set("damage", dmg)
    x = deformed(xyz, 1)
    p.o.setd("xyz", x)
    p.g.set("xyz", x)
    bd = np.zeros(p.o.info("nbonds"))
    bd[::7] = 1.0
    bd[3::11] = 0.25
    p.o.setd("bdamage", bd)
    p.g.set("bdamage", bd)
    p.o.update_geometry()
    p.o.update_force_state()
    p.g.update_geometry()
    p.g.update_force_state()
    for f in ("dis", "dLe", "bstrain", "cs", "bforce", "bdamage"):
        np.testing.assert_array_equal(p.g.get(f), p.o.getd(f), err_msg=f)
    for f in ("dLe_total", "TdLe_total", "cs_sum", "Pin", "stress", "damage_visual"):
        np.testing.assert_array_equal(p.g.get(f), p.o.getd(f), err_msg=f)
    cons = np.zeros(3 * len(xyz), dtype=np.int32)
    cons[::5] = 1
    pex = np.linspace(-1, 1, 3 * len(xyz))
    p.o.seti("cons", cons)
    p.g.set("disp_constraint", cons)
    p.o.setd("Pex", pex)
    p.g.set("Pex", pex)
    p.o.update_rr()
    nr, nR = p.g.update_rr()
    np.testing.assert_array_equal(p.g.get("residual"), p.o.getd("residual"))
    np.testing.assert_array_equal(p.g.get("reaction"), p.o.getd("reaction"))
    assert nr == pytest.approx(p.o.norm("residual"), rel=1e-14)
    assert nR == pytest.approx(p.o.norm("reaction"), rel=1e-14)


# ----------------------------------------------------------------- stiffness assembly
@pytest.mark.parametrize("lat", [O.SC3D, O.HEXAGON2D])
@pytest.mark.parametrize("transpose,diag10", [(1, 1), (0, 0)])
def test_assembly_bitwise(lat, transpose, diag10):
    say("[gpu] assembly", lat, transpose, diag10)
    if lat == O.SC3D:
        xyz, box = sc_cube(5)
        dim = 3
    else:
        xyz, box = hex_plate(10, 10)
        dim = 2
    p = Pair(lat, 0.5, O.ELASTIC, xyz, box)
    p.set_elastic(0.0, False, 200e3, 0.3, 1.0)
    x = deformed(xyz, 2, dim=dim)
    bd = np.zeros(p.o.info("nbonds"))
    bd[::9] = 1.0
    bd[4::13] = 0.5
    for s in (p.o,):
        s.setd("xyz", x)
        s.setd("bdamage", bd)
        s.mode(cg_transpose=transpose, diag10_literal=diag10)
    p.g.set("xyz", x)
    p.g.set("bdamage", bd)
    p.g.mode(cg_transpose=transpose, diag10_literal=diag10)
    p.o.update_geometry()
    p.g.update_geometry()
    p.o.calc_stiffness()
    r, c, v = p.o.K_csr()
    import scipy.sparse as sp
    K = sp.csr_matrix((v, c, r), shape=(len(r) - 1, len(r) - 1))
    A = (K.T if transpose else K).tocsr()
    A.sort_indices()
    for plan in (1, 0):  # the planned kernel and the row-list kernel
        p.g.mode(asm_plan=plan)
        for prec in (1, 0):
            p.g.mode(precision=prec)
            p.g.calc_stiffness()
            gr, gc, gv = p.g.matrix()
            np.testing.assert_array_equal(gr, A.indptr)
            np.testing.assert_array_equal(gc, A.indices)
            exp = A.data if prec == 1 else A.data.astype(np.float32).astype(np.float64)
            np.testing.assert_array_equal(gv, exp)
    p.g.mode(asm_plan=1)  # its diagonal feeds the Dirichlet value ||diag K|| below
    # Dirichlet elimination (stiffness.h:265-294)
    cons = np.zeros(3 * len(xyz), dtype=np.int32)
    cons[[0, 1, 5, 9, 3 * 7 + 1]] = 1
    p.o.seti("cons", cons)
    p.g.set("disp_constraint", cons)
    p.o.update_force_state()
    p.g.update_force_state()
    p.o.update_stiffness_disp_bc()
    r, c, v = p.o.K_csr()
    K = sp.csr_matrix((v, c, r), shape=(len(r) - 1, len(r) - 1))
    A = (K.T if transpose else K).tocsr()
    A.sort_indices()
    p.g.mode(precision=1)
    p.g.calc_stiffness()
    p.g.update_stiffness_disp_bc()
    gr, gc, gv = p.g.matrix()
    np.testing.assert_array_equal(gc, A.indices)
    d = np.abs(gv - A.data)
    assert d.max() <= 1e-12 * np.abs(A.data).max(), d.max()


@pytest.mark.parametrize("lat,r,box", [(O.SQUARE2D, 0.5, [0, 9, 0, 9, 0, 0]), (O.HEXAGON2D, 0.5, [0, 9, 0, 9, 0, 0]),
                                       (O.SC3D, 0.5, [0, 5, 0, 5, 0, 5]), (O.FCC3D, 0.5, [0, 4, 0, 4, 0, 4]),
                                       (O.BCC3D, 0.5, [0, 4, 0, 4, 0, 4])])
def test_planned_assembly_matches_row_list_kernel(lat, r, box):
    """The planned assembly (blocks grouped by kind, one block row per lane) against the
    row-list kernel, which the test above pins to the oracle: every lattice, a deformed state
    with broken and partly damaged bonds, both stores and precisions, bit for bit."""
    import lpmgpu as G

    say("[gpu] planned assembly", lat)
    if lat == O.BCC3D:  # cubic cells of edge 4r/sqrt(3) with their body centres
        a = 4 * r / np.sqrt(3.0)
        c = np.stack(np.meshgrid(*[np.arange(4)] * 3, indexing="ij"), -1).reshape(-1, 3) * a
        xyz = np.concatenate([c, c + 0.5 * a])
    else:
        xyz = G.generate_lattice(lat, box, r)
    g = G.Context(lat, r, O.ELASTIC, xyz)
    g.set_elastic(0.0, lat in (O.SQUARE2D, O.HEXAGON2D), 200e3, 0.3, 1.0, elmap_literal=0)
    rng = np.random.default_rng(lat)
    x = xyz + 1e-2 * r * rng.standard_normal(xyz.shape)
    if lat in (O.SQUARE2D, O.HEXAGON2D):
        x[:, 2] = xyz[:, 2]
    bd = np.where(rng.random(g.size("nbonds")) < 0.1, 1.0, 0.0)
    bd[rng.random(len(bd)) < 0.1] = 0.4
    g.set("xyz", x)
    g.set("bdamage", bd)
    g.update_geometry()
    for transpose, diag10 in ((0, 0), (1, 1)):
        g.mode(cg_transpose=transpose, diag10_literal=diag10)
        for prec in (1, 0):
            g.mode(precision=prec)
            out = []
            for plan in (0, 1):
                g.mode(asm_plan=plan)
                g.calc_stiffness()
                out.append(g.matrix())
            # every launch shape of the planned kernel (the tuner's candidates) writes the same blocks
            g.mode(asm_tune=0)
            for w in (1, 2, 4):
                g.mode(asm_wpb=w)
                g.calc_stiffness()
                out.append(g.matrix())
            g.mode(asm_tune=1)
            for o in out[1:]:
                for a, b in zip(out[0], o):
                    np.testing.assert_array_equal(b, a)
    g.close()


# ----------------------------------------------------------------- CG and one Newton step
def _bar(model=O.ELASTIC, n=5):
    xyz, box = sc_cube(n)
    p = Pair(O.SC3D, 0.5, model, xyz, box)
    p.set_elastic(0.0, False, 200e3, 0.3, 1.0)
    bottom, top = faces(xyz, 2)
    return p, xyz, bottom, top


def test_cg_fp64_matches_dense_solution():
    say("[gpu] cg fp64")
    p, xyz, bottom, top = _bar()
    ls_o, ls_g = p.load(disp=[(bottom, "x", 0.0), (bottom, "y", 0.0), (bottom, "z", 0.0), (top, "z", 0.01)])
    p.o.apply_bc(ls_o)
    p.g.apply_bc(ls_g)
    p.o.mode(cg_transpose=0, diag10_literal=0, cg_tol=1e-14)
    p.g.mode(cg_transpose=0, diag10_literal=0, cg_tol=1e-14, precision=1, cg_maxit=5000)
    for s in (p.o, p.g):
        s.update_geometry()
        s.update_force_state()
        s.update_rr()
    p.o.calc_stiffness()
    p.o.update_stiffness_disp_bc()
    p.g.calc_stiffness()
    p.g.update_stiffness_disp_bc()
    K = p.o.K_dense()
    rhs = p.o.getd("residual")
    np.testing.assert_array_equal(p.g.get("residual"), rhs)
    it, err = p.g.solve_linear_system()
    d = p.g.get("disp")
    d_ref = np.linalg.solve(K, rhs)
    assert err <= 1e-14 and it > 0
    assert rel(d, d_ref) < 1e-11
    # positions moved by d (solver.h:97-105)
    x_exp = xyz.copy()
    x_exp[top, 2] += 0.01
    np.testing.assert_array_equal(p.g.get("xyz").reshape(-1, 3), x_exp + d.reshape(-1, 3))


@pytest.mark.parametrize("precision", [0, 1])
def test_static_step_equilibrium_parity(precision):
    """One static load step (solver_static.h:97-149) with Dirichlet elimination, both sides
    converged to a tight residual: displacements and reactions agree to 1e-8 relative."""
    say("[gpu] static step precision", precision)
    p, xyz, bottom, top = _bar(n=6)
    disp = [(bottom, "x", 0.0), (bottom, "y", 0.0), (bottom, "z", 0.0), (top, "z", 0.02)]
    ls_o, ls_g = p.load(disp=disp)
    p.o.mode(dirichlet_static=1, cg_tol=1e-13, tol_NR=1e-12, max_NR_iter=50)
    p.g.mode(dirichlet_static=1, cg_tol=1e-6 if precision == 0 else 1e-13, tol_NR=1e-12, max_NR_iter=50,
             precision=precision, cg_transpose=0, diag10_literal=0, cg_maxit=4000)
    for s in (p.o, p.g):
        s.update_geometry()
        s.update_force_state()
    assert p.o.solve_step_static(ls_o, -1)
    assert p.g.solve_step_static(ls_g, -1)
    u_o = p.o.getd("xyz") - xyz.ravel()
    u_g = p.g.get("xyz") - xyz.ravel()
    say("   disp rel err", rel(u_g, u_o), "newton", p.g.stat("newton_iters"), "cg", p.g.stat("cg_iters"))
    assert rel(u_g, u_o) < 1e-8
    assert rel(p.g.get("reaction"), p.o.getd("reaction")) < 1e-8
    # constrained particles did not move
    np.testing.assert_array_equal(p.g.get("xyz").reshape(-1, 3)[bottom], xyz[bottom])


_HEADLINE_ORACLE = {}


def _headline_oracle_trajectory():
    """The FP64 oracle's displacements and reactions after each of 5 load steps of the 12^3 bar
    (reference start, solved tightly); computed once and shared by the settings below."""
    if not _HEADLINE_ORACLE:
        p, xyz, bottom, top = _bar(n=12)
        L = xyz[:, 2].max() - xyz[:, 2].min()
        disp = [(bottom, "x", 0.0), (bottom, "y", 0.0), (bottom, "z", 0.0), (top, "z", 0.005 * L / 20)]
        p.o.mode(dirichlet_static=1, cg_tol=1e-13, tol_NR=1e-11, max_NR_iter=50)
        p.o.update_geometry()
        p.o.update_force_state()
        traj = []
        for _ in range(5):
            ls_o, _ = p.load(disp=disp)
            assert p.o.solve_step_static(ls_o, -1)
            traj.append((p.o.getd("xyz") - xyz.ravel(), p.o.getd("reaction")))
        _HEADLINE_ORACLE["traj"] = traj
    return _HEADLINE_ORACLE["traj"]


@pytest.mark.parametrize("forcing,precond,predictor", [(0, 0, 0), (1, 0, 0), (1, 2, 0), (1, 2, 1), (1, 0, 1), (1, 2, 2)])
def test_headline_settings_equilibrium_parity(forcing, precond, predictor):
    """The headline benchmark's solver settings (FP32 matrix/SpMV/CG with rtol 1e-3, optional
    inexact-Newton forcing, Jacobi or two-level preconditioner, optional linear or quadratic
    load-step predictor, Newton tol 1e-10) on a 12^3 SC bar in uniaxial tension, 5 load steps: displacements and
    reactions within 1e-8 of the FP64 oracle (reference start, no predictor) solved tightly."""
    say("[gpu] headline settings forcing", forcing, "precond", precond, "predictor", predictor)
    traj = _headline_oracle_trajectory()
    xyz, box = sc_cube(12)
    bottom, top = faces(xyz, 2)
    L = xyz[:, 2].max() - xyz[:, 2].min()
    g = G.Context(G.SC3D, 0.5, G.ELASTIC, xyz)
    g.set_elastic(0.0, False, 200e3, 0.3, 1.0, elmap_literal=0)
    g.mode(dirichlet_static=1, cg_tol=1e-3, cg_forcing=forcing, tol_NR=1e-10, max_NR_iter=30, precision=0,
           cg_transpose=0, diag10_literal=0, cg_maxit=20000, precond=precond, coarse_cells=4 if precond == 2 else 0,
           predictor=predictor)
    ls = g.load_step(disp=[(bottom, "x", 0.0), (bottom, "y", 0.0), (bottom, "z", 0.0), (top, "z", 0.005 * L / 20)])
    g.update_geometry()
    g.update_force_state()
    for k, (u_o, R_o) in enumerate(traj):
        g.reset_stats()
        assert g.solve_step_static(ls, -1)
        u_g = g.get("xyz") - xyz.ravel()
        say("   step", k, "disp rel err", rel(u_g, u_o), "reaction rel err", rel(g.get("reaction"), R_o),
            "newton", g.stat("newton_iters"), "cg", g.stat("cg_iters"))
        assert rel(u_g, u_o) < 1e-8
        assert rel(g.get("reaction"), R_o) < 1e-8


@pytest.mark.parametrize("lattice", ["sc", "hex"])
@pytest.mark.parametrize("precision", [0, 1])
def test_two_level_preconditioner(lattice, precision):
    """precond 2 (Jacobi + rigid-body aggregate coarse correction) reaches the oracle's
    equilibrium to 1e-8 and needs fewer CG iterations than Jacobi on the same step."""
    say("[gpu] two-level", lattice, "precision", precision)
    iters = {}
    for pc in (0, 2):
        if lattice == "sc":
            p, xyz, lo, hi = _bar(n=12)
            axes = "xyz"
        else:
            xyz, box = hex_plate(40, 40)
            p = Pair(O.HEXAGON2D, 0.5, O.ELASTIC, xyz, box)
            p.set_elastic(0.0, True, 69e3, 0.3, 1.0)
            lo, hi = faces(xyz, 1)
            axes = "xy"
        disp = [(lo, a, 0.0) for a in axes] + [(hi, axes[-1], 0.02)]
        ls_o, ls_g = p.load(disp=disp)
        p.o.mode(dirichlet_static=1, cg_tol=1e-13, tol_NR=1e-11, max_NR_iter=50)
        p.g.mode(dirichlet_static=1, cg_tol=1e-6 if precision == 0 else 1e-12, tol_NR=1e-11, max_NR_iter=50,
                 precision=precision, cg_transpose=0, diag10_literal=0, cg_maxit=20000, precond=pc, coarse_cells=4)
        for s in (p.o, p.g):
            s.update_geometry()
            s.update_force_state()
        assert p.o.solve_step_static(ls_o, -1)
        assert p.g.solve_step_static(ls_g, -1)
        u_o = p.o.getd("xyz") - xyz.ravel()
        u_g = p.g.get("xyz") - xyz.ravel()
        iters[pc] = p.g.stat("cg_iters")
        say("   precond", pc, "ncoarse", p.g.get_mode("ncoarse"), "disp rel err", rel(u_g, u_o), "cg", iters[pc])
        assert rel(u_g, u_o) < 1e-8
        assert rel(p.g.get("reaction"), p.o.getd("reaction")) < 1e-8
    assert iters[2] < iters[0]


_BAR13 = {}


def _bar13_oracle():
    """The oracle's equilibrium of the SC 13^3 bar step below (computed once for the module)."""
    if not _BAR13:
        p, xyz, lo, hi = _bar(n=13)
        ls_o, _ = p.load(disp=[(lo, a, 0.0) for a in "xyz"] + [(hi, "z", 0.02)])
        p.o.mode(dirichlet_static=1, cg_tol=1e-13, tol_NR=1e-11, max_NR_iter=50)
        p.o.update_geometry()
        p.o.update_force_state()
        assert p.o.solve_step_static(ls_o, -1)
        _BAR13["u"] = p.o.getd("xyz") - xyz.ravel()
    return _BAR13["u"]


@pytest.mark.parametrize("cells", [3, 7, 12])
def test_two_level_restriction_workgroup_sizes(cells):
    """The restriction's workgroup (64, 256 or 1024 threads, the fewest that cover the largest
    aggregate in one round of 4 members per thread) on SC 13^3 with aggregates of 3, 7 and 12
    spacings (up to 64, 512 and 2197 members): the equilibrium matches the oracle to 1e-8 in
    FP32 and FP64, and the coarse correction cuts the CG iterations against Jacobi."""
    say("[gpu] two-level restriction, aggregate edge", cells)
    u_o = _bar13_oracle()
    for precision in (0, 1):
        iters = {}
        for pc in (0, 2):
            p, xyz, lo, hi = _bar(n=13)
            _, ls_g = p.load(disp=[(lo, a, 0.0) for a in "xyz"] + [(hi, "z", 0.02)])
            p.g.mode(dirichlet_static=1, cg_tol=1e-6 if precision == 0 else 1e-12, tol_NR=1e-11, max_NR_iter=50,
                     precision=precision, cg_transpose=0, diag10_literal=0, cg_maxit=20000, precond=pc,
                     coarse_cells=cells)
            p.g.update_geometry()
            p.g.update_force_state()
            assert p.g.solve_step_static(ls_g, -1)
            u_g = p.g.get("xyz") - xyz.ravel()
            iters[pc] = p.g.stat("cg_iters")
            say(f"   precision {precision} precond {pc} ncoarse {p.g.get_mode('ncoarse')} "
                f"disp rel err {rel(u_g, u_o):.2e} cg {iters[pc]}")
            assert rel(u_g, u_o) < 1e-8
        assert iters[2] < iters[0]


@pytest.mark.parametrize("cells", [1, 2])
def test_two_level_dependent_coarse_modes(cells):
    """Aggregates of one or two lattice spacings on a hex plate whose top row has its y DoF
    prescribed: single-particle aggregates have a zero rotation mode and top-row aggregates a
    rotation mode that is a multiple of their x translation once y is masked. The inversion must
    drop those modes: the CG still converges in fewer iterations than Jacobi and reaches the
    oracle's equilibrium."""
    say("[gpu] two-level dependent modes, coarse_cells", cells)
    xyz, box = hex_plate(12, 10)
    iters = {}
    for pc in (0, 2):
        p = Pair(O.HEXAGON2D, 0.5, O.ELASTIC, xyz, box)
        p.set_elastic(0.0, True, 69e3, 0.3, 1.0)
        lo, hi = faces(xyz, 1)
        ls_o, ls_g = p.load(disp=[(lo, "x", 0.0), (lo, "y", 0.0), (hi, "y", 0.005)])
        p.o.mode(dirichlet_static=1, cg_tol=1e-13, tol_NR=1e-11, max_NR_iter=50)
        p.g.mode(dirichlet_static=1, cg_tol=1e-12, tol_NR=1e-11, max_NR_iter=50, precision=1, cg_transpose=0,
                 diag10_literal=0, cg_maxit=5000, precond=pc, coarse_cells=cells)
        for s in (p.o, p.g):
            s.update_geometry()
            s.update_force_state()
        assert p.o.solve_step_static(ls_o, -1)
        assert p.g.solve_step_static(ls_g, -1)
        iters[pc] = p.g.stat("cg_iters")
        err = rel(p.g.get("xyz") - xyz.ravel(), p.o.getd("xyz") - xyz.ravel())
        say("   precond", pc, "ncoarse", p.g.get_mode("ncoarse"), "disp rel err", err, "cg", iters[pc])
        assert err < 1e-8
    assert iters[2] < iters[0]


def test_two_level_auto_aggregates_hex_122():
    """A 122 x 122 hex plate (14,884 particles) with automatically sized aggregates, FP32 CG:
    every Newton iteration's CG converges in a few hundred iterations (an aggregate one lattice
    row thick at the top edge once made the coarse matrix singular here)."""
    say("[gpu] two-level auto aggregates hex 122x122")
    xyz, box = hex_plate(122, 122)
    y = xyz[:, 1]
    lo, hi = np.nonzero(y < y.min() + 1e-6)[0], np.nonzero(y > y.max() - 1e-6)[0]
    g = G.Context(G.HEXAGON2D, 0.5, G.ELASTIC, xyz)
    g.set_elastic(0.0, True, 110e3, 0.34, 1.0, elmap_literal=0)
    g.mode(dirichlet_static=1, cg_transpose=0, diag10_literal=0, precision=0, precond=2, cg_tol=1e-3, cg_forcing=1,
           tol_NR=1e-10, max_NR_iter=30, cg_maxit=5000)
    ls = g.load_step(disp=[(lo, "x", 0.0), (lo, "y", 0.0), (hi, "y", 2e-4 * (y.max() - y.min()))])
    g.update_geometry()
    g.update_force_state()
    for _ in range(2):
        g.reset_stats()
        assert g.solve_step_static(ls, -1)
        say("   newton", g.stat("newton_iters"), "cg", g.stat("cg_iters"), "ncoarse", g.get_mode("ncoarse"))
        assert g.stat("newton_iters") <= 6
        assert g.stat("cg_iters") < 1000


def test_literal_modes_reach_the_same_equilibrium():
    """Oracle in the reference's literal CG semantics (K^T, Q2) and GPU with a symmetric K both
    converge to the same equilibrium: the fixed point does not depend on K."""
    say("[gpu] literal modes")
    xyz, box = hex_plate(12, 10)
    p = Pair(O.HEXAGON2D, 0.5, O.ELASTIC, xyz, box)
    p.set_elastic(0.0, False, 69e3, 0.3, 1.0)
    bottom, top = faces(xyz, 1)
    ls_o, ls_g = p.load(disp=[(bottom, "x", 0.0), (bottom, "y", 0.0), (top, "y", 0.01)])
    p.o.mode(dirichlet_static=1, cg_tol=1e-13, cg_maxit=20000, tol_NR=1e-11, max_NR_iter=60)
    p.g.mode(dirichlet_static=1, cg_tol=1e-6, tol_NR=1e-11, max_NR_iter=60, cg_maxit=3000)
    for s in (p.o, p.g):
        s.update_geometry()
        s.update_force_state()
    assert p.o.solve_step_static(ls_o, -1)
    assert p.g.solve_step_static(ls_g, -1)
    u_o = p.o.getd("xyz") - xyz.ravel()
    u_g = p.g.get("xyz") - xyz.ravel()
    say("   disp rel err", rel(u_g, u_o))
    assert rel(u_g, u_o) < 1e-8


# ----------------------------------------------------------------- damage and state
def test_bond_breaking_flags_bitwise():
    say("[gpu] bond breaking")
    xyz, box = hex_plate(12, 12)
    p = Pair(O.HEXAGON2D, 0.5, O.ELASTIC, xyz, box)
    p.set_elastic(0.0, True, 69e3, 0.3, 3e-3)
    bottom, top = faces(xyz, 1)
    ls_o, ls_g = p.load(disp=[(bottom, "x", 0.0), (bottom, "y", 0.0), (top, "y", 0.05)])
    # symmetric K on both sides: after bonds break the literal K^T / Q2 CG of the oracle stalls
    p.o.mode(dirichlet_static=1, cg_tol=1e-13, tol_NR=1e-11, max_NR_iter=60, cg_transpose=0, diag10_literal=0)
    p.g.mode(dirichlet_static=1, cg_tol=1e-13, tol_NR=1e-11, max_NR_iter=60, cg_maxit=20000, cg_transpose=0,
             diag10_literal=0, precision=1)
    for s in (p.o, p.g):
        s.update_geometry()
        s.update_force_state()
    ok_o = p.o.solve_step_static(ls_o, -1)
    ok_g = p.g.solve_step_static(ls_g, -1)
    assert ok_o == ok_g
    bo, bg = p.o.getd("bdamage"), p.g.get("bdamage")
    say("   broken", int((bo == 1).sum()), int((bg == 1).sum()))
    assert (bo == 1).sum() > 0
    np.testing.assert_array_equal(bg == 1.0, bo == 1.0)
    assert rel(p.g.get("xyz") - xyz.ravel(), p.o.getd("xyz") - xyz.ravel()) < 1e-8


def test_elastic_damage_nonlocal_step():
    say("[gpu] elastic damage")
    xyz, box = hex_plate(12, 12)
    p = Pair(O.HEXAGON2D, 0.5, O.ELASTIC_DAMAGE, xyz, box)
    p.set_elastic_damage(1.2, True, 3.2e3, 0.28, 0.05, 1.0, 10.0, 0.99)
    bottom, top = faces(xyz, 1)
    t = np.zeros(len(xyz), dtype=np.int32)
    t[bottom] = 4
    p.set_type(t)
    p.o.search_nonlocal(1.5)
    p.g.search_nonlocal(1.5)
    np.testing.assert_array_equal(p.g.get("nl_ptr"), p.o.geti("nlptr"))
    np.testing.assert_array_equal(p.g.get("nl_idx"), p.o.geti("nlidx"))
    ls_o, ls_g = p.load(disp=[(bottom, "x", 0.0), (bottom, "y", 0.0), (top, "y", 0.2)])
    p.o.mode(dirichlet_static=1, cg_tol=1e-13, tol_NR=1e-11, max_NR_iter=80, cg_transpose=0, diag10_literal=0)
    p.g.mode(dirichlet_static=1, cg_tol=1e-13, tol_NR=1e-11, max_NR_iter=80, cg_maxit=20000, cg_transpose=0,
             diag10_literal=0, precision=1)
    for s in (p.o, p.g):
        s.update_geometry()
        s.update_force_state()
        s.update_state_var()
    for k in range(2):
        ok_o = p.o.solve_step_static(ls_o, 4)
        ok_g = p.g.solve_step_static(ls_g, 4)
        assert ok_o == ok_g
    Do, Dg = p.o.getd("damage"), p.g.get("damage")
    say("   damage max", Do.max(), "diff", np.abs(Dg - Do).max())
    assert Do.max() > 0
    assert rel(Dg, Do) < 1e-8
    np.testing.assert_array_equal(Dg == 1.0, Do == 1.0)
    assert rel(p.g.get("xyz") - xyz.ravel(), p.o.getd("xyz") - xyz.ravel()) < 1e-8


def test_j2_state_update():
    say("[gpu] j2")
    xyz, box = sc_cube(5)
    p = Pair(O.SC3D, 0.5, O.J2, xyz, box)
    p.set_j2(0.0, False, 70e3, 0.33, 50.0, 0.3, 2e3, 0.0, 1.0, 1.0)
    bottom, top = faces(xyz, 2)
    ls_o, ls_g = p.load(disp=[(bottom, "x", 0.0), (bottom, "y", 0.0), (bottom, "z", 0.0), (top, "z", 0.01)])
    p.o.mode(dirichlet_static=1, cg_tol=1e-13, tol_NR=1e-11, max_NR_iter=80, cg_transpose=0, diag10_literal=0)
    p.g.mode(dirichlet_static=1, cg_tol=1e-13, tol_NR=1e-11, max_NR_iter=80, cg_maxit=20000, cg_transpose=0,
             diag10_literal=0, precision=1)
    for s in (p.o, p.g):
        s.update_geometry()
        s.update_force_state()
    for k in range(3):
        assert p.o.solve_step_static(ls_o, -1) == p.g.solve_step_static(ls_g, -1)
    svo, svg = p.o.getd("state_var"), p.g.get("state_var")
    say("   alpha max", svo.reshape(-1, 7)[:, 0].max(), "diff", np.abs(svg - svo).max())
    assert svo.reshape(-1, 7)[:, 0].max() > 0
    assert rel(svg, svo) < 1e-8
    assert rel(p.g.get("dLp"), p.o.getd("dLp")) < 1e-8
    assert rel(p.g.get("xyz") - xyz.ravel(), p.o.getd("xyz") - xyz.ravel()) < 1e-8


def test_hcf_fatigue_cycles():
    say("[gpu] hcf fatigue")
    xyz, box = hex_plate(10, 16)
    p = Pair(O.HEXAGON2D, 0.5, O.HCF, xyz, box)
    p.set_hcf(4.0, True, 110e3, 0.34, 5.206e-5, 2.019, 0.5)
    left = np.nonzero(xyz[:, 0] < 0.5)[0]
    bottom, top = faces(xyz, 1)
    t = np.zeros(len(xyz), dtype=np.int32)
    t[top] = 4
    t[bottom] = 4
    p.set_type(t)
    p.o.search_nonlocal(1.5)
    p.g.search_nonlocal(1.5)
    p.o.mode(cg_tol=1e-13, tol_NR=1e-10, max_NR_iter=60, cg_transpose=0, diag10_literal=0)
    p.g.mode(cg_tol=1e-13, tol_NR=1e-10, max_NR_iter=60, cg_maxit=20000, cg_transpose=0, diag10_literal=0, precision=1)
    for s in (p.o, p.g):
        s.update_geometry()
        s.update_force_state()
    fmax = 400.0
    ls_o, ls_g = p.load(disp=[(left, "x", 0.0), (bottom, "y", 0.0)], force=[(top, 0.0, 0.1 * fmax, 0.0)])
    p.o.solve_problem_fatigue_static([ls_o])
    p.g.solve_problem_fatigue_static([ls_g])
    spec = [0.1 * fmax] + [fmax, 0.1 * fmax] * 3
    No = p.o.solve_problem_cyclic(1, 1, spec, [left, top, bottom], 1.0, 4)
    Ng = p.g.solve_problem_cyclic(G.LOAD_UNIAXIAL_FORCE, 1, spec, [left, top, bottom], 1.0, 4)
    assert No == Ng == 3
    Do, Dg = p.o.getd("damage"), p.g.get("damage")
    say("   damage max", Do.max(), "diff", np.abs(Dg - Do).max())
    assert Do.max() > 0
    assert rel(Dg, Do) < 1e-8
    svo, svg = p.o.getd("state_var").reshape(-1, 7)[:, :3], p.g.get("state_var").reshape(-1, 7)[:, :3]
    np.testing.assert_allclose(svg, svo, rtol=1e-8, atol=1e-12)


def test_step_halving_path():
    """max_NR_iter reached -> reset positions/state, halve the step, re-insert it
    (solver_static.h:78-87); Pex is not rolled back (literal)."""
    say("[gpu] halving")
    p, xyz, bottom, top = _bar(n=4)
    ls_o, ls_g = p.load(disp=[(bottom, "x", 0.0), (bottom, "y", 0.0), (bottom, "z", 0.0)],
                        force=[(top, 0.0, 0.0, 20.0)])
    p.o.mode(dirichlet_static=1, cg_tol=1e-13, tol_NR=1e-9, max_NR_iter=2)
    p.g.mode(dirichlet_static=1, cg_tol=1e-6, tol_NR=1e-9, max_NR_iter=2, cg_maxit=2000)
    for s in (p.o, p.g):
        s.update_geometry()
        s.update_force_state()
    ok = p.g.solve_step_static(ls_g, -1)
    assert not ok
    p.g.reset(True)
    np.testing.assert_array_equal(p.g.get("xyz"), xyz.ravel())
    assert p.g.get("Pex").reshape(-1, 3)[top, 2].sum() == pytest.approx(20.0)
