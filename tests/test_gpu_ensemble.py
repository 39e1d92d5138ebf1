"""Specimen ensembles: several independent particle systems of uneven size in one context
(one launch per kernel for all of them), each with its own Newton test. Every specimen must
reach the equilibrium the FP64 oracle computes for it alone; the specimens are placed on top of
each other in space, so any cross-talk (bonds, nonlocal neighbours, coarse aggregates) would
show."""
import sys

import numpy as np
import pytest

import oracle as O
from scenarios import G, hex_plate, rel, sc_cube

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _union(parts):
    """Concatenate specimens: positions, bond CSR with shifted ids, and particle offsets."""
    xyz, bp, bs, p2, first = [], [0], [], [], [0]
    for x, (b_ptr, b_split, b_p2) in parts:
        off = first[-1]
        boff = bp[-1]
        xyz.append(x)
        bp.extend(list(np.asarray(b_ptr[1:]) + boff))
        bs.extend(list(np.asarray(b_split) + boff))
        p2.extend(list(np.asarray(b_p2) + off))
        first.append(off + len(x))
    return (np.vstack(xyz), (np.array(bp, dtype=np.int32), np.array(bs, dtype=np.int32), np.array(p2, dtype=np.int32)),
            np.array(first, dtype=np.int64))


@pytest.mark.parametrize("precond,precision", [(0, 1), (2, 0)])
def test_hex_ensemble_matches_single_specimen_oracle(precond, precision):
    say("[gpu] ensemble hex precond", precond, "precision", precision)
    sizes = [(12, 10), (24, 16), (16, 30), (40, 40)]
    pulls = [0.001, 0.002, 0.0015, 0.001]
    parts, cases = [], []
    for (nx, ny), pull in zip(sizes, pulls):
        x, box = hex_plate(nx, ny)
        parts.append((x, G.build_bonds(G.HEXAGON2D, 0.5, x)))
        cases.append((x, box, pull))
    xyz, bonds, first = _union(parts)
    g = G.Context(G.HEXAGON2D, 0.5, G.ELASTIC, xyz, bonds=bonds)
    g.set_specimens(first)
    g.set_elastic(0.0, True, 69e3, 0.3, 1.0, elmap_literal=0)
    g.mode(dirichlet_static=1, cg_transpose=0, diag10_literal=0, precision=precision, precond=precond, coarse_cells=4,
           cg_tol=1e-12 if precision == 1 else 1e-5, tol_NR=1e-10, max_NR_iter=50, cg_maxit=50000)
    disp = []
    for s, (x, box, pull) in enumerate(cases):
        ids = np.arange(first[s], first[s + 1])
        y = x[:, 1]
        lo, hi = ids[y < y.min() + 1e-6], ids[y > y.max() - 1e-6]
        disp += [(lo, "x", 0.0), (lo, "y", 0.0), (hi, "y", pull)]
    ls = g.load_step(disp=disp)
    g.update_geometry()
    g.update_force_state()
    steps = 3
    for _ in range(steps):
        assert g.solve_step_static(ls, -1)
    its = g.specimen_stat("newton_iters")
    assert np.all(g.specimen_stat("converged") == 1)
    xg = g.get("xyz").reshape(-1, 3)
    for s, (x, box, pull) in enumerate(cases):
        o = O.System(O.HEXAGON2D, 0.5, O.ELASTIC, x, box)
        o.mode(elmap_literal=0, dirichlet_static=1, cg_tol=1e-12, tol_NR=1e-10, max_NR_iter=50)
        o.set_elastic(0.0, True, 69e3, 0.3, 1.0)
        y = x[:, 1]
        lo, hi = np.nonzero(y < y.min() + 1e-6)[0], np.nonzero(y > y.max() - 1e-6)[0]
        o.update_geometry()
        o.update_force_state()
        for _ in range(steps):
            assert o.solve_step_static(O.LoadStep().add_disp(lo, "x", 0.0).add_disp(lo, "y", 0.0).add_disp(hi, "y", pull), -1)
        u_o = o.getd("xyz").reshape(-1, 3) - x
        u_g = xg[first[s]:first[s + 1]] - x
        err = rel(u_g, u_o)
        say(f"   specimen {s} ({len(x)} particles): newton {its[s]:.0f} disp rel err {err:.3e}")
        assert err < 1e-8


@pytest.mark.parametrize("precision", [1, 0])
def test_hcf_ensemble_cycles_match_single_specimen_oracle(precision):
    """Three HCF specimens of different sizes, each with its own fatigue parameters (A, B, d) and
    force amplitude, through a fatigue static step and three load cycles (R = 0.1) in one
    context (lpm_solve_problem_cyclic_specimens). Each specimen's damage, HCF state and
    positions must match the oracle run on that specimen alone: FP64 CG (Jacobi) to 1e-8, and the
    bench's FP32 settings (two-level preconditioner, CG tolerance 1e-3 with forcing, the same FP64
    Newton test) to 1e-8 in positions and 1e-7 in damage and state."""
    say("[gpu] HCF ensemble cycles, precision", precision)
    cases = [((10, 16), (5.206e-5, 2.019, 0.5), 400.0), ((14, 12), (8e-5, 2.019, 0.5), 500.0),
             ((8, 20), (3e-5, 2.1, 0.45), 300.0)]
    parts, per = [], []
    for s, ((nx, ny), abd, fmax) in enumerate(cases):
        x, box = hex_plate(nx, ny)
        parts.append((x, G.build_bonds(G.HEXAGON2D, 0.5, x)))
        left = np.nonzero(x[:, 0] < x[:, 0].min() + 0.5)[0]
        y = x[:, 1]
        bottom, top = np.nonzero(y < y.min() + 1e-6)[0], np.nonzero(y > y.max() - 1e-6)[0]
        t = np.zeros(len(x), dtype=np.int32)
        t[top] = 4
        t[bottom] = 4
        per.append((x, box, abd, fmax, left, bottom, top, t))
    xyz, bonds, first = _union(parts)
    g = G.Context(G.HEXAGON2D, 0.5, G.HCF, xyz, bonds=bonds)
    g.set_specimens(first)
    g.set("type", np.concatenate([p[7] for p in per]))
    g.set_hcf(4.0, True, 110e3, 0.34, 5.206e-5, 2.019, 0.5, elmap_literal=0)
    mat = g.get("mat").reshape(-1, 10)  # per-specimen fatigue parameters (A, B, d) per particle
    for s, p in enumerate(per):
        mat[first[s]:first[s + 1], 0:3] = p[2]
    g.set("mat", mat)
    g.search_nonlocal(1.5)
    if precision == 1:
        g.mode(precision=1, cg_tol=1e-12, tol_NR=1e-10, max_NR_iter=60, cg_maxit=5000, cg_transpose=0, diag10_literal=0)
    else:
        g.mode(precision=0, precond=2, coarse_cells=4, cg_tol=1e-3, cg_forcing=1, tol_NR=1e-10, max_NR_iter=60,
               cg_maxit=50000, cg_transpose=0, diag10_literal=0, dirichlet_static=1)
    bar = 1e-8 if precision == 1 else 1e-7
    g.update_geometry()
    g.update_force_state()
    disp, force, groups = [], [], []
    for s, p in enumerate(per):
        off = first[s]
        left, bottom, top = p[4] + off, p[5] + off, p[6] + off
        disp += [(left, "x", 0.0), (bottom, "y", 0.0)]
        force += [(top, 0.0, 0.1 * p[3], 0.0)]
        groups.append([left, top, bottom])
    g.solve_problem_fatigue_static([g.load_step(disp=disp, force=force)])
    spec = [0.1] + [1.0, 0.1] * 3
    Ng = g.solve_problem_cyclic_specimens(G.LOAD_UNIAXIAL_FORCE, 1, spec, groups, [p[3] for p in per], 1.0, 4)
    assert Ng == 3
    Dg = g.get("damage")
    svg = g.get("state_var").reshape(-1, 7)[:, :3]
    xg = g.get("xyz").reshape(-1, 3)
    for s, (x, box, abd, fmax, left, bottom, top, t) in enumerate(per):
        o = O.System(O.HEXAGON2D, 0.5, O.HCF, x, box)
        o.mode(elmap_literal=0, cg_tol=1e-13, tol_NR=1e-10, max_NR_iter=60, cg_transpose=0, diag10_literal=0)
        o.seti("type", t)
        o.set_hcf(4.0, True, 110e3, 0.34, *abd)
        o.search_nonlocal(1.5)
        o.update_geometry()
        o.update_force_state()
        o.solve_problem_fatigue_static([O.LoadStep().add_disp(left, "x", 0.0).add_disp(bottom, "y", 0.0)
                                        .add_force(top, 0.0, 0.1 * fmax, 0.0)])
        No = o.solve_problem_cyclic(1, 1, [v * fmax for v in spec], [left, top, bottom], 1.0, 4)
        assert No == 3
        Do = o.getd("damage")
        sl = slice(first[s], first[s + 1])
        say(f"   specimen {s} ({len(x)} particles): damage max {Do.max():.4e} diff {np.abs(Dg[sl] - Do).max():.3e}")
        assert Do.max() > 0
        assert rel(Dg[sl], Do) < bar
        np.testing.assert_allclose(svg[sl], o.getd("state_var").reshape(-1, 7)[:, :3], rtol=bar, atol=1e-12)
        assert rel((xg[sl] - x).ravel(), o.getd("xyz") - x.ravel()) < 1e-8


def test_batched_adjoint_matches_single_specimen_adjoints():
    """One adjoint solve on an ensemble, with the group holding every specimen's top row, gives
    F = sum of the specimens' reactions; no bond joins two specimens, so the gradient on specimen
    s's bonds is d(F_s)/d(Kn, Tv) of that specimen alone. It must match the adjoint of the
    specimen's own context (itself checked against oracle finite differences elsewhere)."""
    say("[gpu] batched adjoint")
    sizes = [(12, 10), (20, 14), (9, 16)]
    pulls = [0.001, 0.0015, 0.002]
    parts, cases = [], []
    for (nx, ny), pull in zip(sizes, pulls):
        x, box = hex_plate(nx, ny)
        parts.append((x, G.build_bonds(G.HEXAGON2D, 0.5, x)))
        cases.append((x, pull))
    xyz, bonds, first = _union(parts)

    def run(ctx, specs):
        ctx.set_elastic(0.0, True, 69e3, 0.3, 1.0, elmap_literal=0)
        ctx.mode(dirichlet_static=1, cg_transpose=0, diag10_literal=0, precision=1, cg_tol=1e-12, tol_NR=1e-11,
                 max_NR_iter=50, cg_maxit=50000)
        disp, tops = [], []
        for off, x, pull in specs:
            y = x[:, 1]
            ids = np.arange(off, off + len(x))
            lo, hi = ids[y < y.min() + 1e-6], ids[y > y.max() - 1e-6]
            disp += [(lo, "x", 0.0), (lo, "y", 0.0), (hi, "y", pull)]
            tops.append(hi)
        ls = ctx.load_step(disp=disp)
        ctx.update_geometry()
        ctx.update_force_state()
        for _ in range(2):
            assert ctx.solve_step_static(ls, -1)
        return ctx.adjoint_reaction(ctx.group(np.concatenate(tops)), "y")

    g = G.Context(G.HEXAGON2D, 0.5, G.ELASTIC, xyz, bonds=bonds)
    g.set_specimens(first)
    F, gKn, gTv, nref, rres = run(g, [(first[s], x, pull) for s, (x, pull) in enumerate(cases)])
    say("   ensemble F", F, "refinements", nref, "rel residual", rres)
    assert rres < 1e-9
    bptr = g.get("bond_ptr")
    Fs = 0.0
    for s, (x, pull) in enumerate(cases):
        c1 = G.Context(G.HEXAGON2D, 0.5, G.ELASTIC, x)
        F1, k1, t1, _, r1 = run(c1, [(0, x, pull)])
        Fs += F1
        b0, b1 = bptr[first[s]], bptr[first[s + 1]]
        ek, et = rel(gKn[b0:b1], k1), rel(gTv[b0:b1], t1)
        say(f"   specimen {s}: F {F1:.10e} dKn rel err {ek:.2e} dTv rel err {et:.2e}")
        assert ek < 1e-8 and et < 1e-8
    assert F == pytest.approx(Fs, rel=1e-10)


def test_ensemble_predictor_reaches_the_same_equilibrium():
    """Load-step predictor (linear and quadratic) on a specimen ensemble with small strain increments (1e-4 to 3e-4 of
    each specimen's height per step, as in the ensemble benchmark): every specimen converges to
    Newton 1e-10, whose multiplier is taken at the reference's start, and ends within 1e-8 of the
    run without the predictor."""
    say("[gpu] ensemble with the load-step predictor")
    sizes, strains = [(40, 46), (70, 80), (90, 104)], [1e-4, 3e-4, 2e-4]
    parts, cases = [], []
    for (nx, ny), e in zip(sizes, strains):
        x, box = hex_plate(nx, ny)
        parts.append((x, G.build_bonds(G.HEXAGON2D, 0.5, x)))
        cases.append((x, e * (x[:, 1].max() - x[:, 1].min())))
    xyz, bonds, first = _union(parts)
    out = []
    for pred in (0, 1, 2):
        g = G.Context(G.HEXAGON2D, 0.5, G.ELASTIC, xyz, bonds=bonds)
        g.set_specimens(first)
        g.set_elastic(0.0, True, 110e3, 0.34, 1.0, elmap_literal=0)
        g.mode(dirichlet_static=1, cg_transpose=0, diag10_literal=0, precision=0, precond=2, cg_tol=1e-3, cg_forcing=1,
               tol_NR=1e-10, max_NR_iter=30, cg_maxit=50000, predictor=pred)
        disp = []
        for s, (x, pull) in enumerate(cases):
            ids = np.arange(first[s], first[s + 1])
            y = x[:, 1]
            lo, hi = ids[y < y.min() + 1e-6], ids[y > y.max() - 1e-6]
            disp += [(lo, "x", 0.0), (lo, "y", 0.0), (hi, "y", pull)]
        ls = g.load_step(disp=disp)
        g.update_geometry()
        g.update_force_state()
        for k in range(5):
            assert g.solve_step_static(ls, -1), (pred, k, g.specimen_stat("newton_iters"))
        say("   predictor", pred, "Newton per specimen (last step)", g.specimen_stat("newton_iters"))
        out.append(g.get("xyz").reshape(-1, 3) - xyz)
    for s in range(len(cases)):
        for p in (1, 2):
            err = rel(out[p][first[s]:first[s + 1]], out[0][first[s]:first[s + 1]])
            say(f"   specimen {s}, predictor {p}: displacement rel diff {err:.2e}")
            assert err < 1e-8
