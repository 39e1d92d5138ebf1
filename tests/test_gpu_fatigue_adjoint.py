"""Sensitivity of a displacement objective to the HCF law's parameters (A, B, d per particle)
through one fatigue damage update, on the GPU, against central differences of the FP64 oracle
running the same sequence: a loading and an unloading fatigue step, each followed by the state
update (the cycle's energy extremes), the fatigue damage update with a fixed cycle step, and the
reloading step whose equilibrium the objective reads."""
import sys

import numpy as np
import pytest

import oracle as O
from scenarios import G, faces, hex_plate

pytestmark = pytest.mark.gpu

A0, B0, D0 = 5.206e-5, 2.019, 0.5
DT = 2e8  # cycles of the update: damage up to about 0.1, none clamped
FMAX = 400.0


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


def _oracle(xyz, box, left, bottom, top, t, A, dB=0.0, dd=0.0):
    """Top-edge y displacement sum after the sequence, per-particle A (array)."""
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

    assert s.solve_step_fatigue(mk(FMAX))
    s.update_state_var()
    assert s.solve_step_fatigue(mk(-0.9 * FMAX))
    s.update_state_var()
    s.update_fatigue_damage(DT, 4)
    D = s.getd("damage")
    assert s.solve_step_fatigue(mk(0.9 * FMAX))
    return (s.getd("xyz").reshape(-1, 3) - xyz)[top, 1].sum(), D


def test_fatigue_parameter_sensitivity_matches_central_differences():
    say("[gpu] HCF parameter sensitivity through one fatigue update")
    xyz, box, left, bottom, top, t = _setup()
    n = len(xyz)
    g = G.Context(G.HEXAGON2D, 0.5, G.HCF, xyz, bonds="device")
    g.set_hcf(4.0, True, 110e3, 0.34, A0, B0, D0, elmap_literal=0)
    g.set("type", t)
    g.search_nonlocal(1.5)
    g.mode(cg_tol=1e-13, tol_NR=1e-10, max_NR_iter=60, cg_maxit=20000, precision=1, cg_transpose=0, diag10_literal=0)
    g.update_geometry()
    g.update_force_state()

    def step(f):
        return g.load_step(disp=[(left, "x", 0.0), (bottom, "y", 0.0)], force=[(top, 0.0, f, 0.0)])

    ls = [step(FMAX), step(-0.9 * FMAX), step(0.9 * FMAX)]
    assert g.solve_step_fatigue(ls[0])
    g.update_state_var()
    assert g.solve_step_fatigue(ls[1])
    g.update_state_var()
    g.update_fatigue_damage(DT, 4)
    Dg = g.get("damage")
    assert g.solve_step_fatigue(ls[2])
    w = np.zeros(2 * n)
    w[2 * top + 1] = 1.0
    Gv, gA, gB, gd, nref, rres = g.adjoint_fatigue_params(w_x=w)
    say(f"   G {Gv:.12e} refinements {nref} rel residual {rres:.1e}")
    assert rres < 1e-9
    A = np.full(n, A0)
    G0, Do = _oracle(xyz, box, left, bottom, top, t, A)
    say("   damage max", Do.max(), "GPU vs oracle", np.abs(Dg - Do).max())
    np.testing.assert_allclose(Dg, Do, rtol=1e-7, atol=1e-15)
    assert Do.max() > 0.05 and Do.max() < 0.5
    assert Gv - xyz[top, 1].sum() == pytest.approx(G0, rel=1e-8)
    assert np.all(gA[t == 4] == 0)
    rng = np.random.default_rng(3)
    r = rng.uniform(-1, 1, n)
    h, hb = 1e-3, 1e-4
    cases = {
        "A scaled": (np.sum(gA * A), dict(A=A * (1 + h)), dict(A=A * (1 - h)), h),
        "A along a random direction": (np.sum(gA * A * r), dict(A=A * (1 + h * r)), dict(A=A * (1 - h * r)), h),
        "B shifted": (np.sum(gB), dict(A=A, dB=hb), dict(A=A, dB=-hb), hb),
        "d shifted": (np.sum(gd), dict(A=A, dd=hb), dict(A=A, dd=-hb), hb),
    }
    for name, (adj, up, dn, hh) in cases.items():
        fd = (_oracle(xyz, box, left, bottom, top, t, **up)[0] - _oracle(xyz, box, left, bottom, top, t, **dn)[0]) / (2 * hh)
        err = abs(adj - fd) / abs(fd)
        say(f"   {name}: adjoint {adj:.10e} fd {fd:.10e} rel {err:.2e}")
        assert err < 1e-5


def test_batched_fatigue_parameter_sensitivity_matches_single_specimens():
    """Three HCF specimens of different sizes, fatigue parameters and force amplitudes in one
    context, through the same loading, unloading, fatigue update and reloading sequence. One
    sensitivity call with the objective summed over every specimen's top edge gives each
    specimen's dG_s/d(A, B, d): no bond or nonlocal neighbour joins two specimens. Each slice must
    match the sensitivity of the specimen's own context (checked against oracle finite
    differences above)."""
    from test_gpu_ensemble import _union

    say("[gpu] batched HCF parameter sensitivity, 3 specimens")
    cases = [((10, 16), (A0, B0, D0), FMAX), ((14, 12), (8e-5, 2.019, 0.5), 500.0), ((8, 20), (3e-5, 2.1, 0.45), 300.0)]
    per, parts = [], []
    for (nx, ny), abd, fmax in cases:
        x, box = hex_plate(nx, ny)
        parts.append((x, G.build_bonds(G.HEXAGON2D, 0.5, x)))
        left = np.nonzero(x[:, 0] < x[:, 0].min() + 0.5)[0]
        bottom, top = faces(x, 1)
        t = np.zeros(len(x), dtype=np.int32)
        t[top] = 4
        t[bottom] = 4
        per.append((x, abd, fmax, left, bottom, top, t))

    def run(ctx, specs, abds):
        ctx.set("type", np.concatenate([p[6] for p in specs]))
        ctx.set_hcf(4.0, True, 110e3, 0.34, A0, B0, D0, elmap_literal=0)
        mat = ctx.get("mat").reshape(-1, 10)
        off = 0
        for p, abd in zip(specs, abds):
            mat[off:off + len(p[0]), 0:3] = abd
            off += len(p[0])
        ctx.set("mat", mat)
        ctx.search_nonlocal(1.5)
        ctx.mode(cg_tol=1e-13, tol_NR=1e-10, max_NR_iter=60, cg_maxit=20000, precision=1, cg_transpose=0,
                 diag10_literal=0)
        ctx.update_geometry()
        ctx.update_force_state()
        w = np.zeros(2 * off)

        def step(scale):
            disp, force, o = [], [], 0
            for x, abd, fmax, left, bottom, top, t in specs:
                disp += [(left + o, "x", 0.0), (bottom + o, "y", 0.0)]
                force += [(top + o, 0.0, scale * fmax, 0.0)]
                o += len(x)
            return ctx.load_step(disp=disp, force=force)

        o = 0
        for p in specs:
            w[2 * (p[5] + o) + 1] = 1.0
            o += len(p[0])
        ls = [step(1.0), step(-0.9), step(0.9)]
        assert ctx.solve_step_fatigue(ls[0])
        ctx.update_state_var()
        assert ctx.solve_step_fatigue(ls[1])
        ctx.update_state_var()
        ctx.update_fatigue_damage(DT, 4)
        assert ctx.solve_step_fatigue(ls[2])
        return ctx.adjoint_fatigue_params(w_x=w)

    xyz, bonds, first = _union(parts)
    g = G.Context(G.HEXAGON2D, 0.5, G.HCF, xyz, bonds=bonds)
    g.set_specimens(first)
    Gv, gA, gB, gd, nref, rres = run(g, per, [p[1] for p in per])
    say(f"   ensemble G {Gv:.12e} refinements {nref} rel residual {rres:.1e}")
    assert rres < 1e-9
    Gs = 0.0
    for s, p in enumerate(per):
        c1 = G.Context(G.HEXAGON2D, 0.5, G.HCF, p[0], bonds="device")
        G1, a1, b1, d1, _, r1 = run(c1, [p], [p[1]])
        assert r1 < 1e-9
        Gs += G1
        sl = slice(first[s], first[s + 1])
        errs = [np.linalg.norm(u[sl] - v) / np.linalg.norm(v) for u, v in ((gA, a1), (gB, b1), (gd, d1))]
        say(f"   specimen {s} ({len(p[0])} particles): G {G1:.10e} rel err A {errs[0]:.1e} B {errs[1]:.1e} "
            f"d {errs[2]:.1e}")
        assert np.linalg.norm(a1) > 0
        assert max(errs) < 1e-7
    assert Gv == pytest.approx(Gs, rel=1e-10)
