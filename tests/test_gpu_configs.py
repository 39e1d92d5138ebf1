"""Benchmark configurations other than the headline, at the sizes their parity checks allow.

Config 1: 2D hexagonal plate of 100 rows of 100 particles (10,000 particles from the restated
generator with this box), 1 mm lattice constant, E = 69 GPa,
nu = 0.3, plane stress, bottom edge clamped, top edge pulled +0.001 mm per step for 10 steps,
FP64 on the GPU against the FP64 oracle (converged-equilibrium protocol: both sides run Newton
to 1e-10 instead of the configuration's 1e-6, so that the comparison measures the
implementation, not the stopping point).

Config 4 (reduced to 40 x 40): a hex plate with an edge crack of a quarter of the width at
mid-height (bonds whose segment crosses the crack line removed, as the reference's crack
geometry does, geometry/geo1_CT_configuration.cpp:95-108), bond breaking at a critical stretch
of 1e-3, mode-I opening over several steps: broken-bond flags bit-exact and displacements within
1e-8 of the oracle; then the adjoint d(top reaction)/d(Kn, Tv) at the final state (frozen
broken-bond set) against central differences of the oracle."""
import sys

import numpy as np
import pytest

import oracle as O
from scenarios import G, hex_plate, rel

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _csr_to_list(bp, bs, p2):
    bp, bs = np.asarray(bp, dtype=np.int64), np.asarray(bs, dtype=np.int64)
    p1 = np.repeat(np.arange(len(bp) - 1), np.diff(bp))
    lay = (np.arange(bp[-1]) >= bs[p1]).astype(np.int64)
    return lay, p1, np.asarray(p2)


@pytest.mark.parametrize("precond,precision", [(0, 1), (2, 1), (2, 0)])
def test_config1_hex_plate(precond, precision):
    say("[gpu] config 1 hex 100x100 precond", precond, "precision", precision)
    xyz, box = hex_plate(100, 100)
    assert len(xyz) == 10000
    y = xyz[:, 1]
    lo = np.nonzero(y < y.min() + 1.0)[0]  # bottom edge: y < ymin + 2r
    hi = np.nonzero(y > y.max() - 1e-6)[0]
    o = O.System(O.HEXAGON2D, 0.5, O.ELASTIC, xyz, box)
    o.mode(elmap_literal=0, dirichlet_static=1, cg_tol=1e-11, tol_NR=1e-10, max_NR_iter=50)
    o.set_elastic(0.0, True, 69e3, 0.3, 1.0)
    g = G.Context(G.HEXAGON2D, 0.5, G.ELASTIC, xyz)
    g.set_elastic(0.0, True, 69e3, 0.3, 1.0, elmap_literal=0)
    g.mode(dirichlet_static=1, cg_transpose=0, diag10_literal=0, precision=precision, precond=precond,
           cg_tol=1e-11 if precision == 1 else 1e-5, tol_NR=1e-10, max_NR_iter=50, cg_maxit=50000)
    ls_g = g.load_step(disp=[(lo, "x", 0.0), (lo, "y", 0.0), (hi, "y", 0.001)])
    for s in (o, g):
        s.update_geometry()
        s.update_force_state()
    for k in range(10):
        ls_o = O.LoadStep().add_disp(lo, "x", 0.0).add_disp(lo, "y", 0.0).add_disp(hi, "y", 0.001)
        assert o.solve_step_static(ls_o, -1)
        assert g.solve_step_static(ls_g, -1)
    u_o = o.getd("xyz") - xyz.ravel()
    u_g = g.get("xyz") - xyz.ravel()
    say("   disp rel err", rel(u_g, u_o), "reaction rel err", rel(g.get("reaction"), o.getd("reaction")),
        "cg", g.stat("cg_iters"), "newton", g.stat("newton_iters"))
    assert rel(u_g, u_o) < 1e-8
    assert rel(g.get("reaction"), o.getd("reaction")) < 1e-8


INC = 0.0015  # opening per step at each edge: the crack starts at step 4 and grows stably at step 5


def _cracked_plate(n=40):
    xyz, box = hex_plate(n, n)
    bp, bs, p2 = G.build_bonds(G.HEXAGON2D, 0.5, xyz)
    lay, p1, pp = _csr_to_list(bp, bs, p2)
    y0 = 0.5 * (xyz[:, 1].min() + xyz[:, 1].max()) + 0.25  # between two rows
    a = xyz[:, 0].min() + 0.25 * (xyz[:, 0].max() - xyz[:, 0].min())
    ya, yb = xyz[p1, 1], xyz[pp, 1]
    cross = (ya - y0) * (yb - y0) < 0
    t = np.where(cross, (y0 - ya) / np.where(cross, yb - ya, 1.0), 0.0)
    xc = xyz[p1, 0] + t * (xyz[pp, 0] - xyz[p1, 0])
    keep = ~(cross & (xc < a))
    lay, p1, pp = lay[keep], p1[keep], pp[keep]
    # back to CSR (same order: per particle, layer 0 then layer 1)
    n_ = len(xyz)
    bp2 = np.zeros(n_ + 1, dtype=np.int64)
    np.add.at(bp2, p1 + 1, 1)
    bp2 = np.cumsum(bp2).astype(np.int32)
    bs2 = (bp2[:-1] + np.bincount(p1[lay == 0], minlength=n_)).astype(np.int32)
    return xyz, box, (lay, p1, pp), (bp2, bs2, pp.astype(np.int32))


def _oracle_crack(xyz, box, bl, steps, crit, scale_kn=(1.0, 1.0), scale_tv=1.0, bdamage=None, one_shot=None):
    o = O.System(O.HEXAGON2D, 0.5, O.ELASTIC, xyz, box, bonds=bl)
    o.mode(elmap_literal=0, dirichlet_static=1, cg_tol=1e-13, tol_NR=1e-10, max_NR_iter=60)
    o.set_elastic(0.0, True, 69e3, 0.3, crit)
    layer = o.geti("layer")
    o.setd("Kn", o.getd("Kn") * np.where(layer == 0, scale_kn[0], scale_kn[1]))
    o.setd("Tv", o.getd("Tv") * scale_tv)
    if bdamage is not None:
        o.setd("bdamage", bdamage)
    y = xyz[:, 1]
    lo = np.nonzero(y < y.min() + 1e-6)[0]
    hi = np.nonzero(y > y.max() - 1e-6)[0]
    o.update_geometry()
    o.update_force_state()
    if one_shot is not None:
        ls = O.LoadStep().add_disp(lo, "x", 0.0).add_disp(lo, "y", -one_shot).add_disp(hi, "x", 0.0).add_disp(hi, "y", one_shot)
        o.solve_step_static(ls, -1)
    else:
        for _ in range(steps):
            ls = O.LoadStep().add_disp(lo, "x", 0.0).add_disp(lo, "y", -INC).add_disp(hi, "x", 0.0).add_disp(hi, "y", INC)
            assert o.solve_step_static(ls, -1)
    return o, lo, hi


def test_config4_cracked_plate_breaking_and_adjoint():
    say("[gpu] config 4 cracked hex plate (40 x 40)")
    xyz, box, bl, csr = _cracked_plate(40)
    steps, crit = 5, 1e-3
    o, lo, hi = _oracle_crack(xyz, box, bl, steps, crit)
    g = G.Context(G.HEXAGON2D, 0.5, G.ELASTIC, xyz, bonds=csr)
    g.set_elastic(0.0, True, 69e3, 0.3, crit, elmap_literal=0)
    g.mode(dirichlet_static=1, cg_transpose=0, diag10_literal=0, precision=1, cg_tol=1e-13, tol_NR=1e-10,
           max_NR_iter=60, cg_maxit=50000)
    ls = g.load_step(disp=[(lo, "x", 0.0), (lo, "y", -INC), (hi, "x", 0.0), (hi, "y", INC)])
    g.update_geometry()
    g.update_force_state()
    for _ in range(steps):
        assert g.solve_step_static(ls, -1)
    bd_o, bd_g = o.getd("bdamage"), g.get("bdamage")
    nbroken = int(np.sum(bd_o > 0.5))
    say("   broken bonds", nbroken, "mismatches", int(np.sum(bd_o != bd_g)))
    assert nbroken > 0
    np.testing.assert_array_equal(bd_g, bd_o)
    u_o = o.getd("xyz") - xyz.ravel()
    u_g = g.get("xyz") - xyz.ravel()
    say("   disp rel err", rel(u_g, u_o))
    assert rel(u_g, u_o) < 1e-8
    # adjoint of the top reaction at the final state, broken-bond set frozen
    F, gKn, gTv, nref, rres = g.adjoint_reaction(g.group(hi), "y")
    say("   F", F, "refinements", nref, "rel residual", rres)
    assert rres < 1e-9
    total = INC * steps
    frozen = dict(bdamage=bd_o, one_shot=total)
    F0 = _oracle_crack(xyz, box, bl, 0, 1e9, **frozen)[0].getd("Pin").reshape(-1, 3)[hi, 1].sum()
    assert F == pytest.approx(F0, rel=1e-8)
    layer = o.geti("layer")
    kn, tv = o.getd("Kn"), o.getd("Tv")
    h = 1e-5

    def Fo(**kw):
        return _oracle_crack(xyz, box, bl, 0, 1e9, **frozen, **kw)[0].getd("Pin").reshape(-1, 3)[hi, 1].sum()

    for name, adj, up, dn in (
        ("Kn layer 0", np.sum(gKn[layer == 0] * kn[layer == 0]), dict(scale_kn=(1 + h, 1.0)), dict(scale_kn=(1 - h, 1.0))),
        ("Kn layer 1", np.sum(gKn[layer == 1] * kn[layer == 1]), dict(scale_kn=(1.0, 1 + h)), dict(scale_kn=(1.0, 1 - h))),
        ("Tv", np.sum(gTv * tv), dict(scale_tv=1 + h), dict(scale_tv=1 - h)),
    ):
        fd = (Fo(**up) - Fo(**dn)) / (2 * h)
        err = abs(adj - fd) / abs(fd)
        say(f"   {name}: adjoint {adj:.10e} fd {fd:.10e} rel {err:.2e}")
        assert err < 1e-5
