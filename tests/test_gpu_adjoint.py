"""Tangent and adjoint sensitivities on the GPU against central finite differences of the FP64
oracle: the exact force Jacobian-vector product, and d(reaction)/d(Kn, Tv) of a converged load
step (damage and broken bonds frozen), per layer and per bond."""
import sys

import numpy as np
import pytest

import oracle as O
from scenarios import Pair, faces, hex_plate, rel, sc_cube

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _setup(lattice):
    if lattice == "sc":
        xyz, box = sc_cube(5)
        p = Pair(O.SC3D, 0.5, O.ELASTIC, xyz, box)
        p.set_elastic(0.0, False, 200e3, 0.3, 1.0)
        lo, hi = faces(xyz, 2)
        axes, dim = "xyz", 3
    else:
        xyz, box = hex_plate(10, 10)
        p = Pair(O.HEXAGON2D, 0.5, O.ELASTIC, xyz, box)
        p.set_elastic(0.0, True, 69e3, 0.3, 1.0)
        lo, hi = faces(xyz, 1)
        axes, dim = "xy", 2
    return p, xyz, lo, hi, axes, dim


@pytest.mark.parametrize("lattice", ["sc", "hex"])
def test_force_jvp_matches_central_difference(lattice):
    say("[gpu] force jvp", lattice)
    p, xyz, lo, hi, axes, dim = _setup(lattice)
    rng = np.random.default_rng(11)
    x = xyz + 3e-3 * rng.standard_normal(xyz.shape)
    if dim == 2:
        x[:, 2] = xyz[:, 2]
    p.g.set("xyz", x)
    p.g.update_geometry()
    p.g.update_force_state()
    v = rng.standard_normal(len(xyz) * dim)
    jv = p.g.force_jvp(v)
    h = 1e-6
    V = np.zeros_like(x)
    V[:, :dim] = v.reshape(-1, dim)

    def pin(xx):
        p.o.setd("xyz", xx)
        p.o.update_geometry()
        p.o.update_force_state()
        return p.o.getd("Pin").reshape(-1, 3)[:, :dim].ravel().copy()

    fd = (pin(x + h * V) - pin(x - h * V)) / (2 * h)
    say("   jvp rel err", rel(jv, fd))
    assert rel(jv, fd) < 1e-7


def _oracle_reaction(xyz, box, lattice, lo, hi, axes, dim, scale_kn0=1.0, scale_kn1=1.0, scale_tv=1.0, bond=None,
                     bond_scale=1.0):
    lat = O.SC3D if lattice == "sc" else O.HEXAGON2D
    s = O.System(lat, 0.5, O.ELASTIC, xyz, box)
    s.mode(elmap_literal=0, dirichlet_static=1, cg_tol=1e-14, tol_NR=1e-12, max_NR_iter=60)
    if lattice == "sc":
        s.set_elastic(0.0, False, 200e3, 0.3, 1.0)
    else:
        s.set_elastic(0.0, True, 69e3, 0.3, 1.0)
    layer = s.geti("layer")
    kn, tv = s.getd("Kn"), s.getd("Tv")
    kn = kn * np.where(layer == 0, scale_kn0, scale_kn1)
    tv = tv * scale_tv
    if bond is not None:
        kn[bond] *= bond_scale
    s.setd("Kn", kn)
    s.setd("Tv", tv)
    ls = O.LoadStep()
    for a in axes:
        ls.add_disp(lo, a, 0.0)
    ls.add_disp(hi, axes[-1], 0.02)
    s.update_geometry()
    s.update_force_state()
    s.solve_step_static(ls, -1)
    return s.getd("Pin").reshape(-1, 3)[hi, dim - 1].sum()


@pytest.mark.parametrize("lattice", ["sc", "hex"])
def test_adjoint_reaction_gradient_matches_central_difference(lattice):
    say("[gpu] adjoint", lattice)
    p, xyz, lo, hi, axes, dim = _setup(lattice)
    g = p.g
    g.mode(dirichlet_static=1, cg_transpose=0, diag10_literal=0, precision=1, cg_tol=1e-13, tol_NR=1e-12,
           max_NR_iter=60, cg_maxit=20000)
    ls = g.load_step(disp=[(lo, a, 0.0) for a in axes] + [(hi, axes[-1], 0.02)])
    g.update_geometry()
    g.update_force_state()
    assert g.solve_step_static(ls, -1)
    F, gKn, gTv, nref, rres = g.adjoint_reaction(g.group(hi), axes[-1])
    say("   F", F, "refinements", nref, "rel residual", rres)
    assert rres < 1e-9
    F0 = _oracle_reaction(xyz, p.box, lattice, lo, hi, axes, dim)
    assert F == pytest.approx(F0, rel=1e-9)
    layer = p.o.geti("layer")
    kn = p.o.getd("Kn")
    tv = p.o.getd("Tv")
    h = 1e-5
    # (adjoint dF/dln(theta), oracle kwargs up, down, relative step)
    cases = {
        "Kn layer 0": (np.sum(gKn[layer == 0] * kn[layer == 0]), dict(scale_kn0=1 + h), dict(scale_kn0=1 - h), h),
        "Kn layer 1": (np.sum(gKn[layer == 1] * kn[layer == 1]), dict(scale_kn1=1 + h), dict(scale_kn1=1 - h), h),
        "Tv": (np.sum(gTv * tv), dict(scale_tv=1 + h), dict(scale_tv=1 - h), h),
    }
    # one bond on its own: the largest Kn sensitivity
    b = int(np.argmax(np.abs(gKn)))
    hb = 1e-3
    cases[f"Kn of bond {b}"] = (gKn[b] * kn[b], dict(bond=b, bond_scale=1 + hb), dict(bond=b, bond_scale=1 - hb), hb)
    for name, (adj, up, dn, hh) in cases.items():
        fd = (_oracle_reaction(xyz, p.box, lattice, lo, hi, axes, dim, **up) -
              _oracle_reaction(xyz, p.box, lattice, lo, hi, axes, dim, **dn)) / (2 * hh)
        err = abs(adj - fd) / abs(fd)
        say(f"   {name}: adjoint {adj:.10e} fd {fd:.10e} rel {err:.2e}")
        assert err < 1e-5
