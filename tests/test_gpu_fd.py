"""Finite-difference stiffness (StiffnessMode::FiniteDifference, stiffness.h:103-154 with
resumeParticleState, particle.h:183-195) on the GPU against the oracle.

The GPU perturbs one colour class of particles per pass and restores positions exactly; the
oracle's fd_exact mode computes the same quotient sequentially, so the FP64 blocks must match
bit for bit, and the FP32 store must be their rounding. The oracle's literal mode (restores by
(x + h) - h) differs from that by the effect of one-ulp position drift divided by h, which
tests/test_oracle.py bounds on the CPU; the static step in FD mode is compared with the literal
oracle at the equilibrium tolerance."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from scenarios import Pair, faces, hex_plate, rel, sc_cube

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _prepared(lat, model):
    """A deformed lattice with some broken and partly damaged bonds (and particle damage for the
    damage models), its state stored so that the last converged copies equal the current ones."""
    if lat == O.SC3D:
        xyz, box = sc_cube(4)
        dim = 3
    else:
        xyz, box = hex_plate(7, 7)
        dim = 2
    p = Pair(lat, 0.5, model, xyz, box)
    if model == O.ELASTIC:
        p.set_elastic(0.0, dim == 2, 200e3, 0.3, 1.0)
    elif model == O.ELASTIC_DAMAGE:
        p.set_elastic_damage(1.0, True, 69e3, 0.3, 1e9, 1.0, 10.0, 0.99)
    else:
        p.set_hcf(1.0, True, 110e3, 0.34, 5.206e-5, 2.019, 0.5)
    rng = np.random.default_rng(11)
    x = xyz + 2e-3 * rng.standard_normal(xyz.shape)
    if dim == 2:
        x[:, 2] = xyz[:, 2]
    nb = p.o.info("nbonds")
    bd = np.zeros(nb)
    bd[::9] = 1.0
    bd[4::13] = 0.5
    p.o.setd("xyz", x)
    p.g.set("xyz", x)
    p.o.setd("bdamage", bd)
    p.g.set("bdamage", bd)
    if model != O.ELASTIC:
        D = np.where(np.arange(len(xyz)) % 5 == 0, 0.3, 0.0)
        p.o.setd("damage", D)
        p.g.set("damage", D)
    for s in (p.o, p.g):
        s.update_geometry()
        s.update_force_state()
        s.store()
    return p, x, dim


@pytest.mark.parametrize("lat,model", [(O.SC3D, O.ELASTIC), (O.HEXAGON2D, O.ELASTIC), (O.HEXAGON2D, O.ELASTIC_DAMAGE),
                                       (O.HEXAGON2D, O.HCF)])
@pytest.mark.parametrize("transpose", [0, 1])
def test_fd_blocks_bitwise(lat, model, transpose):
    say("[gpu] fd blocks", lat, model, transpose)
    p, x, dim = _prepared(lat, model)
    p.o.mode(stiffness_fd=1, fd_exact=1, cg_transpose=transpose)
    p.o.calc_stiffness()
    r, c, v = p.o.K_csr()
    K = sp.csr_matrix((v, c, r), shape=(len(r) - 1, len(r) - 1))
    A = (K.T if transpose else K).tocsr()
    A.sort_indices()
    p.g.mode(stiffness_fd=1, cg_transpose=transpose)
    for prec in (1, 0):
        p.g.mode(precision=prec)
        p.g.calc_stiffness()
        gr, gc, gv = p.g.matrix()
        np.testing.assert_array_equal(gr, A.indptr)
        np.testing.assert_array_equal(gc, A.indices)
        exp = A.data if prec == 1 else A.data.astype(np.float32).astype(np.float64)
        np.testing.assert_array_equal(gv, exp)
    assert p.g.get_mode("fd_colors") > 1
    # positions come back exactly; the restore leaves the state at its last converged copy
    np.testing.assert_array_equal(p.g.get("xyz"), x.ravel())
    for f in ("damage", "bdamage", "dLp"):
        np.testing.assert_array_equal(p.g.get(f), p.o.getd(f), err_msg=f)


def test_fd_static_step_matches_literal_oracle():
    """A static step with FD stiffness (SolverStatic::solveProblemStep, solver_static.h:97-149)
    against the literal oracle: the equilibrium is the same within 1e-8."""
    say("[gpu] fd static step")
    xyz, box = sc_cube(4)
    p = Pair(O.SC3D, 0.5, O.ELASTIC, xyz, box)
    p.set_elastic(0.0, False, 200e3, 0.3, 1.0)
    bottom, top = faces(xyz, 2)
    ls_o, ls_g = p.load(disp=[(bottom, "x", 0.0), (bottom, "y", 0.0), (bottom, "z", 0.0), (top, "z", 0.01)])
    p.o.mode(dirichlet_static=1, stiffness_fd=1, cg_tol=1e-14, tol_NR=1e-11, max_NR_iter=50)
    p.g.mode(dirichlet_static=1, stiffness_fd=1, precision=1, cg_tol=1e-13, tol_NR=1e-11, max_NR_iter=50, cg_maxit=5000)
    for s in (p.o, p.g):
        s.update_geometry()
        s.update_force_state()
    assert p.o.solve_step_static(ls_o, -1)
    assert p.g.solve_step_static(ls_g, -1)
    u_o = p.o.getd("xyz") - xyz.ravel()
    u_g = p.g.get("xyz") - xyz.ravel()
    err = rel(u_g, u_o)
    say("  fd step displacement rel err", err)
    assert err < 1e-8
    assert rel(p.g.get("reaction"), p.o.getd("reaction")) < 1e-8


def test_fd_refused_for_j2():
    """J2Plasticity's FD pass carries the whole return-mapping increment (the state update of
    stiffness.h:127-128 runs on the previous evaluation's stress): the mode is refused."""
    import lpmgpu as G

    xyz, box = sc_cube(3)
    g = G.Context(O.SC3D, 0.5, O.J2, xyz)
    with pytest.raises(RuntimeError, match="J2Plasticity"):
        g.mode(stiffness_fd=1)
