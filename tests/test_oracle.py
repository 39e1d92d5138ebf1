"""Pins of the FP64 oracle (oracle/lpm_oracle.cpp) against what fixes its answers
independently: the reference's own K dump, the reference's Fortran closed forms,
brute force, continuum elasticity, conservation laws and dense linear algebra.
CPU only."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def sc_block(n=5, r=0.5, ptype=O.ELASTIC):
    """SC lattice with n particles per side (box chosen so that Q11's offset keeps n per axis)."""
    L = 2 * r * (n - 1)
    a = math.sqrt(3 * L * L)
    # y, z rows start at -a + k*2r; shift the box so that exactly n rows fall inside
    off = (-a) % (2 * r)
    box = [0, L, off, off + L, off, off + L]
    xyz = O.generate(O.SC3D, box, r)
    assert len(xyz) == n ** 3, len(xyz)
    return O.System(O.SC3D, r, ptype, xyz, box), xyz, box


def hex_corner_plate(nrow=8, r=0.2):
    """Hex plate whose even rows start r to the left of the odd rows (particle 0 at the corner)."""
    for b0 in np.arange(-0.5, 0.5, 0.01):
        box = [b0, b0 + 2 * r * nrow, 0.0, 2 * r * math.sqrt(3) / 2 * (nrow - 0.5), 0, 1]
        xyz = O.generate(O.HEXAGON2D, box, r)
        ys = np.unique(np.round(xyz[:, 1], 9))
        r1 = xyz[np.isclose(xyz[:, 1], ys[0])]
        r2 = xyz[np.isclose(xyz[:, 1], ys[1])]
        if len(r1) == len(r2) and r2[:, 0].min() < r1[:, 0].min():
            return xyz, box
    raise AssertionError("no corner configuration found")


# ----------------------------------------------------------------- a1/a2 Kn, Tv
@pytest.mark.parametrize("E,mu,r", [(200e3, 0.3, 0.5), (70e3, 0.33, 0.25)])
def test_kntv_sc_matches_fortran_closed_form(E, mu, r):
    # benchmark/SC_Hailong/VCPM_3D_elasticity_afem_cg_pc_2.F90:40-42
    s, _, _ = sc_block(3, r)
    s.mode(elmap_literal=0)
    s.set_elastic(0.0, False, E, mu, 1.0)
    Kn, Tv, lay = s.getd("Kn"), s.getd("Tv"), s.geti("layer")
    kn = r * E / (2 * (1 + mu))
    tv = r * E * (4 * mu - 1) / (36 * (1 + mu) * (1 - 2 * mu))
    np.testing.assert_allclose(Kn[lay == 0], kn, rtol=1e-13)
    np.testing.assert_allclose(Kn[lay == 1], kn, rtol=1e-13)
    np.testing.assert_allclose(Tv, tv, rtol=1e-12)


def test_kntv_fcc_matches_fortran_closed_form():
    # benchmark/FCC_Hailong/VCPM_3D_elasticity_iso_afem_fcc2.F90:53-55
    r, E, mu = 0.5, 70e3, 0.33
    xyz = O.generate(O.FCC3D, [0, 3, 0, 3, 0, 3], r)
    s = O.System(O.FCC3D, r, O.ELASTIC, xyz)
    s.mode(elmap_literal=0)
    s.set_elastic(0.0, False, E, mu, 1.0)
    Kn, Tv, lay = s.getd("Kn"), s.getd("Tv"), s.geti("layer")
    np.testing.assert_allclose(Kn[lay == 0], math.sqrt(2) * r * E / (2 * (1 + mu)), rtol=1e-13)
    np.testing.assert_allclose(Kn[lay == 1], math.sqrt(2) * r * E / (8 * (1 + mu)), rtol=1e-13)
    np.testing.assert_allclose(Tv, math.sqrt(2) * r * E * (4 * mu - 1) / (48 * (1 + mu) * (1 - 2 * mu)), rtol=1e-12)


def test_literal_mapping_is_the_transpose():
    # Eigen::Map of a row-major-written array is read column-major (unit_cell.h:66)
    xyz, box = hex_corner_plate(6)
    s = O.System(O.HEXAGON2D, 0.2, O.ELASTIC, xyz, box)
    E, mu = 69e3, 0.3
    C11 = E * (1 - mu) / (1 + mu) / (1 - 2 * mu)
    C12 = E * mu / (1 + mu) / (1 - 2 * mu)
    C44 = E / 2 / (1 + mu)
    data = np.array([math.sqrt(3) / 12, -math.sqrt(3) / 12, 0, -math.sqrt(3) / 144, math.sqrt(3) / 48, 0, 0, 0, 1.0])
    for lit, M in ((1, data.reshape(3, 3).T), (0, data.reshape(3, 3))):
        s.mode(elmap_literal=lit)
        s.set_elastic(0.0, False, E, mu, 1.0)
        kt = M @ np.array([C11, C12, C44])
        np.testing.assert_allclose(s.getd("Kn")[0], kt[0], rtol=1e-13)
        np.testing.assert_allclose(s.getd("Tv")[0], kt[1], rtol=1e-13)
    # the literal mapping makes the lattice's bulk stiffness negative (Kn + 6 Tv < 0)
    s.mode(elmap_literal=1)
    s.set_elastic(0.0, False, E, mu, 1.0)
    assert s.getd("Kn")[0] + 6 * s.getd("Tv")[0] < 0


# ----------------------------------------------------------------- golden: the reference's K dump
def test_k_row_matches_reference_swap_dump():
    """Particle 0's K row printed by the reference (.test_k.txt.swp) for a hex plate, E=69 GPa,
    nu=0.3, plane strain. The dump was taken in a slightly deformed state, so entries agree to
    ~1e-5 relative; any dropped, mis-signed or transposed term is off by >= Tv*cs^2 ~ 80."""
    fx = json.load(open(os.path.join(HERE, "golden", "swap_k_row0.json")))
    xyz, box = hex_corner_plate(8)
    s = O.System(O.HEXAGON2D, 0.2, O.ELASTIC, xyz, box)
    cp, ci = s.geti("cptr"), s.geti("cidx")
    row = [int(j) for j in ci[cp[0]:cp[1]]]
    # map the plate's ids onto the dump's ids (dump rows hold 100 particles)
    per_row = int(np.sum(np.isclose(xyz[:, 1], xyz[0, 1])))
    dump_ids = sorted(int(k) for k in fx["blocks"])
    mine = sorted((j // per_row) * 100 + (j % per_row) for j in row)
    assert mine == dump_ids
    s.mode(elmap_literal=0)
    s.set_elastic(0.0, False, 69e3, 0.3, 1.0)
    s.update_geometry()
    s.update_force_state()
    worst = 0.0
    for j in row:
        K = s.local_stiffness(0, j)[:2, :2]
        D = np.array(fx["blocks"][str((j // per_row) * 100 + (j % per_row))])[:2, :2]
        worst = max(worst, np.abs(K - D).max())
        # Q2: the diagonal block's [1][0] entry is never written
        if j == 0:
            assert K[1, 0] == 0.0 and D[1, 0] == 0.0
    assert worst < 0.2, worst
    # the literal (column-major) mapping does not reproduce the dump
    s.mode(elmap_literal=1)
    s.set_elastic(0.0, False, 69e3, 0.3, 1.0)
    s.update_geometry()
    K = s.local_stiffness(0, 0)[:2, :2]
    assert np.abs(K - np.array(fx["blocks"]["0"])[:2, :2]).max() > 1000


# ----------------------------------------------------------------- a27/a11 topology
@pytest.mark.parametrize("lat,box,r,nl", [
    (O.HEXAGON2D, [0, 6, 0, 6, 0, 1], 0.25, (6, 6)),
    (O.SC3D, [0, 3, 0, 3, 0, 3], 0.25, (6, 12)),
    (O.FCC3D, [0, 2.5, 0, 2.5, 0, 2.5], 0.25, (12, 6)),
    (O.SQUARE2D, [0, 4, 0, 4, 0, 1], 0.25, (4, 4)),
])
def test_bonds_cell_list_equals_bruteforce(lat, box, r, nl):
    xyz = O.generate(lat, box, r)
    a = O.System(lat, r, O.ELASTIC, xyz, box, bruteforce=False)
    b = O.System(lat, r, O.ELASTIC, xyz, box, bruteforce=True)
    for f in ("bptr", "bsplit", "p2", "layer", "opp", "cptr", "cidx"):
        np.testing.assert_array_equal(a.geti(f), b.geti(f), err_msg=f)
    np.testing.assert_array_equal(a.getd("dis0"), b.getd("dis0"))
    # brute force restated in numpy: all pairs with d < 1.01*cutoff[1], layer split at 1.01*cutoff[0]
    cell = {O.HEXAGON2D: (2 * r, 2 * math.sqrt(3) * r), O.SC3D: (2 * r, 2 * math.sqrt(2) * r),
            O.FCC3D: (2 * r, 2 * math.sqrt(2) * r), O.SQUARE2D: (2 * r, 2 * math.sqrt(2) * r)}[lat]
    D = np.sqrt(((xyz[:, None, :] - xyz[None, :, :]) ** 2).sum(-1))
    bp, p2, lay = a.geti("bptr"), a.geti("p2"), a.geti("layer")
    for i in range(len(xyz)):
        js = np.nonzero((D[i] < 1.01 * cell[1]) & (np.arange(len(xyz)) != i))[0]
        l0 = [j for j in js if D[i, j] < 1.01 * cell[0]]
        l1 = [j for j in js if D[i, j] >= 1.01 * cell[0]]
        assert list(p2[bp[i]:bp[i + 1]]) == l0 + l1
        assert list(lay[bp[i]:bp[i + 1]]) == [0] * len(l0) + [1] * len(l1)
    # interior particles have the lattice's full neighbour counts
    nb = a.geti("nb")
    assert nb.max() == nl[0] + nl[1]
    cnt0 = np.array([np.sum(lay[bp[i]:bp[i + 1]] == 0) for i in range(len(xyz))])
    assert cnt0.max() == nl[0]


@pytest.mark.parametrize("lat,box,r,full", [
    (O.HEXAGON2D, [0, 6, 0, 6, 0, 1], 0.25, 31),
    (O.SC3D, [0, 3, 0, 3, 0, 3], 0.25, 61),
    (O.FCC3D, [0, 4, 0, 4, 0, 4], 0.25, 61),
])
def test_conns_two_hop_per_layer(lat, box, r, full):
    xyz = O.generate(lat, box, r)
    s = O.System(lat, r, O.ELASTIC, xyz, box)
    bp, bs, p2 = s.geti("bptr"), s.geti("bsplit"), s.geti("p2")
    cp, ci, nlq = s.geti("cptr"), s.geti("cidx"), s.geti("nconn_largeq")
    n = len(xyz)
    for i in range(n):
        exp = set()
        for lo_hi in ((bp[i], bs[i]), (bs[i], bp[i + 1])):
            layer0 = lo_hi[0] == bp[i]
            for b in range(*lo_hi):
                k = p2[b]
                exp.add(k)
                kl = (bp[k], bs[k]) if layer0 else (bs[k], bp[k + 1])
                exp.update(p2[kl[0]:kl[1]])
        got = list(ci[cp[i]:cp[i + 1]])
        assert got == sorted(exp)
        assert nlq[i] == sum(1 for j in got if j >= i)
    # nconn of an interior particle = nneighbors_AFEM + 1 (unit_cell.h:71,89,106)
    assert (cp[1:] - cp[:-1]).max() == full


def test_generators_spacing_and_box():
    for lat, box, r in ((O.HEXAGON2D, [0, 5, 0, 5, 0, 1], 0.25), (O.SC3D, [0, 3, 0, 3, 0, 3], 0.25),
                        (O.FCC3D, [0, 3, 0, 3, 0, 3], 0.25), (O.SQUARE2D, [0, 3, 0, 3, 0, 1], 0.25)):
        x = O.generate(lat, box, r)
        assert len(x) > 8
        lo = np.array(box[0::2])
        hi = np.array(box[1::2])
        dim = 2 if lat in (O.HEXAGON2D, O.SQUARE2D) else 3
        assert np.all(x[:, :dim] >= lo[:dim]) and np.all(x[:, :dim] <= hi[:dim])
        D = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)) + np.eye(len(x)) * 1e9
        np.testing.assert_allclose(D.min(), 2 * r, rtol=1e-12)
    # rotation matrices are orthonormal in every convention
    for flag in (0, 1, 2):
        R = O.rmatrix(flag, [0.3, 1.1, 2.5])
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-14)


# ----------------------------------------------------------------- a3-a9 force state
def test_zero_force_at_rest_and_newton_third_law():
    s, xyz, _ = sc_block(5)
    s.mode(elmap_literal=0)
    s.set_elastic(0.0, False, 200e3, 0.3, 1.0)
    s.update_geometry()
    s.update_force_state()
    assert np.all(s.getd("Pin") == 0.0)
    assert np.all(s.getd("bforce") == 0.0)
    rng = np.random.default_rng(1)
    s.setd("xyz", xyz + 1e-3 * rng.standard_normal(xyz.shape))
    s.update_geometry()
    s.update_force_state()
    P = s.getd("Pin").reshape(-1, 3)
    assert np.abs(P).max() > 1.0
    np.testing.assert_allclose(P.sum(0), 0.0, atol=1e-9 * np.abs(P).max())


def test_rigid_motion_invariance():
    s, xyz, _ = sc_block(4)
    s.mode(elmap_literal=0)
    s.set_elastic(0.0, False, 200e3, 0.3, 1.0)
    rng = np.random.default_rng(2)
    x = xyz + 1e-3 * rng.standard_normal(xyz.shape)
    s.setd("xyz", x)
    s.update_geometry()
    s.update_force_state()
    P0 = s.getd("Pin").reshape(-1, 3)
    R = O.rmatrix(0, [0.4, -0.7, 1.3])
    s.setd("xyz", x @ R.T + np.array([3.0, -1.0, 0.5]))
    s.update_geometry()
    s.update_force_state()
    P1 = s.getd("Pin").reshape(-1, 3)
    np.testing.assert_allclose(P1, P0 @ R.T, atol=1e-9 * np.abs(P0).max())


@pytest.mark.parametrize("lat", [O.SC3D, O.HEXAGON2D])
def test_uniform_strain_reproduces_elastic_constants(lat):
    """The LPM mapping is built so that an interior particle under uniform strain carries the
    continuum stress: uniaxial strain e in x gives s11 = C11 e and s22 = C12 e."""
    E, mu, e = 200e3, 0.3, 1e-7
    if lat == O.SC3D:
        s, xyz, _ = sc_block(7)
    else:
        xyz, box = hex_corner_plate(12, 0.5)
        s = O.System(lat, 0.5, O.ELASTIC, xyz, box)
    s.mode(elmap_literal=0)
    s.set_elastic(0.0, False, E, mu, 1.0)
    x = xyz.copy()
    x[:, 0] *= (1 + e)
    s.setd("xyz", x)
    s.update_geometry()
    s.update_force_state()
    nb = s.geti("nb")
    full = nb.max()
    bp = s.geti("bptr")
    p2 = s.geti("p2")
    # a particle whose bonds' partners all have full neighbourhoods is in the uniform interior
    interior = [i for i in range(len(xyz)) if nb[i] == full and all(nb[j] == full for j in p2[bp[i]:bp[i + 1]])]
    assert interior
    st = s.getd("stress").reshape(-1, 6)[interior]
    C11 = E * (1 - mu) / (1 + mu) / (1 - 2 * mu)
    C12 = E * mu / (1 + mu) / (1 - 2 * mu)
    np.testing.assert_allclose(st[:, 0], C11 * e, rtol=1e-5)
    np.testing.assert_allclose(st[:, 1], C12 * e, rtol=1e-5)
    if lat == O.SC3D:
        np.testing.assert_allclose(st[:, 2], C12 * e, rtol=1e-5)
    np.testing.assert_allclose(st[:, 3:], 0.0, atol=1e-9 * C11 * e)


def test_interior_bond_force_under_dilation():
    """Uniform dilation e: dLe = e*d_l for every bond, so f = 2 e d_l (Kn + n_l Tv)."""
    r, e = 0.5, 1e-6
    s, xyz, _ = sc_block(6, r)
    s.mode(elmap_literal=0)
    s.set_elastic(0.0, False, 200e3, 0.3, 1.0)
    s.setd("xyz", xyz * (1 + e))
    s.update_geometry()
    s.update_force_state()
    nb, bp, lay = s.geti("nb"), s.geti("bptr"), s.geti("layer")
    Kn, Tv, f = s.getd("Kn"), s.getd("Tv"), s.getd("bforce")
    i = int(np.nonzero(nb == 18)[0][0])
    for b in range(bp[i], bp[i + 1]):
        d = 2 * r if lay[b] == 0 else 2 * math.sqrt(2) * r
        nl = 6 if lay[b] == 0 else 12
        np.testing.assert_allclose(f[b], 2 * e * d * (Kn[b] + nl * Tv[b]), rtol=1e-8)


def test_residual_and_reaction_layout():
    s, xyz, _ = sc_block(3)
    s.mode(elmap_literal=0)
    s.set_elastic(0.0, False, 200e3, 0.3, 1.0)
    cons = np.zeros(3 * len(xyz), dtype=np.int32)
    cons[[0, 4, 8, 11]] = 1
    s.seti("cons", cons)
    pex = np.arange(3 * len(xyz), dtype=float)
    s.setd("Pex", pex)
    rng = np.random.default_rng(3)
    s.setd("xyz", xyz + 1e-3 * rng.standard_normal(xyz.shape))
    s.update_geometry()
    s.update_force_state()
    s.update_rr()
    P = s.getd("Pin")
    np.testing.assert_array_equal(s.getd("residual"), (1 - cons) * (pex - P))
    np.testing.assert_array_equal(s.getd("reaction"), P[cons == 1])


# ----------------------------------------------------------------- a12-a16 stiffness and CG
def test_analytic_block_close_to_finite_difference():
    """derivative.h:5 says the analytic blocks drop small terms; at rest they must agree with the
    forward difference of the force (stiffness.h:103-154) to the FD truncation level."""
    s, xyz, _ = sc_block(5)
    s.mode(elmap_literal=0)
    s.set_elastic(0.0, False, 200e3, 0.3, 1.0)
    s.update_geometry()
    s.update_force_state()
    s.store()
    cp, ci = s.geti("cptr"), s.geti("cidx")
    i = 62  # interior
    scale = np.abs(s.local_stiffness(i, i)).max()
    for j in ci[cp[i]:cp[i + 1]]:
        Ka = s.local_stiffness(i, int(j))
        s.mode(stiffness_fd=1)
        Kf = s.local_stiffness(i, int(j))
        s.mode(stiffness_fd=0)
        assert np.abs(Ka - Kf).max() < 1e-4 * scale, (j, Ka, Kf)


def test_fd_restore_drift_is_bounded():
    """localStiffnessFD restores a moved particle by (x + h) - h (stiffness.h:138-140): exact
    unless x + h crosses into the next binade, where the particle is left one ulp off and later
    blocks see it. The oracle's fd_exact mode (the parallel evaluation's quotient) restores x
    exactly: equal blocks on a lattice without such coordinates, and blocks within
    |K| * ulp(x) / h where a coordinate sits within h below a power of two."""
    xyz = O.generate(O.HEXAGON2D, [0, 7, 0, 6, 0, 1], 0.5)

    def fd(x, exact):
        s = O.System(O.HEXAGON2D, 0.5, O.ELASTIC, xyz)
        s.mode(elmap_literal=0)
        s.set_elastic(0.0, True, 69e3, 0.3, 1.0)
        s.setd("xyz", x)
        s.update_geometry()
        s.update_force_state()
        s.store()
        s.mode(stiffness_fd=1, fd_exact=int(exact))
        s.calc_stiffness()
        return s.getd("kval"), s.getd("xyz")

    rng = np.random.default_rng(2)
    x = xyz + 1e-3 * rng.standard_normal(xyz.shape)
    x[:, 2] = 0
    h = 1e-6 * 0.5
    assert not np.any(((x + h) - h) != x)  # no coordinate within h below a power of two
    ke, xe = fd(x, True)
    kl, xl = fd(x, False)
    np.testing.assert_array_equal(kl, ke)
    np.testing.assert_array_equal(xl, xe)
    k = int(np.argmin(np.abs(x[:, 0] - 2.0)))
    x[k, 0] = 2.0 - 1e-7  # (x + h) - h = x - 2^-52
    ke, xe = fd(x, True)
    kl, xl = fd(x, False)
    np.testing.assert_array_equal(xe, x.ravel())
    assert xl[3 * k] == x[k, 0] - 2.0 ** -52       # the literal restore drifted
    assert np.any(kl != ke)
    assert np.abs(kl - ke).max() <= 1e-8 * np.abs(ke).max()


def _small_problem(eliminate=True):
    s, xyz, _ = sc_block(4)
    s.mode(elmap_literal=0)
    s.set_elastic(0.0, False, 200e3, 0.3, 1.0)
    z = xyz[:, 2]
    bottom = np.nonzero(z < z.min() + 0.1)[0]
    top = np.nonzero(z > z.max() - 0.1)[0]
    ls = O.LoadStep()
    for ax in "xyz":
        ls.add_disp(bottom, ax, 0.0)
    ls.add_disp(top, "z", 0.01)
    s.apply_bc(ls)
    s.update_geometry()
    s.update_force_state()
    s.update_rr()
    s.calc_stiffness()
    if eliminate:
        s.update_stiffness_disp_bc()
    return s, ls, bottom, top


@pytest.mark.parametrize("transpose", [1, 0])
def test_cg_converges_to_dense_solution(transpose):
    s, _, _, _ = _small_problem()
    # symmetric diagonal blocks: K is then exactly symmetric at rest and CG reaches 1e-14
    s.mode(cg_transpose=transpose, cg_tol=1e-14, diag10_literal=0)
    s.calc_stiffness()
    s.update_stiffness_disp_bc()
    K = s.K_dense()
    r = s.getd("residual")
    assert s.cg() == 1
    d = s.getd("disp")
    A = K.T if transpose else K
    d_ref = np.linalg.solve(A, r)
    np.testing.assert_allclose(d, d_ref, rtol=0, atol=1e-10 * np.abs(d_ref).max())
    assert 0 < s.info("cg_iters") <= 2 * len(r)


def test_literal_k_asymmetry_is_only_q2():
    """At rest the analytic K is symmetric except for the diagonal block's unwritten [1][0]
    entry (derivative.h:38)."""
    s, _, _, _ = _small_problem(eliminate=False)
    K = s.K_dense()
    A = K - K.T
    n = K.shape[0] // 3
    mask = np.zeros_like(K, dtype=bool)
    for i in range(n):
        mask[3 * i + 1, 3 * i] = mask[3 * i, 3 * i + 1] = True
    assert np.abs(A[~mask]).max() <= 1e-12 * np.abs(K).max()
    assert np.abs(A[mask]).max() > 1.0
    s.mode(diag10_literal=0)
    s.calc_stiffness()
    K = s.K_dense()
    assert np.abs(K - K.T).max() <= 1e-12 * np.abs(K).max()


def test_dirichlet_elimination():
    s, _, bottom, top = _small_problem(eliminate=False)
    K0 = s.K_dense()
    s.update_stiffness_disp_bc()
    K1 = s.K_dense()
    cons = s.geti("cons")
    el = np.nonzero(cons[np.repeat(np.arange(len(cons) // 3), 3) * 3 + np.tile([0, 1, 2], len(cons) // 3)] == 1)[0]
    nd = np.linalg.norm(np.diag(K0))
    for R in el:
        assert np.all(np.delete(K1[R], R) == 0) and np.all(np.delete(K1[:, R], R) == 0)
        assert K1[R, R] == pytest.approx(nd, rel=1e-14)
    keep = np.setdiff1d(np.arange(K0.shape[0]), el)
    np.testing.assert_array_equal(K1[np.ix_(keep, keep)], K0[np.ix_(keep, keep)])


def test_static_step_equilibrium_and_reaction_balance():
    s, xyz, _ = sc_block(4)
    s.mode(elmap_literal=0, dirichlet_static=1, cg_tol=1e-12, tol_NR=1e-10)
    s.set_elastic(0.0, False, 200e3, 0.3, 1.0)
    z = xyz[:, 2]
    bottom = np.nonzero(z < z.min() + 0.1)[0]
    top = np.nonzero(z > z.max() - 0.1)[0]
    ls = O.LoadStep()
    for ax in "xyz":
        ls.add_disp(bottom, ax, 0.0)
    ls.add_force(top, 0.0, 0.0, 50.0)
    s.update_geometry()
    s.update_force_state()
    assert s.solve_step_static(ls, -1)
    P = s.getd("Pin").reshape(-1, 3)
    Pex = s.getd("Pex").reshape(-1, 3)
    # reactions at the clamped face balance the applied load (sum of Pin over all particles is 0)
    np.testing.assert_allclose(P[bottom, 2].sum(), -50.0, rtol=1e-8)
    free = np.setdiff1d(np.arange(len(P)), bottom)
    assert np.abs((Pex - P)[free]).max() < 1e-8 * 50
    # constrained particles did not drift (Dirichlet DoFs eliminated)
    np.testing.assert_array_equal(s.getd("xyz").reshape(-1, 3)[bottom], xyz[bottom])


def test_step_halving_keeps_pex_literal():
    """A step that cannot converge is halved and re-run (solver_static.h:78-87); Pex is not
    rolled back (Q7), positions are."""
    s, xyz, _ = sc_block(3)
    s.mode(elmap_literal=0, dirichlet_static=1, cg_tol=1e-12, max_NR_iter=1, tol_NR=1e-14)
    s.set_elastic(0.0, False, 200e3, 0.3, 1.0)
    top = np.nonzero(xyz[:, 2] > xyz[:, 2].max() - 0.1)[0]
    bottom = np.nonzero(xyz[:, 2] < xyz[:, 2].min() + 0.1)[0]
    ls = O.LoadStep()
    for ax in "xyz":
        ls.add_disp(bottom, ax, 0.0)
    ls.add_force(top, 0.0, 0.0, 8.0)
    s.update_geometry()
    s.update_force_state()
    assert not s.solve_step_static(ls, -1)
    s.reset(True)
    np.testing.assert_array_equal(s.getd("xyz"), xyz.ravel())
    pex = s.getd("Pex").reshape(-1, 3)
    np.testing.assert_allclose(pex[top, 2].sum(), 8.0)  # still applied after the reset


# ----------------------------------------------------------------- a20-a24 damage / state
def test_func_phi_closed_form():
    assert O.func_phi(0.0, 2.0) == 1.0
    assert O.func_phi(2.0, 2.0) == 0.25
    assert O.func_phi(4.0, 2.0) == pytest.approx(1 / (1 + 256) ** 2, rel=1e-15)


def test_bond_breaking_flags():
    s, xyz, _ = sc_block(4)
    s.mode(elmap_literal=0)
    s.set_elastic(0.0, False, 200e3, 0.3, 1e-3)
    x = xyz.copy()
    x[:, 0] *= np.where(x[:, 0] > 1.0, 1.002, 1.0)  # stretch the bonds that cross x = 1
    s.setd("xyz", x)
    s.update_geometry()
    s.update_force_state()
    bs = s.getd("bstrain")
    expect = bs >= 1e-3
    assert expect.any() and not expect.all()
    assert s.update_static_damage(-1)
    np.testing.assert_array_equal(s.getd("bdamage") == 1.0, expect)
    # broken bonds carry no force and no geometry afterwards (bond.h:32-38)
    s.update_geometry()
    s.update_force_state()
    assert np.all(s.getd("bforce")[expect] == 0.0)
    assert np.all(s.getd("cs").reshape(-1, 3)[expect] == 0.0)
    assert not s.update_static_damage(-1)


def test_strain_energy_density_matches_continuum():
    """Elastic-damage energy (particle_elastic_damage.h:89-104) of an interior particle under
    uniaxial strain e equals the continuum 1/2 C11 e^2."""
    E, mu, e = 200e3, 0.3, 1e-6
    s, xyz, _ = sc_block(7, ptype=O.ELASTIC_DAMAGE)
    s.mode(elmap_literal=0)
    s.set_elastic_damage(1.0, False, E, mu, 1e9, 1.0, 10.0, 0.99)
    x = xyz.copy()
    x[:, 0] *= (1 + e)
    s.setd("xyz", x)
    s.update_geometry()
    s.update_force_state()
    s.update_state_var()
    nb, bp, p2 = s.geti("nb"), s.geti("bptr"), s.geti("p2")
    i = [k for k in range(len(xyz)) if nb[k] == 18 and all(nb[j] == 18 for j in p2[bp[k]:bp[k + 1]])][0]
    C11 = E * (1 - mu) / (1 + mu) / (1 - 2 * mu)
    U = s.getd("state_var").reshape(-1, 7)[i, 0]
    np.testing.assert_allclose(U, 0.5 * C11 * e * e, rtol=1e-5)
    # below the damage surface k0 no damage rate
    assert s.getd("Ddot_local")[i] == 0.0


def test_hcf_state_sentinel_and_reset():
    """HCF min-energy sentinel is INT_MAX stored as a double (particle_fatigue_hcf.h:27-28) and is
    re-armed after each fatigue damage update (:156-157)."""
    xyz = O.generate(O.HEXAGON2D, [0, 3, 0, 3, 0, 1], 0.25)
    s = O.System(O.HEXAGON2D, 0.25, O.HCF, xyz)
    s.mode(elmap_literal=0)
    s.set_hcf(1.0, True, 110e3, 0.34, 5.206e-5, 2.019, 0.5)
    sv = s.getd("state_var").reshape(-1, 7)
    assert np.all(sv[:, 1] == 2147483647.0) and np.all(sv[:, 0] == 0) and np.all(sv[:, 2] == 0)
    rng = np.random.default_rng(4)
    s.setd("xyz", xyz + 1e-3 * rng.standard_normal(xyz.shape))
    s.update_geometry()
    s.update_force_state()
    s.update_state_var()
    sv = s.getd("state_var").reshape(-1, 7)
    U = np.array([s.strain_energy(i) for i in range(len(xyz))])
    np.testing.assert_array_equal(sv[:, 1], U)
    np.testing.assert_array_equal(sv[:, 2], U)
    # a single energy level has no range: no damage rate yet
    assert np.all(s.getd("Ddot_local") == 0)
    s.setd("xyz", xyz + 2e-3 * rng.standard_normal(xyz.shape))
    s.update_geometry()
    s.update_force_state()
    s.update_state_var()
    sv = s.getd("state_var").reshape(-1, 7)
    U2 = np.array([s.strain_energy(i) for i in range(len(xyz))])
    np.testing.assert_array_equal(sv[:, 1], np.minimum(U, U2))
    np.testing.assert_array_equal(sv[:, 2], np.maximum(U, U2))
    assert np.all(s.getd("Ddot_local") > 0)
    s.search_nonlocal(1.5)
    s.update_fatigue_damage(1.0, -1)
    sv = s.getd("state_var").reshape(-1, 7)
    assert np.all(sv[:, 1] == 2147483647.0) and np.all(sv[:, 2] == 0)
    assert np.all(s.getd("damage") > 0)


def test_nonlocal_average_of_a_constant_is_the_constant():
    """The nonlocal rate is a phi*V_m-weighted mean (solver_static.h:36-51): a uniform local rate
    averages to itself; a particle with no nonlocal neighbour gets 0/0 = NaN (Q9)."""
    xyz = O.generate(O.HEXAGON2D, [0, 3, 0, 3, 0, 1], 0.25)
    s = O.System(O.HEXAGON2D, 0.25, O.ELASTIC_DAMAGE, xyz)
    s.mode(elmap_literal=0)
    s.set_elastic_damage(1.0, True, 3.2e3, 0.28, 0.1936, 1.0, 10.0, 0.99)
    s.search_nonlocal(1.5)
    s.setd("Ddot_local", np.full(len(xyz), 0.125))
    s.update_static_damage(-1)
    np.testing.assert_allclose(s.getd("Ddot_nonlocal"), 0.125, rtol=1e-14)
    s2 = O.System(O.HEXAGON2D, 0.25, O.ELASTIC_DAMAGE, xyz)
    s2.update_static_damage(-1)
    assert np.all(np.isnan(s2.getd("Ddot_nonlocal")))


def test_dump_format(tmp_path):
    s, xyz, box = sc_block(3)
    p = str(tmp_path / "a.dump")
    s.write_dump(p, 7)
    lines = open(p).read().splitlines()
    assert lines[0] == "ITEM: TIMESTEP" and lines[1] == "7" and lines[3] == str(len(xyz))
    assert lines[8].startswith("ITEM: ATOMS id type x y z dx dy dz s11")
    assert len(lines) == 9 + len(xyz)
    assert len(lines[9].split()) == 17


def test_j2_return_mapping_textbook_cases():
    """J2 radial return with linear mixed hardening (particle_j2plasticity.h:78-140) against
    textbook facts, with the stress set directly: von Mises sigma_eq of uniaxial stress s is |s|
    and of pure shear t is sqrt(3) t; the plastic multiplier is (sigma_eq - sigma_y)/(3G + H);
    uniaxial plastic flow is (1, -1/2, -1/2) dlambda; the back stress moves by 2/3 xi H depsp;
    triaxiality is 1/3 in uniaxial tension and 0 in shear; a bond of length dis along x
    lengthens plastically by dis * depsp_xx; below yield nothing changes."""
    s, xyz, box = sc_block(4, ptype=O.J2)
    s.mode(elmap_literal=0)
    s.set_j2(0.0, False, 200e3, 0.3, 300.0, 0.5, 2000.0, 0.2, 1.0)
    E, mu, sy, xi, H, A = 200e3, 0.3, 300.0, 0.5, 2000.0, 0.2
    G = E / (2 * (1 + mu))
    s.update_geometry()
    s.update_force_state()
    n = len(xyz)
    st = np.zeros((n, 6))
    st[0, 0] = 450.0            # uniaxial tension above yield
    st[1, 5] = 200.0            # pure shear (Voigt xy), sqrt(3)*200 = 346.4 > 300
    st[2, 0] = 100.0            # uniaxial, elastic
    s.setd("stress", st.ravel())
    dlp0 = s.getd("dLp").copy()
    s.update_state_var()
    sv = s.getd("state_var").reshape(n, 7)
    Dd = s.getd("Ddot_local")
    dl_u = (450.0 - sy) / (3 * G + H)
    dl_s = (np.sqrt(3.0) * 200.0 - sy) / (3 * G + H)
    assert sv[0, 0] == pytest.approx(dl_u, rel=1e-13)
    np.testing.assert_allclose(sv[0, 1:4], 2.0 / 3.0 * xi * H * dl_u * np.array([1.0, -0.5, -0.5]), rtol=1e-12)
    assert Dd[0] == pytest.approx(dl_u * (1 + A / 3.0), rel=1e-12)
    assert sv[1, 0] == pytest.approx(dl_s, rel=1e-13)
    # shear flow: depsp_xy = 1.5 dlambda t / sigma_eq = dlambda sqrt(3)/2 (Voigt stores the tensor component)
    assert sv[1, 6] == pytest.approx(2.0 / 3.0 * xi * H * dl_s * np.sqrt(3.0) / 2.0, rel=1e-12)
    assert Dd[1] == pytest.approx(dl_s, rel=1e-12)
    assert sv[2, 0] == 0.0 and Dd[2] == 0.0
    # plastic bond lengthening of particle 0's bonds: dis * (n . depsp . n) with depsp = dl (1, -1/2, -1/2)
    bptr = s.geti("bptr")
    cs = s.getd("cs").reshape(-1, 3)
    dis = s.getd("dis")
    b = np.arange(bptr[0], bptr[1])
    exp = dis[b] * dl_u * (cs[b, 0] ** 2 - 0.5 * cs[b, 1] ** 2 - 0.5 * cs[b, 2] ** 2)
    np.testing.assert_allclose(s.getd("dLp")[b] - dlp0[b], exp, rtol=1e-12, atol=1e-18)


# ----------------------------------------------------------------- geometry tools (holes, notches, images)
def test_geometry_cut_shapes_closed_form():
    """The tools' inCircle / inNotch / inReverseNotch on points whose answer the geometry fixes:
    circle boundary and just outside; the notch is the wedge of half-angle 30 degrees opening to
    +x from the tip, clipped to |y - ty| <= half width; the reverse notch opens to -x."""
    s3 = 1 / math.sqrt(3)
    pts = np.array([[3.0, 0.0, 0], [3.0 + 1e-9, 0.0, 0], [0.0, 0.0, 0],            # circle r=3 at origin
                    [11.0, 5.0, 0], [9.0, 5.0, 0], [11.0, 5.0 + 0.9 * s3, 0],         # notch tip (10, 5), hw 2
                    [11.0, 5.0 + 1.1 * s3, 0], [20.0, 7.5, 0], [20.0, 6.5, 0],
                    [29.0, 5.0, 0], [31.0, 5.0, 0]])                                   # reverse notch tip (30, 5)
    removed_circle = O.cut_particles(pts, [(0, 0.0, 0.0, 3.0)]) == 0
    assert list(removed_circle[:3]) == [True, False, True]
    removed_notch = O.cut_particles(pts, [(1, 10.0, 5.0, 2.0)]) == 0
    # inside the wedge and the band; behind the tip; inside the wedge; outside the wedge; outside the band; inside
    assert list(removed_notch[3:9]) == [True, False, True, False, False, True]
    removed_rev = O.cut_particles(pts, [(2, 30.0, 5.0, 2.0)]) == 0
    assert list(removed_rev[9:11]) == [True, False]


def _seg_point_distance(a, b, c):
    ab, ac = b - a, c - a
    t = np.clip(np.dot(ac, ab) / np.dot(ab, ab), 0.0, 1.0)
    return np.linalg.norm(a + t * ab - c)


def test_search_neighbor_cut_against_exact_geometry():
    """searchNeighbor of the tools: with no cut it equals createBonds (same bonds, same per-layer
    order); with a hole, a bond is dropped when its segment passes well inside the circle and
    kept when it stays outside (the 10-midpoint rule decides only grazing segments, skipped)."""
    xyz = O.generate(O.SQUARE2D, [0, 20, 0, 20, 0, 1], 0.5)
    lay, p1, p2 = O.search_neighbor_cut(O.SQUARE2D, 0.5, xyz, [])
    s = O.System(O.SQUARE2D, 0.5, O.ELASTIC, xyz, bruteforce=True)
    bp, bsp, q2 = s.geti("bptr"), s.geti("bsplit"), s.geti("p2")
    for i in range(len(xyz)):
        sel = p1 == i
        np.testing.assert_array_equal(p2[sel][lay[sel] == 0], q2[bp[i]:bsp[i]])
        np.testing.assert_array_equal(p2[sel][lay[sel] == 1], q2[bsp[i]:bp[i + 1]])
    c, r = np.array([10.3, 9.7, 0.0]), 4.2
    kept = xyz  # the bond rule alone (the tools also drop the particles inside, see above)
    lay, p1, p2 = O.search_neighbor_cut(O.SQUARE2D, 0.5, kept, [(0, c[0], c[1], r)])
    present = set(zip(p1.tolist(), p2.tolist()))
    full = O.search_neighbor_cut(O.SQUARE2D, 0.5, kept, [])
    checked = 0
    for i, j in zip(full[1].tolist(), full[2].tolist()):
        d = _seg_point_distance(kept[i], kept[j], c)
        if d < r - 0.2:
            assert (i, j) not in present
            checked += 1
        elif d > r:
            assert (i, j) in present
            checked += 1
    assert checked > 0.9 * len(full[0]) and len(present) < len(full[0])


def test_image_particles_from_reference_image():
    """createPlate of the image-to-dump notebook on the reference's G6a_XZ_128 image: one particle
    per white pixel, at the pixel centre with rows counted from the top, in row-major order."""
    import lpmgpu as G

    img = np.load(os.path.join(HERE, "golden", "G6a_XZ_128.npy"))
    pixel = 0.0061
    xyz = G.image_particles(img, pixel)
    rows, cols = np.nonzero(img == 255)
    assert len(xyz) == len(rows) == 14749
    np.testing.assert_allclose(xyz[:, 0], (cols + 0.5) * pixel, rtol=0, atol=1e-15)
    np.testing.assert_allclose(xyz[:, 1], (img.shape[0] - rows - 0.5) * pixel, rtol=0, atol=1e-15)
    assert np.all(xyz[:, 2] == 0)
