"""Closed-form pins of the oracle's damage, plasticity-damage and fatigue laws, of the
fatigue cycle driver and of the static step-halving driver. CPU only.

Each expected value below is computed by hand from the reference's formulas applied to a
state whose inputs are known in closed form, never by calling the oracle's own helpers:
 - the strain energy of a fully bonded SC particle under a uniform dilation e is the
   continuum value 9/2 K e^2 with K = E / (3 (1 - 2 nu)) (the SC bond constants of the
   Fortran closed form reproduce the bulk modulus exactly);
 - rates, damage increments and thresholds use values that are exact in binary, so the
   expected numbers are exact too;
 - the cycle driver's visited cycles are listed by hand for each time map.
Every check is chosen so that a plausible slip (swapped exponents, a dropped term, '>' for
'>=', a missing duplicate step, a jump by the wrong amount) changes the asserted value.
"""
import math

import numpy as np
import pytest

import oracle as O

E, NU, R = 200e3, 0.3, 0.5
K_BULK = E / (3 * (1 - 2 * NU))


def sc5(ptype):
    """SC 5^3 (27 fully bonded interior particles), box shifted so that 5 rows fall inside."""
    n, r = 5, R
    L = 2 * r * (n - 1)
    a = math.sqrt(3 * L * L)
    off = (-a) % (2 * r)
    box = [0, L, off, off + L, off, off + L]
    xyz = O.generate(O.SC3D, box, r)
    assert len(xyz) == n ** 3
    return O.System(O.SC3D, r, ptype, xyz, box), xyz


def dilate(s, xyz, e):
    """Uniform dilation x -> (1 + e) x; returns the ids of the fully bonded particles and the
    closed-form energy density 9/2 K e^2 they carry (particle_elastic_damage.h:89-104)."""
    s.setd("xyz", xyz * (1 + e))
    s.update_geometry()
    s.update_force_state()
    full = np.nonzero(s.geti("nb") == 18)[0]
    assert len(full) == 27
    return full, 4.5 * K_BULK * e * e


def self_only_nonlocal(s):
    """Nonlocal lists that hold only the particle itself (cutoff below the lattice spacing), so
    that the nonlocal rate equals the local one: phi(0) V_m / (phi(0) V_m), exact for
    power-of-two rates."""
    s.search_nonlocal(0.5)
    nlp = s.geti("nlptr")
    assert np.all(np.diff(nlp) == 1)


# ------------------------------------------------------------------ ElasticDamage rate law
def test_elastic_damage_rate_law_closed_form():
    """updateParticleStateVariables (particle_elastic_damage.h:150-161):
    U = energy; delta = U - U_last; f = U - k1 D - k0; Ddot = delta / k1 if f > 0 and delta > 0."""
    s, xyz = sc5(O.ELASTIC_DAMAGE)
    k0, k1 = 0.25, 4.0
    s.mode(elmap_literal=0)
    s.set_elastic_damage(1.0, False, E, NU, k0, k1, 10.0, 0.99)
    e = 1e-3
    full, U = dilate(s, xyz, e)
    assert U == pytest.approx(0.75, rel=1e-12)
    # per-particle cases: (U_last, D, expected Ddot)
    cases = [
        (0.3, 0.05, (U - 0.3) / k1),   # f = 0.3 > 0, delta > 0: Ddot = delta / k1 = 0.1125
        (0.3, 0.0, (U - 0.3) / k1),    # D does not enter the rate itself
        (0.3, 0.15, 0.0),              # f = U - 0.6 - 0.25 < 0 although U - k0 > 0 (k1 D term)
        (0.9, 0.0, 0.0),               # delta < 0 although f > 0
        (0.0, 0.0, U / k1),            # from zero: Ddot = U / k1 = 0.1875
    ]
    sv_last = np.zeros(7 * len(xyz))
    D = np.zeros(len(xyz))
    for (ul, d, _), i in zip(cases, full):
        sv_last[7 * i] = ul
        D[i] = d
    s.setd("state_var_last", sv_last)
    s.setd("damage", D)
    s.update_state_var()
    sv = s.getd("state_var").reshape(-1, 7)
    dd = s.getd("Ddot_local")
    for (ul, d, exp), i in zip(cases, full):
        assert sv[i, 0] == pytest.approx(U, rel=1e-10)
        if exp == 0.0:
            assert dd[i] == 0.0, (ul, d, dd[i])
        else:
            assert dd[i] == pytest.approx(exp, rel=1e-9), (ul, d, dd[i], exp)


def test_elastic_damage_surface_is_strict():
    """f = U - k1 D - k0 must be > 0 (particle_elastic_damage.h:157): with k0 = U - k1 D to the
    last bit the rate is 0, one ulp less k0 gives delta / k1."""
    s, xyz = sc5(O.ELASTIC_DAMAGE)
    s.mode(elmap_literal=0)
    e = 1e-3
    s.set_elastic_damage(1.0, False, E, NU, 0.0, 2.0, 10.0, 0.99)
    full, _ = dilate(s, xyz, e)
    s.update_state_var()
    U = s.getd("state_var").reshape(-1, 7)[full[0], 0]
    i = full[0]
    for k0, expect_zero in ((U, True), (np.nextafter(U, 0.0), False)):
        s.set_elastic_damage(1.0, False, E, NU, k0, 2.0, 10.0, 0.99, pid=int(i))
        s.setd("damage", np.zeros(len(xyz)))
        s.setd("state_var_last", np.zeros(7 * len(xyz)))
        s.update_state_var()
        d = s.getd("Ddot_local")[i]
        if expect_zero:
            assert d == 0.0
        else:
            assert d == U / 2.0


# ------------------------------------------------------------------ damage accumulation (a22, a23)
@pytest.mark.parametrize("ptype", [O.ELASTIC_DAMAGE, O.J2])
def test_damage_accumulation_threshold_and_change_flag(ptype):
    """updateParticleStaticDamage (particle_elastic_damage.h:163-173, particle_j2plasticity.h:
    142-152): D += Ddot_nonlocal; D >= threshold -> 1; returns |D - D_last| > 1e-3."""
    thr = 0.5

    def run(D0, Dlast, rate):
        s, xyz = sc5(ptype)
        s.mode(elmap_literal=0)
        if ptype == O.ELASTIC_DAMAGE:
            s.set_elastic_damage(1.0, False, E, NU, 1e9, 1.0, 10.0, thr)
        else:
            s.set_j2(1.0, False, E, NU, 300.0, 0.0, 2e3, 0.0, 1.0)
            s.set_j2_damage_threshold(thr)
        self_only_nonlocal(s)
        n = len(xyz)
        s.setd("damage", np.full(n, D0))
        s.setd("damage_last", np.full(n, Dlast))
        s.setd("Ddot_local", np.full(n, rate))
        changed = s.update_static_damage(-1)
        return changed, s.getd("damage"), s.getd("Ddot_nonlocal")

    ch, D, dn = run(0.25, 0.25, 0.125)
    assert np.all(dn == 0.125)
    assert np.all(D == 0.375) and ch                      # plain accumulation, flag set
    ch, D, _ = run(0.25, 0.25, 0.25)
    assert np.all(D == 1.0) and ch                        # 0.5 reaches the threshold: '>='
    ch, D, _ = run(0.25, 0.25, 0.125 + 2.0 ** -10)
    assert np.all(D == 0.375 + 2.0 ** -10)                # just below the threshold stays
    ch, D, _ = run(0.25, 0.25, 2.0 ** -11)
    assert np.all(D == 0.25 + 2.0 ** -11) and not ch      # |D - D_last| = 4.9e-4 <= 1e-3
    ch, D, _ = run(0.25, 0.25 - 2.0 ** -9, 2.0 ** -11)
    assert ch                                             # compared with D_last, not the old D
    ch, D, _ = run(0.25, 0.25, 2.0 ** -9)
    assert ch                                             # 1.95e-3 > 1e-3


def test_hcf_has_no_static_damage():
    """ParticleFatigueHCF keeps the base class's updateParticleStaticDamage (particle.h:63):
    the static damage update leaves damage alone and reports no change."""
    s, xyz = sc5(O.HCF)
    s.mode(elmap_literal=0)
    s.set_hcf(1.0, False, E, NU, 1e-3, 2.0, 0.5)
    self_only_nonlocal(s)
    s.setd("damage", np.full(len(xyz), 0.25))
    s.setd("Ddot_local", np.full(len(xyz), 0.125))
    assert not s.update_static_damage(-1)
    assert np.all(s.getd("damage") == 0.25)


# ------------------------------------------------------------------ HCF rate law and update (a24)
def test_hcf_rate_law_closed_form():
    """ParticleFatigueHCF::updateParticleStateVariables (particle_fatigue_hcf.h:127-145):
    sv1 = min(U, sv1), sv2 = max(U, sv2), sv0 = max(sv2, sv0), Ddot = A sv0^B (sv2 - sv1)^d."""
    A, B, d = 3e-3, 2.019, 0.5
    s, xyz = sc5(O.HCF)
    s.mode(elmap_literal=0)
    s.set_hcf(1.0, False, E, NU, A, B, d)
    full, U = dilate(s, xyz, 1e-3)
    # (sv0, sv1, sv2) before -> expected (sv0, sv1, sv2) after and Ddot
    cases = [
        ((10.0, 0.5, 0.6), (10.0, 0.5, U), A * 10.0 ** B * (U - 0.5) ** d),   # sv0 kept: it is the max
        ((0.7, 0.1, 0.3), (U, 0.1, U), A * U ** B * (U - 0.1) ** d),           # sv0 <- sv2
        ((0.2, 2147483647.0, 0.0), (U, U, U), 0.0),                             # first visit: no range
        ((2.0, 0.9, 1.5), (2.0, U, 1.5), A * 2.0 ** B * (1.5 - U) ** d),       # U is the new minimum
    ]
    sv = np.zeros((len(xyz), 7))
    sv[:, 1] = 2147483647.0
    for (old, _, _), i in zip(cases, full):
        sv[i, :3] = old
    s.setd("state_var", sv.ravel())
    s.update_state_var()
    out = s.getd("state_var").reshape(-1, 7)
    dd = s.getd("Ddot_local")
    for (old, new, rate), i in zip(cases, full):
        np.testing.assert_allclose(out[i, :3], new, rtol=1e-10)
        if rate == 0.0:
            assert dd[i] == 0.0
        else:
            assert dd[i] == pytest.approx(rate, rel=1e-9), (old, dd[i], rate)
    # the exponents are not interchangeable: B and d swapped would give this instead
    swapped = A * 10.0 ** d * (U - 0.5) ** B
    assert abs(swapped - cases[0][2]) > 0.5 * cases[0][2]


def test_hcf_fatigue_update_dn_threshold_and_rearm():
    """ParticleFatigueHCF::updateParticleFatigueDamage (particle_fatigue_hcf.h:147-160) through
    SolverFatigue::updateFatigueDamage (solver_fatigue.h:75-126): D += Ddot_nonlocal * dN;
    D >= threshold -> 1; sv1 = INT_MAX, sv2 = 0 (sv0 kept); returns |D - D_last| > 1e-3.
    Particles of the undamaged type are skipped entirely (no damage, no re-arm)."""
    s, xyz = sc5(O.HCF)
    n = len(xyz)
    s.mode(elmap_literal=0)
    s.set_hcf(1.0, False, E, NU, 1e-3, 2.0, 0.5, dthr=0.25)
    self_only_nonlocal(s)
    typ = np.zeros(n, dtype=np.int32)
    typ[:5] = 1
    s.seti("type", typ)
    sv = np.zeros((n, 7))
    sv[:, 0], sv[:, 1], sv[:, 2] = 3.0, 0.5, 0.75
    s.setd("state_var", sv.ravel())
    D0 = np.full(n, 0.125)
    D0[10:20] = 0.25 - 7 * 2.0 ** -6    # reaches the threshold 0.25 exactly with dN = 7
    s.setd("damage", D0)
    s.setd("damage_last", D0)
    rate = np.full(n, 2.0 ** -10)
    rate[10:20] = 2.0 ** -6
    s.setd("Ddot_local", rate)
    changed = s.update_fatigue_damage(7.0, 1)
    D = s.getd("damage")
    out = s.getd("state_var").reshape(-1, 7)
    assert np.all(D[:5] == 0.125)                                   # undamaged type untouched
    assert np.all(out[:5, 1] == 0.5) and np.all(out[:5, 2] == 0.75)  # ... and not re-armed
    assert np.all(D[20:] == 0.125 + 7 * 2.0 ** -10)                 # dN multiplies the rate
    assert np.all(D[10:20] == 1.0)                                  # '>=' the threshold
    assert np.all(out[5:, 1] == 2147483647.0) and np.all(out[5:, 2] == 0.0)
    assert np.all(out[:, 0] == 3.0)                                 # sv0 (the max) is kept
    assert changed                                                  # the clamped ones moved by > 1e-3
    # without the clamped particles: 7 * 2^-10 = 6.8e-3 > 1e-3 changes; 2^-10 alone does not
    s.setd("damage", np.full(n, 0.125))
    s.setd("damage_last", np.full(n, 0.125))
    s.setd("Ddot_local", np.full(n, 2.0 ** -10))
    assert not s.update_fatigue_damage(1.0, 1)
    assert np.all(s.getd("damage")[5:] == 0.125 + 2.0 ** -10)


# ------------------------------------------------------------------ cycle driver (a25)
def hex_plate(nx=8, ny=8, r=0.5):
    box = [0.0, 2 * r * nx, 0.0, 2 * r * math.sqrt(3) / 2 * ny, 0.0, 1.0]
    return O.generate(O.HEXAGON2D, box, r)


def cyclic_groups(xyz):
    x, y = xyz[:, 0], xyz[:, 1]
    left = np.nonzero(x < x.min() + 1e-6)[0]
    top = np.nonzero(y > y.max() - 1e-6)[0]
    bottom = np.nonzero(y < y.min() + 1e-6)[0]
    return left, top, bottom


SPECTRUM = np.array([0.0] + sum(([10.0 + k, 1.0 + 0.5 * k] for k in range(40)), []))  # low, (hi, low) x 40


@pytest.mark.parametrize("tmode,tau,visits", [
    (0, 1.0, list(range(40))),              # dN = int(1 / 1) = 1
    (0, 0.5, list(range(0, 40, 2))),        # dN = 2
    (0, 0.3, list(range(0, 40, 3))),        # dN = int(3.33) = 3
    (0, 0.05, [0, 20]),                     # dN = 20: cycles 0 and 20, then N = 40 ends the run
    # exponential map dN = int(1/tau exp(t/tau)) with t = 0, 1, 2, ...: tau = 2 gives
    # int(0.5) = 0, int(0.82) = 0, int(1.36) = 1, int(2.24) = 2, int(3.69) = 3, int(6.09) = 6,
    # int(10.04) = 10, int(16.56) = 16 -> N = 0, 0, 0, 1, 3, 6, 12, 22, then 38 + ... ends
    (1, 2.0, [0, 0, 0, 1, 3, 6, 12, 22, 38]),
])
def test_cycle_driver_visits_and_load_sequence(tmode, tau, visits):
    """SolverFatigue::solveProblemCyclic + generateLoadCycle + solveProblemOneCycle
    (solver_fatigue.h:128-241): cycle N applies f[2N] -> f[2N+1] -> f[2N+2] in n_interval equal
    substeps each way, then N += dN. With LoadUniaxialForce (load_step_fatigue.h:22-35) the top
    group's Pex is the sum of the applied increments, so it names the visited cycles."""
    xyz = hex_plate()
    s = O.System(O.HEXAGON2D, 0.5, O.HCF, xyz)
    s.mode(elmap_literal=0, dirichlet_static=1, cg_tol=1e-12, tol_NR=1e-8, max_NR_iter=50)
    s.set_hcf(1.0, True, 110e3, 0.34, 0.0, 2.019, 0.5)   # A = 0: no damage, elastic cycling
    s.update_geometry()
    s.update_force_state()
    left, top, bottom = cyclic_groups(xyz)
    n_interval = 2
    N = s.solve_problem_cyclic(1, n_interval, SPECTRUM, [left, top, bottom], tau, -1, tmode=tmode)
    n_cycles = 0.5 * (len(SPECTRUM) - 1)
    # the final N is the first multiple past the last visit
    assert N >= n_cycles and visits[-1] < n_cycles
    applied = sum(SPECTRUM[2 * v + 2] - SPECTRUM[2 * v] for v in visits)
    pex = s.getd("Pex").reshape(-1, 3)
    np.testing.assert_allclose(pex[top, 1].sum(), applied, rtol=1e-12, atol=1e-12)
    assert np.all(pex[np.setdiff1d(np.arange(len(xyz)), top)] == 0.0)
    # 2 n_interval substeps per visited cycle, each one Newton solve
    assert s.info("newton_log_len") == 2 * n_interval * len(visits)
    cons = s.geti("cons").reshape(-1, 3)
    assert np.all(cons[left, 0] == 1) and np.all(cons[bottom, 1] == 1)


@pytest.mark.parametrize("ltype", [0, 1, 2, 3])
def test_fatigue_load_templates(ltype):
    """The four load templates (load_step_fatigue.h:22-85) through the cycle driver, one cycle
    (spectrum 0 -> 3 -> 1, n_interval 2, tau 1): which groups are constrained on which axes, which
    move, and which carry the force. Net increment per cycle f[2] - f[0] = 1."""
    xyz = hex_plate()
    s = O.System(O.HEXAGON2D, 0.5, O.HCF, xyz)
    s.mode(elmap_literal=0, dirichlet_static=1, cg_tol=1e-12, tol_NR=1e-8, max_NR_iter=50)
    s.set_hcf(1.0, True, 110e3, 0.34, 0.0, 2.019, 0.5)
    s.update_geometry()
    s.update_force_state()
    left, top, bottom = cyclic_groups(xyz)
    groups = [top, bottom] if ltype == 2 else [left, top, bottom]
    if ltype == 3:
        mid = np.nonzero(np.abs(xyz[:, 1] - np.median(xyz[:, 1])) < 0.3)[0]
        groups = [mid, top, bottom]
    spec = [0.0, 3e-3, 1e-3] if ltype in (0, 2) else [0.0, 30.0, 10.0]
    N = s.solve_problem_cyclic(ltype, 2, spec, groups, 1.0, -1)
    assert N == 1
    net = spec[2] - spec[0]
    x = s.getd("xyz").reshape(-1, 3)
    pex = s.getd("Pex").reshape(-1, 3)
    cons = s.geti("cons").reshape(-1, 3)
    if ltype == 0:    # LoadUniaxialDisp: left x 0, top y += dy, bottom y 0
        np.testing.assert_allclose(x[top, 1] - xyz[top, 1], net, rtol=1e-9)
        assert np.all(cons[left, 0] == 1) and np.all(cons[top, 1] == 1) and np.all(cons[bottom, 1] == 1)
        assert np.all(x[bottom, 1] == xyz[bottom, 1]) and np.all(pex == 0)
    elif ltype == 1:  # LoadUniaxialForce: left x 0, top Pex_y += fy / |top|, bottom y 0
        np.testing.assert_allclose(pex[top, 1], net / len(top), rtol=1e-12)
        assert np.all(cons[left, 0] == 1) and np.all(cons[bottom, 1] == 1)
        assert np.all(cons[top, 1] == 0) and np.all(cons[np.setdiff1d(top, left)] == 0)
    elif ltype == 2:  # LoadDogBoneDisp(0, y, 0): group1 x, y += dy, z; group2 x, y, z fixed
        np.testing.assert_allclose(x[top, 1] - xyz[top, 1], net, rtol=1e-9)
        assert np.all(x[top, 0] == xyz[top, 0]) and np.all(x[bottom] == xyz[bottom])
        assert np.all(cons[top] == 1) and np.all(cons[bottom] == 1) and np.all(pex == 0)
    else:             # LoadCTForce: mid y 0, top and bottom x 0, top +fy, bottom -fy
        g0, g1, g2 = groups
        np.testing.assert_allclose(pex[g1, 1], net / len(g1), rtol=1e-12)
        np.testing.assert_allclose(pex[g2, 1], -net / len(g2), rtol=1e-12)
        assert np.all(cons[g0, 1] == 1) and np.all(cons[g1, 0] == 1) and np.all(cons[g2, 0] == 1)
        assert np.all(cons[g1, 1] == 0) and np.all(cons[g2, 1] == 0)


# ------------------------------------------------------------------ static halving driver (a19)
def _halving_setup():
    """A clamped SC 4^3 bar at equilibrium under a first top displacement. The step under test
    (top displacement D plus a small force F on the second layer) then runs with max_NR_iter = 2:
    'if (++ni > max_NR_iter) return max_NR_iter' and solveProblemStep's 'n_newton >=
    max_NR_iter' (solver.h:72-73, solver_static.h:119-120) make a step pass only when one
    Newton iteration meets the test."""
    n, r = 4, 0.5
    L = 2 * r * (n - 1)
    a = math.sqrt(3 * L * L)
    off = (-a) % (2 * r)
    box = [0, L, off, off + L, off, off + L]
    xyz = O.generate(O.SC3D, box, r)
    s = O.System(O.SC3D, r, O.ELASTIC, xyz, box)
    s.mode(elmap_literal=0, dirichlet_static=1, cg_tol=1e-14, tol_NR=1e-11, max_NR_iter=50, cg_transpose=0,
           diag10_literal=0)
    s.set_elastic(0.0, False, E, NU, 1.0)
    z = xyz[:, 2]
    bottom = np.nonzero(z < z.min() + 1e-6)[0]
    top = np.nonzero(z > z.max() - 1e-6)[0]
    mid = np.nonzero(np.abs(z - np.sort(np.unique(z))[1]) < 1e-6)[0]
    s.update_geometry()
    s.update_force_state()
    pre = O.LoadStep()
    for ax in "xyz":
        pre.add_disp(bottom, ax, 0.0)
    pre.add_disp(top, "z", 0.02)
    assert s.solve_step_static(pre, -1)
    return s, xyz, bottom, top, mid


def _ratio_after_one_iteration(s, step):
    """The Newton test ratio ||r_1|| / max(||r_0||, ||R_0||) after one iteration of the step
    (solver.h:53-89 with the assembly of solver_static.h:110-113), from the current state."""
    s.apply_bc(step)
    s.calc_stiffness()
    s.update_stiffness_disp_bc()
    s.update_geometry()
    s.update_force_state()
    s.update_rr()
    mult = max(s.norm("residual"), s.norm("reaction"))
    s.solve_linear_system()
    s.update_geometry()
    s.update_force_state()
    s.update_rr()
    return s.norm("residual") / mult


def _halving_run(rollback):
    D, F = 0.2, 1e-6
    s, xyz, bottom, top, mid = _halving_setup()

    def step(scale):
        ls = O.LoadStep()
        for ax in "xyz":
            ls.add_disp(bottom, ax, 0.0)
        ls.add_disp(top, "z", D * scale)
        ls.add_force(mid, 0.0, 0.0, F * scale)
        return ls

    x_conv = s.getd("xyz").copy()
    pex0 = s.getd("Pex").copy()
    # one-iteration ratios of D, and of D/2 twice in a row (the second from the first's result)
    rD = _ratio_after_one_iteration(s, step(1.0))
    s.reset(True)
    s.setd("Pex", pex0)
    r1 = _ratio_after_one_iteration(s, step(0.5))
    r2 = _ratio_after_one_iteration(s, step(0.5))
    s.reset(True)
    s.setd("Pex", pex0)
    # resetStateVar restores positions and state, not the bond geometry (particle.h:128-148):
    # recompute it so that the run starts where the measurement of D started
    s.update_geometry()
    s.update_force_state()
    assert max(r1, r2) < 0.5 * rD, (rD, r1, r2)
    tol = math.sqrt(max(r1, r2) * rD)   # D needs two iterations, each D/2 one
    s.mode(max_NR_iter=2, tol_NR=tol, rollback_pex=int(rollback))
    s.solve_problem_static([step(1.0)], -1)
    return s, x_conv, pex0, top, mid, D, F


def test_static_halving_sequence_and_pex():
    """SolverStatic::solveProblem (solver_static.h:64-95): a failing step is reset to the last
    converged state (positions included), halved, and a copy of the halved step is inserted, so
    [D] runs as D (fails), D/2, D/2. Pex is not rolled back (Q7): the force of the failed attempt
    stays, so F + F/2 + F/2 = 2F ends up applied."""
    s, x_conv, pex0, top, mid, D, F = _halving_run(False)
    assert s.info("n_steps") == 2 and s.info("n_restarts") == 1
    assert list(s.geti("newton_log")[-3:]) == [2, 1, 1]   # D failed (max_NR_iter), D/2 twice in 1
    x = s.getd("xyz").reshape(-1, 3)
    xc = x_conv.reshape(-1, 3)
    np.testing.assert_allclose(x[top, 2] - xc[top, 2], D, rtol=1e-13)   # D/2 + D/2 from the reset start
    pex = s.getd("Pex").reshape(-1, 3)
    np.testing.assert_allclose(pex[mid, 2].sum() - pex0.reshape(-1, 3)[mid, 2].sum(), 2 * F, rtol=1e-12)


def test_static_halving_with_pex_rollback():
    """The same run with the Q7 switch rollback_pex=1: the failed attempt's force is removed, so
    F/2 + F/2 = F is applied; the step list is unchanged."""
    s, x_conv, pex0, top, mid, D, F = _halving_run(True)
    assert s.info("n_steps") == 2
    assert list(s.geti("newton_log")[-3:]) == [2, 1, 1]
    pex = s.getd("Pex").reshape(-1, 3)
    np.testing.assert_allclose(pex[mid, 2].sum() - pex0.reshape(-1, 3)[mid, 2].sum(), F, rtol=1e-12)
