"""Adjoints of general linear objectives and over checkpointed load paths, on the GPU, against
central finite differences of the FP64 oracle:
- load-step checkpoints: a restored state replays the next step bit for bit;
- G = w_x . x (a displacement field) and G = reaction: dG/dKn, dG/dTv and dG/dbdamage
  (per layer, and one bond's damage);
- the torch.autograd.Function of the C++ harness (lpm_autograd.load_path) over a three-step
  path with G_k = reaction_k + w_x . x_k: Kn.grad, Tv.grad and bdamage.grad against the
  oracle's differences of the whole path."""
import sys

import numpy as np
import pytest

import oracle as O
from scenarios import Pair, faces, hex_plate, rel

pytestmark = pytest.mark.gpu

STEP = 0.02  # top edge displacement per load step


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _plate():
    xyz, box = hex_plate(10, 10)
    p = Pair(O.HEXAGON2D, 0.5, O.ELASTIC, xyz, box)
    p.set_elastic(0.0, True, 69e3, 0.3, 1.0)
    lo, hi = faces(xyz, 1)
    g = p.g
    g.mode(dirichlet_static=1, cg_transpose=0, diag10_literal=0, precision=1, cg_tol=1e-13, tol_NR=1e-12,
           max_NR_iter=60, cg_maxit=20000)
    ls = g.load_step(disp=[(lo, "x", 0.0), (lo, "y", 0.0), (hi, "y", STEP)])
    g.update_geometry()
    g.update_force_state()
    return p, xyz, lo, hi, ls


def _weights(xyz, lo, hi, seed=5):
    """w_pin: the top edge's y reaction; w_x: random weights on the free DoFs (x of the top
    edge is free)."""
    n = len(xyz)
    w_pin = np.zeros(2 * n)
    w_pin[2 * hi + 1] = 1.0
    rng = np.random.default_rng(seed)
    w_x = rng.standard_normal(2 * n)
    w_x[2 * lo] = w_x[2 * lo + 1] = 0.0
    w_x[2 * hi + 1] = 0.0
    return w_pin, w_x


def _oracle_path(xyz, box, lo, hi, nsteps, w_pin, w_x, coef, kn_scale=(1.0, 1.0), tv_scale=1.0, bd=None):
    """sum_k coef[k] (w_pin . Pin_k + w_x . u_k) over nsteps oracle load steps from rest, with the
    displacements u = x - x0 (the same derivatives as positions, without their rounding);
    add sum_k coef[k] w_x . x0 for the value of the position objective."""
    s = O.System(O.HEXAGON2D, 0.5, O.ELASTIC, xyz, box)
    s.mode(elmap_literal=0, dirichlet_static=1, cg_tol=1e-14, tol_NR=1e-12, max_NR_iter=60)
    s.set_elastic(0.0, True, 69e3, 0.3, 1.0)
    layer = s.geti("layer")
    s.setd("Kn", s.getd("Kn") * np.where(layer == 0, kn_scale[0], kn_scale[1]))
    s.setd("Tv", s.getd("Tv") * tv_scale)
    if bd is not None:
        s.setd("bdamage", bd)
    s.update_geometry()
    s.update_force_state()
    tot = 0.0
    for k in range(nsteps):
        ls = O.LoadStep().add_disp(lo, "x", 0.0).add_disp(lo, "y", 0.0).add_disp(hi, "y", STEP)
        assert s.solve_step_static(ls, -1)
        pin = s.getd("Pin").reshape(-1, 3)[:, :2].ravel()
        u = (s.getd("xyz").reshape(-1, 3) - xyz)[:, :2].ravel()
        tot += coef[k] * (w_pin @ pin + w_x @ u)
    return tot


def test_force_vjp_is_the_transpose():
    """<w, J v> = <J^T w, v> for the exact force Jacobian where it is not symmetric: random
    partial damage per directed bond and five bonds broken without their reverse."""
    say("[gpu] force vjp: transpose of the jvp")
    p, xyz, lo, hi, ls = _plate()
    g = p.g
    rng = np.random.default_rng(2)
    nb = g.size("nbonds")
    bd = np.where(rng.random(nb) < 0.2, 0.8 * rng.random(nb), 0.0)
    bd[rng.integers(0, nb, 5)] = 1.0
    g.set("bdamage", bd)
    x = xyz + 2e-3 * rng.standard_normal(xyz.shape)
    x[:, 2] = xyz[:, 2]
    g.set("xyz", x)
    g.update_geometry()
    g.update_force_state()
    nd = 2 * len(xyz)
    for _ in range(3):
        v, w = rng.standard_normal(nd), rng.standard_normal(nd)
        a, b = w @ g.force_jvp(v), g.force_vjp(w) @ v
        asym = abs(w @ g.force_jvp(v) - v @ g.force_jvp(w))
        say(f"   <w,Jv> {a:.12e} <J^T w,v> {b:.12e} rel {abs(a - b) / abs(a):.1e}; <w,Jv> - <v,Jw> {asym:.3e}")
        assert abs(a - b) <= 1e-12 * abs(a)
        assert asym > 1e-4 * abs(a)  # the case is not symmetric


def test_checkpoint_restore_replays_bitwise():
    say("[gpu] checkpoints: save / restore / replay")
    p, xyz, lo, hi, ls = _plate()
    g = p.g
    g.checkpoint_save(0)
    x0 = g.get("xyz")
    states = []
    for k in range(3):
        assert g.solve_step_static(ls, -1)
        g.checkpoint_save(k + 1)
        states.append((g.get("xyz"), g.get("Pin"), g.get("bforce")))
    g.checkpoint_restore(2)
    np.testing.assert_array_equal(g.get("xyz"), states[1][0])
    assert g.solve_step_static(ls, -1)
    for a, b in zip((g.get("xyz"), g.get("Pin"), g.get("bforce")), states[2]):
        np.testing.assert_array_equal(a, b)
    g.checkpoint_restore(0)
    np.testing.assert_array_equal(g.get("xyz"), x0)
    with pytest.raises(Exception, match="empty"):
        g.checkpoint_restore(7)
    g.checkpoint_release()
    with pytest.raises(Exception, match="empty"):
        g.checkpoint_restore(1)


def test_adjoint_linear_displacement_field_and_damage():
    say("[gpu] adjoint of a displacement field and of the reaction, with bond damage")
    p, xyz, lo, hi, ls = _plate()
    g = p.g
    # a partly damaged bond pair, so that the damage derivative is taken away from zero too
    bd0 = np.zeros(g.size("nbonds"))
    opp = g.get("opp")
    b0 = int(np.nonzero(p.o.geti("layer") == 0)[0][len(xyz) // 2 * 3])
    bd0[b0] = bd0[opp[b0]] = 0.3
    g.set("bdamage", bd0)
    g.update_geometry()
    g.update_force_state()
    assert g.solve_step_static(ls, -1)
    w_pin, w_x = _weights(xyz, lo, hi)
    layer = p.o.geti("layer")
    kn, tv = g.get("Kn"), g.get("Tv")
    for name, wp, wx in (("displacement field", None, w_x), ("reaction", w_pin, None)):
        G, gKn, gTv, gBd, nref, rres = g.adjoint_linear(w_pin=wp, w_x=wx)
        say(f"   {name}: G {G:.10e} refinements {nref} rel residual {rres:.1e}")
        assert rres < 1e-9
        wpo = np.zeros_like(w_pin) if wp is None else wp
        wxo = np.zeros_like(w_x) if wx is None else wx
        G0 = _oracle_path(xyz, p.box, lo, hi, 1, wpo, wxo, [1.0], bd=bd0) + wxo @ xyz[:, :2].ravel()
        assert G == pytest.approx(G0, rel=1e-9, abs=1e-9 * np.abs(wxo).sum())
        # homogeneity in (Kn, Tv) at fixed boundary displacements: degree 0 for positions,
        # degree 1 for reactions (Euler's identity)
        euler = np.sum(gKn * kn) + np.sum(gTv * tv)
        scale = np.sum(np.abs(gKn * kn)) + np.sum(np.abs(gTv * tv))
        say(f"   Euler: sum(dG/dtheta theta) {euler:.6e} (scale {scale:.3e})")
        if wp is None:
            assert abs(euler) <= 1e-7 * scale
        else:
            assert abs(euler - G) <= 1e-7 * abs(G)
        # the displacement field moves by ~1e-3 of a spacing per unit of relative stiffness: steps
        # of 1e-3 keep the oracle's round-off small beside the difference (truncation ~1e-6)
        h, hb = (1e-5, 1e-4) if wp is not None else (1e-3, 1e-3)

        def bump(bond, dh):
            v = bd0.copy()
            v[bond] += dh
            return v

        b = int(np.argmax(np.abs(gBd)))
        cases = {
            "Kn layer 0": (np.sum(gKn[layer == 0] * kn[layer == 0]), dict(kn_scale=(1 + h, 1.0), bd=bd0),
                           dict(kn_scale=(1 - h, 1.0), bd=bd0), h),
            "Tv": (np.sum(gTv * tv), dict(tv_scale=1 + h, bd=bd0), dict(tv_scale=1 - h, bd=bd0), h),
            f"bdamage of bond {b}": (gBd[b], dict(bd=bump(b, hb)), dict(bd=bump(b, -hb)), hb),
            f"bdamage of bond {b0} (d = 0.3)": (gBd[b0], dict(bd=bump(b0, hb)), dict(bd=bump(b0, -hb)), hb),
        }
        for cname, (adj, up, dn, hh) in cases.items():
            fd = (_oracle_path(xyz, p.box, lo, hi, 1, wpo, wxo, [1.0], **up) -
                  _oracle_path(xyz, p.box, lo, hi, 1, wpo, wxo, [1.0], **dn)) / (2 * hh)
            err = abs(adj - fd) / abs(fd)
            say(f"   {name}, {cname}: adjoint {adj:.10e} fd {fd:.10e} rel {err:.2e}")
            assert err < 1e-5


def test_autograd_load_path_matches_oracle_differences():
    torch = pytest.importorskip("torch")
    import lpm_autograd as A

    say("[gpu] torch.autograd.Function over a 3-step load path")
    p, xyz, lo, hi, ls = _plate()
    g = p.g
    g.checkpoint_save(0)
    w_pin, w_x = _weights(xyz, lo, hi)
    coef = [0.3, -0.5, 1.0]
    kn = torch.tensor(g.get("Kn"), dtype=torch.float64, requires_grad=True)
    tv = torch.tensor(g.get("Tv"), dtype=torch.float64, requires_grad=True)
    bd = torch.zeros(g.size("nbonds"), dtype=torch.float64, requires_grad=True)
    steps = [ls.address() for _ in range(3)]
    G = A.load_path(kn, tv, bd, g.handle, steps, 0, torch.tensor(w_pin), torch.tensor(w_x))
    x_end = g.get("xyz")
    loss = (G * torch.tensor(coef, dtype=torch.float64)).sum()
    loss.backward()
    np.testing.assert_array_equal(g.get("xyz"), x_end)  # backward leaves the end state
    Go = [_oracle_path(xyz, p.box, lo, hi, k + 1, w_pin, w_x, [0.0] * k + [1.0]) + w_x @ xyz[:, :2].ravel()
          for k in range(3)]
    say("   G", G.detach().numpy(), "oracle", np.array(Go))
    assert rel(G.detach().numpy(), np.array(Go)) < 1e-9
    layer = p.o.geti("layer")
    knv, tvv = kn.detach().numpy(), tv.detach().numpy()
    gk, gt, gb = kn.grad.numpy(), tv.grad.numpy(), bd.grad.numpy()
    h, hb = 1e-5, 1e-4
    b = int(np.argmax(np.abs(gb)))
    bd_up, bd_dn = np.zeros(len(gb)), np.zeros(len(gb))
    bd_up[b], bd_dn[b] = hb, -hb
    cases = {
        "Kn layer 1": (np.sum(gk[layer == 1] * knv[layer == 1]), dict(kn_scale=(1.0, 1 + h)), dict(kn_scale=(1.0, 1 - h)), h),
        "Tv": (np.sum(gt * tvv), dict(tv_scale=1 + h), dict(tv_scale=1 - h), h),
        f"bdamage of bond {b}": (gb[b], dict(bd=bd_up), dict(bd=bd_dn), hb),
    }
    for name, (adj, up, dn, hh) in cases.items():
        fd = (_oracle_path(xyz, p.box, lo, hi, 3, w_pin, w_x, coef, **up) -
              _oracle_path(xyz, p.box, lo, hi, 3, w_pin, w_x, coef, **dn)) / (2 * hh)
        err = abs(adj - fd) / abs(fd)
        say(f"   {name}: autograd {adj:.10e} fd {fd:.10e} rel {err:.2e}")
        assert err < 1e-5
    g.checkpoint_release()
