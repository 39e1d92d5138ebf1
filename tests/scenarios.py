"""Shared test scenarios: the same particle system, material and loads set up on the
oracle (FP64 CPU restatement) and on the HIP library, so that tests compare like
with like. Small sizes (the oracle must finish in seconds)."""
import math

import numpy as np

import oracle as O

try:
    import lpmgpu as G
except Exception:  # the CPU suite can run without the product library
    G = None


def sc_cube(n=6, r=0.5):
    """SC lattice with n particles per side; the box is shifted so that the generator's
    offset (rows start at -a, utilities.h:174-195) keeps n rows in y and z."""
    L = 2 * r * (n - 1)
    a = math.sqrt(3 * L * L)
    off = (-a) % (2 * r)
    box = [0, L, off, off + L, off, off + L]
    xyz = O.generate(O.SC3D, box, r)
    assert len(xyz) == n ** 3
    return xyz, box


def hex_plate(nx=12, ny=12, r=0.5):
    box = [0.0, 2 * r * nx, 0.0, 2 * r * math.sqrt(3) / 2 * ny, 0.0, 1.0]
    xyz = O.generate(O.HEXAGON2D, box, r)
    return xyz, box


def faces(xyz, axis, tol=1e-6):
    c = xyz[:, axis]
    return np.nonzero(c < c.min() + tol)[0], np.nonzero(c > c.max() - tol)[0]


class Pair:
    """The same scenario on the oracle (self.o) and on the GPU (self.g)."""

    def __init__(self, lattice, r, model, xyz, box, gpu=True):
        self.lattice, self.r, self.model = lattice, r, model
        self.xyz, self.box = xyz, box
        self.o = O.System(lattice, r, model, xyz, box)
        self.g = G.Context(lattice, r, model, xyz) if gpu else None
        self.steps_o, self.steps_g = [], []

    def both(self, f):
        f(self.o)
        if self.g is not None:
            f(self.g)

    def set_elastic(self, L, ps, E, mu, crit, literal=0):
        self.o.mode(elmap_literal=literal)
        self.o.set_elastic(L, ps, E, mu, crit)
        if self.g is not None:
            self.g.set_elastic(L, ps, E, mu, crit, elmap_literal=literal)

    def set_elastic_damage(self, L, ps, E, mu, k0, k1, ct, dthr, literal=0):
        self.o.mode(elmap_literal=literal)
        self.o.set_elastic_damage(L, ps, E, mu, k0, k1, ct, dthr)
        if self.g is not None:
            self.g.set_elastic_damage(L, ps, E, mu, k0, k1, ct, dthr, elmap_literal=literal)

    def set_j2(self, L, ps, E, mu, sy, xi, H, A, crit, thr, literal=0):
        self.o.mode(elmap_literal=literal)
        self.o.set_j2(L, ps, E, mu, sy, xi, H, A, crit)
        self.o.set_j2_damage_threshold(thr)
        if self.g is not None:
            self.g.set_j2(L, ps, E, mu, sy, xi, H, A, crit, damage_threshold=thr, elmap_literal=literal)

    def set_hcf(self, L, ps, E, mu, A, B, d, literal=0):
        self.o.mode(elmap_literal=literal)
        self.o.set_hcf(L, ps, E, mu, A, B, d)
        if self.g is not None:
            self.g.set_hcf(L, ps, E, mu, A, B, d, elmap_literal=literal)

    def set_type(self, t):
        t = np.asarray(t, dtype=np.int32)
        self.o.seti("type", t)
        if self.g is not None:
            self.g.set("type", t)

    def load(self, disp=(), force=()):
        ls = O.LoadStep()
        for bc in disp:
            ids, flag, step = bc[:3]
            mode = bc[3] if len(bc) > 3 else O.RELATIVE
            ls.add_disp(ids, flag, step, mode)
        for bc in force:
            ids, fx, fy, fz = bc[:4]
            mode = bc[4] if len(bc) > 4 else O.RELATIVE
            ls.add_force(ids, fx, fy, fz, mode)
        lg = self.g.load_step(disp, force) if self.g is not None else None
        return ls, lg


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)
