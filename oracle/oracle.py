"""ctypes wrapper of the FP64 oracle (oracle/lpm_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke() check of
__graft_entry__.py and by the cpu_baseline leg of bench.py. The product
library never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liblpm_oracle.so")
_lib = None

# enums mirroring include/lpm.h:56-96
SQUARE2D, HEXAGON2D, SC3D, FCC3D, BCC3D = 0, 1, 2, 3, 4
ELASTIC, ELASTIC_DAMAGE, J2, HCF = 0, 1, 2, 3
RELATIVE, ABSOLUTE = 0, 1

_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.c_void_p
_S = ctypes.c_char_p
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile the oracle with its Makefile (g++)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "ora_generate": (_I, [_I, _DP, _D, _I, _DP, _DP, _I]),
            "ora_rmatrix": (None, [_I, _DP, _DP]),
            "ora_create": (_P, [_I, _D, _I, _I, _DP, _DP, _I]),
            "ora_create_with_bonds": (_P, [_I, _D, _I, _I, _DP, _DP, _I, _IP, _IP, _IP]),
            "ora_destroy": (None, [_P]),
            "ora_info": (_I, [_P, _S]),
            "ora_len": (ctypes.c_long, [_P, _S]),
            "ora_get_d": (_I, [_P, _S, _DP]),
            "ora_set_d": (_I, [_P, _S, _DP]),
            "ora_get_i": (_I, [_P, _S, _IP]),
            "ora_set_i": (_I, [_P, _S, _IP]),
            "ora_set_mode": (_I, [_P, _S, _D]),
            "ora_set_elastic": (None, [_P, _I, _D, _I, _D, _D, _D]),
            "ora_set_elastic_c": (None, [_P, _I, _D, _D, _D, _D, _D]),
            "ora_set_elastic_damage": (None, [_P, _I, _D, _I, _D, _D, _D, _D, _D, _D]),
            "ora_set_j2": (None, [_P, _I, _D, _I, _D, _D, _D, _D, _D, _D, _D]),
            "ora_set_j2_damage_threshold": (None, [_P, _I, _D]),
            "ora_set_hcf": (None, [_P, _I, _D, _I, _D, _D, _D, _D, _D, _D, _D]),
            "ora_search_nonlocal": (None, [_P, _D]),
            "ora_cut_particles": (None, [_I, _DP, _I, _IP, _DP, ctypes.POINTER(ctypes.c_ubyte)]),
            "ora_search_neighbor_cut": (_I, [_I, _D, _I, _DP, _I, _IP, _DP, _I, _IP, _IP, _IP]),
            "ora_update_geometry": (None, [_P]),
            "ora_update_force_state": (None, [_P]),
            "ora_update_rr": (None, [_P]),
            "ora_norm": (_D, [_P, _S]),
            "ora_calc_stiffness": (None, [_P]),
            "ora_update_stiffness_disp_bc": (None, [_P]),
            "ora_cg": (_I, [_P]),
            "ora_cg_error": (_D, [_P]),
            "ora_solve_linear_system": (None, [_P]),
            "ora_newton": (_I, [_P]),
            "ora_update_state_var": (None, [_P]),
            "ora_update_static_damage": (_I, [_P, _I]),
            "ora_update_fatigue_damage": (_I, [_P, _D, _I]),
            "ora_store": (None, [_P]),
            "ora_reset": (None, [_P, _I]),
            "ora_write_dump": (None, [_P, _S, _I]),
            "ora_ls_new": (_P, []),
            "ora_ls_free": (None, [_P]),
            "ora_ls_add_disp": (None, [_P, _I, _IP, _I, _I, _D]),
            "ora_ls_add_force": (None, [_P, _I, _IP, _I, _D, _D, _D]),
            "ora_apply_bc": (None, [_P, _P]),
            "ora_solve_step_static": (_I, [_P, _P, _I]),
            "ora_solve_step_fatigue": (_I, [_P, _P]),
            "ora_solve_problem_static": (None, [_P, _I, ctypes.POINTER(_P), _I]),
            "ora_solve_problem_fatigue_static": (None, [_P, _I, ctypes.POINTER(_P)]),
            "ora_solve_problem_cyclic": (_I, [_P, _I, _I, _I, _DP, _I, _IP, _IP, _I, _D, _I, _I]),
            "ora_local_stiffness": (None, [_P, _I, _I, _DP]),
            "ora_func_phi": (_D, [_D, _D]),
            "ora_strain_energy": (_D, [_P, _I]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_DP)


def _ip(a):
    return a.ctypes.data_as(_IP)


def generate(lattice, box, radius, eulerflag=0, angles=(0.0, 0.0, 0.0)):
    """Lattice points of createPlateSQ2D/HEX2D, createCuboidSC3D/FCC3D (utilities.h:84-301)."""
    L = lib()
    box = np.ascontiguousarray(box, dtype=np.float64)
    ang = np.ascontiguousarray(angles, dtype=np.float64)
    n = L.ora_generate(lattice, _dp(box), radius, eulerflag, _dp(ang), None, 0)
    out = np.zeros((n, 3))
    L.ora_generate(lattice, _dp(box), radius, eulerflag, _dp(ang), _dp(out), n)
    return out


def _cuts(cuts):
    """cuts: [(kind, x, y, size)] with kind 0 circle, 1 notch, 2 reverse notch."""
    kind = np.ascontiguousarray([c[0] for c in cuts], dtype=np.int32)
    par = np.ascontiguousarray([c[1:4] for c in cuts], dtype=np.float64).reshape(-1)
    return kind, par


def cut_particles(xyz, cuts):
    """The geometry tools' isValid over particles (1 = kept)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    kind, par = _cuts(cuts)
    keep = np.zeros(len(xyz), dtype=np.uint8)
    lib().ora_cut_particles(len(xyz), _dp(xyz), len(kind), _ip(kind), _dp(par),
                            keep.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
    return keep


def search_neighbor_cut(lattice, radius, xyz, cuts):
    """The geometry tools' searchNeighbor (all pairs): (layer, p1, p2) in the tools' order."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    kind, par = _cuts(cuts)
    z = np.zeros(1, dtype=np.int32)
    m = lib().ora_search_neighbor_cut(lattice, radius, len(xyz), _dp(xyz), len(kind), _ip(kind), _dp(par), 0,
                                      _ip(z), _ip(z), _ip(z))
    lay, p1, p2 = (np.zeros(m, dtype=np.int32) for _ in range(3))
    lib().ora_search_neighbor_cut(lattice, radius, len(xyz), _dp(xyz), len(kind), _ip(kind), _dp(par), m,
                                  _ip(lay), _ip(p1), _ip(p2))
    return lay, p1, p2


def rmatrix(flag, angles):
    R = np.zeros(9)
    a = np.ascontiguousarray(angles, dtype=np.float64)
    lib().ora_rmatrix(flag, _dp(a), _dp(R))
    return R.reshape(3, 3)


class LoadStep:
    """LoadStep / DispBC / ForceBC (include/load_step.h, include/boundary.h)."""

    def __init__(self):
        self.h = lib().ora_ls_new()
        self.disp = []
        self.force = []

    def add_disp(self, ids, flag, step, mode=RELATIVE):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        lib().ora_ls_add_disp(self.h, len(ids), _ip(ids), mode, ord(flag), step)
        self.disp.append((ids, flag, step, mode))
        return self

    def add_force(self, ids, fx, fy, fz, mode=RELATIVE):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        lib().ora_ls_add_force(self.h, len(ids), _ip(ids), mode, fx, fy, fz)
        self.force.append((ids, fx, fy, fz, mode))
        return self

    def __del__(self):
        try:
            lib().ora_ls_free(self.h)
        except Exception:
            pass


class System:
    """A particle assembly of the oracle (Assembly + Solver state)."""

    def __init__(self, lattice, radius, ptype, xyz, box=None, bruteforce=False, bonds=None):
        L = lib()
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        box = np.ascontiguousarray(box if box is not None else np.zeros(6), dtype=np.float64)
        if bonds is None:
            self.h = L.ora_create(lattice, radius, ptype, len(xyz), _dp(xyz), _dp(box), int(bruteforce))
        else:
            lay, p1, p2 = (np.ascontiguousarray(b, dtype=np.int32) for b in bonds)
            self.h = L.ora_create_with_bonds(lattice, radius, ptype, len(xyz), _dp(xyz), _dp(box), len(lay),
                                             _ip(lay), _ip(p1), _ip(p2))
        if not self.h:
            raise ValueError("bond list is not symmetric (a bond has no reverse bond)")
        self.lattice = lattice
        self.ptype = ptype

    def __del__(self):
        try:
            lib().ora_destroy(self.h)
        except Exception:
            pass

    # ---- fields
    def info(self, what):
        return lib().ora_info(self.h, what.encode())

    @property
    def N(self):
        return self.info("N")

    @property
    def dim(self):
        return self.info("dim")

    def getd(self, name):
        n = lib().ora_len(self.h, name.encode())
        if n < 0:
            raise KeyError(name)
        out = np.zeros(n)
        lib().ora_get_d(self.h, name.encode(), _dp(out))
        return out

    def setd(self, name, v):
        n = lib().ora_len(self.h, name.encode())
        v = np.ascontiguousarray(v, dtype=np.float64).ravel()
        assert v.size == n, (name, v.size, n)
        lib().ora_set_d(self.h, name.encode(), _dp(v))

    def geti(self, name):
        n = lib().ora_len(self.h, name.encode())
        if n < 0:
            raise KeyError(name)
        out = np.zeros(n, dtype=np.int32)
        lib().ora_get_i(self.h, name.encode(), _ip(out))
        return out

    def seti(self, name, v):
        n = lib().ora_len(self.h, name.encode())
        v = np.ascontiguousarray(v, dtype=np.int32).ravel()
        assert v.size == n, (name, v.size, n)
        lib().ora_set_i(self.h, name.encode(), _ip(v))

    def mode(self, **kw):
        for k, v in kw.items():
            if lib().ora_set_mode(self.h, k.encode(), float(v)) != 0:
                raise KeyError(k)
        return self

    # ---- material (setParticleProperty)
    def set_elastic(self, L, plane_stress, E, mu, crit, pid=-1):
        lib().ora_set_elastic(self.h, pid, L, int(plane_stress), E, mu, crit)

    def set_elastic_c(self, L, C11, C12, C44, crit, pid=-1):
        lib().ora_set_elastic_c(self.h, pid, L, C11, C12, C44, crit)

    def set_elastic_damage(self, L, plane_stress, E, mu, k0, k1, ct, dthr, pid=-1):
        lib().ora_set_elastic_damage(self.h, pid, L, int(plane_stress), E, mu, k0, k1, ct, dthr)

    def set_j2(self, L, plane_stress, E, mu, sy, xi, H, A, crit, pid=-1):
        lib().ora_set_j2(self.h, pid, L, int(plane_stress), E, mu, sy, xi, H, A, crit)

    def set_j2_damage_threshold(self, thr, pid=-1):
        lib().ora_set_j2_damage_threshold(self.h, pid, thr)

    def set_hcf(self, L, plane_stress, E, mu, A, B, d, dthr=1.0, flr=1.108, pid=-1):
        lib().ora_set_hcf(self.h, pid, L, int(plane_stress), E, mu, A, B, d, dthr, flr)

    # ---- phases
    def search_nonlocal(self, ratio):
        lib().ora_search_nonlocal(self.h, ratio)

    def update_geometry(self):
        lib().ora_update_geometry(self.h)

    def update_force_state(self):
        lib().ora_update_force_state(self.h)

    def update_rr(self):
        lib().ora_update_rr(self.h)

    def norm(self, name):
        return lib().ora_norm(self.h, name.encode())

    def calc_stiffness(self):
        lib().ora_calc_stiffness(self.h)

    def update_stiffness_disp_bc(self):
        lib().ora_update_stiffness_disp_bc(self.h)

    def K_csr(self):
        """(rowptr, col, val) of K over K's rows (row i*dim+r, column j*dim+s)."""
        return self.geti("krow"), self.geti("kcol"), self.getd("kval")

    def K_dense(self):
        import scipy.sparse as sp
        r, c, v = self.K_csr()
        n = len(r) - 1
        return sp.csr_matrix((v, c, r), shape=(n, n)).toarray()

    def cg(self):
        return lib().ora_cg(self.h)

    def cg_error(self):
        return lib().ora_cg_error(self.h)

    def solve_linear_system(self):
        lib().ora_solve_linear_system(self.h)

    def newton(self):
        return lib().ora_newton(self.h)

    def update_state_var(self):
        lib().ora_update_state_var(self.h)

    def update_static_damage(self, undamaged):
        return bool(lib().ora_update_static_damage(self.h, undamaged))

    def update_fatigue_damage(self, dNdt, undamaged):
        return bool(lib().ora_update_fatigue_damage(self.h, dNdt, undamaged))

    def store(self):
        lib().ora_store(self.h)

    def reset(self, reset_xyz):
        lib().ora_reset(self.h, int(reset_xyz))

    def write_dump(self, path, step):
        lib().ora_write_dump(self.h, path.encode(), step)

    def apply_bc(self, ls):
        lib().ora_apply_bc(self.h, ls.h)

    def solve_step_static(self, ls, undamaged=-1):
        return bool(lib().ora_solve_step_static(self.h, ls.h, undamaged))

    def solve_step_fatigue(self, ls):
        return bool(lib().ora_solve_step_fatigue(self.h, ls.h))

    def solve_problem_static(self, steps, undamaged=-1):
        arr = (_P * len(steps))(*[s.h for s in steps])
        lib().ora_solve_problem_static(self.h, len(steps), arr, undamaged)

    def solve_problem_fatigue_static(self, steps):
        arr = (_P * len(steps))(*[s.h for s in steps])
        lib().ora_solve_problem_fatigue_static(self.h, len(steps), arr)

    def solve_problem_cyclic(self, ltype, n_interval, spectrum, groups, tau, undamaged, tmode=0, max_cycles_run=0):
        spec = np.ascontiguousarray(spectrum, dtype=np.float64)
        sizes = np.array([len(g) for g in groups], dtype=np.int32)
        ids = np.ascontiguousarray(np.concatenate([np.asarray(g, dtype=np.int32) for g in groups]), dtype=np.int32)
        return lib().ora_solve_problem_cyclic(self.h, ltype, n_interval, len(spec), _dp(spec), len(groups), _ip(sizes),
                                              _ip(ids), tmode, tau, undamaged, max_cycles_run)

    def local_stiffness(self, i, j):
        out = np.zeros(9)
        lib().ora_local_stiffness(self.h, i, j, _dp(out))
        return out.reshape(3, 3)

    def strain_energy(self, i):
        return lib().ora_strain_energy(self.h, i)


def func_phi(x, L):
    return lib().ora_func_phi(x, L)
