"""ctypes binding of liblpm_mi355x.so (C ABI: include/lpm_mi355x.h).

Used by the tests, bench.py and __graft_entry__.py. It loads the in-tree library
and raises if it is missing: there is no CPU fallback.
"""
import ctypes
import os
import re

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "liblpm_mi355x.so")
HEADER = os.path.join(os.path.dirname(_PKG), "include", "lpm_mi355x.h")

SQUARE2D, HEXAGON2D, SC3D, FCC3D, BCC3D = 0, 1, 2, 3, 4
ELASTIC, ELASTIC_DAMAGE, J2, HCF = 0, 1, 2, 3
RELATIVE, ABSOLUTE = 0, 1
LOAD_UNIAXIAL_DISP, LOAD_UNIAXIAL_FORCE, LOAD_DOGBONE_DISP, LOAD_CT_FORCE = 0, 1, 2, 3

_D, _I, _L, _P, _S = ctypes.c_double, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_char_p
_DP, _IP = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)
_LP = ctypes.POINTER(ctypes.c_int64)


class Topology(ctypes.Structure):
    _fields_ = [("lattice", ctypes.c_int32), ("radius", _D), ("model", ctypes.c_int32), ("n", _L), ("xyz", _DP),
                ("bond_ptr", _IP), ("bond_split", _IP), ("bond_p2", _IP), ("device", ctypes.c_int32)]


class Cut(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("x", _D), ("y", _D), ("size", _D)]


CUT_CIRCLE, CUT_NOTCH, CUT_REVERSE_NOTCH = 0, 1, 2


def _cut_array(cuts):
    arr = (Cut * max(len(cuts), 1))()
    for k, (kind, x, y, size) in enumerate(cuts):
        arr[k] = Cut(int(kind), float(x), float(y), float(size))
    return arr


class DispBC(ctypes.Structure):
    _fields_ = [("group", ctypes.c_int32), ("mode", ctypes.c_int32), ("flag", ctypes.c_int32), ("step", _D)]


class ForceBC(ctypes.Structure):
    _fields_ = [("group", ctypes.c_int32), ("mode", ctypes.c_int32), ("fx", _D), ("fy", _D), ("fz", _D)]


class LoadStepC(ctypes.Structure):
    _fields_ = [("n_disp", ctypes.c_int32), ("disp", ctypes.POINTER(DispBC)), ("n_force", ctypes.c_int32),
                ("force", ctypes.POINTER(ForceBC))]

    def address(self):
        """The lpm_load_step* as an integer (for C++ callers such as lpm_autograd)."""
        return ctypes.addressof(self)


_SIGS = {
    "lpm_generate_lattice": (_I, [_I, _DP, _D, _I, _DP, _DP, _L, _LP]),
    "lpm_build_bonds": (_I, [_I, _D, _L, _DP, _IP, _IP, _IP, _L, _LP]),
    "lpm_unit_cell": (_I, [_I, _D, _I, _DP, _DP]),
    "lpm_create": (_I, [ctypes.POINTER(Topology), ctypes.POINTER(_P)]),
    "lpm_destroy": (_I, [_P]),
    "lpm_last_error": (_S, [_P]),
    "lpm_size": (_L, [_P, _S]),
    "lpm_set_mode": (_I, [_P, _S, _D]),
    "lpm_get_mode": (_D, [_P, _S]),
    "lpm_upload": (_I, [_P, _S, _P, _L]),
    "lpm_download": (_I, [_P, _S, _P, _L]),
    "lpm_download_matrix": (_I, [_P, _IP, _IP, _DP, _L]),
    "lpm_set_material": (_I, [_P, _L, _D, _DP, _I, _DP, ctypes.c_int32]),
    "lpm_ce_from_e": (_I, [_I, _D, _D, _DP]),
    "lpm_search_nonlocal": (_I, [_P, _D]),
    "lpm_group_create": (_I, [_P, _L, _IP, _IP]),
    "lpm_apply_bc": (_I, [_P, ctypes.POINTER(LoadStepC)]),
    "lpm_update_geometry": (_I, [_P]),
    "lpm_update_force_state": (_I, [_P]),
    "lpm_update_rr": (_I, [_P, _DP, _DP]),
    "lpm_update_rr_noalloc": (_I, [ctypes.c_int64, _I, _P, _P, _P, _P, _DP, _P]),
    "lpm_calc_stiffness": (_I, [_P]),
    "lpm_update_stiffness_disp_bc": (_I, [_P]),
    "lpm_solve_linear_system": (_I, [_P, _IP, _DP]),
    "lpm_newton_iteration": (_I, [_P, _IP]),
    "lpm_update_state_var": (_I, [_P]),
    "lpm_update_static_damage": (_I, [_P, ctypes.c_int32, _IP]),
    "lpm_update_fatigue_damage": (_I, [_P, _D, ctypes.c_int32, _IP]),
    "lpm_store_state": (_I, [_P]),
    "lpm_reset_state": (_I, [_P, ctypes.c_int32]),
    "lpm_solve_step_static": (_I, [_P, ctypes.POINTER(LoadStepC), ctypes.c_int32, _IP]),
    "lpm_solve_problem_static": (_I, [_P, ctypes.POINTER(LoadStepC), ctypes.c_int32, ctypes.c_int32, _IP]),
    "lpm_solve_step_fatigue": (_I, [_P, ctypes.POINTER(LoadStepC), _IP]),
    "lpm_solve_problem_fatigue_static": (_I, [_P, ctypes.POINTER(LoadStepC), ctypes.c_int32, _IP]),
    "lpm_solve_problem_cyclic": (_I, [_P, ctypes.c_int32, ctypes.c_int32, _DP, _L, _IP, ctypes.c_int32, ctypes.c_int32,
                                      _D, ctypes.c_int32, ctypes.c_int32, _IP]),
    "lpm_stat": (_D, [_P, _S]),
    "lpm_reset_stats": (_I, [_P]),
    "lpm_set_timing": (_I, [_P, _I]),
    "lpm_spmv_bench": (_I, [_P, ctypes.c_int32, _DP]),
    "lpm_copy_bench": (_I, [_P, ctypes.c_int64, ctypes.c_int32, _DP]),
    "lpm_force_jvp": (_I, [_P, _DP, _DP]),
    "lpm_force_vjp": (_I, [_P, _DP, _DP]),
    "lpm_set_specimens": (_I, [_P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]),
    "lpm_specimen_stat": (_I, [_P, _S, _DP]),
    "lpm_solve_problem_cyclic_specimens": (_I, [_P, ctypes.c_int32, ctypes.c_int32, _DP, _L, _IP, _DP, ctypes.c_int32,
                                                _D, ctypes.c_int32, ctypes.c_int32, _IP]),
    "lpm_adjoint_reaction": (_I, [_P, ctypes.c_int32, ctypes.c_int32, _DP, _DP, _DP, ctypes.POINTER(ctypes.c_int32), _DP]),
    "lpm_adjoint_linear": (_I, [_P, _DP, _DP, _DP, _DP, _DP, _DP, ctypes.POINTER(ctypes.c_int32), _DP]),
    "lpm_adjoint_fatigue_params": (_I, [_P, _DP, _DP, _DP, _DP, _DP, _DP, ctypes.POINTER(ctypes.c_int32), _DP]),
    "lpm_fatigue_record": (_I, [_P, ctypes.c_int32]),
    "lpm_adjoint_fatigue_path": (_I, [_P, _DP, _DP, _DP, _DP, _DP, _DP, ctypes.POINTER(ctypes.c_int32), _DP]),
    "lpm_checkpoint_save": (_I, [_P, ctypes.c_int32]),
    "lpm_checkpoint_restore": (_I, [_P, ctypes.c_int32]),
    "lpm_checkpoint_release": (_I, [_P, ctypes.c_int32]),
    "lpm_stream": (_P, [_P]),
    "lpm_dump_async": (_I, [_P, _S, ctypes.c_int32, _DP]),
    "lpm_dump_wait": (_I, [_P]),
    "lpm_format_e4": (_I, [_DP, _L, ctypes.c_char_p, _L]),
    "lpm_create_cut": (_I, [ctypes.POINTER(Topology), ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    "lpm_cut_particles": (_I, [_DP, _L, ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_uint8)]),
    "lpm_image_particles": (_I, [ctypes.POINTER(ctypes.c_uint8), _L, _L, _D, _DP, _L, ctypes.POINTER(_L)]),
}

_lib = None


def lib():
    """Load the in-tree library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def header_symbols():
    """Function names declared in include/lpm_mi355x.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lpm_[a-z_0-9]+)\s*\(", txt)))


class LPMError(RuntimeError):
    pass


def _dp(a):
    return a.ctypes.data_as(_DP)


def _ip(a):
    return a.ctypes.data_as(_IP)


# ---------------------------------------------------------------- host setup (no GPU needed)
def generate_lattice(lattice, box, radius, eulerflag=0, angles=(0.0, 0.0, 0.0)):
    L = lib()
    box = np.ascontiguousarray(box, dtype=np.float64)
    ang = np.ascontiguousarray(angles, dtype=np.float64)
    n = ctypes.c_int64(0)
    _check(L.lpm_generate_lattice(lattice, _dp(box), radius, eulerflag, _dp(ang), None, 0, ctypes.byref(n)))
    out = np.zeros((n.value, 3))
    _check(L.lpm_generate_lattice(lattice, _dp(box), radius, eulerflag, _dp(ang), _dp(out), n.value, ctypes.byref(n)))
    return out


def build_bonds(lattice, radius, xyz):
    L = lib()
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = len(xyz)
    nb = ctypes.c_int64(0)
    _check(L.lpm_build_bonds(lattice, radius, n, _dp(xyz), None, None, None, 0, ctypes.byref(nb)))
    bp = np.zeros(n + 1, dtype=np.int32)
    bs = np.zeros(n, dtype=np.int32)
    p2 = np.zeros(max(nb.value, 1), dtype=np.int32)
    _check(L.lpm_build_bonds(lattice, radius, n, _dp(xyz), _ip(bp), _ip(bs), _ip(p2), nb.value, ctypes.byref(nb)))
    return bp, bs, p2[:nb.value]


def unit_cell(lattice, radius, literal=1):
    out = np.zeros(8)
    m = np.zeros(9)
    _check(lib().lpm_unit_cell(lattice, radius, literal, _dp(out), _dp(m)))
    return out, m.reshape(3, 3)


def cut_particles(xyz, cuts):
    """1 for the particles in no cut (the geometry tools' isValid), else 0."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    keep = np.zeros(len(xyz), dtype=np.uint8)
    arr = _cut_array(cuts)
    _check(lib().lpm_cut_particles(_dp(xyz), len(xyz), ctypes.cast(arr, ctypes.c_void_p), len(cuts),
                                   keep.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
    return keep


def image_particles(img, pixel):
    """One particle per pixel equal to 255 (generate_dump.ipynb createPlate); img is nrows x ncols uint8."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    n = _L()
    ip = img.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    _check(lib().lpm_image_particles(ip, img.shape[0], img.shape[1], pixel, None, 0, ctypes.byref(n)))
    out = np.zeros((n.value, 3))
    _check(lib().lpm_image_particles(ip, img.shape[0], img.shape[1], pixel, _dp(out), n.value, ctypes.byref(n)))
    return out


def format_e4(x):
    """The dump writer's "%.4e" formatting of x (host code of the library)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    cap = 48 * len(x) + 64
    buf = ctypes.create_string_buffer(cap)
    _check(lib().lpm_format_e4(_dp(x), len(x), buf, cap))
    return buf.value.decode().split("\n")[:-1]


def ce_from_e(plane_stress, E, mu):
    Ce = np.zeros(3)
    _check(lib().lpm_ce_from_e(int(plane_stress), E, mu, _dp(Ce)))
    return Ce


def _check(rc, ctx=None):
    if rc != 0:
        msg = lib().lpm_last_error(ctx)
        raise LPMError(f"lpm error {rc}: {msg.decode() if msg else ''}")


_INT_FIELDS = {"type", "disp_constraint", "nb", "bond_ptr", "bond_split", "bond_p2", "opp", "conn_ptr", "conn_idx",
               "nl_ptr", "nl_idx"}


class Context:
    """One particle system on one GPU (an Assembly + Solver of the reference)."""

    def __init__(self, lattice, radius, model, xyz, bonds=None, device=0, cuts=None):
        """bonds: (bond_ptr, bond_split, bond_p2) lists, None for the host search (lpm_build_bonds),
        or "device" for the search inside lpm_create on the GPU; cuts (with bonds="device"):
        [(kind, x, y, size)] holes and notches whose bonds the search drops (lpm_create_cut)."""
        L = lib()
        self.xyz0 = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        if isinstance(bonds, str):
            if bonds != "device":
                raise ValueError("bonds must be None, 'device' or (bond_ptr, bond_split, bond_p2)")
            t = Topology(lattice, radius, model, len(self.xyz0), _dp(self.xyz0), None, None, None, device)
        else:
            if bonds is None:
                bonds = build_bonds(lattice, radius, self.xyz0)
            self._bonds = tuple(np.ascontiguousarray(b, dtype=np.int32) for b in bonds)
            t = Topology(lattice, radius, model, len(self.xyz0), _dp(self.xyz0), *(_ip(b) for b in self._bonds), device)
        h = _P()
        if cuts:
            if bonds != "device":
                raise ValueError("cuts apply to the device bond search (bonds='device')")
            arr = _cut_array(cuts)
            rc = L.lpm_create_cut(ctypes.byref(t), ctypes.cast(arr, ctypes.c_void_p), len(cuts), ctypes.byref(h))
        else:
            rc = L.lpm_create(ctypes.byref(t), ctypes.byref(h))
        if rc != 0:
            raise LPMError(f"lpm_create failed ({rc}): {L.lpm_last_error(None).decode()}")
        self.h = h
        self.lattice, self.radius, self.model = lattice, radius, model
        self.n = len(self.xyz0)
        self.dim = int(L.lpm_size(self.h, b"dim"))
        self._groups = {}

    def close(self):
        if getattr(self, "h", None):
            lib().lpm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        """The lpm_ctx* as an integer (for C++ callers such as lpm_autograd)."""
        return self.h.value

    def _chk(self, rc):
        if rc != 0:
            raise LPMError(f"lpm error {rc}: {lib().lpm_last_error(self.h).decode()}")

    def size(self, what):
        return int(lib().lpm_size(self.h, what.encode()))

    def mode(self, **kw):
        for k, v in kw.items():
            self._chk(lib().lpm_set_mode(self.h, k.encode(), float(v)))
        return self

    def get_mode(self, name):
        return lib().lpm_get_mode(self.h, name.encode())

    # ---- transfers
    def _count(self, name):
        per = {"xyz": 3, "xyz_initial": 3, "xyz_last": 3, "Pin": 3, "Pex": 3, "stress": 6, "state_var": 7,
               "state_var_last": 7, "mat": 10, "dLe_total": 2, "TdLe_total": 2, "cs_sum": 6, "disp_constraint": 3}
        n, nb = self.n, self.size("nbonds")
        if name in ("damage", "damage_last", "damage_visual", "Ddot_local", "Ddot_nonlocal", "nonlocal_L", "type", "nb",
                    "bond_split"):
            return n
        if name in per:
            return per[name] * n
        if name == "cs":
            return 3 * nb
        if name in ("dis_initial", "dis", "dis_last", "dLe", "dLp", "dLp_last", "bstrain", "Kn", "Tv", "bforce",
                    "bforce_last", "bdamage", "bdamage_last", "bond_p2", "opp"):
            return nb
        if name in ("bond_ptr", "conn_ptr"):
            return n + 1
        if name == "nl_ptr":
            return self.size("nl_ptr")
        if name == "conn_idx":
            return self.size("nconn")
        if name == "nl_idx":
            return self.size("nnl")
        if name in ("residual", "disp"):
            return self.dim * n
        raise KeyError(name)

    def get(self, name):
        if name == "reaction":
            k = lib().lpm_download(self.h, b"reaction", None, 0)
            if k < 0:
                self._chk(k)
            out = np.zeros(k)
            self._chk(lib().lpm_download(self.h, b"reaction", out.ctypes.data_as(_P), k))
            return out
        cnt = self._count(name)
        out = np.zeros(cnt, dtype=np.int32 if name in _INT_FIELDS else np.float64)
        self._chk(lib().lpm_download(self.h, name.encode(), out.ctypes.data_as(_P), cnt))
        return out

    def set(self, name, v):
        dt = np.int32 if name in _INT_FIELDS else np.float64
        v = np.ascontiguousarray(v, dtype=dt).ravel()
        self._chk(lib().lpm_upload(self.h, name.encode(), v.ctypes.data_as(_P), v.size))

    def matrix(self):
        """(rowptr, col, val) of the matrix handed to CG (K, or K^T when cg_transpose)."""
        nnz = self.size("knnz")
        nd = self.dim * self.n
        r = np.zeros(nd + 1, dtype=np.int32)
        c = np.zeros(nnz, dtype=np.int32)
        v = np.zeros(nnz)
        self._chk(lib().lpm_download_matrix(self.h, _ip(r), _ip(c), _dp(v), nnz))
        return r, c, v

    # ---- material (setParticleProperty)
    def set_material(self, Ce, params, nonlocal_L=0.0, elmap_literal=1, pid=-1):
        Ce = np.ascontiguousarray(Ce, dtype=np.float64)
        p = np.ascontiguousarray(params, dtype=np.float64)
        self._chk(lib().lpm_set_material(self.h, pid, nonlocal_L, _dp(Ce), int(elmap_literal), _dp(p), len(p)))

    def set_elastic(self, L, plane_stress, E, mu, crit, elmap_literal=1, pid=-1):
        self.set_material(ce_from_e(plane_stress, E, mu), [crit], L, elmap_literal, pid)

    def set_elastic_damage(self, L, plane_stress, E, mu, k0, k1, ct, dthr, elmap_literal=1, pid=-1):
        self.set_material(ce_from_e(plane_stress, E, mu), [k0, k1, dthr, ct], L, elmap_literal, pid)

    def set_j2(self, L, plane_stress, E, mu, sy, xi, H, A, crit, damage_threshold=0.0, elmap_literal=1, pid=-1):
        self.set_material(ce_from_e(plane_stress, E, mu), [E, mu, sy, xi, H, A, damage_threshold, crit], L,
                          elmap_literal, pid)

    def set_hcf(self, L, plane_stress, E, mu, A, B, d, dthr=1.0, flr=1.108, elmap_literal=1, pid=-1):
        self.set_material(ce_from_e(plane_stress, E, mu), [A, B, d, E, mu, dthr, flr], L, elmap_literal, pid)

    # ---- groups and loads
    def group(self, ids):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        key = ids.tobytes()
        if key not in self._groups:
            g = ctypes.c_int32(0)
            self._chk(lib().lpm_group_create(self.h, len(ids), _ip(ids), ctypes.byref(g)))
            self._groups[key] = g.value
        return self._groups[key]

    def load_step(self, disp=(), force=()):
        """disp: [(ids, flag, step[, mode])], force: [(ids, fx, fy, fz[, mode])] -> LoadStepC (keeps arrays alive)."""
        d = (DispBC * max(len(disp), 1))()
        for k, bc in enumerate(disp):
            ids, flag, step = bc[:3]
            mode = bc[3] if len(bc) > 3 else RELATIVE
            d[k] = DispBC(self.group(ids), mode, ord(flag), step)
        f = (ForceBC * max(len(force), 1))()
        for k, bc in enumerate(force):
            ids, fx, fy, fz = bc[:4]
            mode = bc[4] if len(bc) > 4 else RELATIVE
            f[k] = ForceBC(self.group(ids), mode, fx, fy, fz)
        ls = LoadStepC(len(disp), d, len(force), f)
        ls._keep = (d, f)
        return ls

    # ---- phases
    def search_nonlocal(self, ratio):
        self._chk(lib().lpm_search_nonlocal(self.h, ratio))

    def apply_bc(self, ls):
        self._chk(lib().lpm_apply_bc(self.h, ctypes.byref(ls)))

    def update_geometry(self):
        self._chk(lib().lpm_update_geometry(self.h))

    def update_force_state(self):
        self._chk(lib().lpm_update_force_state(self.h))

    def update_rr(self):
        a, b = _D(), _D()
        self._chk(lib().lpm_update_rr(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def calc_stiffness(self):
        self._chk(lib().lpm_calc_stiffness(self.h))

    def update_stiffness_disp_bc(self):
        self._chk(lib().lpm_update_stiffness_disp_bc(self.h))

    def solve_linear_system(self):
        it, err = ctypes.c_int32(0), _D()
        self._chk(lib().lpm_solve_linear_system(self.h, ctypes.byref(it), ctypes.byref(err)))
        return it.value, err.value

    def newton(self):
        ni = ctypes.c_int32(0)
        self._chk(lib().lpm_newton_iteration(self.h, ctypes.byref(ni)))
        return ni.value

    def update_state_var(self):
        self._chk(lib().lpm_update_state_var(self.h))

    def update_static_damage(self, undamaged=-1):
        a = ctypes.c_int32(0)
        self._chk(lib().lpm_update_static_damage(self.h, undamaged, ctypes.byref(a)))
        return bool(a.value)

    def update_fatigue_damage(self, dNdt, undamaged=-1):
        a = ctypes.c_int32(0)
        self._chk(lib().lpm_update_fatigue_damage(self.h, dNdt, undamaged, ctypes.byref(a)))
        return bool(a.value)

    def store(self):
        self._chk(lib().lpm_store_state(self.h))

    def reset(self, reset_xyz):
        self._chk(lib().lpm_reset_state(self.h, int(reset_xyz)))

    def solve_step_static(self, ls, undamaged=-1):
        ok = ctypes.c_int32(0)
        self._chk(lib().lpm_solve_step_static(self.h, ctypes.byref(ls), undamaged, ctypes.byref(ok)))
        return bool(ok.value)

    def solve_problem_static(self, steps, undamaged=-1):
        arr = (LoadStepC * len(steps))(*steps)
        n = ctypes.c_int32(0)
        self._chk(lib().lpm_solve_problem_static(self.h, arr, len(steps), undamaged, ctypes.byref(n)))
        return n.value

    def solve_step_fatigue(self, ls):
        ok = ctypes.c_int32(0)
        self._chk(lib().lpm_solve_step_fatigue(self.h, ctypes.byref(ls), ctypes.byref(ok)))
        return bool(ok.value)

    def solve_problem_fatigue_static(self, steps):
        arr = (LoadStepC * len(steps))(*steps)
        n = ctypes.c_int32(0)
        self._chk(lib().lpm_solve_problem_fatigue_static(self.h, arr, len(steps), ctypes.byref(n)))
        return n.value

    def solve_problem_cyclic(self, ltype, n_interval, spectrum, groups, tau, undamaged, tmode=0, max_cycles_run=0):
        spec = np.ascontiguousarray(spectrum, dtype=np.float64)
        g = np.array([self.group(x) for x in groups], dtype=np.int32)
        out = ctypes.c_int32(0)
        self._chk(lib().lpm_solve_problem_cyclic(self.h, ltype, n_interval, _dp(spec), len(spec), _ip(g), len(g), tmode,
                                                 tau, undamaged, max_cycles_run, ctypes.byref(out)))
        return out.value

    def solve_problem_cyclic_specimens(self, ltype, n_interval, spectrum, groups, scale, tau, undamaged, tmode=0,
                                       max_cycles_run=0):
        """groups: per specimen, the template's groups (id arrays); scale: per-specimen amplitude factors."""
        spec = np.ascontiguousarray(spectrum, dtype=np.float64)
        g = np.array([self.group(x) for gs in groups for x in gs], dtype=np.int32)
        sc = np.ascontiguousarray(scale, dtype=np.float64)
        out = ctypes.c_int32(0)
        self._chk(lib().lpm_solve_problem_cyclic_specimens(self.h, ltype, n_interval, _dp(spec), len(spec), _ip(g),
                                                           _dp(sc), tmode, tau, undamaged, max_cycles_run,
                                                           ctypes.byref(out)))
        return out.value

    # ---- stats
    def stat(self, name):
        return lib().lpm_stat(self.h, name.encode())

    def reset_stats(self):
        self._chk(lib().lpm_reset_stats(self.h))

    def set_timing(self, on):
        self._chk(lib().lpm_set_timing(self.h, int(on)))

    def spmv_bench(self, reps):
        ms = _D()
        self._chk(lib().lpm_spmv_bench(self.h, reps, ctypes.byref(ms)))
        return ms.value

    def set_specimens(self, first):
        """Declare an ensemble: specimen s owns particle ids first[s] .. first[s+1]-1."""
        first = np.ascontiguousarray(first, dtype=np.int64)
        self._nspec = len(first) - 1
        self._chk(lib().lpm_set_specimens(self.h, self._nspec, first.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))

    def specimen_stat(self, name):
        out = np.zeros(self._nspec)
        self._chk(lib().lpm_specimen_stat(self.h, name.encode(), out.ctypes.data_as(_DP)))
        return out

    def adjoint_linear(self, w_pin=None, w_x=None):
        """dG/dKn, dG/dTv, dG/dbdamage per bond for G = w_pin . Pin + w_x . x (DoF-order weights,
        either may be None). Returns (G, dKn, dTv, dbdamage, refinements, relative residual)."""
        nd = self.n * self.size("dim")
        nb = self.size("nbonds")
        w = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (w_pin, w_x)]
        for a in w:
            assert a is None or a.size == nd
        gk, gt, gd = np.zeros(nb), np.zeros(nb), np.zeros(nb)
        G, rr, it = _D(), _D(), ctypes.c_int32()
        ptr = [None if a is None else a.ctypes.data_as(_DP) for a in w]
        self._chk(lib().lpm_adjoint_linear(self.h, ptr[0], ptr[1], gk.ctypes.data_as(_DP), gt.ctypes.data_as(_DP),
                                           gd.ctypes.data_as(_DP), ctypes.byref(G), ctypes.byref(it), ctypes.byref(rr)))
        return G.value, gk, gt, gd, it.value, rr.value

    def adjoint_fatigue_params(self, w_pin=None, w_x=None):
        """dG/dA, dG/dB, dG/dd per particle (HCF law) through the last fatigue damage update, for
        G = w_pin . Pin + w_x . x at the current equilibrium. Returns (G, dA, dB, dd, refinements,
        relative residual)."""
        nd = self.n * self.size("dim")
        w = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (w_pin, w_x)]
        for a in w:
            assert a is None or a.size == nd
        ga, gb, gd = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        G, rr, it = _D(), _D(), ctypes.c_int32()
        ptr = [None if a is None else a.ctypes.data_as(_DP) for a in w]
        self._chk(lib().lpm_adjoint_fatigue_params(self.h, ptr[0], ptr[1], ga.ctypes.data_as(_DP), gb.ctypes.data_as(_DP),
                                                   gd.ctypes.data_as(_DP), ctypes.byref(G), ctypes.byref(it),
                                                   ctypes.byref(rr)))
        return G.value, ga, gb, gd, it.value, rr.value

    def fatigue_record(self, on=True):
        """Start (dropping any earlier record) or stop recording the fatigue load path."""
        self._chk(lib().lpm_fatigue_record(self.h, 1 if on else 0))

    def adjoint_fatigue_path(self, w_pin=None, w_x=None):
        """dG/dA, dG/dB, dG/dd per particle (HCF law) through the whole recorded fatigue path, for
        G = w_pin . Pin + w_x . x at the current equilibrium. Returns (G, dA, dB, dd, transposed
        solves, largest relative residual)."""
        nd = self.n * self.size("dim")
        w = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (w_pin, w_x)]
        for a in w:
            assert a is None or a.size == nd
        ga, gb, gd = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        G, rr, ns = _D(), _D(), ctypes.c_int32()
        ptr = [None if a is None else a.ctypes.data_as(_DP) for a in w]
        self._chk(lib().lpm_adjoint_fatigue_path(self.h, ptr[0], ptr[1], ga.ctypes.data_as(_DP), gb.ctypes.data_as(_DP),
                                                 gd.ctypes.data_as(_DP), ctypes.byref(G), ctypes.byref(ns),
                                                 ctypes.byref(rr)))
        return G.value, ga, gb, gd, ns.value, rr.value

    def checkpoint_save(self, slot):
        self._chk(lib().lpm_checkpoint_save(self.h, int(slot)))

    def checkpoint_restore(self, slot):
        self._chk(lib().lpm_checkpoint_restore(self.h, int(slot)))

    def checkpoint_release(self, slot=-1):
        self._chk(lib().lpm_checkpoint_release(self.h, int(slot)))

    def force_vjp(self, w):
        """Transposed product (dPin/dx)^T w, DoF order."""
        nd = self.n * self.size("dim")
        w = np.ascontiguousarray(w, dtype=np.float64).ravel()
        assert w.size == nd
        out = np.zeros(nd)
        self._chk(lib().lpm_force_vjp(self.h, w.ctypes.data_as(_DP), out.ctypes.data_as(_DP)))
        return out

    def force_jvp(self, v):
        """Exact directional derivative of the internal force: (dPin/dx) v, DoF order."""
        nd = self.n * self.size("dim")
        v = np.ascontiguousarray(v, dtype=np.float64).ravel()
        assert v.size == nd
        out = np.zeros(nd)
        self._chk(lib().lpm_force_jvp(self.h, v.ctypes.data_as(_DP), out.ctypes.data_as(_DP)))
        return out

    def adjoint_reaction(self, group, axis):
        """dF/dKn and dF/dTv per bond for F = sum of Pin[axis] over the group (a group id from
        load_step / group()). Returns (F, dKn, dTv, refinements, relative residual)."""
        nb = self.size("nbonds")
        gk, gt = np.zeros(nb), np.zeros(nb)
        F = _D()
        it = ctypes.c_int32()
        rr = _D()
        ax = "xyz".index(axis) if isinstance(axis, str) else int(axis)
        self._chk(lib().lpm_adjoint_reaction(self.h, int(group), ax, gk.ctypes.data_as(_DP), gt.ctypes.data_as(_DP),
                                             ctypes.byref(F), ctypes.byref(it), ctypes.byref(rr)))
        return F.value, gk, gt, it.value, rr.value

    def dump_async(self, path, step, box):
        """Append a writeDump frame for the current state to `path` from a writer thread."""
        b = np.ascontiguousarray(box, dtype=np.float64)
        self._chk(lib().lpm_dump_async(self.h, str(path).encode(), int(step), _dp(b)))

    def dump_wait(self):
        self._chk(lib().lpm_dump_wait(self.h))

    def copy_bench(self, nbytes, reps):
        ms = _D()
        self._chk(lib().lpm_copy_bench(self.h, int(nbytes), int(reps), ctypes.byref(ms)))
        return ms.value
