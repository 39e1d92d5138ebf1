"""lpm_update_rr_noalloc (the residual and norms of lpm_update_rr on caller-owned device arrays,
partial sums in the library's static device array) against lpm_update_rr on the same state:
residual and norms bit for bit, on two calls in a row, the second with a smaller input.

The caller's device arrays are allocated by the test through the HIP runtime the library itself
links (libamdhip64.so.7, already loaded with it), so no second runtime enters the process."""
import ctypes

import numpy as np
import pytest

import lpmgpu as G
import oracle as O
from scenarios import sc_cube

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    G.lib()  # loads the library and with it its HIP runtime
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def _device(hip, a):
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
    assert hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D) == 0
    return p


def _state(nside, seed):
    xyz, _ = sc_cube(nside)
    g = G.Context(O.SC3D, 0.5, O.ELASTIC, xyz)
    g.set_elastic(0.0, False, 200e3, 0.3, 1.0, elmap_literal=0)
    rng = np.random.default_rng(seed)
    g.set("xyz", xyz + 1e-3 * rng.standard_normal(xyz.shape))
    cons = (rng.random(3 * len(xyz)) < 0.2).astype(np.int32)
    pex = rng.standard_normal(3 * len(xyz))
    g.set("disp_constraint", cons)
    g.set("Pex", pex)
    g.update_geometry()
    g.update_force_state()
    return g, pex, cons


def test_noalloc_residual_matches_update_rr():
    print("[gpu] no-alloc residual", flush=True)
    L = G.lib()
    hip = _hip()
    for nside, seed in ((12, 1), (7, 2)):   # the second call is smaller
        g, pex, cons = _state(nside, seed)
        nr, nR = g.update_rr()
        r_ref = g.get("residual")
        n = g.n
        pin = g.get("Pin")
        d_pex, d_pin, d_cons = _device(hip, pex), _device(hip, pin), _device(hip, cons)
        res = np.full(3 * n, np.nan)
        d_res = _device(hip, res)
        norms = np.zeros(2)
        try:
            rc = L.lpm_update_rr_noalloc(n, 3, d_pex, d_pin, d_cons, d_res,
                                         norms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None)
            assert rc == 0, L.lpm_last_error(None).decode()
            assert hip.hipMemcpy(res.ctypes.data_as(ctypes.c_void_p), d_res, res.nbytes, D2H) == 0
        finally:
            for p in (d_pex, d_pin, d_cons, d_res):
                hip.hipFree(p)
        np.testing.assert_array_equal(res, r_ref)
        assert norms[0] == nr and norms[1] == nR
        print(f"   n={n}: residual and norms identical ({nr:.6e}, {nR:.6e})", flush=True)
        g.close()
