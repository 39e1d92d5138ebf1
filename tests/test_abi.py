"""CPU checks of the product library: it loads, exports every symbol the C header
declares, and its host-side setup code (lattice generators, unit cell, bond search)
reproduces the oracle bit for bit. No GPU is touched."""
import math

import numpy as np
import pytest

import lpmgpu as G
import oracle as O


def test_library_exports_every_declared_symbol():
    L = G.lib()
    syms = G.header_symbols()
    assert len(syms) >= 35
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


CASES = [
    (O.HEXAGON2D, [0, 6, 0, 5, 0, 1], 0.25, 0, (0, 0, 0)),
    (O.HEXAGON2D, [0, 6, 0, 5, 0, 1], 0.25, 0, (0, 0, 0.3)),
    (O.SQUARE2D, [0, 4, 0, 4, 0, 1], 0.2, 0, (0, 0, 0)),
    (O.SC3D, [0, 3, 0, 3, 0, 3], 0.25, 0, (0, 0, 0)),
    (O.FCC3D, [0, 3, 0, 3, 0, 3], 0.25, 0, (0.4, 1.1, 2.0)),
    (O.FCC3D, [0, 3, 0, 3, 0, 3], 0.25, 2, (0.4, 1.1, 2.0)),
    (O.SC3D, [0, 3, 0, 3, 0, 3], 0.25, 1, (0.2, 0.3, 0.5)),
]


@pytest.mark.parametrize("lat,box,r,flag,ang", CASES)
def test_generators_match_oracle_bitwise(lat, box, r, flag, ang):
    a = G.generate_lattice(lat, box, r, flag, ang)
    b = O.generate(lat, box, r, flag, ang)
    assert a.shape == b.shape and len(a) > 0
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("lat,box,r,flag,ang", CASES)
def test_bond_search_matches_oracle_bitwise(lat, box, r, flag, ang):
    xyz = O.generate(lat, box, r, flag, ang)
    bp, bs, p2 = G.build_bonds(lat, r, xyz)
    s = O.System(lat, r, O.ELASTIC, xyz, bruteforce=True)
    np.testing.assert_array_equal(bp, s.geti("bptr"))
    np.testing.assert_array_equal(bs, s.geti("bsplit"))
    np.testing.assert_array_equal(p2, s.geti("p2"))


def test_unit_cell_and_ce():
    for lat, r in ((G.HEXAGON2D, 0.2), (G.SC3D, 0.5), (G.FCC3D, 0.3), (G.SQUARE2D, 0.1), (G.BCC3D, 0.4)):
        out, lit = G.unit_cell(lat, r, literal=1)
        _, rowm = G.unit_cell(lat, r, literal=0)
        np.testing.assert_array_equal(lit, rowm.T)
        assert out[5] == 2 * r
    out, _ = G.unit_cell(G.SC3D, 0.5)
    assert out[0] == 3 and out[1] == 18 and out[4] == 60 and out[7] == 1.0
    out, _ = G.unit_cell(G.HEXAGON2D, 0.5)
    assert out[0] == 2 and out[1] == 12 and out[6] == pytest.approx(math.sqrt(3))
    E, mu = 69e3, 0.3
    np.testing.assert_allclose(G.ce_from_e(False, E, mu), [E * (1 - mu) / (1 + mu) / (1 - 2 * mu),
                                                           E * mu / (1 + mu) / (1 - 2 * mu), E / 2 / (1 + mu)], rtol=0)


def test_bad_topology_is_rejected_without_launching():
    """An asymmetric bond list is refused on the host (the reference would dereference end(),
    particle.h:207-212); so is a bad lattice id. No GPU is needed to get these errors."""
    xyz = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    bp = np.array([0, 1, 1], dtype=np.int32)
    bs = np.array([1, 1], dtype=np.int32)
    p2 = np.array([1], dtype=np.int32)
    with pytest.raises(G.LPMError, match="reverse bond"):
        G.Context(G.SC3D, 0.5, G.ELASTIC, xyz, bonds=(bp, bs, p2))
    with pytest.raises(G.LPMError, match="lattice"):
        G.Context(17, 0.5, G.ELASTIC, xyz, bonds=(bp, bs, p2))


def test_dump_number_formatting_matches_printf():
    """The dump writer's "%.4e" (an exact fast path with a printf fallback near ties) against the
    C library's own snprintf, on values of every magnitude, exact ties, values one ulp either side
    of ties and decade boundaries, signed zeros, infinities and NaN."""
    import ctypes
    import ctypes.util

    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.snprintf.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_double]

    def printf_e4(v):
        buf = ctypes.create_string_buffer(64)
        libc.snprintf(buf, 64, b"%.4e", float(v))
        return buf.value.decode()

    rng = np.random.default_rng(0)
    vals = [0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -1.0, 5e-324, 1.7976931348623157e308]
    vals += list(rng.standard_normal(20000) * 10.0 ** rng.integers(-30, 30, 20000))
    vals += list(rng.uniform(-1, 1, 20000))
    for e in range(-25, 26):
        for k in (10000, 12345, 99999, 50000, 99995, 10005):
            for frac in (0.0, 0.5):
                t = (k + frac) * 10.0 ** (e - 4)
                vals += [t, np.nextafter(t, np.inf), np.nextafter(t, -np.inf), -t]
        p = 10.0 ** e
        vals += [p, np.nextafter(p, 0), np.nextafter(p, np.inf), 9.99995 * p, 9.99994999 * p, 9.999950001 * p]
    vals = np.array(vals, dtype=np.float64)
    got = G.format_e4(vals)
    want = [printf_e4(v) for v in vals]
    bad = [(v, g, w) for v, g, w in zip(vals, got, want) if g != w]
    assert not bad, bad[:10]


def test_noalloc_entry_point_refuses_sizes_beyond_its_array():
    """lpm_update_rr_noalloc checks sizes on the host before any device call: above
    LPM_NOALLOC_MAX_DOFS it returns LPM_E_LIMIT with a message (no GPU needed)."""
    import ctypes

    import lpmgpu as G

    L = G.lib()
    p = ctypes.c_void_p(8)
    rc = L.lpm_update_rr_noalloc(8388608 // 3 + 1, 3, p, p, p, p, None, None)
    assert rc == -6 and b"exceed the limit" in L.lpm_last_error(None)
    assert L.lpm_update_rr_noalloc(10, 4, p, p, p, p, None, None) == -1
