"""Function-level parity of the engine's physics routines, on the CPU.

The engine compiled for the host (tests/host_engine.cpp) evaluates, on the
grids of functions_grid.py and for every element of the table: the exact
radiative and ionisation DCS models (dcs_bremsstrahlung_SSR,
dcs_pair_production_SSR, dcs_photonuclear_DRSS, dcs_ionisation through
the reference's dcs_get, src/pumas.c:11058-15170), the tabulated DCS with
its exact model fallbacks (dcs_evaluate, src/pumas.c:11243-11347), the
dilogarithm (math_dilog, src/pumas.c:12609-12686) and the polar angle
samplers of discrete energy losses (src/pumas.c:11736-11976, Brent's
minimiser included), with the same Philox stream as the oracle. Every
value must be the reference's bit for bit (and every polar sample must
consume the same number of draws). tests/test_gpu_functions.py runs the
same comparison on the GPU build.
"""
import ctypes

import numpy as np
import pytest

import pumas_amd as pa

import functions_grid as G
from test_host_engine import build_host


@pytest.fixture(scope="module")
def host():
    lib = build_host()
    lib.host_functions.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_uint64, ctypes.c_int64]
    lib.host_functions.restype = None
    return lib


def engine_values(lib, which, ip, iel, a, b, seed=0):
    out = np.zeros(len(a))
    draws = np.zeros(len(a), np.uint32)
    lib.host_functions(which, ip, iel, a.ctypes.data, b.ctypes.data,
                       out.ctypes.data, draws.ctypes.data, seed, len(a))
    return out, draws


def reference_dcs(ref, which, ip, iel, K, q):
    f = ref.dcs_exact if which == 0 else ref.dcs_evaluate
    return np.array([f(ip, iel, k, x) for k, x in zip(K, q)])


def assert_same(got, want, what):
    bad = (got.view(np.int64) != want.view(np.int64)) & \
        ~(np.isnan(got) & np.isnan(want))
    assert not bad.any(), (f"{what}: {int(bad.sum())} of {len(got)} differ, "
                           f"first at {np.nonzero(bad)[0][:5]}: "
                           f"{got[bad][:3]} vs {want[bad][:3]}")


@pytest.mark.parametrize("physics", ["standard_c005", "iron_c001"])
@pytest.mark.parametrize("which", [0, 1], ids=["exact", "tabulated"])
def test_dcs_bit_exact(host, request, physics, which):
    ref = request.getfixturevalue("ref_standard" if physics ==
                                  "standard_c005" else "ref_iron")
    assert host.host_load(pa.physics_path(physics).encode()) == 0
    K, q = G.dcs_grid()
    for ip in G.PROCESSES:
        for iel in range(n_elements(physics)):
            got, _ = engine_values(host, which, ip, iel, K, q)
            want = reference_dcs(ref, which, ip, iel, K, q)
            assert np.count_nonzero(want) > 0
            assert_same(got, want, f"{G.PROCESSES[ip]} element {iel}")


def test_dilog_bit_exact(host, ref_standard):
    assert host.host_load(pa.physics_path("standard_c005").encode()) == 0
    x, y = G.dilog_grid()
    got, _ = engine_values(host, 2, 0, 0, x, y)
    want = np.array([ref_standard.lib.ref_dilog(v) for v in x])
    assert_same(got, want, "dilog")


@pytest.mark.parametrize("physics", ["standard_c005", "iron_c001"])
def test_polar_samplers_bit_exact(host, request, physics):
    ref = request.getfixturevalue("ref_standard" if physics ==
                                  "standard_c005" else "ref_iron")
    assert host.host_load(pa.physics_path(physics).encode()) == 0
    ki, kf = G.polar_grid()
    for ip in G.PROCESSES:
        for iel in range(n_elements(physics)):
            got, d = engine_values(host, 3, ip, iel, ki, kf, seed=5)
            want, dw = ref.polar(ip, iel, ki, kf, seed=5)
            assert_same(got, want, f"polar {G.PROCESSES[ip]} element {iel}")
            np.testing.assert_array_equal(d, dw)


def n_elements(physics):
    """Elements of a PTAB table (header: magic[8], version, n_entries,
    imeta[] with PTAB_N_ELEMENTS at index 3)."""
    import struct
    data = open(pa.physics_path(physics), "rb").read(16 + 4 * 4)
    return struct.unpack_from("<4i", data, 16)[3]


def _row_probes(table, rng):
    """Values in [table[0], table[-1]): every entry, its neighbours in
    floating point, the midpoints of the rows and random values."""
    t = np.asarray(table, float)
    x = np.concatenate([t, np.nextafter(t, np.inf), np.nextafter(t, -np.inf),
                        0.5 * (t[:-1] + t[1:]),
                        rng.uniform(t[0], t[-1], 4096)])
    return x[(x >= t[0]) & (x < t[-1])]


def test_fresh_row_lookups(host):
    """The row lookups of a history without cached rows (closed form CSDA)
    return the bracketing row: the largest i with table[i] <= x, computed
    here with numpy's sorted search, on increasing tables of 2 to 1000 rows,
    tables with repeated values, and the kinetic energy grid of both
    physics (whose lookups go through the coarse index). The search with a
    first round of 15 pivots over the whole table (the device reads them
    with scalar loads when a wavefront shares the table) is checked too."""
    host.host_rows.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                               ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_int64]
    host.host_rows.restype = None
    host.host_k_grid.argtypes = [ctypes.c_void_p, ctypes.c_int]
    host.host_k_grid.restype = ctypes.c_int
    rng = np.random.default_rng(5)

    def rows(which, table, x):
        table = np.ascontiguousarray(table, float)
        out = np.full(len(x), -7, np.int32)
        host.host_rows(which, table.ctypes.data, len(table), x.ctypes.data,
                       out.ctypes.data, len(x))
        return out

    tables = [np.cumsum(rng.uniform(0.01, 1.0, n))
              for n in (2, 3, 15, 16, 17, 18, 33, 146, 257, 300, 1000)]
    plateau = np.cumsum(rng.uniform(0.01, 1.0, 146))
    plateau[40:60] = plateau[40]
    plateau[100:103] = plateau[100]
    tables += [plateau, np.logspace(-3, 9, 146)]
    for t in tables:
        x = _row_probes(t, rng)
        want = np.searchsorted(t, x, side="right") - 1
        np.testing.assert_array_equal(rows(0, t, x), want,
                                      err_msg=f"{len(t)} rows")
        np.testing.assert_array_equal(rows(2, t, x), want,
                                      err_msg=f"{len(t)} rows, uniform")
    for physics in ("standard_c005", "iron_c001"):
        assert host.host_load(pa.physics_path(physics).encode()) == 0
        K = np.zeros(4096)
        n = host.host_k_grid(K.ctypes.data, len(K))
        K = K[:n]
        x = _row_probes(K, rng)
        x = np.concatenate([x, np.exp(rng.uniform(np.log(K[1]),
                                                  np.log(K[-1]), 20000))])
        x = x[(x >= K[0]) & (x < K[-1])]
        want = np.searchsorted(K, x, side="right") - 1
        np.testing.assert_array_equal(rows(1, K, x), want, err_msg=physics)

