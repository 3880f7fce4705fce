"""Function-level parity of the engine's physics routines on the GPU.

The same comparison as test_functions.py (exact and tabulated DCS of every
process and element, the dilogarithm, the DEL polar angle samplers with
their random draws), with the routines built for gfx950 (test library
tests/kernels/probe.hip, on the PTAB table copied to the device): every
value must be the reference's bit for bit.
"""
import ctypes
import os

import numpy as np
import pytest

import pumas_amd as pa

import functions_grid as G
from test_functions import assert_same, n_elements, reference_dcs

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PROBE = os.path.join(HERE, "_build", "libpumas_probe.so")


@pytest.fixture(scope="module")
def probe():
    if not os.path.exists(PROBE):
        pytest.fail(f"{PROBE} is missing: run make -C tests/kernels")
    lib = ctypes.CDLL(PROBE)
    lib.probe_load.argtypes = [ctypes.c_char_p]
    lib.probe_load.restype = ctypes.c_int
    lib.probe_functions.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_uint64, ctypes.c_size_t]
    lib.probe_functions.restype = ctypes.c_int
    return lib


def device_values(lib, which, ip, iel, a, b, seed=0):
    out = np.zeros(len(a))
    draws = np.zeros(len(a), np.uint32)
    assert lib.probe_functions(which, ip, iel, a.ctypes.data, b.ctypes.data,
                               out.ctypes.data, draws.ctypes.data, seed,
                               len(a)) == 0
    return out, draws


@pytest.mark.parametrize("physics", ["standard_c005", "iron_c001"])
@pytest.mark.parametrize("which", [0, 1], ids=["exact", "tabulated"])
def test_dcs_bit_exact(probe, request, physics, which):
    ref = request.getfixturevalue("ref_standard" if physics ==
                                  "standard_c005" else "ref_iron")
    assert probe.probe_load(pa.physics_path(physics).encode()) == 0
    K, q = G.dcs_grid()
    for ip in G.PROCESSES:
        for iel in range(n_elements(physics)):
            print(f"dcs {which} {physics} process {ip} element {iel}",
                  flush=True)
            got, _ = device_values(probe, which, ip, iel, K, q)
            want = reference_dcs(ref, which, ip, iel, K, q)
            assert_same(got, want, f"{G.PROCESSES[ip]} element {iel}")


def test_dilog_bit_exact(probe, ref_standard):
    assert probe.probe_load(pa.physics_path("standard_c005").encode()) == 0
    x, y = G.dilog_grid()
    got, _ = device_values(probe, 2, 0, 0, x, y)
    want = np.array([ref_standard.lib.ref_dilog(v) for v in x])
    assert_same(got, want, "dilog")


@pytest.mark.parametrize("physics", ["standard_c005", "iron_c001"])
def test_polar_samplers_bit_exact(probe, request, physics):
    ref = request.getfixturevalue("ref_standard" if physics ==
                                  "standard_c005" else "ref_iron")
    assert probe.probe_load(pa.physics_path(physics).encode()) == 0
    ki, kf = G.polar_grid()
    for ip in G.PROCESSES:
        for iel in range(n_elements(physics)):
            print(f"polar {physics} process {ip} element {iel}", flush=True)
            got, d = device_values(probe, 3, ip, iel, ki, kf, seed=5)
            want, dw = ref.polar(ip, iel, ki, kf, seed=5)
            assert_same(got, want, f"polar {G.PROCESSES[ip]} element {iel}")
            np.testing.assert_array_equal(d, dw)
