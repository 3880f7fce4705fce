"""The engine's math functions on the GPU against the host C library.

The transport engine calls pumas_amd/csrc/libm.h for every transcendental
function of the FP64 path. Built for gfx950 (test library
tests/kernels/probe.hip), each must return the host C library's bits on
every argument tried (1e7 per function, over the ranges the transport
uses and the whole range). The device library's own functions (OCML),
which the engine no longer calls on the FP64 path, are compared the same
way and their mismatch rates printed: they are the reason for libm.h.
sqrt and division are correctly rounded on both sides and must agree.
"""
import ctypes
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PROBE = os.path.join(HERE, "_build", "libpumas_probe.so")

F_EXP, F_LOG, F_POW, F_SIN, F_COS, F_SINCOS_S, F_SINCOS_C, F_LOG10 = range(8)
F_LOGF, F_EXPF, F_SQRT, F_DIV, F_ATAN = range(8, 13)
F_OCML = 100
N = 10_000_000


@pytest.fixture(scope="module")
def probe():
    if not os.path.exists(PROBE):
        pytest.fail(f"{PROBE} is missing: run make -C tests/kernels")
    lib = ctypes.CDLL(PROBE)
    P = ctypes.c_void_p
    lib.probe_math.argtypes = [ctypes.c_int, P, P, P, ctypes.c_size_t]
    lib.probe_math.restype = ctypes.c_int
    lib.probe_host_math.argtypes = [ctypes.c_int, P, P, P, ctypes.c_size_t]
    lib.probe_host_math.restype = None
    return lib


def arguments(which, n=N, seed=0):
    rng = np.random.default_rng(seed + which)
    q = n // 4
    y = np.zeros(n)
    if which in (F_EXP,):
        x = np.concatenate([rng.uniform(-745, 745, q), rng.uniform(-20, 20, q),
                            rng.uniform(-1, 1, q), -np.abs(rng.normal(0, 50, n - 3 * q))])
    elif which in (F_LOG, F_LOG10):
        x = np.concatenate([rng.random(q) + 1e-300, rng.uniform(0.9, 1.1, q),
                            np.exp(rng.uniform(-30, 30, q)),
                            rng.integers(1, 0x7ff0000000000000, n - 3 * q,
                                         dtype=np.int64).view(np.float64)])
    elif which == F_POW:
        x = np.concatenate([rng.random(q) * 10, np.exp(rng.uniform(-50, 50, q)),
                            rng.random(q), rng.random(n - 3 * q) * 1e4])
        y = np.concatenate([rng.uniform(-3, 3, q), rng.uniform(-10, 10, q),
                            1.0 / (1.0 - rng.random(q) * 3),
                            rng.uniform(-2, 2, n - 3 * q)])
    elif which in (F_SIN, F_COS, F_SINCOS_S, F_SINCOS_C, F_ATAN):
        x = np.concatenate([rng.uniform(-0.9, 0.9, q), rng.uniform(-2.5, 2.5, q),
                            rng.uniform(-7, 7, q), rng.uniform(-1e8, 1e8, n - 3 * q)])
    elif which in (F_LOGF, F_EXPF):
        # every 2^32 / n-th float bit pattern (finite ones)
        u = (np.arange(n, dtype=np.uint64) * ((1 << 32) // n)).astype(np.uint32)
        x = u.view(np.float32).astype(np.float64)
        x = x[np.isfinite(x)]
        y = np.zeros(len(x))
    elif which == F_SQRT:
        x = np.concatenate([rng.random(q), np.exp(rng.uniform(-700, 700, n - q))])
    elif which == F_DIV:
        x = np.exp(rng.uniform(-300, 300, n))
        y = np.exp(rng.uniform(-300, 300, n)) * rng.choice([-1, 1], n)
    return np.ascontiguousarray(x), np.ascontiguousarray(y)


def evaluate(probe, which, x, y):
    out = np.empty_like(x)
    assert probe.probe_math(which, x.ctypes.data, y.ctypes.data,
                            out.ctypes.data, len(x)) == 0
    return out


def host(probe, which, x, y):
    out = np.empty_like(x)
    probe.probe_host_math(which, x.ctypes.data, y.ctypes.data,
                          out.ctypes.data, len(x))
    return out


def mismatches(a, b):
    return int(np.count_nonzero((a.view(np.int64) != b.view(np.int64)) &
                                ~(np.isnan(a) & np.isnan(b))))


CASES = [(F_EXP, "exp"), (F_LOG, "log"), (F_POW, "pow"), (F_SIN, "sin"),
         (F_COS, "cos"), (F_SINCOS_S, "sincos (sin)"),
         (F_SINCOS_C, "sincos (cos)"), (F_LOG10, "log10"), (F_LOGF, "logf"),
         (F_EXPF, "expf")]


@pytest.mark.gpu
@pytest.mark.parametrize("which,name", CASES)
def test_engine_math_matches_c_library(probe, which, name):
    x, y = arguments(which)
    ref = host(probe, which, x, y)
    got = evaluate(probe, which, x, y)
    bad = mismatches(got, ref)
    ocml_which = (F_OCML + which) if which not in (F_SINCOS_S, F_SINCOS_C) \
        else F_OCML + (F_SIN if which == F_SINCOS_S else F_COS)
    ocml = mismatches(evaluate(probe, ocml_which, x, y), ref)
    print(f"{name}: engine {bad} / {len(x)} differ from the C library; "
          f"device library {ocml} ({ocml / len(x):.2e})", flush=True)
    assert bad == 0


@pytest.mark.gpu
@pytest.mark.parametrize("which,name", [(F_SQRT, "sqrt"), (F_DIV, "division")])
def test_correctly_rounded_operations(probe, which, name):
    x, y = arguments(which)
    bad = mismatches(evaluate(probe, which, x, y), host(probe, which, x, y))
    print(f"{name}: {bad} / {len(x)} differ", flush=True)
    assert bad == 0


@pytest.mark.gpu
def test_atan_rate_reported(probe):
    """atan (EHS screening set-up) stays the device library's: a wrong last
    bit on every call moved no history beyond 1e-12 in a host experiment.
    Its mismatch rate is reported, not asserted."""
    x, y = arguments(F_ATAN)
    bad = mismatches(evaluate(probe, F_ATAN, x, y), host(probe, F_ATAN, x, y))
    print(f"atan: device library {bad} / {len(x)} ({bad / len(x):.2e})",
          flush=True)
