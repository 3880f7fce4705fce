"""The engine's math functions (pumas_amd/csrc/libm.h) against the host C
library, on the CPU: the engine compiled for the host must return the C
library's bits (float logf / expf on all 2^32 inputs,
double functions on 1e8 random arguments each). The GPU build of the same
functions is compared in test_gpu_libm.py."""
import ctypes
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "liblibm_check.so")


@pytest.fixture(scope="module")
def lib():
    os.makedirs(BUILD, exist_ok=True)
    src = os.path.join(HERE, "libm_check.cpp")
    csrc = os.path.join(REPO, "pumas_amd", "csrc")
    deps = [src, os.path.join(csrc, "libm.h"),
            os.path.join(csrc, "libm_tables.h")]
    if (not os.path.exists(LIB)) or any(
            os.path.getmtime(d) > os.path.getmtime(LIB) for d in deps):
        subprocess.check_call(
            ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp",
             "-ffp-contract=off", "-mfma", "-I", csrc, "-o",
             f"{LIB}.{os.getpid()}.tmp", src])
        os.replace(f"{LIB}.{os.getpid()}.tmp", LIB)
    lib = ctypes.CDLL(LIB)
    lib.check_float.restype = ctypes.c_long
    lib.check_float.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.c_long]
    lib.check_double.restype = ctypes.c_long
    lib.check_double.argtypes = [ctypes.c_int, ctypes.c_long,
                                 ctypes.c_uint64]
    return lib


@pytest.mark.parametrize("which,name", [(0, "logf"), (1, "expf")])
def test_float_functions_match_c_library(lib, which, name):
    assert lib.check_float(which, 1, 0) == 0, name


@pytest.mark.parametrize("which,name", [(0, "exp"), (1, "log"), (2, "pow"),
                                        (3, "sin"), (4, "cos"),
                                        (5, "sincos"), (6, "log10")])
def test_double_functions_match_c_library(lib, which, name):
    assert lib.check_double(which, 100_000_000, 12345 + which) == 0, name
