"""The C ABI library loads and exports what include/pumas_gpu.h declares;
argument checks that fail before any device work (CPU only)."""
import ctypes
import os
import re

import pytest

from conftest import REPO

import pumas_amd as pa
from pumas_amd import _abi

HEADER = os.path.join(REPO, "include", "pumas_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(pumas_gpu_\w+)\s*\(", text)
    return sorted(set(names))


def test_header_declarations_are_listed():
    assert declared_functions() == sorted(_abi.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(pa.LIBRARY_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_version_and_defaults():
    lib = pa.library()
    assert lib.pumas_gpu_version().decode().startswith("pumas_hip")
    c = _abi.GpuContext()
    lib.pumas_gpu_context_default(ctypes.byref(c))
    d = pa.Context()
    assert (c.mode.energy_loss, c.mode.decay, c.mode.direction,
            c.mode.scattering, c.event, c.accuracy) == (
        d.energy_loss, d.decay, d.direction, d.scattering, d.event,
        d.accuracy)


def test_struct_layouts():
    assert ctypes.sizeof(_abi.GpuContext) == 4 * 4 + 4 + 4 + 4 * 8 + 8
    assert pa.STATE_DTYPE.itemsize == 104
    assert pa.STATE_DTYPE.fields["position"][1] == 48
    assert pa.STATE_DTYPE.fields["decayed"][1] == 96


def test_upload_errors_without_device(tmp_path):
    with pytest.raises(pa.PumasError) as e:
        pa.Physics(tmp_path / "missing.ptab")
    assert e.value.code == pa.Return.PATH_ERROR
    bad = tmp_path / "bad.ptab"
    bad.write_bytes(b"NOTAPTAB" + bytes(200))
    with pytest.raises(pa.PumasError) as e:
        pa.Physics(bad)
    assert e.value.code == pa.Return.FORMAT_ERROR


def test_transport_null_arguments():
    lib = pa.library()
    ctx = pa.Context().to_c()
    rc = lib.pumas_gpu_transport(None, ctypes.byref(ctx), None, None, 0, 0,
                                 0, None, None, None, None)
    assert rc == pa.Return.VALUE_ERROR


def test_bench_refuses_several_gpus():
    """One GPU per box here: bench.py --gpus N > 1 exits with a message and
    a non zero status instead of launching anything."""
    import json
    import subprocess
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    r = subprocess.run([sys.executable, os.path.join(here, "..", "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_ctypes_mirrors_match_the_header(tmp_path):
    """The ctypes structures of the binding have the sizes and member
    offsets that the C compiler gives the header's structures."""
    checks = {
        "GpuOptions": ("struct pumas_gpu_options",
                       ["precision", "recorder", "tail_del", "lds_tables",
                        "lanes_per_wave"]),
        "GpuStats": ("struct pumas_gpu_stats", ["rounds", "lds_bytes"]),
        "GpuOutputs": ("struct pumas_gpu_outputs", ["rc"]),
        "TallySpec": ("struct pumas_gpu_tally_spec", ["cos_max"]),
        "RequestStatus": ("struct pumas_gpu_request_status", ["message"]),
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>',
             '#include "pumas_gpu.h"', 'int main(void) {']
    for name, (cname, fields) in checks.items():
        lines.append(f'printf("{name} size %zu\\n", sizeof({cname}));')
        for f in fields:
            lines.append(f'printf("{name} {f} %zu\\n", '
                         f'offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    import subprocess
    subprocess.check_call(["gcc", "-std=c99", "-I",
                           os.path.join(REPO, "include"), "-o", str(exe),
                           str(src)])
    out = subprocess.check_output([str(exe)], text=True).split("\n")
    for line in filter(None, out):
        name, what, value = line.split()
        cls = getattr(_abi, name)
        if what == "size":
            assert ctypes.sizeof(cls) == int(value), line
        else:
            assert getattr(cls, what).offset == int(value), line


class ReferenceContext(ctypes.Structure):
    """Public layout of the reference's struct pumas_context
    (include/pumas.h:636-672): four host pointers, then the value part."""
    _fields_ = [("medium", ctypes.c_void_p), ("random", ctypes.c_void_p),
                ("recorder", ctypes.c_void_p), ("user_data", ctypes.c_void_p),
                ("mode", _abi.ContextMode), ("event", ctypes.c_int),
                ("limit", _abi.ContextLimit), ("accuracy", ctypes.c_double)]


def test_context_from_reference():
    """The value part of a reference context is read at the reference's
    offsets (no device work)."""
    lib = pa.library()
    r = ReferenceContext()
    r.medium, r.random, r.recorder, r.user_data = 1, 2, 3, 4
    r.mode.energy_loss, r.mode.decay = pa.Mode.STRAGGLED, pa.Mode.RANDOMISED
    r.mode.direction, r.mode.scattering = pa.Mode.BACKWARD, pa.Mode.MIXED
    r.event = int(pa.Event.LIMIT_DISTANCE | pa.Event.MEDIUM)
    r.limit.energy, r.limit.distance = 0.5, 1234.5
    r.limit.grammage, r.limit.time = 6.5e5, 3e-6
    r.accuracy = 0.02
    c = _abi.GpuContext()
    lib.pumas_gpu_context_from_reference(ctypes.byref(r), ctypes.byref(c))
    assert (c.mode.energy_loss, c.mode.decay, c.mode.direction,
            c.mode.scattering) == (2, 1, 1, 1)
    assert c.event == r.event
    assert (c.limit.energy, c.limit.distance, c.limit.grammage,
            c.limit.time) == (0.5, 1234.5, 6.5e5, 3e-6)
    assert c.accuracy == 0.02


def test_c_caller_compiles_against_reference_header():
    """tests/abi_check.c includes the reference's include/pumas.h before
    pumas_gpu.h and static-asserts the shared layouts and enum values; it
    is built (tests/kernels/Makefile) wherever the reference is present."""
    ref = "/root/reference/include/pumas.h"
    if not os.path.exists(ref):
        pytest.skip("reference header not present")
    exe = os.path.join(REPO, "tests", "_build", "abi_check")
    import subprocess
    subprocess.check_call(["make", "-s", "-C",
                           os.path.join(REPO, "tests", "kernels"),
                           "../_build/abi_check"])
    assert os.access(exe, os.X_OK)
