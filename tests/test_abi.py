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
