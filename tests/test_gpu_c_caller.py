"""A plain C program calls libpumas_hip on golden cases.

tests/abi_check.c (built by tests/kernels/Makefile against the reference's
include/pumas.h, which it includes before pumas_gpu.h) fills a reference
`struct pumas_context`, converts it with pumas_gpu_context_from_reference
and transports the golden initial states through pumas_gpu_transport_host.
Its outputs must be the reference's golden results at the FP64 tolerance
(1e-12 on every history, same events, media, step and draw counts).
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import load_golden
from parity import assert_parity

import make_golden
import pumas_amd as pa
from pumas_amd import STATE_DTYPE

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "_build", "abi_check")
CASES = {c[0]: c for c in make_golden.cases()}


class Result:
    pass


@pytest.mark.parametrize("name", ["slabs_straggled", "slabs_backward",
                                  "limits_forward"])
def test_c_caller_golden(tmp_path, name):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is missing: run make -C tests/kernels")
    _, pname, ctx, desc, _ = CASES[name]
    if desc.kind != pa.GeometryKind.SLABS or desc.step_type != pa.Step.RAW:
        pytest.skip("the C caller builds the raw slab stack")
    g = load_golden(name)
    n = len(g["initial"])
    src, out = tmp_path / "states.bin", tmp_path / "out.bin"
    g["initial"].tofile(src)
    args = [EXE, pa.physics_path(pname), str(src), str(out),
            str(int(g["seed"])), str(int(ctx.energy_loss)),
            str(int(ctx.scattering)), str(int(ctx.direction)),
            str(int(ctx.decay)), str(int(ctx.event)),
            repr(ctx.limit_energy), repr(ctx.limit_distance),
            repr(ctx.limit_grammage), repr(ctx.limit_time)]
    print(f"c caller {name}: {n} histories", flush=True)
    p = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0
    raw = np.fromfile(out, np.uint8)
    r = Result()
    off = n * STATE_DTYPE.itemsize
    r.states = raw[:off].view(STATE_DTYPE)
    rest = raw[off:].view(np.int32)
    r.events, r.media = rest[:n], rest[n:3 * n].reshape(n, 2)
    r.steps = rest[3 * n:4 * n].view(np.uint32)
    r.draws = rest[4 * n:5 * n].view(np.uint32)
    r.rc = rest[5 * n:6 * n]
    assert_parity(r, g, 1e-12)
