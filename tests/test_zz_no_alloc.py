"""pumas_gpu_transport_noalloc against pumas_gpu_transport, through the C
ABI: the entry point that allocates nothing (its buffers are the library's
static device array) must give the same states and outputs, bit for bit,
on two calls in a row, the second with a smaller batch (the array keeps
the first call's contents), and must refuse a batch beyond its maximum."""
import ctypes

import numpy as np
import pytest

import pumas_amd as pa
from pumas_amd import _abi
from pumas_amd import workloads as W

pytestmark = pytest.mark.gpu


def _device_call(fn, physics, ctx, geom, states, seed):
    import torch
    n = len(states)
    d = torch.from_numpy(states.view(np.uint8).reshape(-1).copy()).cuda()
    outs = {k: torch.zeros(n * m, dtype=torch.int32, device="cuda")
            for k, m in (("events", 1), ("media", 2), ("steps", 1),
                         ("draws", 1), ("rc", 1))}
    o = _abi.GpuOutputs()
    for k, t in outs.items():
        setattr(o, k, t.data_ptr())
    opts = pa._options(pa.Precision.FP64, 0, 0, 0)
    st = _abi.GpuStats()
    c = ctx.to_c()
    torch.cuda.synchronize()
    rc = fn(physics.handle, ctypes.byref(c), geom.handle, d.data_ptr(), n,
            seed, 0, ctypes.byref(o), ctypes.byref(opts), ctypes.byref(st),
            None)
    torch.cuda.synchronize()
    out = {k: t.cpu().numpy() for k, t in outs.items()}
    out["states"] = d.cpu().numpy().view(_abi.STATE_DTYPE).copy()
    return rc, out, st


@pytest.mark.parametrize("case", ["slabs_forward", "slabs_backward_mixed"])
def test_noalloc_equals_transport(gpu_standard, case):
    lib = pa.library()
    if case == "slabs_forward":
        ctx = W.slabs_forward_context()
        big, small = W.slabs_forward_states(20000, 31), \
            W.slabs_forward_states(3000, 32)
    else:
        ctx = W.slabs_backward_context(energy_loss=pa.Mode.MIXED)
        big, small = W.slabs_backward_states(20000, seed=31), \
            W.slabs_backward_states(3000, seed=32)
    geom = pa.Geometry(gpu_standard, W.slabs_geometry())
    for states, seed in ((big, 31), (small, 32)):
        print(f"noalloc {case}: {len(states)} histories", flush=True)
        rc0, ref, st0 = _device_call(lib.pumas_gpu_transport, gpu_standard,
                                     ctx, geom, states, seed)
        rc1, got, st1 = _device_call(lib.pumas_gpu_transport_noalloc,
                                     gpu_standard, ctx, geom, states, seed)
        assert rc0 == rc1 == 0
        assert got["states"].tobytes() == ref["states"].tobytes()
        for k in ("events", "media", "steps", "draws", "rc"):
            np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
        assert st1.steps == st0.steps and st1.histories == len(states)
        assert st1.kernel_ms == 0.0


def test_noalloc_refuses_large_batches(gpu_standard):
    import torch
    lib = pa.library()
    geom = pa.Geometry(gpu_standard, W.slabs_geometry())
    n = _abi.NOALLOC_MAX_HISTORIES + 1
    d = torch.zeros(n * _abi.STATE_DTYPE.itemsize, dtype=torch.uint8,
                    device="cuda")
    c = W.slabs_forward_context().to_c()
    rc = lib.pumas_gpu_transport_noalloc(gpu_standard.handle, ctypes.byref(c),
                                         geom.handle, d.data_ptr(), n, 0, 0,
                                         None, None, None, None)
    assert rc == int(pa.Return.VALUE_ERROR)
    assert b"exceed" in lib.pumas_gpu_error_message()
    torch.cuda.synchronize()
    assert int(d.sum()) == 0  # nothing ran
