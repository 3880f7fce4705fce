"""Device frame recorder against the reference's recorder.

Frames of a history are recorded in time order by the lane that carries
it; frames of different histories interleave. Grouped by history (stable
order), the device frames must be the reference's: same count, events,
media and history ids, states within 1e-12.
"""
import numpy as np
import pytest

import pumas_amd as pa
from pumas_amd import workloads as W

import make_golden
from test_gpu_parity import RTOL

pytestmark = pytest.mark.gpu


def grouped(frames):
    order = np.lexsort((np.arange(len(frames)), frames["history"]))
    return frames[order]


@pytest.mark.parametrize("period", [1, 5, 0])
@pytest.mark.parametrize("case", ["slabs_straggled", "slabs_backward",
                                  "csda_rock"])
def test_frames_vs_reference_recorder(gpu_standard, ref_standard,
                                      case, period):
    import oracle as orc
    name, pname, ctx, desc, states = next(
        c for c in make_golden.cases() if c[0] == case)
    states = states[:64]
    rec = pa.Recorder(1 << 18, period=period, first_history=0,
                      n_histories=len(states))
    gpu = pa.transport(gpu_standard, ctx, pa.Geometry(gpu_standard, desc),
                       states.copy(), recorder=rec)
    frames, count = rec.frames()
    fin, ref = orc.transport_record(ref_standard, ctx,
                                    orc.RefGeometry(desc),
                                    states, period=period,
                                    capacity=1 << 18)
    assert count == ref["count"] == len(frames)
    f = grouped(frames)
    np.testing.assert_array_equal(f["history"], ref["history"])
    np.testing.assert_array_equal(f["event"], ref["events"])
    np.testing.assert_array_equal(f["medium"], ref["media"])
    for k in ("energy", "distance", "grammage", "time", "weight"):
        np.testing.assert_allclose(f["state"][k], ref["states"][k],
                                   rtol=RTOL, atol=1e-300)
    np.testing.assert_allclose(f["state"]["position"],
                               ref["states"]["position"], rtol=0,
                               atol=1e-6)
    np.testing.assert_allclose(f["state"]["direction"],
                               ref["states"]["direction"], rtol=0,
                               atol=1e-9)
    # Recording changes no result.
    plain = pa.transport(gpu_standard, ctx, pa.Geometry(gpu_standard, desc),
                         states.copy())
    assert (plain.states == gpu.states).all()


def test_history_range_and_capacity(gpu_standard):
    s = W.slabs_forward_states(256, seed=3)
    geom = pa.Geometry(gpu_standard, W.slabs_geometry())
    ctx = W.slabs_forward_context()
    rec = pa.Recorder(1 << 16, period=1, first_history=100, n_histories=3)
    pa.transport(gpu_standard, ctx, geom, s.copy(), recorder=rec)
    frames, count = rec.frames()
    assert set(np.unique(frames["history"]).tolist()) == {100, 101, 102}
    small = pa.Recorder(10, period=1, first_history=0, n_histories=256)
    pa.transport(gpu_standard, ctx, geom, s.copy(), recorder=small)
    frames, count = small.frames()
    assert len(frames) == 10 and count > 10
    small.clear()
    assert small.frames()[1] == 0
