"""Continuous batching server: admission, concurrency, and identical
results for a request alone and in a mix of uneven requests."""
import numpy as np
import pytest

import pumas_amd as pa
from pumas_amd import workloads as W

pytestmark = pytest.mark.gpu

TERRAIN_N, TERRAIN_SPACING = 128, 20.0


@pytest.fixture(scope="module")
def geometries(gpu_standard):
    return {
        "slabs": pa.Geometry(gpu_standard, W.slabs_geometry()),
        "terrain": pa.Geometry(gpu_standard, W.terrain_geometry(
            TERRAIN_N, spacing=TERRAIN_SPACING)),
    }


def mix(seed=0, duration=0.1, nmax=2e4):
    return W.serving_requests(duration, rate=200.0, seed=seed, nmin=1e2,
                              nmax=nmax)


def inputs(req):
    return W.serving_inputs(req, (TERRAIN_N - 1) * TERRAIN_SPACING)


def same(a, b):
    assert np.array_equal(a.states, b.states)
    for k in ("events", "media", "steps", "draws", "rc"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k))


def test_mix_equals_alone(gpu_standard, geometries):
    reqs = mix()
    assert len(reqs) >= 8
    with pa.Server(slots=4) as srv:
        tickets = []
        for r in reqs:
            ctx, s = inputs(r)
            tickets.append(srv.submit(gpu_standard, ctx, geometries[r.geometry],
                                      s, first_history=r.first))
        served = [srv.wait(t) for t in tickets]
    for r, out in zip(reqs, served):
        ctx, s = inputs(r)
        alone = pa.transport(gpu_standard, ctx, geometries[r.geometry], s,
                             first_history=r.first)
        same(out.result, alone)
        assert out.rc == 0
        assert out.latency_ms >= out.queue_ms >= 0.0
    print("requests", len(reqs), "p50 latency %.1f ms" %
          np.median([o.latency_ms for o in served]))


def test_mix_equals_alone_full_sizes(gpu_standard, geometries):
    """Requests of 10^3 to 10^6 muons under a 200 GB device budget, as in
    the serving workload: each result equals the same request run alone."""
    reqs = W.serving_requests(0.06, rate=200.0, seed=5, nmin=1e3, nmax=1e6)
    assert len(reqs) >= 6
    assert max(r.n for r in reqs) > 1e5
    with pa.Server(hbm_budget=200 * 10**9, slots=4) as srv:
        tickets = []
        for r in reqs:
            ctx, s = inputs(r)
            tickets.append(srv.submit(gpu_standard, ctx, geometries[r.geometry],
                                      s, first_history=r.first))
        served = [srv.wait(t) for t in tickets]
    for r, out in zip(reqs, served):
        ctx, s = inputs(r)
        alone = pa.transport(gpu_standard, ctx, geometries[r.geometry], s,
                             first_history=r.first)
        same(out.result, alone)
        assert out.rc == 0
    print("requests", len(reqs), "muons", sum(r.n for r in reqs),
          "p50 latency %.1f ms" % np.median([o.latency_ms for o in served]))


def test_budget_admission(gpu_standard, geometries):
    """A budget that holds one request at a time still completes every
    request, in order of submission, with identical results."""
    reqs = mix(seed=1, duration=0.04, nmax=5e3)
    budget = pa.library().pumas_gpu_server_request_bytes(
        max(r.n for r in reqs), 0) + 1
    with pa.Server(hbm_budget=budget, slots=4) as srv:
        tickets = []
        for r in reqs:
            ctx, s = inputs(r)
            tickets.append(srv.submit(gpu_standard, ctx, geometries[r.geometry],
                                      s, first_history=r.first))
        served = [srv.wait(t) for t in tickets]
    for r, out in zip(reqs, served):
        ctx, s = inputs(r)
        same(out.result, pa.transport(gpu_standard, ctx,
                                      geometries[r.geometry], s,
                                      first_history=r.first))


def test_request_errors(gpu_standard, geometries):
    with pa.Server(hbm_budget=64 << 20, slots=2) as srv:
        ctx, s = inputs(W.ServingRequest(0.0, "slabs", pa.Mode.MIXED,
                                         10**6, 0))
        with pytest.raises(pa.PumasError):
            srv.submit(gpu_standard, ctx, geometries["slabs"], s)
        bad = W.slabs_forward_states(64)
        bad["direction"][3] = (0.0, 0.0, 0.5)
        t_bad = srv.submit(gpu_standard, W.slabs_forward_context(),
                           geometries["slabs"], bad)
        good = W.slabs_forward_states(64, first=64)
        t_good = srv.submit(gpu_standard, W.slabs_forward_context(),
                            geometries["slabs"], good, first_history=64)
        out = srv.wait(t_bad, raise_errors=False)
        assert out.rc == pa.Return.DIRECTION_ERROR
        assert out.result.rc[3] == pa.Return.DIRECTION_ERROR
        assert (np.delete(out.result.rc, 3) == 0).all()
        assert srv.wait(t_good).rc == 0
