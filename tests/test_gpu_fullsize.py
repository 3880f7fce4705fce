"""The full-size geometries and request sizes of the workloads.

* Terrain muography over the 1024 x 1024 heightmap at 5 m: histories of a
  subset of the 4096 lines of sight, FP64, against the reference at 1e-12.
* The 512^3 voxel grid at 2 m (iron physics, 1% cutoff), FP64, against
  the reference at 1e-12 (the reference aborts on some backward histories,
  which the device marks PUMAS_GPU_HISTORY_NO_TARGET, set for set).
* FP32-bulk against the FP64 reference directly: the Gaisser-weighted
  flux of 2 10^5 terrain histories (bias within 0.1%, chi2 of the energy
  histogram), and FP32-bulk against FP64 on 10^6 voxel histories.
* Serving: requests of 10^3 to 10^6 muons, alone and in a mix.
"""
import numpy as np
import pytest

from parity import assert_parity

import pumas_amd as pa
from pumas_amd import workloads as W
from test_gpu_geometries import Subset

pytestmark = pytest.mark.gpu

EXTENT = 1023 * 5.0


@pytest.fixture(scope="module")
def terrain():
    return W.terrain_geometry()


@pytest.fixture(scope="module")
def voxels():
    return W.voxel_geometry_large()


def sightline_states(n_lines, per_line, seed=0, kmin=1e-2, kmax=1e2):
    """per_line events on each of the first n_lines lines of sight of the
    4096 of the terrain workload, K log-uniform in [kmin, kmax] GeV."""
    lines = W.terrain_lines(4096, EXTENT, seed=seed)
    line = np.repeat(np.arange(n_lines), per_line)
    n = len(line)
    u = W.uniform(seed, 7, np.arange(n, dtype=np.uint64), 0)
    rk = np.log(kmax / kmin)
    s = np.zeros(n, pa.STATE_DTYPE)
    s["charge"] = -1.0
    s["energy"] = kmin * np.exp(rk * u)
    s["weight"] = s["energy"] * rk
    s["direction"] = lines["direction"][line]
    s["position"] = lines["position"][line]
    return s


def compare(ref_physics, gpu_physics, ctx, desc, states, seed):
    import oracle as orc
    print(f"reference: {len(states)} histories", flush=True)
    ref = orc.transport(ref_physics, ctx, orc.RefGeometry(desc), states,
                        seed=seed, threads=16)
    print("device", flush=True)
    gpu = pa.transport(gpu_physics, ctx, pa.Geometry(gpu_physics, desc),
                       states.copy(), seed=seed)
    aborted = ref.rc == -1
    flagged = gpu.rc == int(pa.HistoryStatus.NO_TARGET)
    print(f"{int(aborted.sum())} reference aborts, {int(flagged.sum())} "
          f"flagged", flush=True)
    np.testing.assert_array_equal(flagged, aborted)
    keep = ~aborted
    assert_parity(Subset(gpu, keep), Subset(ref, keep), 1e-12)
    return gpu, ref


def test_terrain_full_size_vs_reference(ref_standard, gpu_standard, terrain):
    s = sightline_states(128, 16, seed=0)
    gpu, ref = compare(ref_standard, gpu_standard, W.terrain_context(),
                       terrain, s, seed=5)
    assert (gpu.events == pa.Event.MEDIUM).mean() > 0.99


def test_voxels_full_size_vs_reference(ref_iron, gpu_iron, voxels):
    s = W.voxel_states(1024, 512 * 2.0, seed=7)
    gpu, ref = compare(ref_iron, gpu_iron, W.voxel_context(), voxels, s,
                       seed=7)
    assert (ref.rc != -1).mean() > 0.5


def flux(states):
    c = -states["direction"][:, 2]
    return states["weight"] * W.gaisser_flux(states["energy"], c)


def chi2(a, b, wa, wb, bins):
    ha, _ = np.histogram(a, bins=bins, weights=wa)
    hb, _ = np.histogram(b, bins=bins, weights=wb)
    va, _ = np.histogram(a, bins=bins, weights=wa ** 2)
    vb, _ = np.histogram(b, bins=bins, weights=wb ** 2)
    ok = (va + vb) > 0
    return float(np.sum((ha[ok] - hb[ok]) ** 2 / (va[ok] + vb[ok]))), \
        int(ok.sum())


def test_fp32_bulk_flux_vs_reference_terrain(ref_standard, gpu_standard,
                                             terrain):
    """FP32-bulk on the device against the FP64 reference on the CPU, the
    same 2 10^5 histories and random streams: integrated Gaisser flux bias
    within 0.1% and the energy histogram within statistics."""
    import oracle as orc
    s = sightline_states(500, 400, seed=1)
    ctx = W.terrain_context()
    print("reference FP64", flush=True)
    ref = orc.transport(ref_standard, ctx, orc.RefGeometry(terrain), s,
                        seed=9, threads=16)
    print("device FP32-bulk", flush=True)
    r32 = pa.transport(gpu_standard, ctx, pa.Geometry(gpu_standard, terrain),
                       s.copy(), seed=9, precision=pa.Precision.FP32_BULK)
    ok = (ref.rc == 0) & (r32.rc == 0)
    f64, f32 = flux(ref.states[ok]), flux(r32.states[ok])
    bias = f32.mean() / f64.mean() - 1.0
    sigma = np.std(f32 - f64) / np.sqrt(ok.sum()) / f64.mean()
    c2, ndf = chi2(ref.states["energy"][ok], r32.states["energy"][ok],
                   ref.states["weight"][ok], r32.states["weight"][ok],
                   np.logspace(-2, 4, 31))
    print(f"FP32-bulk vs reference: flux bias {bias:.2e} +- {sigma:.1e}, "
          f"energy chi2/ndf {c2:.1f}/{ndf}", flush=True)
    assert abs(bias) < 1e-3
    assert c2 / ndf < 2.0


def test_fp32_bulk_voxels_flux(gpu_iron, voxels):
    """FP32-bulk against FP64 on 10^6 histories of the voxel workload."""
    s = W.voxel_states(1_000_000, 512 * 2.0, seed=11)
    ctx = W.voxel_context()
    geom = pa.Geometry(gpu_iron, voxels)
    r64 = pa.transport(gpu_iron, ctx, geom, s.copy(), seed=11)
    r32 = pa.transport(gpu_iron, ctx, geom, s.copy(), seed=11,
                       precision=pa.Precision.FP32_BULK)
    ok = (r64.rc == 0) & (r32.rc == 0)
    f64, f32 = flux(r64.states[ok]), flux(r32.states[ok])
    bias = f32.mean() / f64.mean() - 1.0
    sigma = np.std(f32 - f64) / np.sqrt(ok.sum()) / f64.mean()
    print(f"voxels FP32-bulk vs FP64: flux bias {bias:.2e} +- {sigma:.1e} "
          f"({int(ok.sum())} histories)", flush=True)
    assert abs(bias) < 1e-3


def test_serving_full_request_sizes(gpu_standard, terrain):
    """Requests of 10^3 to 10^6 muons (slabs or the full terrain, CSDA or
    mixed) served together equal the same requests alone, bit for bit."""
    geoms = {"slabs": pa.Geometry(gpu_standard, W.slabs_geometry()),
             "terrain": pa.Geometry(gpu_standard, terrain)}
    reqs = W.serving_requests(0.04, rate=200.0, seed=3, nmin=1e3, nmax=1e6)
    assert len(reqs) >= 4
    print(f"{len(reqs)} requests, {sum(r.n for r in reqs)} muons",
          flush=True)
    with pa.Server(slots=4) as srv:
        tickets = []
        for r in reqs:
            ctx, st = W.serving_inputs(r, EXTENT)
            tickets.append(srv.submit(gpu_standard, ctx, geoms[r.geometry],
                                      st, first_history=r.first))
        served = [srv.wait(t) for t in tickets]
    for r, out in zip(reqs, served):
        ctx, st = W.serving_inputs(r, EXTENT)
        alone = pa.transport(gpu_standard, ctx, geoms[r.geometry], st,
                             first_history=r.first)
        assert np.array_equal(out.result.states, alone.states)
        for k in ("events", "media", "steps", "draws", "rc"):
            np.testing.assert_array_equal(getattr(out.result, k),
                                          getattr(alone, k))
    print("max request", max(r.n for r in reqs), "p50 latency %.1f ms" %
          np.median([o.latency_ms for o in served]), flush=True)
