"""GPU transport vs the reference, history by history.

The device engine uses the reference's arithmetic step by step, and its
math functions (pumas_amd/csrc/libm.h) return the bits of the C library
the reference is built against, including the float logf / expf of
dcs_evaluate (src/pumas.c:11298-11346). The bar is therefore the FP64
per-event tolerance on every history: events, media, step and random
draw counts identical (no branch flips), and final energy, distance,
grammage, time, weight, position and direction within 1e-12 relative for
every history (assert_parity prints the measured fraction within 1e-12 and
the largest error; the runs so far measure 0, bit identical results).
"""
import numpy as np
import pytest

from conftest import load_golden
from parity import assert_parity

import make_golden
import pumas_amd as pa
from pumas_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = {c[0]: c for c in make_golden.cases()}
RTOL = 1e-12


@pytest.mark.parametrize("name", sorted(CASES))
def test_golden_fixture(gpu_standard, name):
    print(f"golden fixture {name}", flush=True)
    _, _, ctx, desc, _ = CASES[name]
    g = load_golden(name)
    geom = pa.Geometry(gpu_standard, desc)
    r = pa.transport(gpu_standard, ctx, geom, g["initial"].copy(),
                     seed=int(g["seed"]), max_loop=make_golden.max_loop(name))
    assert_parity(r, g, RTOL)


def _live(ref_physics, gpu_physics, ctx, desc, states, seed=0, threads=8):
    import oracle as orc
    ref = orc.transport(ref_physics, ctx, orc.RefGeometry(desc), states,
                        seed=seed, threads=threads)
    gpu = pa.transport(gpu_physics, ctx, pa.Geometry(gpu_physics, desc),
                       states.copy(), seed=seed)
    return gpu, ref


def test_csda_full_workload_vs_reference(ref_standard, gpu_standard):
    """The whole backward CSDA workload: 10^5 histories, seed 0."""
    s = W.csda_rock_states(100000, seed=0)
    gpu, ref = _live(ref_standard, gpu_standard, W.csda_rock_context(),
                     W.csda_rock_geometry(), s)
    assert_parity(gpu, ref, RTOL)
    flux_gpu = W.integrated_flux(gpu.states)
    flux_ref = W.integrated_flux(ref.states)
    assert abs(flux_gpu / flux_ref - 1) < 1e-12


@pytest.mark.parametrize("seed", [1, 2])
def test_slabs_straggled_vs_reference(ref_standard, gpu_standard, seed):
    s = W.slabs_forward_states(8192, seed=seed)
    gpu, ref = _live(ref_standard, gpu_standard, W.slabs_forward_context(),
                     W.slabs_geometry(), s, seed=seed)
    assert_parity(gpu, ref, RTOL)


def test_slabs_mixed_backward_vs_reference(ref_standard, gpu_standard):
    ctx = W.slabs_backward_context(energy_loss=pa.Mode.MIXED)
    s = W.slabs_backward_states(4096, seed=3)
    gpu, ref = _live(ref_standard, gpu_standard, ctx, W.slabs_geometry(), s,
                     seed=3)
    assert_parity(gpu, ref, RTOL)


def test_slabs_step_check_vs_reference(ref_standard, gpu_standard):
    """Boundary bisection of checked geometry steps (src/pumas.c:6424-6485)."""
    s = W.slabs_forward_states(4096, seed=4)
    gpu, ref = _live(ref_standard, gpu_standard, W.slabs_forward_context(),
                     W.slabs_geometry(step_type=pa.Step.CHECK), s, seed=4)
    assert_parity(gpu, ref, RTOL)


def test_no_scattering_and_disabled_loss(ref_standard, gpu_standard):
    """Straight steps and energy loss disabled with a distance limit."""
    s = W.slabs_forward_states(1024, seed=5)
    for ctx in (W.slabs_forward_context(energy_loss=pa.Mode.MIXED).replace(
                    scattering=pa.Mode.DISABLED),
                pa.Context(energy_loss=pa.Mode.DISABLED,
                           scattering=pa.Mode.MIXED,
                           event=pa.Event.LIMIT_DISTANCE,
                           limit_distance=300.0)):
        gpu, ref = _live(ref_standard, gpu_standard, ctx, W.slabs_geometry(),
                         s, seed=5)
        assert_parity(gpu, ref, RTOL)


def test_edge_cases_vs_reference(ref_standard, gpu_standard):
    """Decayed, outside, non unit direction, null weight, limits already
    reached, and a history with no energy."""
    s = W.slabs_forward_states(8, seed=6)
    s["decayed"][0] = 1
    s["position"][1, 2] = 900.0          # starts outside the stack
    s["direction"][2] *= 1.01            # not a unit vector
    s["weight"][3] = 0.0
    s["distance"][4] = 2e3               # distance limit already reached
    s["energy"][5] = 0.0
    ctx = W.slabs_forward_context().replace(
        event=pa.Event.LIMIT_DISTANCE, limit_distance=1e3)
    import oracle as orc
    ref = orc.transport(ref_standard, ctx, orc.RefGeometry(W.slabs_geometry()),
                        s, seed=6)
    geom = pa.Geometry(gpu_standard, W.slabs_geometry())
    with pytest.raises(pa.PumasError) as e:
        pa.transport(gpu_standard, ctx, geom, s.copy(), seed=6)
    assert e.value.code == pa.Return.DIRECTION_ERROR
    gpu = pa.transport(gpu_standard, ctx, geom, s.copy(), seed=6,
                       raise_errors=False)
    assert gpu.error == pa.Return.DIRECTION_ERROR
    assert ref.rc[2] == pa.Return.DIRECTION_ERROR
    assert_parity(gpu, ref, RTOL)
    assert gpu.events[0] == pa.Event.VERTEX_DECAY
    assert gpu.events[1] == pa.Event.MEDIUM and gpu.media[1, 0] == -1


def test_batch_errors(gpu_standard):
    geom = pa.Geometry(gpu_standard, W.slabs_geometry())
    s = W.slabs_forward_states(4)
    with pytest.raises(pa.PumasError) as e:
        pa.transport(gpu_standard, pa.Context(accuracy=0.0), geom, s.copy())
    assert e.value.code == pa.Return.ACCURACY_ERROR
    r = pa.transport(gpu_standard, pa.Context(), geom, s[:0].copy())
    assert len(r.states) == 0


def test_backward_cutoff_check(gpu_standard):
    """Backward straggled transport needs a cutoff >= 1%
    (src/pumas.c:3170-3177): a table image patched to 0.5% is refused."""
    import ctypes
    import struct
    data = bytearray(open(pa.physics_path("standard_c005"), "rb").read())
    off = 8 + 4 + 4 + 16 * 4 + 2 * 8  # fmeta[PTAB_CUTOFF]
    struct.pack_into("<d", data, off, 0.005)
    lib = pa.library()
    h = ctypes.c_void_p()
    buf = (ctypes.c_char * len(data)).from_buffer(data)
    assert lib.pumas_gpu_physics_upload_memory(buf, len(data),
                                                ctypes.byref(h)) == 0
    ph = pa.Physics.__new__(pa.Physics)
    ph.path, ph._handle = "patched", h
    geom = pa.Geometry(ph, W.slabs_geometry())
    with pytest.raises(pa.PumasError) as e:
        pa.transport(ph, W.slabs_backward_context(), geom,
                     W.slabs_backward_states(4))
    assert e.value.code == pa.Return.ACCURACY_ERROR
    ph.destroy()


def test_batch_split_invariance(gpu_standard):
    """History i depends only on (seed, first_history + i): transporting a
    batch in two halves gives the same results."""
    s = W.slabs_forward_states(2048, seed=9)
    geom = pa.Geometry(gpu_standard, W.slabs_geometry())
    ctx = W.slabs_forward_context()
    whole = pa.transport(gpu_standard, ctx, geom, s.copy(), seed=9)
    a = pa.transport(gpu_standard, ctx, geom, s[:1000].copy(), seed=9)
    b = pa.transport(gpu_standard, ctx, geom, s[1000:].copy(), seed=9,
                     first_history=1000)
    joined = np.concatenate([a.states, b.states])
    assert (joined == whole.states).all()


def test_closed_form_batch_above_stepping_chunk(gpu_standard):
    """A closed form CSDA batch larger than the 2^25 histories of a stepping
    chunk runs as one launch; it equals the same batch transported in two
    calls split at 2^25."""
    n0 = 1 << 25
    s = W.csda_rock_states(n0 + 4096, seed=13)
    geom = pa.Geometry(gpu_standard, W.csda_rock_geometry())
    ctx = W.csda_rock_context()
    whole = pa.transport(gpu_standard, ctx, geom, s.copy(), seed=13)
    a = pa.transport(gpu_standard, ctx, geom, s[:n0].copy(), seed=13)
    b = pa.transport(gpu_standard, ctx, geom, s[n0:].copy(), seed=13,
                     first_history=n0)
    assert whole.total_steps == 0
    assert (np.concatenate([a.states, b.states]) == whole.states).all()
    for k in ("events", "media", "steps", "draws", "rc"):
        np.testing.assert_array_equal(
            np.concatenate([getattr(a, k), getattr(b, k)]),
            getattr(whole, k))


@pytest.mark.parametrize("energy_loss", [pa.Mode.STRAGGLED, pa.Mode.MIXED])
def test_graph_replays_equal_eager_launches(gpu_standard, energy_loss):
    """The rounds between two host checks are replayed as a captured HIP
    graph. With two rounds per check a batch takes many replays, each
    separated by the host's counter copy and synchronisation; every call
    must equal the same transport launched round by round, and repeated
    calls (fresh graphs) must equal each other."""
    s = W.slabs_backward_states(3000, seed=12)
    geom = pa.Geometry(gpu_standard, W.slabs_geometry())
    ctx = W.slabs_backward_context(energy_loss=energy_loss)
    eager = pa.transport(gpu_standard, ctx, geom, s.copy(), seed=12,
                         rounds_per_check=2, eager=True)
    for _ in range(3):
        g = pa.transport(gpu_standard, ctx, geom, s.copy(), seed=12,
                         rounds_per_check=2)
        assert (g.states == eager.states).all()
        for k in ("events", "media", "steps", "draws", "rc"):
            np.testing.assert_array_equal(getattr(g, k), getattr(eager, k))
    g = pa.transport(gpu_standard, ctx, geom, s.copy(), seed=12)
    assert (g.states == eager.states).all()


@pytest.mark.parametrize("energy_loss", [pa.Mode.STRAGGLED, pa.Mode.MIXED])
def test_backward_weights_vs_reference(ref_standard, gpu_standard,
                                       energy_loss):
    """Backward transport with discrete energy losses: the weights carry
    the float DCS evaluations (logf / expf of the C library on both sides)
    and are held to the FP64 tolerance."""
    ctx = W.slabs_backward_context(energy_loss=energy_loss)
    s = W.slabs_backward_states(4096, seed=11)
    gpu, ref = _live(ref_standard, gpu_standard, ctx, W.slabs_geometry(), s,
                     seed=11)
    assert_parity(gpu, ref, RTOL)


@pytest.mark.parametrize("block_size,blocks_per_cu,ehs_batch,tail_del",
                         [(64, 8, 8, 0), (128, 4, 32, 1), (256, 1, 1, -1),
                          (256, 2, 16, -1)])
def test_launch_settings_invariance(gpu_standard, block_size, blocks_per_cu,
                                    ehs_batch, tail_del):
    """Launch shape, EHS batching and in-wavefront tail DELs change the
    interleaving of histories, never their results."""
    s = W.slabs_forward_states(4000, seed=13)
    geom = pa.Geometry(gpu_standard, W.slabs_geometry())
    ctx = W.slabs_forward_context()
    ref = pa.transport(gpu_standard, ctx, geom, s.copy(), seed=13)
    alt = pa.transport(gpu_standard, ctx, geom, s.copy(), seed=13,
                       block_size=block_size, blocks_per_cu=blocks_per_cu,
                       ehs_batch=ehs_batch, tail_del=tail_del)
    assert (alt.states == ref.states).all()
    for k in ("events", "media", "steps", "draws", "rc"):
        np.testing.assert_array_equal(getattr(alt, k), getattr(ref, k))


def test_autotune(gpu_standard):
    import torch
    s = W.slabs_forward_states(50_000, seed=14)
    geom = pa.Geometry(gpu_standard, W.slabs_geometry())
    ctx = W.slabs_forward_context()
    d = torch.from_numpy(s.view(np.uint8).reshape(-1)).cuda()
    before = d.clone()
    best, rate = pa.autotune(gpu_standard, ctx, geom, d.data_ptr(), len(s))
    torch.cuda.synchronize()
    assert torch.equal(d, before)
    assert best["block_size"] in (64, 128, 256, 512)
    assert best["ehs_batch"] in (8, 16, 24, 32)
    assert rate > 0.0
    tuned = pa.transport(gpu_standard, ctx, geom, s.copy(), **best)
    ref = pa.transport(gpu_standard, ctx, geom, s.copy())
    assert (tuned.states == ref.states).all()


@pytest.mark.parametrize("lanes", [1, 7, 64])
def test_lanes_per_wave_invariance(gpu_standard, lanes):
    """Spreading a batch over more wavefronts with fewer histories each
    changes where histories run, never their results."""
    s = W.slabs_forward_states(3000, seed=17)
    geom = pa.Geometry(gpu_standard, W.slabs_geometry())
    ctx = W.slabs_forward_context()
    ref = pa.transport(gpu_standard, ctx, geom, s.copy(), seed=17,
                       lanes_per_wave=64)
    alt = pa.transport(gpu_standard, ctx, geom, s.copy(), seed=17,
                       lanes_per_wave=lanes)
    assert (alt.states == ref.states).all()
    for k in ("events", "media", "steps", "draws", "rc"):
        np.testing.assert_array_equal(getattr(alt, k), getattr(ref, k))
