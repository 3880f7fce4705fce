"""The device engine compiled for the host, bit for bit against the oracle.

tests/host_engine.cpp builds the transport engine headers of the library
with g++ (no GPU, glibc libm) and runs each history with the same Philox
stream as the oracle. With the libm of the oracle, every operation of the
engine must reproduce the reference exactly: final states, events, media,
step and draw counts are compared for equality on the golden cases and on
terrain and voxel geometries. The comparison is with the unmodified reference:
the engine's math functions (pumas_amd/csrc/libm.h) return the C library's
bits, including the float logf / expf of dcs_evaluate
(src/pumas.c:11243-11347).
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import pumas_amd as pa
from pumas_amd import _abi
from pumas_amd import workloads as W

import locals_cases
import make_golden

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "libhost_engine.so")


def build_host(name="libhost_engine.so", defines=()):
    """The engine compiled for the host (g++, glibc; -mfma so that the
    explicit fma() calls of the math functions are single instructions)."""
    os.makedirs(BUILD, exist_ok=True)
    lib_path = os.path.join(BUILD, name)
    src = os.path.join(HERE, "host_engine.cpp")
    csrc = os.path.join(REPO, "pumas_amd", "csrc")
    deps = [src] + [os.path.join(csrc, f) for f in os.listdir(csrc)
                    if f.endswith(".h")]
    if (not os.path.exists(lib_path)) or any(
            os.path.getmtime(d) > os.path.getmtime(lib_path) for d in deps):
        # Built under a name of this process, then renamed: test workers
        # that rebuild together never load a partly written library.
        tmp = f"{lib_path}.{os.getpid()}.tmp"
        subprocess.check_call(
            ["g++", "-O2", "-std=c++17", "-fPIC", "-shared",
             "-ffp-contract=off", "-mfma"] + [f"-D{d}" for d in defines] +
            ["-I", csrc, "-I", os.path.join(REPO, "include"), "-o",
             tmp, src])
        os.replace(tmp, lib_path)
    lib = ctypes.CDLL(lib_path)
    lib.host_load.argtypes = [ctypes.c_char_p]
    lib.host_record.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_ulonglong, ctypes.c_int,
                                ctypes.c_uint64, ctypes.c_uint64]
    lib.host_record.restype = None
    lib.host_set_max_steps.argtypes = [ctypes.c_int]
    lib.host_set_max_steps.restype = None
    lib.host_set_loop_max.argtypes = [ctypes.c_int]
    lib.host_set_loop_max.restype = None
    lib.host_transport.argtypes = [
        ctypes.POINTER(_abi.GpuContext), ctypes.POINTER(_abi.GeometryDesc),
        ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64,
        ctypes.c_uint64] + [ctypes.c_void_p] * 5
    return lib


@pytest.fixture(scope="module")
def host():
    return build_host()


def run_host(lib, physics, ctx, desc, states, seed=0, max_loop=1 << 16):
    assert lib.host_load(pa.physics_path(physics).encode()) == 0
    lib.host_set_loop_max(max_loop)
    s = states.copy()
    n = len(s)
    ev = np.zeros(n, np.int32)
    md = np.zeros((n, 2), np.int32)
    st = np.zeros(n, np.uint32)
    dr = np.zeros(n, np.uint32)
    rc = np.zeros(n, np.int32)
    c = ctx.to_c()
    d = desc.to_c()
    lib.host_transport(ctypes.byref(c), ctypes.byref(d), s.ctypes.data, n,
                       seed, 0, ev.ctypes.data, md.ctypes.data,
                       st.ctypes.data, dr.ctypes.data, rc.ctypes.data)
    return s, ev, md, st, dr, rc


def check_exact(host_out, ref, skip_aborted=False):
    s, ev, md, st, dr, rc = host_out
    ok = (ref.rc != -1) if skip_aborted else np.ones(len(s), bool)
    assert ok.mean() > 0.5
    for f in ("energy", "distance", "grammage", "time", "weight", "position",
              "direction", "decayed", "charge"):
        np.testing.assert_array_equal(s[f][ok], ref.states[f][ok], err_msg=f)
    np.testing.assert_array_equal(ev[ok], ref.events[ok])
    np.testing.assert_array_equal(md[ok], ref.media[ok])
    np.testing.assert_array_equal(st[ok], ref.steps[ok])
    np.testing.assert_array_equal(dr[ok], ref.draws[ok])


@pytest.mark.parametrize("case", [c[0] for c in make_golden.cases()])
def test_golden_cases_bit_exact(host, ref_standard, case):
    import oracle as orc
    name, pname, ctx, desc, states = next(
        c for c in make_golden.cases() if c[0] == case)
    states = states[:1024]
    out = run_host(host, pname, ctx, desc, states,
                   max_loop=make_golden.max_loop(name))
    ref = orc.transport(ref_standard, ctx,
                        orc.RefGeometry(desc), states, seed=0,
                        threads=8)
    check_exact(out, ref)


@pytest.mark.parametrize("case", [c[0] for c in locals_cases.cases()])
def test_non_uniform_locals_bit_exact(host, ref_standard, case):
    """Non uniform media (exponential density profiles with a locals
    step): the engine against the reference with the equivalent
    `pumas_locals_cb`, every history bit identical."""
    import oracle as orc
    name, pname, ctx, desc, states = next(
        c for c in locals_cases.cases() if c[0] == case)
    out = run_host(host, pname, ctx, desc, states)
    ref = orc.transport(ref_standard, ctx, orc.RefGeometry(desc), states,
                        seed=0, threads=8)
    print(f"{case}: {int(ref.steps.sum())} geometry steps, events "
          f"{np.unique(ref.events)}")
    check_exact(out, ref)


def test_non_uniform_locals_change_the_transport(ref_standard):
    """The locals cases exercise the non uniform path: the reference's
    results change when the exponential media are made uniform."""
    import oracle as orc
    for name, pname, ctx, desc, states in locals_cases.cases(128):
        m = desc.media
        uniform = pa.GeometryDescription(
            desc.kind, [(int(a), float(b)) for a, b in
                        zip(m["material"], m["density"])],
            step_type=desc.step_type, z=desc.z)
        a = orc.transport(ref_standard, ctx, orc.RefGeometry(desc), states)
        b = orc.transport(ref_standard, ctx, orc.RefGeometry(uniform), states)
        assert (a.states["position"] != b.states["position"]).any(), name
        assert a.steps.sum() > b.steps.sum(), name


def test_terrain_bit_exact(host, ref_standard):
    import oracle as orc
    desc = W.terrain_geometry(64, spacing=20.0)
    states = W.terrain_states(300, 63 * 20.0)
    ctx = W.terrain_context()
    out = run_host(host, "standard_c005", ctx, desc, states)
    ref = orc.transport(ref_standard, ctx,
                        orc.RefGeometry(desc), states, seed=0,
                        threads=8)
    check_exact(out, ref, skip_aborted=True)


def test_voxels_bit_exact(host, ref_iron):
    """Voxels at a 1% cutoff: the reference aborts (assert(0) of its
    backward target selection, src/pumas.c:6264) on some histories; the
    engine marks exactly those with PUMAS_GPU_HISTORY_NO_TARGET, and every
    other history is bit identical."""
    import oracle as orc
    desc = W.voxel_geometry(32, size=8.0)
    states = W.voxel_states(400, 32 * 8.0)
    ctx = W.voxel_context()
    out = run_host(host, "iron_c001", ctx, desc, states)
    ref = orc.transport(ref_iron, ctx,
                        orc.RefGeometry(desc), states, seed=0,
                        threads=8)
    aborted = ref.rc == -1
    print(f"voxels: {int(aborted.sum())} of {len(states)} reference aborts")
    assert aborted.any()
    rc = out[5]
    np.testing.assert_array_equal(rc == int(pa.HistoryStatus.NO_TARGET),
                                  aborted)
    check_exact(out, ref, skip_aborted=True)


@pytest.mark.parametrize("period", [1, 7, 0])
@pytest.mark.parametrize("case", ["slabs_straggled", "slabs_backward",
                                  "slabs_mixed", "csda_rock"])
def test_recorder_frames_bit_exact(host, ref_standard, case, period):
    """The device recorder's frames (track start and stop, every period-th
    step, DEL pre and post points, medium changes) are the frames of the
    reference's recorder, in order, bit for bit."""
    import oracle as orc
    name, pname, ctx, desc, states = next(
        c for c in make_golden.cases() if c[0] == case)
    states = states[:16]
    cap = 1 << 16
    frames = np.zeros(cap, _abi.FRAME_DTYPE)
    count = ctypes.c_ulonglong(0)
    host.host_record(frames.ctypes.data, ctypes.addressof(count), cap, period,
                     0, len(states))
    try:
        out = run_host(host, pname, ctx, desc, states)
    finally:
        host.host_record(None, None, 0, 0, 0, 0)
    fin, ref = orc.transport_record(ref_standard, ctx,
                                    orc.RefGeometry(desc),
                                    states, period=period, capacity=cap)
    assert (out[0] == fin).all()
    k = count.value
    assert k == ref["count"] and k < cap and k >= 2 * len(states)
    f = frames[:k]
    assert (f["state"] == ref["states"]).all()
    np.testing.assert_array_equal(f["medium"], ref["media"])
    np.testing.assert_array_equal(f["event"], ref["events"])
    np.testing.assert_array_equal(f["history"], ref["history"])


@pytest.mark.parametrize("direction", ["forward", "backward"])
@pytest.mark.parametrize("energy", [np.inf, np.nan])
def test_non_finite_energy_ends_with_value_error(host, direction, energy):
    """A history whose kinetic energy is not a finite number would step
    until the step cap (the reference has no end for it): the engine ends it
    at its first step with PUMAS_RETURN_VALUE_ERROR, and the other histories
    of the batch are unaffected."""
    if direction == "forward":
        ctx, states = W.slabs_forward_context(), W.slabs_forward_states(4, 1)
    else:
        ctx = W.slabs_backward_context()
        states = W.slabs_backward_states(4, seed=1)
    bad = states.copy()
    bad["energy"][1] = energy
    s, ev, md, st, dr, rc = run_host(host, "standard_c005", ctx,
                                     W.slabs_geometry(), bad)
    assert rc[1] == int(pa.Return.VALUE_ERROR)
    assert st[1] <= 1
    ok = run_host(host, "standard_c005", ctx, W.slabs_geometry(), states)
    for i in (0, 2, 3):
        assert rc[i] == 0
        assert st[i] == ok[3][i]
        assert s["energy"][i] == ok[0]["energy"][i]


@pytest.mark.parametrize("physics", ["standard_c005", "iron_c001"])
def test_coarse_kinetic_index_matches_bisection(host, physics):
    """The O(1) bracket of the kinetic energy grid (exponent and leading
    mantissa bits, then a few comparisons) is the bisection's row for every
    grid value, its double precision neighbours, the mid points and 10^6
    log-uniform values."""
    host.host_check_k_index.restype = ctypes.c_long
    host.host_check_k_index.argtypes = [ctypes.c_long]
    assert host.host_load(pa.physics_path(physics).encode()) == 0
    assert host.host_check_k_index(1_000_000) == 0


def test_step_cap_status(host):
    """A history that reaches the step cap ends there with no event and
    rc PUMAS_GPU_HISTORY_STEP_LIMIT (not as a limit event); the others are
    unchanged."""
    ctx, desc = W.slabs_forward_context(), W.slabs_geometry()
    states = W.slabs_forward_states(64, seed=2)
    ok = run_host(host, "standard_c005", ctx, desc, states)
    host.host_set_max_steps(200)
    try:
        capped = run_host(host, "standard_c005", ctx, desc, states)
    finally:
        host.host_set_max_steps(10000000)
    s, ev, md, st, dr, rc = capped
    long = ok[3] > 200
    assert long.any() and (~long).any()
    assert (rc[long] == int(pa.HistoryStatus.STEP_LIMIT)).all()
    assert (ev[long] == int(pa.Event.NONE)).all()
    assert (st[long] == 200).all()
    assert (rc[~long] == 0).all()
    assert (s[~long] == ok[0][~long]).all()
    np.testing.assert_array_equal(ev[~long], ok[1][~long])


def test_loop_limit_status(host):
    """With the loop limit set to 2 iterations, histories whose rejection
    loops give up end with rc PUMAS_GPU_HISTORY_LOOP_LIMIT and no event; the
    histories whose loops never reached the limit are bit identical to a run
    with the default limit."""
    ctx, desc = W.slabs_forward_context(), W.slabs_geometry()
    states = W.slabs_forward_states(128, seed=3)
    ok = run_host(host, "standard_c005", ctx, desc, states)
    out = run_host(host, "standard_c005", ctx, desc, states, max_loop=2)
    flagged = out[5] == int(pa.HistoryStatus.LOOP_LIMIT)
    assert flagged.any() and (~flagged).any()
    assert (out[1][flagged] == int(pa.Event.NONE)).all()
    assert (out[5][~flagged] == 0).all()
    assert (out[0][~flagged] == ok[0][~flagged]).all()


def test_energy_above_tables_loop_limit(host):
    """Forward straggled transport at 1e11 GeV, a hundred times the last
    table row: the elastic hard scattering angle loop of the reference
    accepts so rarely there (the nuclear form factor cuts the large angles)
    that it needs more than 2^25 trials. Run with a loop limit of 2^25, the
    engine ends the history after that many with
    PUMAS_GPU_HISTORY_LOOP_LIMIT and no event."""
    import make_golden as M
    s = M.edge_states(2, [1e11], direction=-1.0, z=W.SLAB_START_Z)
    out = run_host(host, "standard_c005", W.slabs_forward_context(),
                   W.slabs_geometry(), s, max_loop=1 << 25)
    assert (out[5] == int(pa.HistoryStatus.LOOP_LIMIT)).all()
    assert (out[1] == int(pa.Event.NONE)).all()
