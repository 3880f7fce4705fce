#!/usr/bin/env python3
"""Benchmark: batched muon transport on one MI355X.

A step of this benchmark is one pass of `pumas_gpu_transport` over the whole
batch of the forward slab workload: 10^7 muons with kinetic energy
log-uniform in [10, 10^4] GeV and cos(zenith) uniform in [0.5, 1], through
100 m of air, 500 m of standard rock and 200 m of water, straggled energy
loss (class II discrete losses above a 5% cutoff) and mixed multiple
scattering, FP64. Each pass uses fresh random streams (history ids offset by
the pass index) on the same initial states, which are resident in HBM
before the timed region.

The reported value is muon-steps per second (one step = one Monte Carlo
step of pumas_context_transport, i.e. one geometry step proposal); events
per second are reported beside it. The timed passes launch the scheduler's
rounds one by one, with HIP events around each stepping kernel launch (the
roofline's kernel time); after the timed region the same batch is run a
few passes more as HIP graph replays (the library's default launch path,
`graph_replay`) and in FP32-bulk (`fp32_bulk`). The CPU baseline is the reference library
(compiled from its source by oracle/Makefile) with one context per thread on
the host cores, over a bounded sample of the same workload.

Other workloads (`--workload`, not the default line): `csda` — backward
CSDA through 1 km of standard rock, 10^8 histories (closed form, one
muon-step per history); `terrain` — backward straggled muography along
4096 lines of sight (`--lines`, 10^3..10^6 events each) under a 1024 x 1024
heightmap;
`voxels` — backward straggled transport through 512^3 voxels of 2 m at a 1%
cutoff, 10^8 histories (10^7 generated states tiled on the device, each
history with its own id and random stream). `--precision fp32` selects
FP32-bulk.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N]
                       [--workload slabs|csda|terrain|voxels]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

# Algorithmic HBM bytes per Monte Carlo step of the FP64 stepping record:
# struct pumas_state (104 B) plus the per-history stepping carry (96 B),
# read and written once per step.
BYTES_PER_STEP_FP64 = 2 * (104 + 96)
# Closed form CSDA path: the state, read and written once per history.
BYTES_PER_HISTORY_CSDA = 2 * 104
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="slabs",
                    choices=("slabs", "csda", "terrain", "voxels"))
    ap.add_argument("--n", type=int, default=0,
                    help="histories per pass (0: the workload's own, 10^7 "
                         "for slabs, 10^8 for voxels and csda)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target duration of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", choices=("fp64", "fp32"), default="fp64",
                    help="fp64: reference arithmetic; fp32: FP32-bulk")
    ap.add_argument("--graph", action="store_true",
                    help="replay the scheduler rounds as HIP graphs (no "
                         "per-launch timing of the stepping kernel then)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (default launch settings)")
    ap.add_argument("--tune-n", type=int, default=4_000_000,
                    help="histories of the autotuner's sample")
    ap.add_argument("--lines", type=int, default=4096,
                    help="lines of sight of the terrain workload (10^3 to "
                         "10^6 events each)")
    ap.add_argument("--max-steps", type=int, default=1_000_000,
                    help="Monte Carlo steps after which a history ends as "
                         "a limit (the longest histories of these "
                         "workloads take a few thousand)")
    ap.add_argument("--tail-del", type=int, default=0,
                    help="DELs inside the stepping wavefronts once the step "
                         "queue is empty: 1 always, -1 never, 0 in the rounds "
                         "that fit the resident lanes")
    ap.add_argument("--lds-tables", type=int, default=0,
                    help="step tables of the stepping kernel in LDS: 0 "
                         "when they fit, -1 never")
    ap.add_argument("--block-size", type=int, default=0)
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--ehs-batch", type=int, default=0,
                    help="lanes waiting for an elastic hard scattering "
                         "before the wavefront runs them (0: default or "
                         "tuned)")
    ap.add_argument("--serve-seconds", type=float, default=0.25,
                    help="arrival window of the serving run (Poisson "
                         "requests at 200/s); 0 skips it")
    ap.add_argument("--serve-slots", type=int, default=0)
    ap.add_argument("--extra-passes", type=int, default=3,
                    help="passes of each side measurement after the timed "
                         "region of the slab workload: the same batch "
                         "replayed as HIP graphs (the library's default "
                         "launch path) and in FP32-bulk (0 skips them)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run a plain device-to-device copy of this "
                         "many MiB (calibrates HBM byte counters)")
    return ap.parse_args()


def copy_calibration(mib):
    """A plain copy kernel of known size, for scaling HBM byte counters."""
    import torch
    n = mib << 20
    src = torch.empty(n, dtype=torch.uint8, device="cuda:0").fill_(1)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n,
                      "launches": 3}))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_threads():
    """Threads of the CPU baseline: every core of the affinity mask, capped
    by OMP_NUM_THREADS when it is set (a host shared by several GPUs gives
    each GPU a share of its cores and states it there)."""
    affinity = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS")
    threads = affinity
    note = f"all {affinity} cores of the affinity mask"
    if share and share.isdigit() and 0 < int(share) < affinity:
        threads = int(share)
        note = (f"this GPU's share of the host ({threads} cores, "
                f"OMP_NUM_THREADS) of the {affinity} in the affinity mask")
    return threads, note


def cpu_baseline(states, seconds, physics, desc, ctx):
    """Reference CPU path, one context per thread, on a bounded sample."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as orc
    if not orc.available():
        return None
    threads, cores_note = cpu_threads()
    phys = orc.RefPhysics(physics)
    geom = orc.RefGeometry(desc)
    # Calibrate on a small sample, then size the sample to `seconds`.
    n = 256 * threads
    r = orc.transport(phys, ctx, geom, states[:n], seed=0, threads=threads)
    rate = n / max(r.seconds, 1e-3)
    n = int(min(len(states), max(n, rate * seconds)))
    r = orc.transport(phys, ctx, geom, states[:n], seed=0, threads=threads)
    phys.close()
    return {
        "value": r.total_steps / r.seconds,
        "unit": "muon-steps/s",
        "events_per_s": n / r.seconds,
        "cores": threads,
        "cores_note": cores_note,
        "cpu_model": cpu_model(),
        "kind": "reference",
        "sample": f"{n} histories of the same workload ({r.total_steps} "
                  f"steps, {r.seconds:.1f} s)",
    }


def serving(pa, phys, seconds, slots, rate=200.0, seed=0):
    """Serving workload: Poisson arrivals (200/s) over `seconds` of
    requests of 10^U[3,6] muons, forward slabs or backward terrain
    (1024 x 1024 heightmap at 5 m), CSDA or MIXED energy loss, submitted to
    the continuous batching server as they arrive. Latency = submission to
    completion of a request, host buffers included."""
    import threading
    import numpy as np
    from pumas_amd import workloads as W
    reqs = W.serving_requests(seconds, rate=rate, seed=seed)
    if not reqs:
        return None
    log(f"serving: {len(reqs)} requests, {sum(r.n for r in reqs)} muons")
    geoms = {"slabs": pa.Geometry(phys, W.slabs_geometry()),
             "terrain": pa.Geometry(phys, W.terrain_geometry())}
    extent = 1023 * 5.0
    inputs = [W.serving_inputs(r, extent) for r in reqs]
    srv = pa.Server(slots=slots)
    tickets = [None] * len(reqs)
    ready = threading.Semaphore(0)
    t0 = time.perf_counter()

    def submitter():
        for k, (r, (ctx, st)) in enumerate(zip(reqs, inputs)):
            dt = r.arrival - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            tickets[k] = srv.submit(phys, ctx, geoms[r.geometry], st,
                                    first_history=r.first, outputs=False)
            ready.release()

    th = threading.Thread(target=submitter)
    th.start()
    out = []
    for k in range(len(reqs)):
        ready.acquire()
        out.append(srv.wait(tickets[k]))
    th.join()
    wall = time.perf_counter() - t0
    srv.close()
    lat = np.array([o.latency_ms for o in out])
    queue = np.array([o.queue_ms for o in out])
    muons = sum(r.n for r in reqs)
    steps = sum(o.steps for o in out)
    return {
        "requests": len(reqs), "muons": muons, "window_s": seconds,
        "rate_per_s": rate, "slots": slots or 4,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "mean_queue_ms": float(queue.mean()),
        "wall_s": wall, "events_per_s": muons / wall,
        "steps_per_s": steps / wall,
    }


WORKLOADS = {
    "slabs": "forward straggled + mixed scattering, 3 slabs (100 m air / "
             "500 m standard rock / 200 m water), K0 log-uniform [10, 1e4] "
             "GeV, cos(zenith) in [0.5, 1]",
    "csda": "backward CSDA, no scattering, 1 km of standard rock, K_f "
            "log-uniform [1e-2, 1e2] GeV",
    "terrain": "backward straggled + mixed scattering under a 1024 x 1024 "
               "heightmap (5 m, relief 0-800 m, column density 2.0-2.9), "
               "{lines} lines of sight x 10^U[3,6] events, K_f log-uniform "
               "[1e-2, 1e2] GeV",
    "voxels": "backward straggled + mixed scattering through 512^3 voxels "
              "of 2 m (rock / water / air / iron, lognormal density), "
              "cutoff 0.01, K_f log-uniform [0.1, 1e3] GeV",
}


TERRAIN_EXTENT = 1023 * 5.0


def terrain_on_device(W, torch, np, lines, n, dev, pool=10_000_000):
    """States of the terrain workload built on the device, line by line as
    terrain_sightlines: the detector position and direction of each event's
    line, and a final kinetic energy log-uniform in [1e-2, 1e2] GeV from a
    pool of `pool` Philox draws (event j takes draw j mod pool; the first
    `pool` events are those of terrain_sightlines). Returns the states as
    bytes on `dev` and a host sample of 2 10^5 events spread over the
    batch."""
    import math
    from pumas_amd import STATE_DTYPE
    rk = math.log(1e2 / 1e-2)
    m = min(n, pool)
    energy = 1e-2 * np.exp(rk * W.uniform(0, 0, np.arange(m, dtype=np.uint64),
                                          0))
    counts = lines["count"]
    start = np.concatenate([[0], np.cumsum(counts)])

    def fill(out, e, pos, dirs):
        out[:, 0] = -1.0
        out[:, 1] = e
        out[:, 5] = e * rk
        out[:, 6:9] = pos
        out[:, 9:12] = dirs

    e_dev = torch.from_numpy(energy).to(dev)
    pos = torch.from_numpy(lines["position"]).to(dev)
    dirs = torch.from_numpy(lines["direction"]).to(dev)
    line_of = torch.repeat_interleave(
        torch.arange(len(counts), device=dev),
        torch.from_numpy(counts).to(dev))
    states = torch.zeros((n, 13), dtype=torch.float64, device=dev)
    step = 1 << 25
    for a in range(0, n, step):
        b = min(n, a + step)
        j = torch.arange(a, b, device=dev)
        li = line_of[a:b]
        fill(states[a:b], e_dev[j % m], pos[li], dirs[li])
    del line_of
    idx = np.linspace(0, n - 1, 200_000).astype(np.int64)
    li = np.searchsorted(start, idx, side="right") - 1
    sample = np.zeros((len(idx), 13))
    fill(sample, energy[idx % m], lines["position"][li],
         lines["direction"][li])
    return states.view(torch.uint8).reshape(-1), \
        sample.view(STATE_DTYPE).reshape(-1)


def setup(args, pa, W):
    """(physics name, geometry description, context, host states, n)."""
    wl = args.workload
    if wl == "slabs":
        n = args.n or 10_000_000
        log(f"bench: generating {n} initial states")
        return ("standard_c005", W.slabs_geometry(),
                W.slabs_forward_context(), W.slabs_forward_states(n), n)
    if wl == "csda":
        n = args.n or 100_000_000
        return ("standard_c005", W.csda_rock_geometry(),
                W.csda_rock_context(), None, n)
    if wl == "terrain":
        lines = W.terrain_lines(args.lines, TERRAIN_EXTENT)
        n = int(lines["count"].sum())
        log(f"bench: {args.lines} lines of sight, {n} events")
        if n <= 10_000_000:
            states, _ = W.terrain_sightlines(args.lines, TERRAIN_EXTENT)
            return ("standard_c005", W.terrain_geometry(),
                    W.terrain_context(), states, n)
        return ("standard_c005", W.terrain_geometry(), W.terrain_context(),
                lines, n)
    n = args.n or 100_000_000
    m = min(n, 10_000_000)
    log("bench: generating the 512^3 voxel grid")
    t0 = time.perf_counter()
    desc = W.voxel_geometry_large()
    log(f"bench: voxel grid ready ({time.perf_counter() - t0:.0f} s), "
        f"generating {m} states")
    states = W.voxel_states(m, 512 * 2.0)
    log(f"bench: states ready ({time.perf_counter() - t0:.0f} s)")
    return ("iron_c001", desc, W.voxel_context(), states, n)


def main():
    args = parse()
    if args.copy_mib > 0:
        copy_calibration(args.copy_mib)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "runs on one GPU only (--gpus 1); "
                          "batches would shard across GPUs by history id "
                          "range, which is not launched here"}))
        sys.exit(2)

    import numpy as np
    import torch
    import pumas_amd as pa
    from pumas_amd import workloads as W

    pname, desc, ctx, states, n = setup(args, pa, W)
    phys = pa.Physics(pa.physics_path(pname))
    geom = pa.Geometry(phys, desc)
    dev = torch.device("cuda:0")
    if states is None:
        # CSDA: 10^7 generated states tiled on the device (the closed form
        # path draws no random numbers; the copies are independent work).
        base = W.csda_rock_states(min(n, 10_000_000))
        t = torch.from_numpy(base.view(np.uint8).reshape(-1, 104)).to(dev)
        reps = (n + len(base) - 1) // len(base)
        pristine = t.repeat(reps, 1)[:n].reshape(-1).contiguous()
        sample = base
    elif isinstance(states, dict):
        log("bench: building the terrain states on the device")
        pristine, sample = terrain_on_device(W, torch, np, states, n, dev)
    elif len(states) < n:
        # More histories than generated states: the states are tiled on the
        # device. Every history has its own id, hence its own random stream,
        # so the copies are independent histories of the same workload.
        t = torch.from_numpy(states.view(np.uint8).reshape(-1, 104)).to(dev)
        reps = (n + len(states) - 1) // len(states)
        pristine = t.repeat(reps, 1)[:n].reshape(-1).contiguous()
        sample = states
    else:
        pristine = torch.from_numpy(states.view(np.uint8).reshape(-1)).to(dev)
        sample = states
    work = torch.empty_like(pristine)
    stream = torch.cuda.current_stream().cuda_stream
    csda = args.workload == "csda"

    precision = pa.Precision.FP64 if args.precision == "fp64" else \
        pa.Precision.FP32_BULK

    tuned = {"block_size": args.block_size,
             "blocks_per_cu": args.blocks_per_cu,
             "ehs_batch": args.ehs_batch}
    if (not args.no_tune and not csda and args.block_size == 0 and
            args.blocks_per_cu == 0 and args.ehs_batch == 0):
        m = min(n, args.tune_n)
        tuned, rate = pa.autotune(phys, ctx, geom, pristine.data_ptr(), m,
                                  seed=7, precision=precision, stream=stream)
        log(f"bench: autotuned on {m} histories: {tuned} "
            f"({rate:.3e} steps/s on the sample)")

    # The plain method's tally: Gaisser-form flux at cos(theta) = 1 per cm^2,
    # without charge fraction, summed on the device after each pass.
    flux_spec = pa.TallySpec(flux=pa.Flux.GAISSER, charge_fraction=False,
                             scale=1e-4, cos_from_direction=False,
                             cos_theta=1.0) if csda else None
    tallies = []

    # Inputs of the timed passes are copied in before the timed region when
    # they fit in 32 GB (each pass transports its own copy); larger batches
    # restore one buffer at the start of each pass, inside the timed region.
    staged = []
    restore_in_loop = pristine.numel() * args.steps > (32 << 30)

    def one_pass(k, buf=None, prec=None, eager=None):
        if buf is None:
            work.copy_(pristine)
            buf = work
        st = pa.transport_device(
            phys, ctx, geom, buf.data_ptr(), n, seed=0,
            first_history=k * n, stream=stream,
            precision=precision if prec is None else prec,
            block_size=tuned["block_size"],
            blocks_per_cu=tuned["blocks_per_cu"],
            ehs_batch=tuned["ehs_batch"],
            eager=(not args.graph) if eager is None else eager,
            lds_tables=args.lds_tables, max_steps=args.max_steps,
            tail_del=args.tail_del)
        if flux_spec is not None:
            tallies.append(pa.tally_device(buf.data_ptr(), n, spec=flux_spec,
                                           stream=stream))
        return st

    for k in range(args.warmup):
        st = one_pass(1000 + k)
        log(f"bench: warmup pass {k}: {st.kernel_ms:.1f} ms, {st.steps} steps")

    staged = []
    if not restore_in_loop:
        staged = [pristine.clone() for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    passes = []
    for k in range(args.steps):
        st = one_pass(k, None if restore_in_loop else staged[k])
        passes.append(st)
        log(f"bench: pass {k}: {st.kernel_ms:.1f} ms, {st.steps} steps, "
            f"stepping kernel {st.step_kernel_ms:.1f} ms over "
            f"{st.step_launches} launches")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    del staged

    def side(label, prec, eager):
        """Passes of the same batch after the timed region, timed the same
        way (wall clock between two device synchronisations)."""
        if args.extra_passes <= 0 or args.workload != "slabs":
            return None
        k_passes = min(args.extra_passes, args.steps)
        bufs = [pristine.clone() for _ in range(k_passes)] \
            if not restore_in_loop else [None] * k_passes
        one_pass(2000, prec=prec, eager=eager)  # warm-up of this mode
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        res = [one_pass(3000 + j, bufs[j], prec=prec, eager=eager)
               for j in range(k_passes)]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t1
        del bufs
        steps = sum(st.steps for st in res)
        log(f"bench: {label}: {k_passes} passes, {steps / dt:.4e} steps/s")
        return {"value": steps / dt, "unit": "muon-steps/s",
                "events_per_s": k_passes * n / dt,
                "ms_per_step": 1e3 * dt / k_passes, "passes": k_passes}

    graph_replay = side("graph replay", precision, False) \
        if not args.graph else None
    fp32_bulk = side("FP32-bulk", pa.Precision.FP32_BULK, not args.graph) \
        if args.precision == "fp64" else None

    events_per_s = args.steps * n / elapsed
    mean_ms = sum(st.kernel_ms for st in passes) / len(passes)
    if csda:
        # One muon-step per history on the closed form path; the dominant
        # kernel is k_init (one launch per chunk of up to 2^30 histories: a
        # closed form launch keeps no per-history records), timed with HIP
        # events around each launch on its stream.
        total_steps = args.steps * n
        kernel, launches = "k_init", args.steps * ((n + (1 << 30) - 1) >>
                                                   30)
        kernel_ms = sum(st.init_kernel_ms for st in passes)
        unit_bytes = BYTES_PER_HISTORY_CSDA
    else:
        # Roofline of the stepping kernel: algorithmic bytes of its
        # launches over their summed duration (HIP events around each
        # launch, on the stream it runs on).
        total_steps = sum(st.steps for st in passes)
        kernel, launches = "k_step", sum(st.step_launches for st in passes)
        kernel_ms = sum(st.step_kernel_ms for st in passes)
        if kernel_ms <= 0.0:  # graph replays: no per-launch timing
            kernel_ms = sum(st.kernel_ms for st in passes)
        unit_bytes = BYTES_PER_STEP_FP64
    steps_per_s = total_steps / elapsed
    if fp32_bulk:
        fp32_bulk["vs_fp64"] = fp32_bulk["value"] / steps_per_s
    if graph_replay:
        graph_replay["vs_eager"] = graph_replay["value"] / steps_per_s
    achieved = total_steps * unit_bytes / (kernel_ms * 1e-3) / 1e9

    # HBM bytes per launch of the dominant kernel, from the committed
    # FETCH_SIZE / WRITE_SIZE profile of this workload at its default size,
    # used only when the profile was taken with the launch settings timed
    # here.
    traffic = None
    traffic_note = None
    profiled = {"slabs": 10_000_000, "csda": 100_000_000}
    tpath = os.path.join(HERE, "profiles",
                         f"{kernel}_{args.workload}_traffic.json")
    if (profiled.get(args.workload) == n and os.path.exists(tpath) and
            args.precision == "fp64"):
        with open(tpath) as f:
            prof = json.load(f)
        # One entry per profiled launch, or a single profile.
        entries = prof.get("profiles", [prof])
        match = [e for e in entries if e.get("launch") == tuned]
        if match:
            traffic = match[0].get("bytes_per_launch")
            traffic_note = prof.get("note")
        else:
            traffic_note = (f"profiled launches {[e.get('launch') for e in entries]}"
                            f" differ from the timed one {tuned}")

    serve = None
    if args.serve_seconds > 0 and args.workload == "slabs":
        log("bench: serving run")
        serve = serving(pa, phys, args.serve_seconds, args.serve_slots)
        serve["note"] = ("Poisson arrivals at 200/s of 10^U[3,6] muons offer "
                         "~2.9e7 muons/s, above one GPU's capacity: these "
                         "latencies include queueing")
        # The same request mix at half of the throughput just measured: the
        # latency of requests that do not queue behind a backlog.
        mean_n = serve["muons"] / max(serve["requests"], 1)
        rate = 0.5 * serve["events_per_s"] / mean_n
        log(f"bench: serving run at half load ({rate:.2f} requests/s)")
        half = serving(pa, phys, max(2.0, 8.0 / rate), args.serve_slots,
                       rate=rate, seed=1)
        if half:
            half["note"] = "arrival rate set to half the measured capacity"
            serve["half_load"] = half

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (reference, all host cores)")
        cpu = cpu_baseline(sample, args.cpu_seconds, pname, desc, ctx)

    line = {
        "metric": "muon-steps/s & events/s at <=0.1% flux bias vs FP64; "
                  "p50/p99 request latency",
        "value": steps_per_s,
        "unit": "muon-steps/s",
        "events_per_s": events_per_s,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": WORKLOADS[args.workload].format(lines=args.lines),
            "histories": n,
            "precision": args.precision,
            "launch": tuned,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": kernel,
            "kernel_avg_ms": kernel_ms / max(launches, 1),
            "launches_per_pass": launches / args.steps,
            "bytes_per_unit": unit_bytes,
            "traffic_note": traffic_note,
            "measured_bytes_per_unit": (
                traffic * launches / (n * args.steps if csda else
                                      total_steps) if traffic else None),
            "bound_note": (
                "frac is the algorithmic-bytes roofline; the kernels are not "
                "bandwidth bound: the stepping kernel keeps a history in "
                "registers between DEL vertices (a few B of HBM traffic per "
                "step measured) and is bound by VALU issue over divergent "
                "lanes (VALUBusy ~66 %, VALUUtilization ~35 %); the closed "
                "form entry kernel moves exactly its 208 B per history and "
                "is bound by the latency of its dependent table lookups, all "
                "L2 hits (VALUBusy ~32 %); profiles/ holds the counters"),
        },
        "transport_ms_per_pass": mean_ms,
        "history_status": {
            "step_limited": sum(st.step_limited for st in passes),
            "loop_limited": sum(st.loop_limited for st in passes),
            "no_target": sum(st.no_target for st in passes),
            "note": "histories of the timed passes ended with a status "
                    "code (enum pumas_gpu_history_status)"},
        "graph_replay": graph_replay,
        "fp32_bulk": fp32_bulk,
        "cpu_baseline": cpu,
        "serving": serve,
    }
    if tallies:
        t = tallies[-1]
        line["tally"] = {"flux_per_cm2_s_sr": t.mean, "error": t.error,
                         "model": "Gaisser form, cos(theta) = 1",
                         "note": "computed on the device inside each timed "
                                 "pass"}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
