"""pumas_amd: MI355X transport of batches of muon histories.

Python host side of libpumas_hip (include/pumas_gpu.h). It mirrors the
reference's transport surface: a physics object, a context with the
`mode` / `event` / `limit` / `accuracy` fields of `struct pumas_context`
(include/pumas.h:554-672), media described as data instead of callbacks, and
`transport`, the batch counterpart of `pumas_context_transport`
(include/pumas.h:1046-1048). Errors are raised as `PumasError` carrying the
reference's `enum pumas_return` code.

The compute path is the HIP library only: importing this package fails when
`libpumas_hip.so` has not been built.
"""
import ctypes
import dataclasses
import os

import numpy as np

from . import _abi
from ._abi import (Event, Flux, GeometryKind, HistoryStatus, Locals, Mode,
                   Precision, Return, Step, STATE_DTYPE)

__all__ = [
    "Context", "Event", "Flux", "Geometry", "GeometryDescription",
    "Locals", "Medium", "Recorder",
    "GeometryKind", "HistoryStatus", "Mode", "Physics", "Precision", "PumasError", "Return",
    "STATE_DTYPE", "Server", "Step", "Tally", "TallySpec", "library",
    "new_states",
    "autotune", "physics_path", "tally", "tally_device", "transport",
    "transport_device",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
# PUMAS_HIP_LIBRARY selects another build of the same library (e.g. a
# variant compiled with different launch parameters).
LIBRARY_PATH = os.environ.get("PUMAS_HIP_LIBRARY",
                              os.path.join(_HERE, "libpumas_hip.so"))


def physics_path(name):
    """Path of a PTAB physics file shipped with the package."""
    return os.path.join(_HERE, "physics", name + ".ptab")


class PumasError(RuntimeError):
    """A library call returned a code other than PUMAS_RETURN_SUCCESS."""

    def __init__(self, code, message):
        self.code = Return(code)
        super().__init__(f"{self.code.name}: {message}")


_lib = None


def library():
    """The loaded libpumas_hip (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY_PATH):
            raise ImportError(
                f"{LIBRARY_PATH} is missing: build it with "
                "`make -C pumas_amd/csrc`")
        lib = ctypes.CDLL(LIBRARY_PATH)
        _abi.declare(lib)
        _lib = lib
    return _lib


def _check(rc):
    if rc != 0:
        msg = library().pumas_gpu_error_message().decode()
        raise PumasError(rc, msg)


def new_states(n):
    """A zeroed array of `n` struct pumas_state records."""
    return np.zeros(n, dtype=STATE_DTYPE)


class Physics:
    """Physics tables uploaded to the device (pumas_gpu_physics_upload)."""

    def __init__(self, path):
        self.path = str(path)
        self._handle = ctypes.c_void_p()
        _check(library().pumas_gpu_physics_upload(
            self.path.encode(), ctypes.byref(self._handle)))

    @property
    def handle(self):
        return self._handle

    @property
    def n_materials(self):
        return library().pumas_gpu_physics_n_materials(self._handle)

    @property
    def n_energies(self):
        return library().pumas_gpu_physics_n_energies(self._handle)

    @property
    def cutoff(self):
        return library().pumas_gpu_physics_cutoff(self._handle)

    def destroy(self):
        if self._handle:
            library().pumas_gpu_physics_destroy(ctypes.byref(self._handle))

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


@dataclasses.dataclass
class Context:
    """Value fields of `struct pumas_context` (include/pumas.h:636-672),
    with the defaults of pumas_context_create for a muon
    (src/pumas.c:2582-2605)."""
    energy_loss: int = Mode.STRAGGLED
    decay: int = Mode.WEIGHTED
    direction: int = Mode.FORWARD
    scattering: int = Mode.MIXED
    event: int = Event.NONE
    limit_energy: float = 0.0
    limit_distance: float = 0.0
    limit_grammage: float = 0.0
    limit_time: float = 0.0
    accuracy: float = 1e-2

    def replace(self, **changes):
        """A copy with some fields changed."""
        return dataclasses.replace(self, **changes)

    def to_c(self):
        c = _abi.GpuContext()
        c.mode.energy_loss = int(self.energy_loss)
        c.mode.decay = int(self.decay)
        c.mode.direction = int(self.direction)
        c.mode.scattering = int(self.scattering)
        c.event = int(self.event)
        c.limit.energy = self.limit_energy
        c.limit.distance = self.limit_distance
        c.limit.grammage = self.limit_grammage
        c.limit.time = self.limit_time
        c.accuracy = self.accuracy
        return c


def _ptr(array, ctype):
    if array is None:
        return None
    return array.ctypes.data_as(ctypes.POINTER(ctype))


MEDIUM_DTYPE = np.dtype([("material", "i4"), ("density", "f8"),
                         ("locals", "i4"), ("scale", "f8"), ("z_ref", "f8"),
                         ("cos_min", "f8")], align=True)


@dataclasses.dataclass(frozen=True)
class Medium:
    """A medium (struct pumas_gpu_medium): a material index and its local
    properties, what the reference's `pumas_locals_cb` returns
    (include/pumas.h:325-349). Uniform by default; `exponential` gives the
    density profile density * exp(-(z - z_ref) / scale) with a proposed step
    scale / max(|u_z|, cos_min), as the atmospheres of
    examples/pumas/geometry.c:78-97 and tests/test-pumas.c:1871-1884."""
    material: int
    density: float = 0.0
    locals: Locals = Locals.UNIFORM
    scale: float = 0.0
    z_ref: float = 0.0
    cos_min: float = 0.0

    @classmethod
    def exponential(cls, material, density, scale, z_ref=0.0, cos_min=1.0):
        return cls(material, density, Locals.EXPONENTIAL, scale, z_ref,
                   cos_min)

    def row(self):
        return (self.material, self.density, int(self.locals), self.scale,
                self.z_ref, self.cos_min)


def _medium_row(m):
    """A Medium, or a (material, density) pair for a uniform medium."""
    if isinstance(m, Medium):
        return m.row()
    material, density = m
    return (material, density, 0, 0.0, 0.0, 0.0)


class GeometryDescription:
    """Host description of a device geometry (struct
    pumas_gpu_geometry_desc). The numpy arrays are kept alive here."""

    def __init__(self, kind, media, step_type=Step.CHECK, z=None,
                 grid=None, height=None, density=None, label=None,
                 z_top=0.0):
        self.kind = GeometryKind(kind)
        self.step_type = Step(step_type)
        self.media = np.array([_medium_row(m) for m in media],
                              dtype=MEDIUM_DTYPE)
        self.z = None if z is None else np.ascontiguousarray(z, "f8")
        self.grid = grid or {}
        self.height = None if height is None else np.ascontiguousarray(
            height, "f4")
        self.density = None if density is None else np.ascontiguousarray(
            density, "f4")
        self.label = None if label is None else np.ascontiguousarray(
            label, "u1")
        self.z_top = float(z_top)

    def to_c(self):
        d = _abi.GeometryDesc()
        d.kind = int(self.kind)
        d.step_type = int(self.step_type)
        d.media = self.media.ctypes.data_as(ctypes.POINTER(_abi.GpuMedium))
        d.n_media = len(self.media)
        if self.z is not None:
            d.z = _ptr(self.z, ctypes.c_double)
            d.n_layers = len(self.z) - 1
        g = self.grid
        d.nx, d.ny, d.nz = g.get("nx", 0), g.get("ny", 0), g.get("nz", 0)
        d.x0, d.y0, d.z0 = g.get("x0", 0.), g.get("y0", 0.), g.get("z0", 0.)
        d.dx, d.dy, d.dz = g.get("dx", 0.), g.get("dy", 0.), g.get("dz", 0.)
        if self.height is not None:
            d.height = _ptr(self.height, ctypes.c_float)
        if self.density is not None:
            d.density = _ptr(self.density, ctypes.c_float)
        if self.label is not None:
            d.label = _ptr(self.label, ctypes.c_uint8)
        d.z_top = self.z_top
        return d

    @classmethod
    def infinite(cls, material, density=0.0):
        """One infinite uniform medium, as examples/pumas/straight.c:61-74."""
        return cls(GeometryKind.INFINITE, [(material, density)])

    @classmethod
    def slabs(cls, z, media, step_type=Step.RAW):
        """Horizontal layers; medium i fills z in [z[i], z[i+1])."""
        return cls(GeometryKind.SLABS, media, step_type=step_type, z=z)

    @classmethod
    def terrain(cls, height, cell_density, x0, y0, dx, dy, z0, z_top,
                air=(2, 0.0), ground_material=0):
        """Heightmap terrain: `height` [ny][nx] node heights, one rock
        medium per cell with density `cell_density` [ny-1][nx-1]."""
        ny, nx = height.shape
        return cls(GeometryKind.TERRAIN, [air, (ground_material, 0.0)],
                   step_type=Step.CHECK,
                   grid=dict(nx=nx, ny=ny, x0=x0, y0=y0, z0=z0, dx=dx, dy=dy),
                   height=height, density=cell_density, z_top=z_top)

    @classmethod
    def voxels(cls, label, density, media, x0, y0, z0, dx, dy, dz,
               step_type=Step.RAW):
        """Voxel grid: `label` [nz][ny][nx] medium class, `density` the
        voxel density in kg/m^3."""
        nz, ny, nx = label.shape
        return cls(GeometryKind.VOXELS, media, step_type=step_type,
                   grid=dict(nx=nx, ny=ny, nz=nz, x0=x0, y0=y0, z0=z0,
                             dx=dx, dy=dy, dz=dz),
                   label=label, density=density)


class Geometry:
    """Device geometry (pumas_gpu_geometry_create)."""

    def __init__(self, physics, description):
        self.description = description
        self._handle = ctypes.c_void_p()
        self._desc = description.to_c()
        _check(library().pumas_gpu_geometry_create(
            physics.handle, ctypes.byref(self._desc),
            ctypes.byref(self._handle)))

    @property
    def handle(self):
        return self._handle

    def destroy(self):
        if self._handle:
            library().pumas_gpu_geometry_destroy(ctypes.byref(self._handle))

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


@dataclasses.dataclass
class TransportResult:
    states: object
    events: object = None
    media: object = None
    steps: object = None
    draws: object = None
    rc: object = None
    total_steps: int = 0
    kernel_ms: float = 0.0
    lds_bytes: int = 0
    # Histories ended with a status code (enum pumas_gpu_history_status).
    step_limited: int = 0
    loop_limited: int = 0
    no_target: int = 0


@dataclasses.dataclass
class TransportStats:
    """Counters of one pumas_gpu_transport call."""
    steps: int = 0
    histories: int = 0
    kernel_ms: float = 0.0
    step_kernel_ms: float = 0.0
    step_launches: int = 0
    rounds: int = 0
    lds_bytes: int = 0
    step_limited: int = 0
    loop_limited: int = 0
    no_target: int = 0
    init_kernel_ms: float = 0.0

    @classmethod
    def from_c(cls, st):
        return cls(int(st.steps), int(st.histories), float(st.kernel_ms),
                   float(st.step_kernel_ms), int(st.step_launches),
                   int(st.rounds), int(st.lds_bytes), int(st.step_limited),
                   int(st.loop_limited), int(st.no_target),
                   float(st.init_kernel_ms))


def _options(precision, block_size, blocks_per_cu, max_steps,
             rounds_per_check=0, eager=False, ehs_batch=0, recorder=None,
             tail_del=0, lds_tables=0, lanes_per_wave=0, max_loop=0):
    o = _abi.GpuOptions()
    o.precision = int(precision)
    o.block_size = block_size
    o.blocks_per_cu = blocks_per_cu
    o.max_steps = max_steps
    o.rounds_per_check = rounds_per_check
    o.eager = int(bool(eager))
    o.ehs_batch = ehs_batch
    o.recorder = recorder.handle if recorder is not None else None
    o.tail_del = int(tail_del)
    o.lds_tables = int(lds_tables)
    o.lanes_per_wave = int(lanes_per_wave)
    o.max_loop = int(max_loop)
    return o


def transport(physics, context, geometry, states, seed=0, first_history=0,
              outputs=True, precision=Precision.FP64, block_size=0,
              blocks_per_cu=0, max_steps=0, raise_errors=True,
              rounds_per_check=0, eager=False, ehs_batch=0, recorder=None,
              tail_del=0, lds_tables=0, lanes_per_wave=0, max_loop=0):
    """Transport host `states` (a STATE_DTYPE array, updated in place).

    Batch counterpart of pumas_context_transport: history i draws from the
    Philox stream keyed by `seed` at history id `first_history + i`.
    Returns a TransportResult with the per-history outputs when `outputs`.
    A history that raises an error leaves the others unaffected; the first
    error is raised as PumasError, or stored in `result.error` when
    `raise_errors` is false.
    """
    lib = library()
    states = np.ascontiguousarray(states, dtype=STATE_DTYPE)
    n = len(states)
    res = TransportResult(states=states)
    out = _abi.GpuOutputs()
    if outputs:
        res.events = np.zeros(n, np.int32)
        res.media = np.zeros((n, 2), np.int32)
        res.steps = np.zeros(n, np.uint32)
        res.draws = np.zeros(n, np.uint32)
        res.rc = np.zeros(n, np.int32)
        out.events = res.events.ctypes.data
        out.media = res.media.ctypes.data
        out.steps = res.steps.ctypes.data
        out.draws = res.draws.ctypes.data
        out.rc = res.rc.ctypes.data
    stats = _abi.GpuStats()
    opts = _options(precision, block_size, blocks_per_cu, max_steps,
                    rounds_per_check, eager, ehs_batch, recorder, tail_del,
                    lds_tables, lanes_per_wave, max_loop)
    ctx = context.to_c()
    rc = lib.pumas_gpu_transport_host(
        physics.handle, ctypes.byref(ctx), geometry.handle,
        states.ctypes.data, n, seed, first_history, ctypes.byref(out),
        ctypes.byref(opts), ctypes.byref(stats))
    res.error = Return(rc)
    if raise_errors:
        _check(rc)
    res.total_steps = int(stats.steps)
    res.kernel_ms = float(stats.kernel_ms)
    res.lds_bytes = int(stats.lds_bytes)
    res.step_limited = int(stats.step_limited)
    res.loop_limited = int(stats.loop_limited)
    res.no_target = int(stats.no_target)
    return res


def transport_device(physics, context, geometry, states_ptr, n, seed=0,
                     first_history=0, outputs=None, precision=Precision.FP64,
                     block_size=0, blocks_per_cu=0, max_steps=0, stream=None,
                     stats=True, rounds_per_check=0, eager=False,
                     ehs_batch=0, recorder=None, tail_del=0,
                     lds_tables=0, lanes_per_wave=0, max_loop=0):
    """Transport `n` states already in device memory at `states_ptr`.

    `outputs` maps any of "events", "media", "steps", "draws", "rc" to a
    device pointer. `stream` is a hipStream_t handle (int) or None.
    Returns a TransportStats when `stats`, else None.
    """
    lib = library()
    out = _abi.GpuOutputs()
    for k, v in (outputs or {}).items():
        setattr(out, k, v)
    st = _abi.GpuStats()
    opts = _options(precision, block_size, blocks_per_cu, max_steps,
                    rounds_per_check, eager, ehs_batch, recorder, tail_del,
                    lds_tables, lanes_per_wave, max_loop)
    ctx = context.to_c()
    _check(lib.pumas_gpu_transport(
        physics.handle, ctypes.byref(ctx), geometry.handle,
        ctypes.c_void_p(states_ptr), n, seed, first_history,
        ctypes.byref(out), ctypes.byref(opts),
        ctypes.byref(st) if stats else None,
        ctypes.c_void_p(stream) if stream else None))
    if stats:
        return TransportStats.from_c(st)
    return None


@dataclasses.dataclass
class TallySpec:
    """struct pumas_gpu_tally_spec: flux model, selection and histograms."""
    flux: int = Flux.GAISSER
    charge_fraction: bool = True
    scale: float = 1.0
    cos_from_direction: bool = True
    cos_theta: float = 1.0
    event_mask: int = 0
    n_energy_bins: int = 0
    energy_min: float = 1e-2
    energy_max: float = 1e4
    n_cos_bins: int = 0
    cos_min: float = 0.0
    cos_max: float = 1.0

    def to_c(self):
        c = _abi.TallySpec()
        c.flux = int(self.flux)
        c.charge_fraction = int(bool(self.charge_fraction))
        c.scale = self.scale
        c.cos_from_direction = int(bool(self.cos_from_direction))
        c.cos_theta = self.cos_theta
        c.event_mask = int(self.event_mask)
        c.n_energy_bins = self.n_energy_bins
        c.energy_min = self.energy_min
        c.energy_max = self.energy_max
        c.n_cos_bins = self.n_cos_bins
        c.cos_min = self.cos_min
        c.cos_max = self.cos_max
        return c


@dataclasses.dataclass
class Tally:
    """Flux estimator of a batch: mean and standard error of weight x flux
    over the tallied histories, and the energy / cos(zenith) histograms
    (per bin: sum, sum of squares)."""
    sum: float
    sum2: float
    count: int
    energy: object
    cos_theta: object
    n: int

    @property
    def mean(self):
        """Estimator per history of the batch (as examples/pumas/straight.c:
        every history counts in the denominator)."""
        return self.sum / self.n if self.n else 0.0

    @property
    def error(self):
        if self.n < 2:
            return 0.0
        m = self.mean
        return np.sqrt(max(self.sum2 / self.n - m * m, 0.0) / self.n)


def _tally_call(fn, states_arg, events_arg, n, spec):
    c = spec.to_c()
    r = _abi.TallyResult()
    energy = np.zeros((max(spec.n_energy_bins, 0), 2))
    cos_t = np.zeros((max(spec.n_cos_bins, 0), 2))
    r.energy = energy.ctypes.data if spec.n_energy_bins > 0 else None
    r.cos_theta = cos_t.ctypes.data if spec.n_cos_bins > 0 else None
    _check(fn(states_arg, events_arg, n, ctypes.byref(c), ctypes.byref(r)))
    return Tally(float(r.sum), float(r.sum2), int(r.count), energy, cos_t, n)


def tally(states, events=None, spec=None):
    """Tally host `states` (and `events`) on the device."""
    lib = library()
    spec = spec or TallySpec()
    states = np.ascontiguousarray(states, dtype=STATE_DTYPE)
    ev = None
    if events is not None:
        events = np.ascontiguousarray(events, dtype=np.int32)
        ev = events.ctypes.data
    return _tally_call(lib.pumas_gpu_tally_host, states.ctypes.data, ev,
                       len(states), spec)


def tally_device(states_ptr, n, events_ptr=None, spec=None, stream=None):
    """Tally `n` device-resident states at `states_ptr` on `stream`."""
    lib = library()
    spec = spec or TallySpec()

    def fn(s, e, n_, c, r):
        return lib.pumas_gpu_tally(s, e, n_, c, r,
                                   ctypes.c_void_p(stream) if stream else None)
    return _tally_call(fn, ctypes.c_void_p(states_ptr),
                       ctypes.c_void_p(events_ptr) if events_ptr else None,
                       n, spec)


@dataclasses.dataclass
class RequestResult:
    """A served request: its TransportResult and timings."""
    result: TransportResult
    rc: int
    queue_ms: float
    latency_ms: float
    kernel_ms: float
    steps: int
    message: str


class Server:
    """Continuous batching server (pumas_gpu_server): requests run as soon
    as a slot and their device memory are free, concurrently with the
    requests in flight; results are those of the same request alone."""

    def __init__(self, hbm_budget=0, slots=0):
        lib = library()
        o = _abi.ServerOptions()
        o.hbm_budget = int(hbm_budget)
        o.slots = int(slots)
        h = ctypes.c_void_p()
        _check(lib.pumas_gpu_server_create(ctypes.byref(o), ctypes.byref(h)))
        self.handle = h
        self._inflight = {}

    def submit(self, physics, context, geometry, states, seed=0,
               first_history=0, precision=Precision.FP64, outputs=True):
        """Queue host `states` (updated in place); returns a ticket."""
        lib = library()
        states = np.ascontiguousarray(states, dtype=STATE_DTYPE)
        n = len(states)
        res = TransportResult(states=states)
        r = _abi.Request()
        r.physics = physics.handle
        r.geometry = geometry.handle
        r.context = context.to_c()
        r.precision = int(precision)
        r.states = states.ctypes.data
        r.n = n
        r.seed = seed
        r.first_history = first_history
        if outputs:
            res.events = np.zeros(n, np.int32)
            res.media = np.zeros((n, 2), np.int32)
            res.steps = np.zeros(n, np.uint32)
            res.draws = np.zeros(n, np.uint32)
            res.rc = np.zeros(n, np.int32)
            r.outputs.events = res.events.ctypes.data
            r.outputs.media = res.media.ctypes.data
            r.outputs.steps = res.steps.ctypes.data
            r.outputs.draws = res.draws.ctypes.data
            r.outputs.rc = res.rc.ctypes.data
        r.device = 0
        t = ctypes.c_uint64()
        _check(lib.pumas_gpu_server_submit(self.handle, ctypes.byref(r),
                                           ctypes.byref(t)))
        # The arrays stay referenced until the request is waited for.
        self._inflight[t.value] = (res, physics, geometry)
        return t.value

    def done(self, ticket):
        return library().pumas_gpu_server_poll(self.handle, ticket) == 1

    def wait(self, ticket, raise_errors=True):
        lib = library()
        st = _abi.RequestStatus()
        rc = lib.pumas_gpu_server_wait(self.handle, ticket, ctypes.byref(st))
        res = self._inflight.pop(ticket)[0]
        if raise_errors:
            _check(rc)
        res.total_steps = int(st.steps)
        res.kernel_ms = float(st.kernel_ms)
        return RequestResult(res, int(st.rc), float(st.queue_ms),
                             float(st.latency_ms), float(st.kernel_ms),
                             int(st.steps), st.message.decode(errors="replace"))

    def close(self):
        if self.handle:
            library().pumas_gpu_server_destroy(ctypes.byref(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(physics, context, geometry, states_ptr, n, seed=0,
             precision=Precision.FP64, stream=None):
    """Launch settings (block_size, blocks_per_cu, ehs_batch) of the fastest
    candidate on a sample of `n` device-resident states (left unchanged),
    and its rate in Monte Carlo steps per second."""
    lib = library()
    best = _abi.GpuOptions()
    rate = ctypes.c_double()
    ctx = context.to_c()
    _check(lib.pumas_gpu_autotune(
        physics.handle, ctypes.byref(ctx), geometry.handle,
        ctypes.c_void_p(states_ptr), n, seed, int(precision),
        ctypes.byref(best), ctypes.byref(rate),
        ctypes.c_void_p(stream) if stream else None))
    return ({"block_size": best.block_size,
             "blocks_per_cu": best.blocks_per_cu,
             "ehs_batch": best.ehs_batch}, float(rate.value))


class Recorder:
    """Device frame recorder (pumas_gpu_recorder), the counterpart of the
    reference's pumas_recorder: frames of histories with ids in
    [first_history, first_history + n_histories), every `period`-th step
    plus track start/stop, DEL and medium-change points."""

    def __init__(self, capacity, period=1, first_history=0, n_histories=1):
        lib = library()
        h = ctypes.c_void_p()
        _check(lib.pumas_gpu_recorder_create(capacity, period, first_history,
                                             n_histories, ctypes.byref(h)))
        self.handle = h
        self.capacity = capacity

    def frames(self):
        """Recorded frames (FRAME_DTYPE array, in recording order) and the
        number of frames recorded, dropped ones included."""
        lib = library()
        out = np.zeros(self.capacity, _abi.FRAME_DTYPE)
        length = ctypes.c_uint64()
        _check(lib.pumas_gpu_recorder_read(self.handle, out.ctypes.data,
                                           self.capacity,
                                           ctypes.byref(length)))
        return out[:min(length.value, self.capacity)], int(length.value)

    def clear(self):
        _check(library().pumas_gpu_recorder_clear(self.handle))

    def close(self):
        if self.handle:
            library().pumas_gpu_recorder_destroy(ctypes.byref(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
