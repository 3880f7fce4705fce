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
from ._abi import (Event, GeometryKind, Mode, Precision, Return, Step,
                   STATE_DTYPE)

__all__ = [
    "Context", "Event", "Geometry", "GeometryDescription", "GeometryKind", "Mode", "Physics",
    "Precision", "PumasError", "Return", "STATE_DTYPE", "Step", "library",
    "new_states", "physics_path", "transport",
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
                "`python -c 'import __graft_entry__ as g; g.build()'` "
                "(or `make -C pumas_amd/csrc`)")
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


class GeometryDescription:
    """Host description of a device geometry (struct
    pumas_gpu_geometry_desc). The numpy arrays are kept alive here."""

    def __init__(self, kind, media, step_type=Step.CHECK, z=None,
                 grid=None, height=None, density=None, label=None,
                 z_top=0.0):
        self.kind = GeometryKind(kind)
        self.step_type = Step(step_type)
        self.media = np.array(
            [(m, d) for m, d in media],
            dtype=np.dtype([("material", "i4"), ("density", "f8")],
                           align=True))
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


@dataclasses.dataclass
class TransportStats:
    """Counters of one pumas_gpu_transport call."""
    steps: int = 0
    histories: int = 0
    kernel_ms: float = 0.0
    step_kernel_ms: float = 0.0
    step_launches: int = 0
    rounds: int = 0

    @classmethod
    def from_c(cls, st):
        return cls(int(st.steps), int(st.histories), float(st.kernel_ms),
                   float(st.step_kernel_ms), int(st.step_launches),
                   int(st.rounds))


def _options(precision, block_size, blocks_per_cu, max_steps,
             rounds_per_check=0):
    o = _abi.GpuOptions()
    o.precision = int(precision)
    o.block_size = block_size
    o.blocks_per_cu = blocks_per_cu
    o.max_steps = max_steps
    o.rounds_per_check = rounds_per_check
    return o


def transport(physics, context, geometry, states, seed=0, first_history=0,
              outputs=True, precision=Precision.FP64, block_size=0,
              blocks_per_cu=0, max_steps=0, raise_errors=True):
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
    opts = _options(precision, block_size, blocks_per_cu, max_steps)
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
    return res


def transport_device(physics, context, geometry, states_ptr, n, seed=0,
                     first_history=0, outputs=None, precision=Precision.FP64,
                     block_size=0, blocks_per_cu=0, max_steps=0, stream=None,
                     stats=True, rounds_per_check=0):
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
                    rounds_per_check)
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
