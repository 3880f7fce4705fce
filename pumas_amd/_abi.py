"""ctypes and numpy mirrors of the C ABI in include/pumas_gpu.h.

The state and context layouts are those of the reference public header
(include/pumas.h:289-309 for `struct pumas_state`, include/pumas.h:554-594
for the context mode and limits), so that arrays can be shared with the
reference library unchanged.
"""
import ctypes
import enum

import numpy as np


class Mode(enum.IntEnum):
    """enum pumas_mode, include/pumas.h:90-140 (aliased values kept)."""
    DISABLED = -1
    CSDA = 0
    MIXED = 1
    STRAGGLED = 2
    WEIGHTED = 0
    RANDOMISED = 1
    FORWARD = 0
    BACKWARD = 1


class Event(enum.IntFlag):
    """enum pumas_event, include/pumas.h:199-236."""
    NONE = 0
    LIMIT_ENERGY = 1
    LIMIT_DISTANCE = 2
    LIMIT_GRAMMAGE = 4
    LIMIT_TIME = 8
    LIMIT = 15
    MEDIUM = 16
    VERTEX_BREMSSTRAHLUNG = 32
    VERTEX_PAIR_CREATION = 64
    VERTEX_PHOTONUCLEAR = 128
    VERTEX_DELTA_RAY = 256
    VERTEX_DEL = 480
    VERTEX_COULOMB = 512
    VERTEX_DECAY = 1024
    VERTEX = 2016
    WEIGHT = 2048
    START = 4096
    STOP = 8192


class Return(enum.IntEnum):
    """enum pumas_return, include/pumas.h:143-196."""
    SUCCESS = 0
    ACCURACY_ERROR = 1
    END_OF_FILE = 2
    DECAY_ERROR = 3
    DENSITY_ERROR = 4
    DIRECTION_ERROR = 5
    INCOMPLETE_FILE = 6
    INDEX_ERROR = 7
    PHYSICS_ERROR = 8
    INTERNAL_ERROR = 9
    IO_ERROR = 10
    FORMAT_ERROR = 11
    MEDIUM_ERROR = 12
    MEMORY_ERROR = 13
    MODEL_ERROR = 14
    MISSING_LIMIT = 15
    MISSING_RANDOM = 16
    PATH_ERROR = 17
    RAISE_ERROR = 18
    TOO_LONG = 19
    UNDEFINED_MDF = 20
    UNKNOWN_ELEMENT = 21
    UNKNOWN_MATERIAL = 22
    UNKNOWN_PARTICLE = 23
    VALUE_ERROR = 24


class HistoryStatus(enum.IntEnum):
    """enum pumas_gpu_history_status: per-history codes (outputs rc) of
    histories whose end state is not the reference's."""
    STEP_LIMIT = 100
    LOOP_LIMIT = 101
    NO_TARGET = 102


class GeometryKind(enum.IntEnum):
    INFINITE = 0
    SLABS = 1
    TERRAIN = 2
    VOXELS = 3


class Locals(enum.IntEnum):
    """enum pumas_gpu_locals."""
    UNIFORM = 0
    EXPONENTIAL = 1


class Step(enum.IntEnum):
    """enum pumas_step, include/pumas.h:461-476."""
    CHECK = 0
    RAW = 1


class Precision(enum.IntEnum):
    FP64 = 0
    FP32_BULK = 1


# struct pumas_state as a numpy record (104 bytes, 8-byte aligned).
STATE_DTYPE = np.dtype([
    ("charge", "f8"), ("energy", "f8"), ("distance", "f8"),
    ("grammage", "f8"), ("time", "f8"), ("weight", "f8"),
    ("position", "f8", (3,)), ("direction", "f8", (3,)),
    ("decayed", "i4"), ("_pad", "i4"),
])
assert STATE_DTYPE.itemsize == 104


class ContextMode(ctypes.Structure):
    _fields_ = [("energy_loss", ctypes.c_int), ("decay", ctypes.c_int),
                ("direction", ctypes.c_int), ("scattering", ctypes.c_int)]


class ContextLimit(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("distance", ctypes.c_double),
                ("grammage", ctypes.c_double), ("time", ctypes.c_double)]


class GpuContext(ctypes.Structure):
    _fields_ = [("mode", ContextMode), ("event", ctypes.c_int),
                ("limit", ContextLimit), ("accuracy", ctypes.c_double)]


class GpuMedium(ctypes.Structure):
    _fields_ = [("material", ctypes.c_int), ("density", ctypes.c_double),
                ("locals", ctypes.c_int), ("scale", ctypes.c_double),
                ("z_ref", ctypes.c_double), ("cos_min", ctypes.c_double)]


class GeometryDesc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int), ("step_type", ctypes.c_int),
        ("media", ctypes.POINTER(GpuMedium)), ("n_media", ctypes.c_int),
        ("z", ctypes.POINTER(ctypes.c_double)), ("n_layers", ctypes.c_int),
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("x0", ctypes.c_double), ("y0", ctypes.c_double),
        ("z0", ctypes.c_double), ("dx", ctypes.c_double),
        ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("height", ctypes.POINTER(ctypes.c_float)),
        ("density", ctypes.POINTER(ctypes.c_float)),
        ("label", ctypes.POINTER(ctypes.c_uint8)),
        ("z_top", ctypes.c_double),
    ]


class GpuOutputs(ctypes.Structure):
    _fields_ = [("events", ctypes.c_void_p), ("media", ctypes.c_void_p),
                ("steps", ctypes.c_void_p), ("draws", ctypes.c_void_p),
                ("rc", ctypes.c_void_p)]


class GpuOptions(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("block_size", ctypes.c_int),
                ("blocks_per_cu", ctypes.c_int), ("max_steps", ctypes.c_int),
                ("rounds_per_check", ctypes.c_int), ("eager", ctypes.c_int),
                ("ehs_batch", ctypes.c_int), ("recorder", ctypes.c_void_p),
                ("tail_del", ctypes.c_int), ("lds_tables", ctypes.c_int),
                ("lanes_per_wave", ctypes.c_int), ("max_loop", ctypes.c_int)]


class GpuStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_uint64), ("histories", ctypes.c_uint64),
                ("kernel_ms", ctypes.c_double),
                ("step_kernel_ms", ctypes.c_double),
                ("step_launches", ctypes.c_uint64),
                ("rounds", ctypes.c_uint64), ("lds_bytes", ctypes.c_uint64),
                ("step_limited", ctypes.c_uint64),
                ("loop_limited", ctypes.c_uint64),
                ("no_target", ctypes.c_uint64),
                ("init_kernel_ms", ctypes.c_double)]


# Histories per call of pumas_gpu_transport_noalloc
# (PUMAS_GPU_NOALLOC_MAX_HISTORIES).
NOALLOC_MAX_HISTORIES = 1 << 18


class Flux(enum.IntEnum):
    """enum pumas_gpu_flux."""
    NONE = 0
    GAISSER = 1
    GCCLY = 2


TALLY_MAX_BINS = 256


class TallySpec(ctypes.Structure):
    _fields_ = [("flux", ctypes.c_int), ("charge_fraction", ctypes.c_int),
                ("scale", ctypes.c_double),
                ("cos_from_direction", ctypes.c_int),
                ("cos_theta", ctypes.c_double),
                ("event_mask", ctypes.c_uint),
                ("n_energy_bins", ctypes.c_int),
                ("energy_min", ctypes.c_double),
                ("energy_max", ctypes.c_double),
                ("n_cos_bins", ctypes.c_int),
                ("cos_min", ctypes.c_double), ("cos_max", ctypes.c_double)]


class TallyResult(ctypes.Structure):
    _fields_ = [("sum", ctypes.c_double), ("sum2", ctypes.c_double),
                ("count", ctypes.c_uint64), ("energy", ctypes.c_void_p),
                ("cos_theta", ctypes.c_void_p)]


SERVER_MAX_SLOTS = 32


class ServerOptions(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_size_t), ("slots", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [("physics", ctypes.c_void_p), ("geometry", ctypes.c_void_p),
                ("context", GpuContext), ("precision", ctypes.c_int),
                ("states", ctypes.c_void_p), ("n", ctypes.c_size_t),
                ("seed", ctypes.c_uint64), ("first_history", ctypes.c_uint64),
                ("outputs", GpuOutputs), ("device", ctypes.c_int)]


class RequestStatus(ctypes.Structure):
    _fields_ = [("rc", ctypes.c_int), ("queue_ms", ctypes.c_double),
                ("latency_ms", ctypes.c_double),
                ("kernel_ms", ctypes.c_double), ("steps", ctypes.c_uint64),
                ("message", ctypes.c_char * 256)]


FRAME_DTYPE = np.dtype([("state", STATE_DTYPE), ("medium", np.int32),
                        ("event", np.int32), ("history", np.uint64),
                        ("reserved", np.uint64)])
assert FRAME_DTYPE.itemsize == 128


# Every entry point declared by include/pumas_gpu.h.
EXPORTED_SYMBOLS = (
    "pumas_gpu_context_default",
    "pumas_gpu_physics_upload",
    "pumas_gpu_physics_upload_memory",
    "pumas_gpu_physics_destroy",
    "pumas_gpu_physics_n_materials",
    "pumas_gpu_physics_n_energies",
    "pumas_gpu_physics_cutoff",
    "pumas_gpu_geometry_create",
    "pumas_gpu_geometry_destroy",
    "pumas_gpu_transport",
    "pumas_gpu_transport_host",
    "pumas_gpu_transport_noalloc",
    "pumas_gpu_transport_context",
    "pumas_gpu_context_from_reference",
    "pumas_gpu_tally",
    "pumas_gpu_tally_host",
    "pumas_gpu_autotune",
    "pumas_gpu_recorder_create",
    "pumas_gpu_recorder_destroy",
    "pumas_gpu_recorder_clear",
    "pumas_gpu_recorder_read",
    "pumas_gpu_server_create",
    "pumas_gpu_server_destroy",
    "pumas_gpu_server_submit",
    "pumas_gpu_server_poll",
    "pumas_gpu_server_request_bytes",
    "pumas_gpu_server_wait",
    "pumas_gpu_error_message",
    "pumas_gpu_version",
)


def declare(lib):
    """Attach argument and result types to the library functions."""
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    lib.pumas_gpu_context_default.argtypes = [P(GpuContext)]
    lib.pumas_gpu_context_default.restype = None
    lib.pumas_gpu_physics_upload.argtypes = [ctypes.c_char_p, P(vp)]
    lib.pumas_gpu_physics_upload.restype = ctypes.c_int
    lib.pumas_gpu_physics_upload_memory.argtypes = [vp, ctypes.c_size_t, P(vp)]
    lib.pumas_gpu_physics_upload_memory.restype = ctypes.c_int
    lib.pumas_gpu_physics_destroy.argtypes = [P(vp)]
    lib.pumas_gpu_physics_destroy.restype = None
    lib.pumas_gpu_physics_n_materials.argtypes = [vp]
    lib.pumas_gpu_physics_n_materials.restype = ctypes.c_int
    lib.pumas_gpu_physics_n_energies.argtypes = [vp]
    lib.pumas_gpu_physics_n_energies.restype = ctypes.c_int
    lib.pumas_gpu_physics_cutoff.argtypes = [vp]
    lib.pumas_gpu_physics_cutoff.restype = ctypes.c_double
    lib.pumas_gpu_geometry_create.argtypes = [vp, P(GeometryDesc), P(vp)]
    lib.pumas_gpu_geometry_create.restype = ctypes.c_int
    lib.pumas_gpu_geometry_destroy.argtypes = [P(vp)]
    lib.pumas_gpu_geometry_destroy.restype = None
    lib.pumas_gpu_transport.argtypes = [
        vp, P(GpuContext), vp, vp, ctypes.c_size_t, ctypes.c_uint64,
        ctypes.c_uint64, P(GpuOutputs), P(GpuOptions), P(GpuStats), vp]
    lib.pumas_gpu_transport.restype = ctypes.c_int
    lib.pumas_gpu_transport_host.argtypes = [
        vp, P(GpuContext), vp, vp, ctypes.c_size_t, ctypes.c_uint64,
        ctypes.c_uint64, P(GpuOutputs), P(GpuOptions), P(GpuStats)]
    lib.pumas_gpu_transport_host.restype = ctypes.c_int
    lib.pumas_gpu_transport_noalloc.argtypes = [
        vp, P(GpuContext), vp, vp, ctypes.c_size_t, ctypes.c_uint64,
        ctypes.c_uint64, P(GpuOutputs), P(GpuOptions), P(GpuStats), vp]
    lib.pumas_gpu_transport_noalloc.restype = ctypes.c_int
    lib.pumas_gpu_transport_context.argtypes = [
        vp, vp, vp, vp, ctypes.c_size_t, ctypes.c_uint64,
        ctypes.c_uint64, P(GpuOutputs), P(GpuOptions), P(GpuStats), vp]
    lib.pumas_gpu_transport_context.restype = ctypes.c_int
    lib.pumas_gpu_context_from_reference.argtypes = [vp, P(GpuContext)]
    lib.pumas_gpu_context_from_reference.restype = None
    lib.pumas_gpu_tally.argtypes = [vp, vp, ctypes.c_size_t, P(TallySpec),
                                    P(TallyResult), vp]
    lib.pumas_gpu_tally.restype = ctypes.c_int
    lib.pumas_gpu_tally_host.argtypes = [vp, vp, ctypes.c_size_t,
                                         P(TallySpec), P(TallyResult)]
    lib.pumas_gpu_tally_host.restype = ctypes.c_int
    lib.pumas_gpu_recorder_create.argtypes = [
        ctypes.c_size_t, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, P(vp)]
    lib.pumas_gpu_recorder_create.restype = ctypes.c_int
    lib.pumas_gpu_recorder_destroy.argtypes = [P(vp)]
    lib.pumas_gpu_recorder_destroy.restype = None
    lib.pumas_gpu_recorder_clear.argtypes = [vp]
    lib.pumas_gpu_recorder_clear.restype = ctypes.c_int
    lib.pumas_gpu_recorder_read.argtypes = [vp, vp, ctypes.c_size_t,
                                            P(ctypes.c_uint64)]
    lib.pumas_gpu_recorder_read.restype = ctypes.c_int
    lib.pumas_gpu_autotune.argtypes = [
        vp, P(GpuContext), vp, vp, ctypes.c_size_t, ctypes.c_uint64,
        ctypes.c_int, P(GpuOptions), P(ctypes.c_double), vp]
    lib.pumas_gpu_autotune.restype = ctypes.c_int
    lib.pumas_gpu_server_create.argtypes = [P(ServerOptions), P(vp)]
    lib.pumas_gpu_server_create.restype = ctypes.c_int
    lib.pumas_gpu_server_destroy.argtypes = [P(vp)]
    lib.pumas_gpu_server_destroy.restype = None
    lib.pumas_gpu_server_submit.argtypes = [vp, P(Request),
                                            P(ctypes.c_uint64)]
    lib.pumas_gpu_server_submit.restype = ctypes.c_int
    lib.pumas_gpu_server_poll.argtypes = [vp, ctypes.c_uint64]
    lib.pumas_gpu_server_poll.restype = ctypes.c_int
    lib.pumas_gpu_server_request_bytes.argtypes = [ctypes.c_size_t,
                                                   ctypes.c_int]
    lib.pumas_gpu_server_request_bytes.restype = ctypes.c_size_t
    lib.pumas_gpu_server_wait.argtypes = [vp, ctypes.c_uint64,
                                          P(RequestStatus)]
    lib.pumas_gpu_server_wait.restype = ctypes.c_int
    lib.pumas_gpu_error_message.argtypes = []
    lib.pumas_gpu_error_message.restype = ctypes.c_char_p
    lib.pumas_gpu_version.argtypes = []
    lib.pumas_gpu_version.restype = ctypes.c_char_p
