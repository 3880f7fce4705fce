"""Python access to the oracle (test infrastructure only).

The oracle is the reference PUMAS library compiled from its own source by
oracle/Makefile into oracle/_ref/libpumas_ref.so, with the harness of
oracle/harness.c. Only tests, the smoke check and the CPU baseline of
bench.py use it, as a checker.
"""
import ctypes
import os
import shutil
import tempfile

import numpy as np

from pumas_amd import _abi
from pumas_amd._abi import STATE_DTYPE

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LIBRARY = os.path.join(HERE, "_ref", "libpumas_ref.so")
MDF = {
    "standard_c005": (os.path.join(REPO, "pumas_amd", "export",
                                   "materials_standard.xml"), 0.05),
    "iron_c001": (os.path.join(REPO, "pumas_amd", "export",
                               "materials_iron.xml"), 0.01),
}

_libs = {}


def available():
    return os.path.exists(LIBRARY)


def library():
    """The oracle library."""
    if "ref" not in _libs:
        path = LIBRARY
        if not os.path.exists(path):
            raise FileNotFoundError(
                f"{path} is missing; build it with `make -C oracle`")
        lib = ctypes.CDLL(path)
        vp, P = ctypes.c_void_p, ctypes.POINTER
        lib.ref_physics_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p,
                                           ctypes.c_double, P(vp)]
        lib.ref_physics_create.restype = ctypes.c_int
        lib.ref_physics_destroy.argtypes = [vp]
        lib.ref_physics_destroy.restype = None
        lib.ref_physics_export.argtypes = [vp, ctypes.c_char_p]
        lib.ref_physics_export.restype = ctypes.c_int
        lib.ref_geometry_create.argtypes = [P(_abi.GeometryDesc)]
        lib.ref_geometry_create.restype = vp
        lib.ref_geometry_destroy.argtypes = [vp]
        lib.ref_geometry_destroy.restype = None
        lib.ref_transport.argtypes = [
            vp, P(_abi.GpuContext), vp, vp, ctypes.c_int64, ctypes.c_uint64,
            ctypes.c_uint64, vp, vp, vp, vp, vp, ctypes.c_int,
            P(ctypes.c_uint64)]
        lib.ref_transport.restype = ctypes.c_double
        lib.ref_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint32,
                                    ctypes.c_uint64, ctypes.c_uint32]
        lib.ref_uniform.restype = ctypes.c_double
        lib.ref_philox4x32_10.argtypes = [P(ctypes.c_uint32),
                                          P(ctypes.c_uint32),
                                          P(ctypes.c_uint32)]
        lib.ref_philox4x32_10.restype = None
        lib.ref_cel.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_double]
        lib.ref_cel.restype = ctypes.c_double
        lib.ref_dcs_evaluate.argtypes = [vp, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_double]
        lib.ref_dcs_evaluate.restype = ctypes.c_double
        lib.ref_dcs_exact.argtypes = [vp, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_double]
        lib.ref_dcs_exact.restype = ctypes.c_double
        lib.ref_dilog.argtypes = [ctypes.c_double]
        lib.ref_dilog.restype = ctypes.c_double
        lib.ref_polar.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp,
                                  ctypes.c_int64, ctypes.c_uint64, vp, vp]
        lib.ref_polar.restype = ctypes.c_int
        _libs["ref"] = lib
    return _libs["ref"]


class RefPhysics:
    """Reference physics created by pumas_physics_create from the MDF the
    PTAB file of the same name was exported from."""

    def __init__(self, name):
        mdf, cutoff = MDF[name]
        self.name = name
        self.lib = library()
        self._dedx = tempfile.mkdtemp(prefix="pumas_dedx_")
        self.handle = ctypes.c_void_p()
        rc = self.lib.ref_physics_create(mdf.encode(), self._dedx.encode(),
                                         cutoff, ctypes.byref(self.handle))
        if rc != 0:
            raise RuntimeError(f"pumas_physics_create failed ({rc})")

    def export(self, path):
        rc = self.lib.ref_physics_export(self.handle, str(path).encode())
        if rc != 0:
            raise RuntimeError("export failed")

    def cel(self, which, scheme, material, x):
        return self.lib.ref_cel(self.handle, which, scheme, material, x)

    def dcs_evaluate(self, process, element, K, q):
        return self.lib.ref_dcs_evaluate(self.handle, process, element, K, q)

    def dcs_exact(self, process, element, K, q):
        return self.lib.ref_dcs_exact(self.handle, process, element, K, q)

    def polar(self, process, element, ki, kf, seed=0):
        """Polar angles mu of DELs ki[i] -> kf[i] (sample i draws from the
        transport stream of history i) and the draws of each sample."""
        ki = np.ascontiguousarray(ki, np.float64)
        kf = np.ascontiguousarray(kf, np.float64)
        mu = np.zeros(len(ki))
        draws = np.zeros(len(ki), np.uint32)
        assert self.lib.ref_polar(self.handle, process, element,
                                  ki.ctypes.data, kf.ctypes.data, len(ki),
                                  seed, mu.ctypes.data,
                                  draws.ctypes.data) == 0
        return mu, draws

    def close(self):
        if self.handle:
            self.lib.ref_physics_destroy(self.handle)
            self.handle = ctypes.c_void_p()
        shutil.rmtree(self._dedx, ignore_errors=True)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RefGeometry:
    def __init__(self, description):
        self.description = description
        self.lib = library()
        self._desc = description.to_c()
        self.handle = self.lib.ref_geometry_create(ctypes.byref(self._desc))

    def __del__(self):
        try:
            self.lib.ref_geometry_destroy(self.handle)
        except Exception:
            pass


class RefResult:
    pass


def transport(physics, context, geometry, states, seed=0, first=0,
              threads=1):
    """Transport `states` (STATE_DTYPE, copied) with the reference."""
    states = np.array(states, dtype=STATE_DTYPE, copy=True)
    n = len(states)
    r = RefResult()
    r.states = states
    r.events = np.zeros(n, np.int32)
    r.media = np.zeros((n, 2), np.int32)
    r.steps = np.zeros(n, np.uint32)
    r.draws = np.zeros(n, np.uint32)
    r.rc = np.zeros(n, np.int32)
    total = ctypes.c_uint64()
    ctx = context.to_c()
    if geometry.lib is not physics.lib:
        raise ValueError("physics and geometry come from different oracles")
    r.seconds = physics.lib.ref_transport(
        physics.handle, ctypes.byref(ctx), geometry.handle,
        states.ctypes.data, n, seed, first, r.events.ctypes.data,
        r.media.ctypes.data, r.steps.ctypes.data, r.draws.ctypes.data,
        r.rc.ctypes.data, threads, ctypes.byref(total))
    r.total_steps = int(total.value)
    return r


def uniform(seed, stream, history, index):
    return library().ref_uniform(seed, stream, history, index)


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    library().ref_philox4x32_10(c, k, o)
    return tuple(o)


LIBRARY_FLUX = os.path.join(HERE, "_ref", "libflux_ref.so")


def flux(kind, cos_theta, kinetic, charge):
    """The reference's example flux models (examples/pumas/flux.c),
    kind 1 = flux_gaisser, 2 = flux_gccly, evaluated element-wise."""
    if "flux" not in _libs:
        lib = ctypes.CDLL(LIBRARY_FLUX)
        lib.ref_flux_batch.argtypes = [ctypes.c_int, ctypes.c_size_t] + \
            [ctypes.c_void_p] * 4
        lib.ref_flux_batch.restype = ctypes.c_int
        _libs["flux"] = lib
    c = np.ascontiguousarray(cos_theta, dtype=np.float64)
    k = np.ascontiguousarray(kinetic, dtype=np.float64)
    q = np.ascontiguousarray(charge, dtype=np.float64)
    c, k, q = np.broadcast_arrays(c, k, q)
    c, k, q = (np.ascontiguousarray(x) for x in (c, k, q))
    out = np.empty(c.shape)
    rc = _libs["flux"].ref_flux_batch(kind, c.size, c.ctypes.data,
                                      k.ctypes.data, q.ctypes.data,
                                      out.ctypes.data)
    assert rc == 0
    return out


FRAME_FIELDS = ("states", "media", "events", "history")


def transport_record(physics, context, geometry, states, seed=0,
                     first_history=0, period=1, capacity=1 << 20):
    """Transport on one thread with the reference's recorder (period as
    `pumas_recorder.period`); returns (final states, frames) with frames a
    dict of arrays: states, media (index, -1 = outside), events, history."""
    lib = library()
    f = lib.ref_transport_record
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64,
                  ctypes.c_uint64, ctypes.c_int, ctypes.c_int64] + \
        [ctypes.c_void_p] * 4
    f.restype = ctypes.c_int64
    s = np.ascontiguousarray(states, dtype=STATE_DTYPE).copy()
    fs = np.zeros(capacity, STATE_DTYPE)
    fm = np.zeros(capacity, np.int32)
    fe = np.zeros(capacity, np.int32)
    fh = np.zeros(capacity, np.uint64)
    ctx = context.to_c()
    count = f(physics.handle, ctypes.byref(ctx), geometry.handle,
              s.ctypes.data, len(s), seed, first_history, period, capacity,
              fs.ctypes.data, fm.ctypes.data, fe.ctypes.data, fh.ctypes.data)
    k = min(count, capacity)
    return s, {"states": fs[:k], "media": fm[:k], "events": fe[:k],
               "history": fh[:k], "count": count}
