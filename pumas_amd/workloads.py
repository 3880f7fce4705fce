"""Synthetic workloads: initial states, geometries and transport contexts.

Initial states are drawn from the Philox4x32-10 stream 0 (the transport
itself draws from stream 1), keyed by the seed and the history id, so that a
batch is reproducible independently of how it is split:

    key = (seed low, seed high), counter = (draw // 2, 0, id low, id high)
    draw 2k -> words (x0, x1), draw 2k+1 -> words (x2, x3)
    u = ((x1 << 32 | x0) >> 11 + 0.5) * 2^-53

Workload sizes follow the configurations the library is benchmarked on:
  - `csda_rock`: backward CSDA through 1 km of standard rock (the closed
    form path of the reference, examples/pumas/straight.c:133-179);
  - `slabs_forward`: forward straggled transport with multiple scattering
    through 100 m of air, 500 m of standard rock and 200 m of water.
"""
import math

import numpy as np

from . import Context, Event, GeometryDescription, Mode, STATE_DTYPE, Step

MUON_MASS = 0.10565839  # GeV, src/pumas.c:148

# Material indices of the PTAB physics files.
ROCK, WATER, AIR, IRON = 0, 1, 2, 3

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
_MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10 over uint32 arrays (Salmon et al. 2011)."""
    c0, c1, c2, c3 = (np.asarray(x, np.uint32) for x in (c0, c1, c2, c3))
    k0 = np.uint32(k0)
    k1 = np.uint32(k1)
    for _ in range(10):
        p0 = _M0 * c0.astype(np.uint64)
        p1 = _M1 * c2.astype(np.uint64)
        n0 = (p1 >> np.uint64(32)).astype(np.uint32) ^ c1 ^ k0
        n2 = (p0 >> np.uint64(32)).astype(np.uint32) ^ c3 ^ k1
        c1 = (p1 & _MASK32).astype(np.uint32)
        c3 = (p0 & _MASK32).astype(np.uint32)
        c0, c2 = n0, n2
        k0 = np.uint32((int(k0) + int(_W0)) & 0xFFFFFFFF)
        k1 = np.uint32((int(k1) + int(_W1)) & 0xFFFFFFFF)
    return c0, c1, c2, c3


def uniform(seed, stream, history, index):
    """Uniform draw `index` of each history id in `history` (uint64)."""
    history = np.asarray(history, np.uint64)
    lo = (history & _MASK32).astype(np.uint32)
    hi = (history >> np.uint64(32)).astype(np.uint32)
    n = history.shape
    c0 = np.full(n, index >> 1, np.uint32)
    c1 = np.full(n, stream, np.uint32)
    x0, x1, x2, x3 = philox4x32_10(c0, c1, lo, hi, seed & 0xFFFFFFFF,
                                   (seed >> 32) & 0xFFFFFFFF)
    if index & 1:
        w = (x3.astype(np.uint64) << np.uint64(32)) | x2.astype(np.uint64)
    else:
        w = (x1.astype(np.uint64) << np.uint64(32)) | x0.astype(np.uint64)
    return ((w >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53


def _ids(n, first):
    return np.arange(first, first + n, dtype=np.uint64)


# ---------------------------------------------------------------------------
# Backward CSDA through 1 km of standard rock.
# ---------------------------------------------------------------------------
def csda_rock_states(n, seed=0, first=0, kmin=1e-2, kmax=1e2):
    """Final states at the detector: K log-uniform in [kmin, kmax] GeV,
    weight 1/pdf, charge -1, momentum pointing down (0, 0, -1)."""
    ids = _ids(n, first)
    rk = math.log(kmax / kmin)
    u = uniform(seed, 0, ids, 0)
    s = np.zeros(n, STATE_DTYPE)
    s["charge"] = -1.0
    s["energy"] = kmin * np.exp(rk * u)
    s["weight"] = s["energy"] * rk
    s["direction"][:, 2] = -1.0
    return s


def csda_rock_context(distance=1e3):
    return Context(energy_loss=Mode.CSDA, decay=Mode.WEIGHTED,
                   direction=Mode.BACKWARD, scattering=Mode.DISABLED,
                   event=Event.LIMIT_DISTANCE, limit_distance=distance)


def csda_rock_geometry():
    return GeometryDescription.infinite(ROCK)


def gaisser_flux(kinetic, cos_theta=1.0):
    """Gaisser form sea level flux, 0.14 E^-2.7 [1/(1 + 1.1 E cos/115) +
    0.054/(1 + 1.1 E cos/850)] per GeV cm^2 s sr, E the total energy."""
    e = np.asarray(kinetic) + MUON_MASS
    ec = 1.1 * e * cos_theta
    return 0.14 * e ** -2.7 * (1.0 / (1.0 + ec / 115.0) +
                               0.054 / (1.0 + ec / 850.0))


def integrated_flux(states):
    """Monte Carlo estimate of the flux at the detector from backward
    weighted states (mean of w * flux(K_surface))."""
    cos_theta = -states["direction"][:, 2]
    return float(np.mean(states["weight"] *
                         gaisser_flux(states["energy"], cos_theta)))


# ---------------------------------------------------------------------------
# Forward straggled transport through air / rock / water slabs.
# ---------------------------------------------------------------------------
SLAB_Z = (0.0, 200.0, 700.0, 800.0)
SLAB_MEDIA = ((WATER, 0.0), (ROCK, 0.0), (AIR, 0.0))
SLAB_START_Z = 800.0 - 1e-3


def slabs_forward_states(n, seed=0, first=0, kmin=10.0, kmax=1e4,
                         cmin=0.5, cmax=1.0):
    """Initial states at the top of the stack: K log-uniform in
    [kmin, kmax] GeV, cos(zenith) uniform in [cmin, cmax], azimuth uniform,
    going down, charge -1, weight 1."""
    ids = _ids(n, first)
    u0 = uniform(seed, 0, ids, 0)
    u1 = uniform(seed, 0, ids, 1)
    u2 = uniform(seed, 0, ids, 2)
    s = np.zeros(n, STATE_DTYPE)
    s["charge"] = -1.0
    s["energy"] = kmin * np.exp(math.log(kmax / kmin) * u0)
    s["weight"] = 1.0
    ct = cmin + (cmax - cmin) * u1
    st = np.sqrt(1.0 - ct * ct)
    phi = 2.0 * math.pi * u2
    s["direction"][:, 0] = st * np.cos(phi)
    s["direction"][:, 1] = st * np.sin(phi)
    s["direction"][:, 2] = -ct
    s["position"][:, 2] = SLAB_START_Z
    return s


def slabs_forward_context(energy_loss=Mode.STRAGGLED):
    return Context(energy_loss=energy_loss, decay=Mode.WEIGHTED,
                   direction=Mode.FORWARD, scattering=Mode.MIXED,
                   event=Event.NONE)


def slabs_geometry(step_type=Step.RAW):
    return GeometryDescription.slabs(SLAB_Z, SLAB_MEDIA, step_type=step_type)


def slabs_backward_states(n, seed=0, first=0, kmin=1e-1, kmax=1e3,
                          cmin=0.5, cmax=1.0):
    """Final states at a detector at the bottom of the stack: K log-uniform
    in [kmin, kmax] GeV, momentum going down with cos(zenith) uniform in
    [cmin, cmax], weight 1/pdf, charge -1."""
    ids = _ids(n, first)
    u0 = uniform(seed, 0, ids, 0)
    u1 = uniform(seed, 0, ids, 1)
    u2 = uniform(seed, 0, ids, 2)
    rk = math.log(kmax / kmin)
    s = np.zeros(n, STATE_DTYPE)
    s["charge"] = -1.0
    s["energy"] = kmin * np.exp(rk * u0)
    s["weight"] = s["energy"] * rk
    ct = cmin + (cmax - cmin) * u1
    st = np.sqrt(1.0 - ct * ct)
    phi = 2.0 * math.pi * u2
    s["direction"][:, 0] = st * np.cos(phi)
    s["direction"][:, 1] = st * np.sin(phi)
    s["direction"][:, 2] = -ct
    s["position"][:, 2] = 1e-3
    return s


def slabs_backward_context(energy_loss=Mode.STRAGGLED):
    return Context(energy_loss=energy_loss, decay=Mode.WEIGHTED,
                   direction=Mode.BACKWARD, scattering=Mode.MIXED,
                   event=Event.NONE)
