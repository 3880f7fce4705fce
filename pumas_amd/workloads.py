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
import dataclasses
import math

import numpy as np

from . import (Context, Event, GeometryDescription, GeometryKind, Medium,
               Mode, STATE_DTYPE, Step)

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


def uniforms(seed, stream, history, count, chunk=1 << 20):
    """Draws 0..count-1 of each history id: the same values as `uniform`,
    with one Philox block per pair of draws, over chunks of ids."""
    history = np.asarray(history, np.uint64)
    out = [np.empty(history.shape, np.float64) for _ in range(count)]
    for a in range(0, history.size, chunk):
        h = history[a:a + chunk]
        lo = (h & _MASK32).astype(np.uint32)
        hi = (h >> np.uint64(32)).astype(np.uint32)
        for b in range((count + 1) // 2):
            x0, x1, x2, x3 = philox4x32_10(
                np.full(h.shape, b, np.uint32),
                np.full(h.shape, stream, np.uint32), lo, hi,
                seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
            for k, (xl, xh) in enumerate(((x0, x1), (x2, x3))):
                if 2 * b + k >= count:
                    break
                w = (xh.astype(np.uint64) << np.uint64(32)) | \
                    xl.astype(np.uint64)
                out[2 * b + k][a:a + chunk] = \
                    ((w >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
    return out


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
    u0, u1, u2 = uniforms(seed, 0, ids, 3)
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
    u0, u1, u2 = uniforms(seed, 0, ids, 3)
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


# ---------------------------------------------------------------------------
# Non uniform media: standard rock under an exponential atmosphere, the
# setting of examples/pumas/geometry.c (a flat rock layer and an atmosphere
# whose locals callback returns a density falling with altitude and a
# proposed step).
# ---------------------------------------------------------------------------
ATMOSPHERE_Z = (0.0, 200.0, 2000.0)


def atmosphere_geometry(step_type=Step.RAW, scale=500.0, cos_min=0.05):
    """Standard rock (2650 kg/m^3) in z in [0, 200) m under air in
    [200, 2000) m with density 1.205 exp(-(z - 200) / scale) kg/m^3 and the
    proposed step scale / max(|u_z|, cos_min)."""
    air = Medium.exponential(AIR, 1.205, scale, z_ref=200.0, cos_min=cos_min)
    return GeometryDescription.slabs(ATMOSPHERE_Z, [(ROCK, 2650.0), air],
                                     step_type=step_type)


def atmosphere_states(n, seed=0, first=0, kmin=1.0, kmax=1e3, cmin=0.3,
                      cmax=1.0, forward=True):
    """Forward: initial states at the top of the air (z = 1999 m) going
    down; backward: final states at the bottom of the rock (z = 1 mm) going
    down, weight 1/pdf. K log-uniform in [kmin, kmax] GeV, cos(zenith)
    uniform in [cmin, cmax], charge -1."""
    s = (slabs_forward_states if forward else slabs_backward_states)(
        n, seed=seed, first=first, kmin=kmin, kmax=kmax, cmin=cmin,
        cmax=cmax)
    if forward:
        s["position"][:, 2] = 1999.0
    return s


def exponential_infinite_geometry(scale=100.0, cos_min=1.0):
    """One infinite water medium of density 1000 exp(-z / scale) kg/m^3
    (no boundary: every step is set by the physics and the locals)."""
    return GeometryDescription(
        GeometryKind.INFINITE,
        [Medium.exponential(WATER, 1000.0, scale, cos_min=cos_min)])


# ---------------------------------------------------------------------------
# Terrain: heightmap of seeded fractal value noise, backward muography.
# ---------------------------------------------------------------------------
def value_noise_2d(nx, ny, seed, octaves=6, base_cells=4, gain=0.5):
    """Fractal (fBm) value noise on an [ny, nx] node grid, in [0, 1]: sum of
    octaves of bilinearly interpolated random lattices drawn from the
    Philox stream (stream 2, history = octave)."""
    out = np.zeros((ny, nx))
    amp, total = 1.0, 0.0
    for o in range(octaves):
        cells = base_cells * 2 ** o
        ids = np.arange((cells + 1) ** 2, dtype=np.uint64) + np.uint64(
            o << 32)
        lattice = uniform(seed, 2, ids, 0).reshape(cells + 1, cells + 1)
        fx = np.linspace(0, cells, nx, endpoint=True)
        fy = np.linspace(0, cells, ny, endpoint=True)
        ix = np.minimum(fx.astype(int), cells - 1)
        iy = np.minimum(fy.astype(int), cells - 1)
        hx = (fx - ix)[None, :]
        hy = (fy - iy)[:, None]
        l00 = lattice[iy[:, None], ix[None, :]]
        l10 = lattice[iy[:, None], ix[None, :] + 1]
        l01 = lattice[iy[:, None] + 1, ix[None, :]]
        l11 = lattice[iy[:, None] + 1, ix[None, :] + 1]
        out += amp * ((l00 * (1 - hx) + l10 * hx) * (1 - hy) +
                      (l01 * (1 - hx) + l11 * hx) * hy)
        total += amp
        amp *= gain
    out /= total
    return (out - out.min()) / (out.max() - out.min())


TERRAIN_Z0 = -500.0
TERRAIN_ZTOP = 1000.0
TERRAIN_DETECTOR_Z = -100.0


def terrain_geometry(n=1024, spacing=5.0, relief=800.0, seed=0,
                     rho_min=2000.0, rho_max=2900.0, air=(AIR, 0.0)):
    """Heightmap of n x n nodes at `spacing` m with heights in [0, relief] m
    and one standard rock column per cell with a density uniform in
    [rho_min, rho_max] kg/m^3; air above, to TERRAIN_ZTOP (`air`: a Medium,
    or a (material, density) pair)."""
    h = relief * value_noise_2d(n, n, seed)
    ids = np.arange((n - 1) ** 2, dtype=np.uint64)
    rho = rho_min + (rho_max - rho_min) * uniform(seed, 3, ids, 0)
    half = 0.5 * (n - 1) * spacing
    return GeometryDescription.terrain(
        h.astype(np.float32), rho.reshape(n - 1, n - 1).astype(np.float32),
        x0=-half, y0=-half, dx=spacing, dy=spacing, z0=TERRAIN_Z0,
        z_top=TERRAIN_ZTOP, air=air, ground_material=ROCK)


def terrain_states(n, extent, seed=0, first=0, kmin=1e-2, kmax=1e2,
                   cmin=0.5, cmax=1.0):
    """Final states at detectors under the terrain: positions uniform in
    the central half of the map at TERRAIN_DETECTOR_Z, momentum going down
    in a cone cos(zenith) in [cmin, cmax], K log-uniform, weight 1/pdf."""
    ids = _ids(n, first)
    u = uniforms(seed, 0, ids, 5)
    rk = math.log(kmax / kmin)
    s = np.zeros(n, STATE_DTYPE)
    s["charge"] = -1.0
    s["energy"] = kmin * np.exp(rk * u[0])
    s["weight"] = s["energy"] * rk
    ct = cmin + (cmax - cmin) * u[1]
    st = np.sqrt(1.0 - ct * ct)
    phi = 2.0 * math.pi * u[2]
    s["direction"][:, 0] = st * np.cos(phi)
    s["direction"][:, 1] = st * np.sin(phi)
    s["direction"][:, 2] = -ct
    s["position"][:, 0] = 0.5 * extent * (u[3] - 0.5)
    s["position"][:, 1] = 0.5 * extent * (u[4] - 0.5)
    s["position"][:, 2] = TERRAIN_DETECTOR_Z
    return s


def terrain_context():
    return Context(energy_loss=Mode.STRAGGLED, decay=Mode.WEIGHTED,
                   direction=Mode.BACKWARD, scattering=Mode.MIXED,
                   event=Event.NONE)


# ---------------------------------------------------------------------------
# Voxels: rock / water / air / iron by 3-D noise thresholds.
# ---------------------------------------------------------------------------
def value_noise_3d(n, seed, octaves=4, base_cells=2):
    out = np.zeros((n, n, n))
    amp, total = 1.0, 0.0
    for o in range(octaves):
        cells = base_cells * 2 ** o
        ids = np.arange((cells + 1) ** 3, dtype=np.uint64) + np.uint64(
            o << 32)
        lat = uniform(seed, 4, ids, 0).reshape((cells + 1,) * 3)
        f = np.linspace(0, cells, n, endpoint=True)
        i = np.minimum(f.astype(int), cells - 1)
        h = f - i
        hz, hy, hx = h[:, None, None], h[None, :, None], h[None, None, :]
        iz, iy, ix = i[:, None, None], i[None, :, None], i[None, None, :]
        acc = 0.0
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    w = ((hz if dz else 1 - hz) * (hy if dy else 1 - hy) *
                         (hx if dx else 1 - hx))
                    acc = acc + w * lat[iz + dz, iy + dy, ix + dx]
        out += amp * acc
        total += amp
        amp *= 0.5
    return out / total


VOXEL_MATERIALS = ((ROCK, 0.0), (WATER, 0.0), (AIR, 0.0), (IRON, 0.0))
VOXEL_DEFAULT_DENSITY = (2650.0, 1000.0, 1.205, 7874.0)


def voxel_geometry(n=512, size=2.0, seed=0, sigma=0.1):
    """n^3 voxels of `size` m; material classes by thresholds on 3-D value
    noise (air < 0.35 < water < 0.45 < rock < 0.7 < iron) and a lognormal
    density scale of mean one and log-sigma `sigma` per voxel."""
    noise = value_noise_3d(n, seed)
    label = np.full(noise.shape, 0, np.uint8)      # rock
    label[noise < 0.45] = 1                        # water
    label[noise < 0.35] = 2                        # air
    label[noise >= 0.7] = 3                        # iron
    ids = np.arange(n ** 3, dtype=np.uint64)
    u1 = uniform(seed, 5, ids, 0)
    u2 = uniform(seed, 5, ids, 1)
    g = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)
    scale = np.exp(sigma * g - 0.5 * sigma * sigma)
    rho = np.asarray(VOXEL_DEFAULT_DENSITY)[label] * scale.reshape(label.shape)
    half = 0.5 * n * size
    return GeometryDescription.voxels(
        label, rho.astype(np.float32), VOXEL_MATERIALS, x0=-half, y0=-half,
        z0=-half, dx=size, dy=size, dz=size)


def voxel_states(n, extent, seed=0, first=0, kmin=1e-1, kmax=1e3,
                 cmin=0.5, cmax=1.0):
    """Final states at detectors near the bottom of the voxel cube, momentum
    going down in a cone, K log-uniform in [kmin, kmax] GeV."""
    ids = _ids(n, first)
    u = uniforms(seed, 0, ids, 5)
    rk = math.log(kmax / kmin)
    s = np.zeros(n, STATE_DTYPE)
    s["charge"] = -1.0
    s["energy"] = kmin * np.exp(rk * u[0])
    s["weight"] = s["energy"] * rk
    ct = cmin + (cmax - cmin) * u[1]
    st = np.sqrt(1.0 - ct * ct)
    phi = 2.0 * math.pi * u[2]
    s["direction"][:, 0] = st * np.cos(phi)
    s["direction"][:, 1] = st * np.sin(phi)
    s["direction"][:, 2] = -ct
    s["position"][:, 0] = 0.25 * extent * (u[3] - 0.5)
    s["position"][:, 1] = 0.25 * extent * (u[4] - 0.5)
    s["position"][:, 2] = -0.45 * extent
    return s


def voxel_context():
    return Context(energy_loss=Mode.STRAGGLED, decay=Mode.WEIGHTED,
                   direction=Mode.BACKWARD, scattering=Mode.MIXED,
                   event=Event.NONE)


# ---------------------------------------------------------------------------
# Serving: Poisson arrivals of transport requests of uneven sizes.
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class ServingRequest:
    arrival: float      # s after the start of the run
    geometry: str       # "slabs" (forward) or "terrain" (backward)
    energy_loss: int    # Mode.CSDA or Mode.MIXED
    n: int              # muons
    first: int          # first history id (Philox counter)


def serving_requests(duration, rate=200.0, seed=0, nmin=1e3, nmax=1e6):
    """Requests arriving over `duration` s at `rate` per s (exponential
    gaps), each with round(10^U[log nmin, log nmax]) muons, geometry slabs
    or terrain and energy loss CSDA or MIXED with probability 1/2 each.
    History ids are disjoint across requests."""
    out, t, k, first = [], 0.0, 0, 0
    while True:
        u = uniform(seed, 4, np.array([k], np.uint64), 0)[0]
        t += -math.log(u) / rate
        if t >= duration:
            break
        v = [uniform(seed, 4, np.array([k], np.uint64), j)[0]
             for j in (1, 2, 3)]
        n = int(round(10 ** (math.log10(nmin) +
                            (math.log10(nmax) - math.log10(nmin)) * v[0])))
        out.append(ServingRequest(
            t, "slabs" if v[1] < 0.5 else "terrain",
            Mode.CSDA if v[2] < 0.5 else Mode.MIXED, n, first))
        first += n
        k += 1
    return out


def serving_inputs(req, terrain_extent, seed=0):
    """(context, initial states) of a serving request."""
    if req.geometry == "slabs":
        return (slabs_forward_context(energy_loss=req.energy_loss),
                slabs_forward_states(req.n, seed=seed, first=req.first))
    ctx = terrain_context().replace(energy_loss=req.energy_loss)
    return ctx, terrain_states(req.n, terrain_extent, seed=seed,
                               first=req.first)


# ---------------------------------------------------------------------------
# Full-size inputs of the terrain and voxel workloads.
# ---------------------------------------------------------------------------
def terrain_sightlines(n_lines, extent, seed=0, nmin=1e3, nmax=1e6,
                       kmin=1e-2, kmax=1e2, cmin=0.5, cmax=1.0):
    """Batched muography: `n_lines` lines of sight, each a detector under
    the terrain (uniform in the central half of the map) looking up along a
    direction in the cone cos(zenith) in [cmin, cmax], with
    round(10^U[log nmin, log nmax]) events of final kinetic energy
    log-uniform in [kmin, kmax] GeV. Returns (states, line index per state);
    history ids are consecutive."""
    lines = terrain_lines(n_lines, extent, seed, nmin, nmax, cmin, cmax)
    counts = lines["count"]
    n = int(counts.sum())
    line = np.repeat(np.arange(n_lines), counts)
    ids = _ids(n, 0)
    u0 = uniform(seed, 0, ids, 0)
    rk = math.log(kmax / kmin)
    s = np.zeros(n, STATE_DTYPE)
    s["charge"] = -1.0
    s["energy"] = kmin * np.exp(rk * u0)
    s["weight"] = s["energy"] * rk
    s["direction"] = lines["direction"][line]
    s["position"] = lines["position"][line]
    return s, line


def terrain_lines(n_lines, extent, seed=0, nmin=1e3, nmax=1e6, cmin=0.5,
                  cmax=1.0):
    """The lines of sight of terrain_sightlines: per line, its event count,
    detector position and (momentum) direction."""
    lines = np.arange(n_lines, dtype=np.uint64)
    v = [uniform(seed, 6, lines, k) for k in range(5)]
    counts = np.rint(10 ** (math.log10(nmin) + (math.log10(nmax) -
                                                  math.log10(nmin)) * v[0]))
    ct = cmin + (cmax - cmin) * v[1]
    st = np.sqrt(1.0 - ct * ct)
    phi = 2.0 * math.pi * v[2]
    direction = np.stack([st * np.cos(phi), st * np.sin(phi), -ct], axis=1)
    position = np.stack([0.5 * extent * (v[3] - 0.5),
                         0.5 * extent * (v[4] - 0.5),
                         np.full(n_lines, TERRAIN_DETECTOR_Z)], axis=1)
    return {"count": counts.astype(np.int64), "direction": direction,
            "position": position}


def voxel_geometry_large(n=512, size=2.0, seed=0, sigma=0.1):
    """voxel_geometry at full size (512^3), generated with torch's seeded CPU
    generator (vectorised, reproducible on any host): the same noise
    construction (4 octaves of trilinear value noise), thresholds and
    lognormal density scale, other random numbers. The trilinear
    interpolation runs axis by axis (x, then y, then z), so that only the
    last pass touches n^3 values. For throughput runs and the full-size
    tests; the small parity geometries use voxel_geometry."""
    import torch
    g = torch.Generator()
    g.manual_seed(seed)
    out = torch.zeros((n, n, n), dtype=torch.float64)
    tmp = torch.empty_like(out)  # reused: large fresh allocations are slow
    amp, total = 1.0, 0.0
    f_base = torch.linspace(0, 1, n, dtype=torch.float64)
    for o in range(4):
        cells = 2 * 2 ** o
        lat = torch.rand((cells + 1,) * 3, generator=g, dtype=torch.float64)
        f = f_base * cells
        i = torch.clamp(f.long(), max=cells - 1)
        h = f - i
        # lat[z, y, x] -> interpolate x, then y (small arrays), then z.
        ax = lat[:, :, i] * (1 - h) + lat[:, :, i + 1] * h
        ay = ax[:, i, :] * (1 - h)[:, None] + ax[:, i + 1, :] * h[:, None]
        for k, w in ((i, amp * (1 - h)), (i + 1, amp * h)):
            torch.index_select(ay, 0, k, out=tmp)
            tmp.mul_(w[:, None, None])
            out.add_(tmp)
        total += amp
        amp *= 0.5
        del ax, ay
    del tmp
    out /= total
    label = torch.zeros((n, n, n), dtype=torch.uint8)
    label[out < 0.45] = 1
    label[out < 0.35] = 2
    label[out >= 0.7] = 3
    del out
    scale = torch.exp(sigma * torch.randn((n, n, n), generator=g) -
                      0.5 * sigma * sigma).numpy()
    label = label.numpy()
    rho = np.asarray(VOXEL_DEFAULT_DENSITY, np.float32)[label]
    rho *= scale
    half = 0.5 * n * size
    return GeometryDescription.voxels(
        label, rho, VOXEL_MATERIALS, x0=-half, y0=-half,
        z0=-half, dx=size, dy=size, dz=size)
