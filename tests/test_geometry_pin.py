"""The device geometries pinned on the CPU by an independent description.

pumas_amd/csrc/geometry.h (the medium callbacks of both the device engine
and the oracle harness) is built for the host (tests/geometry_check.c) and
compared with a classification of the same geometry descriptions written
here from their definitions with numpy and scipy:

* slabs: medium k for z in [z_k, z_k+1), outside elsewhere;
* terrain (the full 1024 x 1024 heightmap at 5 m): outside beyond the grid
  or [z0, z_top), rock column 1 + iy (nx - 1) + ix below the bilinear
  surface (scipy's linear grid interpolator), air above;
* voxels (the full 512^3 grid at 2 m): medium 1 + (iz ny + iy) nx + ix.

Media agree at 10^6 random points away from the surfaces (|z - h| > 1e-6
m, 1e-9 m from walls). Raw steps (slabs, voxels) are checked by marching
along the ray: the medium is the start one up to 1e-7 m before the proposed
step and another one at it. Checked steps (terrain) never cross a cell
wall before they end, and end on the bilinear surface when they cross it
(brute force root search along 10^5 rays).
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import REPO

from pumas_amd import _abi
from pumas_amd import workloads as W

BUILD = os.path.join(REPO, "tests", "_build")
LIB = os.path.join(BUILD, "libgeometry_check.so")


@pytest.fixture(scope="module")
def geo():
    os.makedirs(BUILD, exist_ok=True)
    src = os.path.join(REPO, "tests", "geometry_check.c")
    hdr = os.path.join(REPO, "pumas_amd", "csrc", "geometry.h")
    if (not os.path.exists(LIB)) or any(os.path.getmtime(f) >
                                        os.path.getmtime(LIB)
                                        for f in (src, hdr)):
        subprocess.check_call(
            ["gcc", "-O2", "-std=c99", "-fPIC", "-shared",
             "-ffp-contract=off", "-I", os.path.join(REPO, "include"), "-I",
             os.path.join(REPO, "pumas_amd", "csrc"), "-o",
             f"{LIB}.{os.getpid()}.tmp", src])
        os.replace(f"{LIB}.{os.getpid()}.tmp", LIB)
    lib = ctypes.CDLL(LIB)
    lib.geo_check_locate.argtypes = [ctypes.POINTER(_abi.GeometryDesc)] + \
        [ctypes.c_void_p] * 2 + [ctypes.c_double, ctypes.c_int64] + \
        [ctypes.c_void_p] * 3
    lib.geo_check_locate.restype = ctypes.c_int
    return lib


def locate(lib, desc, pos, dirs, sgn=1.0):
    pos = np.ascontiguousarray(pos, np.float64)
    dirs = np.ascontiguousarray(dirs, np.float64)
    n = len(pos)
    medium = np.zeros(n, np.int32)
    step = np.zeros(n)
    st = np.zeros(n, np.int32)
    d = desc.to_c()
    assert lib.geo_check_locate(ctypes.byref(d), pos.ctypes.data,
                                dirs.ctypes.data, sgn, n, medium.ctypes.data,
                                step.ctypes.data, st.ctypes.data) == 0
    return medium, step, st


def random_dirs(rng, n, min_axis=0.0):
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    return u


def slab_media(desc, z):
    zb = desc.z
    k = np.searchsorted(zb, z, side="right") - 1
    return np.where((z >= zb[0]) & (z < zb[-1]), k, -1)


def test_slabs(geo):
    rng = np.random.default_rng(1)
    desc = W.slabs_geometry()
    n = 200_000
    pos = np.zeros((n, 3))
    pos[:, :2] = rng.uniform(-1e3, 1e3, (n, 2))
    pos[:, 2] = rng.uniform(-50.0, 850.0, n)
    dirs = random_dirs(rng, n)
    m, s, st = locate(geo, desc, pos, dirs)
    np.testing.assert_array_equal(m, slab_media(desc, pos[:, 2]))
    inside = (m >= 0) & (np.abs(dirs[:, 2]) > 1e-3)
    for frac in np.linspace(0.0, 1.0, 33):
        back = np.maximum(s - 2e-7, 0.0) * frac
        z = pos[:, 2] + back * dirs[:, 2]
        assert (slab_media(desc, z)[inside] == m[inside]).all()
    z_end = pos[:, 2] + s * dirs[:, 2]
    assert (slab_media(desc, z_end)[inside] != m[inside]).all()


@pytest.fixture(scope="module")
def terrain():
    return W.terrain_geometry()          # 1024 x 1024 nodes at 5 m


def terrain_media(desc, pos):
    from scipy.interpolate import RegularGridInterpolator
    g = desc.grid
    nx, ny = g["nx"], g["ny"]
    xs = g["x0"] + g["dx"] * np.arange(nx)
    ys = g["y0"] + g["dy"] * np.arange(ny)
    surface = RegularGridInterpolator((ys, xs), desc.height.astype(np.float64),
                                      method="linear")
    fx = (pos[:, 0] - g["x0"]) / g["dx"]
    fy = (pos[:, 1] - g["y0"]) / g["dy"]
    ix, iy = np.floor(fx).astype(np.int64), np.floor(fy).astype(np.int64)
    inside = (ix >= 0) & (iy >= 0) & (ix <= nx - 2) & (iy <= ny - 2) & \
        (pos[:, 2] >= g["z0"]) & (pos[:, 2] < desc.z_top)
    h = np.full(len(pos), np.nan)
    h[inside] = surface(np.stack([pos[inside, 1], pos[inside, 0]], axis=1))
    medium = np.where(pos[:, 2] < h, 1 + iy * (nx - 1) + ix, 0)
    return np.where(inside, medium, -1), h, fx, fy


def test_terrain_full_grid(geo, terrain):
    rng = np.random.default_rng(2)
    g = terrain.grid
    n = 1_000_000
    half = 0.5 * (g["nx"] - 1) * g["dx"]
    pos = np.stack([rng.uniform(-1.05 * half, 1.05 * half, n),
                    rng.uniform(-1.05 * half, 1.05 * half, n),
                    rng.uniform(g["z0"] - 50.0, terrain.z_top + 50.0, n)], 1)
    dirs = random_dirs(rng, n)
    m, s, st = locate(geo, terrain, pos, dirs)
    want, h, fx, fy = terrain_media(terrain, pos)
    clear = ~(np.abs(pos[:, 2] - h) < 1e-6) & \
        (np.abs(fx - np.round(fx)) > 1e-9) & (np.abs(fy - np.round(fy)) > 1e-9)
    assert clear.mean() > 0.99
    np.testing.assert_array_equal(m[clear], want[clear])
    assert (m == 0).any() and (m > 0).any() and (m < 0).any()
    # A checked step stays in its cell of the grid (the transport checks
    # the medium at its end).
    inside = (m >= 0) & clear
    ix0, iy0 = np.floor(fx), np.floor(fy)
    for frac in np.linspace(0.0, 1.0, 17):
        p = pos + np.maximum(s * frac - 1e-9, 0.0)[:, None] * dirs
        jx = np.floor((p[:, 0] - g["x0"]) / g["dx"])
        jy = np.floor((p[:, 1] - g["y0"]) / g["dy"])
        ok = (jx == ix0) & (jy == iy0)
        assert ok[inside].all()


def test_terrain_surface_crossings(geo, terrain):
    """The proposed step ends on the bilinear surface when it crosses it
    in the cell: brute force search of the first sign change of z - h along
    each ray (129 samples up to the cell wall or the top / bottom plane,
    then bisection on scipy's interpolated surface). No step passes a
    crossing, and a step shorter than the wall and plane distances ends on
    the surface."""
    from scipy.interpolate import RegularGridInterpolator
    rng = np.random.default_rng(4)
    g = terrain.grid
    n = 100_000
    half = 0.5 * (g["nx"] - 1) * g["dx"]
    xs = g["x0"] + g["dx"] * np.arange(g["nx"])
    ys = g["y0"] + g["dy"] * np.arange(g["ny"])
    surface = RegularGridInterpolator(
        (ys, xs), terrain.height.astype(np.float64), method="linear")
    # Points within 3 m of the surface, going downwards from the air and
    # upwards from the rock, so that many rays cross it in their cell.
    pos = np.stack([rng.uniform(-0.9 * half, 0.9 * half, n),
                    rng.uniform(-0.9 * half, 0.9 * half, n),
                    np.zeros(n)], 1)
    h0 = surface(pos[:, 1::-1])
    pos[:, 2] = h0 + rng.uniform(-3.0, 3.0, n)
    dirs = random_dirs(rng, n)
    dirs[:, 2] = -np.sign(pos[:, 2] - h0) * np.abs(dirs[:, 2])
    m, s, st = locate(geo, terrain, pos, dirs)
    assert (st == 0).all() and (m >= 0).all()

    def f(t):
        p = pos + t[:, None] * dirs
        return p[:, 2] - surface(p[:, 1::-1])

    # Distance to the cell walls and to the top (air) or bottom (rock).
    fx = (pos[:, 0] - g["x0"]) / g["dx"]
    fy = (pos[:, 1] - g["y0"]) / g["dy"]
    lim = np.full(n, np.inf)
    for k, fk, d0, dk in ((0, fx, g["x0"], g["dx"]), (1, fy, g["y0"], g["dy"])):
        lo = d0 + np.floor(fk) * dk
        with np.errstate(divide="ignore"):
            t = np.where(dirs[:, k] > 0, (lo + dk - pos[:, k]) / dirs[:, k],
                         (lo - pos[:, k]) / dirs[:, k])
        lim = np.minimum(lim, np.where(np.abs(dirs[:, k]) > 1e-12, t, np.inf))
    plane = np.where(m == 0, (terrain.z_top - pos[:, 2]) / dirs[:, 2],
                     (g["z0"] - pos[:, 2]) / dirs[:, 2])
    lim = np.minimum(lim, np.where(plane > 0, plane, np.inf))
    f0 = f(np.zeros(n))
    side = np.sign(f0)
    ok = np.abs(f0) > 1e-6
    # First sign change on a grid of the segment [0, lim].
    ts = np.linspace(0.0, 1.0, 129)[:, None] * lim[None, :]
    fs = np.stack([f(t) for t in ts])
    changed = np.sign(fs) != side[None, :]
    has = changed.any(axis=0)
    k = np.argmax(changed, axis=0)
    a = np.where(has, ts[np.maximum(k - 1, 0), np.arange(n)], 0.0)
    b = np.where(has, ts[k, np.arange(n)], 0.0)
    for _ in range(60):
        c = 0.5 * (a + b)
        same = np.sign(f(c)) == side
        a = np.where(same, c, a)
        b = np.where(same, b, c)
    s_brute = np.where(has, 0.5 * (a + b), np.inf)
    cross = ok & has
    assert cross.mean() > 0.3
    print(f"{int(cross.sum())} of {n} rays cross the surface in their cell")
    # No proposed step passes a crossing.
    assert (s[ok] <= s_brute[ok] + 1e-7).all()
    # Where the crossing comes before the walls and planes, the step ends on
    # it (a tangent touch without a sign change aside: f(s) ~ 0 there too).
    by_surface = ok & (s < lim - 1e-6)
    fe = f(np.where(by_surface, s, 0.0))
    assert (np.abs(fe[by_surface]) < 1e-7).all()
    np.testing.assert_allclose(s[cross & by_surface],
                               s_brute[cross & by_surface], rtol=0,
                               atol=1e-6)
    assert (cross & ~by_surface & (s_brute < lim - 1e-6)).sum() == 0


@pytest.fixture(scope="module")
def voxels():
    return W.voxel_geometry_large()      # 512^3 voxels of 2 m


def voxel_media(desc, pos):
    g = desc.grid
    f = [(pos[:, k] - g[o]) / g[d] for k, o, d in
         ((0, "x0", "dx"), (1, "y0", "dy"), (2, "z0", "dz"))]
    i = [np.floor(v).astype(np.int64) for v in f]
    inside = (i[0] >= 0) & (i[1] >= 0) & (i[2] >= 0) & (i[0] < g["nx"]) & \
        (i[1] < g["ny"]) & (i[2] < g["nz"])
    flat = (i[2] * g["ny"] + i[1]) * g["nx"] + i[0]
    return np.where(inside, 1 + flat, -1), f


def test_voxels_full_grid(geo, voxels):
    rng = np.random.default_rng(3)
    g = voxels.grid
    n = 1_000_000
    half = 0.5 * g["nx"] * g["dx"]
    pos = rng.uniform(-1.05 * half, 1.05 * half, (n, 3))
    dirs = random_dirs(rng, n)
    m, s, st = locate(geo, voxels, pos, dirs)
    want, f = voxel_media(voxels, pos)
    clear = np.all([np.abs(v - np.round(v)) > 1e-9 for v in f], axis=0)
    np.testing.assert_array_equal(m[clear], want[clear])
    labels = voxels.label.reshape(-1)
    assert len(np.unique(labels[want[want > 0] - 1])) == 4
    # Raw steps: the start voxel up to 1e-7 m before the step, another one
    # at it (rays not within 1e-6 of a voxel edge at the exit).
    inside = (m > 0) & clear
    for frac in np.linspace(0.0, 1.0, 17):
        p = pos + (np.maximum(s - 2e-7, 0.0) * frac)[:, None] * dirs
        assert (voxel_media(voxels, p)[0][inside] == m[inside]).all()
    end, fe = voxel_media(voxels, pos + s[:, None] * dirs)
    edge = np.sum([np.abs(v - np.round(v)) < 1e-6 for v in fe], axis=0) > 1
    ok = inside & ~edge
    assert ok.mean() > 0.8
    assert (end[ok] != m[ok]).all()
