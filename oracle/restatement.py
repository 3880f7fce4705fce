"""CPU restatement of the reference CSDA transport path (numpy).

Test infrastructure only: this module is a checker for the device engine and
is itself pinned (tests/test_oracle.py) against
  - the reference library compiled from its source (oracle/_ref), history
    by history, on the committed golden fixtures;
  - the known answers of the reference test suite for standard rock
    (tests/test-pumas.c:909-986: stopping power, straggling, CSDA range,
    inverse range, elastic path);
  - closed forms: PCHIP reproduces cubics exactly, the bracket search
    agrees with brute force, a forward CSDA history run to rest travels
    exactly its CSDA range, and the backward weight equals
    S(K_f) / S(K_i) exp(-dt / c tau).

It reads the PTAB tables directly (pumas_amd/csrc/ptab.h layout).
"""
import struct

import numpy as np

PTAB_MAGIC = b"PUMASTAB"
_HEADER = struct.Struct("<8sII16i8d")
_ENTRY = struct.Struct("<32sIIQQ")
_DTYPES = {0: "<f8", 1: "<f4", 2: "<i4"}

CSDA, MIXED = 0, 1
EPSILON_X = 3e-3  # src/pumas.c:108


class Tables:
    """PTAB sections as numpy arrays."""

    def __init__(self, path):
        data = open(path, "rb").read()
        head = _HEADER.unpack_from(data, 0)
        if head[0] != PTAB_MAGIC:
            raise ValueError("not a PTAB file")
        n_entries = head[2]
        imeta = head[3:19]
        fmeta = head[19:27]
        self.n_energies, self.n_materials = imeta[0], imeta[1]
        self.n_elements, self.n_components = imeta[3], imeta[4]
        self.n_table_dcs = imeta[6]
        self.ctau, self.mass, self.cutoff = fmeta[0], fmeta[1], fmeta[2]
        self.sections = {}
        off = _HEADER.size
        for _ in range(n_entries):
            name, dtype, _, count, offset = _ENTRY.unpack_from(data, off)
            off += _ENTRY.size
            name = name.rstrip(b"\0").decode()
            self.sections[name] = np.frombuffer(
                data, dtype=_DTYPES[dtype], count=count, offset=offset)
        self.K = self.sections["K"]

    def col2(self, name, scheme, material):
        """Column of a two-scheme table, src/pumas.c:4265-4360."""
        s = min(scheme, MIXED)
        nK, nM = self.n_energies, self.n_materials
        a = self.sections[name]
        return a[(s * nM + material) * nK:(s * nM + material + 1) * nK]

    def col1(self, name, material):
        nK = self.n_energies
        return self.sections[name][material * nK:(material + 1) * nK]

    def b_max(self, scheme, material):
        return self.sections["b_max"][min(scheme, MIXED) * self.n_materials +
                                      material]


def table_index(table, x):
    """Largest i with table[i] <= x (-1 below, n-1 at or above the last
    node): the result of table_index / table_bracket,
    src/pumas.c:4174-4233."""
    return np.searchsorted(table, x, side="right") - 1


def pchip(t, p0, p1, m0, m1):
    """Hermite cubic, math_pchip_interpolate, src/pumas.c:4118-4125."""
    c2 = -3 * (p0 - p1) - 2 * m0 - m1
    c3 = 2 * (p0 - p1) + m0 + m1
    return p0 + t * (m0 + t * (c2 + t * c3))


def interpolate(tx, ty, tm, x):
    """table_interpolate_pchip, src/pumas.c:4146-4157 (x inside table)."""
    i1 = table_index(tx, x)
    i2 = i1 + 1
    dX = tx[i2] - tx[i1]
    t = (x - tx[i1]) / dX
    m1 = tm[i1] * dX
    m2 = np.where(i2 > 1, tm[i2] * dX, m1)
    return pchip(t, ty[i1], ty[i2], m1, m2)


def _piecewise(x, low, high, inside):
    """Evaluate `inside` on the in-table points only."""
    x = np.asarray(x, np.float64)
    out = np.empty_like(x)
    lo = low(x)
    hi = high(x)
    mid = ~(lo[0] | hi[0])
    out[lo[0]] = lo[1][lo[0]] if np.ndim(lo[1]) else lo[1]
    out[hi[0]] = hi[1][hi[0]] if np.ndim(hi[1]) else hi[1]
    if mid.any():
        out[mid] = inside(x[mid])
    return out


def cel_grammage(tb, scheme, material, K):
    """CSDA range, cel_grammage, src/pumas.c:3644-3669."""
    K = np.asarray(K, np.float64)
    imax = tb.n_energies - 1
    X = tb.col2("X", scheme, material)
    a = tb.sections["a_max"][material]
    b = tb.b_max(scheme, material)
    K0 = tb.K[imax]
    K1 = a / b + tb.mass
    with np.errstate(all="ignore"):
        ext = X[imax] + 1. / b * np.log((K + K1) / (K0 + K1))
    return _piecewise(
        K, lambda k: (k < tb.K[0], 0.0), lambda k: (k >= tb.K[imax], ext),
        lambda k: interpolate(tb.K, X, tb.col2("X_dK", scheme, material), k))


def cel_proper_time(tb, scheme, material, K):
    """cel_proper_time, src/pumas.c:3720-3745."""
    K = np.asarray(K, np.float64)
    imax = tb.n_energies - 1
    T = tb.col2("T", scheme, material)
    a = tb.sections["a_max"][material]
    b = tb.b_max(scheme, material)
    E0 = tb.K[imax] + tb.mass
    E1 = K + tb.mass
    with np.errstate(all="ignore"):
        ext = T[imax] + tb.mass / a * np.log(
            (E1 / E0) * (a + b * E0) / (a + b * E1))
    return _piecewise(
        K, lambda k: (k < tb.K[0], 0.0), lambda k: (k >= tb.K[imax], ext),
        lambda k: interpolate(tb.K, T, tb.col2("T_dK", scheme, material), k))


def cel_kinetic_energy(tb, scheme, material, X):
    """Inverse range, cel_kinetic_energy, src/pumas.c:3758-3784."""
    X = np.asarray(X, np.float64)
    imax = tb.n_energies - 1
    tX = tb.col2("X", scheme, material)
    a = tb.sections["a_max"][material]
    b = tb.b_max(scheme, material)
    K0 = tb.K[imax]
    K1 = a / b + tb.mass
    with np.errstate(all="ignore"):
        ext = (K0 + K1) * np.exp(b * (X - tX[imax])) - K1
    return _piecewise(
        X, lambda x: (x < tX[0], 0.0), lambda x: (x >= tX[imax], ext),
        lambda x: interpolate(tX, tb.K, tb.col2("K_dX", scheme, material), x))


def cel_energy_loss(tb, scheme, material, K):
    """Stopping power, cel_energy_loss, src/pumas.c:3796-3814."""
    K = np.asarray(K, np.float64)
    imax = tb.n_energies - 1
    ext = tb.sections["a_max"][material] + tb.b_max(scheme, material) * (
        K + tb.mass)
    return _piecewise(
        K, lambda k: (k < tb.K[0], 0.0), lambda k: (k >= tb.K[imax], ext),
        lambda k: interpolate(tb.K, tb.col2("dE", scheme, material),
                              tb.col2("dE_dK", scheme, material), k))


def cel_straggling(tb, material, K):
    """cel_straggling, src/pumas.c:3826-3846."""
    K = np.asarray(K, np.float64)
    imax = tb.n_energies - 1
    Om = tb.col1("Omega", material)
    r = (K + tb.mass) / (tb.K[imax] + tb.mass)
    return _piecewise(
        K, lambda k: (k < tb.K[0], 0.0),
        lambda k: (k >= tb.K[imax], Om[imax] * r * r),
        lambda k: interpolate(tb.K, Om, tb.col1("Omega_dK", material), k))


def elastic_path(tb, material, K):
    """Hard elastic mean free path, coulomb_ehs_length,
    src/pumas.c:7446-7466 (K in [K[1], K[max]))."""
    K = np.asarray(K, np.float64)
    Lb = tb.col1("Lb", material)
    i1 = table_index(tb.K, K)
    h = (K - tb.K[i1]) / (tb.K[i1 + 1] - tb.K[i1])
    p2 = K * (K + 2. * tb.mass)
    return (Lb[i1] + h * (Lb[i1 + 1] - Lb[i1])) / p2


def transport_csda(tb, states, material, density, backward, limit_distance,
                   weighted_decay=True):
    """transport_with_csda, src/pumas.c:4777-4949, for an infinite uniform
    medium with a distance limit only (no magnetic field). `states` is a
    STATE_DTYPE array, returned updated (copy)."""
    s = np.array(states, copy=True)
    ki = s["energy"].copy()
    di = s["distance"].copy()
    ti = s["time"].copy()
    xi = cel_grammage(tb, CSDA, material, ki)
    Ti = cel_proper_time(tb, CSDA, material, ki)
    xB = density * (limit_distance - di)
    event = np.full(len(s), 2, np.int32)  # PUMAS_EVENT_LIMIT_DISTANCE
    if backward:
        xf = xB + xi
        small = xB < EPSILON_X * xi
        kf = np.where(small, ki + cel_energy_loss(tb, CSDA, material, ki) * xB,
                      cel_kinetic_energy(tb, CSDA, material, xf))
    else:
        inside = xi > xB
        xf = np.where(inside, xi - xB, 0.0)
        small = xB < EPSILON_X * xi
        kf = np.where(small,
                      np.maximum(ki - cel_energy_loss(tb, CSDA, material, ki) *
                                 xB, 0.0),
                      cel_kinetic_energy(tb, CSDA, material, xf))
        kf = np.where(inside, kf, 0.0)
        event = np.where(inside, event, 1)  # PUMAS_EVENT_LIMIT_ENERGY
    distance = np.abs(xf - xi) / density
    s["energy"] = kf
    s["distance"] = np.where(event == 2, limit_distance, di + distance)
    s["grammage"] += distance * density
    s["time"] += np.abs(Ti - cel_proper_time(tb, CSDA, material, kf)) / density
    if backward:
        s["weight"] *= (cel_energy_loss(tb, CSDA, material, kf) /
                        cel_energy_loss(tb, CSDA, material, ki))
    if weighted_decay:
        s["weight"] *= np.exp(-np.abs(ti - s["time"]) / tb.ctau)
    path = -distance if backward else distance
    s["position"] += path[:, None] * s["direction"]
    return s, event
