"""The oracle, pinned before it is trusted (CPU only).

The oracle is the reference library compiled from its source plus the
numpy restatement of its CSDA path. Checks:
  - Philox4x32-10 known-answer vectors (Random123 kat_vectors);
  - the tables uploaded to the device are the reference's own, byte for
    byte (export of a freshly created physics vs the committed PTAB);
  - the reference test suite's known answers for standard rock
    (tests/test-pumas.c:909-986), on the reference and on the restatement;
  - closed forms: PCHIP is exact on cubics, bracketing agrees with brute
    force, CSDA range / weight identities;
  - the restatement reproduces the reference history by history.
"""
import filecmp
import os

import numpy as np
import pytest

from conftest import REPO, load_golden

import restatement as rs
from pumas_amd import workloads as W

PTAB = os.path.join(REPO, "pumas_amd", "physics", "standard_c005.ptab")
ROCK = 0

# Random123 kat_vectors for philox4x32_10: (counter, key) -> output.
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF),
     (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344),
     (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.fixture(scope="module")
def tables():
    return rs.Tables(PTAB)


def test_philox_kat_numpy():
    for ctr, key, out in PHILOX_KAT:
        r = W.philox4x32_10(*[[c] for c in ctr], *key)
        assert tuple(int(x[0]) for x in r) == out


def test_philox_kat_oracle(ref_standard):
    import oracle as orc
    for ctr, key, out in PHILOX_KAT:
        assert orc.philox(ctr, key) == out


def test_uniform_stream_layout(ref_standard):
    import oracle as orc
    ids = np.array([0, 1, 7, 2**33 + 5], np.uint64)
    for index in range(5):
        u = W.uniform(12345, 1, ids, index)
        for i, h in enumerate(ids):
            assert u[i] == orc.uniform(12345, 1, int(h), index)
    u = W.uniform(0, 0, np.arange(100000, dtype=np.uint64), 0)
    assert 0.0 < u.min() and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 5e-3


@pytest.mark.parametrize("name", ["standard_c005", "iron_c001"])
def test_ptab_is_reference_export(name, tmp_path):
    import oracle as orc
    if not orc.available():
        pytest.skip("oracle library not built")
    p = orc.RefPhysics(name)
    out = tmp_path / (name + ".ptab")
    p.export(out)
    p.close()
    committed = os.path.join(REPO, "pumas_amd", "physics", name + ".ptab")
    assert filecmp.cmp(out, committed, shallow=False)


# tests/test-pumas.c:909-951, 976: (quantity, K, expected, tolerance).
KNOWN = [
    ("stopping_power", 1e0, 1.823e-4, 1e-6),
    ("stopping_power", 1e3, 6.610e-4, 1e-7),
    ("straggling", 1e3, 3.743e-4, 1e-7),
    ("range", 1e0, 5.518e3, 1.0),
    ("range", 1e3, 2.449e6, 1e3),
    ("elastic_path", 1e0, 275.89, 1e-2),
]


@pytest.mark.parametrize("quantity,K,expected,tol", KNOWN)
def test_known_answers_restatement(tables, quantity, K, expected, tol):
    f = {
        "stopping_power": lambda k: rs.cel_energy_loss(tables, rs.CSDA,
                                                       ROCK, k),
        "straggling": lambda k: rs.cel_straggling(tables, ROCK, k),
        "range": lambda k: rs.cel_grammage(tables, rs.CSDA, ROCK, k),
        "elastic_path": lambda k: rs.elastic_path(tables, ROCK, k),
    }[quantity]
    assert abs(float(f(np.array([K]))[0]) - expected) <= tol


@pytest.mark.parametrize("quantity,K,expected,tol", KNOWN)
def test_known_answers_reference(ref_standard, quantity, K, expected, tol):
    which = {"stopping_power": 3, "straggling": 4, "range": 0,
             "elastic_path": 11}[quantity]
    assert abs(ref_standard.cel(which, rs.CSDA, ROCK, K) - expected) <= tol


def test_inverse_range_known_answers(tables):
    # tests/test-pumas.c:941-951
    k = rs.cel_kinetic_energy(tables, rs.CSDA, ROCK, np.array([5.518e3, 2.449e6]))
    assert abs(k[0] - 1.0) <= 1e-3
    assert abs(k[1] - 1e3) <= 1.0


def test_pchip_exact_on_cubics():
    rng = np.random.default_rng(1)
    for _ in range(20):
        a = rng.normal(size=4)
        f = lambda x: a[0] + a[1] * x + a[2] * x**2 + a[3] * x**3
        df = lambda x: a[1] + 2 * a[2] * x + 3 * a[3] * x**2
        x0, x1 = sorted(rng.uniform(-2, 2, 2))
        t = np.linspace(0, 1, 11)
        h = x1 - x0
        y = rs.pchip(t, f(x0), f(x1), df(x0) * h, df(x1) * h)
        np.testing.assert_allclose(y, f(x0 + t * h), rtol=1e-12, atol=1e-12)


def test_table_index_brute_force(tables):
    K = tables.K
    rng = np.random.default_rng(2)
    x = np.concatenate([K, 10 ** rng.uniform(-4, 10, 5000), [-1.0]])
    idx = rs.table_index(K, x)
    for xi, i in zip(x, idx):
        brute = max([j for j in range(len(K)) if K[j] <= xi], default=-1)
        assert i == brute


def test_interpolation_matches_reference(ref_standard, tables):
    rng = np.random.default_rng(3)
    K = 10 ** rng.uniform(-3, 9.5, 2000)
    for scheme in (rs.CSDA, rs.MIXED):
        for material in range(3):
            x = rs.cel_grammage(tables, scheme, material, K)
            e = rs.cel_energy_loss(tables, scheme, material, K)
            for i in range(0, len(K), 97):
                assert x[i] == pytest.approx(
                    ref_standard.cel(0, scheme, material, K[i]), rel=1e-15)
                assert e[i] == pytest.approx(
                    ref_standard.cel(3, scheme, material, K[i]), rel=1e-15)
            ki = rs.cel_kinetic_energy(tables, scheme, material, x)
            for i in range(0, len(K), 97):
                assert ki[i] == pytest.approx(
                    ref_standard.cel(2, scheme, material, x[i]), rel=1e-15)


def test_restatement_reproduces_golden_csda(tables):
    g = load_golden("csda_rock")
    final, event = rs.transport_csda(tables, g["initial"], ROCK, 2650.0,
                                     backward=True, limit_distance=1e3)
    assert (event == g["events"]).all()
    for f in ("energy", "distance", "grammage", "time", "weight"):
        np.testing.assert_allclose(final[f], g["final"][f], rtol=1e-14,
                                   atol=0)
    np.testing.assert_allclose(final["position"], g["final"]["position"],
                               rtol=0, atol=1e-12)


def test_csda_closed_forms(tables):
    # Forward CSDA to rest: the history travels exactly its CSDA range
    # (tests/test-pumas.c:2438-2470).
    s = W.csda_rock_states(512, seed=7)
    s["direction"][:, 2] = 1.0
    x0 = rs.cel_grammage(tables, rs.CSDA, ROCK, s["energy"])
    final, event = rs.transport_csda(tables, s, ROCK, 2650.0, backward=False,
                                     limit_distance=1e9)
    assert (event == 1).all() and (final["energy"] == 0).all()
    np.testing.assert_allclose(final["distance"], x0 / 2650.0, rtol=1e-14)
    # Backward weight: w_f / w_i = S(K_f) / S(K_i) exp(-dt / c tau)
    # (tests/test-pumas.c:2732-2785).
    s = W.csda_rock_states(512, seed=8)
    final, _ = rs.transport_csda(tables, s, ROCK, 2650.0, backward=True,
                                 limit_distance=1e3)
    ratio = (rs.cel_energy_loss(tables, rs.CSDA, ROCK, final["energy"]) /
             rs.cel_energy_loss(tables, rs.CSDA, ROCK, s["energy"]) *
             np.exp(-final["time"] / tables.ctau))
    np.testing.assert_allclose(final["weight"] / s["weight"], ratio,
                               rtol=1e-13)


GOLDEN_CASES = ["csda_rock", "slabs_straggled", "slabs_mixed",
                "slabs_csda_check", "slabs_backward", "limits_forward",
                "limits_time_backward"]


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_oracle_reproduces_golden(ref_standard, name):
    """The committed fixtures are what the reference produces here."""
    import oracle as orc
    import make_golden
    case = {c[0]: c for c in make_golden.cases()}[name]
    _, _, ctx, desc, states = case
    g = load_golden(name)
    assert (states == g["initial"]).all()
    r = orc.transport(ref_standard, ctx, orc.RefGeometry(desc), states,
                      seed=int(g["seed"]), threads=4)
    assert (r.states == g["final"]).all()
    for k in ("events", "media", "steps", "draws", "rc"):
        assert (getattr(r, k) == g[k]).all(), k


def test_flux_models_invariants():
    """The reference's example flux models (compiled from
    examples/pumas/flux.c): the charge fractions of both signs add up to the
    total flux; Volkova's cos(theta*) is 1 at the vertical, so that GCCLY
    tends to Gaisser at high energy; the Gaisser flux falls as E^-2.7 far
    below the pion and kaon critical energies' corrections."""
    import oracle as orc
    if not os.path.exists(orc.LIBRARY_FLUX):
        pytest.skip("oracle not built")
    k = np.logspace(-2, 4, 25)
    c = np.linspace(0.1, 1.0, 25)
    for kind in (1, 2):
        fm = orc.flux(kind, c, k, -1.0)
        fp = orc.flux(kind, c, k, 1.0)
        f0 = orc.flux(kind, c, k, 0.0)
        np.testing.assert_allclose(fm + fp, f0, rtol=1e-14)
        np.testing.assert_allclose(fp / fm, 1.2766, rtol=1e-14)
    kh = np.array([1e5, 1e6])
    np.testing.assert_allclose(orc.flux(2, 1.0, kh, 0.0),
                               orc.flux(1, 1.0, kh, 0.0), rtol=1e-3)
    e = np.array([1e-1, 1e-2]) + 0.10566
    f = orc.flux(1, 1.0, e - 0.10566, 0.0)
    np.testing.assert_allclose(f[0] / f[1], (e[0] / e[1]) ** -2.7, rtol=2e-3)
