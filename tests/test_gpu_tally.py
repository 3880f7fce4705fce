"""On-device flux tallies against the reference's example flux models.

The per-history flux values come from examples/pumas/flux.c compiled from
the reference's source (oracle/_ref/libflux_ref.so); sums over histories are
compared at 1e-12 (only the summation order differs), histogram bins are
compared for values placed away from bin edges, so that the bin of every
history is unambiguous.
"""
import numpy as np
import pytest

import pumas_amd as pa
from pumas_amd import workloads as W

pytestmark = pytest.mark.gpu


def random_states(n, seed):
    rng = np.random.default_rng(seed)
    s = pa.new_states(n)
    s["energy"] = 10 ** rng.uniform(-2, 4, n)
    c = rng.uniform(0.05, 1.0, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    st = np.sqrt(1 - c * c)
    s["direction"][:, 0] = st * np.cos(phi)
    s["direction"][:, 1] = st * np.sin(phi)
    s["direction"][:, 2] = -c
    s["charge"] = rng.choice([-1.0, 1.0], n)
    s["weight"] = rng.uniform(0.5, 2.0, n)
    return s


def oracle_flux(kind, states, charge_fraction=True, cos=None):
    import oracle as orc
    c = -states["direction"][:, 2] if cos is None else np.full(len(states), cos)
    q = states["charge"] if charge_fraction else np.zeros(len(states))
    return orc.flux(kind, c, states["energy"], q)


@pytest.mark.parametrize("flux", [pa.Flux.GAISSER, pa.Flux.GCCLY])
def test_flux_sum_vs_reference_models(flux):
    s = random_states(100_000, 1)
    t = pa.tally(s, spec=pa.TallySpec(flux=flux))
    v = s["weight"] * oracle_flux(int(flux), s)
    assert t.count == len(s)
    np.testing.assert_allclose(t.sum, v.sum(), rtol=1e-12)
    np.testing.assert_allclose(t.sum2, (v * v).sum(), rtol=1e-12)


def test_event_selection_and_weights_only():
    s = random_states(50_000, 2)
    ev = np.where(np.arange(len(s)) % 3 == 0, pa.Event.MEDIUM,
                  pa.Event.LIMIT_ENERGY).astype(np.int32)
    t = pa.tally(s, ev, pa.TallySpec(flux=pa.Flux.NONE,
                                      event_mask=int(pa.Event.MEDIUM)))
    sel = ev == pa.Event.MEDIUM
    assert t.count == sel.sum()
    np.testing.assert_allclose(t.sum, s["weight"][sel].sum(), rtol=1e-12)
    t_all = pa.tally(s, ev, pa.TallySpec(flux=pa.Flux.NONE))
    assert t_all.count == len(s)


def test_histograms_vs_reference_models():
    n_e, n_c = 24, 20
    emin, emax = 1e-2, 1e4
    rng = np.random.default_rng(3)
    n = 60_000
    s = random_states(n, 4)
    # Values at 10%..90% of their bin width: no edge ambiguity.
    be = rng.integers(0, n_e, n)
    bc = rng.integers(0, n_c, n)
    le = np.log(emin) + (be + rng.uniform(0.1, 0.9, n)) * \
        (np.log(emax) - np.log(emin)) / n_e
    s["energy"] = np.exp(le)
    c = 0.2 + (bc + rng.uniform(0.1, 0.9, n)) * 0.8 / n_c
    st = np.sqrt(1 - c * c)
    s["direction"][:, 0] = st
    s["direction"][:, 1] = 0.0
    s["direction"][:, 2] = -c
    spec = pa.TallySpec(flux=pa.Flux.GCCLY, n_energy_bins=n_e,
                        energy_min=emin, energy_max=emax, n_cos_bins=n_c,
                        cos_min=0.2, cos_max=1.0)
    t = pa.tally(s, spec=spec)
    v = s["weight"] * oracle_flux(2, s)
    he = np.bincount(be, weights=v, minlength=n_e)
    hc = np.bincount(bc, weights=v, minlength=n_c)
    np.testing.assert_allclose(t.energy[:, 0], he, rtol=1e-12)
    np.testing.assert_allclose(t.energy[:, 1],
                               np.bincount(be, weights=v * v, minlength=n_e),
                               rtol=1e-12)
    np.testing.assert_allclose(t.cos_theta[:, 0], hc, rtol=1e-12)


def test_deterministic():
    s = random_states(200_000, 5)
    spec = pa.TallySpec(n_energy_bins=64, n_cos_bins=32)
    a = pa.tally(s, spec=spec)
    b = pa.tally(s, spec=spec)
    assert a.sum == b.sum and a.sum2 == b.sum2
    assert (a.energy == b.energy).all() and (a.cos_theta == b.cos_theta).all()


def test_csda_rock_flux_vs_reference(gpu_standard, ref_standard):
    """Config of the plain method: backward CSDA through 1 km of standard
    rock, Gaisser-form flux 0.14 E^-2.7 [...] per cm^2 at cos(theta) = 1
    without charge fraction; GPU transport + GPU tally against reference
    transport + reference flux model."""
    import oracle as orc
    n = 100_000
    s0 = W.csda_rock_states(n)
    ctx = W.csda_rock_context()
    r = pa.transport(gpu_standard, ctx, pa.Geometry(
        gpu_standard, W.csda_rock_geometry()), s0.copy())
    spec = pa.TallySpec(flux=pa.Flux.GAISSER, charge_fraction=False,
                        scale=1e-4, cos_from_direction=False, cos_theta=1.0)
    t = pa.tally(r.states, r.events, spec)
    ref = orc.transport(ref_standard, ctx,
                        orc.RefGeometry(W.csda_rock_geometry()), s0, seed=0)
    v = ref.states["weight"] * oracle_flux(1, ref.states, False, 1.0) * 1e-4
    np.testing.assert_allclose(t.sum, v.sum(), rtol=1e-11)
    np.testing.assert_allclose(t.mean, v.mean(), rtol=1e-11)


def test_bad_specs():
    s = random_states(10, 6)
    with pytest.raises(pa.PumasError):
        pa.tally(s, spec=pa.TallySpec(n_energy_bins=1000))
    with pytest.raises(pa.PumasError):
        pa.tally(s, spec=pa.TallySpec(n_energy_bins=4, energy_min=-1.0))
    with pytest.raises(pa.PumasError):
        pa.tally(s, spec=pa.TallySpec(flux=7))
