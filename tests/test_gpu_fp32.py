"""FP32-bulk transport against the FP64 transport (same random sequence).

FP32-bulk evaluates the per-step table lookups, straggling, scattering
angles and exponential path lengths in single precision and keeps the state
accumulators, the range to energy inversion, event weights and geometry in
double precision. It is validated statistically against FP64 on the same
histories: integrated flux bias within 0.1% (backward muography, Gaisser
flux), transmitted fraction and spectra within statistical agreement
(chi2 per degree of freedom of the energy and angle histograms), and the
per-event spread is reported.
"""
import numpy as np
import pytest

import pumas_amd as pa
from pumas_amd import workloads as W

pytestmark = pytest.mark.gpu


def _pair(physics, ctx, desc, states, seed=0):
    geom = pa.Geometry(physics, desc)
    r64 = pa.transport(physics, ctx, geom, states.copy(), seed=seed)
    r32 = pa.transport(physics, ctx, geom, states.copy(), seed=seed,
                       precision=pa.Precision.FP32_BULK)
    return r64, r32


def _chi2(a, b, wa=None, wb=None, bins=None):
    ha, _ = np.histogram(a, bins=bins, weights=wa)
    hb, _ = np.histogram(b, bins=bins, weights=wb)
    va, _ = np.histogram(a, bins=bins, weights=None if wa is None else wa**2)
    vb, _ = np.histogram(b, bins=bins, weights=None if wb is None else wb**2)
    ok = (va + vb) > 0
    chi2 = np.sum((ha[ok] - hb[ok]) ** 2 / (va[ok] + vb[ok]))
    return chi2, int(ok.sum())


def test_fp32_flux_bias_terrain(gpu_standard):
    desc = W.terrain_geometry(256, spacing=20.0)
    s = W.terrain_states(200000, 255 * 20.0, seed=21)
    r64, r32 = _pair(gpu_standard, W.terrain_context(), desc, s, seed=21)
    assert (r32.events == r64.events).mean() > 0.999
    c64 = -r64.states["direction"][:, 2]
    c32 = -r32.states["direction"][:, 2]
    f64 = r64.states["weight"] * W.gaisser_flux(r64.states["energy"], c64)
    f32 = r32.states["weight"] * W.gaisser_flux(r32.states["energy"], c32)
    bias = f32.mean() / f64.mean() - 1.0
    sigma = np.std(f32 - f64) / np.sqrt(len(f64)) / f64.mean()
    rel = np.abs(r32.states["energy"] / r64.states["energy"] - 1.0)
    print(f"fp32 flux bias {bias:.2e} +- {sigma:.1e}; per-event energy "
          f"error median {np.median(rel):.1e}, 90% {np.quantile(rel, 0.9):.1e}"
          f", 99% {np.quantile(rel, 0.99):.1e}")
    assert sigma < 5e-4
    assert abs(bias) < 1e-3
    bins = np.logspace(-2, 4, 31)
    chi2, ndf = _chi2(r64.states["energy"], r32.states["energy"],
                      r64.states["weight"], r32.states["weight"], bins)
    print(f"energy histogram chi2/ndf {chi2:.1f}/{ndf}")
    assert chi2 / ndf < 2.0
    chi2, ndf = _chi2(c64, c32, bins=np.linspace(0.3, 1.0, 36))
    print(f"angle histogram chi2/ndf {chi2:.1f}/{ndf}")
    assert chi2 / ndf < 2.0


def test_fp32_forward_slabs(gpu_standard):
    s = W.slabs_forward_states(200000, seed=22)
    r64, r32 = _pair(gpu_standard, W.slabs_forward_context(),
                     W.slabs_geometry(), s, seed=22)
    n = len(s)
    stop64 = (r64.events == pa.Event.LIMIT_ENERGY).mean()
    stop32 = (r32.events == pa.Event.LIMIT_ENERGY).mean()
    err = np.sqrt(2 * stop64 * (1 - stop64) / n)
    print(f"stopped fraction fp64 {stop64:.5f} fp32 {stop32:.5f} +- {err:.5f}")
    assert abs(stop32 - stop64) < 5 * err + 1e-4
    out64 = r64.events == pa.Event.MEDIUM
    out32 = r32.events == pa.Event.MEDIUM
    bins = np.logspace(-1, 4, 31)
    chi2, ndf = _chi2(r64.states["energy"][out64],
                      r32.states["energy"][out32], bins=bins)
    print(f"exit energy chi2/ndf {chi2:.1f}/{ndf}")
    assert chi2 / ndf < 2.0
    steps = r32.total_steps / r64.total_steps - 1.0
    print(f"steps fp32 / fp64 - 1 = {steps:.2e}")
    assert abs(steps) < 1e-2


@pytest.mark.parametrize("physics", ["gpu_standard", "gpu_iron"])
@pytest.mark.parametrize("precision", [pa.Precision.FP32_BULK,
                                       pa.Precision.FP64])
@pytest.mark.parametrize("tail_del", [1, -1])
@pytest.mark.parametrize("shape", [(256, 2), (512, 1)])
def test_lds_tables_invariance(request, physics, precision, tail_del, shape):
    """The stepping kernel reads the tables of its per-step lookups from a
    copy in LDS or from HBM: the same values, so the same results, bit for
    bit, in both precisions. The copies (FP64 step tables: 109 KB standard
    rock, 163 KB with iron; FP32 mirrors with the FP64 range tables: 83 KB
    and 108 KB) fit once in a CU's LDS: one 512-lane block per CU."""
    phys = request.getfixturevalue(physics)
    s = W.slabs_forward_states(4000, seed=23)
    geom = pa.Geometry(phys, W.slabs_geometry())
    ctx = W.slabs_forward_context()
    kw = dict(seed=23, precision=precision, tail_del=tail_del,
              block_size=shape[0], blocks_per_cu=shape[1])
    hbm = pa.transport(phys, ctx, geom, s.copy(), lds_tables=-1, **kw)
    lds = pa.transport(phys, ctx, geom, s.copy(), lds_tables=1, **kw)
    print(f"LDS copy {lds.lds_bytes} B per block")
    assert hbm.lds_bytes == 0
    if shape[1] == 1:
        assert lds.lds_bytes > 0
    else:
        assert lds.lds_bytes == 0
    assert (lds.states == hbm.states).all()
    for k in ("events", "media", "steps", "draws", "rc"):
        np.testing.assert_array_equal(getattr(lds, k), getattr(hbm, k))
