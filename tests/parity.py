"""Per-history comparison of transported states."""
import numpy as np

FIELDS = ("energy", "distance", "grammage", "time", "weight")


def state_errors(a, b, fields=FIELDS):
    """Largest relative error per history over the scalar fields, the
    position (relative to the larger of |r| and the distance travelled)
    and the direction (absolute)."""
    err = np.zeros(len(a))
    for f in fields:
        x, y = a[f], b[f]
        scale = np.maximum(np.abs(y), 1e-300)
        e = np.where(x == y, 0.0, np.abs(x - y) / scale)
        err = np.maximum(err, e)
    scale = np.maximum(np.linalg.norm(b["position"], axis=1), b["distance"])
    scale = np.maximum(scale, 1e-300)
    e = np.abs(a["position"] - b["position"]).max(axis=1) / scale
    err = np.maximum(err, np.where(scale > 0, e, 0.0))
    err = np.maximum(err, np.abs(a["direction"] - b["direction"]).max(axis=1))
    return err


def weight_errors(a, b):
    x, y = a["weight"], b["weight"]
    return np.where(x == y, 0.0, np.abs(x - y) / np.maximum(np.abs(y), 1e-300))


def assert_parity(gpu, ref, rtol, rtol_strict=1e-12, frac_strict=0.97,
                  counters=True, weight_rtol=None):
    """Events, media, step and draw counts identical; every history within
    `rtol`, and at least `frac_strict` of them within `rtol_strict`. When
    `weight_rtol` is given the weight is checked against it separately."""
    if counters:
        for k in ("events", "media", "steps", "draws", "rc"):
            g, r = getattr(gpu, k), ref[k] if isinstance(ref, dict) else \
                getattr(ref, k)
            bad = np.nonzero(np.atleast_2d((g != r).T).any(axis=0))[0]
            assert len(bad) == 0, f"{k} differ for histories {bad[:10]}"
    final = ref["final"] if isinstance(ref, dict) else ref.states
    fields = FIELDS if weight_rtol is None else FIELDS[:-1]
    err = state_errors(gpu.states, final, fields)
    if weight_rtol is not None:
        werr = weight_errors(gpu.states, final)
        assert werr.max() <= weight_rtol, f"weight error {werr.max()}"
    assert (gpu.states["decayed"] == final["decayed"]).all()
    worst = int(np.argmax(err))
    detail = {f: float(state_errors(gpu.states[worst:worst + 1],
                                    final[worst:worst + 1], (f,))[0])
              for f in FIELDS}
    assert err.max() <= rtol, \
        f"history {worst}: relative error {err[worst]} {detail}"
    strict = float(np.mean(err <= rtol_strict))
    assert strict >= frac_strict, f"only {strict:.4f} within {rtol_strict}"
    return err
