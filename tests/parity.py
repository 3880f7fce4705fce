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


def assert_parity(gpu, ref, rtol, rtol_strict=1e-12, frac_strict=1.0,
                  counters=True, weight_rtol=None, max_flips=0.0):
    """Events, media, step and draw counts identical; every history within
    `rtol`, and at least `frac_strict` of them within `rtol_strict`. When
    `weight_rtol` is given the weight is checked against it separately.
    Up to a fraction `max_flips` of the histories may branch differently
    (different counters, or values beyond the tolerances); they are
    reported and left out of the strict fraction."""
    final = ref["final"] if isinstance(ref, dict) else ref.states
    flipped = np.zeros(len(final), bool)
    if counters:
        for k in ("events", "media", "steps", "draws", "rc"):
            g, r = getattr(gpu, k), ref[k] if isinstance(ref, dict) else \
                getattr(ref, k)
            flipped |= np.atleast_2d((g != r).T).any(axis=0)
    fields = FIELDS if weight_rtol is None else FIELDS[:-1]
    err = state_errors(gpu.states, final, fields)
    werr = weight_errors(gpu.states, final)
    outliers = err > rtol
    if weight_rtol is not None:
        outliers |= werr > weight_rtol
    outliers |= gpu.states["decayed"] != final["decayed"]
    bad = np.nonzero(flipped | outliers)[0]
    if len(bad):
        print(f"parity: {len(bad)} of {len(final)} histories branch "
              f"differently (budget {max_flips:.2%})")
    if len(bad) > max_flips * len(final):
        worst = int(bad[0])
        detail = {f: float(state_errors(gpu.states[worst:worst + 1],
                                        final[worst:worst + 1], (f,))[0])
                  for f in FIELDS}
        raise AssertionError(
            f"{len(bad)} histories differ (first {bad[:10]}); history "
            f"{worst}: flipped={bool(flipped[worst])} errors {detail}")
    keep = ~(flipped | outliers)
    strict = float(np.mean(err[keep] <= rtol_strict)) if keep.any() else 1.0
    print(f"parity: n={len(final)} branch-flips={int(flipped.sum())} "
          f"within {rtol_strict:g}: {float(np.mean(err <= rtol_strict)):.6f} "
          f"max err {float(err.max()) if len(err) else 0.0:.3e} "
          f"max weight err {float(werr.max()) if len(werr) else 0.0:.3e}",
          flush=True)
    assert strict >= frac_strict, f"only {strict:.4f} within {rtol_strict}"
    return err
