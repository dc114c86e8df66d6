"""Shared test utilities: per-variable comparison of two states."""
import numpy as np

from conftest import normwise_rel

VARS = {"ns": ("fluidvars", 0), "vs1": ("fluidvars", 7), "vs2": ("fluidvars", 14),
        "vs3": ("fluidvars", 21), "Ts": ("fluidvars", 28), "rhovs1": ("fluidauxvars", 0),
        "rhoes": ("fluidauxvars", 7)}


def var(st, name, interior=False):
    block, off = VARS[name]
    a = getattr(st, block)[..., off:off + 7]
    if interior:
        a = a[2:-2, 2:-2, 2:-2]
    return a


def compare(got, ref, exact=(), close=(), rtol=1e-12, interior_close=True):
    """Assert bit equality for `exact` variables and normwise rtol for `close` ones,
    species by species; returns the worst relative difference seen."""
    worst = 0.0
    for name in exact:
        a, b = var(got, name), var(ref, name)
        same = np.array_equal(a, b, equal_nan=True)
        if not same:
            bad = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b))))
            i = tuple(bad[0])
            raise AssertionError(f"{name} not bit-identical at {len(bad)} cells, first {i}: "
                                 f"{a[i]!r} vs {b[i]!r} (rel {normwise_rel(a, b):.3e})")
    for name in close:
        a, b = var(got, name, interior_close), var(ref, name, interior_close)
        for s in range(7):
            r = normwise_rel(a[..., s], b[..., s])
            worst = max(worst, r)
            assert r <= rtol, f"{name} species {s + 1}: normwise rel diff {r:.3e} > {rtol:g}"
    return worst
