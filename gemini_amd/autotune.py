"""Per-member choice of kernel variants by measurement.

The library compiles alternative variants of two kernels (the x2/x3 sweep pipeline
depth, the TRBDF2 chunk height); results are identical, speed is not, and which is
faster depends on the grid.  ``autotune`` times every combination on each member with
the library's own kernel timers and keeps the fastest; members of the same shape share
one measurement.  Timing advances the member's state, so the caller re-pushes states
afterwards (``Ensemble.autotune`` does).
"""
from __future__ import annotations

import itertools

KNOBS = {"sweep_pf": (1, 2), "diff_chunk": (4, 8)}
TIMED = ("sweep1", "sweep2", "sweep3", "trbdf2", "pre_convert", "post_convert")


def time_step(tr, dt: float, steps: int = 3) -> float:
    """Milliseconds of the timed kernels per transport step on this member."""
    tr.transport_step(dt)                 # warm-up with the current variants
    tr.synchronize()
    tr.kernel_timing_reset()
    tr.kernel_timing(True)
    for _ in range(steps):
        tr.transport_step(dt)
    tr.kernel_timing(False)
    return sum(tr.kernel_time(k)[0] for k in TIMED) / steps


def best_variant(tr, dt: float, steps: int = 3):
    """(choice, {choice: ms}) for one member, over every combination of KNOBS."""
    results = {}
    for combo in itertools.product(*KNOBS.values()):
        choice = dict(zip(KNOBS, combo))
        for k, v in choice.items():
            tr.set_tuning(k, v)
        results[combo] = time_step(tr, dt, steps)
    best = min(results, key=results.get)
    return dict(zip(KNOBS, best)), results


def autotune(ensemble, states, steps: int = 3):
    """Choose the variants of every member of `ensemble` (members of one grid shape share
    a measurement), then restore each member's state from `states`."""
    chosen = {}
    for i, tr in enumerate(ensemble.trs):
        g = ensemble.grids[i]
        key = (g.kind, g.lx1, g.lx2, g.lx3)
        if key not in chosen:
            chosen[key] = best_variant(tr, float(ensemble.dts[i]), steps)[0]
        for k, v in chosen[key].items():
            tr.set_tuning(k, v)
        ensemble.push(i, states[i])
    return chosen
