"""Ensembles of independent grids (uneven field-line lengths) advanced in one HIP graph.

Every graph replay is checked, member by member and bit for bit, against the same member
stepped eagerly through the implicit library context; eager work and synchronisations
run between the replays.  A member is also checked against the oracle, and kernel
variants chosen per member must not change any result."""
import numpy as np
import pytest

import gemini_amd as ga
from gemini_amd.ensemble import Ensemble
from gemini_amd.transport import Transport, GemhipError
import oracle
from helpers import compare

pytestmark = pytest.mark.gpu

ALL = ("ns", "vs1", "vs2", "vs3", "Ts", "rhovs1", "rhoes")
SHAPES = [(16, 6, 4), (24, 5, 3), (40, 4, 4), (16, 9, 2), (32, 6, 5), (24, 7, 1)]


def members(seed0=20):
    out = []
    for k, shape in enumerate(SHAPES):
        g = ga.cartesian_grid(*shape)
        st = ga.synthetic_state(g, seed=seed0 + k, perturb=shape[2] > 1)
        out.append((g, st, ga.cfl_dt(g, st)))
    return out


def eager_states(ms, nsteps):
    """Each member stepped alone through the implicit context; states after each step."""
    hist = []
    for g, st, dt in ms:
        tr = Transport(g, st.nn, st.Tn)
        tr.set_residency(1)
        tr.push(st)
        per = []
        for _ in range(nsteps):
            tr.transport_step(dt)
            out = st.copy()
            tr.pull(out)
            per.append(out)
        tr.set_residency(0)
        hist.append(per)
    return hist


def test_graph_replays_match_eager_steps():
    ms = members()
    print(f"ensemble of {len(ms)} members", flush=True)
    ref = eager_states(ms, 3)
    ens = Ensemble(ms)
    try:
        for step in range(3):
            ens.run(1)                            # one replay of the captured one-step graph
            for i, (g, st, dt) in enumerate(ms):
                got = st.copy()
                ens.pull(i, got)
                compare(got, ref[i][step], exact=ALL)
        # member 0 against the oracle (the GPU's x**2.5), three steps
        g, st, dt = ms[0]
        r = st.copy()
        with oracle.pow_mode(1):
            for _ in range(3):
                oracle.transport_step(g, r, dt)
        got = st.copy()
        ens.pull(0, got)
        compare(got, r, exact=ALL)
    finally:
        ens.close()


def test_kernel_variants_do_not_change_results():
    ms = members(seed0=40)
    ref = eager_states(ms, 2)
    ens = Ensemble(ms)
    try:
        combos = [(1, 4), (2, 4), (1, 8), (2, 8)]
        for i, tr in enumerate(ens.trs):
            pf, ch = combos[i % len(combos)]
            tr.set_tuning("sweep_pf", pf)
            tr.set_tuning("diff_chunk", ch)
            assert tr.get_tuning("sweep_pf") == pf and tr.get_tuning("diff_chunk") == ch
        ens.run_plain(2)
        for i, (g, st, dt) in enumerate(ms):
            got = st.copy()
            ens.pull(i, got)
            compare(got, ref[i][1], exact=ALL)
        # the autotuner picks per shape and leaves the states as they were
        chosen = ens.autotune([m[1] for m in ms], steps=1)
        assert len(chosen) == len({(g.lx1, g.lx2, g.lx3) for g, _, _ in ms})
        ens.run(2)
        for i, (g, st, dt) in enumerate(ms):
            got = st.copy()
            ens.pull(i, got)
            compare(got, ref[i][1], exact=ALL)
    finally:
        ens.close()


def test_members_of_a_live_graph_cannot_be_destroyed():
    """A captured graph holds its members' HBM and streams: destroying a member is refused
    until the graph is released; autotuning releases graphs captured before it."""
    ms = members()[:2]
    ens = Ensemble(ms)
    try:
        gid = ens.capture(1)
        h = ens.trs[0].handle
        assert ens.lib.gemhip_context_destroy(int(h)) == 2
        assert b"ensemble graph" in ens.lib.gemhip_last_error()
        ens.lib.gemhip_clear_error()
        ens.autotune([st for _, st, _ in ms], steps=1)
        assert ens._graphs == {}
        assert ens.lib.gemhip_ensemble_launch(gid, 1) != 0      # released with the old variants
        ens.run(1)                                               # captured anew
    finally:
        ens.close()


BATCH_SHAPES = [(16, 6, 4), (24, 5, 3), (40, 4, 4), (16, 9, 2), (32, 6, 5), (64, 3, 7),
                (96, 17, 9)]


def batch_members(seed0=60):
    out = []
    for k, shape in enumerate(BATCH_SHAPES):
        g = ga.cartesian_grid(*shape)
        st = ga.synthetic_state(g, seed=seed0 + k, perturb=True)
        out.append((g, st, ga.cfl_dt(g, st)))
    return out


def eager_with_dts(ms, factors):
    """Each member stepped alone through the implicit context, step k with dt * factors[k]."""
    hist = []
    for g, st, dt in ms:
        tr = Transport(g, st.nn, st.Tn)
        tr.set_residency(1)
        tr.push(st)
        per = []
        for f in factors:
            tr.transport_step(dt * f)
            out = st.copy()
            tr.pull(out)
            per.append(out)
        tr.set_residency(0)
        hist.append(per)
    return hist


def _check_members(ens, ms, ref, step, tag):
    for i, (g, st, dt) in enumerate(ms):
        got = st.copy()
        ens.pull(i, got)
        print(f"  {tag}: member {i} {g.shape}", flush=True)
        compare(got, ref[i][step], exact=ALL)


def test_batched_kernels_match_plain_steps():
    """Batched kernels (one launch per kernel for all members) against each member's plain
    steps, bit for bit: an eager batched step, then graph replays under other kernel
    variants and persistent grid sizes, and a replay after the time steps changed in HBM."""
    ms = batch_members()
    print(f"batch of {len(ms)} members", flush=True)
    ref = eager_with_dts(ms, [1.0, 1.0, 1.0, 0.5])
    ens = Ensemble(ms)
    b = ens.batched()
    try:
        b.step(1)
        _check_members(ens, ms, ref, 0, "eager batched step")
        b.set_tuning("sweep_pf", 1)
        b.set_tuning("diff_chunk", 4)
        b.set_tuning("grid", 2)
        b.run(1)
        _check_members(ens, ms, ref, 1, "graph, sweep_pf 1, diff_chunk 4, grid 2")
        b.run(1)                                       # the same graph again
        _check_members(ens, ms, ref, 2, "graph replay")
        b.set_dt(np.array([dt * 0.5 for _, _, dt in ms]))
        b.run(1)                                       # captured before: reads dt from HBM
        _check_members(ens, ms, ref, 3, "graph replay at half the time step")
    finally:
        b.close()
        ens.close()


def test_batch_refuses_members_off_the_fused_path_and_holds_its_members():
    ms = batch_members(seed0=80)[:2]
    g = ga.cartesian_grid(24, 7, 1)                   # lx3 = 1: not on the fused path
    st = ga.synthetic_state(g, seed=3)
    ens = Ensemble(ms + [(g, st, ga.cfl_dt(g, st))])
    try:
        with pytest.raises(GemhipError, match="lx2, lx3 > 1"):
            ens.batched()
        ens.lib.gemhip_clear_error()
    finally:
        ens.close()
    ens = Ensemble(ms)
    b = ens.batched()
    try:
        b.capture(2)
        assert ens.lib.gemhip_context_destroy(int(ens.trs[1].handle)) == 2
        assert b"graph" in ens.lib.gemhip_last_error()     # the graph holds the member
        ens.lib.gemhip_clear_error()
        b._release_graphs()
        assert ens.lib.gemhip_context_destroy(int(ens.trs[1].handle)) == 2
        assert b"batch" in ens.lib.gemhip_last_error()     # and so does the batch
        ens.lib.gemhip_clear_error()
        gid = b.capture(1)
        b.close()                                          # releases its graphs too
        assert ens.lib.gemhip_ensemble_launch(gid, 1) != 0
        ens.lib.gemhip_clear_error()
    finally:
        b.close()
        ens.close()
