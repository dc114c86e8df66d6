"""GPU: the three benchmark configurations beyond the headline grid, at their full sizes.

* 3-D Cartesian 256x192x128 with Gaussian density perturbations: the fused step after one
  step and after 100 steps with a CFL-limited time step from the library's grid rates,
  checked on sampled x1 columns against the oracle's whole step on cuts of half-width 2
  (tests/cutgrid.py; the premise is test_cut_columns_reproduce_the_composite_step);
* the 64-member ensemble of uneven grids, 20 steps in one captured graph: members
  re-run alone are bit-identical, and the smallest member equals the oracle everywhere;
* the differentiation run, 256x128x96 over 50 steps with checkpoints every 10: the HIP
  adjoint against central differences, with the branches of the electron energy path
  that the +-eps runs cross counted per direction.
"""
import numpy as np
import pytest
import torch

import gemini_amd as ga
from gemini_amd.synth_gpu import synthetic_state_device, cfl_dt_device
from gemini_amd.transport import Transport
import oracle
from cutgrid import SubGrid, sub_state, host_state, line_view, sample_columns

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _check_columns(g, before, after, dt, cols, tag):
    """The fused step from ``before`` (vs1i: the interface velocities its ghost fill reads)
    gave ``after``; the oracle's whole step on each column's cut gives the same bits."""
    with oracle.pow_mode(1):
        for c in cols:
            sg = SubGrid(g, c, 2)
            s = sub_state(sg, before)
            oracle.transport_step(sg, s, dt)
            for blk in ("fluidvars", "fluidauxvars"):
                a = line_view(sg, getattr(after, blk), False)[..., :35]
                b = line_view(sg, getattr(s, blk), True)[..., :35]
                if not np.array_equal(a, b):
                    bad = np.argwhere(a != b)[0]
                    raise AssertionError(f"{tag}: column {c} {blk} differs first at {tuple(bad)}: "
                                         f"{a[tuple(bad)]!r} vs {b[tuple(bad)]!r}")


def _cfl_dt(tr, tcfl=0.9, dtmax=0.5):
    """dt_calc's rule on the resident state (the library's grid-rate reduction)."""
    r = tr.grid_rates()
    return min(dtmax, tcfl / max(float(r.max()), 1e-10))


@pytest.mark.timeout(900)
def test_cartesian_256x192x128_100_cfl_steps():
    print("case 3-D Cartesian 256x192x128: generating the state", flush=True)
    g = ga.cartesian_grid(256, 192, 128)
    dev = synthetic_state_device(g, seed=1, device=DEV, perturb=True)
    st = host_state(g, dev)
    del dev
    torch.cuda.empty_cache()
    tr = Transport(g, st.nn, st.Tn, geometry_on_device=True)
    tr.set_residency(1)
    tr.push(st)
    rng = np.random.default_rng(7)
    cols = sample_columns(g, rng, 16, 2) + [(None, 1, 1), (None, g.lx2, g.lx3), (None, 1, g.lx3)]
    # step 1 from the seeded state (interface velocities zero, as allocated)
    dt = _cfl_dt(tr)
    tr.transport_step(dt)
    after = st.copy()
    tr.pull(after)
    print(f"  step 1 (dt {dt:.4g} s): checking {len(cols)} columns", flush=True)
    _check_columns(g, st, after, dt, cols, "step 1")
    # steps 2..99; keep the state entering step 99 for the interface velocities its
    # interface_vels forms (the ghost fill of step 100 reads their x1 boundary planes)
    prev = None
    for k in range(2, 100):
        if k == 99:
            prev = st.copy()
            tr.pull(prev)
        dt = _cfl_dt(tr)
        tr.transport_step(dt)
    tr.pull(st)
    oracle.interface_vels(g, prev)
    st.vs1i[...] = prev.vs1i
    del prev
    dt = _cfl_dt(tr)
    tr.transport_step(dt)
    tr.pull(after)
    tr.set_residency(0)
    for name, (blk, lo) in {"ns": ("fluidvars", 0), "vs1": ("fluidvars", 7),
                            "Ts": ("fluidvars", 28)}.items():
        a = getattr(after, blk)[2:-2, 2:-2, 2:-2, lo:lo + 7]
        assert np.all(np.isfinite(a)), f"{name} not finite after 100 steps"
    print(f"  step 100 (dt {dt:.4g} s): state finite; checking {len(cols)} columns", flush=True)
    _check_columns(g, st, after, dt, cols, "step 100")
    print("  steps 1 and 100 bit-identical to the oracle on every sampled column", flush=True)


@pytest.mark.timeout(900)
def test_ensemble_64_members_20_steps():
    from gemini_amd.ensemble import Ensemble, draw_sizes, device_member
    sizes = draw_sizes(1, 64)
    seeds = [1000 + m for m in range(len(sizes))]
    steps = 20
    print(f"case ensemble of {len(sizes)} members, {sum(a * b * c for a, b, c in sizes)} cells",
          flush=True)
    made = [device_member(sz, sd) for sz, sd in zip(sizes, seeds)]
    grids = [m[0] for m in made]
    trs = [m[1] for m in made]
    dts = [m[2] for m in made]
    small = int(np.argmin([a * b * c for a, b, c in sizes]))
    init_small = host_state(grids[small], made[small][3])
    del made
    torch.cuda.empty_cache()
    ens = Ensemble.from_transports(trs, grids, dts)
    try:
        ens.run(steps)
        print(f"  {steps} steps in one graph launch", flush=True)
        keep = [0, 21, 42, 63, small]
        got = {}
        for i in keep:
            n = grids[i].padded[0] * grids[i].padded[1] * grids[i].padded[2]
            fv = torch.empty(35 * n, dtype=torch.float64, device=DEV)
            aux = torch.empty(14 * n, dtype=torch.float64, device=DEV)
            trs[i].pull_blocks(fv, aux)
            got[i] = (fv, aux)
    finally:
        ens.close()
    for i in keep[:4]:                                  # plain solves of members alone
        g, tr, dt, _ = device_member(sizes[i], seeds[i])
        for _ in range(steps):
            tr.transport_step(dt)
        n = g.padded[0] * g.padded[1] * g.padded[2]
        fv = torch.empty(35 * n, dtype=torch.float64, device=DEV)
        aux = torch.empty(14 * n, dtype=torch.float64, device=DEV)
        tr.pull_blocks(fv, aux)
        tr.close()
        assert torch.equal(fv, got[i][0]) and torch.equal(aux, got[i][1]), f"member {i}"
        print(f"  member {i} {sizes[i]}: bit-identical to its plain solve", flush=True)
    # the smallest member against the oracle, every cell
    g = grids[small]
    ref = init_small
    with oracle.pow_mode(1):
        for _ in range(steps):
            oracle.transport_step(g, ref, dts[small])
    L = g.padded
    fv = got[small][0].cpu().numpy().reshape((35,) + L[::-1]).transpose(3, 2, 1, 0)
    aux = got[small][1].cpu().numpy().reshape((14,) + L[::-1]).transpose(3, 2, 1, 0)
    assert np.array_equal(fv, ref.fluidvars[..., :35], equal_nan=True)
    assert np.array_equal(aux, ref.fluidauxvars[..., :14], equal_nan=True)
    print(f"  member {small} {sizes[small]}: bit-identical to the oracle after {steps} steps",
          flush=True)


@pytest.mark.timeout(900)
def test_adjoint_256x128x96_50_steps_against_central_differences():
    """Directions without a crossed branch must agree with central differences to 1e-6;
    the others are reported with their flip counts."""
    from gemini_amd.adjoint import TransportRun, electron_temperature, interior, branch_flips
    print("case adjoint 256x128x96, 50 steps, checkpoints every 10", flush=True)
    g = ga.cartesian_grid(256, 128, 96)
    st = synthetic_state_device(g, seed=1, device=DEV)
    dt = cfl_dt_device(g, st)
    run = TransportRun(g, st, dt, 50, every=10, geometry_on_device=True)
    run_m = TransportRun(g, st, dt, 50, every=10, geometry_on_device=True)
    del st
    torch.cuda.empty_cache()
    try:
        gen = torch.Generator(device=DEV).manual_seed(1)
        sh = (g.lx3, g.lx2)
        rnd = lambda: torch.rand(sh, generator=gen, device=DEV, dtype=torch.float64)
        k0, q0 = 1.0 + 0.2 * (2.0 * rnd() - 1.0), 1e-5 * rnd()
        kt, qt = 1.0 + 0.2 * (2.0 * rnd() - 1.0), 1e-5 * rnd()
        z = g.x1[2:g.lx1 + 2]
        alts = torch.tensor([int(abs(z - 1e3 * a).argmin())
                             for a in (150, 200, 250, 300, 400, 500, 650, 800)], device=DEV)
        target = interior(run, run.forward(kt, qt))[..., alts].clone()
        loss_of = lambda Te: ((interior(run, Te)[..., alts] - target) ** 2).sum()

        def loss_change(Tp, Tm):
            # loss(Tp) - loss(Tm) formed element by element: the difference of the two
            # sums would lose the change a one-column direction makes to rounding
            a, b = interior(run, Tp)[..., alts], interior(run, Tm)[..., alts]
            return ((a - b) * ((a + b) - 2.0 * target)).sum()
        k = k0.clone().requires_grad_(True)
        q = q0.clone().requires_grad_(True)
        loss_of(electron_temperature(run, k, q)).backward()
        gk, gq = k.grad.clone(), q.grad.clone()
        results = []
        with torch.no_grad():
            for d in range(6):
                dk = torch.randn(sh, generator=gen, device=DEV, dtype=torch.float64)
                dq = 1e-5 * torch.randn(sh, generator=gen, device=DEV, dtype=torch.float64)
                if d >= 2:                                  # confined to one column
                    mask = torch.zeros(sh, dtype=torch.float64, device=DEV)
                    mask[int(torch.randint(0, sh[0], (1,), generator=gen, device=DEV)),
                         int(torch.randint(0, sh[1], (1,), generator=gen, device=DEV))] = 1.0
                    dk, dq = dk * mask, dq * mask
                e = 1e-6
                gp = run.recorded_steps(k0 + e * dk, q0 + e * dq)
                gm = run_m.recorded_steps(k0 - e * dk, q0 - e * dq)
                flips = 0
                for _ in range(50):
                    _, tp = next(gp)
                    _, tm = next(gm)
                    fl = branch_flips(run, tp, tm)
                    flips += fl["x1"] + fl["x2"] + fl["x3"] + fl["floor"] + fl["pivot"]
                    assert fl.get("upwind_x1", 0) + fl.get("upwind_x2", 0) + fl.get("upwind_x3", 0) == 0
                for it in (gp, gm):
                    for _ in it:
                        pass
                fd = loss_change(run.last_Te, run_m.last_Te).item() / (2 * e)
                ad = ((gk * dk).sum() + (gq * dq).sum()).item()
                rel = abs(fd - ad) / abs(fd)
                results.append((d, flips, rel))
                print(f"  direction {d} ({'column' if d >= 2 else 'random'}): relative difference "
                      f"{rel:.2e}, {flips} branch flips", flush=True)
        clean = [r for _, f, r in results if f == 0]
        assert clean, "every direction crossed a branch"
        assert max(clean) <= 1e-6, results
    finally:
        run.close()
        run_m.close()
