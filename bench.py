#!/usr/bin/env python3
"""Benchmark of the MI355X transport step (advection sweeps + TRBDF2, FP64).

Default workload: the production-size closed dipole grid lx1=512 x lx2=256 x lx3=192
(25.2M cells, 7 species), CFL-limited dt, synthetic inputs generated in HBM.  One
"step" is the reference's hot path for one time step (src/libgemini.cpp:155-208
order): state conversions, global ghost fill, x3/x1/x2 MC advection sweeps of the 19
advected field-species, rhov12v1, rhoe2T and the TRBDF2 energy diffusion of all 7
species.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--workload dipole512|cart256|cart128|ensemble64|adjoint]

The default line also carries the whole time step (transport block + source/loss
substep) and the plasma output of the headline state (pack kernel and whole file write,
skipped if the temporary directory lacks the space), and the differentiation form's
adjoint/forward cost ratio, measured on its own configuration (ADJOINT below) with a few
central-difference checks.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0           # MI355X HBM3E
METRIC = "cell·species updates/s (adv+TRBDF2, fp64, 25M pts) vs CPU; adjoint/forward ratio"
UNIT = "cell·species updates/s"

WORKLOADS = {
    "dipole512": dict(kind="dipole", shape=(512, 256, 192), steps=50,
                      desc="closed dipole grid 512x256x192 (25.2M cells), 7 species, fp64",
                      sample=(512, 32, 16)),
    "cart256": dict(kind="cart", shape=(256, 192, 128), steps=100,
                    desc="3-D Cartesian grid 256x192x128 (6.3M cells), 7 species, fp64",
                    sample=(256, 32, 16)),
    "cart128": dict(kind="cart", shape=(128, 64, 1), steps=1,
                    desc="2-D Cartesian grid 128x64x1, 7 species, fp64", sample=(128, 64, 1)),
    # the whole transport block of fluid_adv (advection, clean-up, artificial viscosity,
    # compression, TRBDF2), null cells where the grid dips below 80 km
    "block512": dict(kind="dipole", shape=(512, 256, 192), steps=50, block=True,
                     desc="closed dipole grid 512x256x192 (25.2M cells), 7 species, fp64, "
                          "whole fluid_adv transport block", sample=(512, 32, 16)),
    # a whole time step: the transport block, then the source/loss substep (collisional and
    # chemical sources with exponential time differencing, libgemini.cpp:223-232) with a
    # synthetic neutral atmosphere and ionization rates
    "step512": dict(kind="dipole", shape=(512, 256, 192), steps=30, block=True, sources=True,
                    desc="closed dipole grid 512x256x192 (25.2M cells), 7 species, fp64, "
                         "whole time step: fluid_adv transport block and source/loss substep",
                    sample=(512, 32, 16)),
}

# Algorithmic HBM bytes per cell and launch: each array the kernel must read or write,
# once.  Sweeps: 19 advected field-species read+written, 7 velocity components and 6
# metric arrays read; TRBDF2: Ts, ns (7 species each), h1..h3, nn(1:3), J1 read, Ts
# written; conversions over the padded grid.
BYTES_PER_CELL = {"sweep1": 408, "sweep2": 408, "sweep3": 408, "trbdf2": 224,
                  "pre_convert": 280, "post_convert": 272, "rhov12v1": 160, "rhoe2T": 168,
                  # sources, per interior cell: energy reads ns, rhoes, Ts, vs1-3 (7 each),
                  # nn(1:4), Tn, vn1-3, Qeionize and writes rhoes, Ts; momentum reads ns, Ts,
                  # vs1 (7), rhovs1, vs2, vs3, Q (6), nn(1:4), Tn, vn1, E1, alt, g1, h1-h3 and
                  # writes rhovs1, vs1; mass reads ns (7), Ts (4), vs1-3 of O+, nn(1:6),
                  # Tn, vn1-3, Prionize (6) and writes ns
                  "source_energy": 520, "source_momentum": 568, "source_mass": 296}
# plasma output packing (kernels_output.hip), per interior cell: flagoutput 1 reads ns, vs1,
# Ts (7 each), vs2, vs3 (6 each), J1-J3 and writes 26 real32 values; flagoutput 2 reads 35
# doubles and writes 9 real32; flagoutput 3 reads the electron density and writes it
OUTPUT_BYTES_PER_CELL = {1: 392, 2: 316, 3: 12}
PADDED_KERNELS = {"pre_convert", "post_convert", "rhov12v1", "rhoe2T"}
KERNELS = ["pre_convert", "ghost_fill", "halo_periodic", "save_v1b", "sweep3", "sweep1",
           "sweep2", "post_convert", "rhov12v1", "rhoe2T", "trbdf2", "floor_ghosts",
           "clean_param", "artvisc", "rk2_boundary", "compression", "source_energy",
           "source_momentum", "source_mass"]


ENSEMBLE = dict(members=64, steps=20, seed=1, verify=3,
                desc="64 independent Cartesian grids, lx1 from {64,96,...,512}, lx2 x lx3 "
                     "from 16x16 to 128x96 (seeded draws), 7 species, fp64, 20 steps in one "
                     "HIP graph launch")


ADJOINT = dict(shape=(256, 128, 96), steps=50, every=10, seed=1, fd_dirs=32, eps=1e-6,
               qmax=1e-5, fd_sweep_dirs=3, fd_sweep_eps=(1e-4, 1e-5, 1e-6, 1e-7, 1e-8),
               fd_column_dirs=8,
               alts_km=(150, 200, 250, 300, 400, 500, 650, 800),
               desc="3-D Cartesian 256x128x96 (3.1M cells), 50 advection+TRBDF2 steps, "
                    "gradient of the squared Te mismatch at 8 altitudes against a seeded "
                    "target run w.r.t. the per-column electron conductivity scale and top "
                    "heat flux, checkpoints every 10 steps")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_grid(kind, shape):
    import gemini_amd as ga
    return ga.dipole_grid(*shape) if kind == "dipole" else ga.cartesian_grid(*shape)


def null_cells(g):
    """The reference's null cells (meshobj.f90:586-619): altitude below 80 km."""
    import numpy as np
    alt = np.asarray(g.planes["alt"][0])[2:-2, 2:-2].T                    # (lx1, lx2)
    m = np.empty((g.lx1, g.lx2, g.lx3), dtype=np.int32, order="F")
    m[...] = (alt < 80e3)[:, :, None]
    return m


def cpu_baseline(w, dt, seconds=15.0, block=20):
    """Oracle (C restatement, OpenMP over lines/columns) on a bounded sample grid: blocks of
    `block` steps, each from a fresh copy of the seeded state, with only the steps timed."""
    import numpy as np
    import gemini_amd as ga
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    oracle.set_num_threads(threads)
    g = make_grid(w["kind"], w["sample"])
    st0 = ga.synthetic_state(g, seed=1, perturb=(w["kind"] == "cart" and g.lx3 > 1))
    if w.get("sources"):
        nul = null_cells(g)
        atm = ga.synthetic_atmosphere(g, st0, seed=1)

        def step(s):
            Q = np.zeros((g.lx1, g.lx2, g.lx3, 7), order="F")
            oracle.fluid_adv_block(g, s, dt, nullpts=nul, Q=Q)
            oracle.sources_block(g, s, atm, dt, Q=Q, nullpts=nul)
    elif w.get("block"):
        nul = null_cells(g)
        step = lambda s: oracle.fluid_adv_block(g, s, dt, nullpts=nul)
    else:
        step = lambda s: oracle.transport_step(g, s, dt)
    step(st0.copy())                                  # warm-up
    n, el, finite = 0, 0.0, True
    while el < seconds and n < 400:
        st = st0.copy()
        t0 = time.perf_counter()
        for _ in range(block):
            step(st)
        el += time.perf_counter() - t0
        n += block
        finite = finite and bool(np.isfinite(st.fluidvars).all())
    if not finite:
        log("warning: non-finite values in the CPU sample")
    value = g.ncells * 7 * n / el
    return {"value": value, "unit": UNIT, "cores": threads, "kind": "port",
            "sample": f"oracle (plain-C restatement of the reference path, OpenMP) on a "
                      f"{g.lx1}x{g.lx2}x{g.lx3} {w['kind']} grid with the same field-line "
                      f"length and profiles: {n} steps in {el:.1f} s, in blocks of {block} "
                      f"from the seeded state, after 1 warm-up step"}


def measure_whole_step(tr, g, dt, atm, steps=10, warmup=2):
    """The whole time step -- the fluid_adv transport block, then the source/loss substep --
    on the resident state of the headline run (same grid, HBM-resident inputs)."""
    tr.register_null_cells(null_cells(g))
    tr.enable_sources(atm)

    def step():
        tr.fluid_adv_block(dt)
        tr.sources_block(dt)
    for _ in range(warmup):
        step()
    tr.synchronize()
    tr.kernel_timing_reset()
    tr.kernel_timing(True)
    tr.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    tr.synchronize()
    wall = time.perf_counter() - t0
    tr.kernel_timing(False)
    ms = 1e3 * wall / steps
    kern = {}
    for name in ("source_energy", "source_momentum", "source_mass", "trbdf2"):
        tot, n = tr.kernel_time(name)
        if n:
            kern[name] = round(tot / n, 4)
    return {"workload": WORKLOADS["step512"]["desc"], "steps": steps, "warmup": warmup,
            "ms_per_step": round(ms, 4), "value": g.ncells * 7 / (ms / 1e3), "unit": UNIT,
            "sources_ms": round(sum(kern.get(k, 0.0) for k in
                                    ("source_energy", "source_momentum", "source_mass")), 4),
            "kernel_ms": kern, "state_finite_after_timing": state_is_finite(tr, g)}


def measure_output(tr, g, electrovars, sample_shape, reps=3):
    """plasma_output_nompi_C on the resident state of the headline run: the pack kernel
    (HIP events on the library's stream) and the whole call into a file on local disk
    (pack, one device-to-host copy per dataset overlapped with the HDF5 writes); the
    oracle's packing of the same datasets on a host sample for comparison."""
    import shutil
    import tempfile
    import numpy as np
    import oracle
    from gemini_amd.state import synthetic_state
    ymd, ut = [2015, 4, 13], 12345.678910
    outdir = tempfile.mkdtemp(prefix="gemhip_output_")
    res = {"grid": f"{g.lx1}x{g.lx2}x{g.lx3}", "file_system": outdir.rsplit("/", 1)[0],
           "host_threads": int(os.environ.get("GEMHIP_IO_THREADS") or os.environ.get("OMP_NUM_THREADS")
                               or os.cpu_count())}
    try:
        # raw, then at the reference's deflate level 3 (chunks compressed on the host threads)
        for flag, comp in ((2, 0), (1, 0), (1, 3), (3, 3)):
            tr.register_output(outdir, flag, comp)
            os.remove(tr.plasma_output(None, ymd, ut, 0, electrovars=electrovars))  # staging
            walls = []
            tr.kernel_timing_reset()
            tr.kernel_timing(True)
            for _ in range(reps):
                t0 = time.perf_counter()
                path = tr.plasma_output(None, ymd, ut, 0, electrovars=electrovars)
                walls.append(time.perf_counter() - t0)
                size = os.path.getsize(path)
                os.remove(path)
            tr.kernel_timing(False)
            tot, n = tr.kernel_time("output_pack")
            pack_ms = tot / n
            nb = OUTPUT_BYTES_PER_CELL[flag] * g.ncells
            values = oracle.output_count(g.lx1, g.lx2, g.lx3, flag)
            wall = sorted(walls)[len(walls) // 2]
            res[f"flagoutput{flag}" + (f"_deflate{comp}" if comp and flag == 1 else "")] = {
                "comp_lvl": comp, "pack_ms": round(pack_ms, 4),
                "pack_algorithmic_bytes": nb,
                "pack_GB/s": round(nb / (pack_ms / 1e3) / 1e9, 1),
                "pack_frac": round(nb / (pack_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                "pack_traffic": load_traffic("dipole512", f"output_pack{flag}"),
                "real32_values": values, "file_bytes": size,
                "call_s": round(wall, 3), "call_GB/s_of_real32": round(4 * values / wall / 1e9, 3)}
            log(f"output flagoutput {flag} (deflate {comp}): pack {pack_ms:.3f} ms "
                f"({nb / (pack_ms / 1e3) / 1e9:.0f} GB/s), whole call {wall:.2f} s")
    finally:
        shutil.rmtree(outdir, ignore_errors=True)
    gs = make_grid("dipole", sample_shape)
    st = synthetic_state(gs, seed=1)
    oracle.output_pack(st, gs.gridflag, 1, patch=True)
    t0 = time.perf_counter()
    k = 0
    while time.perf_counter() - t0 < 2.0:
        oracle.output_pack(st, gs.gridflag, 1, patch=True)
        k += 1
    el = time.perf_counter() - t0
    res["cpu_pack"] = {"cells/s": round(gs.ncells * k / el), "cores": 1, "kind": "port",
                       "sample": f"oracle output packing (flagoutput 1) of a {gs.lx1}x{gs.lx2}x"
                                 f"{gs.lx3} dipole state, {k} repetitions in {el:.1f} s"}
    res["gpu_pack_cells/s"] = round(g.ncells / (res["flagoutput1"]["pack_ms"] / 1e3))
    return res


# dt_calc's grid rates (kernels_block.hip k_gridrate), per interior cell: Ts, vs1, vs2, vs3
# of the 7 species and h1, h2, h3 read once
GRIDRATE_BYTES_PER_CELL = 248


def measure_dt_select(tr, g, sample_shape, reps=20):
    """The device half of dt_select_C on the resident state of the headline run: the grid-rate
    reduction (HIP events on the library's stream) and the whole gemhip_grid_rates call
    (reduction, the 21 maxima copied back, synchronisation); the oracle's dt_calc on a host
    sample beside it."""
    import oracle
    from gemini_amd.state import synthetic_state
    tr.grid_rates()
    tr.kernel_timing_reset()
    tr.kernel_timing(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        tr.grid_rates()
    call_ms = 1e3 * (time.perf_counter() - t0) / reps
    tr.kernel_timing(False)
    tot, n = tr.kernel_time("gridrate")
    k_ms = tot / n
    nb = GRIDRATE_BYTES_PER_CELL * g.ncells
    gs = make_grid("dipole", sample_shape)
    st = synthetic_state(gs, seed=1)
    oracle.dt_calc(gs, st, 0.9)
    t0, k = time.perf_counter(), 0
    while time.perf_counter() - t0 < 1.0:
        oracle.dt_calc(gs, st, 0.9)
        k += 1
    el = time.perf_counter() - t0
    return {"kernel_ms": round(k_ms, 4), "algorithmic_bytes": nb,
            "GB/s": round(nb / (k_ms / 1e3) / 1e9, 1),
            "frac": round(nb / (k_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "call_ms": round(call_ms, 4), "calls": reps,
            "cpu_dt_calc": {"cells/s": round(gs.ncells * k / el),
                            "cores": 1,
                            "kind": "port",
                            "sample": f"oracle dt_calc on a {gs.lx1}x{gs.lx2}x{gs.lx3} dipole "
                                      f"state, {k} calls in {el:.1f} s"}}


def state_is_finite(tr, g):
    """Copy the state blocks back into HBM tensors (after timing) and check them."""
    import torch
    n = g.padded[0] * g.padded[1] * g.padded[2]
    fv = torch.empty(35 * n, dtype=torch.float64, device="cuda:0")
    aux = torch.empty(14 * n, dtype=torch.float64, device="cuda:0")
    tr.pull_blocks(fv, aux)
    ok = bool(torch.isfinite(fv).all().item()) and bool(torch.isfinite(aux).all().item())
    del fv, aux
    torch.cuda.empty_cache()
    return ok


def load_traffic(workload, kernel):
    p = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(p) as f:
            d = json.load(f)
        return d[workload][kernel]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def ensemble_member(shape, seed, own_context=True):
    """(grid, Transport with the seeded state pushed, dt) for one ensemble member."""
    from gemini_amd.ensemble import device_member
    g, tr, dt, _ = device_member(shape, seed, own_context)
    return g, tr, dt


def bench_ensemble(args):
    """Batched form: every member advanced `steps` steps by one graph launch; a few members
    re-run alone with the plain API and compared bit for bit."""
    import torch
    from gemini_amd.ensemble import Ensemble
    from gemini_amd.autotune import best_variant
    E = ENSEMBLE
    from gemini_amd.ensemble import draw_sizes
    sizes = draw_sizes(E["seed"], E["members"])
    seeds = [1000 * E["seed"] + m for m in range(len(sizes))]
    t0 = time.perf_counter()
    made = [ensemble_member(sz, sd) for sz, sd in zip(sizes, seeds)]
    grids, trs, dts = zip(*made)
    ens = Ensemble.from_transports(trs, grids, dts)
    log(f"{len(sizes)} members ready in {time.perf_counter() - t0:.1f} s")
    # autotune: every combination of kernel variants timed on each member, fastest kept
    t0 = time.perf_counter()
    choice = {}
    for tr, dt in zip(trs, dts):
        c, _ = best_variant(tr, dt, steps=1)
        for k, v in c.items():
            tr.set_tuning(k, v)
        choice[(c["sweep_pf"], c["diff_chunk"])] = choice.get((c["sweep_pf"], c["diff_chunk"]), 0) + 1
    log(f"autotuned in {time.perf_counter() - t0:.1f} s: {choice}")
    for i, (sz, sd) in enumerate(zip(sizes, seeds)):           # back to the seeded states
        from gemini_amd.synth_gpu import synthetic_state_device
        st = synthetic_state_device(grids[i], seed=sd, device="cuda:0")
        trs[i].push_blocks(st["fluidvars"], st["aux"], st["electrovars"])
        trs[i].set_interface_vels_zero()
    torch.cuda.synchronize()
    steps = E["steps"]
    t0 = time.perf_counter()
    ens.capture(steps)
    log(f"graph of {steps} steps x {len(sizes)} members captured in {time.perf_counter() - t0:.1f} s")
    ens.synchronize()
    t0 = time.perf_counter()
    ens.run(steps)
    graph_s = time.perf_counter() - t0
    # keep a few members' results, then time the same work issued without a graph
    keep = list(range(0, len(sizes), max(1, len(sizes) // E["verify"])))[:E["verify"]]
    kept = {}
    for i in keep:
        n = grids[i].padded[0] * grids[i].padded[1] * grids[i].padded[2]
        fv = torch.empty(35 * n, dtype=torch.float64, device="cuda:0")
        aux = torch.empty(14 * n, dtype=torch.float64, device="cuda:0")
        trs[i].pull_blocks(fv, aux)
        kept[i] = (fv, aux)
    t0 = time.perf_counter()
    ens.run_plain(steps)
    plain_s = time.perf_counter() - t0

    # batched kernels: one persistent launch per kernel of the step for all members
    def reseed():
        from gemini_amd.synth_gpu import synthetic_state_device
        for i, sd in enumerate(seeds):
            st = synthetic_state_device(grids[i], seed=sd, device="cuda:0")
            trs[i].push_blocks(st["fluidvars"], st["aux"], st["electrovars"])
            trs[i].set_interface_vels_zero()
            del st
        torch.cuda.synchronize()
    ens.synchronize()
    reseed()
    batch = ens.batched()
    t0 = time.perf_counter()
    bchoice, btimes = batch.autotune(reseed, steps=2)
    log(f"batched kernels autotuned in {time.perf_counter() - t0:.1f} s: {bchoice}")
    batch.capture(steps)
    batch.synchronize()
    t0 = time.perf_counter()
    batch.run(steps)
    batched_s = time.perf_counter() - t0
    batch_same = 0
    for i in keep:
        n = grids[i].padded[0] * grids[i].padded[1] * grids[i].padded[2]
        fv = torch.empty(35 * n, dtype=torch.float64, device="cuda:0")
        aux = torch.empty(14 * n, dtype=torch.float64, device="cuda:0")
        trs[i].pull_blocks(fv, aux)
        batch_same += int(torch.equal(fv, kept[i][0]) and torch.equal(aux, kept[i][1]))
    log(f"batched graph: {1e3 * batched_s / steps:.3f} ms/step, "
        f"{batch_same}/{len(keep)} kept members equal to the member graph's")
    batch.close()
    ens.close()
    verified = 0
    for i in keep:                                              # plain solve of member i alone
        g, tr, dt = ensemble_member(sizes[i], seeds[i], own_context=True)
        tr.set_interface_vels_zero()
        for _ in range(steps):
            tr.transport_step(dt)
        n = g.padded[0] * g.padded[1] * g.padded[2]
        fv = torch.empty(35 * n, dtype=torch.float64, device="cuda:0")
        aux = torch.empty(14 * n, dtype=torch.float64, device="cuda:0")
        tr.pull_blocks(fv, aux)
        same = torch.equal(fv, kept[i][0]) and torch.equal(aux, kept[i][1])
        verified += int(same)
        if not same:
            log(f"member {i} differs from its plain solve")
        tr.close()
    cells = sum(a * b * c for a, b, c in sizes)
    best_s = min(graph_s, batched_s) if batch_same == len(keep) else graph_s
    value = cells * 7 * steps / best_s
    out = {
        "metric": METRIC, "value": value, "unit": UNIT, "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": round(1e3 * best_s / steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded profiles generated in HBM)",
        "config": {"workload": E["desc"], "members": len(sizes), "cells": cells,
                   "form": "batched kernels" if best_s == batched_s else "member graph",
                   "graph_s": round(graph_s, 4), "plain_issue_s": round(plain_s, 4),
                   "batched_graph_s": round(batched_s, 4),
                   "batched_autotune": bchoice,
                   "batched_equal_to_member_graph": f"{batch_same}/{len(keep)}",
                   "autotune_choices": {f"sweep_pf={k[0]},diff_chunk={k[1]}": v
                                        for k, v in choice.items()},
                   "members_verified_bit_identical": f"{verified}/{len(keep)}"},
    }
    print(json.dumps(out), flush=True)


def bench_adjoint(fd_dirs, column_dirs=None):
    """Differentiation form: forward run with checkpoints, then the hand-written adjoint
    (replay of each segment with the tape, adjoint kernels in reverse), through the
    torch.autograd front end.  Returns the measurements as a dict.

    The central-difference check runs the +eps and -eps forwards in lockstep with the
    adjoint tape recorded, and counts the branches of the electron energy path that differ
    between them (MC limiter choices per sweep, upwind sides, the 100 K floor, the band
    solves' row interchanges): the loss is
    only piecewise smooth, and the adjoint is the derivative of the map with its branches
    frozen.  Besides the configured random directions it checks directions confined to one
    (x2, x3) column."""
    import torch
    from gemini_amd.synth_gpu import synthetic_state_device, cfl_dt_device
    from gemini_amd.adjoint import TransportRun, electron_temperature, interior, branch_flips
    A = ADJOINT
    column_dirs = A["fd_column_dirs"] if column_dirs is None else column_dirs
    t0 = time.perf_counter()
    g = make_grid("cart", A["shape"])
    st = synthetic_state_device(g, seed=A["seed"], device="cuda:0")
    dt = cfl_dt_device(g, st)
    run = TransportRun(g, st, dt, A["steps"], every=A["every"], geometry_on_device=True)
    # a second run of the same setup: the -eps side of the lockstep difference check
    run_m = TransportRun(g, st, dt, A["steps"], every=A["every"], geometry_on_device=True) \
        if fd_dirs or column_dirs else None
    del st
    torch.cuda.empty_cache()
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev).manual_seed(A["seed"])
    sh = (g.lx3, g.lx2)
    rnd = lambda: torch.rand(sh, generator=gen, device=dev, dtype=torch.float64)
    k0, q0 = 1.0 + 0.2 * (2.0 * rnd() - 1.0), A["qmax"] * rnd()
    kt, qt = 1.0 + 0.2 * (2.0 * rnd() - 1.0), A["qmax"] * rnd()
    z = g.x1[2:g.lx1 + 2]
    alts = torch.tensor([int(abs(z - 1e3 * a).argmin()) for a in A["alts_km"]], device=dev)
    target = interior(run, run.forward(kt, qt))[..., alts].clone()

    def loss_of(Te):
        return ((interior(run, Te)[..., alts] - target) ** 2).sum()

    def loss_change(Tp, Tm):
        """loss(Tp) - loss(Tm) as sum (a - b)(a + b - 2 target): the difference is formed
        element by element, not as the difference of two large sums (whose rounding would
        swamp the change a direction confined to one column makes)."""
        a, b = interior(run, Tp)[..., alts], interior(run, Tm)[..., alts]
        return ((a - b) * ((a + b) - 2.0 * target)).sum()

    def loss(k, q):
        return loss_of(electron_temperature(run, k, q))

    # warm-up through the same path (autograd's backward runs on its own device thread;
    # the first library calls there and the adjoint work buffers are set up here)
    kw = k0.clone().requires_grad_(True)
    loss(kw, q0.clone().requires_grad_(True)).backward()
    log(f"adjoint workload ready in {time.perf_counter() - t0:.1f} s, dt = {dt:.4g} s")

    k = k0.clone().requires_grad_(True)
    q = q0.clone().requires_grad_(True)
    run.tr.kernel_timing_reset()
    run.tr.kernel_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L = loss(k, q)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    L.backward()
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    t0 = time.perf_counter()                          # the same backward called directly
    run.backward(torch.zeros(run.N, dtype=torch.float64, device=dev))
    torch.cuda.synchronize()
    t_adj_direct = time.perf_counter() - t0
    run.tr.kernel_timing(False)
    kern = {}
    for name in ("adj_trbdf2", "adj_advection", "trbdf2", "sweep1", "sweep2", "sweep3"):
        tot, n = run.tr.kernel_time(name)
        if n:
            kern[name] = {"ms": round(tot / n, 4), "launches": n}
    gk, gq = k.grad.detach().clone(), q.grad.detach().clone()

    def central_difference(dk, dq, e):
        """(fd, branch flips summed over the run) for direction (dk, dq) and step e."""
        gp = run.recorded_steps(k0 + e * dk, q0 + e * dq)
        gm = run_m.recorded_steps(k0 - e * dk, q0 - e * dq)
        tot = {}
        for _ in range(A["steps"]):
            _, tp = next(gp)
            _, tm = next(gm)
            for key, v in branch_flips(run, tp, tm).items():
                tot[key] = tot.get(key, 0) + v
        for it in (gp, gm):
            for _ in it:
                pass
        fd = loss_change(run.last_Te, run_m.last_Te).item() / (2 * e)
        return fd, tot

    dirs, sweep = [], {}
    with torch.no_grad():
        for d in range(fd_dirs + column_dirs):
            dk = torch.randn(sh, generator=gen, device=dev, dtype=torch.float64)
            dq = A["qmax"] * torch.randn(sh, generator=gen, device=dev, dtype=torch.float64)
            kind = "random"
            if d >= fd_dirs:                               # one (x2, x3) column only
                kind = "column"
                i3, i2 = (int(torch.randint(0, n, (1,), generator=gen, device=dev).item())
                          for n in sh)
                mask = torch.zeros(sh, dtype=torch.float64, device=dev)
                mask[i3, i2] = 1.0
                dk, dq = dk * mask, dq * mask
            ad = ((gk * dk).sum() + (gq * dq).sum()).item()
            # the configured eps on every direction; on the first few, smaller and larger
            # steps too; on column directions eps/10, 10 eps and 100 eps beside it (rounding
            # in the difference grows as 1/eps, truncation as eps^2)
            if d < A["fd_sweep_dirs"]:
                eps_list = A["fd_sweep_eps"]
            elif kind == "column":
                eps_list = [100 * A["eps"], 10 * A["eps"], A["eps"], A["eps"] / 10]
            else:
                eps_list = [A["eps"]]
            by_eps = {}
            for e in eps_list:
                fd, fl = central_difference(dk, dq, e)
                r = abs(fd - ad) / abs(fd)
                nflip = (fl.get("x1", 0) + fl.get("x2", 0) + fl.get("x3", 0) + fl.get("floor", 0)
                         + fl.get("pivot", 0))
                by_eps[f"{e:.0e}"] = {"rel_err": float(f"{r:.3e}"), "flips": nflip}
                if e == A["eps"]:
                    dirs.append({"kind": kind, "rel_err": float(f"{r:.3e}"), "flips": nflip,
                                 "adjoint": float(f"{ad:.6e}"), "fd": float(f"{fd:.6e}"),
                                 "limiter_flips_x3_x1_x2": [fl.get("x3", 0), fl.get("x1", 0),
                                                            fl.get("x2", 0)],
                                 "floor_flips": fl.get("floor", 0),
                                 "pivot_flips": fl.get("pivot", 0),
                                 "upwind_flips": sum(v for kk, v in fl.items()
                                                     if kk.startswith("upwind"))})
            if len(eps_list) > 1:
                dirs[-1]["by_eps"] = by_eps
            if d < A["fd_sweep_dirs"]:
                for ek, v in by_eps.items():
                    sweep.setdefault(ek, []).append(v)
    run.close()
    if run_m is not None:
        run_m.close()
    res = {"forward_ms_per_step": round(1e3 * t_fwd / A["steps"], 4),
           "adjoint_ms_per_step": round(1e3 * t_adj / A["steps"], 4),
           "ratio": round(t_adj / t_fwd, 3), "loss": L.item(),
           "adjoint_ms_per_step_direct_call": round(1e3 * t_adj_direct / A["steps"], 4),
           "checkpoints": run.nseg, "tape_GB": round(run.tape_bytes() / 1e9, 2),
           "kernels_ms": kern,
           "qtop_range_W_m2": [0.0, A["qmax"]],
           "qtop_note": "the configured 0-1e-3 W/m^2 drives the top electron temperatures "
                        "past 1e6 K within a few steps (tests/test_adjoint.py::"
                        "test_configured_top_flux_runs_away), so the fit draws 0-1e-5"}
    if dirs:
        rels = sorted(x["rel_err"] for x in dirs)
        zero = [x["rel_err"] for x in dirs if x["flips"] == 0]
        # the same directions at the best of their swept steps that crosses no branch
        best = [min(v["rel_err"] for v in x["by_eps"].values() if v["flips"] == 0)
                for x in dirs if x["flips"] == 0 and "by_eps" in x]
        res["fd_check"] = {
            "directions": len(dirs), "eps": A["eps"], "max_rel_err": rels[-1],
            "median_rel_err": rels[len(rels) // 2],
            "zero_flip_directions": len(zero),
            "max_rel_err_zero_flip": max(zero) if zero else None,
            "max_rel_err_zero_flip_best_swept_eps": max(best) if best else None,
            "per_direction": dirs, "eps_sweep": sweep}
    log(f"adjoint: forward {res['forward_ms_per_step']} ms/step, adjoint "
        f"{res['adjoint_ms_per_step']} ms/step, ratio {res['ratio']}" +
        (f", FD max rel err {res['fd_check']['max_rel_err']:.2e} over {len(dirs)} directions, "
         f"{res['fd_check']['zero_flip_directions']} without branch flips (max "
         f"{res['fd_check']['max_rel_err_zero_flip']})" if dirs else ""))
    return res, g


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="dipole512",
                    choices=sorted(WORKLOADS) + ["ensemble64", "adjoint"])
    ap.add_argument("--adjoint-fd", type=int, default=None,
                    help="central-difference directions checked by the adjoint measurement "
                         "(default 4 in the default run, 32 in the adjoint workload)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-output", action="store_true",
                    help="skip the plasma-output figures of the default run")
    ap.add_argument("--no-whole-step", action="store_true",
                    help="skip the whole-step (transport block + sources) figure of the default run")
    ap.add_argument("--host-state", action="store_true",
                    help="PCIe-inclusive rate: the state lives in pinned host memory and each "
                         "timed step pushes it to HBM, steps and pulls it back")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (3 launches) to calibrate the "
                         "FETCH_SIZE/WRITE_SIZE counters, then exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 1 << 27                                   # 2^27 doubles = 1 GiB
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the transport "
              f"path shards over x2/x3 across GPUs by design, but the library runs on one.",
              file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on a single rank", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble64":
        bench_ensemble(args)
        return
    if args.workload == "adjoint":
        import torch
        res, g = bench_adjoint(ADJOINT["fd_dirs"] if args.adjoint_fd is None else args.adjoint_fd)
        steps = ADJOINT["steps"]
        ms = res["adjoint_ms_per_step"]
        print(json.dumps({
            "metric": METRIC, "value": g.ncells * 7 / (res["forward_ms_per_step"] / 1e3),
            "unit": UNIT, "n_gpus": 1, "steps": steps, "warmup": 1,
            "ms_per_step": res["forward_ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded profiles generated in HBM)",
            "config": {"workload": ADJOINT["desc"]},
            "adjoint_forward_ratio": res["ratio"], "adjoint": res}), flush=True)
        return
    w = WORKLOADS[args.workload]
    steps = args.steps if args.steps is not None else w["steps"]

    import torch
    from gemini_amd.synth_gpu import synthetic_state_device, cfl_dt_device
    from gemini_amd.transport import Transport

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    log(f"workload {args.workload}: {w['desc']}")
    t0 = time.perf_counter()
    g = make_grid(w["kind"], w["shape"])
    # the 3-D Cartesian configuration carries Gaussian density perturbations
    st = synthetic_state_device(g, seed=args.seed, device="cuda:0",
                                perturb=(w["kind"] == "cart" and g.lx3 > 1))
    dt = cfl_dt_device(g, st)
    torch.cuda.synchronize()
    log(f"inputs ready in {time.perf_counter() - t0:.1f} s, dt = {dt:.4g} s")
    tr = Transport(g, st["nn"], st["Tn"], device=0, geometry_on_device=True)
    tr.set_residency(1)
    tr.push_blocks(st["fluidvars"], st["aux"], st["electrovars"])
    whole_step = (args.workload == "dipole512" and not args.host_state and not args.no_whole_step)
    atm = None
    if w.get("sources") or whole_step:
        from gemini_amd.synth_gpu import synthetic_atmosphere_device
        L1, L2, L3 = g.padded
        ne = st["fluidvars"][6].view(L3, L2, L1)[2:-2, 2:-2, 2:-2]
        atm = synthetic_atmosphere_device(g, st["nn"], st["Tn"], ne, seed=args.seed, device="cuda:0")
    if w.get("sources"):
        tr.enable_sources(atm)
        atm = None
    output = (args.workload == "dipole512" and not args.host_state and not args.no_output)
    electrovars = st["electrovars"] if output else None
    del st
    torch.cuda.empty_cache()
    step = tr.transport_step
    if w.get("block"):
        nul = null_cells(g)
        tr.register_null_cells(nul)
        log(f"{int(nul.sum())} null cells")
        del nul
        step = tr.fluid_adv_block
        args.no_adjoint = True
    if w.get("sources"):
        def step(dt):
            tr.fluid_adv_block(dt)
            tr.sources_block(dt)
    if args.host_state:
        # the caller keeps fluidvars, fluidauxvars and J1 on the host (pinned); every step
        # crosses PCIe twice, as with residency mode 0 around a fused step
        n = g.padded[0] * g.padded[1] * g.padded[2]
        hf = torch.empty(35 * n, dtype=torch.float64, pin_memory=True)
        ha = torch.empty(14 * n, dtype=torch.float64, pin_memory=True)
        tr.pull_blocks(hf, ha)
        inner = step

        def step(dt):
            tr.push_blocks(hf, ha)
            inner(dt)
            tr.pull_blocks(hf, ha)
        args.no_adjoint = True
        args.no_cpu_baseline = True
    log("state in HBM; warm-up")
    for _ in range(args.warmup):
        step(dt)
    tr.synchronize()
    log(f"timing {steps} steps")
    tr.kernel_timing_reset()
    tr.kernel_timing(True)
    tr.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        step(dt)
    tr.synchronize()
    wall = time.perf_counter() - t0
    tr.kernel_timing(False)
    finite = state_is_finite(tr, g)
    if not finite:
        log("warning: non-finite values in the GPU state after the timed steps")
    ms_step = 1e3 * wall / steps
    cells = g.ncells
    padded = g.padded[0] * g.padded[1] * g.padded[2]
    value = cells * 7 / (ms_step / 1e3)
    kern = {}
    for name in KERNELS:
        tot, n = tr.kernel_time(name)
        if n == 0:
            continue
        avg = tot / n
        entry = {"ms": round(avg, 4), "launches_per_step": n / steps}
        if name in BYTES_PER_CELL:
            nb = BYTES_PER_CELL[name] * (padded if name in PADDED_KERNELS else cells)
            entry["GB/s"] = round(nb / (avg / 1e3) / 1e9, 1)
        kern[name] = entry
    dom = max((k for k in kern if k in BYTES_PER_CELL), key=lambda k: kern[k]["ms"] *
              kern[k]["launches_per_step"])
    dom_bytes = BYTES_PER_CELL[dom] * (padded if dom in PADDED_KERNELS else cells)
    achieved = dom_bytes / (kern[dom]["ms"] / 1e3) / 1e9
    traffic = load_traffic(args.workload, dom)
    log(f"{ms_step:.3f} ms/step, {value:.3e} cell-species/s; dominant kernel {dom} "
        f"{kern[dom]['ms']:.3f} ms ({achieved:.0f} GB/s)")
    # the TRBDF2 launch is one kernel for all species; report its step share as well
    step_bytes = 1448 * cells
    whole = None
    if atm is not None:
        log("whole time step: transport block + source/loss substep")
        whole = measure_whole_step(tr, g, dt, atm)
        del atm
        log(f"whole step {whole['ms_per_step']:.3f} ms, sources {whole['sources_ms']:.3f} ms")
    dtsel = None
    if args.workload == "dipole512" and not args.host_state:
        log("dt_select's grid-rate reduction on the resident state")
        dtsel = measure_dt_select(tr, g, w["sample"])
        log(f"grid rates {dtsel['kernel_ms']:.3f} ms ({dtsel['GB/s']:.0f} GB/s), "
            f"call {dtsel['call_ms']:.3f} ms")
    out_res = None
    if output:
        log("plasma output (flagoutput 2, 1, 3) of the resident state")
        import shutil
        import tempfile
        free = shutil.disk_usage(tempfile.gettempdir()).free
        need = 4 * 26 * g.ncells * 2            # the largest file, twice over
        if free < need:
            out_res = {"skipped": f"{free / 1e9:.1f} GB free under {tempfile.gettempdir()}, "
                                  f"{need / 1e9:.1f} GB wanted"}
            log("output figures skipped: " + out_res["skipped"])
        else:
            try:
                out_res = measure_output(tr, g, electrovars, w["sample"])
            except Exception as e:              # the file system must not cost the step figure
                out_res = {"error": f"{type(e).__name__}: {e}"}
                log("output figures failed: " + out_res["error"])
        del electrovars
    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle on host cores)")
        cpu = cpu_baseline(w, dt, args.cpu_seconds)
        if whole is not None:
            log("CPU baseline of the whole step (oracle transport block + sources)")
            whole["cpu_baseline"] = cpu_baseline(WORKLOADS["step512"], dt, 10.0, block=5)
    adj = None
    if not args.no_adjoint:
        del tr
        torch.cuda.empty_cache()
        log("adjoint measurement (" + ADJOINT["desc"] + ")")
        adj, _ = bench_adjoint(ADJOINT["fd_dirs"] if args.adjoint_fd is None else args.adjoint_fd,
                               ADJOINT["fd_column_dirs"])
    out = {
        "metric": METRIC,
        "value": value,
        "unit": UNIT,
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded profiles generated in HBM)",
        "config": {"workload": w["desc"], "grid": f"{w['kind']} {g.lx1}x{g.lx2}x{g.lx3}",
                   "species": 7, "dt_s": dt, "diffusion": "TRBDF2",
                   "step": ("v12rhov1,T2rhoe,ghost fill,sweep3,sweep1,sweep2,rhov12v1,"
                            "clean_param 1-2,VNRicht_artvisc,RK2 boundaries,compression,rhoe2T,"
                            "clean_param 3,TRBDF2,clean_param 3,T2rhoe" + (
                                ",ghost fill,source_loss energy,momentum,mass,clean_param 3,2,1"
                                if w.get("sources") else "")) if w.get("block") else
                           "v12rhov1,T2rhoe,ghost fill,sweep3,sweep1,sweep2,rhov12v1,rhoe2T,"
                           "TRBDF2", "state_finite_after_timing": finite},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": dom_bytes},
        "step_hbm": {"algorithmic_bytes": step_bytes,
                     "GB/s": round(step_bytes / (ms_step / 1e3) / 1e9, 1),
                     "frac": round(step_bytes / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
        "kernels": kern,
        "cpu_baseline": cpu,
        "whole_step": whole,
        "dt_select": dtsel,
        "output": out_res,
        "adjoint_forward_ratio": None if adj is None else adj["ratio"],
        "adjoint": None if adj is None else dict(adj, workload=ADJOINT["desc"]),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
