"""GPU parity of the rest of fluid_adv's transport block: clean_param, VNRicht_artvisc,
RK2_prep_mpi_allspec, RK2_global_boundary_allspec and compression through the C ABI, the
reference's call sequence of src/libgemini.cpp:155-210, and the fused block step, all
bit-identical to the oracle (with its correctly rounded x**2.5)."""
import numpy as np
import pytest

import gemini_amd as ga
from gemini_amd.transport import Transport, GemhipError
import oracle

pytestmark = pytest.mark.gpu


def case(kind, seed=2):
    if kind == "cart2d":
        g = ga.cartesian_grid(64, 16, 1)
    elif kind == "cart":
        g = ga.cartesian_grid(24, 10, 7)
    elif kind == "periodic":
        g = ga.cartesian_grid(20, 6, 8, flagper=True)
    else:
        g = ga.dipole_grid(32, 12, 6)
    st = ga.synthetic_state(g, seed=seed, perturb=g.lx3 > 1)
    alt = np.asarray(g.planes["alt"][0])[2:-2, 2:-2].T          # (lx1, lx2)
    nul = np.ascontiguousarray(np.broadcast_to((alt < 200e3)[:, :, None],
                                               (g.lx1, g.lx2, g.lx3))).astype(np.int32)
    nul = np.asfortranarray(nul)
    return g, st, nul


def advected(g, st, dt):
    """A state as the block sees it after the sweeps (the oracle's advection part)."""
    s = st.copy()
    for name, kw in (("v12rhov1", {}), ("T2rhoe", {}), ("set_global_boundaries_allspec", {}),
                     ("halo_fluidvars", {}), ("interface_vels_allspec", {})):
        oracle.op(g, s, name, **kw)
    for d in (3, 1):
        oracle.op(g, s, "sweep", dt=dt, direction=d, which="allparams")
    oracle.op(g, s, "halo_allparams")
    oracle.op(g, s, "sweep", dt=dt, direction=2, which="allparams")
    oracle.op(g, s, "rhov12v1")
    return s


@pytest.mark.parametrize("kind", ["cart", "periodic", "dipole"])
def test_each_block_entry_point(kind):
    print(f"case {kind}", flush=True)
    g, st, nul = case(kind)
    dt = ga.cfl_dt(g, st)
    base = advected(g, st, dt)
    tr = Transport(g, st.nn, st.Tn)
    tr.register_null_cells(nul)
    Q = np.zeros((g.lx1, g.lx2, g.lx3, 7), order="F")
    for p in (1, 2, 3):
        a, b = base.copy(), base.copy()
        oracle.op(g, a, "clean_param", which=p, nullpts=nul)
        tr.clean_param(b, p)
        assert np.array_equal(a.fluidvars, b.fluidvars), f"clean_param {p}"
    a, b = base.copy(), base.copy()
    for s, run in ((a, "oracle"), (b, "gpu")):
        if run == "oracle":
            oracle.op(g, s, "RK2_prep_mpi_allspec")
            oracle.op(g, s, "RK2_global_boundary_allspec")
            oracle.op(g, s, "VNRicht_artvisc", Q=Q)
            oracle.op(g, s, "compression", dt=dt, Q=Q)
        else:
            tr.RK2_prep_mpi_allspec(s)
            tr.RK2_global_boundary_allspec(s)
            tr.VNRicht_artvisc(s)
            tr.compression(s, dt)
    assert np.array_equal(a.fluidvars, b.fluidvars), "RK2 boundaries"
    assert np.array_equal(a.fluidauxvars, b.fluidauxvars), "compression"
    tr.register_null_cells(None)


@pytest.mark.parametrize("kind", ["cart", "periodic", "dipole", "cart2d"])
def test_block_matches_oracle(kind):
    print(f"case {kind}", flush=True)
    g, st, nul = case(kind)
    dt = ga.cfl_dt(g, st)
    tr = Transport(g, st.nn, st.Tn)
    tr.register_null_cells(nul)
    Q = np.zeros((g.lx1, g.lx2, g.lx3, 7), order="F")
    with oracle.pow_mode(1):
        # the reference's call sequence, one C entry point at a time (residency mode 0)
        ref, got = st.copy(), st.copy()
        tr.set_interface_vels(got)
        for _ in range(2):
            oracle.fluid_adv_block(g, ref, dt, nullpts=nul, Q=Q)
            tr.fluid_adv_block_calls(got, dt)
            assert np.array_equal(got.fluidvars, ref.fluidvars)
            assert np.array_equal(got.fluidauxvars[..., :14], ref.fluidauxvars[..., :14])
        # the fused block on the HBM-resident state
        ref, got = st.copy(), st.copy()
        tr.set_residency(1)
        tr.push(got)
        for _ in range(3):
            oracle.fluid_adv_block(g, ref, dt, nullpts=nul, Q=Q)
            tr.fluid_adv_block(dt)
        tr.pull(got)
        tr.set_residency(0)
    assert np.array_equal(got.fluidvars, ref.fluidvars)
    assert np.array_equal(got.fluidauxvars[..., :14], ref.fluidauxvars[..., :14])
    print(f"  Ts range {ref.Ts.min():.1f} .. {ref.Ts.max():.1f}, "
          f"{int(nul.sum())} null cells", flush=True)
    tr.register_null_cells(None)


def test_block_errors_are_reported():
    g, st, nul = case("cart")
    tr = Transport(g, st.nn, st.Tn)
    s = st.copy()
    with pytest.raises(GemhipError, match="invalid parameter"):
        tr.clean_param(s, 4)
    with pytest.raises(GemhipError, match="residency mode 1"):
        tr.fluid_adv_block(0.1)


@pytest.mark.parametrize("kind", ["cart", "dipole"])
def test_dt_select_matches_oracle(kind):
    """dt_select_C reduces the grid rates on the device; maxima are exact, so the rates and
    every dt of a sequence of calls are bit-identical to the oracle's."""
    g, st, nul = case(kind)
    tr = Transport(g, st.nn, st.Tn)
    tr.register_cfl(0.8, 1)
    tr.set_residency(1)
    tr.push(st.copy())
    _, ref_rates = oracle.dt_calc(g, st, 0.8)
    assert np.array_equal(tr.grid_rates(), ref_rates)
    tr.set_residency(0)
    dt_gpu = dt_ref = 1e-3
    free, _ = oracle.dt_calc(g, st, 0.8)
    for it, (t, tout, tglow) in enumerate([(0.0, 1e9, 1e9), (1.0, 1e9, 1e9),
                                           (2.0, 2.0 + free / 3, 1e9), (3.0, 1e9, 3.0 + free / 4),
                                           (4.0, 4.0, 1e9)], start=1):
        dt_gpu = tr.dt_select(st, t, tout, tglow, dt_gpu)
        dt_ref = oracle.dt_select(g, st, 0.8, 1, t, tout, tglow, it, dt_ref)
        assert dt_gpu == dt_ref, (it, dt_gpu, dt_ref)


def test_grid_rates_when_threads_loop_over_cells():
    """More interior cells (327 680) than the reduction's 1024 x 256 threads: each thread
    takes several cells; the 21 maxima are still the oracle's, bit for bit."""
    g = ga.cartesian_grid(64, 64, 80)
    st = ga.synthetic_state(g, seed=6, perturb=True)
    tr = Transport(g, st.nn, st.Tn)
    tr.set_residency(1)
    tr.push(st.copy())
    _, ref_rates = oracle.dt_calc(g, st, 0.9)
    assert np.array_equal(tr.grid_rates(), ref_rates)
    tr.set_residency(0)
