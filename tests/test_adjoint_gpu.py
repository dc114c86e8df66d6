"""GPU: the fitting extension and the hand-written HIP adjoint, through the C ABI.

* forward with a conductivity scale and a top heat flux: bit-identical to the oracle's
  extension (with the oracle's correctly rounded x**2.5);
* a recorded step (adjoint tape) is the same step;
* the HIP adjoint against the NumPy FP64 discrete adjoint (oracle/adjoint_np.py, itself
  pinned by central differences on the CPU), and against central differences of the
  HIP forward; checkpoint spacing does not change a bit of the gradient;
* the torch.autograd.Function gives the same gradient as the run's backward.
"""
import numpy as np
import pytest
import torch

import gemini_amd as ga
from gemini_amd.adjoint import TransportRun, electron_temperature, interior
from gemini_amd.transport import GemhipError, Transport
import oracle
from oracle import adjoint_np as A

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def case(kind, seed=3):
    if kind == "cart2d":
        g = ga.cartesian_grid(32, 8, 1)
        st = ga.synthetic_state(g, seed=seed)
        dt = min(ga.cfl_dt(g, st), 0.5)
        rng = np.random.default_rng(seed)
        kappa = 1.0 + 0.2 * rng.uniform(-1.0, 1.0, (g.lx2, g.lx3))
        qtop = 1e-5 * rng.uniform(0.0, 1.0, (g.lx2, g.lx3))
        return g, st, dt, kappa, qtop, rng
    if kind == "cart":
        g = ga.cartesian_grid(24, 6, 5)
    elif kind == "periodic":
        g = ga.cartesian_grid(20, 5, 6, flagper=True)
    else:
        g = ga.dipole_grid(24, 6, 4)
    st = ga.synthetic_state(g, seed=seed, perturb=True)
    dt = min(ga.cfl_dt(g, st), 0.5)
    rng = np.random.default_rng(seed)
    kappa = 1.0 + 0.2 * rng.uniform(-1.0, 1.0, (g.lx2, g.lx3))
    qtop = 1e-5 * rng.uniform(0.0, 1.0, (g.lx2, g.lx3)) if g.gridflag == 2 else None
    return g, st, dt, kappa, qtop, rng


def col_tensor(a):
    """(lx2, lx3) numpy -> lx2*lx3 device tensor, x2 fastest (shape (lx3, lx2))."""
    return None if a is None else torch.tensor(np.ascontiguousarray(a.T), device=DEV)


@pytest.mark.parametrize("kind", ["cart", "periodic", "dipole"])
def test_extension_forward_matches_oracle(kind):
    print(f"case {kind}", flush=True)
    g, st, dt, kappa, qtop, _ = case(kind)
    tr = Transport(g, st.nn, st.Tn)
    tr.set_residency(1)
    tr.set_conductivity_scale(kappa)
    tr.set_top_heat_flux(qtop)
    ref, got = st.copy(), st.copy()
    tr.push(got)
    with oracle.pow_mode(1):
        for _ in range(3):
            oracle.transport_step(g, ref, dt, kappa=kappa, qtop=qtop)
            tr.transport_step(dt)
    tr.pull(got)
    assert np.array_equal(got.fluidvars, ref.fluidvars)
    tr.set_conductivity_scale(None)
    tr.set_top_heat_flux(None)
    tr.set_residency(0)


def test_record_step_is_the_transport_step():
    g, st, dt, kappa, qtop, _ = case("cart")
    tr = Transport(g, st.nn, st.Tn)
    tr.set_residency(1)
    tr.set_conductivity_scale(kappa)
    tr.set_top_heat_flux(qtop)
    a, b = st.copy(), st.copy()
    tape = torch.zeros(tr.adjoint_tape_size(), dtype=torch.float64, device=DEV)
    tr.push(a)
    for _ in range(2):
        tr.transport_step(dt)
    tr.pull(a)
    tr.push(b)
    for _ in range(2):
        tr.record_step(dt, tape.data_ptr())
    tr.pull(b)
    assert np.array_equal(a.fluidvars, b.fluidvars)
    # the tape holds the diffusion's input temperature: rhoe2T of the sweep output
    N = int(np.prod(g.padded))
    TD = tape[5 * N:6 * N].cpu().numpy()
    assert np.all(TD >= 100.0) and np.isfinite(TD).all()
    tr.set_conductivity_scale(None)
    tr.set_top_heat_flux(None)
    tr.set_residency(0)


@pytest.mark.parametrize("kind", ["cart", "periodic", "dipole", "cart2d"])
def test_adjoint_matches_numpy_adjoint(kind):
    print(f"case {kind}", flush=True)
    g, st, dt, kappa, qtop, rng = case(kind)
    nsteps = 3
    w = rng.standard_normal(g.padded)
    gk_np, gq_np, lam_np, _ = A.gradient(g, st, dt, nsteps, kappa, qtop, w)
    run = TransportRun(g, st, dt, nsteps, every=2)
    try:
        run.forward(col_tensor(kappa), col_tensor(qtop))
        gk, gq = run.backward(torch.tensor(np.asfortranarray(w).ravel(order="F"), device=DEV))
        gk = gk.cpu().numpy().reshape(g.lx3, g.lx2).T
        rel = np.max(np.abs(gk - gk_np)) / np.max(np.abs(gk_np))
        print(f"  dL/dkappa rel {rel:.2e}", flush=True)
        assert rel <= 1e-9
        if qtop is not None:
            gq = gq.cpu().numpy().reshape(g.lx3, g.lx2).T
            relq = np.max(np.abs(gq - gq_np)) / np.max(np.abs(gq_np))
            print(f"  dL/dq rel {relq:.2e}", flush=True)
            assert relq <= 1e-9
        lam = run.grad_Te0.cpu().numpy().reshape(g.padded, order="F")
        rell = np.max(np.abs(lam - lam_np)) / np.max(np.abs(lam_np))
        print(f"  dL/dTe0 rel {rell:.2e}", flush=True)
        assert rell <= 1e-9
    finally:
        run.close()


def test_checkpoint_spacing_does_not_change_the_gradient():
    g, st, dt, kappa, qtop, rng = case("cart", seed=5)
    w = torch.tensor(rng.standard_normal(int(np.prod(g.padded))), device=DEV)
    out = []
    for every in (1, 3, 7):
        run = TransportRun(g, st, dt, 7, every=every)
        try:
            run.forward(col_tensor(kappa), col_tensor(qtop))
            gk, gq = run.backward(w)
            out.append((gk.cpu().numpy(), gq.cpu().numpy()))
        finally:
            run.close()
    for gk, gq in out[1:]:
        assert np.array_equal(gk, out[0][0]) and np.array_equal(gq, out[0][1])


def test_gradient_matches_central_differences():
    """Loss: squared mismatch of Te at 8 fixed altitudes against a run with other
    parameters; central differences (eps 1e-6) of the HIP forward on 8 directions."""
    g = ga.cartesian_grid(48, 16, 12)
    st = ga.synthetic_state(g, seed=11, perturb=True)
    dt = min(ga.cfl_dt(g, st), 0.5)
    rng = np.random.default_rng(2)
    k0 = torch.tensor(1.0 + 0.2 * rng.uniform(-1, 1, (g.lx3, g.lx2)), device=DEV)
    q0 = torch.tensor(1e-5 * rng.uniform(0, 1, (g.lx3, g.lx2)), device=DEV)
    run = TransportRun(g, st, dt, 12, every=5)
    try:
        alts = torch.tensor(np.linspace(2, g.lx1 - 3, 8).astype(int), device=DEV)
        target = interior(run, run.forward(k0 * 0.9 + 0.1, q0 * 0.5))[..., alts].clone()

        def loss(k, q):
            Te = interior(run, electron_temperature(run, k, q))
            return ((Te[..., alts] - target) ** 2).sum()

        k = k0.clone().requires_grad_(True)
        q = q0.clone().requires_grad_(True)
        L = loss(k, q)
        L.backward()
        gk, gq = k.grad.clone(), q.grad.clone()
        worst = 0.0
        with torch.no_grad():
            for _ in range(8):
                dk = torch.randn_like(k0)
                dq = 1e-5 * torch.randn_like(q0)
                e = 1e-6
                fd = (loss(k0 + e * dk, q0 + e * dq) - loss(k0 - e * dk, q0 - e * dq)).item() / (2 * e)
                ad = ((gk * dk).sum() + (gq * dq).sum()).item()
                worst = max(worst, abs(fd - ad) / abs(fd))
        print(f"  worst relative difference to central differences: {worst:.2e}", flush=True)
        assert worst <= 1e-6
    finally:
        run.close()


def test_adjoint_errors_are_reported():
    g = ga.dipole_grid(24, 6, 4)
    st = ga.synthetic_state(g, seed=1)
    tr = Transport(g, st.nn, st.Tn)
    with pytest.raises(GemhipError):
        tr.set_top_heat_flux(np.zeros((g.lx2, g.lx3)))       # top flux needs gridflag 2
    tr.configure(diffsolvetype=1)
    tr.set_residency(1)
    tr.push(st.copy())
    tr.set_conductivity_scale(np.ones((g.lx2, g.lx3)))
    with pytest.raises(GemhipError):
        tr.transport_step(0.1)                               # extension needs TRBDF2
    tr.set_conductivity_scale(None)
    tr.configure(diffsolvetype=2)
    tr.set_residency(0)
