"""The fitting extension and the NumPy FP64 discrete adjoint, on the CPU.

The NumPy adjoint (oracle/adjoint_np.py) is the checker of the HIP adjoint.  It is
pinned here by central finite differences of the C oracle's forward, and its plain
NumPy restatements (used for complex-step Jacobians) are checked against the oracle.
The extension itself (conductivity scale, Neumann top row with a heat flux) has no
counterpart in the reference; it is pinned by identities the mathematics fixes.
"""
import numpy as np
import pytest

import gemini_amd as ga
import oracle
from oracle import adjoint_np as A


def _case(kind, seed=3):
    if kind == "cart":
        g = ga.cartesian_grid(16, 4, 3)
    elif kind == "periodic":
        g = ga.cartesian_grid(16, 4, 5, flagper=True)
    else:
        g = ga.dipole_grid(20, 4, 3)
    st = ga.synthetic_state(g, seed=seed, perturb=True)
    dt = min(ga.cfl_dt(g, st), 0.5)
    rng = np.random.default_rng(seed)
    kappa = 1.0 + 0.2 * rng.uniform(-1.0, 1.0, (g.lx2, g.lx3))
    qtop = 1e-5 * rng.uniform(0.0, 1.0, (g.lx2, g.lx3)) if g.gridflag == 2 else None
    return g, st, dt, kappa, qtop, rng


@pytest.mark.parametrize("kind", ["cart", "periodic", "dipole"])
def test_numpy_adjoint_matches_central_differences(kind):
    g, st, dt, kappa, qtop, rng = _case(kind)
    nsteps = 3
    w = rng.standard_normal(g.padded)                       # L = sum(w * Te_final)
    gk, gq, _, _ = A.gradient(g, st, dt, nsteps, kappa, qtop, w)

    def loss(k, q):
        return float((A.final_Te(g, st, dt, nsteps, k, q) * w).sum())

    eps = 1e-6
    for _ in range(3):
        dk = rng.standard_normal(kappa.shape)
        dq = None if qtop is None else 1e-5 * rng.standard_normal(kappa.shape)
        qp = None if qtop is None else qtop + eps * dq
        qm = None if qtop is None else qtop - eps * dq
        fd = (loss(kappa + eps * dk, qp) - loss(kappa - eps * dk, qm)) / (2 * eps)
        ad = float((gk * dk).sum() + (0.0 if qtop is None else (gq * dq).sum()))
        scale = max(abs(fd), 1e-3 * (np.abs(gk).sum() + (0 if qtop is None else 1e-5 * np.abs(gq).sum())))
        assert abs(fd - ad) <= 1e-6 * scale, (kind, fd, ad)


def test_numpy_sweep_restatement_matches_oracle():
    rng = np.random.default_rng(0)
    lx = 12
    f = rng.uniform(1.0, 2.0, lx + 4)
    v = rng.uniform(-300.0, 300.0, lx + 1)
    dx = rng.uniform(0.5e3, 2e3, lx + 3)
    dxi = rng.uniform(0.5e3, 2e3, lx)
    ha1 = rng.uniform(0.9, 1.1, lx + 4)
    ha2i = rng.uniform(0.9, 1.1, lx + 1)
    h1i = rng.uniform(0.9, 1.1, lx + 1)
    ref = oracle.advec1D(f, v, 0.7, dx, dxi, ha1, ha2i, h1i)
    got = A.sweep_lines(f[None, :], v[None, :], 0.7, dx, dxi, ha1[None, :], ha2i[None, :],
                        h1i[None, :])[0]
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("flux", [False, True])
def test_numpy_band_systems_match_oracle_trbdf2(flux):
    """Dense solves of the NumPy assembly reproduce the oracle's pivoted band TRBDF2."""
    g = ga.cartesian_grid(24, 1, 1)
    st = ga.synthetic_state(g, seed=2)
    N = g.lx1
    c = (slice(2, N + 2), 2, 2)
    kB, e, gam = A._consts()
    T = st.Ts[..., 6][c].copy()
    col = dict(ne=st.ns[..., 6][c], n1=st.nn[:, 0, 0, 0], n2=st.nn[:, 0, 0, 1],
               n3=st.nn[:, 0, 0, 2], J=st.J1[c], h1=g.h1[c], h2=g.h2[c], h3=g.h3[c],
               dx1=g.dx1, dx1i=g.dx1i, flux=flux)
    kap, q, dt = 1.13, 4e-6, 0.5
    M1, r1, M2, r2T = A.assemble(T[None, :].astype(complex), np.array([kap + 0j]),
                                 np.array([q + 0j]), col, dt, 300.0, 1500.0)
    x1 = np.linalg.solve(M1[0].real, r1[0].real)
    d = np.zeros(N)
    d[1:N - 1] = (4.0 / 3.0) / (dt / 3.0)
    x2 = np.linalg.solve(M2[0].real, r2T[0].real + d * x1)
    lam = A.lambda_e(T, col["ne"], col["n1"], col["n2"], col["n3"]).real * kap
    Cc = ((gam - 1.0) / kB / np.maximum(col["ne"], 1e-5)) / ((col["h1"] * col["h2"]) * col["h3"])
    B = (Cc * ((2.5 * kB / e) * col["J"])) / col["h1"]
    D = ((lam * col["h2"]) * col["h3"]) / col["h1"]
    z = np.zeros(N)
    if flux:
        gtop = (q * col["h1"][-1]) / lam[-1]
        ref = oracle.trbdf2_1d_flux(T, z, B, Cc, D, z, 300.0, gtop, dt, g.dx1, g.dx1i)
    else:
        ref = oracle.trbdf2_1d(T, z, B, Cc, D, z, 300.0, 1500.0, dt, g.dx1, g.dx1i)
    assert np.max(np.abs(x2 - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_neumann_top_row_steady_state():
    """Constant coefficients and a very long step: the BDF2 stage (L-stable) lands on
    the discrete steady state, T(i) = Tmin + g (x1(i) - x1(1)) -- a constant gradient g
    set by the top row, fixed by the bottom value."""
    g = ga.cartesian_grid(20, 1, 1)
    N = g.lx1
    x = g.x1[2:N + 2]
    T0 = np.linspace(400.0, 900.0, N)
    one, z = np.ones(N), np.zeros(N)
    gt = 0.004
    out = oracle.trbdf2_1d_flux(T0, z, z, one * 2.0, one * 3e6, z, 300.0, gt, 1e14, g.dx1, g.dx1i)
    ref = 300.0 + gt * (x - x[0])
    assert np.max(np.abs(out - ref)) <= 1e-8 * np.max(np.abs(ref))


def test_conductivity_scale_is_a_time_rescaling():
    """With no thermoelectric term (J1 = 0), scaling the conductivity by c is the same
    TRBDF2 step as scaling dt by c (both matrices and right-hand sides scale by 1/c)."""
    g = ga.cartesian_grid(24, 1, 1)
    st = ga.synthetic_state(g, seed=4)
    N = g.lx1
    c = (slice(2, N + 2), 2, 2)
    T = st.Ts[..., 6][c].copy()
    z = np.zeros(N)
    Cc = np.linspace(1e3, 5e3, N)
    D = np.linspace(1e-3, 8e-3, N)
    for s in (0.7, 1.3):
        a = oracle.trbdf2_1d(T, z, z, Cc, D * s, z, 300.0, 1500.0, 0.5, g.dx1, g.dx1i)
        b = oracle.trbdf2_1d(T, z, z, Cc, D, z, 300.0, 1500.0, 0.5 * s, g.dx1, g.dx1i)
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))


def test_extension_off_is_the_reference_step():
    """kappa = 1 everywhere multiplies by an exact one: bit-identical to the plain step."""
    g = ga.cartesian_grid(16, 4, 3)
    st = ga.synthetic_state(g, seed=6, perturb=True)
    a, b = st.copy(), st.copy()
    oracle.transport_step(g, a, 0.5)
    oracle.transport_step(g, b, 0.5, kappa=np.ones((g.lx2, g.lx3)))
    assert np.array_equal(a.fluidvars, b.fluidvars)


def test_extension_rejects_unsupported_modes():
    g = ga.dipole_grid(20, 4, 3)
    st = ga.synthetic_state(g, seed=1)
    with pytest.raises(RuntimeError):
        oracle.transport_step(g, st.copy(), 0.1, qtop=np.zeros((g.lx2, g.lx3)))
    with pytest.raises(RuntimeError):
        oracle.transport_step(g, st.copy(), 0.1, diffsolvetype=1, kappa=np.ones((g.lx2, g.lx3)))


def test_configured_top_flux_runs_away():
    """The differentiation workload names a top heat flux of 0-1e-3 W/m^2.  With the
    conductivity lagged at each step's input temperature (as every coefficient of the
    reference's step is), 1e-3 W/m^2 through the top face of the 256-cell altitude grid
    drives the top electron temperatures past 1e6 K within eight steps, while the
    physical topside range, up to 1e-5 W/m^2, stays below 4000 K.  The fit therefore
    draws the flux from 0-1e-5 W/m^2."""
    g = ga.cartesian_grid(256, 2, 2)
    peaks = {}
    for q in (1e-3, 1e-5):
        st = ga.synthetic_state(g, seed=1)
        dt = min(ga.cfl_dt(g, st), 0.5)
        for _ in range(8):
            oracle.transport_step(g, st, dt, kappa=np.ones((2, 2)), qtop=np.full((2, 2), q))
        peaks[q] = float(np.nanmax(st.Ts[2:-2, 2:-2, 2:-2, 6]))
    assert peaks[1e-3] > 1e6 and peaks[1e-5] < 4000.0, peaks


def _minmod(a, b):
    """advec.f90:438-447, and which argument it returned (0: zero, 1: a, 2: b)."""
    if a * b <= 0.0:
        return 0.0, 0
    return (a, 1) if abs(a) < abs(b) else (b, 2)


def test_limiter_codes_follow_the_reference_minmod():
    """The branch codes the adjoint's difference check compares (limiter_codes) are the
    choices advec.f90:396-401 makes, slope by slope, on a field with extrema, plateaus
    and both velocity signs; unread slopes are coded -1."""
    torch = pytest.importorskip("torch")
    from gemini_amd.adjoint import limiter_codes
    rng = np.random.default_rng(4)
    L3, L2, L1 = 7, 6, 12
    f = np.round(rng.uniform(0.0, 4.0, (L3, L2, L1)), 1)       # ties and flat runs
    for axis in (0, 1, 2):
        n = (L3, L2, L1)[axis]
        lx = n - 4
        dx = rng.uniform(0.5, 2.0, lx + 3)
        vshape = [L3 - 4, L2 - 4, L1 - 4]
        vshape[axis] = lx + 1
        v = rng.uniform(-1.0, 1.0, vshape)
        codes = limiter_codes(torch.tensor(f), torch.tensor(dx), torch.tensor(v), axis).numpy()
        g = np.moveaxis(f[tuple(slice(None) if a == axis else slice(2, -2) for a in range(3))], axis, -1)
        vv = np.moveaxis(v, axis, -1)
        for idx in np.ndindex(g.shape[:-1]):
            line, vl = g[idx], vv[idx]
            for i in range(lx + 2):                               # slope(i), i = 0..lx+1
                p = i + 1
                l = (line[p] - line[p - 1]) / dx[i]
                r = (line[p + 1] - line[p]) / dx[i + 1]
                c = (line[p + 1] - line[p - 1]) / (dx[i] + dx[i + 1])
                m, inner = _minmod(2 * l, 2 * r)
                s, outer = _minmod(c, m)
                assert s == oracle.minmod(c, oracle.minmod(2 * l, 2 * r))
                used = (i >= 1 and vl[i - 1] < 0) or (i <= lx and vl[i] >= 0)
                want = 3 * inner + outer if used else -1
                assert codes[idx + (i,)] == want, (axis, idx, i)
