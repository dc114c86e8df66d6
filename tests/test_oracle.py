"""CPU checks that pin the oracle (the C restatement used as the GPU parity checker).

Each check is independent of the oracle's own formulas: the reference's known-answer
test, a real LAPACK (scipy's dgbsv), closed forms of the discrete schemes, invariants
of the limiter and ghost fill, physical constants from their definitions, and exact
arithmetic for x**2.5.
"""
import ctypes as C
import json
import math
import os
from decimal import Decimal, getcontext

import numpy as np
import pytest
from scipy.linalg import lapack

import gemini_amd as ga
import oracle

HERE = os.path.dirname(os.path.abspath(__file__))


# ------------------------------------------------------------------ reference known answers
def _diffusion1d_run():
    """test/diffusion/test_diffusion1D.f90 restated (HDF5 output replaced by arrays)."""
    npts, lt = 256, 100
    x1 = np.array([(ix - 1) / (npts - 1) for ix in range(-1, npts + 3)], dtype=np.float64)
    lx1 = npts
    dx1 = x1[1:lx1 + 4] - x1[0:lx1 + 3]                       # x1(0:lx1+2)-x1(-1:lx1+1)
    x1i = 0.5 * (x1[1:lx1 + 2] + x1[2:lx1 + 3])               # 0.5 single precision: exact
    dx1i = x1i[1:lx1 + 1] - x1i[0:lx1]
    pi = 4.0 * math.atan(1.0)
    Ts0 = np.array([math.sin(2 * pi * v) + math.sin(8 * pi * v) for v in x1])
    TsE, TsB = Ts0[2:lx1 + 2].copy(), Ts0[2:lx1 + 2].copy()
    lam = np.ones(lx1)
    dt = float(np.float32(0.05) / np.float32(64.0)) / (pi * pi) / 1.0
    A = np.zeros(lx1); B = np.zeros(lx1); Cc = np.ones(lx1); D = lam.copy(); E = np.zeros(lx1)
    t = 0.0
    hist = {"TsEuler": [], "TsBDF2": [], "Tstrue": []}
    xin = x1[2:lx1 + 2]
    for it in range(1, lt + 1):
        TsE, _, _ = oracle.backeuler_1d(TsE, A, B, Cc, D, E, 0.0, 0.0, dt, dx1, dx1i)
        TsB = oracle.trbdf2_1d(TsB, A, B, Cc, D, E, 0.0, 0.0, dt, dx1, dx1i)
        t = t + dt
        tru = np.array([math.exp(-((4 * pi ** 2) * 1.0) * t) * math.sin(2 * pi * v)
                        + math.exp(-((64 * pi ** 2) * 1.0) * t) * math.sin(8 * pi * v) for v in xin])
        hist["TsEuler"].append(TsE.copy()); hist["TsBDF2"].append(TsB.copy()); hist["Tstrue"].append(tru)
    return {k: np.array(v).T for k, v in hist.items()}


def test_diffusion1d_known_answers():
    fx = json.load(open(os.path.join(HERE, "golden", "diffusion1d_known_answers.json")))
    h = _diffusion1d_run()
    for key in ("TsEuler", "TsBDF2", "Tstrue"):
        got = h[key][12, 12]
        ref = fx[key]["value"]
        assert np.isclose(got, ref, rtol=fx[key]["rtol"])          # the reference's own bar
        # and far tighter: the fixture was printed with 15 significant digits
        assert abs(got - ref) / abs(ref) < 2e-12, (key, got, ref)
    for key in ("TsEuler", "TsBDF2"):
        err = np.max(np.abs(h[key] - h["Tstrue"]), axis=0)
        assert np.all(err < fx["max_abs_error_bound"]["value"])
    # TRBDF2 is second order: far below backward Euler at the end of the run
    assert np.max(np.abs(h["TsBDF2"][:, -1] - h["Tstrue"][:, -1])) < 1e-4
    assert np.max(np.abs(h["TsEuler"][:, -1] - h["Tstrue"][:, -1])) > 1e-3


# ------------------------------------------------------------------ LAPACK dgbsv
def _band(sub, dia, sup):
    n = len(dia)
    ab = np.zeros((7, n), order="F")
    ab[4, :] = dia
    ab[3, 1:] = sup[:-1]
    ab[5, :-1] = sub[1:]
    return ab


def _systems(rng, count=120):
    out = []
    for _ in range(count):
        n = int(rng.integers(4, 200))
        sub = rng.normal(size=n) * rng.choice([0.1, 1.0, 10.0])
        dia = rng.normal(size=n)
        sup = rng.normal(size=n)
        dia[0], sup[0], sub[-1], dia[-1] = 1.0, 0.0, 0.0, 1.0
        out.append((sub, dia, sup, rng.normal(size=n)))
    # realistic electron columns of the transport path (C1 grid), which pivot heavily
    g = ga.cartesian_grid(128, 4, 1)
    st = ga.synthetic_state(g, seed=3)
    for i2 in range(4):
        out.append(_electron_tr_system(g, st, i2))
    return out


def _electron_tr_system(g, st, i2, dt=0.5):
    """TR-stage band of the electron column (i2, 1) assembled from oracle pieces."""
    lx1 = g.lx1
    T = st.Ts[2:lx1 + 2, i2 + 2, 2, 6]
    n = st.ns[2:lx1 + 2, i2 + 2, 2, 6]
    nn = st.nn[:, i2, 0, :]
    J1 = st.J1[2:lx1 + 2, i2 + 2, 2]
    kB, e = 1.38064852e-23, 1.60217662e-19
    sq = np.sqrt(T)
    lam = (e * 100.0 * 7.7e5) * oracle.eval_pow25(T, 0) / (1.0 + (3.22e4 * T * T) / n * (
        nn[:, 0] * 1.1e-16 * (1 + 5.7e-4 * T) + nn[:, 1] * 2.82e-17 * sq * (1 - 1.21e-4 * T)
        + nn[:, 2] * 2.2e-16 * (1 + 3.6e-2 * sq)))
    Cc = (2.0 / 3.0) / kB / np.maximum(n, 1e-5)
    B = Cc * (2.5 * kB / e * J1)
    Dh = np.zeros(lx1)
    Dh[1:] = 0.5 * (lam[:-1] + lam[1:])
    dx, dxi = g.dx1, g.dx1i
    sub, dia, sup = np.zeros(lx1), np.ones(lx1), np.zeros(lx1)
    for i in range(1, lx1 - 1):
        t1 = Cc[i] * Dh[i] / dxi[i] / dx[i + 1] / 2
        t3 = Cc[i] * Dh[i + 1] / dxi[i] / dx[i + 2] / 2
        t2 = B[i] / (dx[i + 2] + dx[i + 1]) / 2
        sub[i], dia[i], sup[i] = -t1 + t2, 1 / (dt / 2) + t3 + t1, -t3 - t2
    return sub, dia, sup, T.copy()


def test_band_solver_reproduces_lapack_bit_for_bit():
    """With fused multiply-adds in the BLAS updates (as OpenBLAS evaluates them) the
    restated dgbtf2/dgbtrs/dtbsv reproduce scipy's LAPACK dgbsv exactly: same pivots,
    same LU factors, same solution, on random and on heavily pivoting systems."""
    rng = np.random.default_rng(0)
    L = oracle.lib()
    swaps = 0
    try:
        L.orc_set_blas_fma(C.c_int(1))
        for sub, dia, sup, r in _systems(rng):
            ab = _band(sub, dia, sup)
            x, lu, ipiv, info = oracle.gbsv(ab, r)
            lub, piv, xs, infos = lapack.dgbsv(2, 2, ab.copy(), r.copy())
            assert info == infos == 0
            assert np.array_equal(ipiv, piv + 1) and np.array_equal(lu, lub)
            assert np.array_equal(x, xs)
            swaps += int(np.sum(ipiv != np.arange(1, len(dia) + 1)))
    finally:
        L.orc_set_blas_fma(C.c_int(0))
    assert swaps > 1000


def test_band_solver_default_is_netlib_arithmetic():
    """Default (reference BLAS, separately rounded products): same pivots and close to
    LAPACK, while differing in the last bits on pivoting systems."""
    rng = np.random.default_rng(1)
    differ = 0
    for sub, dia, sup, r in _systems(rng, 40)[-4:]:
        ab = _band(sub, dia, sup)
        x, lu, ipiv, info = oracle.gbsv(ab, r)
        lub, piv, xs, infos = lapack.dgbsv(2, 2, ab.copy(), r.copy())
        assert np.array_equal(ipiv, piv + 1)
        assert np.max(np.abs(x - xs) / np.abs(xs)) < 1e-11
        differ += int(not np.array_equal(x, xs))
    assert differ > 0


def test_band_solver_singular_returns_rhs():
    """dgbsv returns INFO>0 and leaves the right-hand side untouched when a pivot is 0."""
    n = 6
    ab = np.zeros((7, n), order="F")
    ab[4, :] = 1.0
    ab[4, 3] = 0.0      # zero diagonal, zero sub-diagonal below it: singular column
    r = np.arange(1.0, n + 1)
    x, lu, ipiv, info = oracle.gbsv(ab, r)
    assert info == 4 and np.array_equal(x, r)


@pytest.mark.parametrize("kind,shape", [("cart", (24, 10, 7)), ("dipole", (20, 9, 5))])
def test_energy_diffusion_cases_pivot(kind, shape):
    """The synthetic states of the parity tests drive LAPACK's partial pivoting: the
    identity Dirichlet row is tiny next to conduction-dominated interior rows, so most
    elimination steps swap rows and the fill-in U2 is in use."""
    g = ga.cartesian_grid(*shape) if kind == "cart" else ga.dipole_grid(*shape)
    st = ga.synthetic_state(g, seed=1, perturb=True)
    dt = ga.cfl_dt(g, st)
    for solver in (2, 1):
        oracle.reset_swap_count()
        oracle.op(g, st.copy(), "energy_diffusion", dt=dt, diffsolvetype=solver)
        rows = g.lx1 * g.lx2 * g.lx3 * 7 * (2 if solver == 2 else 1)
        assert oracle.swap_count() > rows // 4


# ------------------------------------------------------------------ TRBDF2 closed form
@pytest.mark.parametrize("mode", [1, 3, 7])
def test_trbdf2_sine_mode_closed_form(mode):
    """Uniform grid, zero Dirichlet ends, C=D=1: a discrete sine mode is an eigenvector
    with eigenvalue lam = -(4/dx^2) sin^2(k pi dx/2); TR then BDF2 give
    u_h = (1+dt lam/4)/(1-dt lam/4) u0 and u1 = (4 u_h - u0)/(3 - dt lam)."""
    n = 129
    x = (np.arange(-1, n + 3) - 1) / (n - 1)              # ghosted uniform grid on [0,1]
    dx = x[1:n + 4] - x[0:n + 3]
    xi = 0.5 * (x[1:n + 2] + x[2:n + 3])
    dxi = xi[1:n + 1] - xi[0:n]
    h = 1.0 / (n - 1)
    u0 = np.sin(mode * np.pi * x[2:n + 2])
    u0[0] = u0[-1] = 0.0
    dt = 2e-4
    lam = -(4.0 / h ** 2) * np.sin(mode * np.pi * h / 2) ** 2
    uh = (1 + dt * lam / 4) / (1 - dt * lam / 4)
    u1 = (4 * uh - 1.0) / (3 - dt * lam)
    z = np.zeros(n)
    got = oracle.trbdf2_1d(u0, z, z, np.ones(n), np.ones(n), z, 0.0, 0.0, dt, dx, dxi)
    # A dropped term or wrong coefficient would miss by O(dt) or more.  The pivoted solve
    # itself loses a few thousand ulps next to the boundary: LAPACK swaps the unit
    # Dirichlet row below the conduction row, which is ~1e4 times larger.
    assert np.max(np.abs(got[1:-1] - u1 * u0[1:-1])) < 1e-10
    assert np.max(np.abs(got[[0, -1]])) < 1e-10
    be, _, _ = oracle.backeuler_1d(u0, z, z, np.ones(n), np.ones(n), z, 0.0, 0.0, dt, dx, dxi)
    assert np.max(np.abs(be[1:-1] - u0[1:-1] / (1 - dt * lam))) < 1e-10



# ------------------------------------------------------------------ MC advection
def test_minmod_truth_table():
    assert oracle.minmod(1.0, 2.0) == 1.0
    assert oracle.minmod(-3.0, -2.0) == -2.0
    assert oracle.minmod(1.0, -1.0) == 0.0
    assert oracle.minmod(0.0, 5.0) == 0.0
    assert oracle.minmod(2.0, 2.0) == 2.0
    assert oracle.minmod(1e-200, 1e-200) == 0.0       # a*b underflows to 0: limiter closes


def _line(lx, rng, uniform=True):
    x = np.arange(-1, lx + 3, dtype=float) * 1.0 if uniform else np.cumsum(rng.uniform(0.5, 2.0, lx + 4))
    dx = x[1:lx + 4] - x[0:lx + 3]
    xi = 0.5 * (x[1:lx + 2] + x[2:lx + 3])
    return x, dx, xi[1:lx + 1] - xi[0:lx]


def test_advection_free_stream_and_courant_one_shift():
    rng = np.random.default_rng(0)
    lx = 40
    x, dx, dxi = _line(lx, rng)
    one = np.ones(lx + 4)
    # constant state, uniform velocity, unit metric: unchanged
    f = np.full(lx + 4, 3.25)
    out = oracle.advec1D(f, np.full(lx + 1, 0.7), 0.5, dx, dxi, one, one[:lx + 1], one[:lx + 1])
    assert np.array_equal(out, f)
    # Courant number one to the right: each interior cell takes its left neighbour
    f = rng.uniform(1.0, 2.0, lx + 4)
    out = oracle.advec1D(f, np.ones(lx + 1), 1.0, dx, dxi, one, one[:lx + 1], one[:lx + 1])
    assert np.allclose(out[2:lx + 2], f[1:lx + 1], rtol=0, atol=4e-16)
    assert np.array_equal(out[[0, 1, lx + 2, lx + 3]], f[[0, 1, lx + 2, lx + 3]])
    # single cell: copied
    f1 = rng.uniform(size=5)
    assert np.array_equal(oracle.advec1D(f1, np.ones(2), 1.0, np.ones(4), np.ones(1),
                                         np.ones(5), np.ones(2), np.ones(2)), f1)


def test_advection_conserves_and_stays_bounded():
    """Flux form: sum(f*ha1*dx) changes only by the two boundary fluxes; the MC limiter
    keeps a step profile within its initial bounds (no new extrema)."""
    rng = np.random.default_rng(2)
    lx = 64
    x, dx, dxi = _line(lx, rng, uniform=False)
    ha1 = rng.uniform(0.5, 2.0, lx + 4)
    ha2i = rng.uniform(0.5, 2.0, lx + 1)
    h1i = rng.uniform(0.5, 2.0, lx + 1)
    f = np.where(np.arange(lx + 4) < lx // 2, 1.0, 0.2) + 0.0
    v = np.full(lx + 1, 0.3)
    v[: lx // 3] = -0.2
    dt = 0.4
    out = oracle.advec1D(f, v, dt, dx, dxi, ha1, ha2i, h1i)
    mass = lambda a: np.sum(a[2:lx + 2] * ha1[2:lx + 2] * dxi)
    # recompute only the two boundary fluxes to close the budget
    phi = []
    for i in (1, lx + 1):       # constant ends: the limited slopes vanish at both walls
        vi = v[i - 1]
        phi.append((f[i + 1] if vi < 0 else f[i]) * vi)
    budget = mass(f) - dt * (ha2i[lx] * phi[1] - ha2i[0] * phi[0])
    assert abs(mass(out) - budget) < 1e-12 * abs(mass(f))
    # with unit metric factors the limited update is monotone: no new extrema
    one = np.ones(lx + 4)
    out = oracle.advec1D(f, v, dt, dx, dxi, one, one[:lx + 1], one[:lx + 1])
    assert out[2:lx + 2].min() >= 0.2 - 1e-14 and out[2:lx + 2].max() <= 1.0 + 1e-14


# ------------------------------------------------------------------ interface velocities
@pytest.mark.parametrize("gridflag", [0, 1, 2])
def test_interface_velocity_boundary_rules(gridflag):
    g = ga.cartesian_grid(6, 2, 1)
    g.gridflag = gridflag
    st = ga.empty_state(g)
    u = np.array([9.0, 8.0, 5.0, -3.0, 7.0, 1.0, 2.0, -6.0, 4.0, 11.0])   # vs1(-1:8) of a line
    for isp in range(7):
        for i2 in range(6):
            st.vs1[:, i2, 2, isp] = u * (isp + 1)
    oracle.interface_vels(g, st)
    v = st.vs1i[:, 0, 0, 0]
    uu = u[2:8]                                                          # vs1(1:6)
    assert np.array_equal(v[1:6], 0.5 * (uu[:-1] + uu[1:]))
    first = {0: uu[0], 1: min(0.5 * (uu[0] + uu[1]), 0.0), 2: uu[0]}[gridflag]
    last = {0: uu[-1], 1: uu[-1], 2: max(0.5 * (uu[-2] + uu[-1]), 0.0)}[gridflag]
    assert v[0] == first and v[6] == last


# ------------------------------------------------------------------ ghost fill invariants
def test_ghost_fill_extrapolates_profiles():
    """x1 ghosts: exponential density profiles are extrapolated exactly where they decay
    outward and capped where they grow; the ghost temperature equals the edge
    temperature; the ghost velocity is the linear extrapolation through the interface
    velocity; x2 and x3 ghosts copy the edge cell."""
    g = ga.cartesian_grid(8, 3, 4)
    st = ga.synthetic_state(g, seed=4)
    L1 = g.lx1 + 4
    k = np.arange(-1, g.lx1 + 3, dtype=float)
    r = 1.5
    for isp in range(7):
        st.ns[:, :, :, isp] = (1e10 * r ** k)[:, None, None]
    og = oracle.OracleGrid(g)
    oracle.op(g, st, "v12rhov1")
    oracle.op(g, st, "T2rhoe")
    st.vs1i[...] = 0.0
    st.vs1i[0, ...] = 37.0
    st.vs1i[-1, ...] = -11.0
    before = st.copy()
    oracle.op(g, st, "set_global_boundaries_allspec")
    n = st.ns[:, 2, 2, 0]
    assert np.allclose(n[0:2], 1e10 * r ** k[0:2], rtol=1e-15)          # decays outward: exact
    assert n[L1 - 2] == n[L1 - 3] and n[L1 - 1] == n[L1 - 3]              # grows outward: capped
    Tgh = st.rhoes[0:2, 2, 2, 0] / st.ns[0:2, 2, 2, 0]
    Tedge = before.rhoes[2, 2, 2, 0] / before.ns[2, 2, 2, 0]
    assert np.allclose(Tgh, Tedge, rtol=1e-15)
    ms1 = oracle.constants()["ms"][0]
    vgh = st.rhovs1[1, 2, 2, 0] / (st.ns[1, 2, 2, 0] * ms1)
    assert np.isclose(vgh, 2 * 37.0 - before.vs1[2, 2, 2, 0], rtol=1e-14)
    for name in ("ns", "rhoes"):
        a = getattr(st, name)
        assert np.array_equal(a[:, 0, :, :], a[:, 2, :, :]) and np.array_equal(a[:, -1, :, :], a[:, -3, :, :])
        assert np.array_equal(a[:, :, 1, :], a[:, :, 2, :]) and np.array_equal(a[:, :, -2, :], a[:, :, -3, :])


# ------------------------------------------------------------------ constants
def test_thermal_coefficients_from_collision_data():
    """collisions.f90:51-53 precomputes 25/8 kB^2/ms/(Csj*1e-6) (the commented formula at
    collisions.f90:185) and writes the values as single-precision literals."""
    c = oracle.constants()
    kB = c["kB"]
    csj_diag = [0.22, 0.16, 0.17, 0.16, 0.24, 0.90]                 # Csj(isp,isp), :41-49
    for isp in range(6):
        tc = c["thermal_coeff"][isp]
        assert tc == float(np.float32(tc))                           # a widened float literal
        phys = 25.0 / 8.0 * kB ** 2 / c["ms"][isp] / (csj_diag[isp] * 1e-6)
        assert abs(tc - phys) / phys < 4e-3, (isp, tc, phys)
    # electron Spitzer coefficient: 7.7e5 eV cm^-1 s^-1 K^-7/2 in SI
    assert math.isclose(c["elchrg"] * 100.0 * 7.7e5, 7.7e5 * 1.60217662e-19 / 1e-2, rel_tol=1e-15)
    assert c["gammas"].tolist() == [5 / 3, 7 / 5, 7 / 5, 7 / 5, 5 / 3, 5 / 3, 5 / 3]


# ------------------------------------------------------------------ x**2.5
def test_pow25_correctly_rounded_and_close_to_libm():
    rng = np.random.default_rng(5)
    x = np.exp(rng.uniform(np.log(20.0), np.log(1e5), 200000))
    cr = oracle.eval_pow25(x, 1)
    lm = oracle.eval_pow25(x, 0)
    assert np.max(np.abs(cr - lm) / np.spacing(lm)) <= 1.0
    assert np.mean(cr != lm) < 0.005
    getcontext().prec = 50
    for i in np.flatnonzero(cr != lm)[:40]:
        X = Decimal(float(x[i]))
        exact = X * X * X.sqrt()
        assert abs(Decimal(float(cr[i])) - exact) <= abs(Decimal(float(lm[i])) - exact)
