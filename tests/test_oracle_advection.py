"""Pins of the oracle's advection scheme and state conversions, independent of its code.

Each check states a property the mathematics or the physics fixes, not the oracle's
formula retyped:

* the MC-limited scheme with the reference's flux (advec.f90:384-435) is Fromm's scheme
  on smooth monotone data, which translates a quadratic exactly by the coordinate
  distance v dt / h1i; at coordinate Courant number one it is the exact donor-cell shift;
* the sweeps' metric products (advec.f90:283-381): a constant field whose interface
  fluxes ha2i v grow by the same amount across every cell changes by exactly
  dt F c / (dx1i ha1), and the products enter only through that relation;
* the entry points' species counts and tensor ranks (multifluid.f90:49-126) and the
  composite step of the oracle as the same sequence of entry points;
* the conversions (multifluid.f90:179-233): energy density from the degrees of freedom
  of each species, momentum from the species' mass numbers, current-free electrons,
  electron density as the sum of the ions, and the round trips;
* the x1 ghost velocities (advec.f90:52-55, 71-73): linear extrapolation through the
  interface velocity.

``tests/test_oracle_mutants.py`` builds copies of the oracle with one plausible mistake
each and shows that these checks fail on every one of them.
"""
import numpy as np
import pytest

import gemini_amd as ga
import oracle


# ------------------------------------------------------------------ grids
def random_metric_grid(lx1=14, lx2=9, lx3=7, seed=11):
    """Cartesian-shaped grid whose twelve metric arrays are independent random fields
    (no two products of different factors agree anywhere) and whose x2 and x3
    coordinates are non-uniform."""
    g = ga.cartesian_grid(lx1, lx2, lx3)
    rng = np.random.default_rng(seed)
    for name, shape in (("h1", g.padded), ("h2", g.padded), ("h3", g.padded),
                        ("h1x1i", (lx1 + 1, lx2, lx3)), ("h2x1i", (lx1 + 1, lx2, lx3)),
                        ("h3x1i", (lx1 + 1, lx2, lx3)), ("h1x2i", (lx1, lx2 + 1, lx3)),
                        ("h2x2i", (lx1, lx2 + 1, lx3)), ("h3x2i", (lx1, lx2 + 1, lx3)),
                        ("h1x3i", (lx1, lx2, lx3 + 1)), ("h2x3i", (lx1, lx2, lx3 + 1)),
                        ("h3x3i", (lx1, lx2, lx3 + 1))):
        g._full[name] = np.asfortranarray(rng.uniform(0.3, 3.0, shape))
    for ax, lx in ((2, lx2), (3, lx3)):
        x = np.cumsum(rng.uniform(2e3, 9e3, lx + 4))
        setattr(g, f"x{ax}", x)
        dx, dxi = ga.grid.coord_diffs(x, lx)
        setattr(g, f"dx{ax}", dx)
        setattr(g, f"dx{ax}i", dxi)
    return g


# the metric products the reference names per sweep direction and tensor rank
def _cell_product(g, direction, frank):
    """ha1 of the sweep: h1 h2 h3, with h1 squared for rank 1 in x2 and x3
    (advec.f90:302, :334/:337, :370/:373; sweep1 ignores the rank)."""
    if direction == 1 or frank == 0:
        return (g.h1 * g.h2) * g.h3
    return ((g.h1 * g.h1) * g.h2) * g.h3


def _face_product(g, direction, frank):
    """ha2i of the sweep: the two transverse interface factors, with h1 squared for
    rank 1 in x2 and x3 (advec.f90:303, :335/:338, :371/:374)."""
    if direction == 1:
        return g.h2x1i * g.h3x1i
    a, b = (g.h1x2i, g.h3x2i) if direction == 2 else (g.h1x3i, g.h2x3i)
    return (a * a) * b if frank else a * b


def _face_scale(g, direction):
    """h1i of the sweep: the metric factor of the swept axis at its interfaces
    (advec.f90:304, :340, :376)."""
    return {1: g.h1x1i, 2: g.h2x2i, 3: g.h3x3i}[direction]


def _line_index(direction, shape):
    """Index along the swept axis of every interface of a (lx1+1, ...) style array."""
    idx = [np.arange(n) + 1 for n in shape]
    return np.meshgrid(*idx, indexing="ij")[direction - 1].astype(float)


def _interior(a):
    return a[2:-2, 2:-2, 2:-2]


def _spacing(g, direction):
    return {1: (g.dx1, g.dx1i), 2: (g.dx2, g.dx2i), 3: (g.dx3, g.dx3i)}[direction]


def _bcast(v, direction, nd=3):
    s = [1] * nd
    s[direction - 1] = -1
    return v.reshape(s)


# ------------------------------------------------------------------ one line
def _quadratic_line(lx, dx, sign, nu=0.4, hscale=2.5, area=1.7, dt=0.8):
    x = (np.arange(-1, lx + 3) + 4.0) * dx      # x(0) > dx: MC picks the centred slope
    f = x * x
    dxa = np.full(lx + 3, dx)                   # dx1(0:lx+2)
    dxi = np.full(lx, dx)
    v = np.full(lx + 1, sign * nu * hscale * dx / dt)
    out = oracle.advec1D(f, v, dt, dxa, dxi, np.full(lx + 4, hscale * area),
                         np.full(lx + 1, area), np.full(lx + 1, hscale))
    return x, f, out


def test_mc_scheme_translates_a_quadratic_exactly():
    """On a monotone convex profile the MC limiter selects the centred slope, and the
    scheme is then Fromm's, exact for quadratics: one step moves x^2 by the coordinate
    distance nu dx = |v| dt / h1i, in either direction, with cell volume ha1 = h1i ha2i
    as a consistent metric has it.  A plain minmod(l, r) slope, v*h1i for v/h1i,
    exchanged ha1/ha2i or a wrong upwind slope each break the exact translation."""
    for dx in (1.0, 0.37):
        for sign in (1.0, -1.0):
            lx, nu = 40, 0.4
            x, f, out = _quadratic_line(lx, dx, sign, nu=nu)
            want = (x - sign * nu * dx) ** 2
            err = np.max(np.abs(out[2:lx + 2] - want[2:lx + 2])) / np.max(np.abs(want))
            assert err < 1e-14, (dx, sign, err)
            assert np.array_equal(out[[0, 1, lx + 2, lx + 3]], f[[0, 1, lx + 2, lx + 3]])


def _fromm(f, nu):
    """Fromm's scheme for rightward motion at Courant number nu on a uniform grid, as the
    average of the Lax-Wendroff and Beam-Warming updates (interior cells 2..n-3)."""
    fm2, fm1, f0, fp1 = f[:-3], f[1:-2], f[2:-1], f[3:]
    lw = f0 - nu / 2 * (fp1 - fm1) + nu * nu / 2 * (fp1 - 2 * f0 + fm1)
    bw = f0 - nu / 2 * (3 * f0 - 4 * fm1 + fm2) + nu * nu / 2 * (f0 - 2 * fm1 + fm2)
    return 0.5 * (lw + bw)


def test_mc_scheme_is_fromm_on_smooth_monotone_data():
    """Where the centred slope is below twice both one-sided slopes (a smooth monotone
    cubic, away from its inflection) the MC limiter keeps the centred slope, and the
    scheme is Fromm's: the mean of Lax-Wendroff and Beam-Warming.  A cubic tells the
    centred slope from the one-sided ones by a non-constant amount, so limiting with
    minmod(l, r), or MC without its factor 2, changes the update."""
    for dx in (1.0, 0.37):
        for sign in (1.0, -1.0):
            lx, nu, h, area, dt = 40, 0.35, 1.9, 0.6, 0.8
            x = (np.arange(-1, lx + 3) + 4.0) * dx
            f = x ** 3
            v = np.full(lx + 1, sign * nu * h * dx / dt)
            out = oracle.advec1D(f, v, dt, np.full(lx + 3, dx), np.full(lx, dx),
                                 np.full(lx + 4, h * area), np.full(lx + 1, area),
                                 np.full(lx + 1, h))
            want = _fromm(f, nu) if sign > 0 else _fromm(f[::-1], nu)[::-1]
            # interior cells 1..lx are positions 2..lx+1; _fromm covers positions 2..lx+2
            got = out[2:lx + 2]
            ref = want[:lx] if sign > 0 else want[1:lx + 1]
            assert np.max(np.abs(got - ref)) / np.max(np.abs(f)) < 1e-14, (dx, sign)


def test_courant_one_is_an_exact_shift_with_metric_factors():
    """Coordinate Courant number one (v dt / h1i = dx) on rough data, with h1i = 2,
    v = 2, dt = dx = 1 and a consistent volume ha1 = h1i ha2i: the limited correction
    vanishes and each cell takes its upwind neighbour's value."""
    rng = np.random.default_rng(3)
    lx = 50
    f = rng.uniform(1.0, 2.0, lx + 4)
    one = np.ones(lx + 3)
    for sign in (1.0, -1.0):
        out = oracle.advec1D(f, np.full(lx + 1, 2.0 * sign), 1.0, one, one[:lx],
                             np.full(lx + 4, 2.0), np.ones(lx + 1), np.full(lx + 1, 2.0))
        want = f[1:lx + 1] if sign > 0 else f[3:lx + 3]
        assert np.allclose(out[2:lx + 2], want, rtol=0, atol=4e-16 * 2)


# ------------------------------------------------------------------ sweeps
def _sweep_case(g, direction, frank, c=0.05, F=3.0, dt=0.7):
    """Constant field F; interface velocities v = c k / ha2i with k the interface index
    along the line, so the face flux ha2i F v grows by F c across every cell.  The
    slopes vanish, and the update is exactly F - dt F c / (dx1i ha1)."""
    shape = {1: (g.lx1 + 1, g.lx2, g.lx3), 2: (g.lx1, g.lx2 + 1, g.lx3),
             3: (g.lx1, g.lx2, g.lx3 + 1)}[direction]
    P = _face_product(g, direction, frank)
    k = _line_index(direction, shape)
    dxl = _spacing(g, direction)[1]
    c = c * np.min(dxl) / dt                    # changes of order c F / Q
    vi = np.zeros(shape + (7,), order="F")
    fs = np.full(g.padded + (7,), F, order="F")
    for isp in range(7):
        vi[..., isp] = (c * (isp + 1)) * k / P
    oracle.sweep(g, direction, fs, vi, dt, frank, 7)
    Q = _interior(_cell_product(g, direction, frank))
    for isp in range(7):
        want = F - dt * F * (c * (isp + 1)) / _bcast(dxl, direction) / Q
        got = _interior(fs[..., isp])
        assert np.max(np.abs(got - want) / (np.abs(want) + F)) < 1e-13, (direction, frank, isp)
        full = fs[..., isp]
        mask = np.ones(g.padded, bool)
        mask[2:-2, 2:-2, 2:-2] = False
        assert np.all(full[mask] == F)                  # ghost and edge cells untouched
    return fs


@pytest.mark.parametrize("direction", [1, 2, 3])
@pytest.mark.parametrize("frank", [0, 1])
def test_sweep_metric_products(direction, frank):
    """The volume and face products of every sweep direction and tensor rank, on a grid
    whose twelve metric arrays are independent: any other pairing of factors (or an x1
    sweep that honours the rank) changes the update of most cells by O(1)."""
    g = random_metric_grid()
    _sweep_case(g, direction, frank)


@pytest.mark.parametrize("direction", [1, 2, 3])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_sweep_courant_one_reduces_to_donor_cell(direction, sign):
    """Interface velocities at coordinate Courant number one (v = +-h1i dx(k) / dt, with
    dx(k) the centre spacing across interface k) on a smooth monotone field: the MC
    correction vanishes only if the sweep scales v by the swept axis' interface factor
    and uses that interface's spacing, leaving the donor-cell update."""
    g = random_metric_grid(seed=5)
    dt = 0.9
    shape = {1: (g.lx1 + 1, g.lx2, g.lx3), 2: (g.lx1, g.lx2 + 1, g.lx3),
             3: (g.lx1, g.lx2, g.lx3 + 1)}[direction]
    dx, dxi = _spacing(g, direction)
    lx = shape[direction - 1] - 1
    H = _face_scale(g, direction)
    v = sign * H * _bcast(dx[1:lx + 2], direction) / dt       # dx(1:lx+1)
    vi = np.repeat(v[..., None], 7, axis=3).copy(order="F")
    # smooth, monotone, convex along the swept axis (non-zero limited slopes everywhere)
    kk = np.arange(-1, lx + 3, dtype=float)
    prof = 1.0 + 0.05 * kk + 0.004 * kk * kk
    fs = np.asfortranarray(np.broadcast_to(_bcast(prof, direction, 3)[..., None],
                                           g.padded + (7,)).copy())
    f0 = fs[..., 0].copy()

    def take(a, lo, hi):                          # a slice along the swept axis
        s = [slice(None)] * 3
        s[direction - 1] = slice(lo, hi)
        return a[tuple(s)]

    # f along each line from index -1 (position 0) to lx+2, interior across it
    f_ln = f0[tuple(slice(None) if i == direction - 1 else slice(2, -2) for i in range(3))]
    for frank in (0, 1):
        work = fs.copy(order="F")
        oracle.sweep(g, direction, work, vi, dt, frank, 7)
        P = _face_product(g, direction, frank)
        Q = _cell_product(g, direction, frank)
        # donor cell: interface k lies between cells k-1 (position k) and k (k+1), and
        # phi(k) is the upwind cell's value times v(k)
        upw = take(f_ln, 1, lx + 2) if sign > 0 else take(f_ln, 2, lx + 3)
        phi = P * upw * v
        dphi = take(phi, 1, lx + 1) - take(phi, 0, lx)
        change = dt * dphi / _bcast(dxi, direction) / _interior(Q)
        want = _interior(f0) - change
        got = _interior(work[..., 0])
        scale = np.abs(_interior(f0)) + np.abs(change)
        assert np.max(np.abs(got - want) / scale) < 1e-13, (direction, frank, sign)


@pytest.mark.parametrize("direction", [1, 2, 3])
def test_entry_points_sweep_the_reference_species_and_ranks(direction):
    """mass: ns of the six ions as scalars; momentum: rhovs1 of the six ions as the
    parallel component of a vector; energy: rhoes of all seven species as scalars
    (multifluid.f90:49-126).  The electrons' density and momentum are left alone."""
    g = random_metric_grid(seed=7)
    st = ga.empty_state(g)
    shape = {1: (g.lx1 + 1, g.lx2, g.lx3), 2: (g.lx1, g.lx2 + 1, g.lx3),
             3: (g.lx1, g.lx2, g.lx3 + 1)}[direction]
    k = _line_index(direction, shape)
    dxi = _spacing(g, direction)[1]
    dt = 0.6
    c = 0.04 * np.min(dxi) / dt
    for which, field, nsp, frank in (("mass", "ns", 6, 0), ("momentum", "rhovs1", 6, 1),
                                     ("energy", "rhoes", 7, 0)):
        P = _face_product(g, direction, frank)
        vsi = {1: st.vs1i, 2: st.vs2i, 3: st.vs3i}[direction]
        for isp in range(7):
            vsi[..., isp] = c * k / P
        arr = getattr(st, field)
        arr[...] = 2.0
        oracle.op(g, st, "sweep", dt=dt, direction=direction, which=which)
        Q = _interior(_cell_product(g, direction, frank))
        want = 2.0 - dt * 2.0 * c / _bcast(dxi, direction) / Q
        for isp in range(7):
            got = _interior(arr[..., isp])
            if isp < nsp:
                assert np.max(np.abs(got - want) / (np.abs(want) + 2.0)) < 1e-13, (which, isp)
            else:
                assert np.all(got == 2.0), (which, isp)


def test_composite_step_is_the_entry_point_sequence():
    """The oracle's whole transport step equals its entry points called one by one in
    the order of libgemini.cpp:155-208 (so the entry-point pins carry over to it)."""
    g = ga.dipole_grid(20, 8, 6)
    st = ga.synthetic_state(g, seed=9, perturb=True)
    dt = ga.cfl_dt(g, st)
    a, b = st.copy(), st.copy()
    oracle.transport_step(g, a, dt)
    for name, kw in (("v12rhov1", {}), ("T2rhoe", {}), ("set_global_boundaries_allspec", {}),
                     ("halo_fluidvars", {}), ("interface_vels_allspec", {}),
                     ("sweep", dict(direction=3, which="allparams")),
                     ("sweep", dict(direction=1, which="allparams")), ("halo_allparams", {}),
                     ("sweep", dict(direction=2, which="allparams")), ("rhov12v1", {}),
                     ("rhoe2T", {}), ("energy_diffusion", {})):
        oracle.op(g, b, name, dt=dt, **kw)
    for blk in ("fluidvars", "fluidauxvars"):
        x, y = getattr(a, blk)[..., :35], getattr(b, blk)[..., :35]
        assert np.array_equal(x, y, equal_nan=True), blk


# ------------------------------------------------------------------ conversions
AMU_MASS_NUMBERS = [16.0, 30.0, 28.0, 32.0, 14.0, 1.0]        # O+, NO+, N2+, O2+, N+, H+
DEGREES_OF_FREEDOM = [3, 5, 5, 5, 3, 3, 3]                   # atoms 3, diatomic ions 5


def _conv_state(seed=1):
    g = ga.cartesian_grid(10, 6, 5)
    st = ga.synthetic_state(g, seed=seed)
    rng = np.random.default_rng(seed)
    st.vs1[...] = rng.uniform(-300.0, 300.0, st.vs1.shape)
    st.Ts[...] = rng.uniform(100.0, 4000.0, st.Ts.shape)
    return g, st


def test_energy_density_from_degrees_of_freedom():
    """T2rhoe: rhoe = (f/2) n kB T, f = 3 for the atomic ions and electrons and 5 for the
    diatomic NO+, N2+, O2+; rhoe2T inverts it to rounding, and floors at 100 K."""
    g, st = _conv_state()
    kB = oracle.constants()["kB"]
    oracle.op(g, st, "T2rhoe")
    for isp in range(7):
        ratio = st.rhoes[..., isp] / (st.ns[..., isp] * kB * st.Ts[..., isp])
        assert np.allclose(ratio, DEGREES_OF_FREEDOM[isp] / 2.0, rtol=1e-15, atol=0), isp
    T0 = st.Ts.copy()
    st.Ts[...] = 0.0
    oracle.op(g, st, "rhoe2T")
    dense = st.ns >= 1e-5                                         # divisor not floored
    assert np.allclose(st.Ts[dense], T0[dense], rtol=1e-15, atol=0)
    st.rhoes[...] *= 50.0 / 4000.0                               # below 100 K everywhere
    oracle.op(g, st, "rhoe2T")
    assert np.all(st.Ts == 100.0)


def test_momentum_density_from_mass_numbers():
    """v12rhov1: rhovs1 = n m v with the ions' masses their mass numbers in amu and the
    electrons' 5.485799090e-4 amu; rhov12v1 gives the ion velocities back to rounding."""
    g, st = _conv_state(2)
    amu = 1.660539040e-27
    masses = AMU_MASS_NUMBERS + [5.485799090e-4]
    oracle.op(g, st, "v12rhov1")
    for isp in range(7):
        m = st.rhovs1[..., isp] / (st.ns[..., isp] * st.vs1[..., isp])
        assert np.allclose(m, masses[isp] * amu, rtol=1e-15, atol=0), isp
    v0 = st.vs1[..., :6].copy()
    st.vs1[...] = 0.0
    oracle.op(g, st, "rhov12v1")
    dense = st.ns[..., :6] >= 1e-5
    assert np.allclose(st.vs1[..., :6][dense], v0[dense], rtol=1e-15, atol=0)


def test_electrons_carry_no_net_current_and_balance_the_ions():
    """rhov12v1: ne = sum of the six ion densities, and the electron velocity makes the
    parallel current sum_s n_s q_s v_s vanish (the current-free closure of
    multifluid.f90:231-232); densities below 1e-5 use 1e-5 as divisor."""
    g, st = _conv_state(3)
    c = oracle.constants()
    oracle.op(g, st, "v12rhov1")
    st.ns[..., 6] = -1.0
    oracle.op(g, st, "rhov12v1")
    ions = st.ns[..., :6]
    assert np.allclose(st.ns[..., 6], ions.sum(axis=-1), rtol=1e-15, atol=0)
    J = sum(st.ns[..., s] * c["qs"][s] * st.vs1[..., s] for s in range(7))
    Jscale = sum(np.abs(st.ns[..., s] * c["qs"][s] * st.vs1[..., s]) for s in range(6))
    dense = st.ns[..., 6] >= 1e-5
    assert np.max(np.abs(J[dense]) / Jscale[dense]) < 1e-14
    assert np.all(c["qs"][:6] > 0) and c["qs"][6] < 0
    # tenuous cells: the floor on the divisor
    st2 = st.copy()
    st2.ns[3, 3, 3, :6] = 1e-7
    oracle.op(g, st2, "v12rhov1")
    v = st2.vs1[3, 3, 3, :6].copy()
    oracle.op(g, st2, "rhov12v1")
    assert np.allclose(st2.vs1[3, 3, 3, :6], v * 1e-7 / 1e-5, rtol=1e-15, atol=0)


# ------------------------------------------------------------------ x1 ghost velocities
def test_x1_ghost_velocities_extrapolate_linearly_through_the_interface():
    """Both ghost cells of each x1 end: the interface velocity is the mean of the edge
    cell and the first ghost, and the second ghost continues the same line
    (u(-1) - 2 u(0) + u(1) = 0, advec.f90:52-55 and :71-73)."""
    g = ga.cartesian_grid(9, 4, 3)
    st = ga.synthetic_state(g, seed=6)
    rng = np.random.default_rng(6)
    st.vs1[...] = rng.uniform(-200.0, 200.0, st.vs1.shape)
    st.vs1i[0, ...] = rng.uniform(-50.0, 50.0, st.vs1i[0].shape)
    st.vs1i[-1, ...] = rng.uniform(-50.0, 50.0, st.vs1i[-1].shape)
    oracle.op(g, st, "v12rhov1")
    oracle.op(g, st, "T2rhoe")
    before = st.copy()
    oracle.op(g, st, "set_global_boundaries_allspec")
    ms = oracle.constants()["ms"]
    L1 = g.lx1 + 4
    for isp in range(7):
        u = st.rhovs1[:, 2:-2, 2:-2, isp] / (st.ns[:, 2:-2, 2:-2, isp] * ms[isp])
        u1 = before.vs1[:, 2:-2, 2:-2, isp]
        # bottom: ghosts at array 1 (i=0) and 0 (i=-1), edge cell at 2 (i=1)
        vb, vt = st.vs1i[0, :, :, isp], st.vs1i[-1, :, :, isp]
        scale = 400.0
        assert np.max(np.abs(0.5 * (u[1] + u1[2]) - vb)) < 1e-12 * scale
        assert np.max(np.abs(u[0] - 2 * u[1] + u1[2])) < 1e-12 * scale
        assert np.max(np.abs(0.5 * (u[L1 - 2] + u1[L1 - 3]) - vt)) < 1e-12 * scale
        assert np.max(np.abs(u[L1 - 1] - 2 * u[L1 - 2] + u1[L1 - 3])) < 1e-12 * scale
