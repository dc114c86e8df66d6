"""Pins the oracle's source/loss restatement (sources.f90, collisions.f90:58-159,
calculus.in.f90:174-196, multifluid.f90:472-549) by what physics and mathematics fix,
independently of the formulas it restates:

- ETD against the exact solution of df/dt = P - L f (expm1 form), its small-rate branch,
  fixed point and semigroup property;
- charge exchange conserves the ion count, Coulomb exchange conserves energy and momentum
  (up to the rounding of the published rate table), equilibria are fixed points;
- the species mapping of the electron-neutral rates against the thermal conductivity's own
  collision factors;
- the rate fits' branches meet where the reference switches between them, and the
  recombination coefficients take their published values at their reference temperatures;
- exact force terms: gravity alone, and a pressure gradient of an exponential profile.
"""
import math

import numpy as np
import pytest

import gemini_amd as ga
from gemini_amd.state import Atmosphere
import oracle

KB = 1.38064852e-23
AMU = 1.660539040e-27
MS = np.array([16, 30, 28, 32, 14, 1, 5.485799090e-4]) * AMU
MN = np.array([16, 28, 32, 1]) * AMU
GAMMA = np.array([5 / 3, 7 / 5, 7 / 5, 7 / 5, 5 / 3, 5 / 3, 5 / 3])


# ------------------------------------------------------------------ ETD
@pytest.mark.parametrize("x", [1e-3, 0.1, 1.0, 7.5, 40.0])
def test_etd_is_the_exact_solution(x):
    f, P, L = 3.0e10, 2.0e9, 0.8
    dt = x / L
    exact = f * math.exp(-x) - (P / L) * math.expm1(-x)
    assert oracle.etd(f, P, L, dt) == pytest.approx(exact, rel=1e-13, abs=0)


def test_etd_small_rate_branch_fixed_point_and_semigroup():
    f, P = 5.0, 3.0
    assert oracle.etd(f, P, 1e-12, 10.0) == f + P * 10.0          # L dt <= 1e-10
    L = 0.37
    assert oracle.etd(P / L, P, L, 2.0) == pytest.approx(P / L, rel=1e-15, abs=0)
    a = oracle.etd(oracle.etd(f, P, L, 0.7), P, L, 1.9)
    assert a == pytest.approx(oracle.etd(f, P, L, 2.6), rel=1e-13, abs=0)


# ------------------------------------------------------------------ collision rates
def test_electron_neutral_rates_match_the_conductivity_collision_factors():
    """thermal_conduct's electron conductivity holds each neutral's electron collision
    factor; nu_en(T) / (factor(T) sqrt(T)) must then be independent of T for O, N2, O2 --
    a swapped case in maxwell_colln changes the T dependence."""
    g = ga.cartesian_grid(4, 1, 1)
    ratios = {}
    for k in range(3):
        rs = []
        for T in (300.0, 1000.0, 3000.0, 6000.0):
            L = np.array(g.padded)
            Ts = np.full(L, T, order="F")
            ns = np.full(L, 1e11, order="F")
            nn = np.zeros((4, 1, 1, 3), order="F")
            nn[..., k] = 1e14
            J1 = np.zeros(L, order="F")
            lam, beta = np.zeros((4, 1, 1), order="F"), np.zeros((4, 1, 1), order="F")
            og = oracle.OracleGrid(g)
            import ctypes as C
            oracle.lib().orc_thermal_conduct(og.ref, C.c_int(7), oracle.ptr(Ts), oracle.ptr(ns),
                                            oracle.ptr(nn), oracle.ptr(J1), oracle.ptr(lam),
                                            oracle.ptr(beta))
            den = (1.60217662e-19 * 100 * 7.7e5 * T ** 2.5) / lam[0, 0, 0]
            factor = (den - 1.0) * 1e11 / (3.22e4 * T * T)       # nn_k * c_k * shape_k(T)
            nu = oracle.maxwell_colln(7, k + 1, 1e14, 0.0, T)
            rs.append(nu / (factor * math.sqrt(T)))
        ratios[k] = rs
        assert max(rs) / min(rs) - 1 < 1e-9, (k, rs)
    assert len({round(v[0], 3) for v in ratios.values()}) == 3


@pytest.mark.parametrize("T", [400.0, 2500.0])
def test_electron_neutral_rates_are_the_published_expressions(T):
    """Schunk & Nagy (2009) table 4.6 electron-neutral momentum-transfer rates [cm^3/s]."""
    n = 1e17
    expect = {1: 8.9e-11 * (1 + 5.7e-4 * T) * T ** 0.5, 2: 2.33e-11 * (1 - 1.21e-4 * T) * T,
              3: 1.82e-10 * (1 + 3.6e-2 * T ** 0.5) * T ** 0.5, 4: 4.5e-9 * (1 - 1.35e-4 * T) * T ** 0.5}
    for k, v in expect.items():
        assert oracle.maxwell_colln(7, k, n, 300.0, T) == pytest.approx(v * n * 1e-6, rel=1e-13, abs=0)


def test_ion_neutral_nonresonant_rates_are_the_published_table():
    """Schunk & Nagy (2009) table 4.4 coefficients [1e-10 cm^3/s] for a few pairs."""
    for (isp, isp2), c in {(1, 2): 6.82, (2, 1): 2.44, (6, 2): 33.6, (5, 4): 1.45,
                           (3, 4): 0.74, (4, 1): 2.31}.items():
        nu = oracle.maxwell_colln(isp, isp2, 1e18, 800.0, 1200.0)
        assert nu == pytest.approx(c * 1e-10 * 1e18 * 1e-6, rel=1e-7, abs=0)


@pytest.mark.parametrize("pair", [(1, 2), (1, 4), (2, 3), (3, 6), (4, 5), (1, 7), (6, 7)])
def test_coulomb_momentum_conservation_and_limits(pair):
    a, b = pair
    na, nb, Ta, Tb = 3e11, 7e10, 1400.0, 2100.0
    nu_ab, phi, psi = oracle.coulomb_colln(a, b, nb, Ta, Tb, 0.0, 0.0)
    nu_ba, phi2, psi2 = oracle.coulomb_colln(b, a, na, Tb, Ta, 0.0, 0.0)
    assert (phi, psi) == (1.0, 1.0) and (phi2, psi2) == (1.0, 1.0)
    # n_a m_a nu_ab = n_b m_b nu_ba up to the rounding of the published table
    assert na * MS[a - 1] * nu_ab == pytest.approx(nb * MS[b - 1] * nu_ba, rel=0.03, abs=0)
    assert oracle.coulomb_colln(a, a, na, Ta, Ta, 0.0, 0.0) == (0.0, 0.0, 0.0)
    # large drift: Phi -> (3/4) sqrt(pi) / W^3 as erf -> 1 and Psi -> 0
    mred = MS[a - 1] * MS[b - 1] / (MS[a - 1] + MS[b - 1])
    Teff = (MS[b - 1] * Ta + MS[a - 1] * Tb) / (MS[a - 1] + MS[b - 1])
    W = 8.0
    dv = W * math.sqrt(2 * KB * Teff / mred)
    _, phi, psi = oracle.coulomb_colln(a, b, nb, Ta, Tb, dv, 0.0)
    assert psi == pytest.approx(math.exp(-W * W), rel=1e-12, abs=0)
    assert phi == pytest.approx(0.75 * math.sqrt(math.pi) / W ** 3, rel=1e-12, abs=0)
    # the small-drift branch meets the formula at W = 0.1
    dv = 0.1000001 * math.sqrt(2 * KB * Teff / mred)
    _, phi, _ = oracle.coulomb_colln(a, b, nb, Ta, Tb, dv, 0.0)
    assert phi == pytest.approx(1.0, abs=0.01)


# ------------------------------------------------------------------ helpers
def uniform_case(g, ns, Ts, vs1=None, vn=(0.0, 0.0, 0.0), nn=None, Tn=800.0):
    """A state and atmosphere uniform in space with per-species ns, Ts (and vs1)."""
    st = ga.empty_state(g)
    for isp in range(7):
        st.fluidvars[..., isp] = ns[isp]
        st.fluidvars[..., 28 + isp] = Ts[isp]
        if vs1 is not None:
            st.fluidvars[..., 7 + isp] = vs1[isp]
    sh = g.shape
    nnv = np.zeros(sh + (6,), order="F")
    if nn is not None:
        for j, v in enumerate(nn):
            nnv[..., j] = v
    z = lambda v: np.full(sh, v, order="F")
    atm = Atmosphere(nnv, z(Tn), z(vn[0]), z(vn[1]), z(vn[2]), np.zeros(sh + (6,), order="F"), z(0.0))
    st.nn[...] = nnv[..., :3]
    st.Tn[...] = Tn
    return st, atm


NS = [2e11, 5e10, 1e9, 3e10, 1e9, 4e9, 2.88e11]
NN = [3e16, 1e17, 2e16, 1e12, 1e14, 1e13]


# ------------------------------------------------------------------ energy sources
def test_energy_thermal_equilibrium_is_a_fixed_point():
    """Every temperature equal to Tn and every velocity equal to the wind: each species'
    P/L is its own energy density n kB T / (gamma - 1), inelastic electron terms included
    (rotational gain and loss balance, vibrational terms vanish at Te = Tn)."""
    g = ga.cartesian_grid(6, 2, 2)
    T = 950.0
    st, atm = uniform_case(g, NS, [T] * 7, vs1=[12.0] * 7, vn=(12.0, 0.0, 0.0), nn=NN, Tn=T)
    Pr, Lo = oracle.srcs(g, st, atm, "energy")
    for isp in range(7):
        e = NS[isp] * KB * T / (GAMMA[isp] - 1)
        assert Pr[..., isp] / Lo[..., isp] == pytest.approx(np.full(g.shape, e), rel=1e-13, abs=0)


@pytest.mark.parametrize("isp", range(1, 8))
def test_neutral_collisions_relax_every_species_towards_Tn(isp):
    """A species whose only partners are neutrals (every other plasma density zero), at
    rest in the wind: whatever its temperature, P/L is its energy density at Tn -- elastic
    heat transfer and, for electrons above Tn, rotational excitation (vibrational
    excitation only cools then and is clipped) all relax towards the neutral temperature."""
    g = ga.cartesian_grid(5, 1, 1)
    ns = [0.0] * 7
    ns[isp - 1] = 3e10
    Tn = 700.0
    st, atm = uniform_case(g, ns, [2600.0] * 7, vs1=[5.0] * 7, vn=(5.0, 0.0, 0.0), nn=NN, Tn=Tn)
    Pr, Lo = oracle.srcs(g, st, atm, "energy")
    e = ns[isp - 1] * KB * Tn / (GAMMA[isp - 1] - 1)
    assert Pr[0, 0, 0, isp - 1] / Lo[0, 0, 0, isp - 1] == pytest.approx(e, rel=1e-13, abs=0)


def test_coulomb_energy_exchange_conserves_energy():
    """No neutrals, equal velocities, unequal temperatures.  Each species' equation relaxes
    its temperature (dE/dt = n kB/(gamma - 1) dT/dt), so what the exchange conserves is
    the translational energy: the sum of (gamma - 1) (P - L E) vanishes up to the rate
    table's asymmetry; the hottest species cools."""
    g = ga.cartesian_grid(5, 1, 1)
    Ts = [1000.0, 1300.0, 900.0, 1100.0, 1600.0, 2500.0, 3000.0]
    st, atm = uniform_case(g, NS, Ts)
    Pr, Lo = oracle.srcs(g, st, atm, "energy")
    E = np.array([NS[i] * KB * Ts[i] / (GAMMA[i] - 1) for i in range(7)])
    dE = np.array([(GAMMA[i] - 1) * (Pr[0, 0, 0, i] - Lo[0, 0, 0, i] * E[i]) for i in range(7)])
    assert abs(dE.sum()) < 0.03 * np.abs(dE).max()
    assert dE[6] < 0 and dE[0] > 0


def test_energy_friction_heats():
    g = ga.cartesian_grid(5, 1, 1)
    T = 900.0
    st, atm = uniform_case(g, NS, [T] * 7, vs1=[0, 50, -80, 30, 0, 200, 10], nn=NN, Tn=T)
    Pr, Lo = oracle.srcs(g, st, atm, "energy")
    for isp in range(6):
        e = NS[isp] * KB * T / (GAMMA[isp] - 1)
        assert Pr[0, 0, 0, isp] - Lo[0, 0, 0, isp] * e > 0


# ------------------------------------------------------------------ momentum sources
def gravity_free(g):
    planes = dict(g.planes)
    planes["g1"] = (np.zeros_like(planes["g1"][0]), planes["g1"][1])
    return ga.Grid(g.lx1, g.lx2, g.lx3, g.x1, g.x2, g.x3, planes, gridflag=g.gridflag,
                   isglobal=g.isglobal, kind=g.kind)


def test_momentum_drift_equilibrium_is_a_fixed_point():
    """Every species drifting with the neutral wind, uniform pressure, no field, no
    gravity: P/L = n m vn for every ion."""
    g = gravity_free(ga.cartesian_grid(8, 2, 2))
    vn1 = 37.0
    st, atm = uniform_case(g, NS, [900, 1000, 1100, 1200, 1300, 1400, 2000], vs1=[vn1] * 7,
                           vn=(vn1, 0.0, 0.0), nn=NN)
    Pr, Lo = oracle.srcs(g, st, atm, "momentum")
    for isp in range(6):
        assert Pr[..., isp] / Lo[..., isp] == pytest.approx(np.full(g.shape, NS[isp] * MS[isp] * vn1), rel=1e-13, abs=0)


def test_momentum_gravity_alone():
    g = ga.cartesian_grid(8, 2, 2)
    st, atm = uniform_case(g, NS, [1000.0] * 7)
    Pr, _ = oracle.srcs(g, st, atm, "momentum")
    for isp in range(6):
        assert np.array_equal(Pr[..., isp], (NS[isp] * MS[isp]) * g.g1)


def test_momentum_pressure_gradient_of_an_exponential_profile():
    """n = n0 exp(-x1/H) at constant T: log p is linear in x1, so the centred difference is
    exact on any spacing and the ion pressure force is -p/H (interior x1 cells; the
    reference's one-sided ends use dx1(1) and dx1(lx1))."""
    g = gravity_free(ga.cartesian_grid(12, 1, 1))
    H, T = 60e3, 1000.0
    st, atm = uniform_case(g, [1.0] * 7, [T] * 7)
    x1 = g.x1[:, None, None]
    prof = np.exp(-(x1 - 80e3) / H)
    for isp in range(7):
        st.fluidvars[..., isp] = NS[isp] * prof
    Pr, _ = oracle.srcs(g, st, atm, "momentum")
    I = slice(1, g.lx1 - 1)
    e = 1.60217662e-19
    for isp in range(6):
        p = NS[isp] * prof[2:-2] * KB * T
        ne_force = NS[isp] * prof[2:-2] * e * (KB * T / -e) * (-1.0 / H)     # n q Epol1
        expect = -(p * (-1.0 / H)) + ne_force
        got = Pr[I, 0, 0, isp]
        assert got == pytest.approx(expect[I, 0, 0], rel=1e-9, abs=0)


# ------------------------------------------------------------------ continuity sources
def test_charge_exchange_conserves_ions():
    """Without electrons no recombination happens, and every other reaction moves one ion
    from one species to another: sum of P - L n over the ions is zero."""
    g = ga.cartesian_grid(5, 2, 2)
    ns = list(NS[:6]) + [0.0]
    st, atm = uniform_case(g, ns, [1100, 1200, 1300, 1250, 1150, 1050, 2000], vs1=[30] * 7,
                           vn=(10.0, 5.0, -3.0), nn=NN)
    Pr, Lo = oracle.srcs(g, st, atm, "mass")
    net = sum(Pr[..., i] - Lo[..., i] * ns[i] for i in range(6))
    scale = max(np.abs(Pr[..., i]).max() for i in range(6))
    assert np.abs(net).max() < 1e-13 * scale


def test_recombination_coefficients_at_reference_temperatures():
    """Dissociative and radiative recombination [cm^3/s] at their reference temperatures
    (no neutrals, so recombination is the only loss)."""
    g = ga.cartesian_grid(5, 1, 1)
    ne = 1e11
    for Te, expect in ((250.0, {0: 3.7e-12, 5: 3.7e-12}), (300.0, {1: 4.2e-7, 2: 1.8e-7, 3: 1.95e-7})):
        st, atm = uniform_case(g, [1e10] * 6 + [ne], [1000.0] * 6 + [Te])
        _, Lo = oracle.srcs(g, st, atm, "mass")
        for isp, alpha in expect.items():
            assert Lo[0, 0, 0, isp] == pytest.approx(alpha * ne * 1e-6, rel=1e-15, abs=0)
    # O2+ switches fits at 1200 K; the two meet there
    for Te in (1200.0, 1200.0 * (1 + 1e-12)):
        st, atm = uniform_case(g, [1e10] * 6 + [ne], [1000.0] * 6 + [Te])
        _, Lo = oracle.srcs(g, st, atm, "mass")
        assert Lo[0, 0, 0, 3] == pytest.approx(1.95e-7 * 0.25 ** 0.7 * ne * 1e-6, rel=3e-3, abs=0)


def rate(g, reactant_T_species, nn_index, target, Teff):
    """kreac of one reaction at Teff, read from the loss of its ion (no electrons, no
    drift, Tn = Ts = Teff so the reaction's effective temperature is Teff)."""
    Ts = [Teff] * 6 + [1000.0]
    nn = [0.0] * 6
    nn[nn_index] = 1e16
    st, atm = uniform_case(g, [1e10] * 6 + [0.0], Ts, nn=nn, Tn=Teff)
    _, Lo = oracle.srcs(g, st, atm, "mass")
    return Lo[0, 0, 0, target] / (1e16 * 1e-6)


@pytest.mark.parametrize("nn_index,ion,Tswitch,gap", [(1, 0, 3725.0, 0.04828), (2, 0, 4800.0, 3.580e-4),
                                                     (5, 0, 3800.0, 2.524e-4)])
def test_rate_fits_meet_at_their_switch_temperatures(nn_index, ion, Tswitch, gap):
    """The O+ + N2, O2, NO fits switch polynomials at 3725, 4800 and 3800 K; the published
    pieces meet there to 4.8 %, 0.04 % and 0.03 % (a wrong sign or coefficient breaks
    that by orders of magnitude)."""
    g = ga.cartesian_grid(5, 1, 1)
    lo = rate(g, 1, nn_index, ion, Tswitch)
    hi = rate(g, 1, nn_index, ion, Tswitch * (1 + 1e-9))
    assert abs(hi / lo - 1) == pytest.approx(gap, rel=0.002, abs=0)
    assert lo > 0


def test_n2plus_power_laws_meet_at_1500K():
    """N2+ + O (both channels) switches power laws at 1500 K; the published fits meet
    there to 1 % (O+ channel) and 4 % (NO+ channel)."""
    g = ga.cartesian_grid(5, 1, 1)
    lo = rate(g, 3, 0, 2, 1500.0)
    hi = rate(g, 3, 0, 2, 1500.0 * (1 + 1e-9))
    assert hi == pytest.approx(lo, rel=0.05, abs=0)


# ------------------------------------------------------------------ gradient
def test_grad3D1_of_a_linear_function():
    g = ga.cartesian_grid(14, 2, 1)
    x = g.x1[2:-2]
    f = np.asfortranarray(np.broadcast_to((3.0 * x - 5e4)[:, None, None], g.shape))
    d = oracle.grad3D1(g, f)
    assert d[1:-1] == pytest.approx(np.full_like(d[1:-1], 3.0), rel=1e-12, abs=0)
    # ends: one-sided differences over dx1(1) = x1(1) - x1(0) and dx1(lx1)
    assert d[0, 0, 0] == pytest.approx(3.0 * (x[1] - x[0]) / (g.x1[2] - g.x1[1]), rel=1e-12, abs=0)
    assert d[-1, 0, 0] == pytest.approx(3.0 * (x[-1] - x[-2]) / (g.x1[-3] - g.x1[-4]), rel=1e-12, abs=0)
