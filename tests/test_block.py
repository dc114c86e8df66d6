"""The rest of fluid_adv's transport block in the oracle (clean_param, VNRicht_artvisc,
RK2_global_boundary_allspec, compression), pinned on the CPU by closed forms and
invariants that the mathematics fixes."""
import ctypes as C

import numpy as np
import pytest

import gemini_amd as ga
import oracle
from oracle import OracleGrid, _f, lib


class MetricGrid:
    """Cartesian coordinates with explicit metric factors, for the divergence checks."""

    def __init__(self, base, h1, h2, h3):
        self.__dict__.update({k: getattr(base, k) for k in (
            "lx1", "lx2", "lx3", "gridflag", "flagper", "isglobal", "dx1", "dx1i", "dx2",
            "dx2i", "dx3", "dx3i", "x1", "x2", "x3")})
        L = base.padded
        self.h1, self.h2, self.h3 = (np.asfortranarray(np.broadcast_to(h, L)) for h in (h1, h2, h3))
        ones = lambda *s: np.ones(s, order="F")
        for n in ("h1x1i", "h2x1i", "h3x1i"):
            setattr(self, n, ones(self.lx1 + 1, self.lx2, self.lx3))
        for n in ("h1x2i", "h2x2i", "h3x2i"):
            setattr(self, n, ones(self.lx1, self.lx2 + 1, self.lx3))
        for n in ("h1x3i", "h2x3i", "h3x3i"):
            setattr(self, n, ones(self.lx1, self.lx2, self.lx3 + 1))


def _compress(g, st, dt, Q=None):
    Q = np.zeros((g.lx1, g.lx2, g.lx3, 7), order="F") if Q is None else Q
    og = OracleGrid(g)
    fv, fa = st.fluidvars, st.fluidauxvars
    lib().orc_compression(og.ref, C.c_double(dt), _f(fv[..., 7:14]), _f(fv[..., 14:21]), _f(fv[..., 21:28]),
                          _f(Q), _f(fa[..., 7:14]))


def test_compression_of_a_uniformly_stretching_flow():
    """v1 = a x1 has div v = a exactly under centred differences (on any spacing); the two-
    stage update is then RK2 of de/dt = -((gamma-1) e + Q) a."""
    g = ga.cartesian_grid(24, 5, 4)
    st = ga.synthetic_state(g, seed=2, perturb=True)
    a, dt = 2.5e-4, 0.8
    x1 = g.x1[:, None, None]
    st.fluidvars[..., 7:14] = (a * x1)[..., None]
    st.fluidvars[..., 14:28] = 0.0
    e0 = st.fluidauxvars[..., 7:14].copy()
    Q = np.asfortranarray(np.random.default_rng(1).uniform(0, 1e-12, (g.lx1, g.lx2, g.lx3, 7)))
    _compress(g, st, dt, Q)
    gm1 = oracle.constants()["gammas"] - 1.0
    I = (slice(2, -2),) * 3
    e, q = e0[I], Q
    half = e - (dt / 2) * (e * gm1 + q) * a
    ref = e - dt * (half * gm1 + q) * a
    got = st.fluidauxvars[..., 7:14][I]
    assert np.max(np.abs(got - ref) / np.abs(ref)) < 1e-12
    assert np.array_equal(st.fluidauxvars[..., 7:14][:2], e0[:2])    # ghosts untouched


@pytest.mark.parametrize("axis", [1, 2, 3])
def test_divergence_pairs_each_velocity_with_its_metric_product(axis):
    """A flux-free field along axis k (h_i h_j v_k constant, i, j != k) has zero divergence:
    compression leaves the energy density alone; pairing v_k with any other product of
    metric factors would not."""
    base = ga.cartesian_grid(20, 6, 5)
    L1, L2, L3 = base.padded
    ax = [np.arange(L1) + 3.0, np.arange(L2) + 5.0, np.arange(L3) + 4.0]
    shape = [(L1, 1, 1), (1, L2, 1), (1, 1, L3)][axis - 1]
    r = (1e3 + 37.0 * ax[axis - 1]).reshape(shape)               # varies along axis k only
    h = [np.ones((1, 1, 1))] * 3
    h = [r if i != axis - 1 else np.ones((1, 1, 1)) for i in range(3)]
    g = MetricGrid(base, *h)
    st = ga.synthetic_state(base, seed=3)
    st.fluidauxvars[..., 7:14] = np.random.default_rng(3).uniform(1e-9, 2e-9, (L1, L2, L3, 7))
    st.fluidvars[..., 7:28] = 0.0
    v = np.broadcast_to(1e12 / (r * r), (L1, L2, L3))
    st.fluidvars[..., 7 + 7 * (axis - 1):14 + 7 * (axis - 1)] = v[..., None]
    e0 = st.fluidauxvars[..., 7:14].copy()
    _compress(g, st, 1.0)
    I = (slice(2, -2),) * 3 + (slice(None),)
    change = np.max(np.abs(st.fluidauxvars[..., 7:14][I] - e0[I]) / e0[I])
    assert change < 1e-13, change
    # the same field with the metric product of another axis is not flux-free
    h_wrong = [np.ones((1, 1, 1))] * 3
    h_wrong[axis - 1] = r
    other = (axis % 3)
    h_wrong[other] = r
    st2 = st.copy()
    st2.fluidauxvars[..., 7:14] = e0
    st2.fluidvars[..., 7 + 7 * (axis - 1):14 + 7 * (axis - 1)] = v[..., None]
    _compress(MetricGrid(base, *h_wrong), st2, 1.0)
    assert np.max(np.abs(st2.fluidauxvars[..., 7:14][I] - e0[I]) / e0[I]) > 1e-9


def test_artificial_viscosity_scales_and_signs():
    """Q vanishes for expanding flow and for electrons, is linear in the density and
    quadratic in the velocity jump of compressing flow."""
    g = ga.cartesian_grid(16, 3, 2)
    st = ga.synthetic_state(g, seed=1)
    ni = (g.lx1, g.lx2, g.lx3, 7)
    og = OracleGrid(g)
    x1 = g.x1[:, None, None, None]

    def Q_of(a, nscale=1.0):
        s = st.copy()
        s.fluidvars[..., 7:14] = a * x1
        s.fluidvars[..., 0:7] *= nscale
        Q = np.zeros(ni, order="F")
        lib().orc_VNRicht_artvisc(og.ref, _f(s.fluidvars[..., 0:7]), _f(s.fluidvars[..., 7:14]), _f(Q))
        return Q

    assert np.all(Q_of(+1e-3) == 0.0)
    q1, q2, q3 = Q_of(-1e-3), Q_of(-2e-3), Q_of(-1e-3, 2.0)
    assert np.all(q1[..., 6] == 0.0) and np.all(q1[..., :6] > 0.0)
    assert np.allclose(q2[..., :6], 4.0 * q1[..., :6], rtol=1e-12)
    assert np.allclose(q3[..., :6], 2.0 * q1[..., :6], rtol=1e-12)


@pytest.mark.parametrize("kind", ["cart", "dipole"])
def test_clean_param_invariants(kind):
    g = ga.cartesian_grid(16, 4, 3) if kind == "cart" else ga.dipole_grid(24, 4, 3)
    st = ga.synthetic_state(g, seed=4, perturb=True)
    rng = np.random.default_rng(0)
    nul = (rng.uniform(size=(g.lx1, g.lx2, g.lx3)) < 0.1).astype(np.int32)
    I = (slice(2, -2),) * 3
    a = st.copy()
    for p in (1, 2, 3):
        oracle.op(g, a, "clean_param", which=p, nullpts=nul)
    b = a.copy()
    for p in (1, 2, 3):
        oracle.op(g, b, "clean_param", which=p, nullpts=nul)
    assert np.array_equal(a.fluidvars[..., 0:7], b.fluidvars[..., 0:7])      # idempotent
    assert np.array_equal(a.fluidvars[..., 28:35], b.fluidvars[..., 28:35])
    ns, Ts = a.fluidvars[..., 0:7], a.fluidvars[..., 28:35]
    ghost = np.ones(g.padded, bool)
    ghost[I] = False
    assert np.all(ns[ghost] == 1e-5) and np.all(Ts[ghost] == 100.0)
    nmask = np.zeros(g.padded, bool)
    nmask[I] = nul.astype(bool)
    assert np.all(ns[nmask][:, 0] == 1e-20 * 1e-2) and np.all(ns[nmask][:, 1:] == 1e-20)
    assert np.all(Ts[nmask] == 100.0) and np.all(Ts >= 100.0)
    live = ~ghost & ~nmask
    s = np.zeros(g.padded)
    for isp in range(6):
        s = s + ns[..., isp]
    assert np.array_equal(ns[..., 6][live], s[live])
    with pytest.raises(RuntimeError):
        oracle.op(g, a, "clean_param", which=4, nullpts=nul)


def test_rk2_boundary_is_zero_order_hold():
    g = ga.cartesian_grid(12, 4, 3)
    st = ga.synthetic_state(g, seed=5, perturb=True)
    before = st.copy()
    oracle.op(g, st, "RK2_global_boundary_allspec")
    v1, v2, v3 = (st.fluidvars[..., k:k + 7] for k in (7, 14, 21))
    lx1, lx2, lx3 = g.lx1, g.lx2, g.lx3
    I = (slice(2, -2),) * 3
    for v in (v1, v2, v3):
        assert np.array_equal(v[I], before.fluidvars[..., 7:28][I][..., :7] if v is v1 else v[I])
        assert np.array_equal(v[:, 1], v[:, 2]) and np.array_equal(v[:, lx2 + 2], v[:, lx2 + 1])
        assert np.array_equal(v[:, :, 1], v[:, :, 2]) and np.array_equal(v[:, :, lx3 + 2], v[:, :, lx3 + 1])
    assert np.array_equal(v1[1], v1[2]) and np.array_equal(v1[lx1 + 2], v1[lx1 + 1])


def test_dt_calc_closed_form():
    """Uniform temperatures and velocities: the largest grid rate of each species is its
    signal speed over the smallest cell length, with the ion sound speed
    sqrt(kB (Ti + 5/3 Te) / ms) and none for electrons."""
    g = ga.cartesian_grid(20, 5, 4)
    st = ga.synthetic_state(g, seed=1)
    Ti, Te, v1, v2, v3 = 1000.0, 2500.0, 300.0, -200.0, 150.0
    st.fluidvars[..., 28:34] = Ti
    st.fluidvars[..., 34] = Te
    st.fluidvars[..., 7:14] = v1
    st.fluidvars[..., 14:21] = v2
    st.fluidvars[..., 21:28] = v3
    c = oracle.constants()
    dt, rates = oracle.dt_calc(g, st, 0.9)
    dl = [np.min(np.asarray(g.dx1i)), np.min(np.asarray(g.dx2i)), np.min(np.asarray(g.dx3i))]
    for isp in range(7):
        cs = np.sqrt(c["kB"] * (Ti + 5.0 / 3.0 * Te) / c["ms"][isp]) if isp < 6 else 0.0
        want = [(cs + abs(v1)) / dl[0], abs(v2) / dl[1], abs(v3) / dl[2]]
        assert np.allclose(rates[isp], want, rtol=1e-14, atol=0.0)
    assert np.isclose(dt, 0.9 / rates.max(), rtol=1e-15)


def test_dt_select_rules():
    """Output-time clamps, the 1e-6 floor and the factor-2 growth limit from the second
    iteration (libgemini_mpi.f90:467-482, temporal_mpi.f90:53-66)."""
    g = ga.cartesian_grid(16, 4, 3)
    st = ga.synthetic_state(g, seed=2)
    free, _ = oracle.dt_calc(g, st, 0.9)
    sel = lambda **k: oracle.dt_select(g, st, 0.9, k.get("glow", 0), k.get("t", 0.0),
                                       k.get("tout", 1e9), k.get("tglow", 1e9), k.get("it", 1),
                                       k.get("dt", 1e-3))
    assert sel() == free                                   # first iteration: no growth limit
    assert sel(it=2, dt=free / 10) == 2 * (free / 10)      # later: at most twice the last dt
    assert sel(t=10.0, tout=10.0 + free / 3) == (10.0 + free / 3) - 10.0
    assert sel(glow=1, t=1.0, tglow=1.0 + free / 5) == (1.0 + free / 5) - 1.0
    assert sel(glow=0, t=1.0, tglow=1.0 + free / 5) == free
    assert sel(t=5.0, tout=5.0) == 1e-6


@pytest.mark.parametrize("gridflag", [0, 1])
def test_clean_param_velocity_border_rule_on_every_mask(gridflag):
    """clean_param 2's dipole border copy (multifluid.f90:605-657) on every null mask of a
    6-cell field line, against the rule stated through runs of the mask: null cells are
    zeroed; on a closed dipole (gridflag 0) the first null cell from the bottom takes the
    value above it, then the first non-null cell after that run of null cells takes the value
    below it (each search stops at lx1, where nothing is copied); on an inverted open dipole
    (gridflag 1) only the second copy is made, at the first null cell."""
    import itertools
    b = ga.cartesian_grid(6, 1, 1)
    g = ga.Grid(b.lx1, b.lx2, b.lx3, b.x1, b.x2, b.x3, b.planes, gridflag=gridflag,
                isglobal=b.isglobal, kind="cartesian")
    lx1 = g.lx1
    rng = np.random.default_rng(5)
    for bits in itertools.product((0, 1), repeat=lx1):
        mask = np.array(bits, dtype=bool)
        st = ga.synthetic_state(g, seed=1)
        st.fluidvars[..., 7:14] = rng.uniform(-100, 100, st.fluidvars[..., 7:14].shape)
        col = st.fluidvars[:, 2, 2, 7:14]                      # padded: cell i1 at i1 + 1
        want = col.copy()
        want[2:-2][mask] = 0.0
        nul = np.asfortranarray(mask.reshape(lx1, 1, 1).astype(np.int32))
        oracle.op(g, st, "clean_param", which=2, nullpts=nul)
        nulls = np.flatnonzero(mask[:lx1 - 1])                  # searches stop at lx1
        first_null = int(nulls[0]) + 1 if len(nulls) else lx1  # 1-based
        if gridflag == 0:
            run = first_null
            while run < lx1 and mask[run - 1]:
                run += 1
            if first_null != lx1:
                want[first_null + 1] = want[first_null + 2]
            if run != lx1:
                want[run + 1] = want[run]
        elif first_null != lx1:
            want[first_null + 1] = want[first_null]             # from the cell below (a ghost at i1 = 1)
        assert np.array_equal(col, want), bits
