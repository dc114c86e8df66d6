"""GPU parity of the source/loss substep: source_loss_energy_C, _momentum_C, _mass_C and
_allparams_C through the C ABI, and the fused sources block after the transport block,
against the oracle.

The collisional and chemical rates call exp, log, log10, erf, tanh, sin and x**p, which
the device library and glibc evaluate to about an ulp, each rounded its own way; the bar is
therefore the transport path's floating-point tolerance: every updated field within 1e-12
of the oracle, relative to the field's largest magnitude (per species)."""
import numpy as np
import pytest

import gemini_amd as ga
from gemini_amd.transport import Transport, GemhipError
import oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-12
FIELDS = {"ns": (0, "fluidvars"), "vs1": (7, "fluidvars"), "Ts": (28, "fluidvars"),
          "rhovs1": (0, "fluidauxvars"), "rhoes": (7, "fluidauxvars")}


def case(kind, seed=4):
    if kind == "cart2d":
        g = ga.cartesian_grid(64, 16, 1)
    elif kind == "cart":
        g = ga.cartesian_grid(24, 10, 7)
    elif kind == "interior_x1":
        # a subdomain that touches neither x1 end: the pressure gradients use the ghosts
        b = ga.cartesian_grid(20, 6, 5)
        g = ga.Grid(b.lx1, b.lx2, b.lx3, b.x1, b.x2, b.x3, b.planes, gridflag=b.gridflag,
                    isglobal=(False, False) + tuple(b.isglobal[2:]), kind="cartesian")
    else:
        g = ga.dipole_grid(32, 12, 6)
    st = ga.synthetic_state(g, seed=seed, perturb=g.lx3 > 1)
    atm = ga.synthetic_atmosphere(g, st, seed=seed)
    for name in ("v12rhov1", "T2rhoe", "set_global_boundaries_allspec"):
        oracle.op(g, st, name)
    st.electrovars[..., 0] = 1e-3 * np.sin(np.linspace(0, 7, st.electrovars[..., 0].size)).reshape(
        st.electrovars[..., 0].shape, order="F")
    return g, st, atm


def max_rel(a, b, field):
    k, blk = FIELDS[field]
    x = getattr(a, blk)[..., k:k + 7]
    y = getattr(b, blk)[..., k:k + 7]
    scale = np.max(np.abs(x), axis=(0, 1, 2))
    return np.max(np.max(np.abs(x - y), axis=(0, 1, 2)) / np.where(scale > 0, scale, 1.0))


def transport(g, st, atm):
    tr = Transport(g, st.nn, st.Tn)
    tr.enable_sources(atm)
    return tr


@pytest.mark.parametrize("kind", ["cart", "dipole", "cart2d", "interior_x1"])
def test_each_source_loss_entry_point(kind):
    print(f"case {kind}", flush=True)
    g, st, atm = case(kind)
    dt = ga.cfl_dt(g, st)
    tr = transport(g, st, atm)
    Q = np.zeros((g.lx1, g.lx2, g.lx3, 7), order="F")
    oracle.op(g, st.copy(), "VNRicht_artvisc", Q=Q)
    tr.VNRicht_artvisc(st.copy())                           # the library keeps intvars%Q
    assert np.abs(Q).max() > 0
    updated = {"energy": ("Ts", "rhoes"), "momentum": ("vs1", "rhovs1"), "mass": ("ns",)}
    for which, fields in updated.items():
        a, b = st.copy(), st.copy()
        oracle.source_loss(g, a, atm, dt, which, Q=Q)
        tr.source_loss(b, which, dt)
        for f in fields:
            err = max_rel(a, b, f)
            print(f"  {which} {f}: {err:.2e}", flush=True)
            assert err <= RTOL, f"{which} {f} {err:.3e}"
        for f in set(FIELDS) - set(fields):
            assert max_rel(st, b, f) == 0.0, f"{which} changed {f}"
    a, b = st.copy(), st.copy()
    for which in ("energy", "momentum", "mass"):
        oracle.source_loss(g, a, atm, dt, which, Q=Q)
    tr.source_loss(b, "allparams", dt)
    for f in FIELDS:
        assert max_rel(a, b, f) <= RTOL, f"allparams {f}"


def null_cells(g, below):
    alt = np.asarray(g.planes["alt"][0])[2:-2, 2:-2].T          # (lx1, lx2)
    m = np.broadcast_to((alt < below)[:, :, None], (g.lx1, g.lx2, g.lx3))
    return np.asfortranarray(m.astype(np.int32))


@pytest.mark.parametrize("kind,nulls", [("cart", False), ("dipole", False), ("cart", True)])
def test_sources_block_after_transport_block(kind, nulls):
    """One whole time step on the HBM state: gemhip_fluid_adv_block, then
    gemhip_sources_block, against the oracle's transport block and sources block, and the
    reference's sources calls one by one in mode 0 against the fused block; with null
    cells (altitude below 200 km) in one case."""
    print(f"case {kind} nulls={nulls}", flush=True)
    g, st, atm = case(kind)
    nul = null_cells(g, 200e3) if nulls else None
    if nulls:
        assert nul.sum() > 0
    dt = ga.cfl_dt(g, st)
    ref = st.copy()
    Q = np.zeros((g.lx1, g.lx2, g.lx3, 7), order="F")
    with oracle.pow_mode(1):
        oracle.fluid_adv_block(g, ref, dt, Q=Q, nullpts=nul)
    oracle.sources_block(g, ref, atm, dt, Q=Q, nullpts=nul)
    tr = transport(g, st, atm)
    if nulls:
        tr.register_null_cells(nul)
    tr.set_residency(1)
    tr.push(st)
    tr.fluid_adv_block(dt)
    tr.sources_block(dt)
    out = st.copy()
    tr.pull(out)
    for f in FIELDS:
        err = max_rel(ref, out, f)
        print(f"  {f}: {err:.2e}", flush=True)
        assert err <= RTOL, f"{f} {err:.3e}"
    tr.close()
    # the reference's calls of src/libgemini.cpp:155-232 one by one (mode 0)
    tr2 = transport(g, st, atm)
    if nulls:
        tr2.register_null_cells(nul)
    b = st.copy()
    tr2.fluid_adv_block_calls(b, dt)
    tr2.sources_block_calls(b, dt)
    for f in FIELDS:
        assert max_rel(ref, b, f) <= RTOL, f"calls {f}"


def test_sources_need_registration_and_report_errors():
    g, st, atm = case("cart")
    tr = Transport(g, st.nn, st.Tn)
    with pytest.raises(GemhipError, match="atmosphere not registered"):
        tr.source_loss(st.copy(), "energy", 0.1)
    tr.register_atmosphere(atm)
    with pytest.raises(GemhipError, match="gravity not registered"):
        tr.source_loss(st.copy(), "momentum", 0.1)
    tr.register_gravity()
    tr.clear_ionization_arrays()
    a, b = st.copy(), st.copy()
    z = ga.Atmosphere(atm.nn, atm.Tn, atm.vn1, atm.vn2, atm.vn3, 0 * atm.Prionize, 0 * atm.Qeionize)
    oracle.source_loss(g, a, z, 0.1, "mass")
    tr.source_loss(b, "mass", 0.1)
    assert max_rel(a, b, "ns") <= RTOL
