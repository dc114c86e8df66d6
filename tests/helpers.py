"""Shared test utilities: per-variable comparison of two states."""
import numpy as np

from conftest import normwise_rel

VARS = {"ns": ("fluidvars", 0), "vs1": ("fluidvars", 7), "vs2": ("fluidvars", 14),
        "vs3": ("fluidvars", 21), "Ts": ("fluidvars", 28), "rhovs1": ("fluidauxvars", 0),
        "rhoes": ("fluidauxvars", 7)}


def var(st, name, interior=False):
    block, off = VARS[name]
    a = getattr(st, block)[..., off:off + 7]
    if interior:
        a = a[2:-2, 2:-2, 2:-2]
    return a


def compare(got, ref, exact=(), close=(), rtol=1e-12, interior_close=True):
    """Assert bit equality for `exact` variables and normwise rtol for `close` ones,
    species by species; returns the worst relative difference seen."""
    worst = 0.0
    for name in exact:
        a, b = var(got, name), var(ref, name)
        same = np.array_equal(a, b, equal_nan=True)
        if not same:
            bad = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b))))
            i = tuple(bad[0])
            raise AssertionError(f"{name} not bit-identical at {len(bad)} cells, first {i}: "
                                 f"{a[i]!r} vs {b[i]!r} (rel {normwise_rel(a, b):.3e})")
    for name in close:
        a, b = var(got, name, interior_close), var(ref, name, interior_close)
        for s in range(7):
            r = normwise_rel(a[..., s], b[..., s])
            worst = max(worst, r)
            assert r <= rtol, f"{name} species {s + 1}: normwise rel diff {r:.3e} > {rtol:g}"
    return worst


# ------------------------------------------------------------------ host-driver cases
DRIVER_ARRAYS = ("h1", "h2", "h3", "h1x1i", "h2x1i", "h3x1i", "h1x2i", "h2x2i", "h3x2i",
                 "h1x3i", "h2x3i", "h3x3i", "dx1", "dx1i", "dx2", "dx2i", "dx3", "dx3i")


def write_case(d, g, st, dt, nsteps, mode, solver=2, Teinf=1500.0, atm=None):
    """Case files read by tests/drivers/*: case.txt (sizes, flags, step control) and
    in.bin (every input array, Fortran order, little-endian float64, concatenated).  With an
    atmosphere (modes 4 and 5) in.bin continues with nn(:,:,:,1:6), vn1, vn2, vn3,
    Prionize, Qeionize, the grid's alt and g1."""
    import os
    isg = " ".join(str(int(bool(v))) for v in g.isglobal)
    xtype = 1 if g.kind == "cartesian" else 2
    with open(os.path.join(d, "case.txt"), "w") as f:
        f.write(f"{g.lx1} {g.lx2} {g.lx3} {xtype} {int(g.gridflag)} {int(bool(g.flagper))} {isg} "
                f"{solver} {nsteps} {mode} {dt!r} {Teinf!r}\n")
    parts = [np.asarray(getattr(g, n), dtype="<f8") for n in DRIVER_ARRAYS]
    parts += [st.nn[..., :3], st.Tn, st.fluidvars, st.fluidauxvars, st.electrovars,
              st.vs1i, st.vs2i, st.vs3i]
    if atm is not None:
        parts += [atm.nn, atm.vn1, atm.vn2, atm.vn3, atm.Prionize, atm.Qeionize, g.alt, g.g1]
    with open(os.path.join(d, "in.bin"), "wb") as f:
        for a in parts:
            f.write(np.asarray(a, dtype="<f8").tobytes(order="F"))


def read_case_output(d, st):
    """Load out.bin (fluidvars, rhovs1/rhoes, vs1i, vs2i, vs3i) into a copy of st."""
    import os
    out = st.copy()
    raw = np.fromfile(os.path.join(d, "out.bin"), dtype="<f8")
    pos = 0
    for name, sl in (("fluidvars", None), ("fluidauxvars", 14), ("vs1i", None), ("vs2i", None),
                     ("vs3i", None)):
        a = getattr(out, name)
        view = a[..., :sl] if sl else a
        n = view.size
        view[...] = raw[pos:pos + n].reshape(view.shape, order="F")
        pos += n
    assert pos == raw.size
    return out
