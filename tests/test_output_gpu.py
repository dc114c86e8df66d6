"""GPU parity of plasma_output_nompi_C (src/io/io_nompi.f90:183-268): the HDF5 file the
library writes (the pack kernel's real32 datasets, copied to the host and written through
the HDF5 C library) holds bit for bit the oracle's values, under the reference's dataset
names, types and shapes, with its header datasets.  NaNs are compared by position (the
device and the host may give a NaN another payload).  The files are read back with the
tests' own HDF5 access (tests/h5lite.py)."""
import numpy as np
import pytest

import gemini_amd as ga
from gemini_amd.transport import Transport, GemhipError, read_initial_conditions
import oracle
import h5lite

pytestmark = pytest.mark.gpu

YMD, UTSEC, IDENT = [2015, 4, 13], 12345.678910, 42
LIMS = ((80e3, 1e6), (-2.5, 3.25), (0.0, 1.0e5))


def grid(kind):
    if kind == "dipole":
        return ga.dipole_grid(32, 12, 6)
    if kind == "cart_x2line":      # 2-D runs: a singleton x3 or x2
        return ga.cartesian_grid(40, 9, 1)
    if kind == "cart_x3line":
        return ga.cartesian_grid(40, 1, 9)
    b = ga.cartesian_grid(24, 10, 7)
    if kind == "inverted":   # gridflag 1: the potential slab is the x1 = 1 plane
        return ga.Grid(b.lx1, b.lx2, b.lx3, b.x1, b.x2, b.x3, b.planes, gridflag=1,
                       isglobal=b.isglobal, kind="cartesian")
    return b


def state(g, seed=5):
    st = ga.synthetic_state(g, seed=seed, perturb=True)
    rng = np.random.default_rng(seed)
    ev = st.electrovars
    ev[...] = rng.normal(size=ev.shape)
    ev[..., 3:6] *= 1e-6
    ev[..., 6] *= 1e3
    fv = st.fluidvars
    # real32 edge cases: subnormal, underflow, overflow, -0, NaN, a tie
    fv[3, 3, 3, 0] = 1e-45
    fv[4, 3, 3, 0] = 3e-39
    fv[5, 3, 3, 1] = 1e-50
    fv[3, 4, 3, 7] = 4e38
    fv[4, 4, 3, 9] = -0.0
    fv[3, 3, 4, 30] = np.nan
    fv[4, 3, 4, 28] = 1.0 + 2.0 ** -24
    return st


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(a[~na].view(np.uint32), b[~nb].view(np.uint32))


def check_file(path, st, gridflag, flag, root=False):
    """root: gemini.bin's file (plasma_output_hdf5.f90:10-65): no patch datasets, TEall."""
    want = oracle.output_pack(st, gridflag, flag, patch=not root)
    header = {"/flagoutput": ("i", ()), "/time/ymd": ("i", (3,)), "/time/UThour": ("f", ())}
    if root:
        want = {("/TEall" if k == "/Teall" else k): v for k, v in want.items()}
    else:
        header.update({"/x1lims": ("f", (2,)), "/x2lims": ("f", (2,)), "/x3lims": ("f", (2,)),
                       "/patchID": ("i", ())})
    with h5lite.File(path) as f:
        assert f.names() == sorted(list(header) + list(want))
        for name, (kind, shape) in header.items():
            assert f.info(name) == (kind, 4, True, shape), name
        assert f.read("/flagoutput") == flag
        assert list(f.read("/time/ymd")) == YMD
        assert f.read("/time/UThour").view(np.uint32) == np.float32(UTSEC / 3600.0).view(np.uint32)
        if not root:
            for k, lim in enumerate(LIMS):
                assert np.array_equal(f.read(f"/x{k + 1}lims"), np.array(lim, dtype=np.float32))
            assert f.read("/patchID") == IDENT
        for name, arr in want.items():
            assert f.info(name) == ("f", 4, True, arr.shape), name
            got = f.read(name)
            assert same(got, arr), f"{name}: differs from the oracle"


@pytest.mark.parametrize("kind", ["cart", "dipole", "inverted", "cart_x2line", "cart_x3line"])
@pytest.mark.parametrize("flag", [1, 2, 3])
@pytest.mark.parametrize("mode", [0, 1])
def test_output_file_matches_oracle(tmp_path, kind, flag, mode):
    print(f"case {kind} flagoutput {flag} residency {mode}", flush=True)
    g = grid(kind)
    st = state(g)
    tr = Transport(g, st.nn, st.Tn)
    tr.register_output(str(tmp_path), flag, 0 if kind in ("cart", "inverted") else 3)
    tr.set_residency(mode)
    if mode:
        tr.push(st)
    path = tr.plasma_output(st, YMD, UTSEC, IDENT, *LIMS)
    assert path == f"{tmp_path}/{oracle.filestem(YMD, UTSEC)}_{IDENT}.h5"
    check_file(path, st, g.gridflag, flag)


@pytest.mark.parametrize("flag", [1, 2, 3])
@pytest.mark.parametrize("mode", [0, 1])
def test_root_output_file_matches_oracle(tmp_path, flag, mode):
    """gemini.bin's own output file (output_plasma), as gemini3d.compare reads it."""
    print(f"case dipole root writer flagoutput {flag} residency {mode}", flush=True)
    g = grid("dipole")
    st = state(g)
    tr = Transport(g, st.nn, st.Tn)
    tr.register_output(str(tmp_path), flag, 3)
    tr.set_residency(mode)
    if mode:
        tr.push(st)
    path = tr.output_plasma(st, YMD, UTSEC)
    assert path == f"{tmp_path}/{oracle.filestem(YMD, UTSEC)}.h5"
    check_file(path, st, g.gridflag, flag, root=True)
    tr.set_residency(0)


def test_output_reads_the_resident_state(tmp_path):
    """With the state in HBM the file is written from there, after steps the host block
    has not seen."""
    g = grid("dipole")
    st = ga.synthetic_state(g, seed=2, perturb=True)
    for name in ("v12rhov1", "T2rhoe"):
        oracle.op(g, st, name)
    tr = Transport(g, st.nn, st.Tn)
    tr.register_output(str(tmp_path), 1, 0)
    tr.set_residency(1)
    tr.push(st)
    tr.fluid_adv_block(ga.cfl_dt(g, st))
    path = tr.plasma_output(st, YMD, UTSEC, IDENT, *LIMS)
    after = st.copy()
    tr.pull(after)
    assert not np.array_equal(after.fluidvars, st.fluidvars)
    check_file(path, after, g.gridflag, 1)


def test_output_reads_back_as_initial_conditions(tmp_path):
    g = grid("cart")
    st = state(g, seed=7)
    st.fluidvars[..., 30][np.isnan(st.fluidvars[..., 30])] = 1000.0
    tr = Transport(g, st.nn, st.Tn)
    tr.register_output(str(tmp_path), 1, 3)
    path = tr.plasma_output(st, YMD, UTSEC, IDENT, *LIMS)
    simsize = str(tmp_path / "simsize.h5")
    h5lite.write(simsize, {"/lx": np.array([g.lx1, g.lx2, g.lx3], dtype=np.int32)})
    back = ga.empty_state(g)
    read_initial_conditions(simsize, path, back)
    with np.errstate(over="ignore"):
        for blk in (0, 7, 28):
            a = st.fluidvars[2:-2, 2:-2, 2:-2, blk:blk + 7].astype(np.float32).astype(np.float64)
            b = back.fluidvars[2:-2, 2:-2, 2:-2, blk:blk + 7]
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), blk
    top = st.electrovars[g.lx1 + 1, 2:-2, 2:-2, 6].astype(np.float32).astype(np.float64)
    for i1 in range(2, g.lx1 + 2):
        assert np.array_equal(back.electrovars[i1, 2:-2, 2:-2, 6], top)


def test_output_errors(tmp_path):
    g = grid("cart")
    st = state(g)
    # a context of its own: the implicit one keeps the registration of earlier tests, as it
    # keeps the rest of the configuration
    tr = Transport(g, st.nn, st.Tn, own_context=True)
    with pytest.raises(GemhipError, match="gemhip_register_output"):
        tr.plasma_output(st, YMD, UTSEC, IDENT, *LIMS)
    tr.register_output(str(tmp_path / "no_such_dir"), 1, 0)
    with pytest.raises(GemhipError, match="could not create"):
        tr.plasma_output(st, YMD, UTSEC, IDENT, *LIMS)
    tr.register_output(str(tmp_path), 1, 0)
    with pytest.raises(GemhipError, match="utsec2filestem"):
        tr.plasma_output(st, YMD, -3.0, IDENT, *LIMS)
    path = tr.plasma_output(st, YMD, UTSEC, IDENT, *LIMS)   # the context still works
    check_file(path, st, g.gridflag, 1)
    tr.close()


def test_compressed_output_in_many_chunks(tmp_path):
    """x1-x2 planes over 1 MiB are stored in chunks of whole x1 rows, the last one partial;
    the chunks are shuffled and deflated on the host's threads and written pre-filtered,
    and any HDF5 reader decompresses them to the packed values."""
    print("case cart 600x500x3 flagoutput 1 deflate 3", flush=True)
    b = ga.cartesian_grid(600, 500, 3)
    st = ga.synthetic_state(b, seed=9)
    st.electrovars[..., 3:7] = np.random.default_rng(9).normal(size=st.electrovars[..., 3:7].shape)
    tr = Transport(b, st.nn, st.Tn)
    tr.register_output(str(tmp_path), 1, 3)
    tr.set_residency(1)
    tr.push(st)
    path = tr.plasma_output(None, YMD, UTSEC, IDENT, *LIMS, electrovars=st.electrovars)
    check_file(path, st, b.gridflag, 1)
