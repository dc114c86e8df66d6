"""Plasma output and initial-condition input in the reference's HDF5 formats (CPU part).

* The output file stem (src/temporal/timeutils.f90:133-196) of the oracle and of the
  library's host-only gemhip_output_filename against the golden strings of
  test/temporal/test_formats.f90:21-53 and hand-derived date carries and roundings.
* The oracle's restatement of plasma_output_nompi's field datasets (src/io/io_nompi.f90:
  183-268) pinned by what the arithmetic fixes: every written value is the nearest real32
  of the double (numpy's conversion), the ion averages reduce to closed forms on states
  built so that every product and partial sum is exact, a -0 sum starts from +0, and the
  potential slab is the top x1 plane of the grid's orientation.
* The library's initial-condition reader (gemhip_read_initial_conditions, io_nompi.f90:
  271-345) against files written here through the HDF5 C library (tests/h5lite.py):
  interiors land bit-exactly, ghost cells are zeroed, the potential is extended along x1,
  and malformed inputs are rejected.
* The grid size and coordinate readers (reader_hdf5.f90:64-92, readgrid_hdf5.f90:11-39):
  the coordinates written come back bit for bit, and the reference's errors are raised.
"""
import os

import numpy as np
import pytest

import gemini_amd as ga
from gemini_amd import transport as T
import oracle
import h5lite

# test/temporal/test_formats.f90:21-53 (the 86400 case is given as an integer there; the
# C binding passes real(wp) seconds, for which the same carry applies)
GOLDEN = [([2015, 4, 13], 12345.678910, "20150413_12345.680000"),
          ([2015, 4, 13], 345.678911, "20150413_00345.680000"),
          ([2013, 2, 20], 60.0, "20130220_00060.000000"),
          ([2015, 4, 13], 86400.0, "20150414_00000.000000")]
# date carries by the Gregorian leap rule (timeutils.f90:13-27) and 10 ms rounding
CARRIES = [([2016, 2, 28], 86400.0, "20160229_00000.000000"),
           ([2015, 2, 28], 86400.0, "20150301_00000.000000"),
           ([2000, 2, 28], 86400.0, "20000229_00000.000000"),
           ([2100, 2, 28], 86400.0, "21000301_00000.000000"),
           ([2016, 12, 31], 86400.0, "20170101_00000.000000"),
           ([2015, 4, 13], 0.004, "20150413_00000.000000"),
           ([2015, 4, 13], 0.006, "20150413_00000.010000"),
           ([2015, 4, 13], 12.125, "20150413_00012.130000"),   # 1212.5 rounds away from zero
           ([2015, 4, 13], 86399.994, "20150413_86399.990000"),
           ([2015, 4, 13], 86399.996, "20150414_00000.000000"),
           # just below zero the seconds truncate to 0 and the negative microsecond count
           # overflows the I6.6 edit descriptor, which prints asterisks
           ([2015, 4, 13], -0.01, "20150413_00000.******"),
           ([2015, 4, 13], -0.004, "20150413_00000.000000")]


@pytest.mark.parametrize("ymd,ut,stem", GOLDEN + CARRIES)
def test_file_stem_known_answers(ymd, ut, stem):
    assert oracle.filestem(ymd, ut) == stem
    assert T.output_filename("/out/dir", ymd, ut, 17) == f"/out/dir/{stem}_17.h5"
    assert T.output_filename("/o", ymd, ut, -5) == f"/o/{stem}_-5.h5"     # I0 of a negative
    assert T.output_filename_root("/out/dir", ymd, ut) == f"/out/dir/{stem}.h5"


@pytest.mark.parametrize("ut", [-1.0, -5.0, float("nan"), 1e12, 2.0e6])
def test_file_stem_rejects_times_the_reference_stops_on(ut):
    assert oracle.filestem([2015, 4, 13], ut) is None
    with pytest.raises(T.GemhipError, match="utsec2filestem"):
        T.output_filename("/o", [2015, 4, 13], ut, 0)


def test_file_stem_library_matches_oracle_on_random_times():
    rng = np.random.default_rng(3)
    for _ in range(400):
        ymd = [int(rng.integers(1700, 2400)), int(rng.integers(1, 13)), int(rng.integers(1, 29))]
        ut = float(rng.uniform(0, 86400.0)) if rng.random() < 0.9 else float(rng.integers(0, 8640001)) / 100
        ident = int(rng.integers(-5, 100000))
        assert T.output_filename(".", ymd, ut, ident) == f"./{oracle.filestem(ymd, ut)}_{ident}.h5"


# ---------------------------------------------------------------- output restatement pins
def _state(lx1=6, lx2=4, lx3=3, seed=0):
    g = ga.cartesian_grid(lx1, lx2, lx3)
    st = ga.synthetic_state(g, seed=seed)
    rng = np.random.default_rng(seed)
    st.electrovars[...] = rng.normal(size=st.electrovars.shape) * 1e-6
    return g, st


def _bits(a):
    return np.asarray(a, dtype=np.float32).view(np.uint32)


def _interior(a):
    """Interior of a ghosted Fortran field (..., lsp) as the C-order dataset array."""
    return np.ascontiguousarray(a[2:-2, 2:-2, 2:-2].T)


def test_each_value_is_the_nearest_real32():
    g, st = _state()
    fv = st.fluidvars
    # values at the edges of real32: subnormal, underflow to zero, overflow, -0, NaN
    fv[2, 2, 2, 0] = 1e-45
    fv[3, 2, 2, 0] = 3e-39
    fv[4, 2, 2, 0] = 1e-50
    fv[2, 3, 2, 7] = 4e38
    fv[3, 3, 2, 7] = -0.0
    fv[2, 2, 3, 28] = np.nan
    fv[3, 2, 3, 28] = 1.0 + 2.0 ** -24          # halfway: ties to even
    fv[4, 2, 3, 28] = 1.0 + 3 * 2.0 ** -24
    out = oracle.output_pack(st, g.gridflag, 1, patch=False)
    for name, k in (("/nsall", 0), ("/vs1all", 7), ("/Tsall", 28)):
        with np.errstate(over="ignore"):
            want = _interior(fv[..., k:k + 7]).astype(np.float32)
        assert np.array_equal(_bits(out[name]), _bits(want)), name
    for j, name in ((3, "/J1all"), (4, "/J2all"), (5, "/J3all")):
        want = _interior(st.electrovars[..., j]).astype(np.float32)
        assert np.array_equal(_bits(out[name]), _bits(want)), name
    out2 = oracle.output_pack(st, g.gridflag, 2, patch=True)
    out3 = oracle.output_pack(st, g.gridflag, 3, patch=True)
    ne = _interior(fv[..., 6]).astype(np.float32)
    assert np.array_equal(_bits(out2["/neall"]), _bits(ne))
    assert np.array_equal(_bits(out3["/neall"]), _bits(ne))
    assert np.array_equal(_bits(out2["/Teall"]), _bits(_interior(fv[..., 34]).astype(np.float32)))


def _exact_species_state(seed):
    """ns of the six ions 2^-1..2^-5, 2^-5 of ne = 2^37 (sum exactly ne) and per-cell
    velocities and temperatures with few significant bits, so every product and partial
    sum of the averages is exact."""
    g, st = _state(seed=seed)
    rng = np.random.default_rng(seed + 10)
    fv = st.fluidvars
    ne = 2.0 ** 37
    frac = [2.0 ** -1, 2.0 ** -2, 2.0 ** -3, 2.0 ** -4, 2.0 ** -5, 2.0 ** -5]
    for isp in range(6):
        fv[..., isp] = ne * frac[isp]
    fv[..., 6] = ne
    shape = fv.shape[:3]
    comps = {}
    for blk, name in ((7, "v1"), (14, "v2"), (21, "v3"), (28, "T")):
        v = rng.integers(-4000, 4000, size=shape) * 0.25
        if name == "T":
            v = np.abs(v) + 100.0
        comps[name] = v
        for isp in range(6):
            fv[..., blk + isp] = v
        fv[..., blk + 6] = 1e4 + v          # electrons differ: an average including them fails
    return g, st, comps


def test_averages_reduce_to_the_common_ion_value():
    g, st, comps = _exact_species_state(1)
    o1 = oracle.output_pack(st, g.gridflag, 1, patch=True)
    o2 = oracle.output_pack(st, g.gridflag, 2, patch=True)
    for key, name in (("v2", "/v2avgall"), ("v3", "/v3avgall")):
        want = _interior(comps[key]).astype(np.float32)
        assert np.array_equal(_bits(o1[name]), _bits(want)), name
        assert np.array_equal(_bits(o2[name]), _bits(want)), name
    assert np.array_equal(_bits(o2["/v1avgall"]), _bits(_interior(comps["v1"]).astype(np.float32)))
    assert np.array_equal(_bits(o2["/Tavgall"]), _bits(_interior(comps["T"]).astype(np.float32)))


@pytest.mark.parametrize("ion", range(6))
def test_average_of_a_single_ion_is_its_density_share(ion):
    g, st, comps = _exact_species_state(2)
    fv = st.fluidvars
    for isp in range(6):
        fv[..., isp] = 0.0
    fv[..., ion] = 2.0 ** 36                 # half of ne
    fv[..., 14 + ion] = 3.0 * fv[..., 14 + ion] + 0.5
    o = oracle.output_pack(st, g.gridflag, 2, patch=True)
    want = _interior((3.0 * comps["v2"] + 0.5) / 2).astype(np.float32)
    assert np.array_equal(_bits(o["/v2avgall"]), _bits(want))


def test_negative_zero_products_sum_to_positive_zero():
    g, st, _ = _exact_species_state(3)
    st.fluidvars[..., 21:27] = -0.0
    o = oracle.output_pack(st, g.gridflag, 1, patch=True)
    assert np.all(_bits(o["/v3avgall"]) == 0)


@pytest.mark.parametrize("gridflag", [0, 1, 2])
def test_potential_slab_is_the_top_x1_plane(gridflag):
    g, st = _state(lx1=7, lx2=5, lx3=4)
    L1, L2, L3 = st.electrovars.shape[:3]
    i1, i2, i3 = np.meshgrid(np.arange(L1) - 1, np.arange(L2) - 1, np.arange(L3) - 1, indexing="ij")
    st.electrovars[..., 6] = 1000.0 * i1 + i2 + 0.125 * i3
    top = 1 if gridflag == 1 else g.lx1
    for flag in (1, 2, 3):
        phi = oracle.output_pack(st, gridflag, flag, patch=True)["/Phiall"]
        i2c, i3c = np.meshgrid(np.arange(1, g.lx2 + 1), np.arange(1, g.lx3 + 1), indexing="xy")
        assert np.array_equal(phi, (1000.0 * top + i2c + 0.125 * i3c).astype(np.float32))


def test_patch_writer_reads_the_currents_from_the_ghost_origin():
    """plasma_output_nompi declares J1-J3 dimension(1:,1:,1:) (io_nompi.f90:192) and is
    handed the ghosted fields (libgemini.f90:549-551): its cell (i1, i2, i3) is the
    ghosted field at (i1-2, i2-2, i3-2).  The root writer keeps ghosted bounds
    (plasma_output.f90:46) and writes the interior.  Everything else is the same."""
    g, st = _state(lx1=7, lx2=5, lx3=4)
    L1, L2, L3 = st.electrovars.shape[:3]
    i1, i2, i3 = np.meshgrid(np.arange(L1) - 1, np.arange(L2) - 1, np.arange(L3) - 1, indexing="ij")
    for j in (3, 4, 5):
        st.electrovars[..., j] = 1000.0 * i1 + 10.0 * i2 + 0.125 * i3 + j
    c1, c2, c3 = np.meshgrid(np.arange(1, g.lx1 + 1), np.arange(1, g.lx2 + 1),
                             np.arange(1, g.lx3 + 1), indexing="ij")
    for flag in (1, 2):
        patch = oracle.output_pack(st, g.gridflag, flag, patch=True)
        root = oracle.output_pack(st, g.gridflag, flag, patch=False)
        for j, name in ((3, "/J1all"), (4, "/J2all"), (5, "/J3all")):
            for out, shift in ((patch, 2), (root, 0)):
                want = 1000.0 * (c1 - shift) + 10.0 * (c2 - shift) + 0.125 * (c3 - shift) + j
                assert np.array_equal(out[name], np.ascontiguousarray(want.T).astype(np.float32))
        for k in patch:
            if k not in ("/J1all", "/J2all", "/J3all"):
                assert np.array_equal(_bits(patch[k]), _bits(root[k])), k


def test_other_flags_write_the_full_set():
    g, st = _state()
    full = oracle.output_pack(st, g.gridflag, 1, patch=True)
    for flag in (0, 4, -1):
        o = oracle.output_pack(st, g.gridflag, flag, patch=True)
        assert list(o) == list(full)
        for k in full:
            assert np.array_equal(_bits(o[k]), _bits(full[k]))
    n, m = g.lx1 * g.lx2 * g.lx3, g.lx2 * g.lx3
    assert oracle.output_count(g.lx1, g.lx2, g.lx3, 1) == 26 * n + m
    assert oracle.output_count(g.lx1, g.lx2, g.lx3, 2) == 9 * n + m
    assert oracle.output_count(g.lx1, g.lx2, g.lx3, 3) == n + m


# ---------------------------------------------------------------- initial conditions
def _garbage_state(g):
    st = ga.empty_state(g)
    st.fluidvars[...] = np.nan
    st.electrovars[...] = -7.0
    return st


def _ic_arrays(g, seed=0):
    rng = np.random.default_rng(seed)
    shp = (7, g.lx3, g.lx2, g.lx1)
    return {"/nsall": rng.uniform(1e6, 1e12, shp), "/vs1all": rng.normal(size=shp) * 300,
            "/Tsall": rng.uniform(200, 3000, shp), "/Phiall": rng.normal(size=(g.lx3, g.lx2)) * 1e3}


def _check_loaded(st, data, phislab):
    fv = st.fluidvars
    for name, blk in (("/nsall", 0), ("/vs1all", 7), ("/Tsall", 28)):
        a = fv[..., blk:blk + 7]
        assert np.array_equal(_interior(a), np.asarray(data[name], dtype=np.float64)), name
        ghost = a.copy()
        ghost[2:-2, 2:-2, 2:-2] = 0
        assert np.all(ghost == 0), f"{name} ghost cells not zeroed"
    assert np.all(np.isnan(fv[..., 14:28])), "vs2/vs3 must be left alone"
    phi = st.electrovars[..., 6]
    for i1 in range(2, phi.shape[0] - 2):
        assert np.array_equal(phi[i1, 2:-2, 2:-2].T, phislab)
    ghost = phi.copy()
    ghost[2:-2, 2:-2, 2:-2] = 0
    assert np.all(ghost == 0)
    assert np.all(st.electrovars[..., :6] == -7.0), "only the potential is written"


@pytest.mark.parametrize("sizes", ["lx123", "lxs", "lx"])
def test_initial_conditions_land_in_the_interior(tmp_path, sizes):
    g = ga.cartesian_grid(9, 5, 4)
    lx = np.array([g.lx1, g.lx2, g.lx3], dtype=np.int32)
    size_sets = {"lx123": {"/lx1": lx[0], "/lx2": lx[1], "/lx3": lx[2]},
                 "lxs": {"/lxs": lx}, "lx": {"/lx": lx}}
    simsize, ic = str(tmp_path / "simsize.h5"), str(tmp_path / "initial_conditions.h5")
    h5lite.write(simsize, size_sets[sizes])
    data = _ic_arrays(g)
    # the velocity stored as real32, as some writers do: read back widened exactly
    h5lite.write(ic, data, file_types={"/vs1all": "f4"})
    data["/vs1all"] = data["/vs1all"].astype(np.float32).astype(np.float64)
    st = _garbage_state(g)
    T.read_initial_conditions(simsize, ic, st)
    _check_loaded(st, data, data["/Phiall"])


@pytest.mark.parametrize("shape", [(8, 1, 6), (8, 6, 1)])
def test_one_dimensional_potential(tmp_path, shape):
    g = ga.cartesian_grid(*shape)
    simsize, ic = str(tmp_path / "s.h5"), str(tmp_path / "i.h5")
    h5lite.write(simsize, {"/lx": np.array(shape, dtype=np.int32)})
    data = _ic_arrays(g, seed=1)
    line = np.arange(max(g.lx2, g.lx3), dtype=np.float64) * 2.5 - 3.0
    data["/Phiall"] = line
    h5lite.write(ic, data)
    st = _garbage_state(g)
    T.read_initial_conditions(simsize, ic, st)
    _check_loaded(st, data, line.reshape(g.lx3, g.lx2))


def test_absent_potential_is_zero(tmp_path):
    g = ga.cartesian_grid(6, 3, 2)
    simsize, ic = str(tmp_path / "s.h5"), str(tmp_path / "i.h5")
    h5lite.write(simsize, {"/lx": np.array([6, 3, 2], dtype=np.int32)})
    data = _ic_arrays(g, seed=2)
    del data["/Phiall"]
    h5lite.write(ic, data)
    st = _garbage_state(g)
    T.read_initial_conditions(simsize, ic, st)
    _check_loaded(st, data, np.zeros((2, 3)))


def test_malformed_inputs_are_rejected(tmp_path):
    g = ga.cartesian_grid(6, 3, 2)
    simsize, ic = str(tmp_path / "s.h5"), str(tmp_path / "i.h5")
    data = _ic_arrays(g, seed=3)
    h5lite.write(ic, data)
    st = _garbage_state(g)
    h5lite.write(simsize, {"/lx": np.array([6, 3, 3], dtype=np.int32)})
    with pytest.raises(T.GemhipError, match="same size as the grid"):
        T.read_initial_conditions(simsize, ic, st)
    h5lite.write(simsize, {"/nothing": np.int32(1)})
    with pytest.raises(T.GemhipError, match="no grid size"):
        T.read_initial_conditions(simsize, ic, st)
    with pytest.raises(T.GemhipError, match="could not open"):
        T.read_initial_conditions(str(tmp_path / "missing.h5"), ic, st)
    h5lite.write(simsize, {"/lx": np.array([6, 3, 2], dtype=np.int32)})
    bad = dict(data)
    bad["/Tsall"] = data["/Tsall"][:6]
    h5lite.write(ic, bad)
    with pytest.raises(T.GemhipError, match="Tsall must have the shape"):
        T.read_initial_conditions(simsize, ic, st)
    bad = dict(data)
    bad["/Phiall"] = data["/Phiall"].T.copy()
    h5lite.write(ic, bad)
    with pytest.raises(T.GemhipError, match="Phiall must have the shape"):
        T.read_initial_conditions(simsize, ic, st)
    bad["/Phiall"] = np.zeros(3)                # 1-D on a grid with lx2, lx3 > 1
    h5lite.write(ic, bad)
    with pytest.raises(T.GemhipError, match="1-D /Phiall"):
        T.read_initial_conditions(simsize, ic, st)


def test_oracle_output_file_reads_back_as_initial_conditions(tmp_path):
    """An output file (real32 nsall/vs1all/Tsall and the potential slab) is a valid input:
    the interiors come back as the real32 values widened, the potential as its slab."""
    g, st = _state(lx1=8, lx2=5, lx3=3, seed=4)
    out = oracle.output_pack(st, g.gridflag, 1, patch=True)
    f = str(tmp_path / "out.h5")
    h5lite.write(f, out)
    simsize = str(tmp_path / "s.h5")
    h5lite.write(simsize, {"/lx1": np.int32(8), "/lx2": np.int32(5), "/lx3": np.int32(3)})
    back = _garbage_state(g)
    T.read_initial_conditions(simsize, f, back)
    data = {k: out[k].astype(np.float64) for k in ("/nsall", "/vs1all", "/Tsall")}
    _check_loaded(back, data, out["/Phiall"].astype(np.float64))


# ---------------------------------------------------------------- grid size and coordinates
@pytest.mark.parametrize("kind", ["cartesian", "dipole"])
def test_grid_files_give_the_grid_back(tmp_path, kind):
    """simsize.h5 and simgrid.h5 (src/io/reader_hdf5.f90:64-92, readgrid_hdf5.f90:11-39)
    read back to the coordinates written, bit for bit; the mesh built on them is the
    generator's own."""
    g = ga.cartesian_grid(12, 5, 4) if kind == "cartesian" else ga.dipole_grid(16, 6, 3)
    d = tmp_path / "inputs"
    d.mkdir()
    h5lite.write(str(d / "simsize.h5"), {"/lx1": np.int32(g.lx1), "/lx2": np.int32(g.lx2),
                                         "/lx3": np.int32(g.lx3)})
    h5lite.write(str(d / "simgrid.h5"), {"/x1": g.x1, "/x2": g.x2, "/x3": g.x3,
                                         "/glonctr": np.float64(212.5), "/glatctr": np.float64(65.1)})
    assert T.read_simsize(str(d / "simsize.h5")) == (g.lx1, g.lx2, g.lx3)
    for path in (str(d), str(d / "simgrid.h5")):       # directory or file, as the reference
        x1, x2, x3, ctr = T.read_grid_coords(path, (g.lx1, g.lx2, g.lx3))
        assert np.array_equal(x1, g.x1) and np.array_equal(x2, g.x2) and np.array_equal(x3, g.x3)
        assert ctr == (212.5, 65.1)
    back = T.grid_from_files(str(d / "simsize.h5"), str(d), kind)
    for name in T.Transport.GEOM + ("dx1", "dx1i", "dx2", "dx2i", "dx3", "dx3i", "alt", "g1"):
        assert np.array_equal(getattr(back, name), getattr(g, name)), name
    assert (back.gridflag, tuple(back.isglobal)) == (g.gridflag, tuple(g.isglobal))


def test_grid_file_errors(tmp_path):
    g = ga.cartesian_grid(8, 4, 3)
    f = str(tmp_path / "simgrid.h5")
    h5lite.write(f, {"/x1": g.x1, "/x2": g.x2, "/x3": g.x3})
    with pytest.raises(T.GemhipError, match="please specify /glonctr and /glatctr"):
        T.read_grid_coords(f, (8, 4, 3))
    h5lite.write(f, {"/x1": g.x1[2:-2], "/x2": g.x2, "/x3": g.x3, "/glonctr": np.float64(0),
                     "/glatctr": np.float64(0)})
    with pytest.raises(T.GemhipError, match="/x1 must hold 12 values"):
        T.read_grid_coords(f, (8, 4, 3))
    h5lite.write(f, {"/x1": g.x1, "/x2": g.x2, "/x3": g.x3, "/glonctr": np.zeros(2),
                     "/glatctr": np.float64(0)})
    with pytest.raises(T.GemhipError, match="glonctr must be one value"):
        T.read_grid_coords(f, (8, 4, 3))
    with pytest.raises(T.GemhipError, match="could not open"):
        T.read_grid_coords(str(tmp_path / "nowhere"), (8, 4, 3))


# ---------------------------------------------------------------- the writer on its own
@pytest.mark.parametrize("level,threads", [(0, "1"), (3, "1"), (3, "3"), (6, "16")])
def test_writer_chunks_and_threads(tmp_path, level, threads):
    """The library's HDF5 writer (tests/drivers/h5io_check runs it on host arrays): datasets
    whose 1 MiB chunks end partly outside them, compressed on 1 to 16 threads, read back
    to the values written, with the shuffle+deflate pipeline declared in the file."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "drivers", "bin", "h5io_check")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (gemini_amd.build.build_drivers)")
    f = str(tmp_path / "w.h5")
    r = subprocess.run([exe, f, str(level)], env=dict(os.environ, GEMHIP_IO_THREADS=threads),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(f + ".raw", dtype=np.float32)
    a, b = raw[:3 * 700 * 600].reshape(3, 700, 600), raw[3 * 700 * 600:].reshape(5, 3, 450, 530)
    with h5lite.File(f) as h:
        assert h.names() == ["/a", "/c", "/g/b", "/i"]
        assert np.array_equal(h.read("/a"), a) and np.array_equal(h.read("/g/b"), b)
        assert h.read("/c").tolist() == [[3.5]] and h.read("/i") == 7
    h5dump = "/opt/conda/bin/h5dump"        # the HDF5 tools' own view of the storage layout
    dump = subprocess.run([h5dump, "-p", "-H", f], capture_output=True, text=True) \
        if os.path.exists(h5dump) else None
    if dump is not None and dump.returncode == 0:
        text = dump.stdout
        if level:
            assert "CHUNKED ( 1, 436, 600 )" in text and "CHUNKED ( 1, 1, 450, 530 )" in text
            assert "PREPROCESSING SHUFFLE" in text and f"DEFLATE {{ LEVEL {level} }}" in text
        else:
            assert "CONTIGUOUS" in text and "DEFLATE" not in text


# ---------------------------------------------------------------- dateinc
def test_dateinc_known_answers():
    """test/temporal/test_formats.f90:35-49: a 10 ms step, then 60 one-second steps."""
    for inc in (oracle.dateinc, T.dateinc):
        ymd, ut = inc(0.01, [2013, 2, 20], 18000.0)
        assert oracle.filestem(ymd, ut) == "20130220_18000.010000"
        assert T.output_filename_root(".", ymd, ut) == "./20130220_18000.010000.h5"
        for _ in range(60):
            ymd, ut = inc(1.0, ymd, ut)
        assert oracle.filestem(ymd, ut) == "20130220_18060.010000"


@pytest.mark.parametrize("ymd,ut,dt,want", [
    ([2016, 12, 31], 86399.5, 1.0, ([2017, 1, 1], 0.5)),
    ([2016, 2, 28], 86000.0, 500.0, ([2016, 2, 29], 100.0)),
    ([2015, 2, 28], 86000.0, 500.0, ([2015, 3, 1], 100.0)),
    ([2015, 4, 30], 0.0, 86400.0, ([2015, 5, 1], 0.0)),
    ([2015, 4, 13], 100.0, 0.0, ([2015, 4, 13], 100.0))])
def test_dateinc_carries_days(ymd, ut, dt, want):
    assert oracle.dateinc(dt, ymd, ut) == want
    assert T.dateinc(dt, ymd, ut) == want


@pytest.mark.parametrize("ymd,ut,dt", [([2015, 4, 0], 0.0, 1.0), ([2015, 4, 1], -1.0, 1.0),
                                       ([2015, 4, 1], 0.0, -1.0), ([2015, 4, 1], 0.0, 86401.0)])
def test_dateinc_rejects_what_the_reference_stops_on(ymd, ut, dt):
    assert oracle.dateinc(dt, ymd, ut) is None
    with pytest.raises(T.GemhipError):
        T.dateinc(dt, ymd, ut)


def test_dateinc_library_matches_oracle():
    rng = np.random.default_rng(5)
    for _ in range(300):
        ymd = [int(rng.integers(1700, 2400)), int(rng.integers(1, 13)), int(rng.integers(1, 29))]
        ut, dt = float(rng.uniform(0, 86400)), float(rng.uniform(0, 86400))
        assert T.dateinc(dt, ymd, ut) == oracle.dateinc(dt, ymd, ut)


def test_library_leaves_the_hosts_hdf5_error_handler_alone():
    """HDF5's automatic error printing is process-wide; the library turns it off only while
    its own calls run, so a host's handler is the same before and after a failing call.
    Run in a fresh process, so the library's first HDF5 call (its set-up) is inside it."""
    import subprocess
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    code = (
        "import ctypes as C, sys\n"
        f"sys.path[:0] = [{os.path.dirname(here)!r}, {here!r}]\n"
        "import h5lite\n"
        "from gemini_amd import transport as T\n"
        "L = h5lite.lib()\n"
        "handler = C.cast(L.H5Eprint2, C.c_void_p)\n"
        "L.H5Eset_auto2(h5lite.hid_t(0), handler, None)\n"
        "try:\n"
        "    T.read_simsize('/nonexistent/simsize.h5')\n"
        "except T.GemhipError:\n"
        "    pass\n"
        "f, d = C.c_void_p(), C.c_void_p()\n"
        "L.H5Eget_auto2(h5lite.hid_t(0), C.byref(f), C.byref(d))\n"
        "sys.exit(0 if f.value == handler.value and d.value is None else 1)\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
