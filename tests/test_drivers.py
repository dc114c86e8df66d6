"""Host drivers in the reference's host languages, on the GPU: a Fortran program through
the iso_c_binding module gemini_amd/fortran/gemhip_bind.f90 and a C++ program through
include/gemini3d_hip.h, each calling the reference-named entry points in the order of
src/libgemini.cpp:155-208 (mode 0) or the fused step (mode 1), and the same for the whole
transport block of src/libgemini.cpp:155-210 (modes 2 and 3).  Their results must be
bit-identical to the oracle's (with the GPU's x**2.5).  Modes 4 and 5 follow the block with
the sources substep of :223-232, call by call and fused; its collision rates call exp, erf
and log, so those runs are held to 1e-12 (normwise, per species) instead.  Every run ends
with the host's own dateinc_C, plasma_output_nompi_C and gemhip_output_plasma calls: the
files, named for the advanced clock, must hold the driver's final state bit for bit."""
import os
import subprocess

import numpy as np
import pytest

import gemini_amd as ga
import oracle
from helpers import compare, write_case, read_case_output
import h5lite

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "drivers", "bin")
ALL = ("ns", "vs1", "vs2", "vs3", "Ts", "rhovs1", "rhoes")


@pytest.mark.parametrize("driver", ["fortran_driver", "capi_driver"])
@pytest.mark.parametrize("kind,shape,mode", [("cart", (16, 6, 5), 0), ("dipole", (20, 9, 5), 0),
                                             ("dipole", (20, 9, 5), 1), ("cart", (16, 6, 5), 2),
                                             ("dipole", (20, 9, 5), 3), ("dipole", (20, 9, 5), 4),
                                             ("cart", (16, 6, 5), 5)])
def test_host_driver_matches_oracle(tmp_path, driver, kind, shape, mode):
    exe = os.path.join(BIN, driver)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (gemini_amd.build.build_drivers)")
    print(f"case {driver} {kind} {shape} mode {mode}", flush=True)
    g = ga.cartesian_grid(*shape) if kind == "cart" else ga.dipole_grid(*shape)
    st = ga.synthetic_state(g, seed=3, perturb=True)
    dt = ga.cfl_dt(g, st)
    nsteps = 2
    atm = ga.synthetic_atmosphere(g, st, seed=3) if mode >= 4 else None
    write_case(str(tmp_path), g, st, dt, nsteps, mode, atm=atm)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = read_case_output(str(tmp_path), st)
    ref = st.copy()
    with oracle.pow_mode(1):
        for _ in range(nsteps):
            if mode >= 2:
                Q = np.zeros((g.lx1, g.lx2, g.lx3, 7), order="F")
                oracle.fluid_adv_block(g, ref, dt, Q=Q)
                if mode >= 4:
                    oracle.sources_block(g, ref, atm, dt, Q=Q)
            else:
                oracle.transport_step(g, ref, dt)
    if mode >= 4:
        print(f"  worst normwise rel {compare(got, ref, close=ALL, interior_close=False):.2e}")
    else:
        compare(got, ref, exact=ALL)
    if mode in (0, 2):
        for a in ("vs1i", "vs2i", "vs3i"):
            assert np.array_equal(getattr(got, a), getattr(ref, a)), a
    # the host's plasma_output_nompi_C call: the file holds the driver's final state
    flag = 1 + mode % 3
    ymd, ut = oracle.dateinc(dt, [2015, 4, 13], 12345.67891)      # the host's dateinc_C
    stem = oracle.filestem(ymd, ut)
    path = os.path.join(str(tmp_path), f"{stem}_{mode}.h5")
    want = oracle.output_pack(got, g.gridflag, flag, patch=True)
    with h5lite.File(path) as f:
        assert f.read("/flagoutput") == flag and f.read("/patchID") == mode
        for name, arr in want.items():
            a = f.read(name)
            assert a.shape == arr.shape and np.array_equal(a.view(np.uint32), arr.view(np.uint32)), name
    # and gemini.bin's own file (gemhip_output_plasma): plain date name, TEall
    root = os.path.join(str(tmp_path), f"{stem}.h5")
    want = oracle.output_pack(got, g.gridflag, flag, patch=False)
    with h5lite.File(root) as f:
        assert "/patchID" not in f.names()
        for name, arr in want.items():
            a = f.read("/TEall" if name == "/Teall" else name)
            assert np.array_equal(a.view(np.uint32), arr.view(np.uint32)), name
