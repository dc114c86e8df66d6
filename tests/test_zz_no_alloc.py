"""GPU: the allocation-free energy diffusion (gemhip_energy_diffusion_noalloc) through the
C ABI, against energy_diffusion_C on the same inputs, bit for bit over every ghosted
cell of the seven temperature fields: two calls in a row on one GPU (they share the static
work space), the second on a smaller grid; then a grid too large for the work space is
refused with an error and no launch."""
import ctypes as C

import numpy as np
import pytest
import torch

import gemini_amd as ga
from gemini_amd.transport import Transport

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _device(a):
    """HBM copy of a Fortran-ordered host block, same bytes."""
    return torch.tensor(np.asfortranarray(a).ravel(order="F"), device=DEV)


def _noalloc(tr, Ts, ns, J1, dt):
    tr._sel()
    f = tr.lib.gemhip_energy_diffusion_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]
    torch.cuda.synchronize()
    rc = f(Ts.data_ptr(), ns.data_ptr(), J1.data_ptr(), dt)
    tr.lib.gemhip_synchronize()
    return rc


@pytest.mark.parametrize("solver", [2, 1])
def test_noalloc_matches_energy_diffusion(solver):
    for shape, kind in (((64, 20, 9), "dipole"), ((24, 10, 7), "cart")):   # second smaller
        print(f"case {kind} {shape} solver {solver}", flush=True)
        g = ga.dipole_grid(*shape) if kind == "dipole" else ga.cartesian_grid(*shape)
        st = ga.synthetic_state(g, seed=5, perturb=kind == "cart")
        dt = ga.cfl_dt(g, st)
        tr = Transport(g, st.nn, st.Tn, own_context=True)
        try:
            tr.configure(solver, 1500.0)
            ref = st.copy()
            tr.energy_diffusion(ref, dt)                       # energy_diffusion_C
            Ts = _device(st.fluidvars[..., 28:35])
            ns = _device(st.fluidvars[..., 0:7])
            J1 = _device(st.electrovars[..., 3])
            assert _noalloc(tr, Ts, ns, J1, dt) == 0, tr.lib.gemhip_last_error()
            got = Ts.cpu().numpy().reshape((7,) + g.padded[::-1]).transpose(3, 2, 1, 0)
            want = ref.fluidvars[..., 28:35]
            assert np.array_equal(got, want, equal_nan=True), \
                f"{kind} {shape}: {np.sum(got != want)} values differ"
            assert np.array_equal(ns.cpu().numpy(), np.asfortranarray(st.fluidvars[..., 0:7]).ravel(order="F"))
        finally:
            tr.close()


def test_noalloc_refuses_grids_beyond_its_work_space():
    g = ga.cartesian_grid(512, 64, 32)       # 224 column groups of 512 rows: ~530 MB
    st = ga.empty_state(g)
    tr = Transport(g, st.nn, st.Tn, own_context=True)
    try:
        z = torch.zeros(7 * g.padded[0] * g.padded[1] * g.padded[2], dtype=torch.float64, device=DEV)
        rc = _noalloc(tr, z, z, z, 0.1)
        assert rc == 4
        assert b"work space" in tr.lib.gemhip_last_error()
    finally:
        tr.close()
