"""Python front end of the HIP transport library (ctypes over include/gemini3d_hip.h).

``Transport`` mirrors the hot-path interface the reference exposes to its drivers
(src/libgemini.f90:731-1075 ``*_in`` wrappers and the C shims of
src/libgemini_c.f90:445-848): the same operation names, acting in place on the
reference's state blocks (a :class:`gemini_amd.state.State`).  Every call goes
through the C ABI into libgemhip.so; there is no other implementation behind it,
and a missing or unloadable library raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .grid import Grid
from .state import State

HERE = os.path.dirname(os.path.abspath(__file__))
# GEMHIP_LIBRARY selects another build of the same library (e.g. one compiled with
# different tile or pipeline-depth settings)
LIB_PATH = os.environ.get("GEMHIP_LIBRARY") or os.path.join(HERE, "libgemhip.so")

_dp = C.POINTER(C.c_double)
_lib = None


class GemhipError(RuntimeError):
    pass


def load_library(path: str = LIB_PATH):
    """Load libgemhip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise GemhipError(f"{path} is missing: run `python -m gemini_amd.build` first")
    lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
    for name in ("gemhip_init", "gemhip_set_grid_size", "gemhip_register_grid",
                 "gemhip_register_neutrals", "gemhip_register_config", "gemhip_set_residency",
                 "gemhip_state_push", "gemhip_state_pull", "gemhip_get_interface_vels",
                 "gemhip_set_interface_vels", "gemhip_transport_step", "gemhip_device_state",
                 "gemhip_synchronize", "gemhip_kernel_timing", "gemhip_kernel_time",
                 "gemhip_kernel_timing_reset", "gemhip_eval_math"):
        getattr(lib, name).restype = C.c_int
    for name in ("gemhip_context_create", "gemhip_context_select", "gemhip_context_current",
                 "gemhip_context_destroy", "gemhip_set_tuning", "gemhip_get_tuning",
                 "gemhip_ensemble_capture", "gemhip_ensemble_launch", "gemhip_ensemble_release"):
        getattr(lib, name).restype = C.c_int
    for name in ("gemhip_register_null_cells", "gemhip_fluid_adv_block", "gemhip_register_cfl",
                 "gemhip_grid_rates"):
        getattr(lib, name).restype = C.c_int
    lib.gemhip_fluid_adv_block.argtypes = [C.c_double]
    for name in ("gemhip_sources_block", "gemhip_register_atmosphere", "gemhip_register_gravity",
                 "gemhip_set_ionization_rates"):
        getattr(lib, name).restype = C.c_int
    lib.gemhip_sources_block.argtypes = [C.c_double]
    lib.gemhip_register_cfl.argtypes = [C.c_double, C.c_int]
    for name in ("gemhip_register_conductivity_scale", "gemhip_register_top_heat_flux",
                 "gemhip_checkpoint_save", "gemhip_checkpoint_load",
                 "gemhip_adjoint_record_step", "gemhip_adjoint_step"):
        getattr(lib, name).restype = C.c_int
    for name in ("gemhip_checkpoint_size", "gemhip_adjoint_tape_size"):
        getattr(lib, name).restype = C.c_long
    lib.gemhip_adjoint_record_step.argtypes = [C.c_double, C.c_void_p]
    lib.gemhip_adjoint_step.argtypes = [C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p]
    for name in ("gemhip_batch_create", "gemhip_batch_set_dt", "gemhip_batch_set_tuning",
                 "gemhip_batch_step", "gemhip_batch_synchronize", "gemhip_batch_capture",
                 "gemhip_batch_release"):
        getattr(lib, name).restype = C.c_int
    lib.gemhip_batch_set_tuning.argtypes = [C.c_int, C.c_char_p, C.c_int]
    lib.gemhip_set_tuning.argtypes = [C.c_char_p, C.c_int]
    lib.gemhip_get_tuning.argtypes = [C.c_char_p]
    lib.gemhip_last_error.restype = C.c_char_p
    lib.gemhip_stream.restype = C.c_void_p
    lib.gemhip_transport_step.argtypes = [C.c_double]
    lib.gemhip_register_config.argtypes = [C.c_int, C.c_double]
    lib.gemhip_kernel_time.argtypes = [C.c_char_p, _dp, C.POINTER(C.c_long)]
    lib.gemhip_eval_math.argtypes = [C.c_int, _dp, _dp, C.c_long]
    lib.gemhip_output_plasma.restype = C.c_int
    lib.gemhip_output_plasma.argtypes = [C.POINTER(C.c_int), C.c_double, C.c_void_p, C.c_void_p]
    lib.gemhip_output_filename_root.restype = C.c_int
    lib.gemhip_output_filename_root.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.c_double,
                                                C.c_char_p, C.c_long]
    for name in ("gemhip_register_output", "gemhip_output_filename",
                 "gemhip_read_initial_conditions", "gemhip_read_simsize",
                 "gemhip_read_grid_coords"):
        getattr(lib, name).restype = C.c_int
    lib.gemhip_read_simsize.argtypes = [C.c_char_p] + [C.POINTER(C.c_int)] * 3
    lib.gemhip_read_grid_coords.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp,
                                            _dp, _dp]
    lib.gemhip_register_output.argtypes = [C.c_char_p, C.c_int, C.c_int]
    lib.gemhip_output_filename.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.c_double, C.c_int,
                                           C.c_char_p, C.c_long]
    lib.gemhip_read_initial_conditions.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int,
                                                   C.c_int, _dp, _dp]
    _lib = lib
    return lib


def _host_call(name, *args):
    """Call a host-only entry point (no device context needed) and raise on its error."""
    lib = load_library()
    lib.gemhip_set_error_mode(1)
    rc = getattr(lib, name)(*args)
    msg = lib.gemhip_last_error().decode()
    if msg or rc:
        lib.gemhip_clear_error()
        raise GemhipError(msg or f"error code {rc}")
    return rc


def output_filename(outdir: str, ymd, UTsec: float, identifier: int) -> str:
    """The file plasma_output_nompi_C writes: outdir/YYYYMMDD_SSSSS.ffffff_<id>.h5
    (src/temporal/timeutils.f90:133-196, src/io/io_nompi.f90:215-221)."""
    buf = C.create_string_buffer(4096)
    _host_call("gemhip_output_filename", outdir.encode(), (C.c_int * 3)(*[int(v) for v in ymd]),
               float(UTsec), int(identifier), buf, len(buf))
    return buf.value.decode()


def dateinc(dt: float, ymd, UTsec: float):
    """dateinc_C: (ymd, UTsec) advanced by dt (src/temporal/timeutils.f90:79-107)."""
    y = (C.c_int * 3)(*[int(v) for v in ymd])
    u = C.c_double(UTsec)
    lib = load_library()
    lib.gemhip_set_error_mode(1)
    lib.dateinc_C(C.byref(C.c_double(dt)), y, C.byref(u))
    msg = lib.gemhip_last_error().decode()
    if msg:
        lib.gemhip_clear_error()
        raise GemhipError(msg)
    return [y[0], y[1], y[2]], u.value


def output_filename_root(outdir: str, ymd, UTsec: float) -> str:
    """The file gemini.bin's output_plasma writes: outdir/YYYYMMDD_SSSSS.ffffff.h5
    (src/io/plasma_output_hdf5.f90:17)."""
    buf = C.create_string_buffer(4096)
    _host_call("gemhip_output_filename_root", outdir.encode(),
               (C.c_int * 3)(*[int(v) for v in ymd]), float(UTsec), buf, len(buf))
    return buf.value.decode()


def read_simsize(path: str):
    """(lx1, lx2, lx3) of a simsize file (src/io/reader_hdf5.f90:64-92)."""
    n = [C.c_int(0) for _ in range(3)]
    _host_call("gemhip_read_simsize", path.encode(), *[C.byref(v) for v in n])
    return tuple(v.value for v in n)


def read_grid_coords(path: str, sizes):
    """Ghosted coordinates x1, x2, x3 and (glonctr, glatctr) of a simgrid file or of the
    directory holding simgrid.h5 (src/numerical/grid/readgrid_hdf5.f90:11-39)."""
    lx1, lx2, lx3 = sizes
    xs = [np.empty(n + 4) for n in (lx1, lx2, lx3)]
    ctr = [C.c_double(0.0), C.c_double(0.0)]
    _host_call("gemhip_read_grid_coords", path.encode(), lx1, lx2, lx3,
               *[x.ctypes.data_as(_dp) for x in xs], *[C.cast(C.byref(c), _dp) for c in ctr])
    return xs[0], xs[1], xs[2], (ctr[0].value, ctr[1].value)


def grid_from_files(simsize: str, simgrid: str, kind: str, flagper: bool = False) -> Grid:
    """A Grid on the coordinates of the reference's input files.  The metric factors,
    altitude and gravity come from this package's mesh formulas (grid.cartesian_mesh,
    grid.dipole_mesh), not from the reference's mesh objects."""
    from .grid import cartesian_mesh, dipole_mesh
    x1, x2, x3, _ = read_grid_coords(simgrid, read_simsize(simsize))
    if kind == "cartesian":
        return cartesian_mesh(x1, x2, x3, flagper)
    if kind == "dipole":
        return dipole_mesh(x1, x2, x3)
    raise GemhipError(f"unknown mesh kind {kind!r}")


def read_initial_conditions(simsize: str, icfile: str, st: State):
    """Fill ns, vs1, Ts and the potential of st from the reference's input files
    (simsize.h5, initial_conditions.h5: src/io/io_nompi.f90:271-345)."""
    lx1, lx2, lx3 = (d - 4 for d in st.fluidvars.shape[:3])
    _host_call("gemhip_read_initial_conditions", simsize.encode(), icfile.encode(), lx1, lx2, lx3,
               _p(st.fluidvars), _p(st.electrovars))


def _p(a):
    """double* of a float64 numpy array in Fortran order, or of a contiguous float64
    torch tensor (host or device; the library copies with hipMemcpyDefault)."""
    if isinstance(a, np.ndarray):
        if a.dtype != np.float64 or not a.flags.f_contiguous:
            raise GemhipError("arrays must be float64 in Fortran order")
        return a.ctypes.data_as(_dp)
    if a.dtype.__repr__() != "torch.float64" or not a.is_contiguous():
        raise GemhipError("tensors must be contiguous float64")
    return C.cast(C.c_void_p(a.data_ptr()), _dp)


def _addr(a) -> int:
    """Address of a float64 numpy array (Fortran order) or contiguous float64 torch tensor."""
    return C.cast(_p(a), C.c_void_p).value


class _Ref:
    """A C pointer variable whose address is passed as double** (by reference)."""

    def __init__(self, a: np.ndarray | None):
        self.a = a
        self.v = C.c_void_p(a.ctypes.data if a is not None else None)

    @property
    def ref(self):
        return C.byref(self.v)


class Transport:
    """Drop-in hot path: holds the registered mesh, neutrals and config in the library."""

    XTYPE = {"cartesian": 1, "dipole": 2}

    GEOM = ("h1", "h2", "h3", "h1x1i", "h2x1i", "h3x1i", "h1x2i", "h2x2i", "h3x2i",
            "h1x3i", "h2x3i", "h3x3i")

    def __init__(self, grid: Grid, nn, Tn, device: int = 0, diffsolvetype: int = 2,
                 Teinf: float = 1500.0, geometry_on_device: bool = False,
                 own_context: bool = False):
        """nn (lx1,lx2,lx3,>=3) and Tn (lx1,lx2,lx3): Fortran numpy arrays, or torch
        tensors holding the same bytes.  With geometry_on_device the x3-invariant metric
        planes are expanded to 3-D arrays in HBM by torch instead of on the host.  With
        own_context the instance gets a library context of its own (ensemble members);
        otherwise it uses the implicit one the reference API acts on."""
        self.lib = load_library()
        self.g = grid
        self.handle = self.lib.gemhip_context_create() if own_context else 0
        self._sel()
        self.lib.gemhip_set_error_mode(1)
        self._call("gemhip_init", C.c_int(device))
        self._call("gemhip_set_grid_size", grid.lx1, grid.lx2, grid.lx3)
        if geometry_on_device:
            geo = [grid.device_array(n, f"cuda:{device}") for n in self.GEOM]
        else:
            geo = [np.asfortranarray(getattr(grid, n), dtype=np.float64) for n in self.GEOM]
        self._keep = geo + [np.ascontiguousarray(getattr(grid, n), dtype=np.float64) for n in
                            ("dx1", "dx1i", "dx2", "dx2i", "dx3", "dx3i")]
        isg = (C.c_int * 6)(*[int(bool(v)) for v in grid.isglobal])
        self.xtype = self.XTYPE[grid.kind]
        self._call("gemhip_register_grid", self.xtype, int(grid.gridflag), int(bool(grid.flagper)),
                   isg, *[_p(a) for a in self._keep])
        self.register_neutrals(nn, Tn)
        self.configure(diffsolvetype, Teinf)
        self._xt = C.c_int(self.xtype)
        self._null = C.c_void_p(None)
        self.residency = 0
        self._call("gemhip_set_residency", 0)
        self._outdir = None

    # ------------------------------------------------------------------ plumbing
    def _sel(self):
        if self.lib.gemhip_context_select(int(self.handle)) != 0:
            raise GemhipError(f"library context {self.handle} is gone")

    def close(self):
        """Release an own context (HBM buffers, stream)."""
        if self.handle:
            if self.lib.gemhip_context_destroy(int(self.handle)) != 0:
                raise GemhipError(self.lib.gemhip_last_error().decode())
            self.handle = None

    def set_tuning(self, knob: str, value: int):
        self._call("gemhip_set_tuning", knob.encode(), int(value))

    def get_tuning(self, knob: str) -> int:
        self._sel()
        return int(self.lib.gemhip_get_tuning(knob.encode()))

    def _check(self, rc=0):
        msg = self.lib.gemhip_last_error().decode()
        if msg or rc:
            self.lib.gemhip_clear_error()
            raise GemhipError(msg or f"error code {rc}")

    def _call(self, name, *args):
        self._sel()
        rc = getattr(self.lib, name)(*args)
        self._check(rc)
        return rc

    def register_neutrals(self, nn, Tn):
        if isinstance(nn, np.ndarray):
            nn = np.asfortranarray(nn[..., :3], dtype=np.float64)
            Tn = np.asfortranarray(Tn, dtype=np.float64)
        self._call("gemhip_register_neutrals", _p(nn), _p(Tn))
        self._keep_neut = (nn, Tn)

    def configure(self, diffsolvetype=2, Teinf=1500.0):
        self.diffsolvetype, self.Teinf = diffsolvetype, Teinf
        self._call("gemhip_register_config", int(diffsolvetype), float(Teinf))

    def set_residency(self, mode: int):
        self._call("gemhip_set_residency", int(mode))
        self.residency = mode

    def push(self, st: State):
        """Upload the state blocks and its interface velocities (the reference keeps the
        latter in intvars, which persists between steps)."""
        self._call("gemhip_state_push", _p(st.fluidvars), _p(st.fluidauxvars), _p(st.electrovars))
        self.set_interface_vels(st)
        self._pushed = st

    def push_blocks(self, fluidvars, fluidauxvars=None, electrovars=None):
        """Push raw blocks (numpy or torch) laid out like the reference's; None leaves that
        block's copy in HBM as it is."""
        opt = lambda a: None if a is None else _p(a)
        self._call("gemhip_state_push", _p(fluidvars), opt(fluidauxvars), opt(electrovars))

    def pull_blocks(self, fluidvars, fluidauxvars):
        """Copy the resident state into raw blocks (numpy or torch, host or HBM)."""
        self._call("gemhip_state_pull", _p(fluidvars), _p(fluidauxvars))

    def pull(self, st: State):
        self._call("gemhip_state_pull", _p(st.fluidvars), _p(st.fluidauxvars))

    def get_interface_vels(self, st: State):
        self._call("gemhip_get_interface_vels", _p(st.vs1i), _p(st.vs2i), _p(st.vs3i))

    def set_interface_vels_zero(self):
        """Interface velocities all zero, as the reference allocates them (libgemini.f90:288)."""
        n1 = (self.g.lx1 + 1) * self.g.lx2 * self.g.lx3 * 7
        n2 = self.g.lx1 * (self.g.lx2 + 1) * self.g.lx3 * 7
        n3 = self.g.lx1 * self.g.lx2 * (self.g.lx3 + 1) * 7
        z = [np.zeros(n, order="F") for n in (n1, n2, n3)]
        self._call("gemhip_set_interface_vels", *[_p(a) for a in z])

    def set_interface_vels(self, st: State):
        self._call("gemhip_set_interface_vels", _p(st.vs1i), _p(st.vs2i), _p(st.vs3i))

    def synchronize(self):
        self._call("gemhip_synchronize")

    @property
    def stream(self) -> int:
        self._sel()
        return self.lib.gemhip_stream()

    def device_state(self):
        f, a, j = C.POINTER(C.c_double)(), C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
        self._call("gemhip_device_state", C.byref(f), C.byref(a), C.byref(j))
        return C.cast(f, C.c_void_p).value, C.cast(a, C.c_void_p).value, C.cast(j, C.c_void_p).value

    def eval_math(self, op: int, x: np.ndarray) -> np.ndarray:
        """Device evaluation of a scalar routine (0 x**2.5, 1 sqrt, 2 1/x, 3 device pow;
        4 a/b through the reciprocal 1/b, for x of shape (n, 2) holding the pairs)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        n = x.shape[0] if op == 4 else x.size
        if op == 4 and x.shape != (n, 2):
            raise ValueError("op 4 takes an (n, 2) array of (numerator, divisor) pairs")
        y = np.empty(n)
        self._call("gemhip_eval_math", int(op), x.ctypes.data_as(_dp), y.ctypes.data_as(_dp),
                   n)
        return y

    def kernel_timing(self, enable: bool):
        self._call("gemhip_kernel_timing", int(enable))

    def kernel_time(self, name: str):
        ms, n = C.c_double(), C.c_long()
        self._sel()
        self.lib.gemhip_kernel_time(name.encode(), C.byref(ms), C.byref(n))
        return ms.value, n.value

    def kernel_timing_reset(self):
        self._call("gemhip_kernel_timing_reset")

    # ------------------------------------------------------------------ reference operations
    def _fa(self, st: State):
        return _Ref(st.fluidvars), _Ref(st.fluidauxvars)

    def _void(self, name, *args):
        self._sel()
        getattr(self.lib, name)(*args)
        self._check()

    def v12rhov1(self, st):
        f, a = self._fa(st)
        self._void("v12rhov1_C", f.ref, a.ref)

    def T2rhoe(self, st):
        f, a = self._fa(st)
        self._void("T2rhoe_C", f.ref, a.ref)

    def rhoe2T(self, st):
        f, a = self._fa(st)
        self._void("rhoe2T_C", f.ref, a.ref)

    def rhov12v1(self, st):
        f, a = self._fa(st)
        self._void("rhov12v1_C", f.ref, a.ref)

    def set_global_boundaries_allspec(self, st, lsp=7):
        f, a = self._fa(st)
        n = C.c_int(lsp)
        self._void("set_global_boundaries_allspec_C", C.byref(self._xt), C.byref(self._null),
                   f.ref, a.ref, C.byref(self._null), C.byref(n))

    def interface_vels_allspec(self, st, lsp=7):
        f = _Ref(st.fluidvars)
        n = C.c_int(lsp)
        self._void("interface_vels_allspec_C", C.byref(self._xt), C.byref(self._null), f.ref,
                   C.byref(self._null), C.byref(n))

    def halo_fluidvars(self, st):
        f, a = self._fa(st)
        self._void("halo_fluidvars_C", C.byref(self._xt), C.byref(self._null), f.ref, a.ref)

    def halo_allparams(self, st):
        f, a = self._fa(st)
        self._void("halo_allparams_C", C.byref(self._xt), C.byref(self._null), f.ref, a.ref)

    def halo_interface_vels_allspec(self, st, lsp=7):
        f = _Ref(st.fluidvars)
        n = C.c_int(lsp)
        self._void("halo_interface_vels_allspec_C", C.byref(self._xt), C.byref(self._null), f.ref,
                   C.byref(n))

    def sweep(self, st, direction: int, which: str, dt: float):
        """sweep{direction}_allspec_{which}_C; which in mass, momentum, energy, allparams."""
        f, a = self._fa(st)
        d = C.c_double(dt)
        name = f"sweep{direction}_allparams_C" if which == "allparams" else \
            f"sweep{direction}_allspec_{which}_C"
        self._void(name, f.ref, a.ref, C.byref(self._null), C.byref(self._xt),
                   C.byref(self._null), C.byref(d))

    def energy_diffusion(self, st, dt: float):
        f, e = _Ref(st.fluidvars), _Ref(st.electrovars)
        d = C.c_double(dt)
        self._void("energy_diffusion_C", C.byref(self._null), C.byref(self._xt),
                   C.byref(self._null), f.ref, e.ref, C.byref(self._null), C.byref(d))

    def clean_param(self, st, iparm: int):
        """clean_param_C: 1 densities, 2 vs1, 3 temperatures."""
        f = _Ref(st.fluidvars)
        i = C.c_int(iparm)
        self._void("clean_param_C", C.byref(i), C.byref(self._xt), C.byref(self._null), f.ref)

    def VNRicht_artvisc(self, st):
        f = _Ref(st.fluidvars)
        self._void("VNRicht_artvisc_C", f.ref, C.byref(self._null))

    def compression(self, st, dt: float):
        f, a = self._fa(st)
        d = C.c_double(dt)
        self._void("compression_C", f.ref, a.ref, C.byref(self._null), C.byref(self._xt),
                   C.byref(self._null), C.byref(d))

    def RK2_prep_mpi_allspec(self, st):
        f = _Ref(st.fluidvars)
        self._void("RK2_prep_mpi_allspec_C", C.byref(self._xt), C.byref(self._null), f.ref)

    def RK2_global_boundary_allspec(self, st):
        f = _Ref(st.fluidvars)
        self._void("RK2_global_boundary_allspec_C", C.byref(self._xt), C.byref(self._null), f.ref)

    def register_null_cells(self, nullpts):
        """Null-cell mask (lx1, lx2, lx3) (nonzero: null), or None for none."""
        if nullpts is None:
            self._call("gemhip_register_null_cells", None)
            return
        m = np.asfortranarray(nullpts, dtype=np.int32)
        self._call("gemhip_register_null_cells", m.ctypes.data_as(C.POINTER(C.c_int)))

    def register_cfl(self, tcfl: float, flagglow: int = 0):
        self._call("gemhip_register_cfl", float(tcfl), int(flagglow))

    def grid_rates(self) -> np.ndarray:
        """dt_calc's 21 grid rates of the resident state, shape (7, 3)."""
        r = np.zeros(21)
        self._call("gemhip_grid_rates", r.ctypes.data_as(_dp))
        return r.reshape(7, 3)

    def dt_select(self, st, t: float, tout: float, tglowout: float, dt: float) -> float:
        """dt_select_C: the next time step from dt (the previous one)."""
        f, a = self._fa(st)
        args = [C.c_double(v) for v in (t, tout, tglowout, dt)]
        self._void("dt_select_C", C.byref(self._null), C.byref(self._xt), C.byref(self._null),
                   f.ref, a.ref, *[C.byref(v) for v in args])
        return args[3].value

    # ------------------------------------------------------------ source/loss substep
    def register_atmosphere(self, atm):
        """intvars%atmos of an :class:`gemini_amd.state.Atmosphere` (numpy) or a mapping of
        torch tensors with the same bytes (nn, Tn, vn1, vn2, vn3); also the neutrals of the
        energy diffusion."""
        get = (lambda n: atm[n]) if isinstance(atm, dict) else (lambda n: getattr(atm, n))
        arrs = [get(n) for n in ("nn", "Tn", "vn1", "vn2", "vn3")]
        arrs = [np.asfortranarray(a, dtype=np.float64) if isinstance(a, np.ndarray) else a
                for a in arrs]
        self._call("gemhip_register_atmosphere", *[_p(a) for a in arrs])
        self._keep_neut = tuple(arrs)

    def register_gravity(self):
        """The grid's altitude (ghosted) and gravity along x1 (interior)."""
        alt = np.asfortranarray(self.g.alt, dtype=np.float64)
        g1 = np.asfortranarray(self.g.g1, dtype=np.float64)
        self._call("gemhip_register_gravity", _p(alt), _p(g1))

    def set_ionization_rates(self, Prionize, Qeionize):
        """intvars%Prionize (lx1,lx2,lx3,6) and Qeionize (lx1,lx2,lx3), numpy or torch."""
        conv = lambda a: np.asfortranarray(a, dtype=np.float64) if isinstance(a, np.ndarray) else a
        P, Q = conv(Prionize), conv(Qeionize)
        self._call("gemhip_set_ionization_rates", _p(P), _p(Q))

    def enable_sources(self, atm):
        """Register everything the source/loss substep reads: atmosphere, gravity and the
        ionization rates of ``atm``."""
        self.register_atmosphere(atm)
        self.register_gravity()
        get = (lambda n: atm[n]) if isinstance(atm, dict) else (lambda n: getattr(atm, n))
        self.set_ionization_rates(get("Prionize"), get("Qeionize"))

    def clear_ionization_arrays(self):
        self._void("clear_ionization_arrays_C", C.byref(self._null))

    def source_loss(self, st, which: str, dt: float):
        """source_loss_{energy,momentum,mass,allparams}_C."""
        f, a, e = _Ref(st.fluidvars), _Ref(st.fluidauxvars), _Ref(st.electrovars)
        d = C.c_double(dt)
        self._void(f"source_loss_{which}_C", f.ref, a.ref, e.ref, C.byref(self._null),
                   C.byref(self._xt), C.byref(self._null), C.byref(d))

    def sources_block(self, dt: float):
        """The sources substep of src/libgemini.cpp:223-232 on the HBM-resident state."""
        self._call("gemhip_sources_block", C.c_double(dt))

    def sources_block_calls(self, st, dt: float):
        """The reference's calls of src/libgemini.cpp:223-232 one by one."""
        self.set_global_boundaries_allspec(st, 7)
        for w in ("energy", "momentum", "mass"):
            self.source_loss(st, w, dt)
        for iparm in (3, 2, 1):
            self.clean_param(st, iparm)

    def fluid_adv_block(self, dt: float):
        """The whole transport block of src/libgemini.cpp:155-210 on the HBM-resident state
        (residency mode 1)."""
        self._call("gemhip_fluid_adv_block", C.c_double(dt))

    def fluid_adv_block_calls(self, st, dt: float):
        """The reference's calls of src/libgemini.cpp:155-210 one by one."""
        self.v12rhov1(st)
        self.T2rhoe(st)
        self.set_global_boundaries_allspec(st, 7)
        self.halo_fluidvars(st)
        self.interface_vels_allspec(st, 7)
        for d in (3, 1):
            for w in ("mass", "momentum", "energy"):
                self.sweep(st, d, w, dt)
        self.halo_allparams(st)
        for w in ("mass", "momentum", "energy"):
            self.sweep(st, 2, w, dt)
        self.rhov12v1(st)
        self.clean_param(st, 1)
        self.clean_param(st, 2)
        self.VNRicht_artvisc(st)
        self.halo_fluidvars(st)
        self.RK2_global_boundary_allspec(st)
        self.compression(st, dt)
        self.rhoe2T(st)
        self.clean_param(st, 3)
        self.energy_diffusion(st, dt)
        self.clean_param(st, 3)
        self.T2rhoe(st)

    def fluid_adv_hotpath(self, st, dt: float):
        """The hot-path calls of src/libgemini.cpp:155-208 one by one, then rhoe2T before
        the energy diffusion, as the oracle's transport step orders them."""
        self.v12rhov1(st)
        self.T2rhoe(st)
        self.set_global_boundaries_allspec(st, 7)
        self.halo_fluidvars(st)
        self.interface_vels_allspec(st, 7)
        for d in (3, 1):
            for w in ("mass", "momentum", "energy"):
                self.sweep(st, d, w, dt)
        self.halo_allparams(st)
        for w in ("mass", "momentum", "energy"):
            self.sweep(st, 2, w, dt)
        self.rhov12v1(st)
        self.rhoe2T(st)
        self.energy_diffusion(st, dt)

    def transport_step(self, dt: float):
        """Fused step on the HBM-resident state (residency mode 1)."""
        self._call("gemhip_transport_step", C.c_double(dt))

    # ------------------------------------------------------------------ fitting extension
    def _column_param(self, a):
        if a is None:
            return None
        if isinstance(a, np.ndarray):
            a = np.asfortranarray(np.broadcast_to(a, (self.g.lx2, self.g.lx3)), dtype=np.float64)
            return a.reshape(-1, order="F")
        return a.contiguous()

    def set_conductivity_scale(self, kappa):
        """Per-column scale (lx2, lx3) of the electron thermal conductivity; None removes it.
        A torch tensor is read as lx2*lx3 values with x2 fastest."""
        k = self._column_param(kappa)
        self._call("gemhip_register_conductivity_scale", None if k is None else _p(k))

    def set_top_heat_flux(self, qtop):
        """Per-column downward electron heat flux [W/m^2] through the top face (gridflag 2),
        replacing the Dirichlet top row; None restores it."""
        q = self._column_param(qtop)
        self._call("gemhip_register_top_heat_flux", None if q is None else _p(q))

    def register_output(self, outdir: str, flagoutput: int = 1, comp_lvl: int = 3):
        """cfg%outdir and cfg%flagoutput for plasma_output; comp_lvl is the deflate level of
        the field datasets (the reference's phys_consts comp_lvl is 3; 0 stores them raw)."""
        self._call("gemhip_register_output", outdir.encode(), int(flagoutput), int(comp_lvl))
        self._outdir = outdir

    def plasma_output(self, st, ymd, UTsec: float, identifier: int = 0, x1lims=(0.0, 0.0),
                      x2lims=(0.0, 0.0), x3lims=(0.0, 0.0), electrovars=None) -> str:
        """plasma_output_nompi_C: write the state to outdir/<date stem>_<identifier>.h5 and
        return the path.  With the state resident in HBM the fluid fields come from there
        (st may then be None); J1-J3 and the potential come from st.electrovars, or from
        electrovars (a host array or device tensor holding the electrovars block)."""
        ev = electrovars if electrovars is not None else st.electrovars
        f = C.c_void_p(None if st is None else _addr(st.fluidvars))
        e = C.c_void_p(_addr(ev))
        y = (C.c_int * 3)(*[int(v) for v in ymd])
        lims = [(C.c_double * 2)(*[float(v) for v in l]) for l in (x1lims, x2lims, x3lims)]
        self._void("plasma_output_nompi_C", C.byref(self._null), y, C.byref(C.c_double(UTsec)),
                   C.byref(f), C.byref(e), C.byref(C.c_int(identifier)), *lims)
        if self._outdir is None:
            raise GemhipError("the output directory was registered outside this Transport; "
                              "call register_output")
        return output_filename(self._outdir, ymd, UTsec, identifier)

    def output_plasma(self, st, ymd, UTsec: float, electrovars=None) -> str:
        """gemini.bin's output file (src/io/plasma_output.f90:76-136 and
        plasma_output_hdf5.f90:10-65): outdir/<date stem>.h5, the datasets of plasma_output
        with TEall for flagoutput 2 and no patch datasets.  Arguments as plasma_output's."""
        ev = electrovars if electrovars is not None else st.electrovars
        f = C.c_void_p(None if st is None else _addr(st.fluidvars))
        e = C.c_void_p(_addr(ev))
        y = (C.c_int * 3)(*[int(v) for v in ymd])
        self._call("gemhip_output_plasma", y, C.c_double(UTsec), C.byref(f), C.byref(e))
        if self._outdir is None:
            raise GemhipError("the output directory was registered outside this Transport; "
                              "call register_output")
        return output_filename_root(self._outdir, ymd, UTsec)

    def checkpoint_size(self) -> int:
        """doubles in a restart checkpoint (fluidvars and the x1 boundary velocities)."""
        self._sel()
        return int(self.lib.gemhip_checkpoint_size())

    def checkpoint_save(self, dst):
        self._call("gemhip_checkpoint_save", _p(dst))

    def checkpoint_load(self, src):
        self._call("gemhip_checkpoint_load", _p(src))

    def adjoint_tape_size(self) -> int:
        """doubles one recorded step keeps for the adjoint."""
        self._sel()
        return int(self.lib.gemhip_adjoint_tape_size())

    def record_step(self, dt: float, tape_ptr: int):
        """transport_step that also writes the adjoint tape at device address tape_ptr."""
        self._call("gemhip_adjoint_record_step", C.c_double(dt), C.c_void_p(tape_ptr))

    def adjoint_step(self, dt: float, tape_ptr: int, lam_ptr: int, gk_ptr: int = 0,
                     gq_ptr: int = 0):
        """dL/dTe after a recorded step -> before it (in place at lam_ptr); accumulates
        dL/dkappa and dL/dq per column at gk_ptr, gq_ptr (0: not wanted)."""
        self._call("gemhip_adjoint_step", C.c_double(dt), C.c_void_p(tape_ptr),
                   C.c_void_p(lam_ptr), C.c_void_p(gk_ptr or None), C.c_void_p(gq_ptr or None))
