"""ctypes front end of the oracle -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/liboracle.so`` (built from gemini_oracle.c by ``oracle/Makefile``)
and exposes the CPU restatement of the reference's transport path on the
numpy arrays of :mod:`gemini_amd.grid` / :mod:`gemini_amd.state`.  Only the
tests, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use it, and
only as the checker; the product library never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)


class _OrcGrid(C.Structure):
    _fields_ = [("lx1", C.c_int), ("lx2", C.c_int), ("lx3", C.c_int),
                ("gridflag", C.c_int), ("flagper", C.c_int), ("isglobal", C.c_int * 6)] + \
               [(n, _dp) for n in ("h1", "h2", "h3", "h1x1i", "h2x1i", "h3x1i",
                                   "h1x2i", "h2x2i", "h3x2i", "h1x3i", "h2x3i", "h3x3i",
                                   "dx1", "dx1i", "dx2", "dx2i", "dx3", "dx3i")]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []), check=True)
    return LIB_PATH


def _load(path):
    lib_ = C.CDLL(path)
    lib_.orc_minmod.restype = C.c_double
    lib_.orc_minmod.argtypes = [C.c_double, C.c_double]
    for name in ("orc_gbsv", "orc_energy_diffusion", "orc_transport_step", "orc_num_threads",
                 "orc_energy_diffusion_ext", "orc_transport_step_ext", "orc_clean_param",
                 "orc_fluid_adv_block"):
        getattr(lib_, name).restype = C.c_int
    return lib_


def lib():
    """The oracle library: liboracle.so, or the build that $GEMINI_ORACLE_LIB names (the
    mutation tests run the pins against copies with one planted mistake each)."""
    global _lib
    if _lib is None:
        path = os.environ.get("GEMINI_ORACLE_LIB")
        if not path:
            path = LIB_PATH
            if not os.path.exists(path):
                build()
        _lib = _load(path)
    return _lib


def ptr(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous), a.flags
    return a.ctypes.data_as(_dp)


class OracleGrid:
    """Keeps the numpy arrays alive while the C struct points into them."""

    def __init__(self, g):
        self.g = g
        s = _OrcGrid()
        s.lx1, s.lx2, s.lx3 = g.lx1, g.lx2, g.lx3
        s.gridflag = int(g.gridflag)
        s.flagper = int(bool(g.flagper))
        for k in range(6):
            s.isglobal[k] = int(bool(g.isglobal[k]))
        self._keep = []
        for n in ("h1", "h2", "h3", "h1x1i", "h2x1i", "h3x1i", "h1x2i", "h2x2i", "h3x2i",
                  "h1x3i", "h2x3i", "h3x3i", "dx1", "dx1i", "dx2", "dx2i", "dx3", "dx3i"):
            a = np.asfortranarray(getattr(g, n), dtype=np.float64)
            self._keep.append(a)
            setattr(s, n, ptr(a))
        self.s = s

    @property
    def ref(self):
        return C.byref(self.s)


def _f(a):
    assert a.dtype == np.float64 and a.flags.f_contiguous
    return ptr(a)


def _opt(a):
    """double* of an optional per-column parameter array (lx2, lx3), None -> NULL."""
    if a is None:
        return None, None
    a = np.asfortranarray(a, dtype=np.float64)
    return a, ptr(a)


def transport_step(g, st, dt, diffsolvetype=2, Teinf=1500.0, kappa=None, qtop=None):
    """Advance ``st`` (in place) by one oracle transport step; returns st.

    kappa (lx2, lx3) scales the electron conductivity per column and qtop (lx2, lx3) is
    a downward electron heat flux through the top face [W/m^2] (the fitting extension);
    None keeps the reference behaviour."""
    og = OracleGrid(g)
    ka, kp = _opt(kappa)
    qa, qp = _opt(qtop)
    rc = lib().orc_transport_step_ext(og.ref, C.c_double(dt), _f(st.fluidvars),
                                      _f(st.fluidauxvars), _f(st.electrovars), _f(st.vs1i),
                                      _f(st.vs2i), _f(st.vs3i), _f(st.nn), _f(st.Tn),
                                      C.c_int(diffsolvetype), C.c_double(Teinf), kp, qp)
    if rc != 0:
        raise RuntimeError(f"oracle transport step failed ({rc})")
    return st


def set_pow_mode(mode: int):
    """0: x**2.5 via libm pow as the reference; 1: correctly rounded x**2.5."""
    lib().orc_set_pow_mode(C.c_int(mode))


def eval_pow25(x, mode):
    """x**2.5 elementwise: mode 0 libm pow, mode 1 correctly rounded."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib().orc_eval_pow25(ptr(x), ptr(y), C.c_long(x.size), C.c_int(mode))
    return y


class pow_mode:
    """Context manager selecting the oracle's x**2.5 evaluation."""

    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        set_pow_mode(self.mode)

    def __exit__(self, *exc):
        set_pow_mode(0)


def reset_swap_count():
    lib().orc_reset_swap_count()


def swap_count() -> int:
    """Row interchanges made by the band solves since reset_swap_count()."""
    f = lib().orc_swap_count
    f.restype = C.c_long
    return int(f())


def set_num_threads(n: int):
    os.environ["OMP_NUM_THREADS"] = str(n)
    try:
        C.CDLL("libgomp.so.1").omp_set_num_threads(C.c_int(n))
    except OSError:
        pass


# ---- finer-grained entry points used by the unit tests ----

def constants():
    ms, qs, gm, tc = (np.zeros(7), np.zeros(7), np.zeros(7), np.zeros(6))
    kB, e = C.c_double(), C.c_double()
    lib().orc_constants(ptr(ms), ptr(qs), ptr(gm), C.byref(kB), C.byref(e), ptr(tc))
    return dict(ms=ms, qs=qs, gammas=gm, kB=kB.value, elchrg=e.value, thermal_coeff=tc)


def minmod(a, b):
    return lib().orc_minmod(C.c_double(a), C.c_double(b))


def advec1D(f, v, dt, dx, dxi, ha1, ha2i, h1i):
    lx = len(f) - 4
    out = np.zeros(lx + 4)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (f, v, dx, dxi, ha1, ha2i, h1i)]
    lib().orc_advec1D_MC_curv(C.c_int(lx), ptr(arrs[0]), ptr(arrs[1]), C.c_double(dt),
                              ptr(arrs[2]), ptr(arrs[3]), ptr(arrs[4]), ptr(arrs[5]),
                              ptr(arrs[6]), ptr(out))
    return out


def gbsv(ab, b, kl=2):
    """LAPACK-band solve; ab (3*kl+1, n) Fortran band array, returns (x, lu, ipiv, info)."""
    ab = np.array(ab, dtype=np.float64, order="F", copy=True)
    x = np.array(b, dtype=np.float64, copy=True)
    n = ab.shape[1]
    ipiv = np.zeros(n, dtype=np.int32)
    info = lib().orc_gbsv(C.c_int(n), C.c_int(kl), ptr(ab), ptr(x),
                          ipiv.ctypes.data_as(C.POINTER(C.c_int)))
    return x, ab, ipiv, info


def _col(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def trbdf2_1d(Ts, A, B, Cc, D, E, Tmin, Tmax, dt, dx1, dx1i):
    lx1 = len(Ts)
    out = np.zeros(lx1)
    a = [_col(v) for v in (Ts, A, B, Cc, D, E, dx1, dx1i)]
    lib().orc_TRBDF21D(C.c_int(lx1), ptr(a[0]), ptr(a[1]), ptr(a[2]), ptr(a[3]), ptr(a[4]),
                       ptr(a[5]), C.c_double(Tmin), C.c_double(Tmax), C.c_double(dt),
                       ptr(a[6]), ptr(a[7]), ptr(out))
    return out


def trbdf2_1d_flux(Ts, A, B, Cc, D, E, Tmin, gtop, dt, dx1, dx1i):
    """TRBDF2 with the Neumann top row (T(lx1)-T(lx1-1))/dx1(lx1) = gtop."""
    lx1 = len(Ts)
    out = np.zeros(lx1)
    a = [_col(v) for v in (Ts, A, B, Cc, D, E, dx1, dx1i)]
    lib().orc_TRBDF21D_flux(C.c_int(lx1), ptr(a[0]), ptr(a[1]), ptr(a[2]), ptr(a[3]), ptr(a[4]),
                            ptr(a[5]), C.c_double(Tmin), C.c_double(gtop), C.c_double(dt),
                            ptr(a[6]), ptr(a[7]), ptr(out))
    return out


def backeuler_1d(Ts, A, B, Cc, D, E, Tmin, Tmax, dt, dx1, dx1i):
    lx1 = len(Ts)
    out = np.zeros(lx1)
    coeffs = np.zeros((lx1, 3), order="F")
    rhs = np.zeros(lx1)
    a = [_col(v) for v in (Ts, A, B, Cc, D, E, dx1, dx1i)]
    lib().orc_backEuler1D(C.c_int(lx1), ptr(a[0]), ptr(a[1]), ptr(a[2]), ptr(a[3]), ptr(a[4]),
                          ptr(a[5]), C.c_double(Tmin), C.c_double(Tmax), C.c_double(dt),
                          ptr(a[6]), ptr(a[7]), ptr(out), ptr(coeffs), ptr(rhs))
    return out, coeffs, rhs


def interface_vels(g, st, lsp=7):
    og = OracleGrid(g)
    lib().orc_interface_vels_allspec(og.ref, _f(st.fluidvars[..., 7:14]),
                                     _f(st.fluidvars[..., 14:21]), _f(st.fluidvars[..., 21:28]),
                                     _f(st.vs1i), _f(st.vs2i), _f(st.vs3i), C.c_int(lsp))
    return st


def sweep(g, direction, fs, vsi, dt, frank, lsp):
    """fs: (L1,L2,L3,7) Fortran array, modified in place."""
    og = OracleGrid(g)
    lib().orc_sweep_allspec(og.ref, C.c_int(direction), _f(fs), _f(vsi), C.c_double(dt),
                            C.c_int(frank), C.c_int(lsp))
    return fs


def _views(st):
    fv, fa = st.fluidvars, st.fluidauxvars
    return dict(ns=fv[..., 0:7], vs1=fv[..., 7:14], vs2=fv[..., 14:21], vs3=fv[..., 21:28],
                Ts=fv[..., 28:35], rhovs1=fa[..., 0:7], rhoes=fa[..., 7:14],
                J1=st.electrovars[..., 3])


def _mask(nullpts):
    """int* of a null-cell mask (lx1, lx2, lx3), None -> NULL."""
    if nullpts is None:
        return None
    m = np.asfortranarray(nullpts, dtype=np.int32)
    _mask.keep = m
    return m.ctypes.data_as(C.POINTER(C.c_int))


def fluid_adv_block(g, st, dt, nullpts=None, Q=None, diffsolvetype=2, Teinf=1500.0):
    """The transport block of src/libgemini.cpp:155-210 on ``st`` (in place); Q is the
    (lx1, lx2, lx3, 7) artificial-viscosity workspace (allocated when None)."""
    og = OracleGrid(g)
    if Q is None:
        Q = np.zeros((g.lx1, g.lx2, g.lx3, 7), order="F")
    rc = lib().orc_fluid_adv_block(og.ref, C.c_double(dt), _f(st.fluidvars), _f(st.fluidauxvars),
                                   _f(st.electrovars), _f(st.vs1i), _f(st.vs2i), _f(st.vs3i),
                                   _f(st.nn), _f(st.Tn), C.c_int(diffsolvetype),
                                   C.c_double(Teinf), _mask(nullpts), _f(Q))
    if rc != 0:
        raise RuntimeError(f"oracle transport block failed ({rc})")
    # libgemini.cpp:210 converts back to energy density for the sources step
    lib().orc_T2rhoe(og.ref, _f(st.fluidvars[..., 0:7]), _f(st.fluidvars[..., 28:35]),
                     _f(st.fluidauxvars[..., 7:14]))
    return st


def dt_calc(g, st, tcfl):
    """temporal_mpi.f90 dt_calc on one process: (dt, rates (7, 3))."""
    og = OracleGrid(g)
    r = np.zeros(21)
    f = lib().orc_dt_calc
    f.restype = C.c_double
    v = _views(st)
    dt = f(og.ref, _f(v["Ts"]), _f(v["vs1"]), _f(v["vs2"]), _f(v["vs3"]), C.c_double(tcfl), ptr(r))
    return dt, r.reshape(7, 3)


def dt_select(g, st, tcfl, flagglow, t, tout, tglowout, it, dt):
    """dt_select with dt_comm on one process; it is the iteration count (get_it)."""
    og = OracleGrid(g)
    v = _views(st)
    d = C.c_double(dt)
    lib().orc_dt_select(og.ref, _f(v["Ts"]), _f(v["vs1"]), _f(v["vs2"]), _f(v["vs3"]),
                        C.c_double(tcfl), C.c_int(flagglow), C.c_double(t), C.c_double(tout),
                        C.c_double(tglowout), C.c_int(it), C.byref(d))
    return d.value


def op(g, st, name, dt=None, lsp=7, direction=None, which=None, diffsolvetype=2,
       Teinf=1500.0, kappa=None, qtop=None, nullpts=None, Q=None):
    """Apply one reference operation of the transport path to ``st`` in place.

    Names follow the reference's C entry points: v12rhov1, T2rhoe, rhoe2T, rhov12v1,
    set_global_boundaries_allspec, halo_fluidvars, halo_allparams,
    halo_interface_vels_allspec, interface_vels_allspec, sweep (direction 1-3, which
    mass/momentum/energy/allparams) and energy_diffusion.
    """
    og = OracleGrid(g)
    L = lib()
    v = {k: _f(a) for k, a in _views(st).items()}
    if name == "v12rhov1":
        L.orc_v12rhov1(og.ref, v["ns"], v["vs1"], v["rhovs1"])
    elif name == "T2rhoe":
        L.orc_T2rhoe(og.ref, v["ns"], v["Ts"], v["rhoes"])
    elif name == "rhoe2T":
        L.orc_rhoe2T(og.ref, v["ns"], v["rhoes"], v["Ts"])
    elif name == "rhov12v1":
        L.orc_rhov12v1(og.ref, v["ns"], v["rhovs1"], v["vs1"])
    elif name == "set_global_boundaries_allspec":
        L.orc_set_global_boundaries_allspec(og.ref, v["ns"], v["rhovs1"], v["vs1"], v["vs2"],
                                            v["vs3"], v["rhoes"], _f(st.vs1i), C.c_int(lsp))
    elif name in ("halo_fluidvars", "halo_allparams"):
        names = ["ns", "rhovs1", "rhoes"] + (["vs2", "vs3"] if name == "halo_fluidvars" else [])
        for k in names:
            L.orc_halo_periodic(og.ref, v[k], C.c_int(7), C.c_int(2))
    elif name == "halo_interface_vels_allspec":
        for k in ("vs2", "vs3"):
            L.orc_halo_periodic(og.ref, v[k], C.c_int(lsp), C.c_int(1))
    elif name == "interface_vels_allspec":
        L.orc_interface_vels_allspec(og.ref, v["vs1"], v["vs2"], v["vs3"], _f(st.vs1i),
                                     _f(st.vs2i), _f(st.vs3i), C.c_int(lsp))
    elif name == "sweep":
        vsi = {1: st.vs1i, 2: st.vs2i, 3: st.vs3i}[direction]
        parts = {"mass": [("ns", 0, 6)], "momentum": [("rhovs1", 1, 6)],
                 "energy": [("rhoes", 0, 7)]}
        todo = parts["mass"] + parts["momentum"] + parts["energy"] if which == "allparams" \
            else parts[which]
        for field, frank, nsp in todo:
            L.orc_sweep_allspec(og.ref, C.c_int(direction), v[field], _f(vsi), C.c_double(dt),
                                C.c_int(frank), C.c_int(nsp))
    elif name == "clean_param":
        field = {1: "ns", 2: "vs1", 3: "Ts"}.get(which, "ns")
        rc = L.orc_clean_param(og.ref, C.c_int(which), v[field], _mask(nullpts))
        if rc != 0:
            raise RuntimeError(f"oracle clean_param failed ({rc})")
    elif name == "VNRicht_artvisc":
        L.orc_VNRicht_artvisc(og.ref, v["ns"], v["vs1"], _f(Q))
    elif name == "RK2_prep_mpi_allspec":
        for k in ("vs1", "vs2", "vs3"):
            L.orc_halo_periodic(og.ref, v[k], C.c_int(7), C.c_int(1))
    elif name == "RK2_global_boundary_allspec":
        L.orc_RK2_global_boundary_allspec(og.ref, v["vs1"], v["vs2"], v["vs3"])
    elif name == "compression":
        L.orc_compression(og.ref, C.c_double(dt), v["vs1"], v["vs2"], v["vs3"], _f(Q), v["rhoes"])
    elif name == "energy_diffusion":
        ka, kp = _opt(kappa)
        qa, qp = _opt(qtop)
        rc = L.orc_energy_diffusion_ext(og.ref, C.c_double(dt), v["ns"], v["Ts"], v["J1"],
                                        _f(st.nn), _f(st.Tn), C.c_int(diffsolvetype),
                                        C.c_double(Teinf), kp, qp)
        if rc != 0:
            raise RuntimeError(f"oracle energy diffusion failed ({rc})")
    else:
        raise KeyError(name)
    return st


# ---------------------------------------------------------------- source/loss step
class _OrcAtmos(C.Structure):
    _fields_ = [(n, _dp) for n in ("nn", "Tn", "vn1", "vn2", "vn3", "Prionize", "Qeionize",
                                   "alt", "g1")]


class OracleAtmos:
    """orc_atmos over an :class:`gemini_amd.state.Atmosphere` and the grid's alt and g1."""

    def __init__(self, g, atm):
        s = _OrcAtmos()
        self._keep = []
        arrays = dict(nn=atm.nn, Tn=atm.Tn, vn1=atm.vn1, vn2=atm.vn2, vn3=atm.vn3,
                      Prionize=atm.Prionize, Qeionize=atm.Qeionize, alt=g.alt, g1=g.g1)
        for n, a in arrays.items():
            a = np.asfortranarray(a, dtype=np.float64)
            self._keep.append(a)
            setattr(s, n, ptr(a))
        self.s = s

    @property
    def ref(self):
        return C.byref(self.s)


def etd(f, P, L, dt):
    fn = lib().orc_ETD_uncoupled
    fn.restype = C.c_double
    return fn(C.c_double(f), C.c_double(P), C.c_double(L), C.c_double(dt))


def maxwell_colln(isp, isp2, nn2, Tn, Tsp):
    fn = lib().orc_maxwell_colln
    fn.restype = C.c_double
    return fn(C.c_int(isp), C.c_int(isp2), C.c_double(nn2), C.c_double(Tn), C.c_double(Tsp))


def coulomb_colln(isp, isp2, ns2, Ts1, Ts2, v1, v2):
    """(nusj, Phisj, Psisj) at one cell."""
    out = np.zeros(3)
    lib().orc_coulomb_colln(C.c_int(isp), C.c_int(isp2), C.c_double(ns2), C.c_double(Ts1),
                            C.c_double(Ts2), C.c_double(v1), C.c_double(v2), ptr(out))
    return tuple(out)


def grad3D1(g, f):
    """grad3D1_curv_23 of an interior (lx1, lx2, lx3) array over bounds 1..lx1."""
    og = OracleGrid(g)
    f = np.asfortranarray(f, dtype=np.float64)
    out = np.zeros_like(f, order="F")
    lib().orc_grad3D1_interior(og.ref, ptr(f), ptr(out))
    return out


def srcs(g, st, atm, which, Q=None):
    """Pr, Lo (lx1, lx2, lx3, 7) of srcsEnergy / srcsMomentum / srcsContinuity."""
    og, oa = OracleGrid(g), OracleAtmos(g, atm)
    v = {k: _f(a) for k, a in _views(st).items()}
    Pr = np.zeros((g.lx1, g.lx2, g.lx3, 7), order="F")
    Lo = np.zeros_like(Pr, order="F")
    L = lib()
    if which == "energy":
        L.orc_srcs_energy(og.ref, oa.ref, v["ns"], v["vs1"], v["vs2"], v["vs3"], v["Ts"], ptr(Pr), ptr(Lo))
    elif which == "momentum":
        Q = np.zeros((g.lx1, g.lx2, g.lx3, 7), order="F") if Q is None else Q
        L.orc_srcs_momentum(og.ref, oa.ref, v["ns"], v["vs1"], v["vs2"], v["vs3"], v["Ts"],
                            _f(st.electrovars[..., 0]), _f(Q), ptr(Pr), ptr(Lo))
    elif which == "mass":
        L.orc_srcs_continuity(og.ref, oa.ref, v["ns"], v["vs1"], v["vs2"], v["vs3"], v["Ts"],
                              ptr(Pr), ptr(Lo))
    else:
        raise KeyError(which)
    return Pr, Lo


def source_loss(g, st, atm, dt, which, Q=None):
    """source_loss_{energy,momentum,mass} (multifluid.f90:130-175) on ``st`` in place; Q is
    the (lx1, lx2, lx3, 7) artificial viscosity the momentum sources read."""
    og, oa = OracleGrid(g), OracleAtmos(g, atm)
    v = {k: _f(a) for k, a in _views(st).items()}
    L = lib()
    if which == "energy":
        L.orc_source_loss_energy(og.ref, oa.ref, C.c_double(dt), v["ns"], v["vs1"], v["vs2"],
                                 v["vs3"], v["Ts"], v["rhoes"])
    elif which == "momentum":
        Q = np.zeros((g.lx1, g.lx2, g.lx3, 7), order="F") if Q is None else Q
        L.orc_source_loss_momentum(og.ref, oa.ref, C.c_double(dt), v["ns"], v["vs1"], v["vs2"],
                                   v["vs3"], v["Ts"], _f(st.electrovars[..., 0]), _f(Q), v["rhovs1"])
    elif which == "mass":
        L.orc_source_loss_mass(og.ref, oa.ref, C.c_double(dt), v["ns"], v["vs1"], v["vs2"],
                               v["vs3"], v["Ts"])
    else:
        raise KeyError(which)
    return st


def sources_block(g, st, atm, dt, Q=None, nullpts=None):
    """The sources substep of src/libgemini.cpp:223-232 on ``st`` in place."""
    og, oa = OracleGrid(g), OracleAtmos(g, atm)
    Q = np.zeros((g.lx1, g.lx2, g.lx3, 7), order="F") if Q is None else Q
    fn = lib().orc_sources_block
    fn.restype = C.c_int
    rc = fn(og.ref, oa.ref, C.c_double(dt), _f(st.fluidvars), _f(st.fluidauxvars),
            _f(st.electrovars), _f(st.vs1i), _f(Q), _mask(nullpts))
    if rc != 0:
        raise RuntimeError(f"oracle sources block failed ({rc})")
    return st


# ---- plasma output (src/io/io_nompi.f90:183-268, src/temporal/timeutils.f90:151-196) ----
def filestem(ymd, UTsec):
    """The 21-character date stem of an output file, or None where the reference stops."""
    buf = C.create_string_buffer(32)
    rc = lib().orc_filestem((C.c_int * 3)(*[int(v) for v in ymd]), C.c_double(UTsec), buf)
    return None if rc else buf.value.decode()


def dateinc(dtsec, ymd, UTsec):
    """(ymd, UTsec) advanced by dtsec (timeutils.f90:79-107), or None where the reference
    stops."""
    y = (C.c_int * 3)(*[int(v) for v in ymd])
    u = C.c_double(UTsec)
    if lib().orc_dateinc(C.c_double(dtsec), y, C.byref(u)):
        return None
    return [y[0], y[1], y[2]], u.value


def output_count(lx1, lx2, lx3, flagoutput):
    lib().orc_output_count.restype = C.c_long
    return int(lib().orc_output_count(lx1, lx2, lx3, int(flagoutput)))


def output_pack(st, gridflag, flagoutput, patch):
    """The real32 field datasets the output writers put in the file for state st, as a dict
    name -> array in the file's C-order shape (Fortran shape reversed), in file order.
    patch: the patch writer plasma_output_nompi, whose J1-J3 are the ghosted fields read
    from their first element (io_nompi.f90:192); otherwise the root writer (interior J)."""
    lx1, lx2, lx3 = (d - 4 for d in st.fluidvars.shape[:3])
    n, m = lx1 * lx2 * lx3, lx2 * lx3
    out = np.empty(output_count(lx1, lx2, lx3, flagoutput), dtype=np.float32)
    lib().orc_output_pack(lx1, lx2, lx3, int(gridflag), int(flagoutput), int(bool(patch)),
                          ptr(st.fluidvars),
                          ptr(st.electrovars), out.ctypes.data_as(C.c_void_p))
    if flagoutput == 2:
        names = [("neall", 0), ("v1avgall", 0), ("Tavgall", 0), ("Teall", 0), ("J1all", 0),
                 ("J2all", 0), ("J3all", 0), ("v2avgall", 0), ("v3avgall", 0)]
    elif flagoutput == 3:
        names = [("neall", 0)]
    else:
        names = [("nsall", 7), ("vs1all", 7), ("Tsall", 7), ("J1all", 0), ("J2all", 0),
                 ("J3all", 0), ("v2avgall", 0), ("v3avgall", 0)]
    res, off = {}, 0
    for name, nsp in names:
        k = (nsp or 1) * n
        shape = (nsp, lx3, lx2, lx1) if nsp else (lx3, lx2, lx1)
        res["/" + name] = out[off:off + k].reshape(shape)
        off += k
    res["/Phiall"] = out[off:off + m].reshape(lx3, lx2)
    assert off + m == out.size
    return res
