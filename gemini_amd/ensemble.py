"""Ensembles of independent grids advanced together on one GPU.

Each member is a :class:`gemini_amd.transport.Transport` with a library context of its
own (its grid, state in HBM, scratch and stream).  ``Ensemble.run`` captures the members'
transport steps into one HIP graph, in which the members run concurrently, and launches
it; ``Ensemble.run_plain`` issues the same steps member by member.  Results are the same
bit for bit: the graph only changes the order in which independent kernels are issued.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .transport import Transport, GemhipError, load_library


def draw_sizes(seed: int, n: int):
    """Seeded member sizes of the batched configuration: lx1 from {64, 96, ..., 512},
    lx2 from 16-128, lx3 from 16-96."""
    rng = np.random.default_rng(seed)
    return [(int(rng.choice(np.arange(64, 513, 32))), int(rng.integers(16, 129)),
             int(rng.integers(16, 97))) for _ in range(n)]


def device_member(shape, seed: int, own_context: bool = True, device: int = 0):
    """(grid, Transport with the seeded state generated and pushed in HBM, dt, state
    blocks) for one member: a Cartesian grid of ``shape``, the CFL time step of its
    state.  The interface velocities start at zero, as the reference allocates them."""
    from .grid import cartesian_grid
    from .synth_gpu import synthetic_state_device, cfl_dt_device
    g = cartesian_grid(*shape)
    st = synthetic_state_device(g, seed=seed, device=f"cuda:{device}")
    dt = cfl_dt_device(g, st)
    tr = Transport(g, st["nn"], st["Tn"], device=device, geometry_on_device=True,
                   own_context=own_context)
    tr.set_residency(1)
    tr.push_blocks(st["fluidvars"], st["aux"], st["electrovars"])
    tr.set_interface_vels_zero()
    return g, tr, dt, st


class Ensemble:
    def __init__(self, members, device: int = 0, diffsolvetype: int = 2, Teinf: float = 1500.0):
        """members: sequence of (grid, state, dt).  States are pushed to HBM (residency
        mode 1); use pull(i) to read one back."""
        self.lib = load_library()
        self.grids = [m[0] for m in members]
        self.dts = np.array([float(m[2]) for m in members], dtype=np.float64)
        self.trs = []
        for g, st, dt in members:
            tr = Transport(g, st.nn, st.Tn, device=device, diffsolvetype=diffsolvetype,
                           Teinf=Teinf, own_context=True)
            tr.set_residency(1)
            tr.push(st)
            self.trs.append(tr)
        self._graphs = {}

    @classmethod
    def from_transports(cls, trs, grids, dts):
        """Members already set up (own contexts, residency mode 1, state pushed)."""
        self = cls.__new__(cls)
        self.lib = load_library()
        self.grids = list(grids)
        self.dts = np.array([float(d) for d in dts], dtype=np.float64)
        self.trs = list(trs)
        self._graphs = {}
        return self

    def __len__(self):
        return len(self.trs)

    def _handles(self):
        return (C.c_int * len(self.trs))(*[t.handle for t in self.trs])

    def capture(self, nsteps: int) -> int:
        """Capture nsteps steps of every member into one graph (cached per nsteps)."""
        if nsteps in self._graphs:
            return self._graphs[nsteps]
        self.trs[0]._sel()
        gid = self.lib.gemhip_ensemble_capture(len(self.trs), self._handles(),
                                               self.dts.ctypes.data_as(C.POINTER(C.c_double)),
                                               int(nsteps))
        if gid <= 0:
            msgs = [t.lib.gemhip_last_error().decode() for t in self.trs]
            raise GemhipError("ensemble capture failed: " + "; ".join(m for m in msgs if m))
        self._graphs[nsteps] = gid
        return gid

    def run(self, nsteps: int, wait: bool = True):
        """Advance every member nsteps steps with one graph launch."""
        gid = self.capture(nsteps)
        if self.lib.gemhip_ensemble_launch(gid, int(wait)) != 0:
            raise GemhipError("ensemble graph launch failed")

    def run_plain(self, nsteps: int):
        """The same steps issued member after member, without a graph."""
        for _ in range(nsteps):
            for tr, dt in zip(self.trs, self.dts):
                tr.transport_step(float(dt))
        for tr in self.trs:
            tr.synchronize()

    def autotune(self, states, steps: int = 3):
        """Choose each member's kernel variants by measurement, then re-push `states`.
        Graphs captured before hold the old variants, so they are released."""
        from .autotune import autotune
        self._release_graphs()
        return autotune(self, states, steps)

    def _release_graphs(self):
        for gid in self._graphs.values():
            self.lib.gemhip_ensemble_release(gid)
        self._graphs.clear()

    def synchronize(self):
        for tr in self.trs:
            tr.synchronize()

    def push(self, i: int, st):
        self.trs[i].push(st)

    def pull(self, i: int, st):
        self.trs[i].pull(st)

    def batched(self) -> "Batch":
        """The members advanced by batched kernels (see :class:`Batch`)."""
        return Batch(self)

    def close(self):
        self._release_graphs()
        for tr in self.trs:
            tr.close()
        self.trs = []


class Batch:
    """An ensemble's members advanced by batched kernels: each kernel of the fused
    transport step runs once for all members, a persistent grid walking a table of every
    member's blocks, instead of once per member.  Members must take the fused step's path
    (global sides, x3 not periodic, lx2 and lx3 above 1) with TRBDF2 diffusion and no
    fitting extension.  Results are bit-identical to the members' plain steps.

    ``step(n)`` issues n steps on the batch's stream; ``run(n)`` launches n steps captured
    in one HIP graph (cached per n); the knobs ``sweep_pf``, ``diff_chunk`` (kernel
    variants) and ``grid`` (persistent blocks per CU) change speed only."""

    KNOBS = {"sweep_pf": (1, 2), "diff_chunk": (4, 8), "grid": (1, 2)}

    def __init__(self, ensemble: Ensemble):
        self.ens = ensemble
        self.lib = ensemble.lib
        ensemble.trs[0]._sel()
        self.id = self.lib.gemhip_batch_create(len(ensemble.trs), ensemble._handles())
        if self.id <= 0:
            raise GemhipError("batch creation failed: " + self.lib.gemhip_last_error().decode())
        self._graphs = {}
        self.tuning = {"sweep_pf": 2, "diff_chunk": 8, "grid": 1}
        self.set_dt(ensemble.dts)

    def _sel(self):
        self.ens.trs[0]._sel()          # errors are reported in the first member's context

    def _check(self, rc, what):
        if rc != 0:
            raise GemhipError(f"{what} failed: " + self.lib.gemhip_last_error().decode())

    def set_dt(self, dts):
        """Per-member time steps; captured graphs read them from HBM at every launch."""
        d = np.ascontiguousarray(np.asarray(dts, dtype=np.float64))
        if d.shape != (len(self.ens),):
            raise ValueError("one time step per member")
        self._sel()
        self._check(self.lib.gemhip_batch_set_dt(self.id, d.ctypes.data_as(C.POINTER(C.c_double))),
                    "batch dt")
        self.ens.dts = d.copy()

    def set_tuning(self, knob: str, value: int):
        self._sel()
        self._check(self.lib.gemhip_batch_set_tuning(self.id, knob.encode(), int(value)),
                    "batch tuning")
        if self.tuning.get(knob) != value:
            self._release_graphs()           # graphs hold the variants they were captured with
        self.tuning[knob] = value

    def step(self, nsteps: int = 1, wait: bool = True):
        self._sel()
        for _ in range(nsteps):
            self._check(self.lib.gemhip_batch_step(self.id), "batch step")
        if wait:
            self.synchronize()

    def capture(self, nsteps: int) -> int:
        if nsteps not in self._graphs:
            self._sel()
            gid = self.lib.gemhip_batch_capture(self.id, int(nsteps))
            if gid <= 0:
                raise GemhipError("batch capture failed: " + self.lib.gemhip_last_error().decode())
            self._graphs[nsteps] = gid
        return self._graphs[nsteps]

    def run(self, nsteps: int, wait: bool = True):
        gid = self.capture(nsteps)
        self._sel()
        self._check(self.lib.gemhip_ensemble_launch(gid, int(wait)), "batch graph launch")

    def synchronize(self):
        self._sel()
        self._check(self.lib.gemhip_batch_synchronize(self.id), "batch synchronize")

    def autotune(self, restore, steps: int = 2):
        """Time every combination of KNOBS over `steps` batched steps, keep the fastest,
        then call ``restore()`` to put the members' states back (timing advances them)."""
        import itertools
        import time
        times = {}
        for combo in itertools.product(*self.KNOBS.values()):
            for k, v in zip(self.KNOBS, combo):
                self.set_tuning(k, v)
            self.step(1)                                   # warm-up of this variant
            t0 = time.perf_counter()
            self.step(steps)
            times[combo] = (time.perf_counter() - t0) / steps
        best = min(times, key=times.get)
        for k, v in zip(self.KNOBS, best):
            self.set_tuning(k, v)
        restore()
        return dict(zip(self.KNOBS, best)), times

    def _release_graphs(self):
        for gid in self._graphs.values():
            self.lib.gemhip_ensemble_release(gid)
        self._graphs.clear()

    def close(self):
        if self.id > 0:
            self._sel()
            self._release_graphs()
            self.lib.gemhip_batch_release(self.id)
            self.id = 0
