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

    def close(self):
        self._release_graphs()
        for tr in self.trs:
            tr.close()
        self.trs = []
