"""Gradients of a run of transport steps through the library's hand-written HIP adjoint.

The fitting extension gives the electron energy path two per-column parameters: a scale
``kappa`` on the electron thermal conductivity and a downward heat flux ``qtop`` [W/m^2]
through the top face (gridflag 2).  :class:`TransportRun` advances a fixed initial state
``nsteps`` steps for given parameters and returns the final electron temperature;
``backward`` returns dL/dkappa and dL/dqtop for any dL/dTe of that temperature.

Checkpointing is two-level: ``forward`` stores a restart checkpoint (fluidvars and the x1
boundary interface velocities, everything the next step reads) every ``every`` steps;
``backward`` walks the segments from the last, restores the segment's checkpoint,
re-runs its steps while recording the adjoint tape in HBM, then runs the adjoint steps
in reverse.  Memory: nsteps/every checkpoints plus one segment's tape; cost: one extra
forward step per step plus the adjoint kernels.  The replayed steps are the same kernels
on the same inputs, so the recorded trajectory is bit-identical to the forward one.

:func:`electron_temperature` is the ``torch.autograd.Function`` entry point.
"""
from __future__ import annotations

import math

import torch

from .grid import Grid
from .state import State
from .transport import GemhipError, Transport

TE_FIELD = 34           # fluidvars field of the electron temperature (Ts, species 7)


class TransportRun:
    """nsteps transport steps of one initial state, differentiable in kappa and qtop.

    kappa and qtop are torch float64 tensors on the run's device with lx2*lx3 elements,
    x2 fastest (e.g. shape (lx3, lx2)); qtop may be None (the reference's Dirichlet top).
    """

    def __init__(self, grid: Grid, state, dt: float, nsteps: int, every: int = 10,
                 device: int = 0, Teinf: float = 1500.0, geometry_on_device: bool = False):
        """state: a :class:`State`, or a mapping of device blocks (fluidvars, aux,
        electrovars, nn, Tn) laid out like the reference's, whose interface velocities
        start at zero as the reference allocates them."""
        if nsteps < 1 or every < 1:
            raise ValueError("nsteps and every must be positive")
        self.g, self.dt, self.nsteps, self.every = grid, float(dt), int(nsteps), int(every)
        self.dev = torch.device(f"cuda:{device}")
        blocks = not isinstance(state, State)
        nn, Tn = (state["nn"], state["Tn"]) if blocks else (state.nn, state.Tn)
        self.tr = Transport(grid, nn, Tn, device=device, Teinf=Teinf, own_context=True,
                            geometry_on_device=geometry_on_device)
        self.tr.set_residency(1)
        if blocks:
            self.tr.push_blocks(state["fluidvars"], state["aux"], state["electrovars"])
            self.tr.set_interface_vels_zero()
        else:
            self.tr.push(state)
        self.N = grid.padded[0] * grid.padded[1] * grid.padded[2]
        self.ncol = grid.lx2 * grid.lx3
        f64 = dict(dtype=torch.float64, device=self.dev)
        self.init = torch.empty(self.tr.checkpoint_size(), **f64)
        self.tr.checkpoint_save(self.init)
        self.nseg = math.ceil(self.nsteps / self.every)
        self.ckpt = torch.empty((self.nseg, self.init.numel()), **f64)
        self.tsize = self.tr.adjoint_tape_size()
        self.tape = torch.empty(min(self.every, self.nsteps) * self.tsize, **f64)
        self._final = torch.empty_like(self.init)
        self.generation = 0
        self.has_q = False

    def close(self):
        self.tr.close()

    def tape_bytes(self) -> int:
        return self.tape.numel() * 8

    def _params(self, kappa, qtop):
        torch.cuda.synchronize(self.dev)
        for name, t in (("kappa", kappa), ("qtop", qtop)):
            if t is not None and (t.numel() != self.ncol or t.dtype != torch.float64
                                  or t.device != self.dev):
                raise GemhipError(f"{name} must be {self.ncol} float64 values on {self.dev}")
        self.tr.set_conductivity_scale(None if kappa is None else kappa.detach().contiguous())
        self.tr.set_top_heat_flux(None if qtop is None else qtop.detach().contiguous())
        self.has_q = qtop is not None

    def forward(self, kappa, qtop=None) -> torch.Tensor:
        """Electron temperature after nsteps (ghosted field, N values, x1 fastest)."""
        self._params(kappa, qtop)
        self.tr.checkpoint_load(self.init)
        for k in range(self.nsteps):
            if k % self.every == 0:
                self.tr.checkpoint_save(self.ckpt[k // self.every])
            self.tr.transport_step(self.dt)
        self.tr.checkpoint_save(self._final)       # synchronises the library stream
        self.generation += 1
        return self._final[TE_FIELD * self.N:(TE_FIELD + 1) * self.N].clone()

    def backward(self, grad_Te: torch.Tensor):
        """dL/dkappa, dL/dqtop (lx2*lx3 each; None without a heat flux) for the last
        forward, given dL/dTe of its result."""
        lam = grad_Te.detach().to(device=self.dev, dtype=torch.float64).contiguous().clone()
        if lam.numel() != self.N:
            raise GemhipError(f"dL/dTe must have {self.N} values")
        gk = torch.zeros(self.ncol, dtype=torch.float64, device=self.dev)
        gq = torch.zeros(self.ncol, dtype=torch.float64, device=self.dev)
        torch.cuda.synchronize(self.dev)
        base, step = self.tape.data_ptr(), self.tsize * 8
        for seg in reversed(range(self.nseg)):
            k0, k1 = seg * self.every, min((seg + 1) * self.every, self.nsteps)
            self.tr.checkpoint_load(self.ckpt[seg])
            for k in range(k0, k1):
                self.tr.record_step(self.dt, base + (k - k0) * step)
            for k in reversed(range(k0, k1)):
                self.tr.adjoint_step(self.dt, base + (k - k0) * step, lam.data_ptr(),
                                     gk.data_ptr(), gq.data_ptr())
        self.tr.synchronize()
        self.grad_Te0 = lam                        # dL/dTe of the initial state
        return gk, (gq if self.has_q else None)


    # ------------------------------------------------------------------ branch records
    def tape_fields(self, tape: torch.Tensor) -> dict:
        """Views of one recorded step's tape: the electron energy density entering the
        x3, x1 and x2 sweeps (ghosted, (L3, L2, L1)), the electron density at step start
        and at the diffusion, the temperature entering the diffusion, and the electron
        interface velocities of the three directions."""
        L1, L2, L3 = self.g.padded
        lx1, lx2, lx3 = self.g.shape
        N = self.N
        f = lambda k: tape[k * N:(k + 1) * N].view(L3, L2, L1)
        o1 = 6 * N
        o2 = o1 + (lx1 + 1) * lx2 * lx3
        o3 = o2 + lx1 * (lx2 + 1) * lx3
        o4 = o3 + lx1 * lx2 * (lx3 + 1)
        npiv = lx2 * lx3 * 2 * ((lx1 + 31) // 32)
        return {"NE0": f(0), "ES3": f(1), "ES1": f(2), "ES2": f(3), "NE1": f(4), "TD": f(5),
                "V1": tape[o1:o2].view(lx3, lx2, lx1 + 1),
                "V2": tape[o2:o3].view(lx3, lx2 + 1, lx1),
                "V3": tape[o3:o4].view(lx3 + 1, lx2, lx1),
                # row interchanges of the electron columns' TR and BDF2 factorisations
                "PIV": tape[o4:o4 + (npiv + 1) // 2].view(torch.int32)[:npiv]}

    def recorded_steps(self, kappa, qtop=None):
        """Run the forward of ``kappa``, ``qtop`` step by step with the adjoint tape
        recorded, yielding (step, tape) after each step; the final electron temperature
        is in ``self.last_Te`` when the generator is exhausted.  The steps are the
        forward's kernels, so the trajectory is the forward's bit for bit."""
        self._params(kappa, qtop)
        self.tr.checkpoint_load(self.init)
        slot = self.tape[:self.tsize]
        for k in range(self.nsteps):
            self.tr.record_step(self.dt, slot.data_ptr())
            self.tr.synchronize()
            yield k, slot
            torch.cuda.synchronize(self.dev)       # the tape slot is read before reuse
        self.tr.checkpoint_save(self._final)
        self.generation += 1
        self.last_Te = self._final[TE_FIELD * self.N:(TE_FIELD + 1) * self.N].clone()


def limiter_codes(f: torch.Tensor, dx: torch.Tensor, v: torch.Tensor, axis: int):
    """The MC limiter's branch at every slope one sweep of the ghosted field f uses.

    f is (L3, L2, L1), axis 0 (x3), 1 (x2) or 2 (x1) the swept one; dx the coordinate
    differences dx(0:lx+2) of that axis; v the interface velocities (1:lx+1 along the
    axis, interior across it).  advec.f90:396-401 forms slope(i) = minmod(c, minmod(2l, 2r))
    for i = 0..lx+1; the code is 3*inner + outer, with inner 0 (2l 2r <= 0), 1 (|2l| <
    |2r|: 2l) or 2 (2r), and outer 0 (c m <= 0), 1 (|c| < |m|: c) or 2 (m).  Slopes that
    no flux reads (the upwind side of each interface, :416-421) are coded -1."""
    g = f[2:-2, 2:-2, :] if axis == 2 else (f[2:-2, :, 2:-2] if axis == 1 else f[:, 2:-2, 2:-2])
    g = g.movedim(axis, -1)
    lx = g.shape[-1] - 4
    fm, f0, fp = g[..., 0:lx + 2], g[..., 1:lx + 3], g[..., 2:lx + 4]
    dxl, dxr = dx[0:lx + 2], dx[1:lx + 3]
    l2 = 2.0 * ((f0 - fm) / dxl)
    r2 = 2.0 * ((fp - f0) / dxr)
    c = (fp - fm) / (dxl + dxr)
    inner = torch.where(l2 * r2 <= 0, 0, torch.where(l2.abs() < r2.abs(), 1, 2))
    m = torch.where(inner == 0, torch.zeros_like(l2), torch.where(inner == 1, l2, r2))
    outer = torch.where(c * m <= 0, 0, torch.where(c.abs() < m.abs(), 1, 2))
    code = (3 * inner + outer).to(torch.int8)
    vv = v.movedim(axis, -1)                       # interfaces k = 1..lx+1 at positions 0..lx
    used = torch.zeros_like(code, dtype=torch.bool)
    used[..., 1:] |= vv < 0                        # slope(i), i = 1..lx+1, read for v(i) < 0
    used[..., :lx + 1] |= vv >= 0                  # slope(i-1) read for v(i) >= 0
    return torch.where(used, code, torch.full_like(code, -1))


def branch_flips(run: TransportRun, tape_a: torch.Tensor, tape_b: torch.Tensor) -> dict:
    """Branches of one step that differ between two recorded trajectories of the same run
    setup: MC limiter choices of the electron energy sweeps (per direction), upwind sides
    (the velocities do not depend on the fitting parameters, so these stay 0), cells
    where the 100 K floor of rhoe2T holds the temperature entering the diffusion in
    exactly one of the two, and row interchanges of the electron columns' band solves."""
    a, b = run.tape_fields(tape_a), run.tape_fields(tape_b)
    dev = tape_a.device
    dxs = {k: torch.as_tensor(getattr(run.g, f"dx{k}"), device=dev) for k in (1, 2, 3)}
    out = {}
    for name, key, vkey, axis, dk in (("x3", "ES3", "V3", 0, 3), ("x1", "ES1", "V1", 2, 1),
                                      ("x2", "ES2", "V2", 1, 2)):
        if run.g.shape[dk - 1] == 1:
            out[name] = 0
            continue
        ca = limiter_codes(a[key], dxs[dk], a[vkey], axis)
        cb = limiter_codes(b[key], dxs[dk], b[vkey], axis)
        out[name] = int((ca != cb).sum().item())
        out["upwind_" + name] = int(((a[vkey] < 0) != (b[vkey] < 0)).sum().item())
    fa, fb = a["TD"] == 100.0, b["TD"] == 100.0
    out["floor"] = int((fa != fb).sum().item())
    # row interchanges of the band solves (LAPACK's partial pivoting) that differ
    x = torch.bitwise_xor(a["PIV"], b["PIV"])
    bits = torch.stack([(x >> k) & 1 for k in range(32)])
    out["pivot"] = int(bits.sum().item())
    return out


class _ElectronTemperature(torch.autograd.Function):
    @staticmethod
    def forward(ctx, kappa, qtop, run):
        Te = run.forward(kappa, qtop)
        ctx.run, ctx.gen = run, run.generation
        ctx.kshape = kappa.shape
        ctx.qshape = None if qtop is None else qtop.shape
        return Te

    @staticmethod
    def backward(ctx, grad_Te):
        if ctx.run.generation != ctx.gen:
            raise RuntimeError("the run was advanced again before this backward")
        gk, gq = ctx.run.backward(grad_Te)
        return gk.view(ctx.kshape), (None if gq is None else gq.view(ctx.qshape)), None


def electron_temperature(run: TransportRun, kappa, qtop=None) -> torch.Tensor:
    """Final electron temperature of ``run`` as a differentiable function of kappa, qtop."""
    return _ElectronTemperature.apply(kappa, qtop, run)


def interior(run: TransportRun, Te: torch.Tensor) -> torch.Tensor:
    """(lx3, lx2, lx1) interior view of a ghosted field returned by the run."""
    L1, L2, L3 = run.g.padded
    return Te.view(L3, L2, L1)[2:-2, 2:-2, 2:-2]
