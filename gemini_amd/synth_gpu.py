"""Synthetic transport state generated directly in HBM with torch (benchmark inputs).

Same profiles as :func:`gemini_amd.state.synthetic_state` (Chapman-layer ions,
smooth temperatures, smooth drifts, tapered parallel current, exponential
neutrals), drawn from a seeded ``torch.Generator`` on the device so the
25M-cell configurations need no host memory.  Tensors are C-order with the
species/field index first, e.g. ``fluidvars`` is (35, L3, L2, L1): the same bytes
as the reference's Fortran ``fluidvars(-1:lx1+2,-1:lx2+2,-1:lx3+2,35)``.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .grid import Grid

LSP = 7
MINDENS = 1.0e-100
# every ion species carries a 10 m^-3 floor (negligible against 1e9-1e12 m^-3 plasma), so
# no species is so rare that ratios of advected quantities (vs1 = rhovs1/(ms ns)) run away
# in a transport-only run without the reference's source and clean-up substeps
BACKGROUND = 10.0


def _chapman(z, zpk, H):
    zeta = (z - zpk) / H
    return torch.exp(0.5 * (1.0 - zeta - torch.exp(-zeta)))


def _sig(z, z0, w):
    return 0.5 * (1.0 + torch.tanh((z - z0) / w))


def _uniform(gen, lo, hi, shape, dev):
    return torch.rand(shape, generator=gen, device=dev, dtype=torch.float64) * (hi - lo) + lo


def _smooth_field(gen, shape, out, scale, dev, nmodes=3, noise=0.05):
    u = [torch.linspace(0.0, 1.0, n, device=dev, dtype=torch.float64) for n in shape]
    out.zero_()
    for _ in range(nmodes):
        k = _uniform(gen, 0.3, 2.0, (3,), dev)
        ph = _uniform(gen, 0.0, 2 * math.pi, (3,), dev)
        parts = [torch.sin(2 * math.pi * k[i] * u[i] + ph[i]) for i in range(3)]
        amp = _uniform(gen, 0.5, 1.0, (1,), dev)
        plane = amp * parts[1][:, None] * parts[2][None, :]
        out += parts[0][:, None, None] * plane[None, :, :]
    out *= (1.0 - noise) * scale / out.abs().max().clamp_min(1e-30)
    out += (noise * scale) * _uniform(gen, -1.0, 1.0, shape, dev)
    return out


def _gaussian_perturbations(g: Grid, gen, FV, dev, nblobs: int = 6):
    """Six 3-D Gaussian blobs (widths 20-50 km, centres in the interior) raising the ion
    densities by up to 20 %."""
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
    x1, x2, x3 = t(g.x1)[None, None, :], t(g.x2)[None, :, None], t(g.x3)[:, None, None]
    pert = torch.zeros(FV.shape[1:], device=dev, dtype=torch.float64)
    for _ in range(nblobs):
        c = [float(_uniform(gen, x[2], x[n + 1], (1,), dev))
             for x, n in ((g.x1, g.lx1), (g.x2, g.lx2), (g.x3, g.lx3))]
        w = float(_uniform(gen, 20e3, 50e3, (1,), dev))
        pert += torch.exp(-((x1 - c[0]) ** 2 + (x2 - c[1]) ** 2 + (x3 - c[2]) ** 2) / (2 * w * w))
    fac = 1.0 + 0.2 * pert.clamp_max(1.0)
    for isp in range(6):
        FV[isp].mul_(fac)


def synthetic_state_device(g: Grid, seed: int = 0, device="cuda", perturb: bool = False):
    """Dict of device tensors: fluidvars (35,L3,L2,L1), aux (14,L3,L2,L1) (rhovs1,
    rhoes left zero), electrovars (7,L3,L2,L1) with J1 in [3], nn (3,lx3,lx2,lx1),
    Tn (lx3,lx2,lx1).  ``perturb`` adds 3-D Gaussian density perturbations (amplitude
    20 %, widths 20-50 km), as the 3-D Cartesian configuration asks."""
    dev = torch.device(device)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    L1, L2, L3 = g.padded
    sh = (L3, L2, L1)
    f64 = torch.float64
    FV = torch.empty((35,) + sh, device=dev, dtype=f64)
    z = torch.as_tensor(g.alt_plane(), device=dev)                  # (L2, L1)
    nmax = 10.0 ** _uniform(gen, 11.0, 12.0, (L2, 1), dev)
    nF = nmax * _chapman(z, 300e3, 60e3)
    nE = 0.05 * nmax * _chapman(z, 110e3, 12e3)
    fO = _sig(z, 180e3, 25e3)
    fH = _sig(z, 1000e3, 150e3)
    plasmasphere = 3e8 * torch.exp(-(z - 600e3).clamp_min(0.0) / 6000e3) * _sig(z, 500e3, 100e3)
    mol = nF * (1.0 - fO) + nE
    prof = [nF * fO * (1.0 - fH), 0.55 * mol, 0.03 * mol, 0.40 * mol, 0.02 * nF * fO,
            nF * fO * fH + plasmasphere]
    for isp in range(6):
        noise = torch.randn(sh, generator=gen, device=dev, dtype=f64)
        noise.mul_(0.05).exp_().mul_((prof[isp] + BACKGROUND)[None])
        torch.clamp_min(noise, MINDENS, out=FV[isp])
        del noise
    if perturb:
        _gaussian_perturbations(g, gen, FV, dev)
    torch.sum(FV[0:6], dim=0, out=FV[6])
    Tn2 = 1000.0 - 800.0 * torch.exp(-(z - 80e3).clamp_min(0.0) / 40e3)
    Ti = Tn2 + (1500.0 - Tn2) * _sig(z, 300e3, 100e3)
    Te = Tn2 + (3000.0 - Tn2) * _sig(z, 200e3, 80e3)
    for isp in range(6):
        FV[28 + isp].copy_(Ti[None].expand(sh))
    FV[34].copy_(Te[None].expand(sh))
    for isp in range(LSP - 1):
        _smooth_field(gen, sh, FV[7 + isp], 200.0, dev)
    flux = torch.zeros(sh, device=dev, dtype=f64)
    for isp in range(6):
        flux.addcmul_(FV[isp], FV[7 + isp])
    torch.div(flux, FV[6].clamp_min(1e-5), out=FV[13])
    del flux
    _smooth_field(gen, sh, FV[14], 500.0, dev)
    _smooth_field(gen, sh, FV[21], 500.0, dev)
    for isp in range(1, LSP):
        FV[14 + isp].copy_(FV[14])
        FV[21 + isp].copy_(FV[21])
    electrovars = torch.zeros((7,) + sh, device=dev, dtype=f64)
    J1 = electrovars[3]
    _smooth_field(gen, sh, J1, 1e-6, dev)
    J1.mul_((FV[6] / 1e10).clamp_max(1.0))
    zi = z[2:g.lx2 + 2, 2:g.lx1 + 2]
    Hs = lambda H0, z0: H0 + 0.06 * (zi - z0).clamp_min(0.0)
    nn = torch.empty((3, g.lx3, g.lx2, g.lx1), device=dev, dtype=f64)
    for k, a in enumerate((3e17 * torch.exp(-(zi - 100e3) / Hs(8e3, 100e3)),
                           1e19 * torch.exp(-(zi - 80e3) / Hs(7e3, 80e3)),
                           2.5e18 * torch.exp(-(zi - 80e3) / Hs(6.5e3, 80e3)))):
        nn[k].copy_(a[None].expand(g.lx3, g.lx2, g.lx1))
    Tn = Tn2[2:g.lx2 + 2, 2:g.lx1 + 2][None].expand(g.lx3, g.lx2, g.lx1).contiguous()
    aux = torch.zeros((14,) + sh, device=dev, dtype=f64)
    return dict(fluidvars=FV, aux=aux, electrovars=electrovars, nn=nn, Tn=Tn)


def cfl_dt_device(g: Grid, st, tcfl: float = 0.9, dtmax: float = 0.5) -> float:
    """temporal.f90:12-49 evaluated on the device tensors, capped at dtmax."""
    kB = 1.38064852e-23
    amu = 1.660539040e-27
    ms = np.array([16, 30, 28, 32, 14, 1, 5.485799090e-4]) * amu
    dev = st["fluidvars"].device
    sl = (slice(2, g.lx3 + 2), slice(2, g.lx2 + 2), slice(2, g.lx1 + 2))
    FV = st["fluidvars"]
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
    h = lambda n: t(g.planes[n][0])[2:g.lx2 + 2, 2:g.lx1 + 2][None]
    dl1 = t(g.dx1i)[None, None, :] * h("h1")
    dl2 = t(g.dx2i)[None, :, None] * h("h2")
    dl3 = t(g.dx3i)[:, None, None] * h("h3")
    Te = FV[34][sl]
    rate = 0.0
    for isp in range(LSP):
        vsnd = torch.sqrt(kB * FV[28 + isp][sl] / ms[isp] + 5.0 / 3.0 * kB * Te / ms[isp]) \
            if isp < LSP - 1 else 0.0
        rate = max(rate, float(((vsnd + FV[7 + isp][sl].abs()) / dl1).max()),
                   float((FV[14 + isp][sl].abs() / dl2).max()),
                   float((FV[21 + isp][sl].abs() / dl3).max()))
    return min(dtmax, tcfl / rate) if rate > 0 else dtmax


def synthetic_atmosphere_device(g: Grid, nn3, Tn, ne, seed: int = 0, device="cuda"):
    """Neutral atmosphere and ionization rates for the source/loss step, as torch tensors.

    nn3 (3, lx3, lx2, lx1) and Tn (lx3, lx2, lx1) are the state's O, N2, O2 densities and
    neutral temperature (the ones the energy diffusion uses); H, N and NO are added as
    exponential profiles, winds are smooth fields of tens of m/s and the ionization rates
    a Chapman production layer (floored at 1e-5 m^-3 s^-1 as the reference floors them)
    with the Swartz and Nisbet (1973) thermal-electron heating the reference's eheating
    (ionization.f90:281-306) gives for it at electron density ne (lx3, lx2, lx1, the
    interior cells).  Returns nn (6, lx3, lx2, lx1): O, N2, O2,
    H, N, NO; Tn, vn1, vn2, vn3, Qeionize (lx3, lx2, lx1); Prionize (6, lx3, lx2, lx1):
    C-order bytes of the reference's Fortran intvars arrays.
    """
    dev = torch.device(device)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed + 7919)
    f64 = torch.float64
    sh = (g.lx3, g.lx2, g.lx1)
    z = torch.as_tensor(g.alt_plane(), device=dev)[2:g.lx2 + 2, 2:g.lx1 + 2]   # (lx2, lx1)
    nn = torch.empty((6,) + sh, device=dev, dtype=f64)
    nn[0:3].copy_(torch.as_tensor(nn3, device=dev).reshape((3,) + sh))
    zc = z.clamp_min(80e3)
    for k, prof in ((3, 1e11 * torch.exp(-(zc - 500e3) / 2000e3)),        # H
                    (4, 5e12 * torch.exp(-(zc - 150e3) / 40e3)),          # N
                    (5, 4e14 * torch.exp(-(zc - 100e3) / 8e3))):          # NO
        noise = torch.randn(sh, generator=gen, device=dev, dtype=f64)
        nn[k].copy_(noise.mul_(0.05).exp_().mul_(prof[None]))
    winds = []
    for scale in (10.0, 60.0, 60.0):
        w = torch.empty(sh, device=dev, dtype=f64)
        winds.append(_smooth_field(gen, sh, w, scale, dev))
    Prionize = torch.empty((6,) + sh, device=dev, dtype=f64)
    q = 2e9 * _chapman(z, 140e3, 25e3)[None].expand(sh)
    for k, frac in enumerate((0.45, 0.15, 0.25, 0.12, 0.03, 1e-4)):
        noise = torch.randn(sh, generator=gen, device=dev, dtype=f64)
        Prionize[k].copy_((noise.mul_(0.05).exp_() * (frac * q)).clamp_min(1e-5))
    R = torch.log(torch.as_tensor(ne, device=dev).reshape(sh) / (nn[1] + nn[2] + 0.1 * nn[0]))
    avgenergy = torch.exp(-(12.75 + 6.941 * R + 1.166 * R ** 2 + 0.08034 * R ** 3 + 0.001996 * R ** 4))
    Qeionize = (1.60217662e-19 * avgenergy) * Prionize.sum(0)
    del R, avgenergy
    return dict(nn=nn, Tn=torch.as_tensor(Tn, device=dev).reshape(sh).contiguous(),
                vn1=winds[0], vn2=winds[1], vn3=winds[2], Prionize=Prionize, Qeionize=Qeionize)
