"""NumPy FP64 discrete adjoint of the electron energy path -- TEST INFRASTRUCTURE ONLY.

The checker for the HIP adjoint (gemini_amd/csrc/kernels_adjoint.hip).  Only the
tests use it.  It is built so that it shares no derivative code with the HIP kernels:

* forward states come from the C oracle, one reference operation at a time, with the
  fitting extension (electron conductivity scale ``kappa`` and top heat flux ``qtop`` per
  (x2, x3) column);
* the local Jacobians of the two non-trivial maps come from complex-step differentiation
  of plain NumPy restatements, with every branch decided on the real parts (the frozen
  branches of the discrete adjoint):
    - one MC-limited sweep line (advec.f90:384-435), all lines of a direction at once;
    - the TRBDF2 band matrices and right-hand sides of one column
      (PDEparabolic.f90:19-218 with collisions.f90:162-199 and diffusion.f90:44-76);
* the transposed TRBDF2 systems are solved densely (numpy.linalg.solve);
* the copy-type maps (ghost fill advec.f90:21-149, periodic wrap mpihalo.f90:147-152,
  T2rhoe / rhoe2T multifluid.f90:179-202) are transposed by hand, one line each.

Central finite differences of the C oracle's forward pin it (tests/test_adjoint.py).
"""
from __future__ import annotations

import numpy as np

from . import constants, op, transport_step

LSP = 7
MINDENSDIV = 1.0e-5
H_CS = 1e-30          # complex-step size relative to each input's magnitude


def _consts():
    c = constants()
    return c["kB"], c["elchrg"], c["gammas"][LSP - 1]


# ---------------------------------------------------------------- forward with a tape
def record_step(g, st, dt, kappa=None, qtop=None, Teinf=1500.0):
    """One oracle transport step on ``st`` (in place) in the order of orc_transport_step,
    keeping the electron fields the adjoint needs."""
    e = LSP - 1
    t = {}
    op(g, st, "v12rhov1")
    op(g, st, "T2rhoe")
    t["NE0"] = st.fluidvars[..., e].copy(order="F")
    op(g, st, "set_global_boundaries_allspec")
    op(g, st, "halo_fluidvars")
    op(g, st, "interface_vels_allspec")
    t["ES3"] = st.fluidauxvars[..., 7 + e].copy(order="F")
    t["V1"] = st.vs1i[..., e].copy(order="F")
    t["V2"] = st.vs2i[..., e].copy(order="F")
    t["V3"] = st.vs3i[..., e].copy(order="F")
    for w in ("mass", "momentum", "energy"):
        op(g, st, "sweep", dt=dt, direction=3, which=w)
    t["ES1"] = st.fluidauxvars[..., 7 + e].copy(order="F")
    for w in ("mass", "momentum", "energy"):
        op(g, st, "sweep", dt=dt, direction=1, which=w)
    op(g, st, "halo_allparams")
    t["ES2"] = st.fluidauxvars[..., 7 + e].copy(order="F")
    for w in ("mass", "momentum", "energy"):
        op(g, st, "sweep", dt=dt, direction=2, which=w)
    op(g, st, "rhov12v1")
    op(g, st, "rhoe2T")
    t["NE1"] = st.fluidvars[..., e].copy(order="F")
    t["TD"] = st.fluidvars[..., 28 + e].copy(order="F")
    op(g, st, "energy_diffusion", dt=dt, kappa=kappa, qtop=qtop, Teinf=Teinf)
    return t


# ---------------------------------------------------------------- sweeps
def _minmod(a, b):
    """advec.f90:438-447 on the real parts; returns the chosen complex operand."""
    zero = (a.real * b.real) <= 0.0
    first = np.abs(a.real) < np.abs(b.real)
    return np.where(zero, 0.0 * a, np.where(first, a, b))


def sweep_lines(f, v, dt, dx, dxi, ha1, ha2i, h1i):
    """advec1D_MC_curv over many lines: f, ha1 (nl, lx+4); v, ha2i, h1i (nl, lx+1);
    dx (lx+3) = dx(0:lx+2), dxi (lx) = dxi(1:lx).  Complex f is allowed."""
    lx = f.shape[1] - 4
    F = lambda k: f[:, k + 1]
    slope = [None] * (lx + 2)
    for k in range(0, lx + 2):
        l2 = 2.0 * ((F(k) - F(k - 1)) / dx[k])
        r2 = 2.0 * ((F(k + 1) - F(k)) / dx[k + 1])
        c = (F(k + 1) - F(k - 1)) / (dx[k] + dx[k + 1])
        slope[k] = _minmod(c, _minmod(l2, r2))
    phi = [None] * (lx + 2)
    for i in range(1, lx + 2):
        vv = v[:, i - 1]
        hi = h1i[:, i - 1]
        neg = F(i) * vv - ((0.5 * vv) * (dx[i] + (vv / hi) * dt)) * slope[i]
        pos = F(i - 1) * vv + ((0.5 * vv) * (dx[i] - (vv / hi) * dt)) * slope[i - 1]
        phi[i] = np.where(vv < 0.0, neg, pos)
    out = f.copy()
    for i in range(1, lx + 1):
        out[:, i + 1] = F(i) - ((dt * (ha2i[:, i] * phi[i + 1] - ha2i[:, i - 1] * phi[i]))
                                / dxi[i - 1]) / ha1[:, i + 1]
    return out


def _line_geometry(g, d):
    """Lines of direction d through the interior: a function mapping a ghosted field to
    (nl, lx+4) lines and back, and the per-line metric arrays of the energy sweep."""
    lx1, lx2, lx3 = g.lx1, g.lx2, g.lx3
    h1, h2, h3 = (np.asarray(getattr(g, n)) for n in ("h1", "h2", "h3"))
    ha1_field = (h1 * h2) * h3
    if d == 1:
        take = lambda a: a[:, 2:lx2 + 2, 2:lx3 + 2].transpose(1, 2, 0).reshape(lx2 * lx3, lx1 + 4)
        face = lambda a: np.asarray(a).transpose(1, 2, 0).reshape(lx2 * lx3, lx1 + 1)
        ha2i = face(g.h2x1i) * face(g.h3x1i)
        h1i = face(g.h1x1i)
        dx, dxi, lx = np.asarray(g.dx1), np.asarray(g.dx1i), lx1

        def put(a, lines):
            a[:, 2:lx2 + 2, 2:lx3 + 2] = lines.reshape(lx2, lx3, lx1 + 4).transpose(2, 0, 1)
    elif d == 2:
        take = lambda a: a[2:lx1 + 2, :, 2:lx3 + 2].transpose(0, 2, 1).reshape(lx1 * lx3, lx2 + 4)
        face = lambda a: np.asarray(a).transpose(0, 2, 1).reshape(lx1 * lx3, lx2 + 1)
        ha2i = face(g.h1x2i) * face(g.h3x2i)
        h1i = face(g.h2x2i)
        dx, dxi, lx = np.asarray(g.dx2), np.asarray(g.dx2i), lx2

        def put(a, lines):
            a[2:lx1 + 2, :, 2:lx3 + 2] = lines.reshape(lx1, lx3, lx2 + 4).transpose(0, 2, 1)
    else:
        take = lambda a: a[2:lx1 + 2, 2:lx2 + 2, :].reshape(lx1 * lx2, lx3 + 4)
        face = lambda a: np.asarray(a).reshape(lx1 * lx2, lx3 + 1)
        ha2i = face(g.h1x3i) * face(g.h2x3i)
        h1i = face(g.h3x3i)
        dx, dxi, lx = np.asarray(g.dx3), np.asarray(g.dx3i), lx3

        def put(a, lines):
            a[2:lx1 + 2, 2:lx2 + 2, :] = lines.reshape(lx1, lx2, lx3 + 4)
    return take, face, put, take(ha1_field), ha2i, h1i, dx, dxi, lx


def sweep_adjoint(g, d, f, vface, lam, dt):
    """Transpose of the energy sweep along d at input f: returns the adjoint before it."""
    take, face, put, ha1, ha2i, h1i, dx, dxi, lx = _line_geometry(g, d)
    out = lam.copy(order="F")
    if lx <= 1:
        return out
    fl = take(np.asarray(f)).astype(np.complex128)
    vl = face(vface)
    lo = take(np.asarray(lam))
    n = lx + 4
    J = np.zeros((fl.shape[0], n, n))              # J[line, out, in]
    for k in range(n):
        h = H_CS * max(np.abs(fl[:, k].real).max(), 1.0)
        fp = fl.copy()
        fp[:, k] += 1j * h
        J[:, :, k] = sweep_lines(fp, vl, dt, dx, dxi, ha1, ha2i, h1i).imag / h
    put(out, np.einsum("lok,lo->lk", J, lo))
    return out


# ---------------------------------------------------------------- copy-type maps
def halo_adjoint(g, lam):
    """mpihalo.f90:147-152 (periodic x3, depth 2, interior x2) transposed."""
    lx2, lx3 = g.lx2, g.lx3
    s = slice(2, lx2 + 2)
    for h in range(2):
        lam[:, s, lx3 - 1 + h + 1] += lam[:, s, h]          # ghosts -1, 0 <- lx3-1, lx3
        lam[:, s, h] = 0.0
    for h in range(2):
        lam[:, s, 1 + h + 1] += lam[:, s, lx3 + 1 + h + 1]  # ghosts lx3+1, lx3+2 <- 1, 2
        lam[:, s, lx3 + 1 + h + 1] = 0.0
    return lam


def fill_adjoint(g, lam, ne0):
    """advec.f90:42-132 for the electron energy density, transposed in reverse order."""
    lx1, lx2, lx3 = g.lx1, g.lx2, g.lx3
    I = lambda i: i + 1                      # Fortran index (lower bound -1) -> numpy
    isg = [bool(v) for v in g.isglobal]
    if not g.flagper:
        if isg[5]:
            lam[:, :, I(lx3)] += lam[:, :, I(lx3 + 1)] + lam[:, :, I(lx3 + 2)]
            lam[:, :, I(lx3 + 1)] = 0.0
            lam[:, :, I(lx3 + 2)] = 0.0
        if isg[4]:
            lam[:, :, I(1)] += lam[:, :, I(0)] + lam[:, :, I(-1)]
            lam[:, :, I(0)] = 0.0
            lam[:, :, I(-1)] = 0.0
    if isg[3]:
        lam[:, I(lx2), :] += lam[:, I(lx2 + 1), :] + lam[:, I(lx2 + 2), :]
        lam[:, I(lx2 + 1), :] = 0.0
        lam[:, I(lx2 + 2), :] = 0.0
    if isg[2]:
        lam[:, I(1), :] += lam[:, I(0), :] + lam[:, I(-1), :]
        lam[:, I(0), :] = 0.0
        lam[:, I(-1), :] = 0.0
    # densities as the x1 fill leaves them before it fills the energy density
    n = np.array(ne0, copy=True, order="F")
    inner = (slice(None), slice(2, lx2 + 2), slice(2, lx3 + 2))
    ni = n[inner]
    if isg[0]:
        coeff = ni[I(2)] / ni[I(3)]
        n0 = np.minimum(coeff * ni[I(1)], ni[I(1)])
        ni[I(0)] = n0
        ni[I(-1)] = np.minimum(coeff * n0, n0)
    if isg[1]:
        coeff = ni[I(lx1 - 1)] / ni[I(lx1 - 2)]
        nP1 = np.minimum(coeff * ni[I(lx1)], ni[I(lx1)])
        ni[I(lx1 + 1)] = nP1
        ni[I(lx1 + 2)] = np.minimum(coeff * nP1, nP1)
    n[inner] = ni
    if isg[1]:
        lam[I(lx1)] += (n[I(lx1 + 1)] / n[I(lx1)]) * lam[I(lx1 + 1)] + \
            (n[I(lx1 + 2)] / n[I(lx1)]) * lam[I(lx1 + 2)]
        lam[I(lx1 + 1)] = 0.0
        lam[I(lx1 + 2)] = 0.0
    if isg[0]:
        lam[I(1)] += (n[I(0)] / n[I(1)]) * lam[I(0)] + (n[I(-1)] / n[I(1)]) * lam[I(-1)]
        lam[I(0)] = 0.0
        lam[I(-1)] = 0.0
    return lam


# ---------------------------------------------------------------- TRBDF2 column
def lambda_e(T, ne, n1, n2, n3):
    """Electron thermal conductivity (collisions.f90:185-195); complex T allowed."""
    _, e, _ = _consts()
    sq = np.sqrt(T)
    br = ((n1 * 1.1e-16) * (1.0 + 5.7e-4 * T) + ((n2 * 2.82e-17) * sq) * (1.0 - 1.21e-4 * T)) \
        + (n3 * 2.2e-16) * (1.0 + 3.6e-2 * sq)
    den = 1.0 + ((3.22e4 * (T * T)) / ne) * br
    return (((e * 100.0) * 7.7e5) * T ** 2.5) / den


def assemble(T, kap, q, col, dt, Tmin, Tmax_dirichlet):
    """TR and BDF2 band systems of one electron column as dense matrices.

    T (nb, N) complex batch; kap, q (nb,) complex.  col: dict of the column's real inputs
    (ne, n1, n2, n3, J, h1, h2, h3 of length N, dx1 (N+3), dx1i (N), flux flag).
    Returns M1, r1, M2 (nb, N, N), r2T (nb, N): the BDF2 rhs without its TR term."""
    kB, e, gam = _consts()
    nb, N = T.shape
    h1, h2, h3 = col["h1"], col["h2"], col["h3"]
    dx1, dx1i = col["dx1"], col["dx1i"]
    lam = lambda_e(T, col["ne"], col["n1"], col["n2"], col["n3"]) * kap[:, None]
    Cc = ((gam - 1.0) / kB / np.maximum(col["ne"], MINDENSDIV)) / ((h1 * h2) * h3)
    beta = (((5.0 / 2.0) * kB) / e) * col["J"]
    B = (Cc * beta) / h1
    D = ((lam * h2) * h3) / h1
    Dh = np.zeros_like(D)
    Dh[:, 1:] = 0.5 * (D[:, :-1] + D[:, 1:])          # Dh(i), i = 2..N at index i-1
    M1 = np.zeros((nb, N, N), dtype=complex)
    M2 = np.zeros((nb, N, N), dtype=complex)
    r1 = np.zeros((nb, N), dtype=complex)
    r2T = np.zeros((nb, N), dtype=complex)
    M1[:, 0, 0] = M2[:, 0, 0] = 1.0
    r1[:, 0] = r2T[:, 0] = Tmin
    for i in range(2, N):                              # rows 2..N-1, index i-1
        k = i - 1
        s1 = ((Cc[k] * Dh[:, k]) / dx1i[k]) / dx1[i]
        s3 = ((Cc[k] * Dh[:, k + 1]) / dx1i[k]) / dx1[i + 1]
        s2 = B[k] / (dx1[i + 1] + dx1[i])
        t1, t2, t3 = s1 / 2.0, s2 / 2.0, s3 / 2.0
        a, b, c = -t1 + t2, (1.0 / (dt / 2.0) + t3) + t1, -t3 - t2
        M1[:, k, k - 1], M1[:, k, k], M1[:, k, k + 1] = a, b, c
        r1[:, k] = ((T[:, k] / (dt / 2.0) - a * T[:, k - 1]) - (t3 + t1) * T[:, k]) - c * T[:, k + 1]
        M2[:, k, k - 1] = -s1 + s2
        M2[:, k, k] = (1.0 / (dt / 3.0) + s3) + s1
        M2[:, k, k + 1] = -s3 - s2
        r2T[:, k] = -(((1.0 / 3.0) * T[:, k]) / (dt / 3.0))
    if col["flux"]:
        gtop = (q * h1[N - 1]) / lam[:, N - 1]
        for M in (M1, M2):
            M[:, N - 1, N - 2] = -1.0 / dx1[N]
            M[:, N - 1, N - 1] = 1.0 / dx1[N]
        r1[:, N - 1] = r2T[:, N - 1] = gtop
    else:
        M1[:, N - 1, N - 1] = M2[:, N - 1, N - 1] = 1.0
        r1[:, N - 1] = r2T[:, N - 1] = Tmax_dirichlet
    return M1, r1, M2, r2T


def trbdf2_column_adjoint(T, kap, q, col, dt, Tmin, Tmax, lam_out):
    """Adjoint of Te'' = max(TRBDF2(T), 100) for one column: (dL/dT, dL/dkappa, dL/dq)."""
    N = T.size
    # inputs p = (T(1..N), kappa, q), complex-step perturbed one per batch entry
    scale = np.concatenate([np.abs(T), [abs(kap), max(abs(q), 1e-3)]])
    hs = H_CS * np.maximum(scale, 1.0e-300)
    nb = N + 3                                   # entry 0: unperturbed
    Tb = np.tile(T.astype(complex), (nb, 1))
    kb = np.full(nb, kap, dtype=complex)
    qb = np.full(nb, q, dtype=complex)
    for p in range(N):
        Tb[p + 1, p] += 1j * hs[p]
    kb[N + 1] += 1j * hs[N]
    qb[N + 2] += 1j * hs[N + 1]
    M1, r1, M2, r2T = assemble(Tb, kb, qb, col, dt, Tmin, Tmax)
    M1r, r1r, M2r, r2Tr = M1[0].real, r1[0].real, M2[0].real, r2T[0].real
    x1 = np.linalg.solve(M1r, r1r)
    c43 = (4.0 / 3.0) / (dt / 3.0)
    dr2dx1 = np.zeros(N)
    dr2dx1[1:N - 1] = c43
    x2 = np.linalg.solve(M2r, r2Tr + dr2dx1 * x1)
    mu = np.where(x2 >= 100.0, lam_out, 0.0)
    w2 = np.linalg.solve(M2r.T, mu)
    w1 = np.linalg.solve(M1r.T, dr2dx1 * w2)
    dM1 = M1[1:].imag / hs[:, None, None]
    dr1 = r1[1:].imag / hs[:, None]
    dM2 = M2[1:].imag / hs[:, None, None]
    dr2 = r2T[1:].imag / hs[:, None]
    grad = (w2[None, :] * (dr2 - dM2 @ x2)).sum(1) + (w1[None, :] * (dr1 - dM1 @ x1)).sum(1)
    return grad[:N], grad[N], grad[N + 1]


def diffusion_adjoint(g, tape, lam, dt, kappa, qtop, J1, nn, Tn, Teinf=1500.0):
    """Electron TRBDF2 + floor transposed on every interior column; ghosts pass through."""
    lx1, lx2, lx3 = g.lx1, g.lx2, g.lx3
    gk = np.zeros((lx2, lx3))
    gq = np.zeros((lx2, lx3))
    I = slice(2, lx1 + 2)
    for i3 in range(lx3):
        for i2 in range(lx2):
            c = (I, i2 + 2, i3 + 2)
            col = dict(ne=tape["NE1"][c], n1=nn[:, i2, i3, 0], n2=nn[:, i2, i3, 1],
                       n3=nn[:, i2, i3, 2], J=J1[c], h1=np.asarray(g.h1)[c],
                       h2=np.asarray(g.h2)[c], h3=np.asarray(g.h3)[c],
                       dx1=np.asarray(g.dx1), dx1i=np.asarray(g.dx1i), flux=qtop is not None)
            if g.gridflag == 0:
                Tmin, Tmax = Tn[0, i2, i3], Tn[lx1 - 1, i2, i3]
            elif g.gridflag == 1:
                Tmin, Tmax = Teinf, Tn[lx1 - 1, i2, i3]
            else:
                Tmin, Tmax = Tn[0, i2, i3], Teinf
            kap = 1.0 if kappa is None else kappa[i2, i3]
            q = 0.0 if qtop is None else qtop[i2, i3]
            gT, gk[i2, i3], gq[i2, i3] = trbdf2_column_adjoint(
                tape["TD"][c], kap, q, col, dt, Tmin, Tmax, lam[c])
            lam[c] = gT
    return lam, gk, gq


# ---------------------------------------------------------------- one step, a run
def adjoint_step(g, tape, lam, dt, kappa, qtop, J1, nn, Tn, Teinf=1500.0):
    """dL/dTe after the step -> dL/dTe before it; returns (lam, dL/dkappa, dL/dq)."""
    kB, _, gam = _consts()
    lam = np.array(lam, copy=True, order="F")
    lam, gk, gq = diffusion_adjoint(g, tape, lam, dt, kappa, qtop, J1, nn, Tn, Teinf)
    lam = np.where(tape["TD"] > 100.0,
                   lam * (((gam - 1.0) / kB) / np.maximum(tape["NE1"], MINDENSDIV)), 0.0)
    lam = np.asfortranarray(lam)
    lam = sweep_adjoint(g, 2, tape["ES2"], tape["V2"], lam, dt)
    if g.flagper:
        lam = halo_adjoint(g, lam)
    lam = sweep_adjoint(g, 1, tape["ES1"], tape["V1"], lam, dt)
    lam = sweep_adjoint(g, 3, tape["ES3"], tape["V3"], lam, dt)
    if g.flagper:
        lam = halo_adjoint(g, lam)
    lam = fill_adjoint(g, lam, tape["NE0"])
    lam = lam * ((tape["NE0"] * kB) / (gam - 1.0))
    return np.asfortranarray(lam), gk, gq


def gradient(g, st0, dt, nsteps, kappa, qtop, lam_final, Teinf=1500.0):
    """dL/dkappa, dL/dqtop (lx2, lx3) and dL/dTe0 for L with dL/dTe(final) = lam_final."""
    st = st0.copy()
    tapes = [record_step(g, st, dt, kappa, qtop, Teinf) for _ in range(nsteps)]
    lam = np.asfortranarray(lam_final, dtype=np.float64)
    gk = np.zeros((g.lx2, g.lx3))
    gq = np.zeros((g.lx2, g.lx3))
    for t in reversed(tapes):
        lam, a, b = adjoint_step(g, t, lam, dt, kappa, qtop, st0.J1, st0.nn, st0.Tn, Teinf)
        gk += a
        gq += b
    return gk, gq, lam, st


def final_Te(g, st0, dt, nsteps, kappa, qtop, Teinf=1500.0):
    """Electron temperature (ghosted field) after nsteps oracle steps."""
    st = st0.copy()
    for _ in range(nsteps):
        transport_step(g, st, dt, Teinf=Teinf, kappa=kappa, qtop=qtop)
    return st.fluidvars[..., 28 + LSP - 1].copy(order="F")
