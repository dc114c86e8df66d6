"""Grid input contract for the transport path and seeded synthetic grid generators.

The transport kernels read the same geometry the reference's ``curvmesh`` object
carries (src/numerical/coord/meshobj.f90:29-208):

* cell-centred metric factors ``h1, h2, h3`` with two ghost cells,
  shape ``(lx1+4, lx2+4, lx3+4)`` (meshobj.f90:464),
* interface metric factors ``hNx1i`` ``(lx1+1, lx2, lx3)``, ``hNx2i``
  ``(lx1, lx2+1, lx3)``, ``hNx3i`` ``(lx1, lx2, lx3+1)`` (meshobj.f90:467-473),
* coordinate differences ``dxN(0:lxN+2)`` and ``dxNi(1:lxN)`` computed exactly as
  ``calc_coord_diffs`` does (meshobj.f90:356-389),
* ``gridflag`` (meshobj.f90:549-570), ``flagper`` and the ``isglobal*`` flags
  (grid.f90:448-507).

All 3-D arrays are Fortran-ordered float64 (x1 fastest) so their memory matches
the reference's and can be handed to the C ABI unchanged.

Two generators build the synthetic grids the benchmark configurations name:
a Cartesian altitude grid (all metric factors 1, as meshobj_cart.f90:127-146)
and a closed dipole grid whose metric factors follow the analytic dipole forms
of meshobj_dipole.f90:323-360 (hq = r^3/Re^2/sqrt(1+3cos^2 th),
hp = Re sin^3 th/sqrt(1+3cos^2 th), hphi = r sin th).  The real reference reads
its grid from HDF5 files; those readers are outside this path.
"""
from __future__ import annotations

import numpy as np

RE = 6371.0e3  # phys_consts.in.f90:34


PLANE_FIELDS = ("h1", "h2", "h3", "h1x1i", "h2x1i", "h3x1i", "h1x2i", "h2x2i", "h3x2i",
                "h1x3i", "h2x3i", "h3x3i", "alt", "g1")
GM_EARTH = 6.67408e-11 * 5.9722e24          # Gconst*Me, phys_consts.in.f90:27,36


class Grid:
    """Mesh contract of the transport path.

    Every 3-D array of the synthetic meshes is invariant along x3, so a grid keeps
    each as an (n2, n1) C-order plane plus its x3 extent and materialises the
    Fortran-ordered (n1, n2, n3) array on first access (``grid.h1``), or on the GPU
    with :meth:`device_array`.
    """

    def __init__(self, lx1, lx2, lx3, x1, x2, x3, planes, gridflag=2, flagper=False,
                 isglobal=(True,) * 6, kind="cartesian"):
        self.lx1, self.lx2, self.lx3 = int(lx1), int(lx2), int(lx3)
        self.x1, self.x2, self.x3 = (np.asarray(v, dtype=np.float64) for v in (x1, x2, x3))
        self.planes = planes                  # name -> (plane (n2, n1), n3)
        self.gridflag = int(gridflag)
        self.flagper = bool(flagper)
        self.isglobal = tuple(bool(v) for v in isglobal)
        self.kind = kind
        self._full = {}
        self.dx1, self.dx1i = coord_diffs(self.x1, self.lx1)
        self.dx2, self.dx2i = coord_diffs(self.x2, self.lx2)
        self.dx3, self.dx3i = coord_diffs(self.x3, self.lx3)

    def __getattr__(self, name):
        if name in PLANE_FIELDS:
            full = self.__dict__["_full"]
            if name not in full:
                plane, n3 = self.__dict__["planes"][name]
                full[name] = _bcast_planes(plane, n3)
            return full[name]
        raise AttributeError(name)

    def device_array(self, name, device):
        """torch tensor (n3, n2, n1) on ``device`` whose memory is the Fortran array."""
        import torch
        plane, n3 = self.planes[name]
        p = torch.as_tensor(np.ascontiguousarray(plane), device=device)
        return p.unsqueeze(0).expand(n3, *p.shape).contiguous()

    @property
    def shape(self):
        return (self.lx1, self.lx2, self.lx3)

    @property
    def padded(self):
        return (self.lx1 + 4, self.lx2 + 4, self.lx3 + 4)

    @property
    def ncells(self):
        return self.lx1 * self.lx2 * self.lx3

    def alt_plane(self):
        """(L2, L1) altitude plane (altitudes do not depend on x3)."""
        return np.ascontiguousarray(self.planes["alt"][0])


def coord_diffs(x, lx):
    """dx(0:lx+2) and dxi(1:lx) from ghosted centres x(-1:lx+2) (meshobj.f90:371-386)."""
    x = np.asarray(x, dtype=np.float64)
    dx = x[1:lx + 4] - x[0:lx + 3]                 # x(0:lx+2) - x(-1:lx+1)
    xi = 0.5 * (x[1:lx + 2] + x[2:lx + 3])         # x1i(1:lx+1) = 0.5*(x(0:lx)+x(1:lx+1))
    dxi = xi[1:lx + 1] - xi[0:lx]                  # x1i(2:lx+1) - x1i(1:lx)
    return np.ascontiguousarray(dx), np.ascontiguousarray(dxi)


def isglobal_flags(x1, x2, x3, lx1, lx2, lx3, lims1, lims2, lims3):
    """grid.f90:448-507 for a (sub)grid with ghosted centres x1,x2,x3 (index 0 is -1)."""
    def lo(x, lims):
        return abs(x[2] - lims[0]) < abs(x[3] - x[2])
    def hi(x, lx, lims):
        return abs(x[lx + 1] - lims[1]) < abs(x[lx + 1] - x[lx])
    return (lo(x1, lims1), hi(x1, lx1, lims1), lo(x2, lims2), hi(x2, lx2, lims2),
            lo(x3, lims3), hi(x3, lx3, lims3))


def _ghosted_uniform(lx, dx, start=0.0):
    return start + dx * (np.arange(-1, lx + 3, dtype=np.float64) - 1)


def _bcast_planes(plane_c, n3):
    """(n2, n1) C-order plane repeated n3 times -> Fortran-contiguous (n1, n2, n3) array."""
    out = np.empty((n3,) + plane_c.shape, dtype=np.float64)
    out[:] = plane_c[None, :, :]
    return out.T


def _tanh_stretched_alt(lx1, zmin=80e3, zmax=1000e3, dzmin=1e3, dzmax=40e3):
    """Cell centres from zmin to zmax whose spacing rises from dzmin to dzmax (tanh)."""
    n = lx1 - 1
    i = np.arange(n, dtype=np.float64)
    w = max(n / 8.0, 1.0)

    def spacing(i0):
        return dzmin + (dzmax - dzmin) * 0.5 * (1.0 + np.tanh((i - i0) / w))

    lo_, hi_ = -4.0 * n, 5.0 * n
    target = zmax - zmin
    if spacing(lo_).sum() < target:      # even the widest spacing is too short: scale
        dz = spacing(lo_) * (target / spacing(lo_).sum())
    elif spacing(hi_).sum() > target:
        dz = spacing(hi_) * (target / spacing(hi_).sum())
    else:
        for _ in range(200):              # bisection on the transition index
            mid = 0.5 * (lo_ + hi_)
            if spacing(mid).sum() > target:
                lo_ = mid
            else:
                hi_ = mid
        dz = spacing(0.5 * (lo_ + hi_))
        dz *= target / dz.sum()
    z = np.empty(lx1 + 4)
    z[2] = zmin
    z[3:lx1 + 2] = zmin + np.cumsum(dz)
    z[1] = z[2] - dz[0]
    z[0] = z[1] - dz[0]
    z[lx1 + 2] = z[lx1 + 1] + dz[-1]
    z[lx1 + 3] = z[lx1 + 2] + dz[-1]
    return z


def cartesian_grid(lx1, lx2, lx3, dx2=5e3, dx3=5e3, zmin=80e3, zmax=1000e3,
                   flagper=False):
    """Cartesian altitude grid, gridflag 2 (bottom first), all metric factors 1."""
    x1 = _tanh_stretched_alt(lx1, zmin, zmax)
    x2 = _ghosted_uniform(lx2, dx2)
    x3 = _ghosted_uniform(lx3, dx3)
    return cartesian_mesh(x1, x2, x3, flagper)


def cartesian_mesh(x1, x2, x3, flagper=False):
    """The Cartesian mesh on ghosted coordinates x1 (altitude, m; lx1+4 values), x2 and x3
    (m): metric factors 1, altitude x1, gravity -G Me / (Re + x1)^2 along x1."""
    x1, x2, x3 = (np.asarray(x, dtype=np.float64) for x in (x1, x2, x3))
    lx1, lx2, lx3 = x1.size - 4, x2.size - 4, x3.size - 4
    L1, L2, L3 = lx1 + 4, lx2 + 4, lx3 + 4
    ones = lambda n2, n1, n3: (np.ones((n2, n1)), n3)
    planes = {"h1": ones(L2, L1, L3), "h2": ones(L2, L1, L3), "h3": ones(L2, L1, L3),
              "alt": (np.ascontiguousarray(np.broadcast_to(x1[None, :], (L2, L1))), L3)}
    # gravity along x1 (up), interior cells: -G Me / r^2
    g1 = -GM_EARTH / (RE + x1[2:lx1 + 2]) ** 2
    planes["g1"] = (np.ascontiguousarray(np.broadcast_to(g1[None, :], (lx2, lx1))), lx3)
    for n in ("h1x1i", "h2x1i", "h3x1i"):
        planes[n] = ones(lx2, lx1 + 1, lx3)
    for n in ("h1x2i", "h2x2i", "h3x2i"):
        planes[n] = ones(lx2 + 1, lx1, lx3)
    for n in ("h1x3i", "h2x3i", "h3x3i"):
        planes[n] = ones(lx2, lx1, lx3 + 1)
    lims = lambda x, lx: (x[2], x[lx + 1])           # grid.f90:144-146
    flags = isglobal_flags(x1, x2, x3, lx1, lx2, lx3, lims(x1, lx1), lims(x2, lx2), lims(x3, lx3))
    alt = planes["alt"][0]
    gridflag = _gridflag(alt[2, 2], alt[2, 3], alt[2, lx1 + 1])
    return Grid(lx1, lx2, lx3, x1, x2, x3, planes, gridflag=gridflag, flagper=flagper,
                isglobal=flags, kind="cartesian")


def _gridflag(alt1, alt2, altN):
    """meshobj.f90:563-569 from alt(1,1,1), alt(2,1,1), alt(lx1,1,1)."""
    if abs(alt1 - altN) < 100e3:
        return 0
    if alt1 > alt2:
        return 1
    return 2


def _dipole_r_theta(q, p):
    """Solve q = cos(th)/u^2, p = u/sin^2(th) (u = r/Re) for r and theta."""
    q = np.asarray(q, dtype=np.float64)
    p = np.asarray(p, dtype=np.float64)
    q, p = np.broadcast_arrays(q, p)
    lo = np.zeros_like(p)
    hi = p.copy()
    for _ in range(80):                   # f(u) = q^2 u^4 + u/p - 1 is increasing on (0, p]
        mid = 0.5 * (lo + hi)
        fm = q * q * mid ** 4 + mid / p - 1.0
        hi = np.where(fm > 0, mid, hi)
        lo = np.where(fm > 0, lo, mid)
    u = 0.5 * (lo + hi)
    for _ in range(3):                    # Newton polish
        fu = q * q * u ** 4 + u / p - 1.0
        u = u - fu / (4 * q * q * u ** 3 + 1.0 / p)
    costh = np.clip(q * u * u, -1.0, 1.0)
    return u * RE, np.arccos(costh)


def _h_dipole(r, th):
    c = np.cos(th)
    s = np.sin(th)
    den = np.sqrt(1 + 3 * c ** 2)
    return r ** 3 / RE ** 2 / den, RE * s ** 3 / den, r * s


def dipole_grid(lx1, lx2, lx3, Lmin=1.2, Lmax=5.0, altfoot=80e3, phispan=np.radians(20.0)):
    """Closed dipole grid (gridflag 0): x1 = q, x2 = p (L shell), x3 = phi.

    The q range is the one the innermost shell spans between its 80 km footpoints,
    so every cell lies above the footpoint altitude and no null points occur.
    """
    x1, x2, x3 = _dipole_coords(lx1, lx2, lx3, Lmin, Lmax, altfoot, phispan)
    return dipole_mesh(x1, x2, x3)


def _dipole_coords(lx1, lx2, lx3, Lmin, Lmax, altfoot, phispan):
    uf = (RE + altfoot) / RE
    qmax = np.sqrt(1.0 - uf / Lmin) / uf ** 2
    # q centres: sinh stretch to refine near the footpoints, ghosts by edge spacing
    s = np.linspace(-1.0, 1.0, lx1)
    qc = qmax * np.sinh(2.0 * s) / np.sinh(2.0)
    x1 = np.empty(lx1 + 4)
    x1[2:lx1 + 2] = qc
    d0, d1 = qc[1] - qc[0], qc[-1] - qc[-2]
    x1[1], x1[0] = qc[0] - d0, qc[0] - 2 * d0
    x1[lx1 + 2], x1[lx1 + 3] = qc[-1] + d1, qc[-1] + 2 * d1
    dp = (Lmax - Lmin) / max(lx2 - 1, 1)
    x2 = Lmin + dp * (np.arange(-1, lx2 + 3, dtype=np.float64) - 1)
    dphi = phispan / max(lx3, 1)
    x3 = dphi * (np.arange(-1, lx3 + 3, dtype=np.float64) - 1)
    return x1, x2, x3


def dipole_mesh(x1, x2, x3):
    """The dipole mesh on ghosted coordinates x1 = q, x2 = p (L shell), x3 = phi: radius
    and colatitude from (q, p), the metric factors, altitude and gravity along x1."""
    x1, x2, x3 = (np.asarray(x, dtype=np.float64) for x in (x1, x2, x3))
    lx1, lx2, lx3 = x1.size - 4, x2.size - 4, x3.size - 4
    L1, L2, L3 = lx1 + 4, lx2 + 4, lx3 + 4
    # cell centres (with ghosts); metric factors do not depend on phi: (x2, x1) planes
    r, th = _dipole_r_theta(x1[None, :], x2[:, None])
    hq, hp, hphi = _h_dipole(r, th)
    planes = {"h1": (hq, L3), "h2": (hp, L3), "h3": (hphi, L3), "alt": (r - RE, L3)}
    # gravity along x1, interior cells (meshobj_dipole.f90:265-282): -G Me / r^2 times
    # er . e1, with e1 the unit vector along grad q = -(2 cos th er + sin th eth)/(r^3 ...)
    ri_, thi_ = r[2:lx2 + 2, 2:lx1 + 2], th[2:lx2 + 2, 2:lx1 + 2]
    c_ = np.cos(thi_)
    planes["g1"] = (np.ascontiguousarray((-GM_EARTH / ri_ ** 2) * (-2.0 * c_ / np.sqrt(1 + 3 * c_ ** 2))), lx3)
    # q interfaces x1i(1:lx1+1) at interior p
    x1i = 0.5 * (x1[1:lx1 + 2] + x1[2:lx1 + 3])
    ri, thi = _dipole_r_theta(x1i[None, :], x2[2:lx2 + 2, None])
    for n, a in zip(("h1x1i", "h2x1i", "h3x1i"), _h_dipole(ri, thi)):
        planes[n] = (a, lx3)
    # p interfaces x2i(1:lx2+1) at interior q
    x2i = 0.5 * (x2[1:lx2 + 2] + x2[2:lx2 + 3])
    rp, thp = _dipole_r_theta(x1[None, 2:lx1 + 2], x2i[:, None])
    for n, a in zip(("h1x2i", "h2x2i", "h3x2i"), _h_dipole(rp, thp)):
        planes[n] = (a, lx3)
    # phi interfaces: copies of the cell-centred values (meshobj_dipole.f90:181-186)
    for n, a in zip(("h1x3i", "h2x3i", "h3x3i"), (hq, hp, hphi)):
        planes[n] = (np.ascontiguousarray(a[2:lx2 + 2, 2:lx1 + 2]), lx3 + 1)
    lims = lambda x, lx: (x[2], x[lx + 1])           # grid.f90:144-146
    flags = isglobal_flags(x1, x2, x3, lx1, lx2, lx3, lims(x1, lx1), lims(x2, lx2), lims(x3, lx3))
    alt = planes["alt"][0]
    gridflag = _gridflag(alt[2, 2], alt[2, 3], alt[2, lx1 + 1])
    return Grid(lx1, lx2, lx3, x1, x2, x3, planes, gridflag=gridflag, flagper=False,
                isglobal=flags, kind="dipole")
