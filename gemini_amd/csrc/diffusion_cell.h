// diffusion_cell.h -- per-cell inputs, thermal conductivity and parabolic coefficients of
// the field-aligned energy diffusion, shared by the forward solver and its adjoint.
//
// Reference: src/collisions/collisions.f90:162-199 (thermal_conduct) and
// src/numerical/diffusion/diffusion.f90:44-49 (diffusion_prep coefficients), in the
// reference's expression order.
#pragma once

#include "gemhip_internal.h"

namespace gemhip {

// inputs of one cell (neutral densities and J1 for electrons only; kap is the column's
// conductivity scale when the fitting extension is active)
template <bool ELEC>
struct CellIn {
  double T, n, h1, h2, h3;
  double n1, n2, n3, J;
  double kap;
};

// (the metric factors come from DevGrid::hcol at index hc, see hcol_index; gcol: the
// column, lx2 (i3-1) + i2-1, for the conductivity scale)
template <bool ELEC, bool EXT = false>
__device__ __forceinline__ void cell_load(const DevGrid& g, const DevState& s, long so, long cb,
                                          long kb, long ni, int r, CellIn<ELEC>& c, long hc,
                                          int gcol = 0) {
  c.T = s.Ts[so + cb + r];
  c.n = s.ns[so + cb + r];
  c.h1 = g.hcol[hc]; c.h2 = g.hcol[g.hstride + hc]; c.h3 = g.hcol[2 * g.hstride + hc];
  if (ELEC) {
    c.n1 = s.nn[kb + r - 1]; c.n2 = s.nn[ni + kb + r - 1]; c.n3 = s.nn[2 * ni + kb + r - 1];
    c.J = s.J1[cb + r];
    if (EXT) c.kap = s.kappa ? s.kappa[gcol] : 1.0;
  }
}

// The diffusion's column groups: 64 columns each, one lane per column.  Neighbouring
// columns start lx1 + 4 cells apart, so with lx1 a multiple of 8 (and lx2 even, which
// keeps the parity across x3 rows) the rows of even and of odd columns lie 4 cells (32
// bytes) apart in their 64-byte sectors.  Then the two groups of each full block of 128
// columns take its even and its odd columns, and all columns of a group share one sector
// phase.  Otherwise (and in the last, partial block) group cg holds columns 64 cg .. 64 cg
// + 63, lanes past the last column repeating it.
__device__ __forceinline__ bool group_interleaved(const DevGrid& g, long cg) {
  return g.lx1 % 8 == 0 && g.lx2 % 2 == 0 && ((cg >> 1) + 1) * 128 <= (long)g.lx2 * g.lx3;
}
__device__ __forceinline__ int group_column(const DevGrid& g, long cg, int lane) {
  if (group_interleaved(g, cg)) return (int)((cg >> 1) * 128 + 2 * lane + (cg & 1));
  return (int)min(cg * 64 + lane, (long)g.lx2 * g.lx3 - 1);
}

// index in DevGrid::hcol of row r (1..lx1) of column gcol: group-major, then row, then lane
__device__ __forceinline__ long hcol_index(const DevGrid& g, int gcol, int r) {
  const long cg2 = (long)(gcol >> 7) * 2;
  const bool il = group_interleaved(g, cg2);
  const long cg = il ? cg2 + (gcol & 1) : gcol >> 6;
  const int lane = il ? (gcol & 127) >> 1 : gcol & 63;
  return (cg * g.lx1 + (r - 1)) * 64 + lane;
}

// conductivity (collisions.f90:162-199), times the column scale under the extension
template <bool ELEC, bool EXT = false>
__device__ __forceinline__ double cell_lambda(const CellIn<ELEC>& c, int isp) {
  const double T = c.T;
  if (!ELEC) return thermal_coeff(isp) * pow25(T);
  const double sq = sqrt(T);
  const double br = ((c.n1 * 1.1e-16) * (1.0 + 5.7e-4 * T) + ((c.n2 * 2.82e-17) * sq) * (1.0 - 1.21e-4 * T))
                    + (c.n3 * 2.2e-16) * (1.0 + 3.6e-2 * sq);
  const double den = 1.0 + ((3.22e4 * (T * T)) / c.n) * br;
  const double lambda = (((elchrg * 100.0) * 7.7e5) * pow25(T)) / den;
  return EXT ? lambda * c.kap : lambda;
}

// parabolic coefficients (diffusion.f90:44-49)
template <bool ELEC, bool EXT = false>
__device__ __forceinline__ void cell_coeffs(const CellIn<ELEC>& c, int isp, double gm1kB,
                                            double& Cr, double& Br, double& Dr) {
  const double lambda = cell_lambda<ELEC, EXT>(c, isp);
  const double beta = ELEC ? (((5.0 / 2.0) * kB) / elchrg) * c.J : 0.0;
  Cr = (gm1kB / fmax_f(c.n, mindensdiv)) / ((c.h1 * c.h2) * c.h3);
  Br = (Cr * beta) / c.h1;
  Dr = ((lambda * c.h2) * c.h3) / c.h1;
}

}  // namespace gemhip
