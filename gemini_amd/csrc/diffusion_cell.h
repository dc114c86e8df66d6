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

// index in DevGrid::hcol of row r (1..lx1) of column gcol
__device__ __forceinline__ long hcol_index(const DevGrid& g, int gcol, int r) {
  return ((long)(gcol >> 6) * g.lx1 + (r - 1)) * 64 + (gcol & 63);
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
