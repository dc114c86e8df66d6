// kernels_sources.hip -- the source/loss substep of a time step (src/libgemini.cpp:223-232):
// source_loss_energy, source_loss_momentum and source_loss_mass (multifluid.f90:130-175,
// 472-549) with their collisional and chemical rates (sources.f90, collisions.f90:58-159)
// and the exponential time differencing of calculus.in.f90:174-196.
//
// Each substep is one pass over the padded grid, one thread per cell.  The interior cells
// form their production and loss rates in registers and advance the conserved quantity;
// every cell (ghosts included, as the reference's whole-array statements do) then converts
// back to temperature / velocity / electron density.  The rates never touch HBM.
//
// Arithmetic follows the reference's expressions (-ffp-contract=off), with liberties that
// cost a few ulp per operation: a divide by a compile-time constant (a mass sum, 3) is a
// multiply by its correctly rounded reciprocal; within a Coulomb pair 1/Teff is formed once
// and gives Teff**-1.5 and W**2, and exp(-W**2) below W = 0.1 is a Taylor series.  The
// transcendental functions (exp, log, log10, erf, tanh, sin, and x**p for the fitted
// exponents) are the device library's, which agree with glibc to about an ulp; x**0.5 is
// the correctly rounded sqrt.  These passes therefore match the reference to rounding (the
// tests hold them to 1e-12 of each field's magnitude) rather than bit for bit.
//
// Ion-ion (Coulomb) terms: Teff, Wsj, Phisj and Psisj of coulomb_colln are symmetric in
// the species pair, so each unordered pair is evaluated once and feeds both species.  The
// pairs run in lexicographic order, which adds each species' terms in the reference's
// isp2 order; self-collision terms are exact zeros and are skipped.  The pair loop is
// rolled in the energy kernel (one copy of exp and erf instead of 21, per-species values
// selected by a wave-uniform index), which keeps it within 256 registers.
#include "gemhip_internal.h"

namespace gemhip {
namespace {

constexpr int TPB = 256;
constexpr double PI = 3.141592653589793;   // 4*atan(1.0_wp), phys_consts.in.f90:20

// collisions.f90:11-47: default-real literals widened to real64
__device__ __forceinline__ double tab_csn(int isp, int isp2) {
  constexpr float t[7][4] = {
      {-1, 6.82e-10f, 6.64e-10f, -1},        {2.44e-10f, 4.34e-10f, 4.27e-10f, 0.69e-10f},
      {2.58e-10f, -1, 4.49e-10f, 0.74e-10f}, {2.31e-10f, 4.13e-10f, -1, 0.65e-10f},
      {4.42e-10f, 7.47e-10f, 7.25e-10f, 1.45e-10f}, {-1, 33.6e-10f, 32.0e-10f, -1},
      {-1, -1, -1, -1}};
  return (double)t[isp - 1][isp2 - 1];
}
__device__ __forceinline__ double tab_c2sn1(int isp, int isp2) {
  constexpr float t[7][4] = {
      {3.67e-11f, 0, 0, 4.63e-12f}, {0, 0, 0, 0}, {0, 5.14e-11f, 0, 0}, {0, 0, 2.59e-11f, 0},
      {0, 0, 0, 0}, {6.61e-11f, 0, 0, 2.65e-10f}, {-1, -1, -1, -1}};
  return (double)t[isp - 1][isp2 - 1];
}
__device__ __forceinline__ double tab_c2sn2(int isp, int isp2) {
  constexpr float t[7][4] = {
      {0.064f, 0, 0, -1}, {0, 0, 0, 0}, {0, 0.069f, 0, 0}, {0, 0, 0.073f, 0},
      {0, 0, 0, 0}, {0.047f, 0, 0, 0.083f}, {-1, -1, -1, -1}};
  return (double)t[isp - 1][isp2 - 1];
}
// phys_consts.in.f90:60: mn = [16,28,32,1]*amu
__device__ __forceinline__ double mn(int isp2) {
  return (isp2 == 1 ? 16.0 : isp2 == 2 ? 28.0 : isp2 == 3 ? 32.0 : 1.0) * amu;
}

// calculus.in.f90:174-196
__device__ __forceinline__ double etd(double f, double P, double L, double dt) {
  const double Ldt = L * dt;
  const double expL = exp(-1.0 * Ldt);
  if (Ldt > 1e-10) return f * expL + (P / L) * (1.0 - expL);
  return f + P * dt;
}

// collisions.f90:58-114 at one cell (nn2 = nn(isp2), Tsp = Ts(isp))
__device__ __forceinline__ double maxwell(int isp, int isp2, double nn2, double Tn, double Tsp) {
  if (isp < LSP) {
    const double csn = tab_csn(isp, isp2);
    if (csn < 0) {
      const double c1 = tab_c2sn1(isp, isp2);
      if (isp == 1 && isp2 == 4) {
        const double Teff = Tn + Tsp / 16.0;
        return ((c1 * sqrt(Teff)) * nn2) * 1e-6;
      }
      const double Teff = 0.5 * (Tn + Tsp);
      const double b = 1.0 - tab_c2sn2(isp, isp2) * log10(Teff);
      return (((c1 * (b * b)) * sqrt(Teff)) * nn2) * 1e-6;
    }
    return (csn * nn2) * 1e-6;
  }
  const double T = Tsp;
  switch (isp2) {
    case 1: return (((8.9e-11 * (1.0 + 5.7e-4 * T)) * sqrt(T)) * nn2) * 1e-6;
    case 2: return (((2.33e-11 * (1.0 - 1.21e-4 * T)) * T) * nn2) * 1e-6;
    case 3: return (((1.82e-10 * (1.0 + 3.6e-2 * sqrt(T))) * sqrt(T)) * nn2) * 1e-6;
    default: return (((4.5e-9 * (1.0 - 1.35e-4 * T)) * sqrt(T)) * nn2) * 1e-6;
  }
}

// Tables of the rolled pair loop.  The pair index is the same on every lane of a wave, so
// these are scalar loads; each entry is the constant expression the unrolled form would
// fold at compile time.  Pairs run (1,2), (1,3), .., (1,7), (2,3), .., (6,7).
__constant__ int c_pair_a[21] = {1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 5, 5, 6};
__constant__ int c_pair_b[21] = {2, 3, 4, 5, 6, 7, 3, 4, 5, 6, 7, 4, 5, 6, 7, 5, 6, 7, 6, 7, 7};
__constant__ double c_ms[7] = {
    16.0 * amu, 30.0 * amu, 28.0 * amu, 32.0 * amu,
    14.0 * amu, 1.0 * amu, 5.485799090e-4 * amu};
__constant__ double c_ms3[7] = {
    (16.0 * amu) / 3.0, (30.0 * amu) / 3.0, (28.0 * amu) / 3.0,
    (32.0 * amu) / 3.0, (14.0 * amu) / 3.0, (1.0 * amu) / 3.0,
    (5.485799090e-4 * amu) / 3.0};
__constant__ double c_nkB_igm1[7] = {
    kB * (1.0 / (5.0 / 3.0 - 1.0)), kB * (1.0 / (7.0 / 5.0 - 1.0)),
    kB * (1.0 / (7.0 / 5.0 - 1.0)), kB * (1.0 / (7.0 / 5.0 - 1.0)),
    kB * (1.0 / (5.0 / 3.0 - 1.0)), kB * (1.0 / (5.0 / 3.0 - 1.0)),
    kB * (1.0 / (5.0 / 3.0 - 1.0))};
__constant__ double c_igm1[7] = {
    1.0 / (5.0 / 3.0 - 1.0), 1.0 / (7.0 / 5.0 - 1.0), 1.0 / (7.0 / 5.0 - 1.0),
    1.0 / (7.0 / 5.0 - 1.0), 1.0 / (5.0 / 3.0 - 1.0), 1.0 / (5.0 / 3.0 - 1.0),
    1.0 / (5.0 / 3.0 - 1.0)};
__constant__ double c_two_msum[21] = {
    2.0 / ((16.0 * amu) + (30.0 * amu)), 2.0 / ((16.0 * amu) + (28.0 * amu)),
    2.0 / ((16.0 * amu) + (32.0 * amu)), 2.0 / ((16.0 * amu) + (14.0 * amu)),
    2.0 / ((16.0 * amu) + (1.0 * amu)), 2.0 / ((16.0 * amu) + (5.485799090e-4 * amu)),
    2.0 / ((30.0 * amu) + (28.0 * amu)), 2.0 / ((30.0 * amu) + (32.0 * amu)),
    2.0 / ((30.0 * amu) + (14.0 * amu)), 2.0 / ((30.0 * amu) + (1.0 * amu)),
    2.0 / ((30.0 * amu) + (5.485799090e-4 * amu)), 2.0 / ((28.0 * amu) + (32.0 * amu)),
    2.0 / ((28.0 * amu) + (14.0 * amu)), 2.0 / ((28.0 * amu) + (1.0 * amu)),
    2.0 / ((28.0 * amu) + (5.485799090e-4 * amu)), 2.0 / ((32.0 * amu) + (14.0 * amu)),
    2.0 / ((32.0 * amu) + (1.0 * amu)), 2.0 / ((32.0 * amu) + (5.485799090e-4 * amu)),
    2.0 / ((14.0 * amu) + (1.0 * amu)), 2.0 / ((14.0 * amu) + (5.485799090e-4 * amu)),
    2.0 / ((1.0 * amu) + (5.485799090e-4 * amu))};
__constant__ double c_inv_msum[21] = {
    1.0 / ((30.0 * amu) + (16.0 * amu)), 1.0 / ((28.0 * amu) + (16.0 * amu)),
    1.0 / ((32.0 * amu) + (16.0 * amu)), 1.0 / ((14.0 * amu) + (16.0 * amu)),
    1.0 / ((1.0 * amu) + (16.0 * amu)), 1.0 / ((5.485799090e-4 * amu) + (16.0 * amu)),
    1.0 / ((28.0 * amu) + (30.0 * amu)), 1.0 / ((32.0 * amu) + (30.0 * amu)),
    1.0 / ((14.0 * amu) + (30.0 * amu)), 1.0 / ((1.0 * amu) + (30.0 * amu)),
    1.0 / ((5.485799090e-4 * amu) + (30.0 * amu)), 1.0 / ((32.0 * amu) + (28.0 * amu)),
    1.0 / ((14.0 * amu) + (28.0 * amu)), 1.0 / ((1.0 * amu) + (28.0 * amu)),
    1.0 / ((5.485799090e-4 * amu) + (28.0 * amu)), 1.0 / ((14.0 * amu) + (32.0 * amu)),
    1.0 / ((1.0 * amu) + (32.0 * amu)), 1.0 / ((5.485799090e-4 * amu) + (32.0 * amu)),
    1.0 / ((1.0 * amu) + (14.0 * amu)), 1.0 / ((5.485799090e-4 * amu) + (14.0 * amu)),
    1.0 / ((5.485799090e-4 * amu) + (1.0 * amu))};
__constant__ double c_csj_1e6[7][7] = {
    {(double)0.22f * 1e-6, (double)0.26f * 1e-6, (double)0.25f * 1e-6, (double)0.26f * 1e-6, (double)0.22f * 1e-6, (double)0.077f * 1e-6, (double)1.87e-3f * 1e-6},
    {(double)0.14f * 1e-6, (double)0.16f * 1e-6, (double)0.16f * 1e-6, (double)0.17f * 1e-6, (double)0.13f * 1e-6, (double)0.042f * 1e-6, (double)9.97e-4f * 1e-6},
    {(double)0.15f * 1e-6, (double)0.17f * 1e-6, (double)0.17f * 1e-6, (double)0.18f * 1e-6, (double)0.14f * 1e-6, (double)0.045f * 1e-6, (double)1.07e-3f * 1e-6},
    {(double)0.13f * 1e-6, (double)0.16f * 1e-6, (double)0.15f * 1e-6, (double)0.16f * 1e-6, (double)0.12f * 1e-6, (double)0.039f * 1e-6, (double)9.347e-4f * 1e-6},
    {(double)0.25f * 1e-6, (double)0.28f * 1e-6, (double)0.28f * 1e-6, (double)0.28f * 1e-6, (double)0.24f * 1e-6, (double)0.088f * 1e-6, (double)2.136e-3f * 1e-6},
    {(double)1.23f * 1e-6, (double)1.25f * 1e-6, (double)1.25f * 1e-6, (double)1.25f * 1e-6, (double)1.23f * 1e-6, (double)0.90f * 1e-6, (double)29.7e-3f * 1e-6},
    {(double)54.5f * 1e-6, (double)54.5f * 1e-6, (double)54.5f * 1e-6, (double)54.5f * 1e-6, (double)54.5f * 1e-6, (double)54.5f * 1e-6, (double)38.537f * 1e-6}};

// mred/(2 kB) of each pair, so that W**2 = (dv**2) mred/(2 kB Teff)
__constant__ double c_mred_2kB[21] = {
#define GEMHIP_MRED_2KB(ma, mb) ((((ma) * amu) * ((mb) * amu)) / (((ma) * amu) + ((mb) * amu))) / (2.0 * kB)
    GEMHIP_MRED_2KB(16.0, 30.0), GEMHIP_MRED_2KB(16.0, 28.0), GEMHIP_MRED_2KB(16.0, 32.0),
    GEMHIP_MRED_2KB(16.0, 14.0), GEMHIP_MRED_2KB(16.0, 1.0), GEMHIP_MRED_2KB(16.0, 5.485799090e-4),
    GEMHIP_MRED_2KB(30.0, 28.0), GEMHIP_MRED_2KB(30.0, 32.0), GEMHIP_MRED_2KB(30.0, 14.0),
    GEMHIP_MRED_2KB(30.0, 1.0), GEMHIP_MRED_2KB(30.0, 5.485799090e-4), GEMHIP_MRED_2KB(28.0, 32.0),
    GEMHIP_MRED_2KB(28.0, 14.0), GEMHIP_MRED_2KB(28.0, 1.0), GEMHIP_MRED_2KB(28.0, 5.485799090e-4),
    GEMHIP_MRED_2KB(32.0, 14.0), GEMHIP_MRED_2KB(32.0, 1.0), GEMHIP_MRED_2KB(32.0, 5.485799090e-4),
    GEMHIP_MRED_2KB(14.0, 1.0), GEMHIP_MRED_2KB(14.0, 5.485799090e-4),
    GEMHIP_MRED_2KB(1.0, 5.485799090e-4)};
#undef GEMHIP_MRED_2KB

// exp(-x) for 0 <= x < 0.01: Taylor series to x**6 (the remainder is below x**7/5040, 2e-18)
__device__ __forceinline__ double exp_neg_small(double x) {
  double r = 1.0 / 720.0;
  r = r * x - 1.0 / 120.0;
  r = r * x + 1.0 / 24.0;
  r = r * x - 1.0 / 6.0;
  r = r * x + 0.5;
  r = r * x - 1.0;
  return r * x + 1.0;
}

// the pair-symmetric part of coulomb_colln (collisions.f90:138-157) for pair pr = (a, b),
// a < b: 1/Teff**1.5, Phisj, Psisj.  1/Teff is formed once: Teff**-1.5 = (1/Teff)
// sqrt(1/Teff) and W**2 = dv**2 (mred/(2 kB)) (1/Teff), each a few ulp from the
// reference's forms.  Below W = 0.1 (W**2 = 0.01), where Phisj = 1, W itself is not needed
// and Psisj = exp(-W**2) comes from its Taylor series.
struct Pair { double iT15, Phi, Psi; };
__device__ __forceinline__ Pair coulomb_pair(int pr, double Ta, double Tb, double va, double vb) {
  const int a = c_pair_a[pr], b = c_pair_b[pr];
  const double ma = c_ms[a - 1], mb = c_ms[b - 1];
  const double Teff = (mb * Ta + ma * Tb) * c_inv_msum[pr];
  const double iTeff = 1.0 / Teff;
  const double dv = va - vb;
  const double W2 = ((dv * dv) * c_mred_2kB[pr]) * iTeff;
  Pair p;
  p.iT15 = iTeff * sqrt(iTeff);
  if (W2 < 0.01) {
    p.Phi = 1.0;
    p.Psi = exp_neg_small(W2);
  } else {
    const double W = sqrt(W2);
    p.Psi = exp(-W2);
    const double iW = 1.0 / W, iW2 = iW * iW;
    p.Phi = (((3.0 / 4.0) * sqrt(PI)) * erf(W)) * (iW2 * iW) - ((3.0 / 2.0) * iW2) * p.Psi;
  }
  return p;
}

struct Cell {
  int i1, i2, i3;
  bool interior;
  long k;          // interior index (i1-1) + lx1*((i2-1) + lx2*(i3-1))
};
__device__ __forceinline__ Cell decode(const DevGrid& g, long c) {
  Cell q;
  q.i1 = (int)(c % g.L1) - 1;
  const long r = c / g.L1;
  q.i2 = (int)(r % g.L2) - 1;
  q.i3 = (int)(r / g.L2) - 1;
  q.interior = q.i1 >= 1 && q.i1 <= g.lx1 && q.i2 >= 1 && q.i2 <= g.lx2 && q.i3 >= 1 && q.i3 <= g.lx3;
  q.k = (long)(q.i1 - 1) + (long)g.lx1 * ((long)(q.i2 - 1) + (long)g.lx2 * (long)(q.i3 - 1));
  return q;
}

// ---- source_loss_energy: srcsEnergy (sources.f90:292-374), Qeionize, ETD, rhoe2T ----
__global__ __launch_bounds__(TPB) void k_src_energy(DevGrid g, DevState s, DevSources a, double dt) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.N) return;
  const Cell q = decode(g, c);
  double n[LSP], e[LSP];
#pragma unroll
  for (int i = 0; i < LSP; ++i) {
    n[i] = s.ns[i * g.N + c];
    e[i] = s.rhoes[i * g.N + c];
  }
  if (q.interior) {
    const long ni = (long)g.lx1 * g.lx2 * g.lx3, k = q.k;
    double T[LSP], v1[LSP], v2[LSP], v3[LSP], P[LSP], L[LSP];
#pragma unroll
    for (int i = 0; i < LSP; ++i) {
      T[i] = s.Ts[i * g.N + c];
      v1[i] = s.vs1[i * g.N + c];
      v2[i] = s.vs2[i * g.N + c];
      v3[i] = s.vs3[i * g.N + c];
    }
    const double Tn = s.Tn[k], vn1 = a.vn1[k], vn2 = a.vn2[k], vn3 = a.vn3[k];
    double nnk[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) nnk[j] = a.nn[j * ni + k];
    // ion-neutral and electron-neutral: heat transfer and friction
#pragma unroll
    for (int isp = 1; isp <= LSP; ++isp) {
      const double m = ms(isp), gm1 = gammas(isp) - 1.0, nm = n[isp - 1] * m;
      const double d1 = v1[isp - 1] - vn1, d2 = v2[isp - 1] - vn2, d3 = v3[isp - 1] - vn3;
      const double dv2 = ((d1 * d1) + (d2 * d2)) + (d3 * d3);
      const double heat = (nm * kB) * (1.0 / gm1), fric = (nm * (1.0 / gm1)) * dv2;
      double Pp = 0.0, Ll = 0.0;
#pragma unroll
      for (int isp2 = 1; isp2 <= 4; ++isp2) {
        const double nu = maxwell(isp, isp2, nnk[isp2 - 1], Tn, T[isp - 1]);
        double fact = nu * (2.0 / (m + mn(isp2)));
        Pp = Pp + (heat * fact) * Tn;
        Ll = Ll + m * fact;
        fact = fact * (mn(isp2) / 3.0);
        Pp = Pp + fric * fact;
      }
      P[isp - 1] = Pp;
      L[isp - 1] = Ll;
    }
    // ion-ion and ion-electron (Coulomb), one unordered pair per iteration
#pragma unroll 1
    for (int pr = 0; pr < 21; ++pr) {
      const int ia = c_pair_a[pr], ib = c_pair_b[pr];
      const Pair p = coulomb_pair(pr, T[ia - 1], T[ib - 1], v1[ia - 1], v1[ib - 1]);
      const double d1 = v1[ia - 1] - v1[ib - 1], d2 = v2[ia - 1] - v2[ib - 1],
                   d3 = v3[ia - 1] - v3[ib - 1];
      const double dv2 = ((d1 * d1) + (d2 * d2)) + (d3 * d3);
      const double psi2 = p.Psi * c_two_msum[pr], phi2 = p.Phi * c_two_msum[pr];
      // species ia with partner ib, then ib with partner ia
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        const int x = side ? ib : ia, y = side ? ia : ib;
        const double nm = n[x - 1] * c_ms[x - 1];
        const double nu = (c_csj_1e6[x - 1][y - 1] * n[y - 1]) * p.iT15;
        double fact = nu * psi2;
        P[x - 1] = P[x - 1] + ((nm * c_nkB_igm1[x - 1]) * fact) * T[y - 1];
        L[x - 1] = L[x - 1] + c_ms[x - 1] * fact;
        fact = (nu * phi2) * c_ms3[y - 1];
        P[x - 1] = P[x - 1] + ((nm * c_igm1[x - 1]) * dv2) * fact;
      }
    }
    // inelastic electron-neutral collisions (sources.f90:350-373)
    const double sfact = (elchrg / kB) * (gammas(LSP) - 1.0);
    const double Te = T[LSP - 1], nn2 = nnk[1], nn3 = nnk[2];
    double nu = (((sfact * 6.9e-14) * nn3) * 1e-6) / sqrt(Te);
    double iePT = nu * Tn, ieLT = nu;
    nu = (((sfact * 2.9e-14) * nn2) * 1e-6) / sqrt(Te);
    iePT = iePT + nu * Tn;
    ieLT = ieLT + nu;
    double f = 1.06e4 + 7.51e3 * tanh(1.10e-3 * (Te - 1800.0));
    const double gg = (3300.0 + 1.233 * (Te - 1000.0)) - (2.056e-4 * (Te - 1000.0)) * (Te - 4000.0);
    double fact = ((((sfact * 2.99e-12) * nn2) * 1e-6) * exp(((f * (Te - 2000.0)) / Te) / 2000.0)) *
                  (exp((-(gg * (Te - Tn)) / Te) / Tn) - 1.0);
    iePT = iePT - fmax_f(fact, 0.0);
    f = 3300.0 - 839.0 * sin(1.91e-4 * (Te - 2700.0));
    fact = ((((sfact * 5.196e-13) * nn3) * 1e-6) * exp(((f * (Te - 700.0)) / Te) / 700.0)) *
           (exp((-(2770.0 * (Te - Tn)) / Te) / Tn) - 1.0);
    iePT = iePT - fmax_f(fact, 0.0);
    P[LSP - 1] = P[LSP - 1] + (((iePT * n[LSP - 1]) * kB) / (gammas(LSP) - 1.0));
    L[LSP - 1] = L[LSP - 1] + ieLT;
    // energy_source_loss_solve (multifluid.f90:474-495)
    P[LSP - 1] = P[LSP - 1] + a.Qeionize[k];
#pragma unroll
    for (int i = 0; i < LSP; ++i) {
      e[i] = etd(e[i], P[i], L[i], dt);
      s.rhoes[i * g.N + c] = e[i];
    }
  }
#pragma unroll
  for (int isp = 1; isp <= LSP; ++isp) {
    const double t = (((gammas(isp) - 1.0) / kB) * e[isp - 1]) / fmax_f(n[isp - 1], mindensdiv);
    s.Ts[(isp - 1) * g.N + c] = fmax_f(t, 100.0);
  }
}

// d(log p)/dx1 / h1 of species isp at interior cell c (gradient.f90:136-146): centred over
// the ghost cells when the subdomain touches neither x1 end, otherwise over the interior
// with one-sided differences at i1 = 1 and lx1 (sources.f90:259-270).  lp holds log p of
// the block's cells (x1 neighbours are neighbouring threads); cells outside the block are
// evaluated directly, with the same expression.
__device__ __forceinline__ double dlogp(const DevGrid& g, const DevState& s, const double* lp,
                                        int isp, long c, int t, int i1, bool ghosted) {
  const long o = (long)(isp - 1) * g.N;
  auto at = [&](int dt_) -> double {
    const int u = t + dt_;
    if (u >= 0 && u < TPB) return lp[(isp - 1) * TPB + u];
    return log((s.ns[o + c + dt_] * kB) * s.Ts[o + c + dt_]);
  };
  if (!ghosted) {
    if (i1 == 1) return ((at(1) - at(0)) / g.dx1[1]) / g.h1[c];
    if (i1 == g.lx1) return ((at(0) - at(-1)) / g.dx1[g.lx1]) / g.h1[c];
  }
  return ((at(1) - at(-1)) / (g.dx1[i1 + 1] + g.dx1[i1])) / g.h1[c];
}

// ---- source_loss_momentum: srcsMomentum_curv (sources.f90:228-289), ETD, velocities ----
__global__ __launch_bounds__(TPB) void k_src_momentum(DevGrid g, DevState s, DevSources a, double dt) {
  __shared__ double lp[LSP * TPB];
  const int t = threadIdx.x;
  const long c = (long)blockIdx.x * blockDim.x + t;
  if (c < g.N) {
#pragma unroll
    for (int i = 0; i < LSP; ++i) lp[i * TPB + t] = log((s.ns[i * g.N + c] * kB) * s.Ts[i * g.N + c]);
  }
  __syncthreads();
  if (c >= g.N) return;
  const Cell q = decode(g, c);
  double n[LSP], r[LSP - 1];
#pragma unroll
  for (int i = 0; i < LSP; ++i) n[i] = s.ns[i * g.N + c];
#pragma unroll
  for (int i = 0; i < LSP - 1; ++i) r[i] = s.rhovs1[i * g.N + c];
  if (q.interior) {
    const long ni = (long)g.lx1 * g.lx2 * g.lx3, k = q.k;
    const int i1 = q.i1;
    const bool ghosted = !(g.isglobal[0] || g.isglobal[1]);
    double T[LSP], v1[LSP], P[LSP - 1], L[LSP - 1];
#pragma unroll
    for (int i = 0; i < LSP; ++i) {
      T[i] = s.Ts[i * g.N + c];
      v1[i] = s.vs1[i * g.N + c];
    }
    const double Tn = s.Tn[k], vn1 = a.vn1[k];
    double nnk[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) nnk[j] = a.nn[j * ni + k];
#pragma unroll
    for (int isp = 1; isp <= LSP - 1; ++isp) {
      const double nm = n[isp - 1] * ms(isp);
      double Pp = 0.0, Ll = 0.0;
#pragma unroll
      for (int isp2 = 1; isp2 <= 4; ++isp2) {
        const double nu = maxwell(isp, isp2, nnk[isp2 - 1], Tn, T[isp - 1]);
        Ll = Ll + nu;
        Pp = Pp + (nm * nu) * vn1;
      }
      P[isp - 1] = Pp;
      L[isp - 1] = Ll;
    }
#pragma unroll
    for (int pr = 0; pr < 21; ++pr) {
      const int ia = c_pair_a[pr], ib = c_pair_b[pr];
      const Pair p = coulomb_pair(pr, T[ia - 1], T[ib - 1], v1[ia - 1], v1[ib - 1]);
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        const int x = side ? ib : ia, y = side ? ia : ib;
        if (x == LSP) continue;
        const double nu = (c_csj_1e6[x - 1][y - 1] * n[y - 1]) * p.iT15;
        L[x - 1] = L[x - 1] + nu * p.Phi;
        P[x - 1] = P[x - 1] + (((n[x - 1] * c_ms[x - 1]) * nu) * p.Phi) * v1[y - 1];
      }
    }
    // geometric factors, ambipolar field, gravity (sources.f90:248-288)
    const double h1 = g.h1[c], h2 = g.h2[c], h3 = g.h3[c];
    const double h1h2h3 = (h1 * h2) * h3;
    const double dsum = g.dx1[i1 + 1] + g.dx1[i1];
    const double dh2dx1 = ((g.h2[c + 1] - g.h2[c - 1]) / dsum) / h1;
    const double dh3dx1 = ((g.h3[c + 1] - g.h3[c - 1]) / dsum) / h1;
    const double Te = T[LSP - 1];
    const double Epol1 = ((kB * Te) / qs(LSP)) * dlogp(g, s, lp, LSP, c, t, i1, ghosted);
    const double E1f = a.alt[c] < 100e3 ? 0.0 : a.E1[c];
    const double g1 = a.g1[k];
#pragma unroll
    for (int isp = 1; isp <= LSP - 1; ++isp) {
      const double nsp = n[isp - 1], m = ms(isp);
      const double pr = (nsp * kB) * T[isp - 1];
      const double ionpress = pr * dlogp(g, s, lp, isp, c, t, i1, ghosted);
      const double* Qs = a.Q + (long)(isp - 1) * ni;
      double gQ;
      if (i1 == 1) gQ = ((Qs[k + 1] - Qs[k]) / g.dx1[1]) / h1;
      else if (i1 == g.lx1) gQ = ((Qs[k] - Qs[k - 1]) / g.dx1[g.lx1]) / h1;
      else gQ = ((Qs[k + 1] - Qs[k - 1]) / dsum) / h1;
      const double w2 = s.vs2[(isp - 1) * g.N + c], w3 = s.vs3[(isp - 1) * g.N + c];
      const double geom = (((((w2 * w2) * h3) * dh2dx1 + ((w3 * w3) * h2) * dh3dx1) * nsp) * m) / h1h2h3;
      const double Pp = ((((P[isp - 1] + (nsp * qs(isp)) * (E1f + Epol1)) - ionpress) - gQ) + geom) +
                        (nsp * m) * g1;
      r[isp - 1] = etd(r[isp - 1], Pp, L[isp - 1], dt);
      s.rhovs1[(isp - 1) * g.N + c] = r[isp - 1];
    }
  }
  // momentum_source_loss_solve (multifluid.f90:510-527), whole padded grid
  double chrgflux = 0.0;
#pragma unroll
  for (int isp = 1; isp <= LSP - 1; ++isp) {
    const double v = r[isp - 1] / (ms(isp) * fmax_f(n[isp - 1], mindensdiv));
    s.vs1[(isp - 1) * g.N + c] = v;
    chrgflux = chrgflux + (n[isp - 1] * qs(isp)) * v;
  }
  const double ve = -(((1.0 / fmax_f(n[LSP - 1], mindensdiv)) / qs(LSP)) * chrgflux);
  s.vs1[(LSP - 1) * g.N + c] = ve;
  s.rhovs1[(LSP - 1) * g.N + c] = (n[LSP - 1] * ms(LSP)) * ve;
}

// kreac fits of srcsContinuity: ((((c0 s1 c1 t) s2 c2 t^2) s3 c3 t^3) s4 c4 t^4)
__device__ __forceinline__ double poly4(double t, double c0, double c1, double c2, double c3,
                                        double c4, bool p1, bool p2, bool p3, bool p4) {
  const double t2 = t * t, t3 = t2 * t, t4 = t2 * t2;
  double r = c0;
  r = p1 ? r + c1 * t : r - c1 * t;
  r = p2 ? r + c2 * t2 : r - c2 * t2;
  r = p3 ? r + c3 * t3 : r - c3 * t3;
  r = p4 ? r + c4 * t4 : r - c4 * t4;
  return r;
}

// ---- source_loss_mass: srcsContinuity (sources.f90:18-225), Prionize, ETD, ne ----
// CLEAN (the sources block, libgemini.cpp:225-231): clean_param 3 and 1 follow the three
// source passes, with clean_param 2 (velocities only) between them, so this last pass
// also applies them to its cell: temperatures floored at 100 K (ghost and null cells at
// 100 K), ion densities floored at mindens with the electrons their sum (ghost cells at
// mindensdiv, null cells at mindensnull).  Its own reads of Ts come first.
template <bool CLEAN>
__global__ __launch_bounds__(TPB) void k_src_mass(DevGrid g, DevState s, DevSources a, double dt,
                                                  const int* nul) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.N) return;
  const Cell q = decode(g, c);
  if (CLEAN) {
    const bool null = q.interior && nul && nul[q.k] != 0;
    if (!q.interior || null) {
#pragma unroll
      for (int i = 0; i < LSP; ++i) {
        s.Ts[i * g.N + c] = 100.0;
        s.ns[i * g.N + c] = !q.interior ? mindensdiv : (i == 0 ? mindensnull * 1e-2 : mindensnull);
      }
      return;
    }
  }
  double n[LSP];
#pragma unroll
  for (int i = 0; i < LSP; ++i) n[i] = s.ns[i * g.N + c];
  if (q.interior) {
    const long ni = (long)g.lx1 * g.lx2 * g.lx3, k = q.k;
    double P[LSP] = {0, 0, 0, 0, 0, 0, 0}, L[LSP] = {0, 0, 0, 0, 0, 0, 0};
    const double Tn = s.Tn[k];
    const double T1 = s.Ts[c], T3 = s.Ts[2 * g.N + c], T6 = s.Ts[5 * g.N + c];
    const double Te = s.Ts[6 * g.N + c], ne = n[LSP - 1];
    const double d1 = s.vs1[c] - a.vn1[k], d2 = s.vs2[c] - a.vn2[k], d3 = s.vs3[c] - a.vn3[k];
    const double dv2 = ((d1 * d1) + (d2 * d2)) + (d3 * d3);
    auto NN = [&](int j) { return a.nn[(long)(j - 1) * ni + k]; };
    const double e07 = (double)0.7f, e085 = (double)0.85f, e039 = (double)0.39f;
    const double e023 = (double)0.23f, em041 = -(double)0.41f, e044 = (double)0.44f;
    const double em02 = -(double)0.2f, e070 = (double)0.70f, e056 = (double)0.56f;
    const double X16 = (((16.0 * amu) / 3.0) / kB) * dv2, X28 = (((28.0 * amu) / 3.0) / kB) * dv2;
    double Teff, kreac = 0.0, b;
    // O+ + N2 -> NO+ + N
    Teff = (28.0 / (16.0 + 28.0)) * ((X16 + T1) - Tn) + Tn;
    if (Teff <= 3725.0)
      kreac = poly4(Teff / 300.0, 1.71676e-12, 7.19934e-13, 1.33276e-13, 9.28213e-15, 6.39557e-16, false, true, false, true);
    if (Teff > 3725.0 && Teff <= 30000.0)
      kreac = poly4(Teff / 300.0, -1.52489e-11, 7.67112e-13, 1.19064e-13, 1.30858e-15, 4.67756e-18, true, true, false, true);
    if (Teff > 30000.0)
      kreac = poly4(100.0, -1.52489e-11, 7.67112e-13, 1.19064e-13, 1.30858e-15, 4.67756e-18, true, true, false, true);
    b = (kreac * NN(2)) * 1e-6;
    P[1] = P[1] + b * n[0]; L[0] = L[0] + b;
    // O+ + O2 -> O2+ + O
    Teff = (32.0 / (16.0 + 32.0)) * ((X16 + T1) - Tn) + Tn;
    if (Teff <= 4800.0)
      kreac = poly4(Teff / 300.0, 2.78932e-11, 6.92612e-12, 8.67684e-13, 3.47251e-14, 5.07097e-16, false, true, false, true);
    if (Teff > 4800.0 && Teff <= 30000.0) {
      const double t = Teff / 300.0;
      kreac = ((-1.74046e-11 + (3.02328e-12 * Teff) / 300.0) - 2.39214e-15 * (t * t)) - 4.02394e-17 * ((t * t) * t);
    }
    if (Teff > 30000.0)
      kreac = ((-1.74046e-11 + 3.02328e-12 * 100.0) - 2.39214e-15 * 10000.0) - 4.02394e-17 * 1000000.0;
    b = (kreac * NN(3)) * 1e-6;
    P[3] = P[3] + b * n[0]; L[0] = L[0] + b;
    // O+ + NO -> NO+ + O
    Teff = (30.0 / (16.0 + 30.0)) * ((X16 + T1) - Tn) + Tn;
    if (Teff <= 3800.0)
      kreac = poly4(Teff / 300.0, 6.40408e-13, 1.33888e-13, 7.65103e-14, 3.11509e-15, 6.62374e-17, false, true, false, true);
    if (Teff > 3800.0 && Teff <= 30000.0)
      kreac = poly4(Teff / 300.0, -7.48312e-13, 2.31502e-13, 3.07160e-14, 2.65436e-16, 7.76665e-19, true, true, false, true);
    if (Teff > 30000.0)
      kreac = poly4(100.0, -7.48312e-13, 2.31502e-13, 3.07160e-14, 2.65436e-16, 7.76665e-19, true, true, false, true);
    b = (kreac * NN(6)) * 1e-6;
    P[1] = P[1] + b * n[0]; L[0] = L[0] + b;
    // O+ + e -> O + hv
    b = ((3.7e-12 * pow(250.0 / Te, e07)) * ne) * 1e-6;
    L[0] = L[0] + b;
    // N2+ + O -> O+ + N2
    Teff = (16.0 / (28.0 + 16.0)) * ((X28 + T3) - Tn) + Tn;
    kreac = Teff <= 1500.0 ? 1e-11 * pow(300.0 / Teff, e023) : 3.6e-12 * pow(300.0 / Teff, em041);
    b = (kreac * NN(1)) * 1e-6;
    P[0] = P[0] + b * n[2]; L[2] = L[2] + b;
    // N+ + O -> O+ + N
    b = (5e-13 * NN(1)) * 1e-6;
    P[0] = P[0] + b * n[4]; L[4] = L[4] + b;
    // H+ + O -> O+ + H
    b = (((6.e-10 * (8.0 / 9.0)) * sqrt((T6 + Tn / 4.0) / (Tn + T6 / 16.0))) * NN(1)) * 1e-6;
    P[0] = P[0] + b * n[5]; L[5] = L[5] + b;
    // O+ + H -> H+ + O
    b = (6.0e-10 * NN(4)) * 1e-6;
    P[5] = P[5] + b * n[0]; L[0] = L[0] + b;
    // O2+ + N2 -> NO+ + NO (the reference takes nn(:,:,:,3))
    b = (5e-16 * NN(3)) * 1e-6;
    P[1] = P[1] + b * n[3]; L[3] = L[3] + b;
    // O2+ + N -> NO+ + O
    b = (1.2e-10 * NN(5)) * 1e-6;
    P[1] = P[1] + b * n[3]; L[3] = L[3] + b;
    // O2+ + NO -> NO+ + O2
    b = (4.6e-10 * NN(6)) * 1e-6;
    P[1] = P[1] + b * n[3]; L[3] = L[3] + b;
    // N2+ + O -> NO+ + N
    kreac = Teff <= 1500.0 ? 1.4e-10 * pow(300.0 / Teff, e044) : 5.2e-11 * pow(300.0 / Teff, em02);
    b = (kreac * NN(1)) * 1e-6;
    P[1] = P[1] + b * n[2]; L[2] = L[2] + b;
    // N2+ + NO -> NO+ + N2
    b = (4.1e-10 * NN(6)) * 1e-6;
    P[1] = P[1] + b * n[2]; L[2] = L[2] + b;
    // N+ + O2 -> NO+ + O
    b = (2.6e-10 * NN(3)) * 1e-6;
    P[1] = P[1] + b * n[4]; L[4] = L[4] + b;
    // NO+ + e -> N + O
    b = ((4.2e-7 * pow(300.0 / Te, e085)) * ne) * 1e-6;
    L[1] = L[1] + b;
    // N2+ + O2 -> O2+ + N2
    Teff = (32.0 / (28.0 + 32.0)) * ((X28 + T3) - Tn) + Tn;
    b = ((5e-11 * (300.0 / Teff)) * NN(3)) * 1e-6;
    P[3] = P[3] + b * n[2]; L[2] = L[2] + b;
    // N2+ + e -> N + N
    b = ((1.8e-7 * pow(300.0 / Te, e039)) * ne) * 1e-6;
    L[2] = L[2] + b;
    // N+ + O2 -> O2+ + N
    b = (3.1e-10 * NN(3)) * 1e-6;
    P[3] = P[3] + b * n[4]; L[4] = L[4] + b;
    // O2+ + e -> O + O
    kreac = Te <= 1200.0 ? 1.95e-7 * pow(300.0 / Te, e070) : 7.38e-8 * pow(1200.0 / Te, e056);
    b = (kreac * ne) * 1e-6;
    L[3] = L[3] + b;
    // N+ + H -> H+ + N
    b = (3.6e-12 * NN(4)) * 1e-6;
    P[5] = P[5] + b * n[4]; L[4] = L[4] + b;
    // H+ + e -> H + hv
    b = ((3.7e-12 * pow(250.0 / Te, e07)) * ne) * 1e-6;
    L[5] = L[5] + b;
    // mass_source_loss_solve (multifluid.f90:532-549)
#pragma unroll
    for (int i = 0; i < LSP - 1; ++i) {
      n[i] = etd(n[i], P[i] + a.Prionize[i * ni + k], L[i], dt);
      if (CLEAN) n[i] = fmax_f(n[i], mindens);
      s.ns[i * g.N + c] = n[i];
    }
    if (CLEAN) {
#pragma unroll
      for (int i = 0; i < LSP; ++i) s.Ts[i * g.N + c] = fmax_f(s.Ts[i * g.N + c], 100.0);
    }
  }
  double sum = 0.0;
#pragma unroll
  for (int i = 0; i < LSP - 1; ++i) sum = sum + n[i];
  s.ns[(LSP - 1) * g.N + c] = sum;
}

int blocks_of(long n) { return (int)((n + TPB - 1) / TPB); }

__global__ void k_zero(double* p, long n) {
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
    p[k] = 0.0;
}

}  // namespace

// zero n doubles with a kernel (a memset node would be unsafe if a caller captures the call)
void launch_zero(double* p, long n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero, dim3((int)min(blocks_of(n), 16384L)), dim3(TPB), 0, st, p, n);
}

void launch_source_loss_energy(const DevGrid& g, const DevState& s, const DevSources& a,
                               double dt, hipStream_t st) {
  timing_begin("source_energy", st);
  hipLaunchKernelGGL(k_src_energy, dim3(blocks_of(g.N)), dim3(TPB), 0, st, g, s, a, dt);
  timing_end("source_energy", st);
}

void launch_source_loss_momentum(const DevGrid& g, const DevState& s, const DevSources& a,
                                 double dt, hipStream_t st) {
  timing_begin("source_momentum", st);
  hipLaunchKernelGGL(k_src_momentum, dim3(blocks_of(g.N)), dim3(TPB), 0, st, g, s, a, dt);
  timing_end("source_momentum", st);
}

void launch_source_loss_mass(const DevGrid& g, const DevState& s, const DevSources& a,
                             double dt, hipStream_t st, const int* clean_nul, bool clean) {
  timing_begin("source_mass", st);
  if (clean)
    hipLaunchKernelGGL(k_src_mass<true>, dim3(blocks_of(g.N)), dim3(TPB), 0, st, g, s, a, dt,
                       clean_nul);
  else
    hipLaunchKernelGGL(k_src_mass<false>, dim3(blocks_of(g.N)), dim3(TPB), 0, st, g, s, a, dt,
                       nullptr);
  timing_end("source_mass", st);
}

}  // namespace gemhip
