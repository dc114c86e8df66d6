/*
 * gemini_oracle.c -- TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline).
 *
 * A plain-C restatement of the reference's multifluid transport hot path,
 * written to reproduce its floating-point results: every expression keeps the
 * reference's evaluation order (Fortran left-to-right, parentheses kept), no
 * reciprocal is substituted for a divide, and the file is compiled with
 * -ffp-contract=off so no multiply-add is fused (the reference is built for
 * the x86-64 baseline, which has no FMA).  x**2.5 is libm pow(), as gfortran
 * emits for Ts**(5.0_wp/2).
 *
 * Reference files followed (paths relative to the reference root):
 *   src/numerical/advection/advec.f90        (advection, ghost fill, interface velocities)
 *   src/numerical/diffusion/PDEparabolic.f90 (TRBDF21D, backEuler1D)
 *   src/numerical/diffusion/diffusion.f90    (diffusion_prep, TRBDF23D/backEuler3D)
 *   src/collisions/collisions.f90:51-53,162-199 (thermal_conduct)
 *   src/multifluid/multifluid.f90:179-233,309-342 (state conversions, energy_diffusion)
 *   src/mpimod/mpihalo.f90:147-152           (single-process periodic x3 wrap)
 *   src/vendor/lapack95/gbsv.in.f90:58-97    (-> LAPACK dgbsv, Netlib dgbtrf/dgbtf2/
 *                                             dgbtrs/dtbsv/dger/dscal/idamax semantics)
 *   src/numerical/constants/phys_consts.in.f90 (constants)
 *   src/multifluid/multifluid.f90:237-302,553-688 (artificial viscosity, compression,
 *                                             clean_param), src/sources/sources_mpi.f90
 *   src/temporal/temporal_mpi.f90:27-165     (dt_calc, dt_select)
 *   src/sources/sources.f90, src/collisions/collisions.f90:11-159,
 *   src/multifluid/multifluid.f90:130-175,472-549, calculus.in.f90:174-196,
 *   gradient.f90:79-149                      (the source/loss substep)
 */
#include "gemini_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LSP 7

/* ---- constants: phys_consts.in.f90:24-53,82-84; collisions.f90:51-53 ---- */
static const double kB = 1.38064852e-23;
static const double elchrg = 1.60217662e-19;
static const double amu = 1.660539040e-27;
static const double mindensdiv = 1.0e-5;

static double ms_(int isp) {              /* [16,30,28,32,14,1,5.485799090e-4]*amu */
  static const double m[LSP] = {16.0, 30.0, 28.0, 32.0, 14.0, 1.0, 5.485799090e-4};
  return m[isp - 1] * amu;
}
static double qs_(int isp) { return (isp < LSP ? 1.0 : -1.0) * elchrg; }
static double gammas_(int isp) {
  return (isp == 2 || isp == 3 || isp == 4) ? 7.0 / 5.0 : 5.0 / 3.0;
}
/* [real(wp) :: 0.1019e-12, ...]: default-real (single precision) literals
   converted to real64, so each coefficient is the float nearest the decimal. */
static double thermal_coeff_(int isp) {
  static const float c[LSP - 1] = {0.1019e-12f, 0.0747e-12f, 0.0754e-12f,
                                   0.0701e-12f, 0.1068e-12f, 0.3986e-12f};
  return (double)c[isp - 1];
}

void orc_constants(double *ms7, double *qs7, double *gammas7, double *kB_, double *elchrg_,
                   double *thermal_coeff6) {
  for (int i = 1; i <= LSP; ++i) {
    ms7[i - 1] = ms_(i);
    qs7[i - 1] = qs_(i);
    gammas7[i - 1] = gammas_(i);
  }
  for (int i = 1; i < LSP; ++i) thermal_coeff6[i - 1] = thermal_coeff_(i);
  *kB_ = kB;
  *elchrg_ = elchrg;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* x**2.5 as the reference evaluates Ts**(5.0_wp/2): libm pow (mode 0, default).  Mode 1
   instead rounds x^2.5 correctly (x^2 exactly via an FMA split times sqrt(x) with one
   Newton correction, in double-double, then one rounding), the method the GPU library
   uses; it lets the tests check the GPU bit for bit apart from libm's pow, which
   returns the other neighbour of the exact value in about 0.2 % of arguments. */
static int pow_mode = 0;
void orc_set_pow_mode(int mode) { pow_mode = mode; }

static double pow25(double x) {
  if (pow_mode == 0 || !(x > 0x1p-400 && x < 0x1p+400)) return pow(x, 2.5);
  const double s = sqrt(x);
  const double e = fma(-s, s, x);
  const double sl = e / (2.0 * s);
  const double qh = x * x;
  const double ql = fma(x, x, -qh);
  const double ph = qh * s;
  const double pl = fma(qh, s, -ph) + (qh * sl + ql * s);
  return ph + pl;
}

/* y = x**2.5 elementwise with the given mode (for the math checks) */
void orc_eval_pow25(const double *x, double *y, long n, int mode) {
  const int saved = pow_mode;
  pow_mode = mode;
  for (long i = 0; i < n; ++i) y[i] = pow25(x[i]);
  pow_mode = saved;
}

static double fmin_f(double a, double b) { return (b < a) ? b : a; }  /* Fortran MIN */
static double fmax_f(double a, double b) { return (b > a) ? b : a; }  /* Fortran MAX */

/* ---- index helpers (Fortran layout, lower bound -1 for ghosted arrays) ---- */
typedef struct { long L1, L2, L3, N; } dims_t;
static dims_t dims_of(const orc_grid *g) {
  dims_t d;
  d.L1 = g->lx1 + 4; d.L2 = g->lx2 + 4; d.L3 = g->lx3 + 4;
  d.N = d.L1 * d.L2 * d.L3;
  return d;
}
#define GIX(d, i1, i2, i3) ((long)((i1) + 1) + (d).L1 * ((long)((i2) + 1) + (d).L2 * (long)((i3) + 1)))

/* ---- advec.f90:438-447 ---- */
double orc_minmod(double a, double b) {
  if (a * b <= 0.0) return 0.0;
  if (fabs(a) < fabs(b)) return a;
  return b;
}

/* ---- advec.f90:384-435 ----
   f(-1:lx+2), v1i(1:lx+1), dx1(0:lx+2), dx1i(1:lx), ha1(-1:lx+2), ha2i(1:lx+1),
   h1i(1:lx+1); out(-1:lx+2).  C arrays are 0-based: f[k] = f(k-1), v1i[k] = v1i(k+1),
   dx1[k] = dx1(k), dx1i[k] = dx1i(k+1). */
void orc_advec1D_MC_curv(int lx, const double *f, const double *v1i, double dt,
                         const double *dx1, const double *dx1i, const double *ha1,
                         const double *ha2i, const double *h1i, double *out) {
#define F(i) f[(i) + 1]
#define HA1(i) ha1[(i) + 1]
#define V(i) v1i[(i) - 1]
#define HA2I(i) ha2i[(i) - 1]
#define H1I(i) h1i[(i) - 1]
#define DX(i) dx1[(i)]
#define DXI(i) dx1i[(i) - 1]
  if (lx > 1) {
    double *slope = (double *)malloc(sizeof(double) * (size_t)(lx + 2));  /* 0:lx+1 */
    double *phi = (double *)malloc(sizeof(double) * (size_t)(lx + 2));    /* 1:lx+1 */
    double lslope = (F(0) - F(-1)) / DX(0);
    for (int i = 0; i <= lx + 1; ++i) {
      double rslope = (F(i + 1) - F(i)) / DX(i + 1);
      double cslope = (F(i + 1) - F(i - 1)) / (DX(i) + DX(i + 1));
      slope[i] = orc_minmod(cslope, orc_minmod(2.0 * lslope, 2.0 * rslope));
      lslope = rslope;
    }
    for (int i = 1; i <= lx + 1; ++i) {
      double v = V(i);
      if (v < 0.0)
        phi[i] = F(i) * v - ((0.5 * v) * (DX(i) + (v / H1I(i)) * dt)) * slope[i];
      else
        phi[i] = F(i - 1) * v + ((0.5 * v) * (DX(i) - (v / H1I(i)) * dt)) * slope[i - 1];
    }
    for (int i = 1; i <= lx; ++i)
      out[i + 1] = F(i) - ((dt * (HA2I(i + 1) * phi[i + 1] - HA2I(i) * phi[i])) / DXI(i)) / HA1(i);
    out[0] = F(-1); out[1] = F(0);
    out[lx + 2] = F(lx + 1); out[lx + 3] = F(lx + 2);
    free(slope);
    free(phi);
  } else {
    for (int i = 0; i < lx + 4; ++i) out[i] = f[i];
  }
#undef F
#undef HA1
#undef V
#undef HA2I
#undef H1I
#undef DX
#undef DXI
}

/* ---- multifluid.f90:179-233: pointwise state conversions over the padded grid ---- */
void orc_v12rhov1(const orc_grid *g, const double *ns, const double *vs1, double *rhovs1) {
  dims_t d = dims_of(g);
  for (int isp = 1; isp <= LSP; ++isp) {
    long o = (long)(isp - 1) * d.N;
    for (long k = 0; k < d.N; ++k) rhovs1[o + k] = (ns[o + k] * ms_(isp)) * vs1[o + k];
  }
}

void orc_T2rhoe(const orc_grid *g, const double *ns, const double *Ts, double *rhoes) {
  dims_t d = dims_of(g);
  for (int isp = 1; isp <= LSP; ++isp) {
    long o = (long)(isp - 1) * d.N;
    double gm1 = gammas_(isp) - 1.0;
    for (long k = 0; k < d.N; ++k) rhoes[o + k] = ((ns[o + k] * kB) * Ts[o + k]) / gm1;
  }
}

void orc_rhov12v1(const orc_grid *g, double *ns, const double *rhovs1, double *vs1) {
  dims_t d = dims_of(g);
#pragma omp parallel for schedule(static)
  for (long k = 0; k < d.N; ++k) {
    double chrgflux = 0.0;
    for (int isp = 1; isp <= LSP - 1; ++isp) {
      long o = (long)(isp - 1) * d.N + k;
      vs1[o] = rhovs1[o] / (ms_(isp) * fmax_f(ns[o], mindensdiv));
      chrgflux = chrgflux + (ns[o] * qs_(isp)) * vs1[o];
    }
    double s = 0.0;                                   /* sum(ns(:,:,:,1:lsp-1),4) */
    for (int isp = 1; isp <= LSP - 1; ++isp) s = s + ns[(long)(isp - 1) * d.N + k];
    long o7 = (long)(LSP - 1) * d.N + k;
    ns[o7] = s;
    vs1[o7] = -(((1.0 / fmax_f(ns[o7], mindensdiv)) / qs_(LSP)) * chrgflux);
  }
}

void orc_rhoe2T(const orc_grid *g, const double *ns, const double *rhoes, double *Ts) {
  dims_t d = dims_of(g);
  for (int isp = 1; isp <= LSP; ++isp) {
    long o = (long)(isp - 1) * d.N;
    double c = (gammas_(isp) - 1.0) / kB;
    for (long k = 0; k < d.N; ++k) {
      double t = (c * rhoes[o + k]) / fmax_f(ns[o + k], mindensdiv);
      Ts[o + k] = fmax_f(t, 100.0);
    }
  }
}

/* ---- advec.f90:21-149: global-boundary ghost fill (one species at a time) ---- */
static void set_global_boundaries_1(const orc_grid *g, int isp, double *ns, double *rhovs1,
                                    const double *vs1, double *vs2, double *vs3,
                                    double *rhoes, const double *v1i) {
  dims_t d = dims_of(g);
  const int lx1 = g->lx1, lx2 = g->lx2, lx3 = g->lx3;
  const long o = (long)(isp - 1) * d.N;
  double *n = ns + o, *r = rhovs1 + o, *e = rhoes + o, *w2 = vs2 + o, *w3 = vs3 + o;
  const double *u = vs1 + o;
  const double msp = ms_(isp);
  /* v1i(1:lx1+1,1:lx2,1:lx3) */
#define V1I(i1, i2, i3) v1i[((i1) - 1) + (long)(lx1 + 1) * (((i2) - 1) + (long)lx2 * ((i3) - 1))]
#define A3(p, i1, i2, i3) (p)[GIX(d, i1, i2, i3)]
  if (g->isglobal[0]) {                                   /* :42-60 */
    for (int i3 = 1; i3 <= lx3; ++i3)
      for (int i2 = 1; i2 <= lx2; ++i2) {
        double coeff = A3(n, 2, i2, i3) / A3(n, 3, i2, i3);
        A3(n, 0, i2, i3) = fmin_f(coeff * A3(n, 1, i2, i3), A3(n, 1, i2, i3));
        A3(n, -1, i2, i3) = fmin_f(coeff * A3(n, 0, i2, i3), A3(n, 0, i2, i3));
      }
    for (int i3 = 1; i3 <= lx3; ++i3)
      for (int i2 = 1; i2 <= lx2; ++i2)
        A3(r, 0, i2, i3) = 2.0 * V1I(1, i2, i3) - A3(u, 1, i2, i3);
    for (int i3 = -1; i3 <= lx3 + 2; ++i3)
      for (int i2 = -1; i2 <= lx2 + 2; ++i2)
        A3(r, -1, i2, i3) = (A3(r, 0, i2, i3) + A3(r, 0, i2, i3)) - A3(u, 1, i2, i3);
    for (int i3 = -1; i3 <= lx3 + 2; ++i3)
      for (int i2 = -1; i2 <= lx2 + 2; ++i2)
        for (int i1 = -1; i1 <= 0; ++i1)
          A3(r, i1, i2, i3) = (A3(r, i1, i2, i3) * A3(n, i1, i2, i3)) * msp;
    for (int i3 = -1; i3 <= lx3 + 2; ++i3)
      for (int i2 = -1; i2 <= lx2 + 2; ++i2) {
        double Tb = A3(e, 1, i2, i3) / A3(n, 1, i2, i3);
        A3(e, 0, i2, i3) = Tb * A3(n, 0, i2, i3);
        A3(e, -1, i2, i3) = Tb * A3(n, -1, i2, i3);
      }
  }
  if (g->isglobal[1]) {                                   /* :61-79 */
    for (int i3 = 1; i3 <= lx3; ++i3)
      for (int i2 = 1; i2 <= lx2; ++i2) {
        double coeff = A3(n, lx1 - 1, i2, i3) / A3(n, lx1 - 2, i2, i3);
        A3(n, lx1 + 1, i2, i3) = fmin_f(coeff * A3(n, lx1, i2, i3), A3(n, lx1, i2, i3));
        A3(n, lx1 + 2, i2, i3) = fmin_f(coeff * A3(n, lx1 + 1, i2, i3), A3(n, lx1 + 1, i2, i3));
      }
    for (int i3 = 1; i3 <= lx3; ++i3)
      for (int i2 = 1; i2 <= lx2; ++i2)
        A3(r, lx1 + 1, i2, i3) = 2.0 * V1I(lx1 + 1, i2, i3) - A3(u, lx1, i2, i3);
    for (int i3 = -1; i3 <= lx3 + 2; ++i3)
      for (int i2 = -1; i2 <= lx2 + 2; ++i2)
        A3(r, lx1 + 2, i2, i3) = (A3(r, lx1 + 1, i2, i3) + A3(r, lx1 + 1, i2, i3)) - A3(u, lx1, i2, i3);
    for (int i3 = -1; i3 <= lx3 + 2; ++i3)
      for (int i2 = -1; i2 <= lx2 + 2; ++i2)
        for (int i1 = lx1 + 1; i1 <= lx1 + 2; ++i1)
          A3(r, i1, i2, i3) = (A3(r, i1, i2, i3) * A3(n, i1, i2, i3)) * msp;
    for (int i3 = -1; i3 <= lx3 + 2; ++i3)
      for (int i2 = -1; i2 <= lx2 + 2; ++i2) {
        double Tb = A3(e, lx1, i2, i3) / A3(n, lx1, i2, i3);
        A3(e, lx1 + 1, i2, i3) = Tb * A3(n, lx1 + 1, i2, i3);
        A3(e, lx1 + 2, i2, i3) = Tb * A3(n, lx1 + 2, i2, i3);
      }
  }
  /* :83-104, x2 edges: copy the edge cell over both ghosts, all x1 and x3 */
  double *x2f[4] = {w2, n, r, e};
  if (g->isglobal[2])
    for (int q = 0; q < 4; ++q)
      for (int i3 = -1; i3 <= lx3 + 2; ++i3)
        for (int i1 = -1; i1 <= lx1 + 2; ++i1) {
          A3(x2f[q], i1, 0, i3) = A3(x2f[q], i1, 1, i3);
          A3(x2f[q], i1, -1, i3) = A3(x2f[q], i1, 1, i3);
        }
  if (g->isglobal[3])
    for (int q = 0; q < 4; ++q)
      for (int i3 = -1; i3 <= lx3 + 2; ++i3)
        for (int i1 = -1; i1 <= lx1 + 2; ++i1) {
          A3(x2f[q], i1, lx2 + 1, i3) = A3(x2f[q], i1, lx2, i3);
          A3(x2f[q], i1, lx2 + 2, i3) = A3(x2f[q], i1, lx2, i3);
        }
  /* :107-132, x3 edges unless periodic: all x1 and x2 */
  double *x3f[4] = {w3, n, r, e};
  if (!g->flagper) {
    if (g->isglobal[4])
      for (int q = 0; q < 4; ++q)
        for (int i2 = -1; i2 <= lx2 + 2; ++i2)
          for (int i1 = -1; i1 <= lx1 + 2; ++i1) {
            A3(x3f[q], i1, i2, 0) = A3(x3f[q], i1, i2, 1);
            A3(x3f[q], i1, i2, -1) = A3(x3f[q], i1, i2, 1);
          }
    if (g->isglobal[5])
      for (int q = 0; q < 4; ++q)
        for (int i2 = -1; i2 <= lx2 + 2; ++i2)
          for (int i1 = -1; i1 <= lx1 + 2; ++i1) {
            A3(x3f[q], i1, i2, lx3 + 1) = A3(x3f[q], i1, i2, lx3);
            A3(x3f[q], i1, i2, lx3 + 2) = A3(x3f[q], i1, i2, lx3);
          }
  }
#undef V1I
#undef A3
}

void orc_set_global_boundaries_allspec(const orc_grid *g, double *ns, double *rhovs1,
                                       const double *vs1, double *vs2, double *vs3,
                                       double *rhoes, const double *vs1i, int lsp) {
  long nv1i = (long)(g->lx1 + 1) * g->lx2 * g->lx3;
  for (int isp = 1; isp <= lsp; ++isp)
    set_global_boundaries_1(g, isp, ns, rhovs1, vs1, vs2, vs3, rhoes,
                            vs1i + (long)(isp - 1) * nv1i);
}

/* ---- mpihalo.f90:147-152: single-process periodic x3 wrap (lid3 == 1) ---- */
void orc_halo_periodic(const orc_grid *g, double *param, int nspec, int lhalo) {
  if (!g->flagper) return;
  dims_t d = dims_of(g);
  const int lx1 = g->lx1, lx2 = g->lx2, lx3 = g->lx3;
  double *tmp = (double *)malloc(sizeof(double) * (size_t)(lx1 + 4) * (size_t)lhalo);
  for (int s = 0; s < nspec; ++s) {
    double *p = param + (long)s * d.N;
    for (int i2 = 1; i2 <= lx2; ++i2) {
      /* param(-1:lx1+2,1:lx2,1-lhalo:0) = param(-1:lx1+2,1:lx2,lx3-(lhalo-1):lx3) */
      for (int h = 0; h < lhalo; ++h)
        for (int i1 = -1; i1 <= lx1 + 2; ++i1)
          tmp[(i1 + 1) + (long)(lx1 + 4) * h] = p[GIX(d, i1, i2, lx3 - (lhalo - 1) + h)];
      for (int h = 0; h < lhalo; ++h)
        for (int i1 = -1; i1 <= lx1 + 2; ++i1)
          p[GIX(d, i1, i2, 1 - lhalo + h)] = tmp[(i1 + 1) + (long)(lx1 + 4) * h];
      /* param(-1:lx1+2,1:lx2,lx3+1:lx3+lhalo) = param(-1:lx1+2,1:lx2,1:lhalo) */
      for (int h = 0; h < lhalo; ++h)
        for (int i1 = -1; i1 <= lx1 + 2; ++i1)
          tmp[(i1 + 1) + (long)(lx1 + 4) * h] = p[GIX(d, i1, i2, 1 + h)];
      for (int h = 0; h < lhalo; ++h)
        for (int i1 = -1; i1 <= lx1 + 2; ++i1)
          p[GIX(d, i1, i2, lx3 + 1 + h)] = tmp[(i1 + 1) + (long)(lx1 + 4) * h];
    }
  }
  free(tmp);
}

/* ---- advec.f90:155-229: interface velocities ---- */
void orc_interface_vels_allspec(const orc_grid *g, const double *vs1, const double *vs2,
                                const double *vs3, double *vs1i, double *vs2i, double *vs3i,
                                int lsp) {
  dims_t d = dims_of(g);
  const int lx1 = g->lx1, lx2 = g->lx2, lx3 = g->lx3;
  const long n1 = (long)(lx1 + 1) * lx2 * lx3, n2 = (long)lx1 * (lx2 + 1) * lx3,
             n3 = (long)lx1 * lx2 * (lx3 + 1);
  for (int isp = 1; isp <= lsp; ++isp) {
    const double *u1 = vs1 + (long)(isp - 1) * d.N, *u2 = vs2 + (long)(isp - 1) * d.N,
                 *u3 = vs3 + (long)(isp - 1) * d.N;
    double *v1 = vs1i + (long)(isp - 1) * n1, *v2 = vs2i + (long)(isp - 1) * n2,
           *v3 = vs3i + (long)(isp - 1) * n3;
#define V1(i1, i2, i3) v1[((i1) - 1) + (long)(lx1 + 1) * (((i2) - 1) + (long)lx2 * ((i3) - 1))]
#define V2(i1, i2, i3) v2[((i1) - 1) + (long)lx1 * (((i2) - 1) + (long)(lx2 + 1) * ((i3) - 1))]
#define V3(i1, i2, i3) v3[((i1) - 1) + (long)lx1 * (((i2) - 1) + (long)lx2 * ((i3) - 1))]
    for (int i3 = 1; i3 <= lx3; ++i3)
      for (int i2 = 1; i2 <= lx2; ++i2) {
        for (int i1 = 2; i1 <= lx1; ++i1)
          V1(i1, i2, i3) = 0.5 * (u1[GIX(d, i1 - 1, i2, i3)] + u1[GIX(d, i1, i2, i3)]);
        if (g->isglobal[0]) {
          if (g->gridflag == 0) V1(1, i2, i3) = u1[GIX(d, 1, i2, i3)];
          else if (g->gridflag == 1) V1(1, i2, i3) = fmin_f(V1(2, i2, i3), 0.0);
          else V1(1, i2, i3) = u1[GIX(d, 1, i2, i3)];
        } else {
          V1(1, i2, i3) = 0.5 * (u1[GIX(d, 0, i2, i3)] + u1[GIX(d, 1, i2, i3)]);
        }
        if (g->isglobal[1]) {
          if (g->gridflag == 0) V1(lx1 + 1, i2, i3) = u1[GIX(d, lx1, i2, i3)];
          else if (g->gridflag == 1) V1(lx1 + 1, i2, i3) = u1[GIX(d, lx1, i2, i3)];
          else V1(lx1 + 1, i2, i3) = fmax_f(V1(lx1, i2, i3), 0.0);
        } else {
          V1(lx1 + 1, i2, i3) = 0.5 * (u1[GIX(d, lx1, i2, i3)] + u1[GIX(d, lx1 + 1, i2, i3)]);
        }
      }
    for (int i3 = 1; i3 <= lx3; ++i3)
      for (int i2 = 1; i2 <= lx2 + 1; ++i2)
        for (int i1 = 1; i1 <= lx1; ++i1)
          V2(i1, i2, i3) = 0.5 * (u2[GIX(d, i1, i2 - 1, i3)] + u2[GIX(d, i1, i2, i3)]);
    for (int i3 = 1; i3 <= lx3 + 1; ++i3)
      for (int i2 = 1; i2 <= lx2; ++i2)
        for (int i1 = 1; i1 <= lx1; ++i1)
          V3(i1, i2, i3) = 0.5 * (u3[GIX(d, i1, i2, i3 - 1)] + u3[GIX(d, i1, i2, i3)]);
#undef V1
#undef V2
#undef V3
  }
}

/* ---- advec.f90:233-381: dimensionally split sweeps ---- */
void orc_sweep_allspec(const orc_grid *g, int dir, double *fs, const double *vsi, double dt,
                       int frank, int lsp) {
  dims_t d = dims_of(g);
  const int lx1 = g->lx1, lx2 = g->lx2, lx3 = g->lx3;
  const int lx = dir == 1 ? lx1 : dir == 2 ? lx2 : lx3;
  const long nvi = dir == 1 ? (long)(lx1 + 1) * lx2 * lx3
                 : dir == 2 ? (long)lx1 * (lx2 + 1) * lx3 : (long)lx1 * lx2 * (lx3 + 1);
  const double *dx = dir == 1 ? g->dx1 : dir == 2 ? g->dx2 : g->dx3;
  const double *dxi = dir == 1 ? g->dx1i : dir == 2 ? g->dx2i : g->dx3i;
  const int nlineA = dir == 1 ? lx2 : lx1, nlineB = dir == 3 ? lx2 : lx3;
  for (int isp = 1; isp <= lsp; ++isp) {
    double *f = fs + (long)(isp - 1) * d.N;
    const double *vi = vsi + (long)(isp - 1) * nvi;
#pragma omp parallel for collapse(2) schedule(static)
    for (int b = 1; b <= nlineB; ++b)
      for (int a = 1; a <= nlineA; ++a) {
        double fl[lx + 4], out[lx + 4], ha1[lx + 4], vl[lx + 1], ha2i[lx + 1], h1i[lx + 1];
        for (int k = -1; k <= lx + 2; ++k) {
          long c = dir == 1 ? GIX(d, k, a, b) : dir == 2 ? GIX(d, a, k, b) : GIX(d, a, b, k);
          fl[k + 1] = f[c];
          double h1 = g->h1[c], h2 = g->h2[c], h3 = g->h3[c];
          if (dir == 1 || frank == 0) ha1[k + 1] = (h1 * h2) * h3;   /* :302, :334, :370 */
          else ha1[k + 1] = ((h1 * h1) * h2) * h3;                     /* :337, :373 */
        }
        for (int k = 1; k <= lx + 1; ++k) {
          long ci;
          if (dir == 1) {
            ci = (k - 1) + (long)(lx1 + 1) * ((a - 1) + (long)lx2 * (b - 1));
            ha2i[k - 1] = g->h2x1i[ci] * g->h3x1i[ci];                  /* :303 */
            h1i[k - 1] = g->h1x1i[ci];
          } else if (dir == 2) {
            ci = (a - 1) + (long)lx1 * ((k - 1) + (long)(lx2 + 1) * (b - 1));
            double p = g->h1x2i[ci];
            ha2i[k - 1] = frank == 0 ? p * g->h3x2i[ci] : (p * p) * g->h3x2i[ci];
            h1i[k - 1] = g->h2x2i[ci];
          } else {
            ci = (a - 1) + (long)lx1 * ((b - 1) + (long)lx2 * (k - 1));
            double p = g->h1x3i[ci];
            ha2i[k - 1] = frank == 0 ? p * g->h2x3i[ci] : (p * p) * g->h2x3i[ci];
            h1i[k - 1] = g->h3x3i[ci];
          }
          vl[k - 1] = vi[ci];
        }
        orc_advec1D_MC_curv(lx, fl, vl, dt, dx, dxi, ha1, ha2i, h1i, out);
        for (int k = -1; k <= lx + 2; ++k) {
          long c = dir == 1 ? GIX(d, k, a, b) : dir == 2 ? GIX(d, a, k, b) : GIX(d, a, b, k);
          f[c] = out[k + 1];
        }
      }
  }
}

/* ---- collisions.f90:162-199 ---- */
void orc_thermal_conduct(const orc_grid *g, int isp, const double *Ts, const double *ns,
                         const double *nn, const double *J1, double *lambda, double *beta) {
  dims_t d = dims_of(g);
  const int lx1 = g->lx1, lx2 = g->lx2, lx3 = g->lx3;
  const long ni = (long)lx1 * lx2 * lx3;
  for (int i3 = 1; i3 <= lx3; ++i3)
    for (int i2 = 1; i2 <= lx2; ++i2)
      for (int i1 = 1; i1 <= lx1; ++i1) {
        long c = GIX(d, i1, i2, i3);
        long k = (i1 - 1) + (long)lx1 * ((i2 - 1) + (long)lx2 * (i3 - 1));
        double T = Ts[c];
        if (isp < LSP) {
          lambda[k] = thermal_coeff_(isp) * pow25(T);
          beta[k] = 0.0;
        } else {
          double n1 = nn[k], n2 = nn[ni + k], n3 = nn[2 * ni + k];
          double sq = sqrt(T);
          double br = ((n1 * 1.1e-16) * (1.0 + 5.7e-4 * T) + ((n2 * 2.82e-17) * sq) * (1.0 - 1.21e-4 * T))
                      + (n3 * 2.2e-16) * (1.0 + 3.6e-2 * sq);
          double den = 1.0 + ((3.22e4 * (T * T)) / ns[c]) * br;
          lambda[k] = (((elchrg * 100.0) * 7.7e5) * pow25(T)) / den;
          beta[k] = (((5.0 / 2.0) * kB) / elchrg) * J1[c];
        }
      }
}

/* ---- LAPACK dgbsv restated (Netlib reference semantics) ----
   ab is the LAPACK band array AB(ldab=2*kl+ku+1, n), column-major, 1-based in the
   macros below; b is the single right-hand side.  Returns INFO. */
/* BLAS arithmetic of the band LU and solve: 0 (default) the Netlib reference BLAS, every
   product and sum rounded separately; 1 fused multiply-add in the dger (factorisation and forward solve) and dtbsv updates, as
   optimised BLAS libraries (OpenBLAS on FMA-capable x86) evaluate them.  Mode 1 exists
   only so the tests can reproduce such a library bit for bit. */
static int blas_fma = 0;
void orc_set_blas_fma(int mode) { blas_fma = mode; }

#define AB(i, j) ab[((i) - 1) + (long)ldab * ((j) - 1)]
static int dgbtf2_(int n, int kl, int ku, double *ab, int ldab, int *ipiv) {
  int info = 0;
  const int kv = ku + kl;
  for (int j = ku + 2; j <= (kv < n ? kv : n); ++j)       /* zero fill-in columns ku+2..kv */
    for (int i = kv - j + 2; i <= kl; ++i) AB(i, j) = 0.0;
  int ju = 1;
  for (int j = 1; j <= n; ++j) {
    if (j + kv <= n)
      for (int i = 1; i <= kl; ++i) AB(i, j + kv) = 0.0;
    const int km = kl < n - j ? kl : n - j;
    int jp = 1;                                             /* idamax: first maximal |x| */
    double dmax = fabs(AB(kv + 1, j));
    for (int i = 2; i <= km + 1; ++i) {
      double v = fabs(AB(kv + i, j));
      if (v > dmax) { jp = i; dmax = v; }
    }
    ipiv[j - 1] = jp + j - 1;
    if (AB(kv + jp, j) != 0.0) {
      int jt = j + ku + jp - 1 < n ? j + ku + jp - 1 : n;
      if (jt > ju) ju = jt;
      if (jp != 1)                                          /* dswap along the row */
        for (int k = 0; k <= ju - j; ++k) {
          double t = AB(kv + jp - k, j + k);
          AB(kv + jp - k, j + k) = AB(kv + 1 - k, j + k);
          AB(kv + 1 - k, j + k) = t;
        }
      if (km > 0) {
        const double rp = 1.0 / AB(kv + 1, j);              /* dscal(km, ONE/pivot) */
        for (int i = 1; i <= km; ++i) AB(kv + 1 + i, j) = rp * AB(kv + 1 + i, j);
        if (ju > j)                                         /* dger(km, ju-j, -ONE, ...) */
          for (int k = 1; k <= ju - j; ++k) {
            const double y = AB(kv + 1 - k, j + k);
            if (y != 0.0) {
              const double temp = -1.0 * y;
              for (int i = 1; i <= km; ++i)
                AB(kv + 1 + i - k, j + k) =
                    blas_fma ? fma(AB(kv + 1 + i, j), temp, AB(kv + 1 + i - k, j + k))
                             : AB(kv + 1 + i - k, j + k) + AB(kv + 1 + i, j) * temp;
            }
          }
      }
    } else if (info == 0) {
      info = j;
    }
  }
  return info;
}

static void dgbtrs_n_(int n, int kl, int ku, const double *ab, int ldab, const int *ipiv,
                      double *b) {
  const int kd = ku + kl + 1;
  if (kl > 0)
    for (int j = 1; j <= n - 1; ++j) {
      const int lm = kl < n - j ? kl : n - j;
      const int l = ipiv[j - 1];
      if (l != j) { double t = b[l - 1]; b[l - 1] = b[j - 1]; b[j - 1] = t; }
      const double y = b[j - 1];                            /* dger(lm, 1, -ONE, ...) */
      if (y != 0.0) {
        const double temp = -1.0 * y;
        for (int i = 1; i <= lm; ++i)
          b[j + i - 1] = blas_fma ? fma(AB(kd + i, j), temp, b[j + i - 1])
                                  : b[j + i - 1] + AB(kd + i, j) * temp;
      }
    }
  const int k = kl + ku, kplus1 = k + 1;                   /* dtbsv('U','N','N') */
  for (int j = n; j >= 1; --j) {
    if (b[j - 1] != 0.0) {
      const int l = kplus1 - j;
      b[j - 1] = b[j - 1] / AB(kplus1, j);
      const double temp = b[j - 1];
      for (int i = j - 1; i >= (1 > j - k ? 1 : j - k); --i)
        b[i - 1] = blas_fma ? fma(-temp, AB(l + i, j), b[i - 1]) : b[i - 1] - temp * AB(l + i, j);
    }
  }
}
#undef AB

/* rows interchanged by the factorisations since the last reset (tests use it to show a
   case exercises pivoting) */
static long swap_count = 0;
void orc_reset_swap_count(void) { swap_count = 0; }
long orc_swap_count(void) { return swap_count; }

int orc_gbsv(int n, int kl, double *ab, double *b, int *ipiv) {
  const int ldab = 3 * kl + 1;                  /* gbsv.in.f90:72,88: KU = LDA-2*KL-1 */
  const int ku = ldab - 2 * kl - 1;
  int info = dgbtf2_(n, kl, ku, ab, ldab, ipiv);  /* dgbtrf picks dgbtf2 for KU <= 64 */
  long sw = 0;
  for (int j = 0; j < n; ++j) sw += ipiv[j] != j + 1;
  if (sw) {
#pragma omp atomic
    swap_count += sw;
  }
  if (info == 0) dgbtrs_n_(n, kl, ku, ab, ldab, ipiv, b);
  return info;
}

/* ---- PDEparabolic.f90:19-218 ----
   Ts,A..E(1:lx1) -> arrays [0..lx1-1]; dx1(0:lx1+2) -> [0..lx1+2]; dx1i(1:lx1) -> [0..]. */
/* top_flux = 0: the reference's Dirichlet row lx1 (T = Tsmaxx1).  top_flux = 1: the
   Neumann row that PDEparabolic.f90:321-326 carries commented out, (T(lx1) - T(lx1-1)) /
   dx1(lx1) = Tsmaxx1, in both stages (an extension for a top heat flux). */
static void trbdf2_core(int lx1, const double *Ts, const double *A, const double *B,
                        const double *C, const double *D, const double *E, double Tsminx1,
                        double Tsmaxx1, double dt, const double *dx1, const double *dx1i,
                        double *out, int top_flux) {
  const int ll = 2;
  double *M = (double *)calloc((size_t)7 * (size_t)lx1, sizeof(double));
  double *Dh = (double *)malloc(sizeof(double) * (size_t)lx1);
  double *TR = (double *)malloc(sizeof(double) * (size_t)lx1);
  int *ipiv = (int *)malloc(sizeof(int) * (size_t)lx1);
#define MM(r, c) M[((r) - 1) + 7L * ((c) - 1)]
#define T_(i) Ts[(i) - 1]
#define A_(i) A[(i) - 1]
#define B_(i) B[(i) - 1]
#define C_(i) C[(i) - 1]
#define E_(i) E[(i) - 1]
#define DH(i) Dh[(i) - 1]
#define DX(i) dx1[(i)]
#define DXI(i) dx1i[(i) - 1]
  DH(1) = 0.0;
  for (int i = 2; i <= lx1; ++i) DH(i) = 0.5 * (D[i - 2] + D[i - 1]);

  /* TR half step */
  int i = 1;
  MM(ll + 3, i) = 1.0; MM(ll + 2, i + 1) = 0.0; MM(ll + 1, i + 2) = 0.0;
  TR[i - 1] = Tsminx1;
  for (i = 2; i <= lx1 - 1; ++i) {
    const double t1 = (((C_(i) * DH(i)) / DXI(i)) / DX(i)) / 2.0;
    const double t3 = (((C_(i) * DH(i + 1)) / DXI(i)) / DX(i + 1)) / 2.0;
    const double t2 = (B_(i) / (DX(i + 1) + DX(i))) / 2.0;
    if (i >= 3) MM(ll + 5, i - 2) = 0.0;
    MM(ll + 4, i - 1) = -t1 + t2;
    MM(ll + 3, i) = ((1.0 / (dt / 2.0) - A_(i) / 2.0) + t3) + t1;
    MM(ll + 2, i + 1) = -t3 - t2;
    if (i <= lx1 - 2) MM(ll + 1, i + 2) = 0.0;
    double r = T_(i) / (dt / 2.0) + E_(i);
    if (i >= 3) r = r - MM(ll + 5, i - 2) * T_(i - 2);
    r = r - MM(ll + 4, i - 1) * T_(i - 1);
    r = r - ((-(A_(i) / 2.0) + t3) + t1) * T_(i);
    r = r - MM(ll + 2, i + 1) * T_(i + 1);
    if (i <= lx1 - 2) r = r - MM(ll + 1, i + 2) * T_(i + 2);
    TR[i - 1] = r;
  }
  i = lx1;
  MM(ll + 5, i - 2) = 0.0;
  if (top_flux) { MM(ll + 4, i - 1) = -1.0 / DX(i); MM(ll + 3, i) = 1.0 / DX(i); }
  else { MM(ll + 4, i - 1) = 0.0; MM(ll + 3, i) = 1.0; }
  TR[i - 1] = Tsmaxx1;
  orc_gbsv(lx1, 2, M, TR, ipiv);

  /* BDF2 step */
  i = 1;
  MM(ll + 3, i) = 1.0; MM(ll + 2, i + 1) = 0.0; MM(ll + 1, i + 2) = 0.0;
  out[i - 1] = Tsminx1;
  for (i = 2; i <= lx1 - 1; ++i) {
    const double s1 = ((C_(i) * DH(i)) / DXI(i)) / DX(i);
    const double s3 = ((C_(i) * DH(i + 1)) / DXI(i)) / DX(i + 1);
    const double s2 = B_(i) / (DX(i + 1) + DX(i));
    if (i >= 3) MM(ll + 5, i - 2) = 0.0;
    MM(ll + 4, i - 1) = -s1 + s2;
    MM(ll + 3, i) = ((1.0 / (dt / 3.0) - A_(i)) + s3) + s1;
    MM(ll + 2, i + 1) = -s3 - s2;
    if (i <= lx1 - 2) MM(ll + 1, i + 2) = 0.0;
    out[i - 1] = (E_(i) - ((1.0 / 3.0) * T_(i)) / (dt / 3.0)) + ((4.0 / 3.0) * TR[i - 1]) / (dt / 3.0);
  }
  i = lx1;
  MM(ll + 5, i - 2) = 0.0;
  if (top_flux) { MM(ll + 4, i - 1) = -1.0 / DX(i); MM(ll + 3, i) = 1.0 / DX(i); }
  else { MM(ll + 4, i - 1) = 0.0; MM(ll + 3, i) = 1.0; }
  out[i - 1] = Tsmaxx1;
  orc_gbsv(lx1, 2, M, out, ipiv);
  free(M); free(Dh); free(TR); free(ipiv);
}

void orc_TRBDF21D(int lx1, const double *Ts, const double *A, const double *B,
                  const double *C, const double *D, const double *E, double Tsminx1,
                  double Tsmaxx1, double dt, const double *dx1, const double *dx1i,
                  double *out) {
  trbdf2_core(lx1, Ts, A, B, C, D, E, Tsminx1, Tsmaxx1, dt, dx1, dx1i, out, 0);
}

void orc_TRBDF21D_flux(int lx1, const double *Ts, const double *A, const double *B,
                       const double *C, const double *D, const double *E, double Tsminx1,
                       double gtop, double dt, const double *dx1, const double *dx1i,
                       double *out) {
  trbdf2_core(lx1, Ts, A, B, C, D, E, Tsminx1, gtop, dt, dx1, dx1i, out, 1);
}

/* ---- PDEparabolic.f90:221-346 (coeffs(lx1,3) and rhs(lx1) optional outputs) ---- */
void orc_backEuler1D(int lx1, const double *Ts, const double *A, const double *B,
                     const double *C, const double *D, const double *E, double Tsminx1,
                     double Tsmaxx1, double dt, const double *dx1, const double *dx1i,
                     double *out, double *coeffs, double *rhs) {
  const int ll = 2;
  double *M = (double *)calloc((size_t)7 * (size_t)lx1, sizeof(double));
  double *Dh = (double *)malloc(sizeof(double) * (size_t)lx1);
  int *ipiv = (int *)malloc(sizeof(int) * (size_t)lx1);
  DH(1) = 0.0;
  for (int i = 2; i <= lx1; ++i) DH(i) = 0.5 * (D[i - 2] + D[i - 1]);
  for (int i = 1; i <= lx1; ++i) out[i - 1] = T_(i) / dt + E_(i);
  int i = 1;
  MM(ll + 3, i) = 1.0; MM(ll + 2, i + 1) = 0.0; MM(ll + 1, i + 2) = 0.0;
  out[i - 1] = Tsminx1;
  for (i = 2; i <= lx1 - 1; ++i) {
    const double s1 = ((C_(i) * DH(i)) / DXI(i)) / DX(i);
    const double s3 = ((C_(i) * DH(i + 1)) / DXI(i)) / DX(i + 1);
    const double s2 = B_(i) / (DX(i + 1) + DX(i));
    if (i >= 3) MM(ll + 5, i - 2) = 0.0;
    MM(ll + 4, i - 1) = -s1 + s2;
    MM(ll + 3, i) = ((1.0 / dt - A_(i)) + s3) + s1;
    MM(ll + 2, i + 1) = -s3 - s2;
    if (i <= lx1 - 2) MM(ll + 1, i + 2) = 0.0;
  }
  i = lx1;
  MM(ll + 5, i - 2) = 0.0; MM(ll + 4, i - 1) = 0.0; MM(ll + 3, i) = 1.0;
  out[i - 1] = Tsmaxx1;
  if (rhs) for (int k = 0; k < lx1; ++k) rhs[k] = out[k];
  orc_gbsv(lx1, 2, M, out, ipiv);
  if (coeffs)
    for (int k = 1; k <= lx1; ++k) {
      coeffs[(k - 1)] = MM(ll + 2, k);
      coeffs[(k - 1) + lx1] = MM(ll + 3, k);
      coeffs[(k - 1) + 2L * lx1] = MM(ll + 4, k);
    }
  free(M); free(Dh); free(ipiv);
#undef MM
#undef T_
#undef A_
#undef B_
#undef C_
#undef E_
#undef DH
#undef DX
#undef DXI
}

/* ---- multifluid.f90:309-342 with diffusion.f90:24-152 ----
   nn(1:lx1,1:lx2,1:lx3,>=3) and Tn(1:lx1,1:lx2,1:lx3) interior arrays; J1 ghosted.
   Returns 0, or -1 for an unsupported diffsolvetype (the reference's error stop).
   Ghost cells of Ts, which the reference leaves undefined (the TRBDF23D result is only
   assigned on 1:lx1 x 1:lx2 x 1:lx3, diffusion.f90:136,149), keep their input values
   and are floored at 100 like every other cell. */
static int energy_diffusion_core(const orc_grid *g, double dt, const double *ns, double *Ts,
                                 const double *J1, const double *nn, const double *Tn,
                                 int diffsolvetype, double Teinf, const double *kappa,
                                 const double *qtop) {
  if (diffsolvetype != 1 && diffsolvetype != 2) return -1;
  if ((kappa || qtop) && diffsolvetype != 2) return -2;
  if (qtop && g->gridflag != 2) return -3;
  dims_t d = dims_of(g);
  const int lx1 = g->lx1, lx2 = g->lx2, lx3 = g->lx3;
  const long ni = (long)lx1 * lx2 * lx3;
  double *lambda = (double *)malloc(sizeof(double) * (size_t)ni);
  double *beta = (double *)malloc(sizeof(double) * (size_t)ni);
  for (int isp = 1; isp <= LSP; ++isp) {
    double *T = Ts + (long)(isp - 1) * d.N;
    const double *n = ns + (long)(isp - 1) * d.N;
    orc_thermal_conduct(g, isp, T, n, nn, J1, lambda, beta);
    const double gm1kB = (gammas_(isp) - 1.0) / kB;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i3 = 1; i3 <= lx3; ++i3)
      for (int i2 = 1; i2 <= lx2; ++i2) {
        double Tc[lx1], A[lx1], B[lx1], C[lx1], D[lx1], E[lx1], res[lx1];
        const long col = (i2 - 1) + (long)lx2 * (i3 - 1);
        const int elec = isp == LSP;
        double lamN = 0.0;
        for (int i1 = 1; i1 <= lx1; ++i1) {
          long c = GIX(d, i1, i2, i3);
          long k = (i1 - 1) + (long)lx1 * ((i2 - 1) + (long)lx2 * (i3 - 1));
          double h1 = g->h1[c], h2 = g->h2[c], h3 = g->h3[c];
          double lam = lambda[k];
          if (elec && kappa) lam = lam * kappa[col];
          if (i1 == lx1) lamN = lam;
          A[i1 - 1] = 0.0;
          C[i1 - 1] = (gm1kB / fmax_f(n[c], mindensdiv)) / ((h1 * h2) * h3);
          B[i1 - 1] = (C[i1 - 1] * beta[k]) / h1;
          D[i1 - 1] = ((lam * h2) * h3) / h1;
          E[i1 - 1] = 0.0;
          Tc[i1 - 1] = T[c];
        }
        long kb = (long)lx1 * ((i2 - 1) + (long)lx2 * (i3 - 1));
        double Tmin, Tmax;
        if (g->gridflag == 0) { Tmin = Tn[kb]; Tmax = Tn[kb + lx1 - 1]; }
        else if (g->gridflag == 1) { Tmax = Tn[kb + lx1 - 1]; Tmin = Teinf; }
        else { Tmin = Tn[kb]; Tmax = Teinf; }
        if (elec && qtop) {
          /* downward heat flux q through the top face: lambda/h1 dT/dx1 = q */
          const double gtop = (qtop[col] * g->h1[GIX(d, lx1, i2, i3)]) / lamN;
          orc_TRBDF21D_flux(lx1, Tc, A, B, C, D, E, Tmin, gtop, dt, g->dx1, g->dx1i, res);
        } else if (diffsolvetype == 2)
          orc_TRBDF21D(lx1, Tc, A, B, C, D, E, Tmin, Tmax, dt, g->dx1, g->dx1i, res);
        else
          orc_backEuler1D(lx1, Tc, A, B, C, D, E, Tmin, Tmax, dt, g->dx1, g->dx1i, res, 0, 0);
        for (int i1 = 1; i1 <= lx1; ++i1) T[GIX(d, i1, i2, i3)] = res[i1 - 1];
      }
    for (long k = 0; k < d.N; ++k) T[k] = fmax_f(T[k], 100.0);
  }
  free(lambda);
  free(beta);
  return 0;
}

int orc_energy_diffusion(const orc_grid *g, double dt, const double *ns, double *Ts,
                         const double *J1, const double *nn, const double *Tn,
                         int diffsolvetype, double Teinf) {
  return energy_diffusion_core(g, dt, ns, Ts, J1, nn, Tn, diffsolvetype, Teinf, 0, 0);
}

int orc_energy_diffusion_ext(const orc_grid *g, double dt, const double *ns, double *Ts,
                             const double *J1, const double *nn, const double *Tn,
                             int diffsolvetype, double Teinf, const double *kappa,
                             const double *qtop) {
  return energy_diffusion_core(g, dt, ns, Ts, J1, nn, Tn, diffsolvetype, Teinf, kappa, qtop);
}

/* ---- multifluid.f90:553-688 clean_param (paramflag 1 density, 2 velocity vs1,
   3 temperature) over the ghosted 7-species field.  nullpts(lx1,lx2,lx3) (nonzero: null
   cell, meshobj.f90:586-619; NULL: no null cells). */
static const double mindens = 1.0e-100, mindensnull = 1.0e-20;   /* phys_consts.in.f90:82-83 */
int orc_clean_param(const orc_grid *g, int iparm, double *param, const int *nullpts) {
  dims_t d = dims_of(g);
  const int lx1 = g->lx1, lx2 = g->lx2, lx3 = g->lx3;
#define P(isp, i1, i2, i3) param[(long)((isp) - 1) * d.N + GIX(d, i1, i2, i3)]
#define NUL(i1, i2, i3) (nullpts && nullpts[((i1) - 1) + (long)lx1 * (((i2) - 1) + (long)lx2 * ((i3) - 1))])
  if (iparm == 1) {
    for (int isp = 1; isp <= LSP - 1; ++isp)
      for (long k = 0; k < d.N; ++k) {
        double *v = param + (long)(isp - 1) * d.N + k;
        *v = fmax_f(*v, mindens);
      }
    for (long k = 0; k < d.N; ++k) {
      double sacc = 0.0;
      for (int isp = 1; isp <= LSP - 1; ++isp) sacc = sacc + param[(long)(isp - 1) * d.N + k];
      param[(long)(LSP - 1) * d.N + k] = sacc;
    }
    for (int isp = 1; isp <= LSP; ++isp)
      for (int i3 = 1; i3 <= lx3; ++i3)
        for (int i2 = 1; i2 <= lx2; ++i2)
          for (int i1 = 1; i1 <= lx1; ++i1)
            if (NUL(i1, i2, i3)) P(isp, i1, i2, i3) = isp == 1 ? mindensnull * 1e-2 : mindensnull;
  } else if (iparm == 2) {
    for (int isp = 1; isp <= LSP; ++isp)
      for (int i3 = 1; i3 <= lx3; ++i3)
        for (int i2 = 1; i2 <= lx2; ++i2)
          for (int i1 = 1; i1 <= lx1; ++i1)
            if (NUL(i1, i2, i3)) P(isp, i1, i2, i3) = 0.0;
    if (g->gridflag == 0) {
      for (int isp = 1; isp <= LSP; ++isp)
        for (int i3 = 1; i3 <= lx3; ++i3)
          for (int i2 = 1; i2 <= lx2; ++i2) {
            int beg = 1;
            while (!NUL(beg, i2, i3) && beg < lx1) ++beg;
            int end = beg;
            while (NUL(end, i2, i3) && end < lx1) ++end;
            if (beg != lx1) P(isp, beg, i2, i3) = P(isp, beg + 1, i2, i3);
            if (end != lx1) P(isp, end, i2, i3) = P(isp, end - 1, i2, i3);
          }
    } else if (g->gridflag == 1) {
      for (int isp = 1; isp <= LSP; ++isp)
        for (int i3 = 1; i3 <= lx3; ++i3)
          for (int i2 = 1; i2 <= lx2; ++i2) {
            int end = 1;
            while (!NUL(end, i2, i3) && end < lx1) ++end;
            if (end != lx1) P(isp, end, i2, i3) = P(isp, end - 1, i2, i3);
          }
    }
    return 0;
  } else if (iparm == 3) {
    for (long k = 0; k < (long)LSP * d.N; ++k) param[k] = fmax_f(param[k], 100.0);
    for (int isp = 1; isp <= LSP; ++isp)
      for (int i3 = 1; i3 <= lx3; ++i3)
        for (int i2 = 1; i2 <= lx2; ++i2)
          for (int i1 = 1; i1 <= lx1; ++i1)
            if (NUL(i1, i2, i3)) P(isp, i1, i2, i3) = 100.0;
  } else {
    return -1;
  }
  /* ghost slabs (paramflag 1 and 3): mindensdiv or 100 */
  const double gv = iparm == 1 ? mindensdiv : 100.0;
  for (int isp = 1; isp <= LSP; ++isp)
    for (int i3 = -1; i3 <= lx3 + 2; ++i3)
      for (int i2 = -1; i2 <= lx2 + 2; ++i2)
        for (int i1 = -1; i1 <= lx1 + 2; ++i1)
          if (i1 < 1 || i1 > lx1 || i2 < 1 || i2 > lx2 || i3 < 1 || i3 > lx3) P(isp, i1, i2, i3) = gv;
#undef P
#undef NUL
  return 0;
}

/* ---- multifluid.f90:237-268 VNRicht_artvisc (xicon = 3, multifluid.f90:43):
   Q(lx1,lx2,lx3,7) interior array */
void orc_VNRicht_artvisc(const orc_grid *g, const double *ns, const double *vs1, double *Q) {
  dims_t d = dims_of(g);
  const int lx1 = g->lx1, lx2 = g->lx2, lx3 = g->lx3;
  const long ni = (long)lx1 * lx2 * lx3;
  const double xicon = 3.0;
  for (int isp = 1; isp <= LSP; ++isp)
    for (int i3 = 1; i3 <= lx3; ++i3)
      for (int i2 = 1; i2 <= lx2; ++i2)
        for (int i1 = 1; i1 <= lx1; ++i1) {
          const long k = (i1 - 1) + (long)lx1 * ((i2 - 1) + (long)lx2 * (i3 - 1)) + (isp - 1) * ni;
          if (isp == LSP) { Q[k] = 0.0; continue; }
          const double *u = vs1 + (long)(isp - 1) * d.N;
          const double vl = 0.5 * (u[GIX(d, i1 - 1, i2, i3)] + u[GIX(d, i1, i2, i3)]);
          const double vr = 0.5 * (u[GIX(d, i1, i2, i3)] + u[GIX(d, i1 + 1, i2, i3)]);
          const double m = fmin_f(vr - vl, 0.0);
          Q[k] = (((ns[(long)(isp - 1) * d.N + GIX(d, i1, i2, i3)] * ms_(isp)) * 0.25) *
                  (xicon * xicon)) * (m * m);
        }
}

/* ---- sources_mpi.f90:55-112 RK2_global_boundary_allspec on one process: zero-order hold
   of vs1 into the x1 ghosts 0 and lx1+1, then of vs1, vs2, vs3 into the x2 ghosts and
   (unless periodic) the x3 ghosts, each over the whole extent of the other indices */
void orc_RK2_global_boundary_allspec(const orc_grid *g, double *vs1, double *vs2, double *vs3) {
  dims_t d = dims_of(g);
  const int lx1 = g->lx1, lx2 = g->lx2, lx3 = g->lx3;
  for (int isp = 1; isp <= LSP; ++isp) {
    double *f[3] = {vs1 + (long)(isp - 1) * d.N, vs2 + (long)(isp - 1) * d.N,
                    vs3 + (long)(isp - 1) * d.N};
    for (int i3 = -1; i3 <= lx3 + 2; ++i3)
      for (int i2 = -1; i2 <= lx2 + 2; ++i2) {
        f[0][GIX(d, 0, i2, i3)] = f[0][GIX(d, 1, i2, i3)];
        f[0][GIX(d, lx1 + 1, i2, i3)] = f[0][GIX(d, lx1, i2, i3)];
      }
    for (int q = 0; q < 3; ++q)
      for (int i3 = -1; i3 <= lx3 + 2; ++i3)
        for (int i1 = -1; i1 <= lx1 + 2; ++i1) f[q][GIX(d, i1, 0, i3)] = f[q][GIX(d, i1, 1, i3)];
    for (int q = 0; q < 3; ++q)
      for (int i3 = -1; i3 <= lx3 + 2; ++i3)
        for (int i1 = -1; i1 <= lx1 + 2; ++i1)
          f[q][GIX(d, i1, lx2 + 1, i3)] = f[q][GIX(d, i1, lx2, i3)];
    if (!g->flagper) {
      for (int q = 0; q < 3; ++q)
        for (int i2 = -1; i2 <= lx2 + 2; ++i2)
          for (int i1 = -1; i1 <= lx1 + 2; ++i1) f[q][GIX(d, i1, i2, 0)] = f[q][GIX(d, i1, i2, 1)];
      for (int q = 0; q < 3; ++q)
        for (int i2 = -1; i2 <= lx2 + 2; ++i2)
          for (int i1 = -1; i1 <= lx1 + 2; ++i1)
            f[q][GIX(d, i1, i2, lx3 + 1)] = f[q][GIX(d, i1, i2, lx3)];
    }
  }
}

/* ---- multifluid.f90:274-302 compression with div3D_curv_23 (div.f90:118-206) over
   (0:lx1+1, 0:lx2+1, 0:lx3+1), of which the interior, centred differences, is used */
void orc_compression(const orc_grid *g, double dt, const double *vs1, const double *vs2,
                     const double *vs3, const double *Q, double *rhoes) {
  dims_t d = dims_of(g);
  const int lx1 = g->lx1, lx2 = g->lx2, lx3 = g->lx3;
  const long ni = (long)lx1 * lx2 * lx3;
  const double *h1 = g->h1, *h2 = g->h2, *h3 = g->h3;
  for (int isp = 1; isp <= LSP; ++isp) {
    const long o = (long)(isp - 1) * d.N;
    const double gm1 = gammas_(isp) - 1.0;
    for (int i3 = 1; i3 <= lx3; ++i3)
      for (int i2 = 1; i2 <= lx2; ++i2)
        for (int i1 = 1; i1 <= lx1; ++i1) {
          const long c = GIX(d, i1, i2, i3);
          long a = GIX(d, i1 + 1, i2, i3), b = GIX(d, i1 - 1, i2, i3);
          double dv = ((h2[a] * h3[a]) * vs1[o + a] - (h2[b] * h3[b]) * vs1[o + b]) /
                      (g->dx1[i1 + 1] + g->dx1[i1]);
          if (lx2 > 1) {
            a = GIX(d, i1, i2 + 1, i3); b = GIX(d, i1, i2 - 1, i3);
            dv = dv + ((h1[a] * h3[a]) * vs2[o + a] - (h1[b] * h3[b]) * vs2[o + b]) /
                          (g->dx2[i2 + 1] + g->dx2[i2]);
          }
          if (lx3 > 1) {
            a = GIX(d, i1, i2, i3 + 1); b = GIX(d, i1, i2, i3 - 1);
            dv = dv + ((h1[a] * h2[a]) * vs3[o + a] - (h1[b] * h2[b]) * vs3[o + b]) /
                          (g->dx3[i3 + 1] + g->dx3[i3]);
          }
          dv = dv / ((h1[c] * h2[c]) * h3[c]);
          const double q = Q[(i1 - 1) + (long)lx1 * ((i2 - 1) + (long)lx2 * (i3 - 1)) + (isp - 1) * ni];
          const double e0 = rhoes[o + c];
          const double half = e0 - ((dt / 2.0) * (e0 * gm1 + q)) * dv;
          rhoes[o + c] = e0 - (dt * (half * gm1 + q)) * dv;
        }
  }
}

/* ---- temporal_mpi.f90:86-165 dt_calc: per species the largest grid rates
   (vsnd + |vs1|)/dl1, |vs2|/dl2, |vs3|/dl3 over the interior, dl_k = dxki * h_k
   (meshobj.f90:436-442); rates[3*(isp-1)+k].  Returns dt = tcfl * min(1/rate). */
double orc_dt_calc(const orc_grid *g, const double *Ts, const double *vs1, const double *vs2,
                   const double *vs3, double tcfl, double *rates) {
  dims_t d = dims_of(g);
  const int lx1 = g->lx1, lx2 = g->lx2, lx3 = g->lx3;
  for (int k = 0; k < 3 * LSP; ++k) rates[k] = 0.0;
  for (int isp = 1; isp <= LSP; ++isp) {
    const long o = (long)(isp - 1) * d.N, oe = (long)(LSP - 1) * d.N;
    double r1 = 0.0, r2 = 0.0, r3 = 0.0;
    for (int i3 = 1; i3 <= lx3; ++i3)
      for (int i2 = 1; i2 <= lx2; ++i2)
        for (int i1 = 1; i1 <= lx1; ++i1) {
          const long c = GIX(d, i1, i2, i3);
          double vsnd = 0.0;
          if (isp < LSP)
            vsnd = sqrt((kB * Ts[o + c]) / ms_(isp) + (((5.0 / 3.0) * kB) * Ts[oe + c]) / ms_(isp));
          r1 = fmax_f((vsnd + fabs(vs1[o + c])) / (g->dx1i[i1 - 1] * g->h1[c]), r1);
          r2 = fmax_f(fabs(vs2[o + c]) / (g->dx2i[i2 - 1] * g->h2[c]), r2);
          r3 = fmax_f(fabs(vs3[o + c]) / (g->dx3i[i3 - 1] * g->h3[c]), r3);
        }
    rates[3 * (isp - 1)] = r1; rates[3 * (isp - 1) + 1] = r2; rates[3 * (isp - 1) + 2] = r3;
  }
  double m = 0.0;
  for (int k = 0; k < 3; ++k) {               /* min(minval(1/gridrate1), minval(..2), ..3) */
    double mk = 0.0;
    for (int isp = 1; isp <= LSP; ++isp) {
      const double r = fmax_f(rates[3 * (isp - 1) + k], 1e-10);
      const double inv = 1.0 / r;
      if (isp == 1 || inv < mk) mk = inv;
    }
    if (k == 0 || mk < m) m = mk;
  }
  return tcfl * m;
}

/* ---- libgemini_mpi.f90:450-483 dt_select with temporal_mpi.f90:27-83 dt_comm on one
   process: dt_calc, the output-time clamps, the 1e-6 floor, then from the second
   iteration on the growth limit dt <= 2 dtprev.  dt is in (previous) and out. */
void orc_dt_select(const orc_grid *g, const double *Ts, const double *vs1, const double *vs2,
                   const double *vs3, double tcfl, int flagglow, double t, double tout,
                   double tglowout, int it, double *dt) {
  double rates[3 * LSP];
  const double dtprev = *dt;
  double v = orc_dt_calc(g, Ts, vs1, vs2, vs3, tcfl, rates);
  if (flagglow != 0 && t + v > tglowout) v = tglowout - t;
  if (t + v > tout) v = tout - t;
  v = fmax_f(v, 1e-6);
  if (it > 1 && v / dtprev > 2.0) v = 2.0 * dtprev;
  *dt = v;
}

/* ---- the transport block of src/libgemini.cpp:155-209 (app/main.f90:277-341): the
   advection step above, then clean_param 1 and 2, VNRicht_artvisc, halo_fluidvars,
   RK2_global_boundary_allspec, compression, rhoe2T, clean_param 3, energy_diffusion,
   clean_param 3.  Q(lx1,lx2,lx3,7) is the caller's workspace (intvars%Q). */
int orc_fluid_adv_block(const orc_grid *g, double dt, double *fluidvars, double *fluidauxvars,
                        const double *electrovars, double *vs1i, double *vs2i, double *vs3i,
                        const double *nn, const double *Tn, int diffsolvetype, double Teinf,
                        const int *nullpts, double *Q) {
  dims_t d = dims_of(g);
  double *ns = fluidvars, *vs1 = fluidvars + 7 * d.N, *vs2 = fluidvars + 14 * d.N,
         *vs3 = fluidvars + 21 * d.N, *Ts = fluidvars + 28 * d.N;
  double *rhovs1 = fluidauxvars, *rhoes = fluidauxvars + 7 * d.N;
  const double *J1 = electrovars + 3 * d.N;
  orc_v12rhov1(g, ns, vs1, rhovs1);
  orc_T2rhoe(g, ns, Ts, rhoes);
  orc_set_global_boundaries_allspec(g, ns, rhovs1, vs1, vs2, vs3, rhoes, vs1i, LSP);
  orc_halo_periodic(g, ns, LSP, 2);
  orc_halo_periodic(g, rhovs1, LSP, 2);
  orc_halo_periodic(g, rhoes, LSP, 2);
  orc_halo_periodic(g, vs2, LSP, 2);
  orc_halo_periodic(g, vs3, LSP, 2);
  orc_interface_vels_allspec(g, vs1, vs2, vs3, vs1i, vs2i, vs3i, LSP);
  orc_sweep_allspec(g, 3, ns, vs3i, dt, 0, 6);
  orc_sweep_allspec(g, 3, rhovs1, vs3i, dt, 1, 6);
  orc_sweep_allspec(g, 3, rhoes, vs3i, dt, 0, 7);
  orc_sweep_allspec(g, 1, ns, vs1i, dt, 0, 6);
  orc_sweep_allspec(g, 1, rhovs1, vs1i, dt, 0, 6);
  orc_sweep_allspec(g, 1, rhoes, vs1i, dt, 0, 7);
  orc_halo_periodic(g, ns, LSP, 2);
  orc_halo_periodic(g, rhovs1, LSP, 2);
  orc_halo_periodic(g, rhoes, LSP, 2);
  orc_sweep_allspec(g, 2, ns, vs2i, dt, 0, 6);
  orc_sweep_allspec(g, 2, rhovs1, vs2i, dt, 1, 6);
  orc_sweep_allspec(g, 2, rhoes, vs2i, dt, 0, 7);
  orc_rhov12v1(g, ns, rhovs1, vs1);
  orc_clean_param(g, 1, ns, nullpts);
  orc_clean_param(g, 2, vs1, nullpts);
  orc_VNRicht_artvisc(g, ns, vs1, Q);
  orc_halo_periodic(g, ns, LSP, 2);
  orc_halo_periodic(g, rhovs1, LSP, 2);
  orc_halo_periodic(g, rhoes, LSP, 2);
  orc_halo_periodic(g, vs2, LSP, 2);
  orc_halo_periodic(g, vs3, LSP, 2);
  orc_RK2_global_boundary_allspec(g, vs1, vs2, vs3);
  orc_compression(g, dt, vs1, vs2, vs3, Q, rhoes);
  orc_rhoe2T(g, ns, rhoes, Ts);
  orc_clean_param(g, 3, Ts, nullpts);
  int rc = orc_energy_diffusion(g, dt, ns, Ts, J1, nn, Tn, diffsolvetype, Teinf);
  orc_clean_param(g, 3, Ts, nullpts);
  return rc;
}

/* ---- one transport step in the order of src/libgemini.cpp:155-186 and :208
   (app/main.f90:277-355): v12rhov1, T2rhoe, set_global_boundaries_allspec,
   halo_fluidvars, interface_vels_allspec, sweep3 (mass, momentum, energy), sweep1,
   halo_allparams, sweep2, rhov12v1, then rhoe2T (multifluid.f90:179-189) to bring the
   advected energy back to temperature, then energy_diffusion.  The compression,
   artificial-viscosity and clean_param substeps between them are outside this path.
   Block layout: fluidvars(...,35) = ns 1-7, vs1 8-14, vs2 15-21, vs3 22-28, Ts 29-35;
   fluidauxvars(...,>=14) = rhovs1 1-7, rhoes 8-14; electrovars(...,4) = J1. */
int orc_transport_step_ext(const orc_grid *g, double dt, double *fluidvars,
                           double *fluidauxvars, const double *electrovars, double *vs1i,
                           double *vs2i, double *vs3i, const double *nn, const double *Tn,
                           int diffsolvetype, double Teinf, const double *kappa,
                           const double *qtop) {
  dims_t d = dims_of(g);
  double *ns = fluidvars, *vs1 = fluidvars + 7 * d.N, *vs2 = fluidvars + 14 * d.N,
         *vs3 = fluidvars + 21 * d.N, *Ts = fluidvars + 28 * d.N;
  double *rhovs1 = fluidauxvars, *rhoes = fluidauxvars + 7 * d.N;
  const double *J1 = electrovars + 3 * d.N;
  orc_v12rhov1(g, ns, vs1, rhovs1);
  orc_T2rhoe(g, ns, Ts, rhoes);
  orc_set_global_boundaries_allspec(g, ns, rhovs1, vs1, vs2, vs3, rhoes, vs1i, LSP);
  orc_halo_periodic(g, ns, LSP, 2);
  orc_halo_periodic(g, rhovs1, LSP, 2);
  orc_halo_periodic(g, rhoes, LSP, 2);
  orc_halo_periodic(g, vs2, LSP, 2);
  orc_halo_periodic(g, vs3, LSP, 2);
  orc_interface_vels_allspec(g, vs1, vs2, vs3, vs1i, vs2i, vs3i, LSP);
  orc_sweep_allspec(g, 3, ns, vs3i, dt, 0, 6);
  orc_sweep_allspec(g, 3, rhovs1, vs3i, dt, 1, 6);
  orc_sweep_allspec(g, 3, rhoes, vs3i, dt, 0, 7);
  orc_sweep_allspec(g, 1, ns, vs1i, dt, 0, 6);
  orc_sweep_allspec(g, 1, rhovs1, vs1i, dt, 0, 6);
  orc_sweep_allspec(g, 1, rhoes, vs1i, dt, 0, 7);
  orc_halo_periodic(g, ns, LSP, 2);
  orc_halo_periodic(g, rhovs1, LSP, 2);
  orc_halo_periodic(g, rhoes, LSP, 2);
  orc_sweep_allspec(g, 2, ns, vs2i, dt, 0, 6);
  orc_sweep_allspec(g, 2, rhovs1, vs2i, dt, 1, 6);
  orc_sweep_allspec(g, 2, rhoes, vs2i, dt, 0, 7);
  orc_rhov12v1(g, ns, rhovs1, vs1);
  orc_rhoe2T(g, ns, rhoes, Ts);
  return energy_diffusion_core(g, dt, ns, Ts, J1, nn, Tn, diffsolvetype, Teinf, kappa, qtop);
}

int orc_transport_step(const orc_grid *g, double dt, double *fluidvars, double *fluidauxvars,
                       const double *electrovars, double *vs1i, double *vs2i, double *vs3i,
                       const double *nn, const double *Tn, int diffsolvetype, double Teinf) {
  return orc_transport_step_ext(g, dt, fluidvars, fluidauxvars, electrovars, vs1i, vs2i, vs3i,
                                nn, Tn, diffsolvetype, Teinf, 0, 0);
}

/* =======================================================================================
 * Source/loss step (multifluid.f90:130-175, 472-549; sources.f90; collisions.f90:58-159;
 * calculus.in.f90:174-196; gradient.f90:79-149).  Pr/Lo are (lx1,lx2,lx3,7) interior
 * arrays; the atmosphere and ionization rates are interior arrays (orc_atmos).
 * ===================================================================================== */

/* collisions.f90:11-47.  [real(wp) :: 6.82e-10, ...] converts default-real (single
   precision) literals, so every coefficient is the float nearest the decimal. */
static const float Csn_f[LSP][4] = {
    {-1, 6.82e-10f, 6.64e-10f, -1},        {2.44e-10f, 4.34e-10f, 4.27e-10f, 0.69e-10f},
    {2.58e-10f, -1, 4.49e-10f, 0.74e-10f}, {2.31e-10f, 4.13e-10f, -1, 0.65e-10f},
    {4.42e-10f, 7.47e-10f, 7.25e-10f, 1.45e-10f}, {-1, 33.6e-10f, 32.0e-10f, -1},
    {-1, -1, -1, -1}};
static const float C2sn1_f[LSP][4] = {
    {3.67e-11f, 0, 0, 4.63e-12f}, {0, 0, 0, 0}, {0, 5.14e-11f, 0, 0}, {0, 0, 2.59e-11f, 0},
    {0, 0, 0, 0}, {6.61e-11f, 0, 0, 2.65e-10f}, {-1, -1, -1, -1}};
static const float C2sn2_f[LSP][4] = {
    {0.064f, 0, 0, -1}, {0, 0, 0, 0}, {0, 0.069f, 0, 0}, {0, 0, 0.073f, 0},
    {0, 0, 0, 0}, {0.047f, 0, 0, 0.083f}, {-1, -1, -1, -1}};
static const float Csj_f[LSP][LSP] = {
    {0.22f, 0.26f, 0.25f, 0.26f, 0.22f, 0.077f, 1.87e-3f},
    {0.14f, 0.16f, 0.16f, 0.17f, 0.13f, 0.042f, 9.97e-4f},
    {0.15f, 0.17f, 0.17f, 0.18f, 0.14f, 0.045f, 1.07e-3f},
    {0.13f, 0.16f, 0.15f, 0.16f, 0.12f, 0.039f, 9.347e-4f},
    {0.25f, 0.28f, 0.28f, 0.28f, 0.24f, 0.088f, 2.136e-3f},
    {1.23f, 1.25f, 1.25f, 1.25f, 1.23f, 0.90f, 29.7e-3f},
    {54.5f, 54.5f, 54.5f, 54.5f, 54.5f, 54.5f, 38.537f}};
static double mn_(int isp2) {              /* phys_consts.in.f90:60 [16,28,32,1]*amu */
  static const double m[4] = {16.0, 28.0, 32.0, 1.0};
  return m[isp2 - 1] * amu;
}
static const double pi_ = 3.141592653589793;   /* 4*atan(1.0_wp), phys_consts.in.f90:20 */

/* ---- calculus.in.f90:174-196 ETD_uncoupled, one element ---- */
double orc_ETD_uncoupled(double f, double P, double L, double dt) {
  const double Ldt = L * dt;
  const double expL = exp(-1.0 * Ldt);
  if (Ldt > 1e-10) return f * expL + (P / L) * (1.0 - expL);
  return f + P * dt;
}

/* ---- collisions.f90:58-114 maxwell_colln at one cell: nn2 = nn(isp2), Tsp = Ts(isp) ---- */
double orc_maxwell_colln(int isp, int isp2, double nn2, double Tn, double Tsp) {
  if (isp < LSP) {
    const double csn = (double)Csn_f[isp - 1][isp2 - 1];
    if (csn < 0) {
      const double c1 = (double)C2sn1_f[isp - 1][isp2 - 1];
      if (isp == 1 && isp2 == 4) {
        const double Teff = Tn + Tsp / 16.0;
        return ((c1 * pow(Teff, 0.5)) * nn2) * 1e-6;
      }
      const double Teff = 0.5 * (Tn + Tsp);
      const double b = 1.0 - (double)C2sn2_f[isp - 1][isp2 - 1] * log10(Teff);
      return (((c1 * (b * b)) * pow(Teff, 0.5)) * nn2) * 1e-6;
    }
    return (csn * nn2) * 1e-6;
  }
  const double T = Tsp;
  switch (isp2) {
    case 1: return (((8.9e-11 * (1.0 + 5.7e-4 * T)) * pow(T, 0.5)) * nn2) * 1e-6;
    case 2: return (((2.33e-11 * (1.0 - 1.21e-4 * T)) * T) * nn2) * 1e-6;
    case 3: return (((1.82e-10 * (1.0 + 3.6e-2 * pow(T, 0.5))) * pow(T, 0.5)) * nn2) * 1e-6;
    default: return (((4.5e-9 * (1.0 - 1.35e-4 * T)) * pow(T, 0.5)) * nn2) * 1e-6;
  }
}

/* ---- collisions.f90:117-159 coulomb_colln at one cell: out = {nusj, Phisj, Psisj} ---- */
void orc_coulomb_colln(int isp, int isp2, double ns2, double Ts1, double Ts2, double v1,
                       double v2, double *out) {
  if (isp == isp2) { out[0] = out[1] = out[2] = 0.0; return; }
  const double m1 = ms_(isp), m2 = ms_(isp2);
  const double Teff = (m2 * Ts1 + m1 * Ts2) / (m2 + m1);
  out[0] = (((double)Csj_f[isp - 1][isp2 - 1] * ns2) * 1e-6) / pow(Teff, 1.5);
  const double mred = (m1 * m2) / (m1 + m2);
  const double W = fabs(v1 - v2) / sqrt(((2.0 * kB) * Teff) / mred);
  const double Psi = exp(-(W * W));
  out[2] = Psi;
  if (W < 0.1) out[1] = 1.0;
  else out[1] = (((3.0 / 4.0) * sqrt(pi_)) * erf(W)) / ((W * W) * W) - ((3.0 / 2.0) / (W * W)) * Psi;
}

#define IIX(i1, i2, i3) ((long)((i1) - 1) + (long)lx1 * ((long)((i2) - 1) + (long)lx2 * (long)((i3) - 1)))

/* ---- sources.f90:292-374 srcsEnergy ---- */
void orc_srcs_energy(const orc_grid *g, const orc_atmos *a, const double *ns, const double *vs1,
                     const double *vs2, const double *vs3, const double *Ts, double *Pr,
                     double *Lo) {
  dims_t d = dims_of(g);
  const int lx1 = g->lx1, lx2 = g->lx2, lx3 = g->lx3;
  const long ni = (long)lx1 * lx2 * lx3;
  const double sfact = (elchrg / kB) * (gammas_(LSP) - 1.0);
#pragma omp parallel for collapse(2) schedule(static)
  for (int i3 = 1; i3 <= lx3; ++i3)
    for (int i2 = 1; i2 <= lx2; ++i2)
      for (int i1 = 1; i1 <= lx1; ++i1) {
        const long c = GIX(d, i1, i2, i3), k = IIX(i1, i2, i3);
        const double Tn = a->Tn[k], vn1 = a->vn1[k], vn2 = a->vn2[k], vn3 = a->vn3[k];
#define S(arr, isp) (arr)[(long)((isp) - 1) * d.N + c]
        for (int isp = 1; isp <= LSP; ++isp) {
          const double n = S(ns, isp), m = ms_(isp), gm1 = gammas_(isp) - 1.0;
          double P = 0.0, L = 0.0;
          for (int isp2 = 1; isp2 <= 4; ++isp2) {
            const double nu = orc_maxwell_colln(isp, isp2, a->nn[(isp2 - 1) * ni + k], Tn, S(Ts, isp));
            double fact = (2.0 * nu) / (m + mn_(isp2));
            P = P + ((((n * m) * kB) / gm1) * fact) * Tn;
            L = L + m * fact;
            fact = (fact * mn_(isp2)) / 3.0;
            const double d1 = S(vs1, isp) - vn1, d2 = S(vs2, isp) - vn2, d3 = S(vs3, isp) - vn3;
            P = P + (((n * m) / gm1) * (((d1 * d1) + (d2 * d2)) + (d3 * d3))) * fact;
          }
          for (int isp2 = 1; isp2 <= LSP; ++isp2) {
            double cc[3];
            orc_coulomb_colln(isp, isp2, S(ns, isp2), S(Ts, isp), S(Ts, isp2), S(vs1, isp),
                              S(vs1, isp2), cc);
            const double m2 = ms_(isp2);
            double fact = ((2.0 * cc[0]) * cc[2]) / (m + m2);
            P = P + ((((n * m) * kB) / gm1) * fact) * S(Ts, isp2);
            L = L + m * fact;
            fact = ((((2.0 * cc[0]) * cc[1]) / (m + m2)) * m2) / 3.0;
            const double d1 = S(vs1, isp) - S(vs1, isp2), d2 = S(vs2, isp) - S(vs2, isp2),
                         d3 = S(vs3, isp) - S(vs3, isp2);
            P = P + (((n * m) / gm1) * (((d1 * d1) + (d2 * d2)) + (d3 * d3))) * fact;
          }
          Pr[(isp - 1) * ni + k] = P;
          Lo[(isp - 1) * ni + k] = L;
        }
        /* inelastic electron collisions, sources.f90:350-373 */
        const double Te = S(Ts, LSP), nn2 = a->nn[ni + k], nn3 = a->nn[2 * ni + k];
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
        const long ke = (LSP - 1) * ni + k;
        Pr[ke] = Pr[ke] + (((iePT * S(ns, LSP)) * kB) / (gammas_(LSP) - 1.0));
        Lo[ke] = Lo[ke] + ieLT;
#undef S
      }
}

/* gradient.f90:79-149 grad3D1_curv_23 over the interior (bounds 1..lx1): one-sided at the
   ends with dx1(1) and dx1(lx1), centred between.  f and out are (lx1,lx2,lx3). */
void orc_grad3D1_interior(const orc_grid *g, const double *f, double *out) {
  dims_t d = dims_of(g);
  const int lx1 = g->lx1, lx2 = g->lx2, lx3 = g->lx3;
  const double *dx1 = g->dx1, *h1 = g->h1;
  for (int i3 = 1; i3 <= lx3; ++i3)
    for (int i2 = 1; i2 <= lx2; ++i2) {
      out[IIX(1, i2, i3)] = ((f[IIX(2, i2, i3)] - f[IIX(1, i2, i3)]) / dx1[1]) / h1[GIX(d, 1, i2, i3)];
      for (int i1 = 2; i1 <= lx1 - 1; ++i1)
        out[IIX(i1, i2, i3)] = ((f[IIX(i1 + 1, i2, i3)] - f[IIX(i1 - 1, i2, i3)]) /
                                (dx1[i1 + 1] + dx1[i1])) / h1[GIX(d, i1, i2, i3)];
      out[IIX(lx1, i2, i3)] = ((f[IIX(lx1, i2, i3)] - f[IIX(lx1 - 1, i2, i3)]) / dx1[lx1]) /
                              h1[GIX(d, lx1, i2, i3)];
    }
}

/* centred x1 derivative at interior cell i1 of a ghosted field (grad3D1 over 0..lx1+1) */
static double dx1_centred(const orc_grid *g, dims_t d, const double *f, int i1, int i2, int i3) {
  return ((f[GIX(d, i1 + 1, i2, i3)] - f[GIX(d, i1 - 1, i2, i3)]) / (g->dx1[i1 + 1] + g->dx1[i1])) /
         g->h1[GIX(d, i1, i2, i3)];
}

/* ---- sources.f90:228-289 srcsMomentum_curv (ions; the electron row is not used) ---- */
void orc_srcs_momentum(const orc_grid *g, const orc_atmos *a, const double *ns, const double *vs1,
                       const double *vs2, const double *vs3, const double *Ts, const double *E1,
                       const double *Q, double *Pr, double *Lo) {
  dims_t d = dims_of(g);
  const int lx1 = g->lx1, lx2 = g->lx2, lx3 = g->lx3;
  const long ni = (long)lx1 * lx2 * lx3;
  const int ghosted = !(g->isglobal[0] || g->isglobal[1]);
  /* log pressure of every species: ghosted (x1 ghosts 0 and lx1+1 used) or interior */
  double *lp = malloc(sizeof(double) * (ghosted ? d.N : ni) * LSP);
  double *glp = malloc(sizeof(double) * ni * LSP);
  double *gQ = malloc(sizeof(double) * ni * LSP);
  for (int isp = 1; isp <= LSP; ++isp) {
    const long o = (long)(isp - 1) * d.N;
    double *l = lp + (isp - 1) * (ghosted ? d.N : ni), *gl = glp + (isp - 1) * ni;
    for (int i3 = 1; i3 <= lx3; ++i3)
      for (int i2 = 1; i2 <= lx2; ++i2)
        for (int i1 = ghosted ? 0 : 1; i1 <= (ghosted ? lx1 + 1 : lx1); ++i1) {
          const long c = GIX(d, i1, i2, i3);
          const double p = (ns[o + c] * kB) * Ts[o + c];
          l[ghosted ? c : IIX(i1, i2, i3)] = log(p);
        }
    if (ghosted) {
      for (int i3 = 1; i3 <= lx3; ++i3)
        for (int i2 = 1; i2 <= lx2; ++i2)
          for (int i1 = 1; i1 <= lx1; ++i1) gl[IIX(i1, i2, i3)] = dx1_centred(g, d, l, i1, i2, i3);
    } else {
      orc_grad3D1_interior(g, l, gl);
    }
    if (isp < LSP) orc_grad3D1_interior(g, Q + (isp - 1) * ni, gQ + (isp - 1) * ni);
  }
#pragma omp parallel for collapse(2) schedule(static)
  for (int i3 = 1; i3 <= lx3; ++i3)
    for (int i2 = 1; i2 <= lx2; ++i2)
      for (int i1 = 1; i1 <= lx1; ++i1) {
        const long c = GIX(d, i1, i2, i3), k = IIX(i1, i2, i3);
#define S(arr, isp) (arr)[(long)((isp) - 1) * d.N + c]
        const double h1h2h3 = (g->h1[c] * g->h2[c]) * g->h3[c];
        const double dh2dx1 = dx1_centred(g, d, g->h2, i1, i2, i3);
        const double dh3dx1 = dx1_centred(g, d, g->h3, i1, i2, i3);
        const double Te = S(Ts, LSP);
        const double Epol1 = ((kB * Te) / qs_(LSP)) * glp[(LSP - 1) * ni + k];
        const double E1f = a->alt[c] < 100e3 ? 0.0 : E1[c];
        const double Tn = a->Tn[k];
        for (int isp = 1; isp <= LSP - 1; ++isp) {
          const double n = S(ns, isp), m = ms_(isp);
          double P = 0.0, L = 0.0;
          for (int isp2 = 1; isp2 <= 4; ++isp2) {
            const double nu = orc_maxwell_colln(isp, isp2, a->nn[(isp2 - 1) * ni + k], Tn, S(Ts, isp));
            L = L + nu;
            P = P + ((n * m) * nu) * a->vn1[k];
          }
          for (int isp2 = 1; isp2 <= LSP; ++isp2) {
            double cc[3];
            orc_coulomb_colln(isp, isp2, S(ns, isp2), S(Ts, isp), S(Ts, isp2), S(vs1, isp),
                              S(vs1, isp2), cc);
            L = L + cc[0] * cc[1];
            P = P + (((n * m) * cc[0]) * cc[1]) * S(vs1, isp2);
          }
          const double pr = (n * kB) * S(Ts, isp);
          const double ionpress = pr * glp[(isp - 1) * ni + k];
          const double v2 = S(vs2, isp), v3 = S(vs3, isp);
          const double geom = (((((v2 * v2) * g->h3[c]) * dh2dx1 + ((v3 * v3) * g->h2[c]) * dh3dx1) * n) * m) / h1h2h3;
          P = ((((P + (n * qs_(isp)) * (E1f + Epol1)) - ionpress) - gQ[(isp - 1) * ni + k]) + geom) +
              (n * m) * a->g1[k];
          Pr[(isp - 1) * ni + k] = P;
          Lo[(isp - 1) * ni + k] = L;
        }
        Pr[(LSP - 1) * ni + k] = 0.0;
        Lo[(LSP - 1) * ni + k] = 0.0;
#undef S
      }
  free(lp); free(glp); free(gQ);
}

/* kreac polynomial c0 + c1 t + c2 t^2 + c3 t^3 + c4 t^4 with the reference's signs folded
   into the coefficients' use: ((((c0 s1 c1*t) s2 c2*t^2) s3 c3*t^3) s4 c4*t^4) */
static double poly4(double t, double c0, double c1, double c2, double c3, double c4, int s1,
                    int s2, int s3, int s4) {
  const double t2 = t * t, t3 = t2 * t, t4 = t2 * t2;
  double r = c0;
  r = s1 > 0 ? r + c1 * t : r - c1 * t;
  r = s2 > 0 ? r + c2 * t2 : r - c2 * t2;
  r = s3 > 0 ? r + c3 * t3 : r - c3 * t3;
  r = s4 > 0 ? r + c4 * t4 : r - c4 * t4;
  return r;
}

/* ---- sources.f90:18-225 srcsContinuity ---- */
void orc_srcs_continuity(const orc_grid *g, const orc_atmos *a, const double *ns,
                         const double *vs1, const double *vs2, const double *vs3,
                         const double *Ts, double *Pr, double *Lo) {
  dims_t d = dims_of(g);
  const int lx1 = g->lx1, lx2 = g->lx2, lx3 = g->lx3;
  const long ni = (long)lx1 * lx2 * lx3;
  const double e07 = (double)0.7f, e085 = (double)0.85f, e039 = (double)0.39f;
  const double e023 = (double)0.23f, em041 = -(double)0.41f, e044 = (double)0.44f;
  const double em02 = -(double)0.2f, e070 = (double)0.70f, e056 = (double)0.56f;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i3 = 1; i3 <= lx3; ++i3)
    for (int i2 = 1; i2 <= lx2; ++i2)
      for (int i1 = 1; i1 <= lx1; ++i1) {
        const long c = GIX(d, i1, i2, i3), k = IIX(i1, i2, i3);
#define S(arr, isp) (arr)[(long)((isp) - 1) * d.N + c]
#define NN(j) a->nn[(long)((j) - 1) * ni + k]
        double P[LSP + 1] = {0}, L[LSP + 1] = {0};
        const double Tn = a->Tn[k], Te = S(Ts, LSP), ne = S(ns, LSP);
        const double d1 = S(vs1, 1) - a->vn1[k], d2 = S(vs2, 1) - a->vn2[k], d3 = S(vs3, 1) - a->vn3[k];
        const double dv2 = ((d1 * d1) + (d2 * d2)) + (d3 * d3);
        double Teff, kreac = 0.0, b;
        /* O+ + N2 -> NO+ + N */
        Teff = (28.0 / (16.0 + 28.0)) * (((((16.0 * amu) / 3.0) / kB) * dv2 + S(Ts, 1)) - Tn) + Tn;
        if (Teff <= 3725.0)
          kreac = poly4(Teff / 300.0, 1.71676e-12, 7.19934e-13, 1.33276e-13, 9.28213e-15, 6.39557e-16, -1, 1, -1, 1);
        if (Teff > 3725.0 && Teff <= 30000.0)
          kreac = poly4(Teff / 300.0, -1.52489e-11, 7.67112e-13, 1.19064e-13, 1.30858e-15, 4.67756e-18, 1, 1, -1, 1);
        if (Teff > 30000.0)
          kreac = poly4(100.0, -1.52489e-11, 7.67112e-13, 1.19064e-13, 1.30858e-15, 4.67756e-18, 1, 1, -1, 1);
        b = (kreac * NN(2)) * 1e-6;
        P[2] = P[2] + b * S(ns, 1); L[1] = L[1] + b;
        /* O+ + O2 -> O2+ + O */
        Teff = (32.0 / (16.0 + 32.0)) * (((((16.0 * amu) / 3.0) / kB) * dv2 + S(Ts, 1)) - Tn) + Tn;
        if (Teff <= 4800.0)
          kreac = poly4(Teff / 300.0, 2.78932e-11, 6.92612e-12, 8.67684e-13, 3.47251e-14, 5.07097e-16, -1, 1, -1, 1);
        if (Teff > 4800.0 && Teff <= 30000.0) {
          const double t = Teff / 300.0;
          kreac = ((-1.74046e-11 + (3.02328e-12 * Teff) / 300.0) - 2.39214e-15 * (t * t)) - 4.02394e-17 * ((t * t) * t);
        }
        if (Teff > 30000.0)
          kreac = ((-1.74046e-11 + 3.02328e-12 * 100.0) - 2.39214e-15 * 10000.0) - 4.02394e-17 * 1000000.0;
        b = (kreac * NN(3)) * 1e-6;
        P[4] = P[4] + b * S(ns, 1); L[1] = L[1] + b;
        /* O+ + NO -> NO+ + O */
        Teff = (30.0 / (16.0 + 30.0)) * (((((16.0 * amu) / 3.0) / kB) * dv2 + S(Ts, 1)) - Tn) + Tn;
        if (Teff <= 3800.0)
          kreac = poly4(Teff / 300.0, 6.40408e-13, 1.33888e-13, 7.65103e-14, 3.11509e-15, 6.62374e-17, -1, 1, -1, 1);
        if (Teff > 3800.0 && Teff <= 30000.0)
          kreac = poly4(Teff / 300.0, -7.48312e-13, 2.31502e-13, 3.07160e-14, 2.65436e-16, 7.76665e-19, 1, 1, -1, 1);
        if (Teff > 30000.0)
          kreac = poly4(100.0, -7.48312e-13, 2.31502e-13, 3.07160e-14, 2.65436e-16, 7.76665e-19, 1, 1, -1, 1);
        b = (kreac * NN(6)) * 1e-6;
        P[2] = P[2] + b * S(ns, 1); L[1] = L[1] + b;
        /* O+ + e -> O + hv */
        b = ((3.7e-12 * pow(250.0 / Te, e07)) * ne) * 1e-6;
        L[1] = L[1] + b;
        /* N2+ + O -> O+ + N2 */
        Teff = (16.0 / (28.0 + 16.0)) * (((((28.0 * amu) / 3.0) / kB) * dv2 + S(Ts, 3)) - Tn) + Tn;
        kreac = Teff <= 1500.0 ? 1e-11 * pow(300.0 / Teff, e023) : 3.6e-12 * pow(300.0 / Teff, em041);
        b = (kreac * NN(1)) * 1e-6;
        P[1] = P[1] + b * S(ns, 3); L[3] = L[3] + b;
        /* N+ + O -> O+ + N */
        b = (5e-13 * NN(1)) * 1e-6;
        P[1] = P[1] + b * S(ns, 5); L[5] = L[5] + b;
        /* H+ + O -> O+ + H */
        Teff = S(Ts, 6);
        b = (((6.e-10 * (8.0 / 9.0)) * pow((Teff + Tn / 4.0) / (Tn + Teff / 16.0), 0.5)) * NN(1)) * 1e-6;
        P[1] = P[1] + b * S(ns, 6); L[6] = L[6] + b;
        /* O+ + H -> H+ + O */
        b = (6.0e-10 * NN(4)) * 1e-6;
        P[6] = P[6] + b * S(ns, 1); L[1] = L[1] + b;
        /* O2+ + N2 -> NO+ + NO (the reference takes nn(:,:,:,3)) */
        b = (5e-16 * NN(3)) * 1e-6;
        P[2] = P[2] + b * S(ns, 4); L[4] = L[4] + b;
        /* O2+ + N -> NO+ + O */
        b = (1.2e-10 * NN(5)) * 1e-6;
        P[2] = P[2] + b * S(ns, 4); L[4] = L[4] + b;
        /* O2+ + NO -> NO+ + O2 */
        b = (4.6e-10 * NN(6)) * 1e-6;
        P[2] = P[2] + b * S(ns, 4); L[4] = L[4] + b;
        /* N2+ + O -> NO+ + N */
        Teff = (16.0 / (28.0 + 16.0)) * (((((28.0 * amu) / 3.0) / kB) * dv2 + S(Ts, 3)) - Tn) + Tn;
        kreac = Teff <= 1500.0 ? 1.4e-10 * pow(300.0 / Teff, e044) : 5.2e-11 * pow(300.0 / Teff, em02);
        b = (kreac * NN(1)) * 1e-6;
        P[2] = P[2] + b * S(ns, 3); L[3] = L[3] + b;
        /* N2+ + NO -> NO+ + N2 */
        b = (4.1e-10 * NN(6)) * 1e-6;
        P[2] = P[2] + b * S(ns, 3); L[3] = L[3] + b;
        /* N+ + O2 -> NO+ + O */
        b = (2.6e-10 * NN(3)) * 1e-6;
        P[2] = P[2] + b * S(ns, 5); L[5] = L[5] + b;
        /* NO+ + e -> N + O */
        b = ((4.2e-7 * pow(300.0 / Te, e085)) * ne) * 1e-6;
        L[2] = L[2] + b;
        /* N2+ + O2 -> O2+ + N2 */
        Teff = (32.0 / (28.0 + 32.0)) * (((((28.0 * amu) / 3.0) / kB) * dv2 + S(Ts, 3)) - Tn) + Tn;
        b = ((5e-11 * (300.0 / Teff)) * NN(3)) * 1e-6;
        P[4] = P[4] + b * S(ns, 3); L[3] = L[3] + b;
        /* N2+ + e -> N + N */
        b = ((1.8e-7 * pow(300.0 / Te, e039)) * ne) * 1e-6;
        L[3] = L[3] + b;
        /* N+ + O2 -> O2+ + N */
        b = (3.1e-10 * NN(3)) * 1e-6;
        P[4] = P[4] + b * S(ns, 5); L[5] = L[5] + b;
        /* O2+ + e -> O + O */
        kreac = Te <= 1200.0 ? 1.95e-7 * pow(300.0 / Te, e070) : 7.38e-8 * pow(1200.0 / Te, e056);
        b = (kreac * ne) * 1e-6;
        L[4] = L[4] + b;
        /* N+ + H -> H+ + N */
        b = (3.6e-12 * NN(4)) * 1e-6;
        P[6] = P[6] + b * S(ns, 5); L[5] = L[5] + b;
        /* H+ + e -> H + hv */
        b = ((3.7e-12 * pow(250.0 / Te, e07)) * ne) * 1e-6;
        L[6] = L[6] + b;
        for (int isp = 1; isp <= LSP; ++isp) {
          Pr[(isp - 1) * ni + k] = P[isp];
          Lo[(isp - 1) * ni + k] = L[isp];
        }
#undef NN
#undef S
      }
}

/* ---- multifluid.f90:160-175, 472-495: srcsEnergy, then per species rhoes <- ETD over the
   interior (electrons with Qeionize added to Pr) and Ts <- rhoe2T over the whole padded grid */
void orc_source_loss_energy(const orc_grid *g, const orc_atmos *a, double dt, const double *ns,
                            const double *vs1, const double *vs2, const double *vs3,
                            double *Ts, double *rhoes) {
  dims_t d = dims_of(g);
  const int lx1 = g->lx1, lx2 = g->lx2, lx3 = g->lx3;
  const long ni = (long)lx1 * lx2 * lx3;
  double *Pr = malloc(sizeof(double) * ni * LSP), *Lo = malloc(sizeof(double) * ni * LSP);
  orc_srcs_energy(g, a, ns, vs1, vs2, vs3, Ts, Pr, Lo);
  for (int isp = 1; isp <= LSP; ++isp)
    for (int i3 = 1; i3 <= lx3; ++i3)
      for (int i2 = 1; i2 <= lx2; ++i2)
        for (int i1 = 1; i1 <= lx1; ++i1) {
          const long k = IIX(i1, i2, i3), c = (long)(isp - 1) * d.N + GIX(d, i1, i2, i3);
          double P = Pr[(isp - 1) * ni + k];
          if (isp == LSP) P = P + a->Qeionize[k];
          rhoes[c] = orc_ETD_uncoupled(rhoes[c], P, Lo[(isp - 1) * ni + k], dt);
        }
  orc_rhoe2T(g, ns, rhoes, Ts);
  free(Pr); free(Lo);
}

/* ---- multifluid.f90:144-157, 498-528: srcsMomentum, then rhovs1 <- ETD (ions, interior),
   vs1 <- rhovs1/(ms max(ns, mindensdiv)) and the electron velocity and momentum from the
   ion charge flux, over the whole padded grid */
void orc_source_loss_momentum(const orc_grid *g, const orc_atmos *a, double dt, const double *ns,
                              double *vs1, const double *vs2, const double *vs3, const double *Ts,
                              const double *E1, const double *Q, double *rhovs1) {
  dims_t d = dims_of(g);
  const int lx1 = g->lx1, lx2 = g->lx2, lx3 = g->lx3;
  const long ni = (long)lx1 * lx2 * lx3;
  double *Pr = malloc(sizeof(double) * ni * LSP), *Lo = malloc(sizeof(double) * ni * LSP);
  orc_srcs_momentum(g, a, ns, vs1, vs2, vs3, Ts, E1, Q, Pr, Lo);
  for (int isp = 1; isp <= LSP - 1; ++isp) {
    for (int i3 = 1; i3 <= lx3; ++i3)
      for (int i2 = 1; i2 <= lx2; ++i2)
        for (int i1 = 1; i1 <= lx1; ++i1) {
          const long k = IIX(i1, i2, i3), c = (long)(isp - 1) * d.N + GIX(d, i1, i2, i3);
          rhovs1[c] = orc_ETD_uncoupled(rhovs1[c], Pr[(isp - 1) * ni + k], Lo[(isp - 1) * ni + k], dt);
        }
    const long o = (long)(isp - 1) * d.N;
    for (long k = 0; k < d.N; ++k) vs1[o + k] = rhovs1[o + k] / (ms_(isp) * fmax_f(ns[o + k], mindensdiv));
  }
  const long o7 = (long)(LSP - 1) * d.N;
  for (long k = 0; k < d.N; ++k) {
    double chrgflux = 0.0;
    for (int isp = 1; isp <= LSP - 1; ++isp) {
      const long o = (long)(isp - 1) * d.N + k;
      chrgflux = chrgflux + (ns[o] * qs_(isp)) * vs1[o];
    }
    vs1[o7 + k] = -(((1.0 / fmax_f(ns[o7 + k], mindensdiv)) / qs_(LSP)) * chrgflux);
    rhovs1[o7 + k] = (ns[o7 + k] * ms_(LSP)) * vs1[o7 + k];
  }
  free(Pr); free(Lo);
}

/* ---- multifluid.f90:130-141, 531-549: srcsContinuity, Pr(:,:,:,1:6) += Prionize, ns <-
   ETD (ions, interior), ns(lsp) <- sum of the ions over the whole padded grid */
void orc_source_loss_mass(const orc_grid *g, const orc_atmos *a, double dt, double *ns,
                          const double *vs1, const double *vs2, const double *vs3,
                          const double *Ts) {
  dims_t d = dims_of(g);
  const int lx1 = g->lx1, lx2 = g->lx2, lx3 = g->lx3;
  const long ni = (long)lx1 * lx2 * lx3;
  double *Pr = malloc(sizeof(double) * ni * LSP), *Lo = malloc(sizeof(double) * ni * LSP);
  orc_srcs_continuity(g, a, ns, vs1, vs2, vs3, Ts, Pr, Lo);
  for (int isp = 1; isp <= LSP - 1; ++isp)
    for (int i3 = 1; i3 <= lx3; ++i3)
      for (int i2 = 1; i2 <= lx2; ++i2)
        for (int i1 = 1; i1 <= lx1; ++i1) {
          const long k = IIX(i1, i2, i3), c = (long)(isp - 1) * d.N + GIX(d, i1, i2, i3);
          const double P = Pr[(isp - 1) * ni + k] + a->Prionize[(isp - 1) * ni + k];
          ns[c] = orc_ETD_uncoupled(ns[c], P, Lo[(isp - 1) * ni + k], dt);
        }
  const long o7 = (long)(LSP - 1) * d.N;
  for (long k = 0; k < d.N; ++k) {
    double s = 0.0;
    for (int isp = 1; isp <= LSP - 1; ++isp) s = s + ns[(long)(isp - 1) * d.N + k];
    ns[o7 + k] = s;
  }
  free(Pr); free(Lo);
}

/* ---- libgemini.cpp:223-232: the sources substep after the ionization rates are known:
   global-boundary ghost fill, energy, momentum and mass source/loss, clean_param 3, 2, 1 */
int orc_sources_block(const orc_grid *g, const orc_atmos *a, double dt, double *fluidvars,
                      double *fluidauxvars, const double *electrovars, const double *vs1i,
                      const double *Q, const int *nullpts) {
  dims_t d = dims_of(g);
  const long B = (long)LSP * d.N;
  double *ns = fluidvars, *vs1 = fluidvars + B, *vs2 = fluidvars + 2 * B, *vs3 = fluidvars + 3 * B,
         *Ts = fluidvars + 4 * B, *rhovs1 = fluidauxvars, *rhoes = fluidauxvars + B;
  orc_set_global_boundaries_allspec(g, ns, rhovs1, vs1, vs2, vs3, rhoes, vs1i, LSP);
  orc_source_loss_energy(g, a, dt, ns, vs1, vs2, vs3, Ts, rhoes);
  orc_source_loss_momentum(g, a, dt, ns, vs1, vs2, vs3, Ts, electrovars, Q, rhovs1);
  orc_source_loss_mass(g, a, dt, ns, vs1, vs2, vs3, Ts);
  if (orc_clean_param(g, 3, Ts, nullpts) || orc_clean_param(g, 2, vs1, nullpts) ||
      orc_clean_param(g, 1, ns, nullpts))
    return -1;
  return 0;
}
#undef IIX

/* ==== plasma output (io_nompi.f90:183-258, timeutils.f90:13-27,110-196) ==== */

/* timeutils.f90:13-27 */
static int days_in_month(int year, int month) {
  static const int md[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
  if (year < 1600 || year > 2500 || month < 1 || month > 12) return -1;
  int d = md[month - 1];
  if (month == 2 && ((year % 4 == 0 && year % 100 != 0) || year % 400 == 0)) d = 29;
  return d;
}

/* timeutils.f90:110-130, iteratively */
static int day_wrap(int *year, int *month, int *day) {
  if (*month < 1 || *day < 1) return -1;
  for (;;) {
    while (*month > 12) { *month -= 12; *year += 1; }
    const int dm = days_in_month(*year, *month);
    if (dm < 0) return -1;
    if (*day <= dm) return 0;
    *day -= dm;
    *month += 1;
  }
}

/* timeutils.f90:151-196, for a real(wp) UTsec: the 21-character file stem
   YYYYMMDD_SSSSS.ffffff, time rounded to the nearest 10 ms (nint rounds half away from
   zero).  Returns 0, or -1 where the reference stops with an error. */
int orc_filestem(const int ymd[3], double UTsec, char out[22]) {
  int year = ymd[0], month = ymd[1], day = ymd[2];
  const double r = round(UTsec * 100);
  if (!(r >= -2147483648.0 && r <= 2147483647.0) || fabs(r * 10) > 2147483647.0) return -1;
  int millisec = (int)r * 10;
  int seconds = millisec / 1000;
  if (seconds < 0 || seconds > 86400) return -1;
  if (seconds == 86400) {
    day += 1;
    seconds = 0;
    millisec = 0;
    if (day_wrap(&year, &month, &day)) return -1;
  }
  /* the reference forms this in default integers, whose wrap-around leaves the true value */
  const long frac = (long)millisec * 1000 - (long)seconds * 1000000;
  if (year < 0 || year > 9999 || month < 0 || month > 99 || day < 0 || day > 99) return -1;
  char buf[64];
  /* a negative count (UTsec just below zero: seconds truncates to 0) overflows the I6.6
     field, which Fortran fills with asterisks */
  if (frac < 0)
    snprintf(buf, sizeof buf, "%04d%02d%02d_%05d.******", year, month, day, seconds);
  else
    snprintf(buf, sizeof buf, "%04d%02d%02d_%05d.%06ld", year, month, day, seconds, frac);
  memcpy(out, buf, 21);
  out[21] = 0;
  return 0;
}

/* timeutils.f90:79-107: returns 0, or -1 where the reference stops */
int orc_dateinc(double dtsec, int ymd[3], double *UTsec) {
  int year = ymd[0], month = ymd[1], day = ymd[2];
  if (day < 1 || *UTsec < 0 || dtsec < 0 || dtsec > 86400) return -1;
  *UTsec = *UTsec + dtsec;
  while (*UTsec >= 86400) {
    *UTsec = *UTsec - 86400.0;
    day = day + 1;
    if (day_wrap(&year, &month, &day)) return -1;
  }
  ymd[0] = year; ymd[1] = month; ymd[2] = day;
  return 0;
}

/* number of real32 values plasma_output_nompi writes for the field datasets, in order:
   flagoutput 2: neall v1avgall Tavgall Teall J1all J2all J3all v2avgall v3avgall Phiall;
   flagoutput 3: neall Phiall; otherwise: nsall vs1all Tsall J1all J2all J3all v2avgall
   v3avgall Phiall (io_nompi.f90:230-258).  Phiall is the (lx2, lx3) topside slab. */
long orc_output_count(int lx1, int lx2, int lx3, int flagoutput) {
  const long n = (long)lx1 * lx2 * lx3, m = (long)lx2 * lx3;
  if (flagoutput == 2) return 9 * n + m;
  if (flagoutput == 3) return n + m;
  return (3L * LSP + 5) * n + m;
}

/* io_nompi.f90:199-213 and the real() conversions of :230-258.  fluidvars is the ghosted
   (lx1+4, lx2+4, lx3+4, 5*lsp) block (ns vs1 vs2 vs3 Ts), electrovars the ghosted
   (..., 7) block (E1 E2 E3 J1 J2 J3 Phi).  Averages are sum over the lsp-1 ions of
   ns*v, accumulated from zero in species order, divided by the electron density.
   With patch set, J1-J3 are read as the patch writer's dimension(1:,1:,1:) dummies see
   the ghosted fields it is handed (io_nompi.f90:192, libgemini.f90:549-551): its
   J(1:lx1,1:lx2,1:lx3) starts at the ghosted block's first element, ghosted index -1 in
   each dimension.  The root writer's J keeps the ghosted bounds (plasma_output.f90:46). */
void orc_output_pack(int lx1, int lx2, int lx3, int gridflag, int flagoutput, int patch,
                     const double *fluidvars, const double *electrovars, float *out) {
  const long L1 = lx1 + 4, L2 = lx2 + 4, L3 = lx3 + 4, Ng = L1 * L2 * L3;
  const long n = (long)lx1 * lx2 * lx3, m = (long)lx2 * lx3, B = (long)LSP * Ng;
  const double *ns = fluidvars, *vs1 = fluidvars + B, *vs2 = fluidvars + 2 * B,
               *vs3 = fluidvars + 3 * B, *Ts = fluidvars + 4 * B;
  const double *J1 = electrovars + 3 * Ng, *J2 = electrovars + 4 * Ng,
               *J3 = electrovars + 5 * Ng, *Phi = electrovars + 6 * Ng;
  float *o = out;
  for (int i3 = 0; i3 < lx3; ++i3)
    for (int i2 = 0; i2 < lx2; ++i2)
      for (int i1 = 0; i1 < lx1; ++i1) {
        const long c = (i1 + 2) + L1 * ((i2 + 2) + L2 * (i3 + 2));
        const long k = i1 + (long)lx1 * (i2 + (long)lx2 * i3);
        const long e = (long)(LSP - 1) * Ng + c;
        const long cj = patch ? i1 + L1 * (i2 + L2 * (long)i3) : c;
        double a2 = 0.0, a3 = 0.0, a1 = 0.0, aT = 0.0;
        for (int isp = 0; isp < LSP - 1; ++isp) {
          const long s = (long)isp * Ng + c;
          a2 = a2 + ns[s] * vs2[s];
          a3 = a3 + ns[s] * vs3[s];
          a1 = a1 + ns[s] * vs1[s];
          aT = aT + ns[s] * Ts[s];
        }
        const double v2avg = a2 / ns[e], v3avg = a3 / ns[e];
        if (flagoutput == 3) {
          o[k] = (float)ns[e];
        } else if (flagoutput == 2) {
          o[k] = (float)ns[e];
          o[n + k] = (float)(a1 / ns[e]);
          o[2 * n + k] = (float)(aT / ns[e]);
          o[3 * n + k] = (float)Ts[e];
          o[4 * n + k] = (float)J1[cj];
          o[5 * n + k] = (float)J2[cj];
          o[6 * n + k] = (float)J3[cj];
          o[7 * n + k] = (float)v2avg;
          o[8 * n + k] = (float)v3avg;
        } else {
          for (int isp = 0; isp < LSP; ++isp) {
            const long s = (long)isp * Ng + c;
            o[isp * n + k] = (float)ns[s];
            o[(LSP + isp) * n + k] = (float)vs1[s];
            o[(2 * LSP + isp) * n + k] = (float)Ts[s];
          }
          o[3 * LSP * n + k] = (float)J1[cj];
          o[(3 * LSP + 1) * n + k] = (float)J2[cj];
          o[(3 * LSP + 2) * n + k] = (float)J3[cj];
          o[(3 * LSP + 3) * n + k] = (float)v2avg;
          o[(3 * LSP + 4) * n + k] = (float)v3avg;
        }
      }
  float *phi = out + orc_output_count(lx1, lx2, lx3, flagoutput) - m;
  const int top = gridflag == 1 ? 0 : lx1 - 1;   /* Phi(1,:,:) inverted, else Phi(lx1,:,:) */
  for (int i3 = 0; i3 < lx3; ++i3)
    for (int i2 = 0; i2 < lx2; ++i2)
      phi[i2 + (long)lx2 * i3] = (float)Phi[(top + 2) + L1 * ((i2 + 2) + L2 * (i3 + 2))];
}
