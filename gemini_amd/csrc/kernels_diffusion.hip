// kernels_diffusion.hip -- field-aligned energy diffusion (TRBDF2 / backward Euler).
//
// Reference: src/multifluid/multifluid.f90:309-342 (energy_diffusion),
// src/collisions/collisions.f90:162-199 (thermal_conduct),
// src/numerical/diffusion/diffusion.f90:24-152 (diffusion_prep, column loops),
// src/numerical/diffusion/PDEparabolic.f90:19-346 (TRBDF21D, backEuler1D) and the
// LAPACK dgbsv call of src/vendor/lapack95/gbsv.in.f90:58-97.
//
// One thread owns one (x2,x3) column of one species and reproduces, operation for
// operation, what the reference computes for it: the conductivity and the band matrix
// entries in the reference's expression order, then LAPACK's banded LU with partial
// pivoting (dgbtf2: idamax pivot search, dscal by the reciprocal pivot, dger update
// skipping zero columns; dgbtrs: row swaps and dger on the right-hand side; dtbsv:
// column-oriented back substitution).  The matrix is tridiagonal stored with kl=ku=2,
// so of the band only the three diagonals and the pivoting fill-in can be non-zero;
// every other term LAPACK forms is a multiply by an exact zero and leaves values
// unchanged, so those terms are not formed.  The divisions stay true IEEE divisions and
// the file is built with -ffp-contract=off, which makes the result bit-identical to the
// reference arithmetic apart from x**2.5 (correctly rounded here, glibc pow there).
//
// The LU factors of a column are needed again in reverse order by the back
// substitution, so the forward sweep writes them to a scratch area laid out
// [column group][row][slot][lane] (64 lanes): each row of a wavefront's 64 columns is
// one contiguous 512-byte access per slot.
#include "gemhip_internal.h"

namespace gemhip {

namespace {

constexpr int NSLOT = 8;   // U0 U1 U2 y t1f t2f t3f TR

struct Coef {
  double K;        // 1/(dt/2) (TR), 1/(dt/3) (BDF2) or 1/dt (backward Euler)
  double dt2, dt3, dt;
  double c13, c43; // 1/3._wp and 4/3._wp
};

struct Col {
  const double *T, *n, *h1, *h2, *h3, *J1, *nn1, *nn2, *nn3;  // pointer to row 0
  double* Tout;
  double* scr;     // scratch of this thread: slot k of row i at scr[((i-1)*NSLOT+k)*64]
  int isp, lx1;
  double gm1kB;
};

__device__ __forceinline__ double& S(const Col& c, int i, int k) {
  return c.scr[((long)(i - 1) * NSLOT + k) * 64];
}

// thermal_conduct (collisions.f90:181-197) and diffusion_prep (diffusion.f90:45-50)
// for row r: returns C, B, D of the parabolic equation.
__device__ __forceinline__ void coeffs_at(const Col& c, int r, double& Cr, double& Br,
                                          double& Dr) {
  const double T = c.T[r], n = c.n[r];
  const double h1 = c.h1[r], h2 = c.h2[r], h3 = c.h3[r];
  double lambda, beta;
  if (c.isp < LSP) {
    lambda = thermal_coeff(c.isp) * pow25(T);
    beta = 0.0;
  } else {
    const double n1 = c.nn1[r - 1], n2 = c.nn2[r - 1], n3 = c.nn3[r - 1];
    const double sq = sqrt(T);
    const double br = ((n1 * 1.1e-16) * (1.0 + 5.7e-4 * T) + ((n2 * 2.82e-17) * sq) * (1.0 - 1.21e-4 * T))
                      + (n3 * 2.2e-16) * (1.0 + 3.6e-2 * sq);
    const double den = 1.0 + ((3.22e4 * (T * T)) / n) * br;
    lambda = (((elchrg * 100.0) * 7.7e5) * pow25(T)) / den;
    beta = (((5.0 / 2.0) * kB) / elchrg) * c.J1[r];
  }
  Cr = (c.gm1kB / fmax_f(n, mindensdiv)) / ((h1 * h2) * h3);
  Br = (Cr * beta) / h1;
  Dr = ((lambda * h2) * h3) / h1;
}

// State of the pivoted elimination between steps: the row at pivot position j holds
// (r0, r1) in columns j, j+1 (column j+2 is always zero there) and right-hand side rb.
struct Elim {
  double r0, r1, rb;
  bool singular;
};

// Step j of dgbtf2 + the forward part of dgbtrs, bringing in original row j+1
// (a = A(j+1,j), b = A(j+1,j+1), cc = A(j+1,j+2), d = rhs(j+1)).  Writes U(j,j..j+2)
// and y(j) to scratch.
__device__ __forceinline__ void elim_step(const Col& c, Elim& e, int j, double a, double b,
                                          double cc, double d) {
  const bool swap = fabs(a) > fabs(e.r0);           // idamax over (r0, a, 0)
  double p0, p1, p2, pd, o0, o1, o2, od;
  if (swap) { p0 = a; p1 = b; p2 = cc; pd = d; o0 = e.r0; o1 = e.r1; o2 = 0.0; od = e.rb; }
  else      { p0 = e.r0; p1 = e.r1; p2 = 0.0; pd = e.rb; o0 = a; o1 = b; o2 = cc; od = d; }
  if (p0 != 0.0) {
    const double rp = 1.0 / p0;                     // dscal(km, ONE/AB(KV+1,J))
    const double l = rp * o0;
    if (p1 != 0.0) o1 = o1 + l * (-p1);             // dger, column j+1
    if (p2 != 0.0) o2 = o2 + l * (-p2);             // dger, column j+2
    if (pd != 0.0) od = od + l * (-pd);             // dgbtrs dger on the rhs
  } else {
    e.singular = true;                              // INFO > 0: dgbsv skips the solve
  }
  S(c, j, 0) = p0; S(c, j, 1) = p1; S(c, j, 2) = p2; S(c, j, 3) = pd;
  e.r0 = o1; e.r1 = o2; e.rb = od;
}

// dtbsv('U','N','N') over the stored factors; out(i) receives x(i).  When the
// factorisation was singular LAPACK leaves the right-hand side in place; rhs(i)
// recomputes it.
template <typename OUT, typename RHS>
__device__ __forceinline__ void back_subst(const Col& c, bool singular, OUT out, RHS rhs) {
  const int N = c.lx1;
  if (singular) {
    for (int i = N; i >= 1; --i) out(i, rhs(i));
    return;
  }
  double x1 = 0.0, x2 = 0.0;   // x(i+1), x(i+2)
  for (int i = N; i >= 1; --i) {
    double acc = S(c, i, 3);
    if (i + 2 <= N && x2 != 0.0) acc = acc - x2 * S(c, i, 2);
    if (i + 1 <= N && x1 != 0.0) acc = acc - x1 * S(c, i, 1);
    if (acc != 0.0) acc = acc / S(c, i, 0);
    out(i, acc);
    x2 = x1;
    x1 = acc;
  }
}

// TRBDF21D (PDEparabolic.f90:19-218) with the boundary values of diffusion_prep
// (diffusion.f90:54-79), or backEuler1D (:221-346) when be != 0.
__global__ __launch_bounds__(64) void k_diffusion(DevGrid g, DevState s, double* scratch,
                                                  Coef tr, Coef bdf, Coef be, int backeuler,
                                                  double Teinf) {
  const long ncol = (long)g.lx2 * g.lx3;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long lane = threadIdx.x & 63;
  const long grp = t >> 6;
  if (t >= ncol * LSP) return;
  const int isp = (int)(t / ncol) + 1;
  const long col = t % ncol;
  const int i2 = (int)(col % g.lx2) + 1, i3 = (int)(col / g.lx2) + 1;
  const int N = g.lx1;
  const long cg = gix(g, 0, i2, i3);                  // cell (0,i2,i3)
  const long so = (long)(isp - 1) * g.N;
  const long ni = (long)g.lx1 * g.lx2 * g.lx3;
  const long kb = (long)g.lx1 * col;                  // Tn/nn index of (1,i2,i3)
  Col c;
  c.T = s.Ts + so + cg;
  c.Tout = s.Ts + so + cg;
  c.n = s.ns + so + cg;
  c.h1 = g.h1 + cg; c.h2 = g.h2 + cg; c.h3 = g.h3 + cg;
  c.J1 = s.J1 + cg;
  c.nn1 = s.nn + kb; c.nn2 = s.nn + ni + kb; c.nn3 = s.nn + 2 * ni + kb;
  c.scr = scratch + grp * (long)N * NSLOT * 64 + lane;
  c.isp = isp;
  c.lx1 = N;
  c.gm1kB = (gammas(isp) - 1.0) / kB;
  const double* dx1 = g.dx1;                          // dx1(0:lx1+2)
  const double* dx1i = g.dx1i - 1;                    // dx1i(1:lx1)
  // diffusion.f90:54-79
  double Tmin, Tmax;
  const double Tn1 = s.Tn[kb], TnN = s.Tn[kb + N - 1];
  if (g.gridflag == 0) { Tmin = Tn1; Tmax = TnN; }
  else if (g.gridflag == 1) { Tmax = TnN; Tmin = Teinf; }
  else { Tmin = Tn1; Tmax = Teinf; }

  // ---- first (or only) stage: matrix from the conductivities, rhs from Ts ----
  const Coef& s1 = backeuler ? be : tr;
  Elim e{1.0, 0.0, Tmin, false};                      // row 1: diag 1, super 0, rhs Tmin
  double Tm = c.T[1], Tc = c.T[2], Tp = 0.0;
  double Cm, Bm, Dm, Cc, Bc, Dc, Cp = 0.0, Bp = 0.0, Dp = 0.0;
  coeffs_at(c, 1, Cm, Bm, Dm);
  coeffs_at(c, 2, Cc, Bc, Dc);
  for (int j = 1; j <= N - 1; ++j) {
    const int r = j + 1;
    double a, b, cc, d;
    if (r < N) {
      Tp = c.T[r + 1];
      coeffs_at(c, r + 1, Cp, Bp, Dp);
      const double Dhr = 0.5 * (Dm + Dc), Dhr1 = 0.5 * (Dc + Dp);
      const double t1f = ((Cc * Dhr) / dx1i[r]) / dx1[r];
      const double t3f = ((Cc * Dhr1) / dx1i[r]) / dx1[r + 1];
      const double t2f = Bc / (dx1[r + 1] + dx1[r]);
      S(c, r, 4) = t1f; S(c, r, 5) = t2f; S(c, r, 6) = t3f;
      if (backeuler) {
        a = -t1f + t2f;
        b = (s1.K + t3f) + t1f;
        cc = -t3f - t2f;
        d = Tc / s1.dt;
      } else {
        const double t1 = t1f / 2.0, t2 = t2f / 2.0, t3 = t3f / 2.0;
        a = -t1 + t2;
        b = (s1.K + t3) + t1;
        cc = -t3 - t2;
        d = ((Tc / s1.dt2 - a * Tm) - (t3 + t1) * Tc) - cc * Tp;
      }
    } else {
      a = 0.0; b = 1.0; cc = 0.0; d = Tmax;
    }
    elim_step(c, e, j, a, b, cc, d);
    Tm = Tc; Tc = Tp;
    Cm = Cc; Bm = Bc; Dm = Dc;
    Cc = Cp; Bc = Bp; Dc = Dp;
  }
  if (e.r0 == 0.0) e.singular = true;
  S(c, N, 0) = e.r0; S(c, N, 1) = 0.0; S(c, N, 2) = 0.0; S(c, N, 3) = e.rb;

  if (backeuler) {
    back_subst(c, e.singular,
               [&](int i, double x) { c.Tout[i] = fmax_f(x, 100.0); },
               [&](int i) {
                 if (i == 1) return Tmin;
                 if (i == N) return Tmax;
                 return c.T[i] / s1.dt;
               });
    return;
  }
  // TR solution into slot 7
  back_subst(c, e.singular, [&](int i, double x) { S(c, i, 7) = x; },
             [&](int i) {
               if (i == 1) return Tmin;
               if (i == N) return Tmax;
               const double t1 = S(c, i, 4) / 2.0, t2 = S(c, i, 5) / 2.0, t3 = S(c, i, 6) / 2.0;
               const double a = -t1 + t2, cc = -t3 - t2;
               return ((c.T[i] / tr.dt2 - a * c.T[i - 1]) - (t3 + t1) * c.T[i]) - cc * c.T[i + 1];
             });

  // ---- BDF2 stage (PDEparabolic.f90:146-216) ----
  Elim f{1.0, 0.0, Tmin, false};
  for (int j = 1; j <= N - 1; ++j) {
    const int r = j + 1;
    double a, b, cc, d;
    if (r < N) {
      const double t1f = S(c, r, 4), t2f = S(c, r, 5), t3f = S(c, r, 6);
      a = -t1f + t2f;
      b = (bdf.K + t3f) + t1f;
      cc = -t3f - t2f;
      d = -((bdf.c13 * c.T[r]) / bdf.dt3) + (bdf.c43 * S(c, r, 7)) / bdf.dt3;
    } else {
      a = 0.0; b = 1.0; cc = 0.0; d = Tmax;
    }
    elim_step(c, f, j, a, b, cc, d);
  }
  if (f.r0 == 0.0) f.singular = true;
  S(c, N, 0) = f.r0; S(c, N, 1) = 0.0; S(c, N, 2) = 0.0; S(c, N, 3) = f.rb;
  back_subst(c, f.singular, [&](int i, double x) { c.Tout[i] = fmax_f(x, 100.0); },
             [&](int i) {
               if (i == 1) return Tmin;
               if (i == N) return Tmax;
               return -((bdf.c13 * c.T[i]) / bdf.dt3) + (bdf.c43 * S(c, i, 7)) / bdf.dt3;
             });
}

}  // namespace

void launch_floor_ghosts(const DevGrid& g, double* Ts, hipStream_t st);

size_t diffusion_scratch_bytes(const DevGrid& g) {
  const long ncol = (long)g.lx2 * g.lx3 * LSP;
  const long groups = (ncol + 63) / 64;
  return (size_t)groups * (size_t)g.lx1 * NSLOT * 64 * sizeof(double);
}

void launch_energy_diffusion(const DevGrid& g, const DevState& s, double dt, int diffsolvetype,
                             double Teinf, double* scratch, hipStream_t st) {
  Coef tr{}, bdf{}, be{};
  tr.dt2 = dt / 2.0; tr.K = 1.0 / tr.dt2; tr.dt = dt;
  bdf.dt3 = dt / 3.0; bdf.K = 1.0 / bdf.dt3; bdf.dt = dt;
  bdf.c13 = 1.0 / 3.0; bdf.c43 = 4.0 / 3.0;
  be.dt = dt; be.K = 1.0 / dt;
  const long nthreads = (long)g.lx2 * g.lx3 * LSP;
  timing_begin("trbdf2", st);
  hipLaunchKernelGGL(k_diffusion, dim3((unsigned)((nthreads + 63) / 64)), dim3(64), 0, st, g, s,
                     scratch, tr, bdf, be, diffsolvetype == 1 ? 1 : 0, Teinf);
  timing_end("trbdf2", st);
  timing_begin("floor_ghosts", st);
  launch_floor_ghosts(g, s.Ts, st);
  timing_end("floor_ghosts", st);
}

}  // namespace gemhip
