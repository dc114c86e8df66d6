// kernels_diffusion.hip -- field-aligned energy diffusion (TRBDF2 / backward Euler).
//
// Reference: src/multifluid/multifluid.f90:309-342 (energy_diffusion),
// src/collisions/collisions.f90:162-199 (thermal_conduct),
// src/numerical/diffusion/diffusion.f90:24-152 (diffusion_prep, column loops),
// src/numerical/diffusion/PDEparabolic.f90:19-346 (TRBDF21D, backEuler1D) and the
// LAPACK dgbsv call of src/vendor/lapack95/gbsv.in.f90:58-97.
//
// Each column (fixed x2, x3, species) is solved by one lane, reproducing operation for
// operation what the reference computes for it: the conductivity and the band matrix
// entries in the reference's expression order, then LAPACK's banded LU with partial
// pivoting (dgbtf2: idamax pivot search, dscal by the reciprocal pivot, dger update
// skipping zero columns; dgbtrs: row swaps and dger on the right-hand side; dtbsv:
// column-oriented back substitution).  The matrix is tridiagonal stored with kl=ku=2,
// so of the band only the three diagonals and the pivoting fill-in can be non-zero;
// every other term LAPACK forms is a multiply by an exact zero and leaves values
// unchanged, so those terms are not formed.  The divisions stay true IEEE divisions and
// the file is built with -ffp-contract=off, which makes the result bit-identical to the
// reference arithmetic apart from x**2.5 (correctly rounded here, libm pow there).
#include "gemhip_internal.h"

namespace gemhip {

namespace {

struct Coef {
  double K;        // 1/(dt/2) (TR), 1/(dt/3) (BDF2) or 1/dt (backward Euler)
  double dt2, dt3, dt;
  double c13, c43; // 1/3._wp and 4/3._wp
};

// State of the pivoted elimination between steps: the row at pivot position j holds
// (r0, r1) in columns j, j+1 (column j+2 is always zero there) and right-hand side rb.
struct Elim {
  double r0, r1, rb;
  bool singular;
};

// ---------------------------------------------------------------- column-group kernel
// One wavefront owns 64 columns of one species.  Each pass over the rows is one of:
//   forward (per stage): the input rows arrive in chunks of R rows; the wave loads a
//     chunk of all 64 columns cooperatively (lane -> (row, column), so every load
//     instruction reads whole 128-byte segments), computes the conductivity and the
//     coefficients C, B, D of each (row, column) in parallel into LDS, and each lane
//     then runs the sequential pivoted elimination of its own column over the chunk;
//   back substitution (per stage): each lane walks its column from row N to 1 with the
//     factors of the next DIFF_PF rows already loaded.
// The BDF2 stage recomputes its coefficients from the inputs (one more read of Ts, ns
// and the metric factors) instead of keeping them from the TR stage.  Scratch per (row,
// lane): the pivot row U0, U1, its right-hand side y, the TR solution and U2, the
// pivoting fill-in, which a column only writes from its first row swap on.  Outputs
// are staged through LDS so global stores are coalesced too.  Blocks are numbered so
// that the seven species of a column group run on one XCD and share its L2.
constexpr int LP = 65;         // padded LDS row (conflict-free transposed access)
#ifndef DIFF_CHUNK
#define DIFF_CHUNK 8           // rows per chunk: 4 x 8 x 65 doubles of LDS per wave
#endif
#ifndef DIFF_PF
#define DIFF_PF 8              // rows of factors in flight in the back substitutions
#endif
enum : int { SU0 = 0, SU1 = 1, SY = 2, STR = 3, SU2 = 4, NSLOT = 5 };

struct Group {
  int N;
  double* scr;                 // this lane's scratch: slot k of row i at scr[((i-1)*NSLOT+k)*64]
};

__device__ __forceinline__ double& S2(const Group& q, int i, int k) {
  return q.scr[((long)(i - 1) * NSLOT + k) * 64];
}

__device__ __forceinline__ void column_of(const DevGrid& g, int gcol, int& i2, int& i3) {
  i2 = gcol % g.lx2 + 1;
  i3 = gcol / g.lx2 + 1;
}

// conductivity and parabolic coefficients of cell (r, i2, i3), species isp
__device__ __forceinline__ void cell_coeffs(const DevGrid& g, const DevState& s, int isp,
                                            long so, long cb, long kb, long ni, int r,
                                            double gm1kB, double& T, double& Cr, double& Br,
                                            double& Dr) {
  T = s.Ts[so + cb + r];
  const double n = s.ns[so + cb + r];
  const double h1 = g.h1[cb + r], h2 = g.h2[cb + r], h3 = g.h3[cb + r];
  double lambda, beta;
  if (isp < LSP) {
    lambda = thermal_coeff(isp) * pow25(T);
    beta = 0.0;
  } else {
    const double n1 = s.nn[kb + r - 1], n2 = s.nn[ni + kb + r - 1], n3 = s.nn[2 * ni + kb + r - 1];
    const double sq = sqrt(T);
    const double br = ((n1 * 1.1e-16) * (1.0 + 5.7e-4 * T) + ((n2 * 2.82e-17) * sq) * (1.0 - 1.21e-4 * T))
                      + (n3 * 2.2e-16) * (1.0 + 3.6e-2 * sq);
    const double den = 1.0 + ((3.22e4 * (T * T)) / n) * br;
    lambda = (((elchrg * 100.0) * 7.7e5) * pow25(T)) / den;
    beta = (((5.0 / 2.0) * kB) / elchrg) * s.J1[cb + r];
  }
  Cr = (gm1kB / fmax_f(n, mindensdiv)) / ((h1 * h2) * h3);
  Br = (Cr * beta) / h1;
  Dr = ((lambda * h2) * h3) / h1;
}

// pivoted elimination step j (dgbtf2 + forward dgbtrs) bringing in row j+1; fsw is the
// first step of this column that swapped rows (U2 is stored from there on)
__device__ __forceinline__ void elim2(const Group& q, Elim& e, int& fsw, int j, double a,
                                      double b, double cc, double d) {
  const bool swap = fabs(a) > fabs(e.r0);
  double p0, p1, p2, pd, o0, o1, o2, od;
  if (swap) { p0 = a; p1 = b; p2 = cc; pd = d; o0 = e.r0; o1 = e.r1; o2 = 0.0; od = e.rb; }
  else      { p0 = e.r0; p1 = e.r1; p2 = 0.0; pd = e.rb; o0 = a; o1 = b; o2 = cc; od = d; }
  if (p0 != 0.0) {
    const double rp = 1.0 / p0;
    const double l = rp * o0;
    if (p1 != 0.0) o1 = o1 + l * (-p1);
    if (p2 != 0.0) o2 = o2 + l * (-p2);
    if (pd != 0.0) od = od + l * (-pd);
  } else {
    e.singular = true;
  }
  if (swap && j < fsw) fsw = j;
  S2(q, j, SU0) = p0; S2(q, j, SU1) = p1; S2(q, j, SY) = pd;
  if (j >= fsw) S2(q, j, SU2) = p2;
  e.r0 = o1; e.r1 = o2; e.rb = od;
}

// dtbsv over the stored factors (column-oriented: row i takes x(i+2)*U2 off first, then
// x(i+1)*U1, and divides by U0; a zero x(j) contributes nothing).  U2 is zero above the
// first swap and is not read there.  A singular factorisation leaves the right-hand side,
// which the caller has stored in the y slot, as the result (dgbsv skips dgbtrs).
template <int PF, typename OUT>
__device__ __forceinline__ void backsub2(const Group& q, bool singular, int fsw, OUT out) {
  const int N = q.N;
  double pu0[PF], pu1[PF], py[PF];
#pragma unroll
  for (int u = 0; u < PF; ++u) {
    const int i = N - u;
    if (i >= 1) { pu0[u] = S2(q, i, SU0); pu1[u] = S2(q, i, SU1); py[u] = S2(q, i, SY); }
  }
  double x1 = 0.0, x2 = 0.0;
  for (int ib = N; ib >= 1; ib -= PF) {
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int i = ib - u;
      if (i < 1) break;
      const double u0 = pu0[u], u1 = pu1[u], y = py[u];
      const int ip = i - PF;
      if (ip >= 1) { pu0[u] = S2(q, ip, SU0); pu1[u] = S2(q, ip, SU1); py[u] = S2(q, ip, SY); }
      double acc = y;
      if (!singular) {
        if (i + 2 <= N && x2 != 0.0 && i >= fsw) acc = acc - x2 * S2(q, i, SU2);
        if (i + 1 <= N && x1 != 0.0) acc = acc - x1 * u1;
        if (acc != 0.0) acc = acc / u0;
      }
      out(i, acc);
      x2 = x1;
      x1 = acc;
    }
  }
}

template <int R>
__global__ __launch_bounds__(64) void k_diffusion(DevGrid g, DevState s, double* scratch,
                                                  Coef tr, Coef bdf, Coef be, int backeuler,
                                                  double Teinf) {
  __shared__ double sT[R][LP], sC[R][LP], sB[R][LP], sD[R][LP];
  const long ncol = (long)g.lx2 * g.lx3;
  const long gps = (ncol + 63) / 64;                 // column groups per species
  const long slot = blockIdx.x / 8;
  const long cg = (slot / LSP) * 8 + blockIdx.x % 8;  // column group (same XCD for all species)
  if (cg >= gps) return;
  const int isp = (int)(slot % LSP) + 1;
  const int lane = threadIdx.x;
  const int col = (int)(cg * 64 + lane);
  const bool valid = col < ncol;
  const int N = g.lx1;
  Group q;
  q.N = N;
  q.scr = scratch + ((long)(isp - 1) * gps + cg) * (long)N * NSLOT * 64 + lane;
  const long so = (long)(isp - 1) * g.N;
  const long ni = (long)g.lx1 * g.lx2 * g.lx3;
  const long colbase0 = cg * 64;                     // first column of the group
  const long kb = (long)g.lx1 * (valid ? col : 0);
  const double gm1kB = (gammas(isp) - 1.0) / kB;
  const double* dx1 = g.dx1;
  const double* dx1i = g.dx1i - 1;
  double Tmin, Tmax;
  {
    const double Tn1 = s.Tn[kb], TnN = s.Tn[kb + N - 1];
    if (g.gridflag == 0) { Tmin = Tn1; Tmax = TnN; }
    else if (g.gridflag == 1) { Tmax = TnN; Tmin = Teinf; }
    else { Tmin = Tn1; Tmax = Teinf; }
  }
  // cooperative mapping of a chunk: lane -> (row lr, column sub-index cs)
  const int lr = lane % R, cs = lane / R;
  constexpr int CPI = 64 / R;                        // columns per load instruction

  // Forward pass of stage 1 (TR, or backward Euler) or stage 2 (BDF2).  mode 0
  // eliminates and stores the factors, mode 1 stores the stage's right-hand side in the
  // y slot (for a singular factorisation), mode 2 (other lanes of that pass) does nothing.
  auto forward = [&](int stage, int mode, Elim& e, int& fsw) {
    const Coef& k1 = stage == 2 ? bdf : (backeuler ? be : tr);
    e = Elim{1.0, 0.0, Tmin, false};
    fsw = N + 1;
    if (mode == 1) S2(q, 1, SY) = Tmin;
    double Tm = 0.0, Tc = 0.0, Cc = 0.0, Bc = 0.0, Dm = 0.0, Dc = 0.0;
    for (int r0 = 1; r0 <= N; r0 += R) {
      double trv[R];                                 // TR of the rows finalised below
      if (stage == 2) {
#pragma unroll
        for (int k = 0; k < R; ++k) {
          const int ro = r0 + k - 1;
          trv[k] = (ro >= 2 && ro <= N - 1) ? S2(q, ro, STR) : 0.0;
        }
      }
      __syncthreads();
      for (int k = 0; k < 64 / CPI; ++k) {
        const int cc = k * CPI + cs;
        const long gcol = colbase0 + cc;
        const int r = r0 + lr;
        if (gcol < ncol && r <= N) {
          int a2, a3;
          column_of(g, (int)gcol, a2, a3);
          double T, C_, B_, D_;
          cell_coeffs(g, s, isp, so, gix(g, 0, a2, a3), (long)g.lx1 * gcol, ni, r, gm1kB, T, C_,
                      B_, D_);
          sT[lr][cc] = T; sC[lr][cc] = C_; sB[lr][cc] = B_; sD[lr][cc] = D_;
        }
      }
      __syncthreads();
      const int rend = min(R, N - r0 + 1);
#pragma unroll
      for (int k = 0; k < R; ++k) {
        if (k >= rend) break;
        const int r = r0 + k;
        const double Tp = sT[k][lane], Cp = sC[k][lane], Bp = sB[k][lane], Dp = sD[k][lane];
        if (r >= 3) {
          const int ro = r - 1;                     // row being finalised (2..N-1)
          const double Dhr = 0.5 * (Dm + Dc), Dhr1 = 0.5 * (Dc + Dp);
          const double t1f = ((Cc * Dhr) / dx1i[ro]) / dx1[ro];
          const double t3f = ((Cc * Dhr1) / dx1i[ro]) / dx1[ro + 1];
          const double t2f = Bc / (dx1[ro + 1] + dx1[ro]);
          double a, b, c, d;
          if (stage == 2) {                         // PDEparabolic.f90:146-216
            a = -t1f + t2f; b = (k1.K + t3f) + t1f; c = -t3f - t2f;
            d = -((k1.c13 * Tc) / k1.dt3) + (k1.c43 * trv[k]) / k1.dt3;
          } else if (backeuler) {                   // PDEparabolic.f90:221-346
            a = -t1f + t2f; b = (k1.K + t3f) + t1f; c = -t3f - t2f; d = Tc / k1.dt;
          } else {                                  // PDEparabolic.f90:19-145
            const double t1 = t1f / 2.0, t2 = t2f / 2.0, t3 = t3f / 2.0;
            a = -t1 + t2; b = (k1.K + t3) + t1; c = -t3 - t2;
            d = ((Tc / k1.dt2 - a * Tm) - (t3 + t1) * Tc) - c * Tp;
          }
          if (mode == 0) elim2(q, e, fsw, ro - 1, a, b, c, d);
          else if (mode == 1) S2(q, ro, SY) = d;
        }
        Tm = Tc; Tc = Tp;
        Cc = Cp; Bc = Bp;
        Dm = Dc; Dc = Dp;
      }
    }
    if (mode == 1) S2(q, N, SY) = Tmax;
    if (mode != 0) return;
    elim2(q, e, fsw, N - 1, 0.0, 1.0, 0.0, Tmax);     // Dirichlet row N
    if (e.r0 == 0.0) e.singular = true;
    S2(q, N, SU0) = e.r0; S2(q, N, SU1) = 0.0; S2(q, N, SY) = e.rb;
  };

  // a forward pass, then the right-hand side again for lanes whose factorisation is singular
  auto factor = [&](int stage, Elim& e, int& fsw) {
    forward(stage, 0, e, fsw);
    if (__any(e.singular)) {
      Elim e2;
      int f2;
      forward(stage, e.singular ? 1 : 2, e2, f2);
    }
  };

  // final temperatures go out through LDS in chunks of R rows
  auto store_out = [&](int i, double x) {
    const int rr0 = ((i - 1) / R) * R + 1;           // chunk start of row i
    sT[i - rr0][lane] = fmax_f(x, 100.0);
    if (i == rr0) {                                   // chunk complete (rows fall to rr0)
      __syncthreads();
      const int rend = min(R, N - rr0 + 1);
      for (int k = 0; k < 64 / CPI; ++k) {
        const int cc = k * CPI + cs;
        const long gcol = colbase0 + cc;
        if (gcol < ncol && lr < rend) {
          int a2, a3;
          column_of(g, (int)gcol, a2, a3);
          s.Ts[so + gix(g, 0, a2, a3) + rr0 + lr] = sT[lr][cc];
        }
      }
      __syncthreads();
    }
  };

  Elim e;
  int fsw;
  factor(1, e, fsw);
  __syncthreads();
  if (backeuler) {
    backsub2<DIFF_PF>(q, e.singular, fsw, store_out);
    return;
  }
  backsub2<DIFF_PF>(q, e.singular, fsw, [&](int i, double x) { S2(q, i, STR) = x; });
  factor(2, e, fsw);
  __syncthreads();
  backsub2<DIFF_PF>(q, e.singular, fsw, store_out);
}

}  // namespace

void launch_floor_ghosts(const DevGrid& g, double* Ts, hipStream_t st);

size_t diffusion_scratch_bytes(const DevGrid& g) {
  const long ncol = (long)g.lx2 * g.lx3;
  const long groups = (ncol + 63) / 64 * LSP;
  return (size_t)groups * (size_t)g.lx1 * NSLOT * 64 * sizeof(double);
}

void launch_energy_diffusion(const DevGrid& g, const DevState& s, double dt, int diffsolvetype,
                             double Teinf, double* scratch, hipStream_t st) {
  Coef tr{}, bdf{}, be{};
  tr.dt2 = dt / 2.0; tr.K = 1.0 / tr.dt2; tr.dt = dt;
  bdf.dt3 = dt / 3.0; bdf.K = 1.0 / bdf.dt3; bdf.dt = dt;
  bdf.c13 = 1.0 / 3.0; bdf.c43 = 4.0 / 3.0;
  be.dt = dt; be.K = 1.0 / dt;
  const long ncol = (long)g.lx2 * g.lx3;
  const long blocks = ((ncol + 63) / 64 + 7) / 8 * 8 * LSP;   // XCD-major numbering
  timing_begin("trbdf2", st);
  hipLaunchKernelGGL(k_diffusion<DIFF_CHUNK>, dim3((unsigned)blocks), dim3(64), 0, st, g, s,
                     scratch, tr, bdf, be, diffsolvetype == 1 ? 1 : 0, Teinf);
  timing_end("trbdf2", st);
  timing_begin("floor_ghosts", st);
  launch_floor_ghosts(g, s.Ts, st);
  timing_end("floor_ghosts", st);
}

}  // namespace gemhip
