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
#include <type_traits>

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
// One wavefront (the whole workgroup) owns 64 columns of one species.  Each pass over
// the rows is one of:
//   forward (per stage): the input rows arrive in chunks of R rows; the wave loads a
//     chunk of all 64 columns cooperatively (lane -> (row, column), so every load
//     instruction reads whole 64-byte segments), computes the conductivity and the
//     coefficients C, B, D of each (row, column) in parallel into LDS, and each lane
//     then runs the sequential pivoted elimination of its own column over the chunk;
//   back substitution (per stage): each lane walks its column from row N to 1, R rows
//     at a time, with the factors of the next R rows already loaded.
// The BDF2 stage recomputes its coefficients from the inputs (one more read of Ts, ns
// and the metric factors) instead of keeping them from the TR stage.  Scratch per (row,
// lane): the pivot row U0, U1, its right-hand side y, the TR solution and U2, the
// pivoting fill-in, which matters only from a column's first row swap on.  Outputs are
// staged through LDS so global stores are coalesced too.  Blocks are numbered so that
// the seven species of a column group run on one XCD and share its L2.
//
// The steady state of every pass issues its memory operations unconditionally (clamped
// indices; stores that must not land are redirected to a slot a later store of the same
// lane overwrites, or that is never read), so the compiler's wait counters stay exact and
// the loads issued ahead stay in flight.  Lanes of a group past the last column solve
// the last column again and write the same values to the same places.
constexpr int LP = 65;         // padded LDS row (conflict-free transposed access)
#ifndef DIFF_CHUNK
#define DIFF_CHUNK 8           // rows per chunk: 4 x 8 x 65 doubles of LDS per wave
#endif
#ifndef DIFF_WAVES_PER_SIMD
#define DIFF_WAVES_PER_SIMD 2  // register budget: 512 / this many VGPRs per lane
#endif
#ifndef DIFF_BATCH_I
#define DIFF_BATCH_I 8         // ion cells per lane whose inputs are loaded together
#endif
#ifndef DIFF_BATCH_E
#define DIFF_BATCH_E 4         // the same for electron cells (nine inputs each)
#endif
enum : int { SU0 = 0, SU1 = 1, SY = 2, STR = 3, SU2 = 4, NSLOT = 5 };

struct Group {
  int N;
  double* scr;                 // this lane's scratch: slot k of row i at scr[((i-1)*NSLOT+k)*64]
};

__device__ __forceinline__ double& S2(const Group& q, int i, int k) {
  return q.scr[((long)(i - 1) * NSLOT + k) * 64];
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ void column_of(const DevGrid& g, int gcol, int& i2, int& i3) {
  i2 = gcol % g.lx2 + 1;
  i3 = gcol / g.lx2 + 1;
}

// inputs of one cell (neutral densities and J1 for electrons only)
template <bool ELEC>
struct CellIn {
  double T, n, h1, h2, h3;
  double n1, n2, n3, J;
};

template <bool ELEC>
__device__ __forceinline__ void cell_load(const DevGrid& g, const DevState& s, long so, long cb,
                                          long kb, long ni, int r, CellIn<ELEC>& c) {
  c.T = s.Ts[so + cb + r];
  c.n = s.ns[so + cb + r];
  c.h1 = g.h1[cb + r]; c.h2 = g.h2[cb + r]; c.h3 = g.h3[cb + r];
  if (ELEC) {
    c.n1 = s.nn[kb + r - 1]; c.n2 = s.nn[ni + kb + r - 1]; c.n3 = s.nn[2 * ni + kb + r - 1];
    c.J = s.J1[cb + r];
  }
}

// conductivity (collisions.f90:162-199) and parabolic coefficients (diffusion.f90)
template <bool ELEC>
__device__ __forceinline__ void cell_coeffs(const CellIn<ELEC>& c, int isp, double gm1kB,
                                            double& Cr, double& Br, double& Dr) {
  const double T = c.T;
  double lambda, beta;
  if (!ELEC) {
    lambda = thermal_coeff(isp) * pow25(T);
    beta = 0.0;
  } else {
    const double sq = sqrt(T);
    const double br = ((c.n1 * 1.1e-16) * (1.0 + 5.7e-4 * T) + ((c.n2 * 2.82e-17) * sq) * (1.0 - 1.21e-4 * T))
                      + (c.n3 * 2.2e-16) * (1.0 + 3.6e-2 * sq);
    const double den = 1.0 + ((3.22e4 * (T * T)) / c.n) * br;
    lambda = (((elchrg * 100.0) * 7.7e5) * pow25(T)) / den;
    beta = (((5.0 / 2.0) * kB) / elchrg) * c.J;
  }
  Cr = (gm1kB / fmax_f(c.n, mindensdiv)) / ((c.h1 * c.h2) * c.h3);
  Br = (Cr * beta) / c.h1;
  Dr = ((lambda * c.h2) * c.h3) / c.h1;
}

// pivoted elimination step j (dgbtf2 + forward dgbtrs) bringing in row j+1, applied only
// when `real`; fsw is the first step of this column that swapped rows.  Every store is
// made: a step that is not real stores into row sj, which the real step sj stores over
// later, and U2 before the first swap goes to row 1's U2 slot, which is read only when
// the first swap is at step 1 (and then every U2 store is a real one).
__device__ __forceinline__ void elim2(const Group& q, Elim& e, int& fsw, int j, int sj, bool real,
                                      double a, double b, double cc, double d) {
  const bool swap = fabs(a) > fabs(e.r0);
  double p0, p1, p2, pd, o0, o1, o2, od;
  if (swap) { p0 = a; p1 = b; p2 = cc; pd = d; o0 = e.r0; o1 = e.r1; o2 = 0.0; od = e.rb; }
  else      { p0 = e.r0; p1 = e.r1; p2 = 0.0; pd = e.rb; o0 = a; o1 = b; o2 = cc; od = d; }
  bool sing = false;
  if (p0 != 0.0) {
    const double rp = 1.0 / p0;
    const double l = rp * o0;
    if (p1 != 0.0) o1 = o1 + l * (-p1);
    if (p2 != 0.0) o2 = o2 + l * (-p2);
    if (pd != 0.0) od = od + l * (-pd);
  } else {
    sing = true;
  }
  const int f2 = (real && swap && j < fsw) ? j : fsw;
  S2(q, sj, SU0) = p0; S2(q, sj, SU1) = p1; S2(q, sj, SY) = pd;
  S2(q, sj >= f2 ? sj : 1, SU2) = p2;
  if (real) {
    fsw = f2;
    e.singular = e.singular || sing;
    e.r0 = o1; e.r1 = o2; e.rb = od;
  }
}

template <int R>
__global__ __launch_bounds__(64, DIFF_WAVES_PER_SIMD) void k_diffusion(DevGrid g, DevState s, double* scratch,
                                                  Coef tr, Coef bdf, Coef be, int backeuler,
                                                  double Teinf);

// The whole solve of one column group; ELEC selects the electron conductivity.
template <int R, bool ELEC>
__device__ __forceinline__ void diffusion_group(const DevGrid& g, const DevState& s,
                                                double* scratch, const Coef& tr, const Coef& bdf,
                                                const Coef& be, int backeuler, double Teinf,
                                                int isp, long cg, long gps,
                                                double (*sT)[LP], double (*sC)[LP],
                                                double (*sB)[LP], double (*sD)[LP]) {
  const long ncol = (long)g.lx2 * g.lx3;
  const int lane = threadIdx.x;
  const long colbase0 = cg * 64;                     // first column of the group
  const int col = (int)min(colbase0 + lane, ncol - 1);
  const int N = g.lx1;
  Group q;
  q.N = N;
  q.scr = scratch + ((long)(isp - 1) * gps + cg) * (long)N * NSLOT * 64 + lane;
  const long so = (long)(isp - 1) * g.N;
  const long ni = (long)g.lx1 * g.lx2 * g.lx3;
  const long kb = (long)g.lx1 * col;
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
  // cooperative mapping of a chunk: lane -> (row lr, column sub-index cs); R loads per lane
  const int lr = lane % R, cs = lane / R;
  constexpr int CPI = 64 / R;                        // columns per load instruction
  constexpr int BATCH = ELEC ? DIFF_BATCH_E : DIFF_BATCH_I;   // cells whose inputs load together
  static_assert(R % BATCH == 0 || R < BATCH, "chunk and batch sizes");
  constexpr int NB = R < BATCH ? R : BATCH;
  int cbk[R];                                        // gix(0, i2, i3) of the columns this lane loads
  int kbk[R];                                        // and their interior offsets lx1*(column)
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int gcol = (int)min(colbase0 + k * CPI + cs, ncol - 1);
    int a2, a3;
    column_of(g, gcol, a2, a3);
    cbk[k] = (int)gix(g, 0, a2, a3);
    kbk[k] = g.lx1 * gcol;
  }

  // load and convert the rows r0..r0+R-1 of the 64 columns into LDS (rows past N repeat N)
  auto load_chunk = [&](int r0) {
    const int r = min(r0 + lr, N);
#pragma unroll
    for (int k0 = 0; k0 < R; k0 += NB) {
      CellIn<ELEC> in[NB];
#pragma unroll
      for (int k = 0; k < NB; ++k) cell_load<ELEC>(g, s, so, cbk[k0 + k], kbk[k0 + k], ni, r, in[k]);
#pragma unroll
      for (int k = 0; k < NB; ++k) {
        double C_, B_, D_;
        cell_coeffs<ELEC>(in[k], isp, gm1kB, C_, B_, D_);
        const int cc = (k0 + k) * CPI + cs;
        sT[lr][cc] = in[k].T; sC[lr][cc] = C_; sB[lr][cc] = B_; sD[lr][cc] = D_;
      }
    }
  };

  // Forward pass of stage 1 (TR, or backward Euler) or stage 2 (BDF2): rows arrive in
  // chunks; row r's arrival completes the coefficients of row ro = r-1, whose elimination
  // is step ro-1.  Rows 1 and 2 (and rows past N) are not real steps.
  auto forward = [&](int stage, Elim& e, int& fsw) {
    const Coef& k1 = stage == 2 ? bdf : (backeuler ? be : tr);
    e = Elim{1.0, 0.0, Tmin, false};
    fsw = N + 1;
    double Tm = 0.0, Tc = 0.0, Cc = 0.0, Bc = 0.0, Dm = 0.0, Dc = 1.0;
    for (int r0 = 1; r0 <= N; r0 += R) {
      double trv[R];                                 // TR of the rows finalised below
      if (stage == 2) {
#pragma unroll
        for (int k = 0; k < R; ++k) trv[k] = S2(q, clampi(r0 + k - 1, 1, N), STR);
      }
      __builtin_amdgcn_wave_barrier();
      load_chunk(r0);
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int r = r0 + k;
        const double Tp = sT[k][lane], Cp = sC[k][lane], Bp = sB[k][lane], Dp = sD[k][lane];
        const int ro = r - 1;                       // row being finalised (2..N-1)
        const bool real = r >= 3 && r <= N;
        const int rc = clampi(ro, 2, N - 1);
        const double Dhr = 0.5 * (Dm + Dc), Dhr1 = 0.5 * (Dc + Dp);
        const double t1f = ((Cc * Dhr) / dx1i[rc]) / dx1[rc];
        const double t3f = ((Cc * Dhr1) / dx1i[rc]) / dx1[rc + 1];
        const double t2f = Bc / (dx1[rc + 1] + dx1[rc]);
        double a, b, c, d;
        if (stage == 2) {                           // PDEparabolic.f90:146-216
          a = -t1f + t2f; b = (k1.K + t3f) + t1f; c = -t3f - t2f;
          d = -((k1.c13 * Tc) / k1.dt3) + (k1.c43 * trv[k]) / k1.dt3;
        } else if (backeuler) {                     // PDEparabolic.f90:221-346
          a = -t1f + t2f; b = (k1.K + t3f) + t1f; c = -t3f - t2f; d = Tc / k1.dt;
        } else {                                    // PDEparabolic.f90:19-145
          const double t1 = t1f / 2.0, t2 = t2f / 2.0, t3 = t3f / 2.0;
          a = -t1 + t2; b = (k1.K + t3) + t1; c = -t3 - t2;
          d = ((Tc / k1.dt2 - a * Tm) - (t3 + t1) * Tc) - c * Tp;
        }
        elim2(q, e, fsw, ro - 1, clampi(ro - 1, 1, N - 1), real, a, b, c, d);
        Tm = Tc; Tc = Tp;
        Cc = Cp; Bc = Bp;
        Dm = Dc; Dc = Dp;
      }
    }
    elim2(q, e, fsw, N - 1, N - 1, true, 0.0, 1.0, 0.0, Tmax);   // Dirichlet row N
    if (e.r0 == 0.0) e.singular = true;
    S2(q, N, SU0) = e.r0; S2(q, N, SU1) = 0.0; S2(q, N, SY) = e.rb;
    S2(q, N >= fsw ? N : 1, SU2) = 0.0;
  };

  // For lanes whose factorisation is singular, dgbsv returns the right-hand side
  // unchanged: store it in the y slot (a slow pass that never runs in practice).
  auto rhs_pass = [&](int stage, bool mine) {
    const Coef& k1 = stage == 2 ? bdf : (backeuler ? be : tr);
    if (mine) S2(q, 1, SY) = Tmin;
    double Tm = 0.0, Tc = 0.0, Cc = 0.0, Bc = 0.0, Dm = 0.0, Dc = 1.0;
    for (int r0 = 1; r0 <= N; r0 += R) {
      __builtin_amdgcn_wave_barrier();
      load_chunk(r0);
      __builtin_amdgcn_wave_barrier();
      for (int k = 0; k < R && r0 + k <= N; ++k) {
        const int r = r0 + k;
        const double Tp = sT[k][lane], Cp = sC[k][lane], Bp = sB[k][lane], Dp = sD[k][lane];
        if (r >= 3 && mine) {
          const int ro = r - 1;
          const double Dhr = 0.5 * (Dm + Dc), Dhr1 = 0.5 * (Dc + Dp);
          const double t1f = ((Cc * Dhr) / dx1i[ro]) / dx1[ro];
          const double t3f = ((Cc * Dhr1) / dx1i[ro]) / dx1[ro + 1];
          const double t2f = Bc / (dx1[ro + 1] + dx1[ro]);
          double d;
          if (stage == 2) {
            d = -((k1.c13 * Tc) / k1.dt3) + (k1.c43 * S2(q, ro, STR)) / k1.dt3;
          } else if (backeuler) {
            d = Tc / k1.dt;
          } else {
            const double t1 = t1f / 2.0, t2 = t2f / 2.0, t3 = t3f / 2.0;
            const double a = -t1 + t2, c = -t3 - t2;
            d = ((Tc / k1.dt2 - a * Tm) - (t3 + t1) * Tc) - c * Tp;
          }
          S2(q, ro, SY) = d;
        }
        Tm = Tc; Tc = Tp;
        Cc = Cp; Bc = Bp;
        Dm = Dc; Dc = Dp;
      }
    }
    if (mine) S2(q, N, SY) = Tmax;
  };

  auto factor = [&](int stage, Elim& e, int& fsw) {
    forward(stage, e, fsw);
    if (__any(e.singular)) rhs_pass(stage, e.singular);
  };

  // dtbsv over the stored factors (column-oriented: row i takes x(i+2)*U2 off first, then
  // x(i+1)*U1, and divides by U0; a zero x(j) contributes nothing).  U2 is zero above the
  // first swap; WU2 = false when no lane of the wave swapped.  A singular factorisation
  // leaves the right-hand side, stored in the y slot, as the result.  Rows are taken in
  // the aligned chunks [kR+1, kR+R] from the top; each chunk's results go to `put`,
  // then `flush(k)` runs.
  auto backsub = [&](auto wu2, bool singular, int fsw, auto put, auto flush) {
    constexpr bool WU2 = decltype(wu2)::value;
    double pu0[R], pu1[R], py[R], pu2[R];
    const int kt = (N - 1) / R;                      // top chunk
    auto fetch = [&](int i, int u) {
      const int ic = clampi(i, 1, N);
      pu0[u] = S2(q, ic, SU0); pu1[u] = S2(q, ic, SU1); py[u] = S2(q, ic, SY);
      if (WU2) pu2[u] = S2(q, ic, SU2);
    };
#pragma unroll
    for (int u = 0; u < R; ++u) fetch(kt * R + 1 + u, u);     // row i lives in slot (i-1) % R
    double x1 = 0.0, x2 = 0.0;
    auto row = [&](int i, int u, bool live) {
      const double u0 = pu0[u], u1 = pu1[u], y = py[u];
      const double u2 = WU2 ? pu2[u] : 0.0;
      fetch(i - R, u);
      double acc = y;
      if (!singular) {
        if (WU2 && i + 2 <= N && x2 != 0.0 && i >= fsw) acc = acc - x2 * u2;
        if (i + 1 <= N && x1 != 0.0) acc = acc - x1 * u1;
        if (acc != 0.0) acc = acc / u0;
      }
      if (live) {
        x2 = x1;
        x1 = acc;
      }
      put(min(i, N), u, acc);
    };
    // top chunk: rows above N are skipped (their put lands on row N before row N's own)
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const int i = kt * R + R - u;
      row(i, R - 1 - u, i <= N);
    }
    flush(kt);
    for (int k = kt - 1; k >= 0; --k) {
#pragma unroll
      for (int u = 0; u < R; ++u) row(k * R + R - u, R - 1 - u, true);
      flush(k);
    }
  };

  // final temperatures go out through LDS, one aligned chunk of R rows at a time
  auto put_out = [&](int i, int u, double x) { sT[(i - 1) % R][lane] = fmax_f(x, 100.0); };
  auto flush_out = [&](int k) {
    __builtin_amdgcn_wave_barrier();
    const int rr = min(k * R + 1 + lr, N);
#pragma unroll
    for (int kk = 0; kk < R; ++kk) {
      const int cc = kk * CPI + cs;
      s.Ts[so + cbk[kk] + rr] = sT[rr - k * R - 1][cc];
    }
    __builtin_amdgcn_wave_barrier();
  };
  auto put_tr = [&](int i, int u, double x) { S2(q, i, STR) = x; };
  auto flush_none = [&](int) {};

  Elim e;
  int fsw;
  factor(1, e, fsw);
  const bool any_swap = __any(fsw <= N);
  if (backeuler) {
    if (any_swap) backsub(std::true_type{}, e.singular, fsw, put_out, flush_out);
    else backsub(std::false_type{}, e.singular, fsw, put_out, flush_out);
    return;
  }
  if (any_swap) backsub(std::true_type{}, e.singular, fsw, put_tr, flush_none);
  else backsub(std::false_type{}, e.singular, fsw, put_tr, flush_none);
  factor(2, e, fsw);
  const bool any_swap2 = __any(fsw <= N);
  if (any_swap2) backsub(std::true_type{}, e.singular, fsw, put_out, flush_out);
  else backsub(std::false_type{}, e.singular, fsw, put_out, flush_out);
}

template <int R>
__global__ __launch_bounds__(64, DIFF_WAVES_PER_SIMD) void k_diffusion(DevGrid g, DevState s, double* scratch,
                                                  Coef tr, Coef bdf, Coef be, int backeuler,
                                                  double Teinf) {
  __shared__ double sT[R][LP], sC[R][LP], sB[R][LP], sD[R][LP];
  const long ncol = (long)g.lx2 * g.lx3;
  const long gps = (ncol + 63) / 64;                 // column groups per species
  const long slot = blockIdx.x / 8;
  const long cg = (slot / LSP) * 8 + blockIdx.x % 8;  // column group (same XCD for all species)
  if (cg >= gps) return;
  const int isp = (int)(slot % LSP) + 1;
  if (isp == LSP)
    diffusion_group<R, true>(g, s, scratch, tr, bdf, be, backeuler, Teinf, isp, cg, gps, sT, sC, sB, sD);
  else
    diffusion_group<R, false>(g, s, scratch, tr, bdf, be, backeuler, Teinf, isp, cg, gps, sT, sC, sB, sD);
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
