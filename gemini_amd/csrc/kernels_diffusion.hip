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

#include "diffusion_cell.h"

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
// One wavefront (the whole workgroup) owns 64 columns of one species, one lane per
// column, and makes four passes over the rows:
//   1 (top-down): the input rows arrive in chunks of R rows.  The wave loads a chunk of
//     all 64 columns cooperatively (lane -> (row, column), so each load instruction reads
//     whole 64-byte segments), computes conductivity and the coefficients C, B, D of each
//     (row, column) in parallel into LDS, and each lane then builds its column's rows
//     and runs the sequential pivoted elimination of the first solve (TR, or backward
//     Euler) with its right-hand side.  For TRBDF2 the same pass also factors the BDF2
//     matrix, which does not depend on the TR solution: its pivot rows, multipliers and
//     row interchanges are stored;
//   2 (bottom-up): back substitution of the first solve (into Ts for backward Euler);
//   3 (top-down, TRBDF2): the BDF2 right-hand side from the input temperatures and the
//     TR solution, taken through the stored interchanges and multipliers (dgbtrs's
//     forward solve);
//   4 (bottom-up, TRBDF2): back substitution of the BDF2 solve into Ts.
// Every value is formed with the operations, and in the order, LAPACK's dgbtf2, dgbtrs
// and dtbsv would use on the reference's band matrix: factorising a matrix before its
// right-hand side is known changes nothing in either.  The pivot rows, right-hand sides
// and solutions live in lane-interleaved HBM scratch (one 512-byte segment per row and
// slot for the wave); outputs are staged through LDS so global stores are coalesced.
// Blocks are numbered so that the seven species of a column group run on one XCD, the
// electron groups (the longest waves) first.
//
// The steady state of every pass issues its memory operations unconditionally (clamped
// indices; stores that must not land go where a later store of the same lane overwrites
// them), so the compiler's wait counters stay exact and loads issued ahead stay in
// flight.  Lanes of a group past the last column solve the last column again and write
// the same values to the same places.
constexpr int LP = 65;         // padded LDS row (conflict-free transposed access)
#ifndef DIFF_CHUNK
#define DIFF_CHUNK 8           // rows per chunk: 4 x 8 x 65 doubles of LDS per wave
#endif
#ifndef DIFF_SPLIT
#define DIFF_SPLIT 0           // 1: TRBDF2 as two launches (passes 1-2, passes 3-4)
#endif
#ifndef DIFF_BATCH_I
#define DIFF_BATCH_I 8         // ion cells per lane whose inputs are loaded together
#endif
#ifndef DIFF_BATCH_E
#define DIFF_BATCH_E 4         // the same for electron cells (nine inputs each)
#endif
// scratch slots per (row, lane)
enum : int {
  SU0 = 0, SU1 = 1, SU2 = 2, SY = 3,   // first solve: pivot row (U0 U1 U2) and its rhs y;
                                       // pass 3 reuses SY for the BDF2 rhs
  STR = 4,                             // TR solution
  SV0 = 5, SV1 = 6, SV2 = 7, SL = 8,   // BDF2 pivot rows and multipliers
  NSLOT = 9                            // (the T part of the BDF2 rhs is formed in pass 3
};                                     // from Ts, which pass 4 is the first to overwrite)

struct Group {
  int N;
  double* scr;                 // this lane's scratch: slot k of row i at scr[((i-1)*NSLOT+k)*64]
  int* swb;                    // BDF2 row interchanges, bit (j&31) of word j>>5 at swb[(j>>5)*64]
};

__device__ __forceinline__ double& S2(const Group& q, int i, int k) {
  return q.scr[((long)(i - 1) * NSLOT + k) * 64];
}
// Scratch traffic is non-temporal (the nt bit: streamed through L2 at low priority).  A
// slot is written once and read once, a pass later; as ordinary accesses the ~256 waves
// of an XCD streaming their pivot rows evicted the input lines that the cooperative loads
// of the next row chunk still needed (those loads read 64-byte halves of 128-byte lines).
__device__ __forceinline__ void put2(const Group& q, int i, int k, double v) {
  __builtin_nontemporal_store(v, &S2(q, i, k));
}
__device__ __forceinline__ double get2(const Group& q, int i, int k) {
  return __builtin_nontemporal_load(&S2(q, i, k));
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ void column_of(const DevGrid& g, int gcol, int& i2, int& i3) {
  i2 = gcol % g.lx2 + 1;
  i3 = gcol / g.lx2 + 1;
}

// Pivoted elimination step j (dgbtf2 with dgbtrs's forward solve) bringing in row j+1 =
// (a, b, cc | d), applied to the state only when `real`.  The row at pivot position j is
// stored at row sj (a step that is not real stores where the real step sj stores later).
__device__ __forceinline__ bool elim_rhs(const Group& q, Elim& e, int j, int sj, bool real,
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
  put2(q, sj, SU0, p0); put2(q, sj, SU1, p1); put2(q, sj, SU2, p2); put2(q, sj, SY, pd);
  if (real) {
    e.singular = e.singular || sing;
    e.r0 = o1; e.r1 = o2; e.rb = od;
  }
  return swap;
}

// The same step on the matrix alone (the BDF2 factorisation): stores the pivot row, the
// multiplier l and the interchange bit that the right-hand side will need in pass 3 (a
// word of 32 steps' bits, stored once its last step, or step N-1, has run).
struct Fact {
  double r0, r1;
  bool singular;
  int word, bits;              // interchange bits of steps 32*word .. 32*word+31
};

__device__ __forceinline__ void elim_mat(const Group& q, Fact& f, int j, int sj, bool real,
                                         double a, double b, double cc) {
  const bool swap = fabs(a) > fabs(f.r0);
  double p0, p1, p2, o0, o1, o2;
  if (swap) { p0 = a; p1 = b; p2 = cc; o0 = f.r0; o1 = f.r1; o2 = 0.0; }
  else      { p0 = f.r0; p1 = f.r1; p2 = 0.0; o0 = a; o1 = b; o2 = cc; }
  bool sing = false;
  double l = 0.0;
  if (p0 != 0.0) {
    const double rp = 1.0 / p0;
    l = rp * o0;
    if (p1 != 0.0) o1 = o1 + l * (-p1);
    if (p2 != 0.0) o2 = o2 + l * (-p2);
  } else {
    sing = true;
  }
  put2(q, sj, SV0, p0); put2(q, sj, SV1, p1); put2(q, sj, SV2, p2); put2(q, sj, SL, l);
  if (real) {
    f.singular = f.singular || sing;
    f.r0 = o1; f.r1 = o2;
    const int w = j >> 5;
    f.bits = (w == f.word ? f.bits : 0) | ((swap ? 1 : 0) << (j & 31));
    f.word = w;
  }
  if ((j & 31) == 31 || j >= q.N - 1) q.swb[f.word * 64] = f.bits;   // the word is complete
}

template <int R, bool BE, bool ELEC, int PH, bool EXT>
__device__ __forceinline__ void diffusion_group(const DevGrid& g, const DevState& s,
                                                double* scratch, int* swbits, int* lst,
                                                const Coef& k1,
                                                const Coef& bdf, double Teinf, int isp,
                                                long cg, long gps, double (*sT)[LP],
                                                double (*sC)[LP], double (*sB)[LP],
                                                double (*sD)[LP]) {
  const long ncol = (long)g.lx2 * g.lx3;
  const int lane = threadIdx.x;
  const int col = group_column(g, cg, lane);
  const int N = g.lx1;
  const long grp = (long)(isp - 1) * gps + cg;
  Group q;
  q.N = N;
  q.scr = scratch + grp * (long)N * NSLOT * 64 + lane;
  q.swb = swbits + grp * (long)((N + 31) / 32) * 64 + lane;
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
  // row N: the reference's Dirichlet row, or under the extension with a top heat flux the
  // Neumann row (T(N) - T(N-1))/dx1(N) = q h1(N)/lambda(N)
  double aN = 0.0, bN = 1.0;
  if (EXT && ELEC && s.qtop) {
    CellIn<ELEC> top;
    cell_load<ELEC, EXT>(g, s, so, gix(g, 0, col % g.lx2 + 1, col / g.lx2 + 1), kb, ni, N, top,
                         hcol_index(g, col, N), col);
    aN = -1.0 / dx1[N];
    bN = 1.0 / dx1[N];
    Tmax = (s.qtop[col] * top.h1) / cell_lambda<ELEC, EXT>(top, isp);
  }
  // cooperative mapping of a chunk: lane -> (row lr, column sub-index cs); R loads per lane
  const int lr = lane % R, cs = lane / R;
  constexpr int CPI = 64 / R;                        // columns per load instruction
  constexpr int BATCH = ELEC ? DIFF_BATCH_E : DIFF_BATCH_I;
  static_assert(R % BATCH == 0 || R < BATCH, "chunk and batch sizes");
  constexpr int NB = R < BATCH ? R : BATCH;
  // Row phase of the chunk windows: a column's row 1 sits 2 cells into its stride, so
  // windows of R rows from row 1 straddle 64-byte sectors, and a sector's other part is
  // fetched again a chunk later (long after the L2 dropped it).  The wave's windows start
  // ph rows early instead, which puts the windows of lane 0's column, and of every column
  // that shares its sector phase (all of an interleaved group's), on sector boundaries.
  // The steps before row 1 are not real.  The electrons' neutral densities are stored without
  // ghosts and are aligned at row 1, so their waves keep phase 0 in pass 1.
  int c2, c3;
  column_of(g, group_column(g, cg, 0), c2, c3);
  const int pos1 = (int)(((unsigned long)(s.Ts + so + gix(g, 0, c2, c3) + 1) >> 3) & 7);
  constexpr int RA = R < 8 ? R : 8;                  // (R a power of two)
  const int ph1 = ELEC ? 0 : pos1 % RA;              // pass 1: windows from row 1 - ph1
  const int ph3 = (pos1 + 1) % RA;                   // pass 3: windows from row 2 - ph3
  const int pho = pos1 % RA;                         // output chunks from row 1 - pho
  int cbk[R];                                        // gix(0, i2, i3) of the columns this lane loads
  int kbk[R];                                        // and their interior offsets lx1*(column)
  int gck[R];                                        // and column numbers (extension only)
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int gcol = group_column(g, cg, k * CPI + cs);
    int a2, a3;
    column_of(g, gcol, a2, a3);
    cbk[k] = (int)gix(g, 0, a2, a3);
    kbk[k] = g.lx1 * gcol;
    gck[k] = gcol;
  }

  // load and convert the rows r0..r0+R-1 of the 64 columns into LDS (rows past N repeat N)
  auto load_chunk = [&](int r0) {
    const int r = clampi(r0 + lr, 1, N);
    // hcol slot of this lane's first column at row r: the group's 64 columns are contiguous
    // (slots past the last column repeat it, as these lanes do)
    const long hrow = ((long)cg * N + (r - 1)) * 64 + cs;
#pragma unroll
    for (int k0 = 0; k0 < R; k0 += NB) {
      CellIn<ELEC> in[NB];
#pragma unroll
      for (int k = 0; k < NB; ++k)
        cell_load<ELEC, EXT>(g, s, so, cbk[k0 + k], kbk[k0 + k], ni, r, in[k],
                             hrow + (k0 + k) * CPI, gck[k0 + k]);
#pragma unroll
      for (int k = 0; k < NB; ++k) {
        double C_, B_, D_;
        cell_coeffs<ELEC, EXT>(in[k], isp, gm1kB, C_, B_, D_);
        const int cc = (k0 + k) * CPI + cs;
        sT[lr][cc] = in[k].T; sC[lr][cc] = C_; sB[lr][cc] = B_; sD[lr][cc] = D_;
      }
    }
  };

  // walk the rows in chunks; row r's arrival completes the coefficients of row ro = r-1
  // (2..N-1), handed to `row(ro, real, t1f, t2f, t3f, Tm, Tc, Tp)`; rows before 3 and rows
  // past N are not real and the handler must leave its state alone for them
  auto rows = [&](auto row) {
    double Tm = 0.0, Tc = 0.0, Cc = 0.0, Bc = 0.0, Dm = 0.0, Dc = 1.0;
    for (int r0 = 1 - ph1; r0 <= N; r0 += R) {
      __builtin_amdgcn_wave_barrier();
      load_chunk(r0);
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int r = r0 + k;
        const double Tp = sT[k][lane], Cp = sC[k][lane], Bp = sB[k][lane], Dp = sD[k][lane];
        const int ro = r - 1;
        const int rc = clampi(ro, 2, N - 1);
        const double Dhr = 0.5 * (Dm + Dc), Dhr1 = 0.5 * (Dc + Dp);
        const double t1f = ((Cc * Dhr) / dx1i[rc]) / dx1[rc];
        const double t3f = ((Cc * Dhr1) / dx1i[rc]) / dx1[rc + 1];
        const double t2f = Bc / (dx1[rc + 1] + dx1[rc]);
        row(ro, r >= 3 && r <= N, t1f, t2f, t3f, Tm, Tc, Tp);
        Tm = Tc; Tc = Tp;
        Cc = Cp; Bc = Bp;
        Dm = Dc; Dc = Dp;
      }
    }
  };

  // rows of the first solve (PDEparabolic.f90:19-145 TR half step, :221-346 backward Euler)
  auto first_row = [&](double t1f, double t2f, double t3f, double Tm, double Tc, double Tp,
                       double& a, double& b, double& c, double& d) {
    if (BE) {
      a = -t1f + t2f; b = (k1.K + t3f) + t1f; c = -t3f - t2f; d = Tc / k1.dt;
    } else {
      const double t1 = t1f / 2.0, t2 = t2f / 2.0, t3 = t3f / 2.0;
      a = -t1 + t2; b = (k1.K + t3) + t1; c = -t3 - t2;
      d = ((Tc / k1.dt2 - a * Tm) - (t3 + t1) * Tc) - c * Tp;
    }
  };

  // dtbsv over stored pivot rows (slots s0 s1 s2, rhs sy), column-oriented: row i takes
  // x(i+2)*U2 off first, then x(i+1)*U1, and divides by U0; a zero x(j) contributes
  // nothing.  Rows are taken in the chunks [kR+1-ph, kR+R-ph] from the top; each chunk's
  // results go to `put` (rows past N and before 1 too, which must land nowhere that
  // matters: rows past N come first and are put on row N), then `flush(k)` runs.
  auto backsub = [&](int s0, int s1, int s2, int sy, bool singular, int ph, auto put,
                     auto flush) {
    double pu0[R], pu1[R], pu2[R], py[R];
    const int kt = (N - 1 + ph) / R;                 // top chunk
    auto fetch = [&](int i, int u) {
      const int ic = clampi(i, 1, N);
      pu0[u] = get2(q, ic, s0); pu1[u] = get2(q, ic, s1); pu2[u] = get2(q, ic, s2);
      py[u] = get2(q, ic, sy);
    };
#pragma unroll
    for (int u = 0; u < R; ++u) fetch(kt * R + 1 - ph + u, u);  // row i: slot (i-1+ph) % R
    double x1 = 0.0, x2 = 0.0;
    auto row = [&](int i, int u, bool live) {
      const double u0 = pu0[u], u1 = pu1[u], u2 = pu2[u], y = py[u];
      fetch(i - R, u);
      double acc = y;
      if (!singular) {
        if (i + 2 <= N && x2 != 0.0) acc = acc - x2 * u2;
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
      const int i = kt * R + R - u - ph;
      row(i, R - 1 - u, i <= N);
    }
    flush(kt);
    for (int k = kt - 1; k >= 0; --k) {
#pragma unroll
      for (int u = 0; u < R; ++u) row(k * R + R - u - ph, R - 1 - u, true);
      flush(k);
    }
  };

  // final temperatures go out through LDS, one chunk of R rows (phase pho) at a time; the
  // slots of rows past N and before 1 are never stored (those lanes store row N or 1 again)
  auto put_out = [&](int, int u, double x) { sT[u][lane] = fmax_f(x, 100.0); };
  auto flush_out = [&](int k) {
    __builtin_amdgcn_wave_barrier();
    const int r0 = k * R + 1 - pho;
    const int rr = clampi(r0 + lr, 1, N);
#pragma unroll
    for (int kk = 0; kk < R; ++kk) {
      const int cc = kk * CPI + cs;
      s.Ts[so + cbk[kk] + rr] = sT[rr - r0][cc];
    }
    __builtin_amdgcn_wave_barrier();
  };

  // the BDF2 factorisation's singular flag, kept between the launches of a split solve
  int* mysing = lst + (((long)(isp - 1) * gps + cg) * 64 + lane);
  Elim e{1.0, 0.0, Tmin, false};
  Fact f{1.0, 0.0, false, 0, 0};
  if (PH == 2) {
    f.singular = *mysing != 0;
  } else {
  // the record of the electron columns' interchanges (fitting extension, adjoint tapes)
  const int nw = (N + 31) / 32;
  int* prec = (EXT && ELEC && s.pivrec) ? s.pivrec + (long)col * 2 * nw : nullptr;
  int tbits = 0;
  auto record_tr = [&](int j, bool swap) {   // TR step j (1..N-1)
    tbits |= (swap ? 1 : 0) << (j & 31);
    if ((j & 31) == 31 || j == N - 1) {
      prec[j >> 5] = tbits;
      tbits = 0;
    }
  };
  // pass 1
  rows([&](int ro, bool real, double t1f, double t2f, double t3f, double Tm, double Tc,
           double Tp) {
    const int j = ro - 1, sj = clampi(j, 1, N - 1);
    double a, b, c, d;
    first_row(t1f, t2f, t3f, Tm, Tc, Tp, a, b, c, d);
    const bool sw = elim_rhs(q, e, j, sj, real, a, b, c, d);
    if (EXT && ELEC && prec && real) record_tr(j, sw);
    if (!BE) {                                       // PDEparabolic.f90:146-216 matrix
      elim_mat(q, f, j, sj, real, -t1f + t2f, (bdf.K + t3f) + t1f, -t3f - t2f);
    }
  });
  {
    const bool sw = elim_rhs(q, e, N - 1, N - 1, true, aN, bN, 0.0, Tmax);   // boundary row N
    if (EXT && ELEC && prec) record_tr(N - 1, sw);
  }
  if (e.r0 == 0.0) e.singular = true;
  S2(q, N, SU0) = e.r0; S2(q, N, SU1) = 0.0; S2(q, N, SU2) = 0.0; S2(q, N, SY) = e.rb;
  if (!BE) {
    elim_mat(q, f, N - 1, N - 1, true, aN, bN, 0.0);
    if (f.r0 == 0.0) f.singular = true;
    S2(q, N, SV0) = f.r0; S2(q, N, SV1) = 0.0; S2(q, N, SV2) = 0.0;
    if (EXT && ELEC && prec)                         // the BDF2 interchanges, as stored
      for (int w = 0; w < nw; ++w) prec[nw + w] = q.swb[w * 64];
  }
  // a singular factorisation leaves dgbsv's right-hand side as the solution: for the
  // lanes concerned, store the first solve's right-hand side in the y slot (a slow pass
  // that never runs on physical inputs)
  if (__any(e.singular)) {
    const bool mine = e.singular;
    if (mine) S2(q, 1, SY) = Tmin;
    rows([&](int ro, bool real, double t1f, double t2f, double t3f, double Tm, double Tc,
             double Tp) {
      double a, b, c, d;
      first_row(t1f, t2f, t3f, Tm, Tc, Tp, a, b, c, d);
      if (mine && real) S2(q, ro, SY) = d;
    });
    if (mine) S2(q, N, SY) = Tmax;
  }

  // pass 2
  __builtin_amdgcn_wave_barrier();
  if (BE) {
    backsub(SU0, SU1, SU2, SY, e.singular, pho, put_out, flush_out);
    return;
  }
  backsub(SU0, SU1, SU2, SY, e.singular, 0, [&](int i, int u, double x) { put2(q, i, STR, x); },
          [&](int) {});

  if (PH == 1) {
    *mysing = f.singular ? 1 : 0;
    return;
  }
  }  // passes 1 and 2

  // pass 3: d(r) = AT(r) + ((4/3) TR(r))/(dt/3) for rows 2..N-1 (Tmin, Tmax at the ends),
  // with AT(r) = -((1/3) T(r))/(dt/3) from the input temperature (its column of Ts, which
  // only pass 4 overwrites), through the interchanges and multipliers of pass 1; singular
  // lanes keep d itself.  The input temperatures come in R-row chunks loaded by the wave
  // cooperatively (as pass 1 loads its inputs) into two alternating LDS buffers, the next
  // chunk's loads in flight while a chunk is used.
  {
    constexpr int PF = R;
    double ptr[PF], pl[PF];
    double tv[R];
    auto issue_T = [&](int r0) {                     // rows r0..r0+R-1 (clamped to 1..N)
      const int rr = clampi(r0 + lr, 1, N);
#pragma unroll
      for (int kk = 0; kk < R; ++kk) tv[kk] = s.Ts[so + cbk[kk] + rr];
    };
    auto store_T = [&](double (*buf)[LP]) {
#pragma unroll
      for (int kk = 0; kk < R; ++kk) buf[lr][kk * CPI + cs] = tv[kk];
    };
    const int t0 = 2 - ph3;                          // first row of chunk 0
    __builtin_amdgcn_wave_barrier();
    issue_T(t0);
    store_T(sT);
    __builtin_amdgcn_wave_barrier();
    issue_T(t0 + R);
    auto fetch = [&](int r, int u) {                 // inputs of step j = r-1
      const int rc = clampi(r, 1, N), jc = clampi(r - 1, 1, N - 1);
      ptr[u] = get2(q, rc, STR); pl[u] = get2(q, jc, SL);
    };
    // the interchange words of a chunk's steps j0 .. j0+PF-1 (at most two), a chunk ahead
    auto words = [&](int j0, int& w0, int& w1) {
      w0 = q.swb[(clampi(j0, 1, N - 1) >> 5) * 64];
      w1 = q.swb[(clampi(j0 + PF - 1, 1, N - 1) >> 5) * 64];
    };
#pragma unroll
    for (int u = 0; u < PF; ++u) fetch(t0 + u, u);
    const bool fs = f.singular;
    S2(q, 1, SY) = Tmin;
    double rb = Tmin;
    int cw0, cw1, nw0, nw1;
    words(t0 - 1, cw0, cw1);
    int ch = 0;                                      // chunk of rows t0 + ch R ..
    for (int j0 = t0 - 1; j0 <= N - 1; j0 += PF, ++ch) {
      double (*tb)[LP] = (ch & 1) ? sC : sT;
      words(j0 + PF, nw0, nw1);
      const int jw0 = clampi(j0, 1, N - 1) >> 5;
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const int j = j0 + u, r = j + 1;
        const double at = -((bdf.c13 * tb[u][lane]) / bdf.dt3), trr = ptr[u], l = pl[u];
        const int bits = (clampi(j, 1, N - 1) >> 5) == jw0 ? cw0 : cw1;
        fetch(r + PF, u);
        const bool live = j >= 1 && j <= N - 1;
        const double d = r >= N ? Tmax : at + (bdf.c43 * trr) / bdf.dt3;
        const bool swap = (bits >> (j & 31)) & 1;
        const double pd = swap ? d : rb;
        double od = swap ? rb : d;
        if (!fs && pd != 0.0) od = od + l * (-pd);
        // steps past N-1 store on row N, which the store after the loop overwrites
        put2(q, live ? (fs ? r : j) : N, SY, fs ? d : pd);
        rb = live ? od : rb;
      }
      cw0 = nw0;
      cw1 = nw1;
      __builtin_amdgcn_wave_barrier();
      store_T((ch & 1) ? sT : sC);                   // chunk ch + 1
      __builtin_amdgcn_wave_barrier();
      issue_T(t0 + (ch + 2) * R);
    }
    S2(q, N, SY) = fs ? Tmax : rb;
  }

  // pass 4
  backsub(SV0, SV1, SV2, SY, f.singular, pho, put_out, flush_out);
}

// PH 0: the whole solve; for a solve split in two launches, PH 1 runs passes 1-2 and
// PH 2 passes 3-4 (shorter waves, so the last wave round leaves less of the GPU idle)
template <int R, bool BE, int PH, bool EXT = false>
__global__ __launch_bounds__(64) void k_diffusion(DevGrid g, DevState s, double* scratch,
                                                  int* swbits, int* lst, Coef k1, Coef bdf,
                                                  double Teinf) {
  __shared__ double sT[R][LP], sC[R][LP], sB[R][LP], sD[R][LP];
  const long ncol = (long)g.lx2 * g.lx3;
  const long gps = (ncol + 63) / 64;                 // column groups per species
  // the electron groups first (their waves are the longest, so the last wave round is
  // shorter), then the six ion species of each run of 8 groups; block b runs on XCD b % 8,
  // and every species of column group cg on XCD cg % 8
  const long ne = (gps + 7) / 8 * 8;
  long cg;
  int isp;
  if ((long)blockIdx.x < ne) {
    cg = blockIdx.x;
    isp = LSP;
  } else {
    const long b = blockIdx.x - ne, slot = b / 8;
    cg = (slot / (LSP - 1)) * 8 + b % 8;
    isp = (int)(slot % (LSP - 1)) + 1;
  }
  if (cg >= gps) return;
  if (isp == LSP)
    diffusion_group<R, BE, true, PH, EXT>(g, s, scratch, swbits, lst, k1, bdf, Teinf, isp, cg,
                                          gps, sT, sC, sB, sD);
  else
    diffusion_group<R, BE, false, PH, EXT>(g, s, scratch, swbits, lst, k1, bdf, Teinf, isp, cg,
                                           gps, sT, sC, sB, sD);
}

}  // namespace

void launch_floor_ghosts(const DevGrid& g, double* Ts, hipStream_t st);

namespace {
// layout of a grid's diffusion work space: the pivot-row slots of every column group, the
// BDF2 interchange words, one singular flag per lane
__host__ __device__ inline long diffusion_groups(const DevGrid& g) {
  return ((long)g.lx2 * g.lx3 + 63) / 64 * LSP;
}
__host__ __device__ inline size_t factor_bytes(const DevGrid& g) {
  return (size_t)diffusion_groups(g) * (size_t)g.lx1 * NSLOT * 64 * sizeof(double);
}
__host__ __device__ inline size_t swap_bytes(const DevGrid& g) {
  return (size_t)diffusion_groups(g) * ((g.lx1 + 31) / 32) * 64 * sizeof(int);
}

// The TRBDF2 solve of every member of a batched step: a persistent grid of one-wave blocks
// taking the members' column groups in turn, numbered as each member's own launch numbers
// them (every member's count is a multiple of 8, so a group keeps its XCD).  The
// coefficients are formed from each member's dt as launch_energy_diffusion forms them.
template <int R>
__global__ __launch_bounds__(64, 2) void kb_diffusion(const BatchMember* M, const long* pre, int n,
                                                   long total) {
  __shared__ double sT[R][LP], sC[R][LP], sB[R][LP], sD[R][LP];
  const long* p = pre + (long)BK_DIFF * (n + 1);
  for (long t = blockIdx.x; t < total; t += gridDim.x) {
    const int m = batch_member(p, n, t);
    const DevGrid& g = M[m].g;
    const DevState& s = M[m].s;
    const long v = t - p[m];
    const long ncol = (long)g.lx2 * g.lx3;
    const long gps = (ncol + 63) / 64;
    const long ne = (gps + 7) / 8 * 8;               // (k_diffusion's numbering)
    long cg;
    int isp;
    if (v < ne) {
      cg = v;
      isp = LSP;
    } else {
      const long b = v - ne, slot = b / 8;
      cg = (slot / (LSP - 1)) * 8 + b % 8;
      isp = (int)(slot % (LSP - 1)) + 1;
    }
    if (cg >= gps) continue;
    const double dt = M[m].dt;
    Coef tr{}, bdf{};
    tr.dt2 = dt / 2.0; tr.K = 1.0 / tr.dt2; tr.dt = dt;
    bdf.dt3 = dt / 3.0; bdf.K = 1.0 / bdf.dt3; bdf.dt = dt;
    bdf.c13 = 1.0 / 3.0; bdf.c43 = 4.0 / 3.0;
    double* scratch = M[m].scratch;
    int* swbits = (int*)((char*)scratch + factor_bytes(g));
    int* lst = (int*)((char*)swbits + swap_bytes(g));
    if (isp == LSP)
      diffusion_group<R, false, true, 0, false>(g, s, scratch, swbits, lst, tr, bdf, M[m].Teinf,
                                                isp, cg, gps, sT, sC, sB, sD);
    else
      diffusion_group<R, false, false, 0, false>(g, s, scratch, swbits, lst, tr, bdf, M[m].Teinf,
                                                 isp, cg, gps, sT, sC, sB, sD);
  }
}
}  // namespace


// the metric factors in column-group order (DevGrid::hcol), one thread per entry; lanes
// past the last column repeat it
__global__ void k_fill_hcol(DevGrid g, double* hcol) {
  const long ncol = (long)g.lx2 * g.lx3;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < g.hstride;
       t += (long)gridDim.x * blockDim.x) {
    const int lane = (int)(t & 63);
    const long q = t >> 6;
    const int r = (int)(q % g.lx1) + 1;
    const long gcol = group_column(g, q / g.lx1, lane);
    const long cb = gix(g, 0, (int)(gcol % g.lx2) + 1, (int)(gcol / g.lx2) + 1);
    hcol[t] = g.h1[cb + r];
    hcol[g.hstride + t] = g.h2[cb + r];
    hcol[2 * g.hstride + t] = g.h3[cb + r];
  }
}

long pivot_record_ints(const DevGrid& g) { return (long)g.lx2 * g.lx3 * 2 * ((g.lx1 + 31) / 32); }

long diffusion_hcol_doubles(const DevGrid& g) {
  return 3 * (((long)g.lx2 * g.lx3 + 63) / 64) * g.lx1 * 64;
}

void launch_fill_hcol(const DevGrid& g, double* hcol, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_hcol, dim3(4096), dim3(256), 0, st, g, hcol);
}

// pivot rows, right-hand sides and solutions, the BDF2 interchange bits, one flag per lane
size_t diffusion_scratch_bytes(const DevGrid& g) {
  return factor_bytes(g) + swap_bytes(g) + (size_t)diffusion_groups(g) * 64 * sizeof(int);
}

namespace {
template <int R>
void diffusion_launch(const DevGrid& g, const DevState& s, int diffsolvetype, double Teinf,
                      double* scratch, const Coef& tr, const Coef& bdf, const Coef& be,
                      hipStream_t st) {
  const long ncol = (long)g.lx2 * g.lx3;
  const long blocks = ((ncol + 63) / 64 + 7) / 8 * 8 * LSP;   // XCD-major numbering
  int* swbits = (int*)((char*)scratch + factor_bytes(g));
  int* lst = (int*)((char*)swbits + swap_bytes(g));
  const dim3 grd((unsigned)blocks), blk(64);
  if (diffsolvetype == 1) {
    hipLaunchKernelGGL((k_diffusion<R, true, 0>), grd, blk, 0, st, g, s, scratch, swbits, lst, be,
                       bdf, Teinf);
  } else if (s.kappa || s.qtop) {                     // fitting extension (TRBDF2 only; the
    // 4-row chunk keeps two waves per SIMD with the extension's extra registers)
    hipLaunchKernelGGL((k_diffusion<4, false, 0, true>), grd, blk, 0, st, g, s, scratch, swbits,
                       lst, tr, bdf, Teinf);
  } else if (DIFF_SPLIT) {
    hipLaunchKernelGGL((k_diffusion<R, false, 1>), grd, blk, 0, st, g, s, scratch, swbits, lst, tr,
                       bdf, Teinf);
    hipLaunchKernelGGL((k_diffusion<R, false, 2>), grd, blk, 0, st, g, s, scratch, swbits, lst, tr,
                       bdf, Teinf);
  } else {
    hipLaunchKernelGGL((k_diffusion<R, false, 0>), grd, blk, 0, st, g, s, scratch, swbits, lst, tr,
                       bdf, Teinf);
  }
}
}  // namespace

void launch_batch_diffusion(const BatchTable& t, int diff_chunk, hipStream_t st) {
  const long total = t.total[BK_DIFF];
  const unsigned b = (unsigned)(total < t.blocks_diff ? total : t.blocks_diff);
  if (diff_chunk == 4)
    hipLaunchKernelGGL(kb_diffusion<4>, dim3(b ? b : 1), dim3(64), 0, st, t.m, t.pre, t.n, total);
  else
    hipLaunchKernelGGL(kb_diffusion<DIFF_CHUNK>, dim3(b ? b : 1), dim3(64), 0, st, t.m, t.pre, t.n,
                       total);
  // (no ghost floor: the batched step's surface conversions floored the ghost temperatures
  // at 100 K, and the solve writes interior rows only)
}

void launch_energy_diffusion(const DevGrid& g, const DevState& s, double dt, int diffsolvetype,
                             double Teinf, double* scratch, hipStream_t st, const Tuning& tune,
                             bool floor_ghosts) {
  Coef tr{}, bdf{}, be{};
  tr.dt2 = dt / 2.0; tr.K = 1.0 / tr.dt2; tr.dt = dt;
  bdf.dt3 = dt / 3.0; bdf.K = 1.0 / bdf.dt3; bdf.dt = dt;
  bdf.c13 = 1.0 / 3.0; bdf.c43 = 4.0 / 3.0;
  be.dt = dt; be.K = 1.0 / dt;
  timing_begin("trbdf2", st);
  if (tune.diff_chunk == 4) diffusion_launch<4>(g, s, diffsolvetype, Teinf, scratch, tr, bdf, be, st);
  else diffusion_launch<DIFF_CHUNK>(g, s, diffsolvetype, Teinf, scratch, tr, bdf, be, st);
  timing_end("trbdf2", st);
  if (!floor_ghosts) return;
  timing_begin("floor_ghosts", st);
  launch_floor_ghosts(g, s.Ts, st);
  timing_end("floor_ghosts", st);
}

}  // namespace gemhip
