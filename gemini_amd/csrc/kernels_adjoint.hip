// kernels_adjoint.hip -- discrete adjoint of the electron energy path of one transport step.
//
// Only the electron temperature depends on the fitting parameters (a per-column scale
// kappa on the electron thermal conductivity and a downward top heat flux q), so the
// adjoint state is one field, lambda = dL/dTe.  The forward step is, for electrons
// (reference order src/libgemini.cpp:155-208):
//   e = ne kB Te / (gamma-1)                           T2rhoe        multifluid.f90:193-202
//   ghost fill of e, periodic x3 wrap                  advec.f90:21-149, mpihalo.f90:147-152
//   e <- S3 e, e <- S1 e, x3 wrap, e <- S2 e           MC-limited sweeps, advec.f90:233-447
//   Te' = max((gamma-1) e / (kB max(ne', nmin)), 100)  rhoe2T        multifluid.f90:179-189
//   Te'' = max(TRBDF2(Te'; kappa, q), 100)             PDEparabolic.f90:19-218
// Each map is differentiated exactly with the branches the forward took (upwind side,
// minmod selection, floors) frozen, and the adjoint runs the maps in reverse:
//   * TRBDF2: the two pivoted band solves are recomputed per column as the forward does,
//     then the transposed systems M2^T w2 = mu and M1^T w1 = (4/dt) w2 are solved with
//     the stored factors (U^T forward substitution, then the row eliminations and
//     interchanges transposed in reverse order), and w1, w2 are chained through the
//     matrix and right-hand-side entries to the conductivity (d lambda_e / dT in closed
//     form), kappa and q;
//   * sweeps: lambda_in(j) = lambda_out(j) + sum over the interface fluxes and limited
//     slopes that read cell j, gathered per cell (no atomics);
//   * ghost fill and wrap: each overwritten ghost hands its adjoint to the cell it was
//     copied or scaled from and is cleared;
//   * conversions: pointwise.
// The inputs each map needs (electron densities, the energy density entering each sweep,
// interface velocities, the temperature entering the diffusion) are taken from a tape
// recorded by the forward step.
#include "diffusion_cell.h"

namespace gemhip {

namespace {

constexpr int TPB = 256;

inline unsigned grid_for(long n) {
  long b = (n + TPB - 1) / TPB;
  return (unsigned)(b > 8192 ? 8192 : b);          // grid-stride loops cover the rest
}
inline unsigned grid_exact(long n) { return (unsigned)((n + TPB - 1) / TPB); }

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ------------------------------------------------------------------ tape helpers
// electron interface velocities (advec.f90:155-229), the values the forward sweeps use
__global__ void k_tape_vels(DevGrid g, const double* vs1, const double* vs2, const double* vs3,
                            double* v1, double* v2, double* v3) {
  const long n1 = (long)(g.lx1 + 1) * g.lx2 * g.lx3;
  const long n2 = (long)g.lx1 * (g.lx2 + 1) * g.lx3;
  const long n3 = (long)g.lx1 * g.lx2 * (g.lx3 + 1);
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n1 + n2 + n3;
       t += (long)gridDim.x * blockDim.x) {
    if (t < n1) {
      const int k = (int)(t % (g.lx1 + 1)) + 1;
      const long line = t / (g.lx1 + 1);
      const int i2 = (int)(line % g.lx2) + 1, i3 = (int)(line / g.lx2) + 1;
      v1[t] = v1i_at(g, vs1 + gix(g, 0, i2, i3), k);
    } else if (t < n1 + n2) {
      const long u = t - n1;
      const int i1 = (int)(u % g.lx1) + 1;
      const long r = u / g.lx1;
      const int k = (int)(r % (g.lx2 + 1)) + 1, i3 = (int)(r / (g.lx2 + 1)) + 1;
      v2[u] = 0.5 * (vs2[gix(g, i1, k - 1, i3)] + vs2[gix(g, i1, k, i3)]);
    } else {
      const long u = t - n1 - n2;
      const int i1 = (int)(u % g.lx1) + 1;
      const long r = u / g.lx1;
      const int i2 = (int)(r % g.lx2) + 1, k = (int)(r / g.lx2) + 1;
      v3[u] = 0.5 * (vs3[gix(g, i1, i2, k - 1)] + vs3[gix(g, i1, i2, k)]);
    }
  }
}

// ------------------------------------------------------------------ pointwise maps
// rhoe2T (with its floor) and the diffusion's floor on ghost cells, transposed:
// lambda <- lambda [Te' > 100] (gamma-1)/kB / max(ne', nmin)
__global__ void k_adj_rhoe2T(DevGrid g, double* lam, const double* TD, const double* NE1) {
  const double c = (gammas(LSP) - 1.0) / kB;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < g.N;
       i += (long)gridDim.x * blockDim.x)
    lam[i] = TD[i] > 100.0 ? lam[i] * (c / fmax_f(NE1[i], mindensdiv)) : 0.0;
}

// T2rhoe transposed: out = in ne kB / (gamma-1)
__global__ void k_adj_T2rhoe(DevGrid g, const double* in, double* out, const double* NE0) {
  const double gm1 = gammas(LSP) - 1.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < g.N;
       i += (long)gridDim.x * blockDim.x)
    out[i] = in[i] * ((NE0[i] * kB) / gm1);
}

// ------------------------------------------------------------------ ghost fill, wrap
// advec.f90:107-132 transposed (x3 ends, every x1 and x2 position)
__global__ void k_adj_bc_x3(DevGrid g, double* lam) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.L1 * g.L2) return;
  const int i1 = (int)(t % g.L1) - 1, i2 = (int)(t / g.L1) - 1;
  if (g.isglobal[5]) {
    const long a = gix(g, i1, i2, g.lx3 + 1), b = gix(g, i1, i2, g.lx3 + 2);
    lam[gix(g, i1, i2, g.lx3)] += lam[a] + lam[b];
    lam[a] = 0.0;
    lam[b] = 0.0;
  }
  if (g.isglobal[4]) {
    const long a = gix(g, i1, i2, 0), b = gix(g, i1, i2, -1);
    lam[gix(g, i1, i2, 1)] += lam[a] + lam[b];
    lam[a] = 0.0;
    lam[b] = 0.0;
  }
}

// advec.f90:83-104 transposed (x2 ends, every x1 and x3 position)
__global__ void k_adj_bc_x2(DevGrid g, double* lam) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.L1 * g.L3) return;
  const int i1 = (int)(t % g.L1) - 1, i3 = (int)(t / g.L1) - 1;
  if (g.isglobal[3]) {
    const long a = gix(g, i1, g.lx2 + 1, i3), b = gix(g, i1, g.lx2 + 2, i3);
    lam[gix(g, i1, g.lx2, i3)] += lam[a] + lam[b];
    lam[a] = 0.0;
    lam[b] = 0.0;
  }
  if (g.isglobal[2]) {
    const long a = gix(g, i1, 0, i3), b = gix(g, i1, -1, i3);
    lam[gix(g, i1, 1, i3)] += lam[a] + lam[b];
    lam[a] = 0.0;
    lam[b] = 0.0;
  }
}

// advec.f90:42-79 transposed for the energy density: e(0) = (e(1)/n(1)) n(0) etc., with
// the densities as the fill leaves them when e is filled (interior columns get their
// density ghosts extrapolated first, advec.f90:45-50)
__global__ void k_adj_bc_x1(DevGrid g, double* lam, const double* NE0) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.L2 * g.L3) return;
  const int i2 = (int)(t % g.L2) - 1, i3 = (int)(t / g.L2) - 1;
  const bool inner = i2 >= 1 && i2 <= g.lx2 && i3 >= 1 && i3 <= g.lx3;
  const long o = gix(g, 0, i2, i3);
  const double* n = NE0 + o;         // n[i1]
  double* l = lam + o;
  const int lx1 = g.lx1;
  if (g.isglobal[1]) {
    double nP1 = n[lx1 + 1], nP2 = n[lx1 + 2];
    if (inner) {
      const double coeff = n[lx1 - 1] / n[lx1 - 2];
      nP1 = fmin_f(coeff * n[lx1], n[lx1]);
      nP2 = fmin_f(coeff * nP1, nP1);
    }
    l[lx1] += (nP1 / n[lx1]) * l[lx1 + 1] + (nP2 / n[lx1]) * l[lx1 + 2];
    l[lx1 + 1] = 0.0;
    l[lx1 + 2] = 0.0;
  }
  if (g.isglobal[0]) {
    double n0 = n[0], nm1 = n[-1];
    if (inner) {
      const double coeff = n[2] / n[3];
      n0 = fmin_f(coeff * n[1], n[1]);
      nm1 = fmin_f(coeff * n0, n0);
    }
    l[1] += (n0 / n[1]) * l[0] + (nm1 / n[1]) * l[-1];
    l[0] = 0.0;
    l[-1] = 0.0;
  }
}

// mpihalo.f90:147-152 (one process along x3, depth 2) transposed; needs lx3 >= 2
__global__ void k_adj_halo(DevGrid g, double* lam) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.L1 * g.lx2) return;
  const int i1 = (int)(t % g.L1) - 1, i2 = (int)(t / g.L1) + 1;
  for (int h = 0; h < 2; ++h) {
    const long gh = gix(g, i1, i2, -1 + h);
    lam[gix(g, i1, i2, g.lx3 - 1 + h)] += lam[gh];
    lam[gh] = 0.0;
  }
  for (int h = 0; h < 2; ++h) {
    const long gh = gix(g, i1, i2, g.lx3 + 1 + h);
    lam[gix(g, i1, i2, 1 + h)] += lam[gh];
    lam[gh] = 0.0;
  }
}

// ------------------------------------------------------------------ sweeps
// advec.f90:384-435 along direction DIR, transposed with the limiter and upwind choices
// the forward made on f.  One thread per ghosted cell; cells that are not on an interior
// line of the direction are left unchanged by the sweep, so their adjoint is copied.
//   out(i) = f(i) - dt (ha2i(i+1) phi(i+1) - ha2i(i) phi(i)) / dxi(i) / ha1(i), 1 <= i <= lx
//   phi(i) = v<0: f(i) v - (v/2)(dx(i) + v dt/h1i(i)) slope(i)
//            v>=0: f(i-1) v + (v/2)(dx(i) - v dt/h1i(i)) slope(i-1)
//   slope(k) = minmod(central, minmod(2 left, 2 right)), 0 <= k <= lx+1
template <int DIR>
__global__ void k_adj_sweep(DevGrid g, const double* lam_out, double* lam_in, const double* f,
                            const double* vi, double dt) {
  const int lx = DIR == 1 ? g.lx1 : DIR == 2 ? g.lx2 : g.lx3;
  const double* dx = DIR == 1 ? g.dx1 : DIR == 2 ? g.dx2 : g.dx3;           // dx[k], k=0..lx+2
  const double* dxi = (DIR == 1 ? g.dx1i : DIR == 2 ? g.dx2i : g.dx3i) - 1;  // dxi[k], k=1..lx
  const long st = DIR == 1 ? 1L : DIR == 2 ? (long)g.L1 : (long)g.L1 * g.L2;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < g.N;
       c += (long)gridDim.x * blockDim.x) {
    const int i1 = (int)(c % g.L1) - 1;
    const int i2 = (int)((c / g.L1) % g.L2) - 1;
    const int i3 = (int)(c / ((long)g.L1 * g.L2)) - 1;
    int j, a, b;
    bool line;
    if (DIR == 1) { j = i1; a = i2; b = i3; line = a >= 1 && a <= g.lx2 && b >= 1 && b <= g.lx3; }
    else if (DIR == 2) { j = i2; a = i1; b = i3; line = a >= 1 && a <= g.lx1 && b >= 1 && b <= g.lx3; }
    else { j = i3; a = i1; b = i2; line = a >= 1 && a <= g.lx1 && b >= 1 && b <= g.lx2; }
    if (!line || lx <= 1) { lam_in[c] = lam_out[c]; continue; }
    auto cell = [&](int k) { return c + (long)(k - j) * st; };
    auto fidx = [&](int i) -> long {                 // interface i of this line
      if (DIR == 1) return (i - 1) + (long)(g.lx1 + 1) * ((a - 1) + (long)g.lx2 * (b - 1));
      if (DIR == 2) return (a - 1) + (long)g.lx1 * ((i - 1) + (long)(g.lx2 + 1) * (b - 1));
      return (a - 1) + (long)g.lx1 * ((b - 1) + (long)g.lx2 * (i - 1));
    };
    // p(k) = dt lam_out(k) / dxi(k) / ha1(k) on the updated cells, k = j-2..j+2
    double P[5];
#pragma unroll
    for (int o = 0; o < 5; ++o) {
      const int k = j - 2 + o;
      P[o] = 0.0;
      if (k >= 1 && k <= lx) {
        const long ck = cell(k);
        P[o] = ((dt * lam_out[ck]) / dxi[k]) / ((g.h1[ck] * g.h2[ck]) * g.h3[ck]);
      }
    }
    // interfaces i = j-1..j+2: flux adjoint and its two upwind coefficients
    double LPHI[4], V[4], CM[4], CP[4];
    bool IV[4];
#pragma unroll
    for (int o = 0; o < 4; ++o) {
      const int i = j - 1 + o;
      IV[o] = i >= 1 && i <= lx + 1;
      LPHI[o] = V[o] = CM[o] = CP[o] = 0.0;
      if (IV[o]) {
        const long fi = fidx(i);
        double ha2i, h1i;
        if (DIR == 1) { ha2i = g.h2x1i[fi] * g.h3x1i[fi]; h1i = g.h1x1i[fi]; }
        else if (DIR == 2) { ha2i = g.h1x2i[fi] * g.h3x2i[fi]; h1i = g.h2x2i[fi]; }
        else { ha2i = g.h1x3i[fi] * g.h2x3i[fi]; h1i = g.h3x3i[fi]; }
        const double v = vi[fi];
        V[o] = v;
        LPHI[o] = ha2i * (P[o + 1] - P[o]);
        CM[o] = -((0.5 * v) * (dx[i] + (v / h1i) * dt));
        CP[o] = (0.5 * v) * (dx[i] - (v / h1i) * dt);
      }
    }
    // f(j-2..j+2) for the limiter choices
    double F[5];
#pragma unroll
    for (int o = 0; o < 5; ++o) {
      const int k = j - 2 + o;
      F[o] = (k >= -1 && k <= lx + 2) ? f[cell(k)] : 0.0;
    }
    double acc = lam_out[c];
    if (IV[1] && V[1] < 0.0) acc += V[1] * LPHI[1];          // phi(j) reads f(j) when v<0
    if (IV[2] && V[2] >= 0.0) acc += V[2] * LPHI[2];         // phi(j+1) reads f(j) when v>=0
#pragma unroll
    for (int o = 0; o < 3; ++o) {                            // slopes k = j-1, j, j+1
      const int k = j - 1 + o;
      if (k < 0 || k > lx + 1) continue;
      double ls = 0.0;                                       // d L / d slope(k)
      if (IV[o] && V[o] < 0.0) ls += LPHI[o] * CM[o];
      if (IV[o + 1] && V[o + 1] >= 0.0) ls += LPHI[o + 1] * CP[o + 1];
      if (ls == 0.0) continue;
      const double fm = F[o], f0 = F[o + 1], fp = F[o + 2];  // f(k-1), f(k), f(k+1)
      const double lsl = (f0 - fm) / dx[k], rsl = (fp - f0) / dx[k + 1];
      const double csl = (fp - fm) / (dx[k] + dx[k + 1]);
      const double l2 = 2.0 * lsl, r2 = 2.0 * rsl;
      int msel;
      double m;
      if (l2 * r2 <= 0.0) { m = 0.0; msel = 0; }
      else if (fabs(l2) < fabs(r2)) { m = l2; msel = 2; }
      else { m = r2; msel = 3; }
      int sel;
      if (csl * m <= 0.0) sel = 0;
      else if (fabs(csl) < fabs(m)) sel = 1;
      else sel = msel;
      // d slope(k) / d f(j) for the chosen form
      if (o == 1) {
        if (sel == 2) acc += (2.0 * ls) / dx[k];
        else if (sel == 3) acc -= (2.0 * ls) / dx[k + 1];
      } else if (o == 0) {                                   // f(j) = f(k+1)
        if (sel == 1) acc += ls / (dx[k] + dx[k + 1]);
        else if (sel == 3) acc += (2.0 * ls) / dx[k + 1];
      } else {                                               // f(j) = f(k-1)
        if (sel == 1) acc -= ls / (dx[k] + dx[k + 1]);
        else if (sel == 2) acc -= (2.0 * ls) / dx[k];
      }
    }
    lam_in[c] = acc;
  }
}

// ------------------------------------------------------------------ TRBDF2
// d lambda_e / dT of the electron conductivity (collisions.f90:185-195)
__device__ __forceinline__ double lambda_e_dT(const CellIn<true>& c, double lam_e) {
  const double T = c.T, sq = sqrt(T);
  const double br = ((c.n1 * 1.1e-16) * (1.0 + 5.7e-4 * T) + ((c.n2 * 2.82e-17) * sq) * (1.0 - 1.21e-4 * T))
                    + (c.n3 * 2.2e-16) * (1.0 + 3.6e-2 * sq);
  const double dbr = (c.n1 * 1.1e-16) * 5.7e-4
                     + (c.n2 * 2.82e-17) * ((0.5 / sq) * (1.0 - 1.21e-4 * T) - sq * 1.21e-4)
                     + (c.n3 * 2.2e-16) * (3.6e-2 * (0.5 / sq));
  const double den = 1.0 + ((3.22e4 * (T * T)) / c.n) * br;
  const double dden = (3.22e4 / c.n) * ((2.0 * T) * br + (T * T) * dbr);
  return lam_e * (2.5 / T - dden / den);
}

// scratch slots per (row, column), column-interleaved: slot k of row i of column col at
// scr[((i-1)*NSA + k)*ncol + col]
enum : int {
  A0 = 0, A1, A2, AY, AL, AS,      // TR: pivot row, rhs -> U^T solve -> w1, multiplier, swap
  B0, B1, B2, BY, BL, BS,          // BDF2: the same, BY: rhs -> mu -> U^T solve -> w2
  X1, X2,                          // TR and BDF2 solutions
  NSA
};

struct Col {
  double* scr;
  long ncol;
  __device__ double& operator()(int i, int k) const {
    return scr[((long)(i - 1) * NSA + k) * ncol];
  }
};

struct AdjArgs {
  const double* TD;      // electron temperature entering the diffusion (ghosted field)
  const double* NE1;     // electron density at the diffusion (ghosted field)
  double* lam;           // in: dL/dTe after the step; out: dL/dTe' (interior rows)
  double* gk;            // += dL/dkappa per column
  double* gq;            // += dL/dq per column (null without a heat flux)
  double dt, Teinf;
};

// pivoted elimination step bringing in row (a, b, cc | d) against the pending row
// (r0, r1 | rb), as dgbtf2 + dgbtrs do it (kernels_diffusion.hip): stores the pivot row,
// its rhs, the multiplier and the interchange
__device__ __forceinline__ bool adj_elim(const Col& q, int j, int s0, double& r0, double& r1,
                                         double* rb, double a, double b, double cc, double d) {
  const bool swap = fabs(a) > fabs(r0);
  double p0, p1, p2, pd, o0, o1, o2, od;
  if (swap) { p0 = a; p1 = b; p2 = cc; pd = d; o0 = r0; o1 = r1; o2 = 0.0; od = rb ? *rb : 0.0; }
  else { p0 = r0; p1 = r1; p2 = 0.0; pd = rb ? *rb : 0.0; o0 = a; o1 = b; o2 = cc; od = d; }
  double l = 0.0;
  const bool sing = p0 == 0.0;
  if (!sing) {
    const double rp = 1.0 / p0;
    l = rp * o0;
    if (p1 != 0.0) o1 = o1 + l * (-p1);
    if (p2 != 0.0) o2 = o2 + l * (-p2);
    if (rb && pd != 0.0) od = od + l * (-pd);
  }
  q(j, s0) = p0; q(j, s0 + 1) = p1; q(j, s0 + 2) = p2;
  if (rb) q(j, s0 + 3) = pd;
  q(j, s0 + 4) = l;
  q(j, s0 + 5) = swap ? 1.0 : 0.0;
  r0 = o1; r1 = o2;
  if (rb) *rb = od;
  return sing;
}

// Every pass below walks the rows of a column in order with a loop-carried recurrence.
// Row data are fetched P rows ahead into a register ring, so the loads of later rows are
// in flight while the recurrence runs (the loop over a full ring is unrolled, its ring
// slots are registers); the last count % P rows are fetched as they come.
template <int P, int DIR, class Raw, class Fetch, class Body>
__device__ __forceinline__ void stream_rows(int first, int count, Fetch fetch, Body body) {
  Raw ring[P];
#pragma unroll
  for (int u = 0; u < P; ++u) fetch(first + DIR * u, ring[u]);
  const int nfull = count / P;
  for (int c = 0; c < nfull; ++c) {
#pragma unroll
    for (int u = 0; u < P; ++u) {
      const int i = first + DIR * (c * P + u);
      const Raw v = ring[u];
      fetch(i + DIR * P, ring[u]);
      body(i, v);
    }
  }
  for (int k = nfull * P; k < count; ++k) {
    const int i = first + DIR * k;
    Raw v;
    fetch(i, v);
    body(i, v);
  }
}

constexpr int PF = 8;            // ring depth of the light passes
constexpr int PFI = 2;           // ring depth of the passes that read the cell inputs

struct R4 { double a, b, c, d; };
struct R5 { double a, b, c, d, e; };

// dtbsv (upper, no transpose) over stored rows s0..s0+2 with rhs slot sy, into slot sx;
// with mask_lam, also mu(i) = dL/dx(i) through the output floor into slot sy
__device__ __forceinline__ void adj_backsub(const Col& q, int N, int s0, int sy, int sx,
                                            const double* lam = nullptr) {
  double x1 = 0.0, x2 = 0.0;                         // x(i+1), x(i+2)
  auto fetch = [&](int i, R5& r) {
    const int c = clampi(i, 1, N);
    r.a = q(c, s0); r.b = q(c, s0 + 1); r.c = q(c, s0 + 2); r.d = q(c, sy);
    r.e = lam ? lam[c] : 0.0;
  };
  stream_rows<PF, -1, R5>(N, N, fetch, [&](int i, const R5& r) {
    double acc = r.d;
    if (i + 2 <= N && x2 != 0.0) acc = acc - x2 * r.c;
    if (i + 1 <= N && x1 != 0.0) acc = acc - x1 * r.b;
    if (acc != 0.0) acc = acc / r.a;
    q(i, sx) = acc;
    if (lam) q(i, sy) = acc >= 100.0 ? r.e : 0.0;
    x2 = x1;
    x1 = acc;
  });
}

// M^T w = rhs for a matrix factored by adj_elim into slots s0..s0+5.  The rhs is slot
// s0+3, or with cw != 0 the interior rows of cw * (slot BY); slot s0+3 ends holding w.
// U^T z = rhs top-down, then the eliminations transposed bottom-up: z(j) -= l_j z(j+1),
// then the interchange.
__device__ __forceinline__ void adj_solveT(const Col& q, int N, int s0, double cw = 0.0) {
  double zm1 = 0.0, zm2 = 0.0;                       // z(i-1), z(i-2)
  double u1m1 = 0.0, u2m1 = 0.0, u2m2 = 0.0;         // U1(i-1), U2(i-1), U2(i-2)
  const int sr = cw != 0.0 ? BY : s0 + 3;
  stream_rows<PF, 1, R4>(1, N, [&](int i, R4& r) {
    const int c = clampi(i, 1, N);
    r.a = q(c, s0); r.b = q(c, s0 + 1); r.c = q(c, s0 + 2); r.d = q(c, sr);
  }, [&](int i, const R4& r) {
    const double rhs = cw == 0.0 ? r.d : (i >= 2 && i <= N - 1 ? cw * r.d : 0.0);
    const double z = ((rhs - u1m1 * zm1) - u2m2 * zm2) / r.a;
    q(i, s0 + 3) = z;
    zm2 = zm1; zm1 = z;
    u2m2 = u2m1;
    u1m1 = r.b; u2m1 = r.c;
  });
  double carry = q(N, s0 + 3);
  stream_rows<PF, -1, R4>(N - 1, N - 1, [&](int j, R4& r) {
    const int c = clampi(j, 1, N);
    r.a = q(c, s0 + 3); r.b = q(c, s0 + 4); r.c = q(c, s0 + 5);
  }, [&](int j, const R4& r) {
    const double a = r.a - r.b * carry;
    if (r.c != 0.0) {
      q(j + 1, s0 + 3) = a;
    } else {
      q(j + 1, s0 + 3) = carry;
      carry = a;
    }
  });
  q(1, s0 + 3) = carry;
}

// one lane per electron column (x2, x3)
__global__ __launch_bounds__(64) void k_adj_trbdf2(DevGrid g, DevState s, AdjArgs A,
                                                    double* scratch) {
  const long ncol = (long)g.lx2 * g.lx3;
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  const int N = g.lx1;
  const int i2 = (int)(col % g.lx2) + 1, i3 = (int)(col / g.lx2) + 1;
  const long cb = gix(g, 0, i2, i3), kb = (long)N * col;
  const long ni = (long)g.lx1 * g.lx2 * g.lx3;
  Col q{scratch + col, ncol};
  DevState se = s;                                   // the diffusion's inputs from the tape
  se.Ts = const_cast<double*>(A.TD);
  se.ns = const_cast<double*>(A.NE1);
  const double* dx1 = g.dx1;
  const double* dx1i = g.dx1i - 1;
  const double gm1kB = (gammas(LSP) - 1.0) / kB;
  const double dt2 = A.dt / 2.0, K1 = 1.0 / dt2;
  const double dt3 = A.dt / 3.0, K2 = 1.0 / dt3;
  const double c13 = 1.0 / 3.0, c43 = 4.0 / 3.0;
  const double kap = s.kappa ? s.kappa[col] : 1.0;
  const bool flux = s.qtop != nullptr;

  // per-row inputs: T, coefficients C, B, D, conductivity without kappa; in pass I also
  // both solutions and both adjoint solutions
  struct Row { double T, C, B, D, lam_e, h1, h23, x1, x2, w1, w2; CellIn<true> in; };
  auto fetch_in = [&](int r, CellIn<true>& c) {
    cell_load<true, true>(g, se, 0, cb, kb, ni, clampi(r, 1, N), c,
                          hcol_index(g, (int)col, clampi(r, 1, N)), (int)col);
  };
  auto make_row = [&](const CellIn<true>& c, Row& w) {
    w.in = c;
    w.in.kap = kap;
    w.T = c.T;
    w.lam_e = cell_lambda<true, false>(w.in, LSP);
    cell_coeffs<true, true>(w.in, LSP, gm1kB, w.C, w.B, w.D);
    w.h1 = c.h1;
    w.h23 = c.h2 * c.h3;
  };
  double Tmin, Tmax;
  {
    const double Tn1 = s.Tn[kb], TnN = s.Tn[kb + N - 1];
    if (g.gridflag == 0) { Tmin = Tn1; Tmax = TnN; }
    else if (g.gridflag == 1) { Tmax = TnN; Tmin = A.Teinf; }
    else { Tmin = Tn1; Tmax = A.Teinf; }
  }
  double aN = 0.0, bN = 1.0, lamN = 0.0, h1N = 0.0;
  if (flux) {
    CellIn<true> c;
    Row top;
    fetch_in(N, c);
    make_row(c, top);
    lamN = top.lam_e * kap;
    h1N = top.h1;
    aN = -1.0 / dx1[N];
    bN = 1.0 / dx1[N];
    Tmax = (s.qtop[col] * h1N) / lamN;
  }
  auto coeffs = [&](int i, const Row& rm, const Row& rc, const Row& rp, double& s1, double& s2,
                    double& s3) {
    const double Dh = 0.5 * (rm.D + rc.D), Dh1 = 0.5 * (rc.D + rp.D);
    s1 = ((rc.C * Dh) / dx1i[i]) / dx1[i];
    s3 = ((rc.C * Dh1) / dx1i[i]) / dx1[i + 1];
    s2 = rc.B / (dx1[i + 1] + dx1[i]);
  };

  // pass A (top-down): both factorisations, the TR right-hand side eliminated alongside;
  // row r's arrival completes row r-1
  bool sing = false;
  double e0 = 1.0, e1 = 0.0, eb = Tmin;              // TR pending row (row 1: identity)
  double f0 = 1.0, f1 = 0.0;                         // BDF2 pending row
  Row rm, rc, rp;
  stream_rows<PFI, 1, CellIn<true>>(1, N, fetch_in, [&](int r, const CellIn<true>& c) {
    rm = rc; rc = rp;
    make_row(c, rp);
    if (r < 3) return;
    const int i = r - 1;
    double s1, s2, s3;
    coeffs(i, rm, rc, rp, s1, s2, s3);
    const double t1 = s1 / 2.0, t2 = s2 / 2.0, t3 = s3 / 2.0;
    const double a = -t1 + t2, b = (K1 + t3) + t1, cc = -t3 - t2;
    const double d = ((rc.T / dt2 - a * rm.T) - (t3 + t1) * rc.T) - cc * rp.T;
    sing |= adj_elim(q, i - 1, A0, e0, e1, &eb, a, b, cc, d);
    sing |= adj_elim(q, i - 1, B0, f0, f1, nullptr, -s1 + s2, (K2 + s3) + s1, -s3 - s2, 0.0);
  });
  sing |= adj_elim(q, N - 1, A0, e0, e1, &eb, aN, bN, 0.0, Tmax);   // boundary row N
  sing |= adj_elim(q, N - 1, B0, f0, f1, nullptr, aN, bN, 0.0, 0.0);
  q(N, A0) = e0; q(N, A1) = 0.0; q(N, A2) = 0.0; q(N, AY) = eb;
  q(N, B0) = f0; q(N, B1) = 0.0; q(N, B2) = 0.0;
  sing |= e0 == 0.0 || f0 == 0.0;
  if (sing) {                                        // no adjoint of a singular solve
    for (int i = 1; i <= N; ++i) A.lam[cb + i] = __builtin_nan("");
    return;
  }

  // pass B: TR solution
  adj_backsub(q, N, A0, AY, X1);
  // pass C: BDF2 right-hand side through the stored eliminations (row 1: Tmin, row N: Tmax)
  {
    double rb = Tmin;
    stream_rows<PF, 1, R4>(1, N - 1, [&](int j, R4& r) {
      const int c = clampi(j, 1, N - 1);
      r.a = A.TD[cb + c + 1]; r.b = q(c + 1, X1); r.c = q(c, BS); r.d = q(c, BL);
    }, [&](int j, const R4& r) {
      const double d = j + 1 >= N ? Tmax : -((c13 * r.a) / dt3) + (c43 * r.b) / dt3;
      const bool swap = r.c != 0.0;
      const double pd = swap ? d : rb;
      double od = swap ? rb : d;
      if (pd != 0.0) od = od + r.d * (-pd);
      q(j, BY) = pd;
      rb = od;
    });
    q(N, BY) = rb;
  }
  // pass D: BDF2 solution, and mu = dL/dx2 through the output floor into BY
  adj_backsub(q, N, B0, BY, X2, A.lam + cb);
  // passes E-F: w2
  adj_solveT(q, N, B0);
  // passes G-H: w1 from dL/dx1 = (4/3)/(dt/3) w2 on the interior rows
  adj_solveT(q, N, A0, c43 / dt3);

  // pass I (top-down): chain w1, w2 through the matrix and rhs entries of rows 2..N-1
  //   dL/dM1(i,k) = -w1(i) (x1(k) + T(k))   (M1 x1 = (2/(dt/2) - M1) T on those rows)
  //   dL/dM2(i,k) = -w2(i) x2(k)
  // to s1 = C Dh(i)/dx1i/dx1(i), s3 = C Dh(i+1)/dx1i/dx1(i+1) (TR uses s/2), then Dh, D,
  // lambda = kappa lambda_e(T), and the direct T terms of both right-hand sides.
  double gk = 0.0, gqc = 0.0;
  double gDh_p = 0.0, gDh_c = 0.0, gDh_n = 0.0;     // dL/dDh(i-1), (i), (i+1) so far
  double gT_m = 0.0, gT_c = 0.0, gT_n = 0.0;         // dL/dT(i-1), (i), (i+1) so far
  // row j complete: D(j) = lambda(j) h2 h3 / h1 enters Dh(j) and Dh(j+1)
  auto finish = [&](int j, const Row& w, double gDh_j, double gDh_j1, double gT_j) {
    const double gD = 0.5 * (gDh_j + gDh_j1);
    double glam = gD * (w.h23 / w.h1);
    if (flux && j == N) {                            // row N rhs: q h1(N) / lambda(N)
      const double gg = w.w1 + w.w2;
      glam += -gg * (Tmax / lamN);
      gqc += gg * (h1N / lamN);
    }
    gk += glam * w.lam_e;
    A.lam[cb + j] = gT_j + glam * (kap * lambda_e_dT(w.in, w.lam_e));
  };
  struct RawI { CellIn<true> in; double x1, x2, w1, w2; };
  stream_rows<PFI, 1, RawI>(1, N, [&](int r, RawI& v) {
    const int c = clampi(r, 1, N);
    fetch_in(c, v.in);
    v.x1 = q(c, X1); v.x2 = q(c, X2); v.w1 = q(c, AY); v.w2 = q(c, BY);
  }, [&](int r, const RawI& v) {
    rm = rc; rc = rp;
    make_row(v.in, rp);
    rp.x1 = v.x1; rp.x2 = v.x2; rp.w1 = v.w1; rp.w2 = v.w2;
    if (r < 3) return;
    const int i = r - 1;
    double s1, s2, s3;
    coeffs(i, rm, rc, rp, s1, s2, s3);
    const double t1 = s1 / 2.0, t2 = s2 / 2.0, t3 = s3 / 2.0;
    const double a = -t1 + t2, cc = -t3 - t2;
    const double w1 = rc.w1, w2 = rc.w2;
    const double G1a = -w1 * (rm.x1 + rm.T), G1b = -w1 * (rc.x1 + rc.T), G1c = -w1 * (rp.x1 + rp.T);
    const double G2a = -w2 * rm.x2, G2b = -w2 * rc.x2, G2c = -w2 * rp.x2;
    const double gs1 = 0.5 * (G1b - G1a) + (G2b - G2a);
    const double gs3 = 0.5 * (G1b - G1c) + (G2b - G2c);
    gDh_c += gs1 * ((rc.C / dx1i[i]) / dx1[i]);
    gDh_n += gs3 * ((rc.C / dx1i[i]) / dx1[i + 1]);
    gT_m += -w1 * a;                                  // row i: -a T(i-1)
    gT_c += w1 * (K1 - (t3 + t1)) - (c13 / dt3) * w2;
    gT_n += -w1 * cc;                                 // row i: -c T(i+1)
    finish(i - 1, rm, gDh_p, gDh_c, gT_m);
    gDh_p = gDh_c; gDh_c = gDh_n; gDh_n = 0.0;
    gT_m = gT_c; gT_c = gT_n; gT_n = 0.0;
  });
  finish(N - 1, rc, gDh_p, gDh_c, gT_m);
  finish(N, rp, gDh_c, 0.0, gT_c);
  if (A.gk) A.gk[col] += gk;
  if (A.gq) A.gq[col] += gqc;
}

}  // namespace

size_t adjoint_scratch_bytes(const DevGrid& g) {
  return (size_t)NSA * g.lx1 * g.lx2 * g.lx3 * sizeof(double);
}

void launch_tape_vels(const DevGrid& g, const DevState& s, double* v1, double* v2, double* v3,
                      hipStream_t st) {
  const long so = (long)(LSP - 1) * g.N;
  const long n = (long)(g.lx1 + 1) * g.lx2 * g.lx3 + (long)g.lx1 * (g.lx2 + 1) * g.lx3 +
                 (long)g.lx1 * g.lx2 * (g.lx3 + 1);
  hipLaunchKernelGGL(k_tape_vels, dim3(grid_for(n)), dim3(TPB), 0, st, g, s.vs1 + so,
                     s.vs2 + so, s.vs3 + so, v1, v2, v3);
}

void launch_adjoint_step(const DevGrid& g, const DevState& s, const AdjTape& t, double dt,
                         double Teinf, double* lam, double* work, double* gk, double* gq,
                         double* scratch, hipStream_t st) {
  const long ncol = (long)g.lx2 * g.lx3;
  const dim3 cells(grid_for(g.N)), blk(TPB);
  AdjArgs A{t.TD, t.NE1, lam, gk, gq, dt, Teinf};
  timing_begin("adj_trbdf2", st);
  hipLaunchKernelGGL(k_adj_trbdf2, dim3((unsigned)((ncol + 63) / 64)), dim3(64), 0, st, g, s, A,
                     scratch);
  timing_end("adj_trbdf2", st);
  timing_begin("adj_advection", st);
  hipLaunchKernelGGL(k_adj_rhoe2T, cells, blk, 0, st, g, lam, t.TD, t.NE1);
  hipLaunchKernelGGL(k_adj_sweep<2>, cells, blk, 0, st, g, lam, work, t.ES2, t.V2, dt);
  const dim3 hal(grid_exact((long)g.L1 * g.lx2));
  if (g.flagper) hipLaunchKernelGGL(k_adj_halo, hal, blk, 0, st, g, work);
  hipLaunchKernelGGL(k_adj_sweep<1>, cells, blk, 0, st, g, work, lam, t.ES1, t.V1, dt);
  hipLaunchKernelGGL(k_adj_sweep<3>, cells, blk, 0, st, g, lam, work, t.ES3, t.V3, dt);
  if (g.flagper) hipLaunchKernelGGL(k_adj_halo, hal, blk, 0, st, g, work);
  else hipLaunchKernelGGL(k_adj_bc_x3, dim3(grid_exact((long)g.L1 * g.L2)), blk, 0, st, g, work);
  hipLaunchKernelGGL(k_adj_bc_x2, dim3(grid_exact((long)g.L1 * g.L3)), blk, 0, st, g, work);
  hipLaunchKernelGGL(k_adj_bc_x1, dim3(grid_exact((long)g.L2 * g.L3)), blk, 0, st, g, work, t.NE0);
  hipLaunchKernelGGL(k_adj_T2rhoe, cells, blk, 0, st, g, work, lam, t.NE0);
  timing_end("adj_advection", st);
}

}  // namespace gemhip
