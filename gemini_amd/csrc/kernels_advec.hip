// kernels_advec.hip -- advection substep of the transport path on gfx950.
//
// Reference: src/numerical/advection/advec.f90 (ghost fill :21-149, interface
// velocities :155-229, sweeps :233-381, MC limiter :384-447),
// src/multifluid/multifluid.f90:49-126 (species/field selection) and :179-233
// (state conversions), src/mpimod/mpihalo.f90:147-152 (periodic x3 wrap).
//
// Every kernel keeps the reference's arithmetic order; the file is built with
// -ffp-contract=off so results are bit-identical to the Fortran/C restatement.
// Arrays are the reference's Fortran layouts (x1 contiguous), so:
//  * x1 sweeps give each (x2,x3) line to one workgroup and stage it in LDS,
//  * x2 and x3 sweeps map threads onto the contiguous x1 index and march each
//    thread along its line, keeping the MC stencil in registers -- every load and
//    store is a coalesced 64-lane access and no LDS or barrier is needed.
#include <algorithm>
#include <cstdint>
#include <initializer_list>

#include "gemhip_internal.h"

namespace gemhip {

namespace {

constexpr int TPB = 256;

inline int blocks_for(long n, int tpb = TPB) {
  long b = (n + tpb - 1) / tpb;
  return (int)(b > 65535L * 16 ? 65535L * 16 : b);
}

// ---------------------------------------------------------------- conversions
// multifluid.f90:193-214 (T2rhoe, v12rhov1) over the whole padded grid
__global__ void k_pre_convert(DevGrid g, DevState s, int mom, int energy) {
  const int isp = blockIdx.y + 1;
  const long o = (long)(isp - 1) * g.N;
  const double msp = ms(isp), gm1 = gammas(isp) - 1.0;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < g.N;
       k += (long)gridDim.x * blockDim.x) {
    const double n = s.ns[o + k];
    if (mom) s.rhovs1[o + k] = (n * msp) * s.vs1[o + k];
    if (energy) s.rhoes[o + k] = ((n * kB) * s.Ts[o + k]) / gm1;
  }
}

// multifluid.f90:218-233 (rhov12v1)
__global__ void k_rhov12v1(DevGrid g, DevState s) {
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < g.N;
       k += (long)gridDim.x * blockDim.x) {
    double chrgflux = 0.0;
#pragma unroll
    for (int isp = 1; isp <= LSP - 1; ++isp) {
      const long o = (long)(isp - 1) * g.N + k;
      const double n = s.ns[o];
      const double v = s.rhovs1[o] / (ms(isp) * fmax_f(n, mindensdiv));
      s.vs1[o] = v;
      chrgflux = chrgflux + (n * qs(isp)) * v;
    }
    double sum = 0.0;
#pragma unroll
    for (int isp = 1; isp <= LSP - 1; ++isp) sum = sum + s.ns[(long)(isp - 1) * g.N + k];
    const long o7 = (long)(LSP - 1) * g.N + k;
    s.ns[o7] = sum;
    s.vs1[o7] = -(((1.0 / fmax_f(sum, mindensdiv)) / qs(LSP)) * chrgflux);
  }
}

// multifluid.f90:179-189 (rhoe2T)
__global__ void k_rhoe2T(DevGrid g, DevState s) {
  const int isp = blockIdx.y + 1;
  const long o = (long)(isp - 1) * g.N;
  const double c = (gammas(isp) - 1.0) / kB;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < g.N;
       k += (long)gridDim.x * blockDim.x) {
    const double t = (c * s.rhoes[o + k]) / fmax_f(s.ns[o + k], mindensdiv);
    s.Ts[o + k] = fmax_f(t, 100.0);
  }
}

// rhov12v1 then rhoe2T of one padded cell: both are pointwise, so doing them cell by cell
// gives what the two whole-array passes give, with ns read once.
// Every input is loaded before the first store (the arrays are distinct, but the compiler
// cannot know that, and would otherwise wait for each species' loads in turn).
__device__ __forceinline__ void post_convert_cell(const DevGrid& g, const DevState& s, long k) {
  double n[LSP], rv[LSP - 1], e[LSP];
#pragma unroll
  for (int isp = 1; isp <= LSP; ++isp) {
    const long o = (long)(isp - 1) * g.N + k;
    if (isp < LSP) {
      n[isp - 1] = s.ns[o];
      rv[isp - 1] = s.rhovs1[o];
    }
    e[isp - 1] = s.rhoes[o];
  }
  double chrgflux = 0.0, sum = 0.0;
#pragma unroll
  for (int isp = 1; isp <= LSP - 1; ++isp) {
    const long o = (long)(isp - 1) * g.N + k;
    const double v = rv[isp - 1] / (ms(isp) * fmax_f(n[isp - 1], mindensdiv));
    s.vs1[o] = v;
    chrgflux = chrgflux + (n[isp - 1] * qs(isp)) * v;
  }
#pragma unroll
  for (int isp = 1; isp <= LSP - 1; ++isp) sum = sum + n[isp - 1];
  n[LSP - 1] = sum;
  const long o7 = (long)(LSP - 1) * g.N + k;
  s.ns[o7] = sum;
  s.vs1[o7] = -(((1.0 / fmax_f(sum, mindensdiv)) / qs(LSP)) * chrgflux);
#pragma unroll
  for (int isp = 1; isp <= LSP; ++isp) {
    const long o = (long)(isp - 1) * g.N + k;
    const double t = (((gammas(isp) - 1.0) / kB) * e[isp - 1]) / fmax_f(n[isp - 1], mindensdiv);
    s.Ts[o] = fmax_f(t, 100.0);
  }
}

__global__ void k_post_convert(DevGrid g, DevState s) {
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < g.N;
       k += (long)gridDim.x * blockDim.x)
    post_convert_cell(g, s, k);
}

// The padded cells outside the interior box, one thread each: first the four x1 ghosts of
// every interior (x2, x3) column, then every cell of the columns with an x2 or x3 ghost
// index.  Returns -1 past the end.
__device__ __forceinline__ int ghost_index(int j, int lx) { return j < 2 ? j - 1 : lx + j - 1; }
__device__ __forceinline__ long surface_cell(const DevGrid& g, long t) {
  const long n1 = 4L * g.lx2 * g.lx3;
  const long ncol = 4L * g.L2 + 4L * g.lx3;
  if (t < n1) {
    const long col = t / 4;
    return gix(g, ghost_index((int)(t % 4), g.lx1), (int)(col % g.lx2) + 1, (int)(col / g.lx2) + 1);
  }
  if (t >= n1 + ncol * g.L1) return -1;
  const long r = t - n1;
  const int i1 = (int)(r % g.L1) - 1;
  const long col = r / g.L1;
  int i2, i3;
  if (col < 4L * g.L2) {
    i2 = (int)(col % g.L2) - 1;
    i3 = ghost_index((int)(col / g.L2), g.lx3);
  } else {
    const long cc = col - 4L * g.L2;
    i3 = (int)(cc % g.lx3) + 1;
    i2 = ghost_index((int)(cc / g.lx3), g.lx2);
  }
  return gix(g, i1, i2, i3);
}

// the fused step's rhov12v1 + rhoe2T of the cells the x2 sweep does not convert
__global__ void k_post_convert_surface(DevGrid g, DevState s) {
  const long c = surface_cell(g, (long)blockIdx.x * blockDim.x + threadIdx.x);
  if (c >= 0) post_convert_cell(g, s, c);
}

// ---------------------------------------------------------------- global ghost fill
// With CONV the fill also stands in for v12rhov1 and T2rhoe (multifluid.f90:193-214),
// which the fused step no longer runs as a pass of their own: every rhovs1 or rhoes value
// the fill reads that the conversions would have written, and that no earlier part of the
// fill has overwritten, is formed on the fly as (n ms) v or ((n kB) T)/(gamma-1), the
// conversions' own expressions, so every ghost value is the one the separate passes
// give.  The interior cells are converted by the x3 sweep when it loads them.  This holds
// with all six sides global and no periodic x3 wrap (the fused step checks).
__device__ __forceinline__ double pre_momentum(double n, double v, int isp) { return (n * ms(isp)) * v; }
__device__ __forceinline__ double pre_energy(double n, double T, int isp) {
  return ((n * kB) * T) / (gammas(isp) - 1.0);
}

// advec.f90:42-79: x1 ends, one thread per padded (x2,x3) column t and species.  The two
// ends mirror each other: end 0 fills cells 0, -1 from cells 1, 2, 3, end 1 fills lx1+1,
// lx1+2 from lx1, lx1-1, lx1-2 (index map m).  Each end's inputs are loaded before the
// stores (the arrays are distinct, which the compiler cannot know), and both ends' before
// either end's stores wherever the far end reads none of the cells the near end writes.
struct X1End { double nm1, n0, n1, n2, n3, u0, u1, s1, r0, vb; };
template <bool CONV>
__device__ __forceinline__ void x1_end_load(const double* n, const double* r, const double* e,
                                            const double* u, const double* T, int lx1, int side,
                                            double vb, X1End& E) {
  auto m = [&](int j) { return side ? lx1 + 1 - j : j; };
  E.nm1 = n[m(-1)]; E.n0 = n[m(0)]; E.n1 = n[m(1)]; E.n2 = n[m(2)]; E.n3 = n[m(3)];
  E.u0 = u[m(0)]; E.u1 = u[m(1)];
  E.s1 = CONV ? T[m(1)] : e[m(1)];
  E.r0 = CONV ? 0.0 : r[m(0)];
  E.vb = vb;
}
template <bool CONV>
__device__ __forceinline__ void x1_end_store(double* n, double* r, double* e, int lx1, int side,
                                             bool inner, int isp, const X1End& E) {
  auto m = [&](int j) { return side ? lx1 + 1 - j : j; };
  const double msp = ms(isp);
  double n0 = E.n0, nm1 = E.nm1, r0;
  if (inner) {
    const double coeff = E.n2 / E.n3;
    n0 = fmin_f(coeff * E.n1, E.n1);
    nm1 = fmin_f(coeff * n0, n0);
    n[m(0)] = n0;
    n[m(-1)] = nm1;
    r0 = 2.0 * E.vb - E.u1;                  // E.vb: v1i at the end's boundary interface
  } else {
    r0 = CONV ? pre_momentum(E.n0, E.u0, isp) : E.r0;
  }
  const double rm = (r0 + r0) - E.u1;
  r[m(-1)] = (rm * nm1) * msp;
  r[m(0)] = (r0 * n0) * msp;
  const double e1 = CONV ? pre_energy(E.n1, E.s1, isp) : E.s1;
  const double Tb = e1 / E.n1;
  e[m(0)] = Tb * n0;
  e[m(-1)] = Tb * nm1;
}
template <bool CONV>
__device__ __forceinline__ void bc_x1_body(const DevGrid& g, const DevState& s, int isp, long t) {
  const int i2 = (int)(t % g.L2) - 1, i3 = (int)(t / g.L2) - 1;
  const bool inner = i2 >= 1 && i2 <= g.lx2 && i3 >= 1 && i3 <= g.lx3;
  const long o = (long)(isp - 1) * g.N + gix(g, 0, i2, i3) - 1;  // index of x1 = -1
  double* n = s.ns + o + 1;        // n[i1] for i1 = -1..lx1+2 -> n[-1..]
  double* r = s.rhovs1 + o + 1;
  double* e = s.rhoes + o + 1;
  const double* u = s.vs1 + o + 1;
  const double* T = s.Ts + o + 1;
  const int lx1 = g.lx1;
  const long vb = 2L * ((i2 - 1) + (long)g.lx2 * (i3 - 1)) + 2L * g.lx2 * g.lx3 * (isp - 1);
  const bool lo = g.isglobal[0], hi = g.isglobal[1];
  const bool early = lx1 >= 3;     // the far end's cells lx1-2.. are not among -1, 0
  X1End A, B;
  if (lo) x1_end_load<CONV>(n, r, e, u, T, lx1, 0, inner ? s.v1b[vb] : 0.0, A);
  if (hi && early) x1_end_load<CONV>(n, r, e, u, T, lx1, 1, inner ? s.v1b[vb + 1] : 0.0, B);
  if (lo) x1_end_store<CONV>(n, r, e, lx1, 0, inner, isp, A);
  if (hi) {
    if (!early) x1_end_load<CONV>(n, r, e, u, T, lx1, 1, inner ? s.v1b[vb + 1] : 0.0, B);
    x1_end_store<CONV>(n, r, e, lx1, 1, inner, isp, B);
  }
}
template <bool CONV>
__global__ void k_bc_x1(DevGrid g, DevState s) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < (long)g.L2 * g.L3) bc_x1_body<CONV>(g, s, blockIdx.y + 1, t);
}

// the value a copying fill reads from cell c: with CONV, momentum (q 2) and energy (q 3) of
// a cell the x1 fill has not written are formed from ns, vs1, Ts
template <bool CONV>
__device__ __forceinline__ double fill_source(const DevState& s, int q, const double* f, long so,
                                              long c, bool filled, int isp) {
  if (!CONV || q < 2 || filled) return f[c];
  const double n = s.ns[so + c];
  return q == 2 ? pre_momentum(n, s.vs1[so + c], isp) : pre_energy(n, s.Ts[so + c], isp);
}

// advec.f90:83-104: x2 ends, every x1 and x3 position t, fields q = vs2, ns, rhovs1, rhoes
template <bool CONV>
__device__ __forceinline__ void bc_x2_body(const DevGrid& g, const DevState& s, int isp, int q,
                                           long t) {
  const int i1 = (int)(t % g.L1) - 1, i3 = (int)(t / g.L1) - 1;
  double* base = q == 0 ? s.vs2 : q == 1 ? s.ns : q == 2 ? s.rhovs1 : s.rhoes;
  const long so = (long)(isp - 1) * g.N;
  double* f = base + so;
  const bool x1ghost = i1 < 1 || i1 > g.lx1;          // written by the x1 fill
  if (g.isglobal[2]) {
    const double v = fill_source<CONV>(s, q, f, so, gix(g, i1, 1, i3), x1ghost, isp);
    f[gix(g, i1, 0, i3)] = v;
    f[gix(g, i1, -1, i3)] = v;
  }
  if (g.isglobal[3]) {
    const double v = fill_source<CONV>(s, q, f, so, gix(g, i1, g.lx2, i3), x1ghost, isp);
    f[gix(g, i1, g.lx2 + 1, i3)] = v;
    f[gix(g, i1, g.lx2 + 2, i3)] = v;
  }
}
template <bool CONV>
__global__ void k_bc_x2(DevGrid g, DevState s) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < (long)g.L1 * g.L3) bc_x2_body<CONV>(g, s, blockIdx.y + 1, blockIdx.z, t);
}

// advec.f90:107-132: x3 ends unless periodic, every x1 and x2 position t
template <bool CONV>
__device__ __forceinline__ void bc_x3_body(const DevGrid& g, const DevState& s, int isp, int q,
                                           long t) {
  const int i1 = (int)(t % g.L1) - 1, i2 = (int)(t / g.L1) - 1;
  double* base = q == 0 ? s.vs3 : q == 1 ? s.ns : q == 2 ? s.rhovs1 : s.rhoes;
  const long so = (long)(isp - 1) * g.N;
  double* f = base + so;
  const bool filled = i1 < 1 || i1 > g.lx1 || i2 < 1 || i2 > g.lx2;   // x1 or x2 fill wrote it
  if (g.isglobal[4]) {
    const double v = fill_source<CONV>(s, q, f, so, gix(g, i1, i2, 1), filled, isp);
    f[gix(g, i1, i2, 0)] = v;
    f[gix(g, i1, i2, -1)] = v;
  }
  if (g.isglobal[5]) {
    const double v = fill_source<CONV>(s, q, f, so, gix(g, i1, i2, g.lx3), filled, isp);
    f[gix(g, i1, i2, g.lx3 + 1)] = v;
    f[gix(g, i1, i2, g.lx3 + 2)] = v;
  }
}
template <bool CONV>
__global__ void k_bc_x3(DevGrid g, DevState s) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < (long)g.L1 * g.L2) bc_x3_body<CONV>(g, s, blockIdx.y + 1, blockIdx.z, t);
}

// ---------------------------------------------------------------- periodic x3 wrap
struct FieldList { double* p[8]; };

// mpihalo.f90:147-152 with one process along x3: every x1 (ghosts included), interior
// x2.  Sources are read before the destinations are written, as Fortran's array
// assignment does, which matters only when lx3 < lhalo.
__global__ void k_halo_periodic(DevGrid g, FieldList fl, int lhalo) {
  const int isp = blockIdx.y + 1;
  double* f = fl.p[blockIdx.z] + (long)(isp - 1) * g.N;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.L1 * g.lx2) return;
  const int i1 = (int)(t % g.L1) - 1, i2 = (int)(t / g.L1) + 1;
  double tmp[2];
  for (int h = 0; h < lhalo; ++h) tmp[h] = f[gix(g, i1, i2, g.lx3 - (lhalo - 1) + h)];
  for (int h = 0; h < lhalo; ++h) f[gix(g, i1, i2, 1 - lhalo + h)] = tmp[h];
  for (int h = 0; h < lhalo; ++h) tmp[h] = f[gix(g, i1, i2, 1 + h)];
  for (int h = 0; h < lhalo; ++h) f[gix(g, i1, i2, g.lx3 + 1 + h)] = tmp[h];
}

// ---------------------------------------------------------------- interface velocities
__global__ void k_vs1i(DevGrid g, DevState s) {
  const int isp = blockIdx.y + 1;
  const long n = (long)(g.lx1 + 1) * g.lx2 * g.lx3;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int k = (int)(t % (g.lx1 + 1)) + 1;
  const long line = t / (g.lx1 + 1);
  const int i2 = (int)(line % g.lx2) + 1, i3 = (int)(line / g.lx2) + 1;
  const double* u = s.vs1 + (long)(isp - 1) * g.N + gix(g, 0, i2, i3);  // u[i1] = vs1(i1)
  const double v = v1i_at(g, u, k);
  s.vs1i[(long)(isp - 1) * n + t] = v;
  const long vb = 2L * (line + (long)g.lx2 * g.lx3 * (isp - 1));
  if (k == 1) s.v1b[vb] = v;
  if (k == g.lx1 + 1) s.v1b[vb + 1] = v;
}

__global__ void k_vs2i(DevGrid g, DevState s) {
  const int isp = blockIdx.y + 1;
  const long n = (long)g.lx1 * (g.lx2 + 1) * g.lx3;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i1 = (int)(t % g.lx1) + 1;
  const long r = t / g.lx1;
  const int k = (int)(r % (g.lx2 + 1)) + 1, i3 = (int)(r / (g.lx2 + 1)) + 1;
  const double* w = s.vs2 + (long)(isp - 1) * g.N;
  s.vs2i[(long)(isp - 1) * n + t] = 0.5 * (w[gix(g, i1, k - 1, i3)] + w[gix(g, i1, k, i3)]);
}

__global__ void k_vs3i(DevGrid g, DevState s) {
  const int isp = blockIdx.y + 1;
  const long n = (long)g.lx1 * g.lx2 * (g.lx3 + 1);
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i1 = (int)(t % g.lx1) + 1;
  const long r = t / g.lx1;
  const int i2 = (int)(r % g.lx2) + 1, k = (int)(r / g.lx2) + 1;
  const double* w = s.vs3 + (long)(isp - 1) * g.N;
  s.vs3i[(long)(isp - 1) * n + t] = 0.5 * (w[gix(g, i1, i2, k - 1)] + w[gix(g, i1, i2, k)]);
}

// only the x1-boundary interface velocities, which the next step's ghost fill reads
__device__ __forceinline__ void save_v1b_body(const DevGrid& g, const DevState& s, int isp, long t) {
  const int i2 = (int)(t % g.lx2) + 1, i3 = (int)(t / g.lx2) + 1;
  const double* u = s.vs1 + (long)(isp - 1) * g.N + gix(g, 0, i2, i3);
  const long vb = 2L * (t + (long)g.lx2 * g.lx3 * (isp - 1));
  s.v1b[vb] = v1i_at(g, u, 1);
  s.v1b[vb + 1] = v1i_at(g, u, g.lx1 + 1);
}
__global__ void k_save_v1b(DevGrid g, DevState s) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < (long)g.lx2 * g.lx3) save_v1b_body(g, s, blockIdx.y + 1, t);
}

// ---------------------------------------------------------------- sweeps
// The fields one sweep launch advances for a range of species: up to three arrays
// (ns, rhovs1, rhoes) sharing the species' velocity and the metric factors, each with
// the tensor rank of its first dimension (advec.f90:313-381: rank 1 squares h1 in the
// x2/x3 metric products).  Species 1..6 carry all three, species 7 (electrons) only
// energy (multifluid.f90:49-126), so the launches for the two groups are separate and
// every wave of a launch runs the same instruction stream.
struct SweepFields {
  double* f[3];
  int frank[3];
  // with all three fields selected, the electron wave (species 7) advances energy only:
  // its three field slots all point at rhoes with rank 0, so the three identical updates
  // write the same values to the same cells
  int elec_energy;
};

// Every memory operation in the steady state of a sweep is unconditional: loads use
// clamped indices (a lane past the end reads a valid cell it does not use) and stores
// either write an updated interior cell or put an unchanged ghost value back.  With no
// branches around memory operations the compiler's wait counters stay exact, so loads
// issued for later steps remain in flight while the current step computes.
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ---------------------------------------------------------------- x1 sweep
// advec.f90:283-309 + :384-435.  Workgroup = one (x2,x3) line, wave w = species w+isp0.
// The wave walks the line in chunks of 64 consecutive cells, lane l at position
// p = 64c + l - 1 holding f(p), f(p+1), the velocity and metric factors at interface /
// cell p, dx(p+1) and dx1i(p-2).  f(p-1), rs(p-1), sl(p-1), phi(p-1) and the other left
// neighbours move one lane up by a DPP wave shift (lane 0 takes lane 63's of the previous
// chunk), and the lane writes the update of cell p-1.  Two register buffers alternate:
// the loads of chunk c+1 are issued before chunk c is computed, and chunk c writes only
// cells chunk c+1 no longer reads.  An odd last chunk runs after the pairs.
__device__ __forceinline__ double lane_up1(double v, double carry) {
  // lane i receives lane i-1's value, lane 0 the carry (DPP wave_shr:1)
  const long long vi = __double_as_longlong(v), ci = __double_as_longlong(carry);
  const int lo = __builtin_amdgcn_update_dpp((int)ci, (int)vi, 0x138, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(ci >> 32), (int)(vi >> 32), 0x138, 0xf, 0xf,
                                             false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double lane63(double v) {
  const long long vi = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)vi, 63);
  const int hi = __builtin_amdgcn_readlane((int)(vi >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int NQ>
struct S1Chunk {
  double f[NQ], fn[NQ];     // f(p), f(p+1)
  double u;                 // vs1(p) (on the fly) or vs1i(p)
  double i1, i2, i3;        // h1x1i, h2x1i, h3x1i at interface p
  double c1, c2, c3;        // h1, h2, h3 at cell p
  double dxp1, dxim;        // dx1(p+1), dx1i(p-2)
};

template <int NQ, bool OTF, int LM = 0>
__device__ __forceinline__ void s1_load(const DevGrid& g, double* const* fp, const double* u,
                                        const double* v1i, long c0, long if0, int p,
                                        S1Chunk<NQ>& d) {
  const int lx1 = g.lx1;
  const int pc = clampi(p, -1, lx1 + 2), pn = clampi(p + 1, -1, lx1 + 2);
  const int pi = clampi(p, 1, lx1 + 1);
  const int pcell = clampi(p, 1, lx1);
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    d.f[q] = fp[q][pc];
    d.fn[q] = fp[q][pn];
  }
  d.u = OTF ? u[pc] : v1i[if0 + pi];
  if (LM) return;                            // (the line's factors are in LDS)
  d.i1 = g.h1x1i[if0 + pi]; d.i2 = g.h2x1i[if0 + pi]; d.i3 = g.h3x1i[if0 + pi];
  d.c1 = g.h1[c0 + pcell]; d.c2 = g.h2[c0 + pcell]; d.c3 = g.h3[c0 + pcell];
  d.dxp1 = g.dx1[clampi(p + 1, 0, lx1 + 2)];
  d.dxim = g.dx1i[clampi(p - 2, 0, lx1 - 1)];
}

template <int NQ>
struct S1Carry {
  double rs[NQ], sl[NQ], ph[NQ], f[NQ];
  double u, h2, ha, dx;
};

template <int NQ, bool OTF, int LM = 0>
__device__ __forceinline__ void s1_chunk(const DevGrid& g, double* const* fp, const double* u,
                                         const double* v1i, long c0, long if0, int c, int lane,
                                         double dt, double vb1, double vbN,
                                         const S1Chunk<NQ>& cur, S1Chunk<NQ>& nxt,
                                         S1Carry<NQ>& k, const double* lm) {
  const int lx1 = g.lx1;
  const int p = 64 * c + lane - 1;
  s1_load<NQ, OTF, LM>(g, fp, u, v1i, c0, if0, p + 64, nxt);              // next chunk
  // h1x1i, h2x1i h3x1i (interface p), (h1 h2) h3 (cell p), dx1(p+1), dx1i(p-2)
  double i1v, h2p, hap, dx_p, dxim;
  double ri1 = 0.0, rha1m = 0.0, rdx = 0.0, rdxs = 0.0, rdxim = 0.0;   // their reciprocals
  const int L = lx1 + 3;
  if (LM) {
    const int pi = clampi(p, 1, lx1 + 1);
    i1v = lm[pi];
    h2p = lm[L + pi];
    hap = lm[2 * L + clampi(p, 1, lx1)];
    dx_p = lm[3 * L + clampi(p + 1, 0, lx1 + 2)];
    dxim = lm[4 * L + clampi(p - 2, 0, lx1 - 1)];
    if (LM == 2) {
      ri1 = lm[5 * L + pi];
      rha1m = lm[6 * L + clampi(p - 1, 1, lx1)];
      rdx = lm[7 * L + clampi(p + 1, 0, lx1 + 2)];
      rdxs = lm[8 * L + clampi(p + 1, 0, lx1 + 2)];
      rdxim = lm[9 * L + clampi(p - 2, 0, lx1 - 1)];
    }
  } else {
    i1v = cur.i1;
    h2p = cur.i2 * cur.i3;
    hap = (cur.c1 * cur.c2) * cur.c3;
    dx_p = cur.dxp1;
    dxim = cur.dxim;
  }
  const bool upd = p >= 2 && p <= lx1 + 1;
  const double um = lane_up1(cur.u, k.u);                                  // vs1(p-1)
  double v;
  if (OTF) v = p == 1 ? vb1 : (p == lx1 + 1 ? vbN : 0.5 * (um + cur.u));
  else v = cur.u;
  const double vq = LM == 2 ? div_rcp(v, i1v, ri1) : v / i1v;
  const double h2m = lane_up1(h2p, k.h2);
  const double ha1m = lane_up1(hap, k.ha);                                 // ha1 of cell p-1
  const double dx_0 = lane_up1(dx_p, k.dx);                                // dx(p)
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const double f0 = cur.f[q], f1 = cur.fn[q];
    const double fm1 = lane_up1(f0, k.f[q]);
    const double rs = LM == 2 ? div_rcp(f1 - f0, dx_p, rdx) : (f1 - f0) / dx_p;  // rs(p)
    const double rsm = lane_up1(rs, k.rs[q]);                              // rs(p-1)
    const double cs = LM == 2 ? div_rcp(f1 - fm1, dx_0 + dx_p, rdxs) : (f1 - fm1) / (dx_0 + dx_p);
    const double sl = minmod(cs, minmod(2.0 * rsm, 2.0 * rs));            // sl(p), p >= 0
    const double slm = lane_up1(sl, k.sl[q]);                              // sl(p-1)
    const double phn = f0 * v - ((0.5 * v) * (dx_0 + vq * dt)) * sl;
    const double phv = fm1 * v + ((0.5 * v) * (dx_0 - vq * dt)) * slm;
    const double ph = v < 0.0 ? phn : phv;                                 // phi(p), p >= 1
    const double phm = lane_up1(ph, k.ph[q]);                              // phi(p-1)
    const double out =
        LM == 2 ? fm1 - div_rcp(div_rcp(dt * (h2p * ph - h2m * phm), dxim, rdxim), ha1m, rha1m)
                : fm1 - ((dt * (h2p * ph - h2m * phm)) / dxim) / ha1m;
    // cells 1..lx1 take the update; the other lanes put the ghost value they cover back
    fp[q][clampi(p - 1, -1, lx1 + 2)] = upd ? out : (p < 0 ? f0 : fm1);
    k.rs[q] = lane63(rs);
    k.sl[q] = lane63(sl);
    k.ph[q] = lane63(ph);
    k.f[q] = lane63(f0);
  }
  k.u = lane63(cur.u);
  k.h2 = lane63(h2p);
  k.ha = lane63(hap);
  k.dx = lane63(dx_p);
}

// line (i2, i3) = (bx + 1, by + 1); under LM the block (k_sweep1) has written the LDS copy
// of the line's factors `lm`, which the walk reads after the barrier below
template <int NQ, bool OTF, int LM = 0>
__device__ __forceinline__ void sweep1_body(const DevGrid& g, const DevState& s,
                                            const SweepFields& F, int isp0, double dt, int bx,
                                            int by, const double* lm = nullptr) {
  const int lx1 = g.lx1;
  const int i2 = bx + 1, i3 = by + 1;
  const int lane = threadIdx.x, isp = isp0 + threadIdx.y;
  const long c0 = gix(g, 0, i2, i3);                                       // cell p -> c0+p
  const long if0 = (long)(lx1 + 1) * ((i2 - 1) + (long)g.lx2 * (i3 - 1)) - 1;  // iface p -> if0+p
  const long so = (long)(isp - 1) * g.N;
  const long nif = (long)(lx1 + 1) * g.lx2 * g.lx3;
  double* fp[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) fp[q] = F.f[q] + so + c0;
  const double* u = s.vs1 + so + c0;                                        // u[p] = vs1(p)
  const double* v1i = s.vs1i + (long)(isp - 1) * nif;
  // boundary interface velocities (advec.f90:171-196), the same for the whole wave
  // (under LM from the planes the step saved from the same vs1 before its x3 sweep: two
  // loads, where v1i_at's branches wait for every load in flight, the factors' included)
  double vb1 = 0.0, vbN = 0.0;
  if (OTF && LM) {
    const long vb = 2L * ((i2 - 1) + (long)g.lx2 * (i3 - 1) + (long)g.lx2 * g.lx3 * (isp - 1));
    vb1 = s.v1b[vb];
    vbN = s.v1b[vb + 1];
  } else if (OTF) {
    vb1 = v1i_at(g, u, 1);
    vbN = v1i_at(g, u, lx1 + 1);
  }
  S1Carry<NQ> k;
#pragma unroll
  for (int q = 0; q < NQ; ++q) k.rs[q] = k.sl[q] = k.ph[q] = k.f[q] = 0.0;
  k.u = k.h2 = 0.0;
  k.ha = k.dx = 1.0;
  const int nchunk = (lx1 + 3 + 63) / 64;                                 // p = -1 .. lx1+1
  S1Chunk<NQ> A, B;
  s1_load<NQ, OTF, LM>(g, fp, u, v1i, c0, if0, lane - 1, A);
  // the first chunk's loads are issued before the barrier after which every wave reads
  // the line's factors, so their latency and the factors' overlap
  if (LM) __syncthreads();
  int c = 0;
  for (; c + 1 < nchunk; c += 2) {
    s1_chunk<NQ, OTF, LM>(g, fp, u, v1i, c0, if0, c, lane, dt, vb1, vbN, A, B, k, lm);
    s1_chunk<NQ, OTF, LM>(g, fp, u, v1i, c0, if0, c + 1, lane, dt, vb1, vbN, B, A, k, lm);
  }
  if (c < nchunk) s1_chunk<NQ, OTF, LM>(g, fp, u, v1i, c0, if0, c, lane, dt, vb1, vbN, A, B, k, lm);
}
// LM 1: the line's metric factors in LDS (dynamic, sweep1_lm_bytes): the block loads
// h1x1i, the products h2x1i h3x1i and (h1 h2) h3, dx1 and dx1i of its line once, in the
// order the walk forms them, instead of every wave loading them for every chunk.  LM 2:
// also the reciprocals of the five divisors the walk divides by (h1x1i, (h1 h2) h3, dx1,
// dx1(p) + dx1(p+1), dx1i), each a correctly rounded 1/b, so that each of the walk's
// thirteen divisions per cell is a multiply and two FMAs (div_rcp) with the same result.
__host__ __device__ inline int sweep1_lm_bytes(int lx1, int lm = 1) {
  return (lm == 2 ? 10 : 5) * (lx1 + 3) * 8;
}
template <int NQ, bool OTF, int LM = 0>
__global__ __launch_bounds__(448) void k_sweep1(DevGrid g, DevState s, SweepFields F, int isp0,
                                                double dt) {
  extern __shared__ double lm[];
  if (LM) {
    const int lx1 = g.lx1, L = lx1 + 3, i2 = blockIdx.x + 1, i3 = blockIdx.y + 1;
    const long c0 = gix(g, 0, i2, i3);
    const long if0 = (long)(lx1 + 1) * ((i2 - 1) + (long)g.lx2 * (i3 - 1)) - 1;
    for (int t = threadIdx.y * blockDim.x + threadIdx.x; t < L; t += blockDim.x * blockDim.y) {
      const int pi = clampi(t, 1, lx1 + 1), pc = clampi(t, 1, lx1);
      const double i1v = g.h1x1i[if0 + pi];
      const double ha = (g.h1[c0 + pc] * g.h2[c0 + pc]) * g.h3[c0 + pc];
      const double dxp = g.dx1[min(t, lx1 + 2)], dxi = g.dx1i[min(t, lx1 - 1)];
      lm[t] = i1v;
      lm[L + t] = g.h2x1i[if0 + pi] * g.h3x1i[if0 + pi];
      lm[2 * L + t] = ha;
      lm[3 * L + t] = dxp;
      lm[4 * L + t] = dxi;
      if (LM == 2) {
        lm[5 * L + t] = 1.0 / i1v;
        lm[6 * L + t] = 1.0 / ha;
        lm[7 * L + t] = 1.0 / dxp;
        lm[8 * L + t] = 1.0 / (g.dx1[min(max(t - 1, 0), lx1 + 2)] + dxp);   // dx(p) + dx(p+1)
        lm[9 * L + t] = 1.0 / dxi;
      }
    }
    // (the barrier is in sweep1_body, after the first chunk's loads)
  }
  if (NQ == 3 && F.elec_energy && isp0 + (int)threadIdx.y == LSP) {
    // the electron wave of an all-species block sweeps its energy only
    SweepFields E{};
    E.f[0] = F.f[2];
    sweep1_body<1, OTF, LM>(g, s, E, isp0, dt, blockIdx.x, blockIdx.y, lm);
  } else {
    sweep1_body<NQ, OTF, LM>(g, s, F, isp0, dt, blockIdx.x, blockIdx.y, lm);
  }
}

// ---------------------------------------------------------------- x2 / x3 sweeps
// advec.f90:313-381 + :384-435.  Thread (i1, line, species): marches along its line
// keeping f(k-1), f(k), f(k+1), the last one-sided slope, limited slope and flux in
// registers for up to three fields that share the velocity and metric factors.
//
// The march is software-pipelined PF steps deep: step k consumes the field values,
// velocity and metric factors loaded PF steps earlier and issues the loads of step
// k+PF.  Field loads run ahead of the in-place stores (step k stores cell k-1 and loads
// cell k+2+PF, which no earlier step has written), so every value read is the original.
// Steps 0 and 1 (no store) are peeled off so every step of the main loop stores.
#ifndef SWEEP_PF
#define SWEEP_PF 2      // x2/x3 march pipeline depth (steps of loads in flight) with 2-3 fields
#endif
#ifndef SWEEP1_LM_MAX
#define SWEEP1_LM_MAX 49152   // x1 sweep: largest LDS copy of a line's metric factors (bytes)
#endif
#ifndef SWEEP1_RCP
#define SWEEP1_RCP 1          // x1 sweep: divisions through the line's reciprocals where they fit
#endif
#ifndef SWEEP1_E7_IN_BLOCK
#define SWEEP1_E7_IN_BLOCK 1  // x1: electrons as a seventh wave of the species block
#endif
#ifndef SWEEP_E7_IN_BLOCK
#define SWEEP_E7_IN_BLOCK 1   // x2/x3: electrons as a seventh wave of the species block
#endif
#ifndef SWEEP23_WAVES
#define SWEEP23_WAVES         // e.g. __attribute__((amdgpu_waves_per_eu(3))) to cap registers
#endif
#ifndef SWEEP_PF1
#define SWEEP_PF1 4     // the same for single-field launches (fewer registers per step)
#endif

// CONV (x3, all fields, the fused step's first sweep): the cells the sweep reads inside the
// line still hold what v12rhov1 and T2rhoe would have converted, so momentum and energy
// are loaded as vs1 and Ts there and converted with the conversions' own expressions
// from the cell's density (the ghost fill has written the ghost cells).  The electron
// wave loads its density in field slot 0 and its momentum in slot 1 without advecting
// them (the electrons' momentum is converted, as v12rhov1 does for every species, but
// not swept), writing the loaded density and the converted momentum back, and advects
// its energy in slot 2.
//
// POST (x2, all fields, the fused step's last sweep): rhov12v1 and rhoe2T
// (multifluid.f90:179-189, 218-233) of every cell the sweep updates, from the values it
// has just computed.  Each ion wave converts its own species (vs1, Ts) and passes its
// density and current term (n qs) vs1 through LDS; after each group of PF steps the
// electron wave sums them in species order for the electron density, velocity and
// temperature of the previous group's cells.  The cells outside the interior box are
// converted by k_post_convert_surface.
// POST: species terms of the cells of two groups of PF steps, [group parity][step][ion][n,
// current][lane], in LDS
template <int PF>
constexpr int sweep_xch_doubles() { return 2 * PF * (LSP - 1) * 2 * 64; }

// block (bx, by): i1 = 64 bx + lane + 1 (lanes past lx1 redo column lx1), line by + 1
// The fused step's sweeps (CONV or POST: blocks of the 7 species waves) share the metric
// factors of each step through LDS: waves 0-5 each load one of hI, hA, hB (interface k)
// and h1, h2, h3 (cell k-1) two groups of steps ahead and store them one group ahead into
// met[group parity][step][factor][lane]; every wave reads them there.
#ifndef SWEEP_SHARE_METRICS
#define SWEEP_SHARE_METRICS 1
#endif

// The fused step's sweeps load each step's three fields and cell velocity by LDS DMA
// (global_load_lds_dwordx4, which holds no registers while in flight) SWEEP_DMA steps
// ahead, into a ring of SWEEP_DMA + 1 slots per wave; a slot is four 64-double rows
// [field 0, field 1, field 2, velocity] and one instruction fills two of them (lanes 0-31
// the 16-byte pairs of the first row, lanes 32-63 of the second).  The compiler does not
// order LDS reads after these loads, so each step waits on the vector memory counter
// itself: that counter retires loads and stores in issue order, and after step k's loads
// (issued at the end of step k - D, or in the prologue) at least sweep_dma_after(...)
// vector memory operations were issued, so waiting until that many remain outstanding
// means step k's rows have landed.  Slot k % (D + 1) is refilled at the end of step
// k + 1, after its rows were read in step k.
#ifndef SWEEP_DMA
#define SWEEP_DMA 6
#endif
template <int D>
constexpr int sweep_ring_doubles() { return LSP * (D + 1) * 4 * 64; }
// metric-factor slots: two groups in registers-then-LDS, or four DMA slots
template <bool DMAV, int PF>
constexpr int sweep_met_slots() { return DMAV && SWEEP_DMA > 0 ? 4 : 2; }
template <int PF, bool DMAV = false>
constexpr int sweep_met_doubles() { return sweep_met_slots<DMAV, PF>() * PF * 6 * 64; }
// vector memory operations issued after the loads of step k: two per step at least; from
// the first step whose D - 1 predecessors all store (k > D), also the step's stores
template <int D>
constexpr int sweep_dma_after(bool steady, int stores) { return (D - 1) * (2 + (steady ? stores : 0)); }
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt holds six bits");
  // gfx9 encoding: vmcnt[3:0], expcnt[6:4] (7: no wait), lgkmcnt[11:8] (15: no wait),
  // vmcnt[5:4] at [15:14]
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
  __asm__ volatile("" ::: "memory");         // no LDS read moves above the wait
}
__device__ __forceinline__ void dma16(const double* src, double* lds) {
  __builtin_amdgcn_global_load_lds((__attribute__((address_space(1))) void*)src,
                                   (__attribute__((address_space(3))) void*)lds, 16, 0, 0);
}
// the four rows of a ring slot at this lane's offset.  Read in assembly (which waits for
// its own reads): an LDS read the compiler sees after a DMA load makes it wait for every
// vector memory operation in flight.
__device__ __forceinline__ void ring_read4(const double* p, double& a, double& b, double& c,
                                           double& d) {
  const unsigned ad = (unsigned)(unsigned long)(__attribute__((address_space(3))) const double*)p;
  __asm__ volatile(
      "ds_read_b64 %0, %4\n\t"
      "ds_read_b64 %1, %4 offset:512\n\t"
      "ds_read_b64 %2, %4 offset:1024\n\t"
      "ds_read_b64 %3, %4 offset:1536\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(a), "=&v"(b), "=&v"(c), "=&v"(d)
      : "v"(ad)
      : "memory");
}
// six rows 1 KiB apart (the metric factors of one step in their DMA slot), likewise
__device__ __forceinline__ void lds_read6(const double* p, double* v) {
  const unsigned ad = (unsigned)(unsigned long)(__attribute__((address_space(3))) const double*)p;
  __asm__ volatile(
      "ds_read_b64 %0, %6\n\t"
      "ds_read_b64 %1, %6 offset:1024\n\t"
      "ds_read_b64 %2, %6 offset:2048\n\t"
      "ds_read_b64 %3, %6 offset:3072\n\t"
      "ds_read_b64 %4, %6 offset:4096\n\t"
      "ds_read_b64 %5, %6 offset:5120\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3]), "=&v"(v[4]), "=&v"(v[5])
      : "v"(ad)
      : "memory");
}
// a read-only grid array (dx, dxi) at a wave-uniform index, by a scalar load: a vector load
// would be the newest vector memory operation and waiting for it would wait for the ring
__device__ __forceinline__ double ld_uniform(const double* p) {
  return *(const __attribute__((address_space(4))) double*)p;
}
// The sweeps' blocks go to the XCDs so that the x1 blocks of a line share one L2:
// workgroups are dispatched to the 8 XCDs round robin, so block b runs on XCD b % 8, and
// XCD x takes lines x, x + 8, ..., each line's x1 blocks one after the other.  The 64-cell
// row segments of adjacent x1 blocks share a 128-byte line at each end (a row starts two
// ghost cells into its line), which x1 blocks on different XCDs each fetched from HBM.
__host__ __device__ inline long sweep_blocks(int nbx, int nlines) {
  return (long)((nlines + 7) / 8) * 8 * nbx;
}
__device__ __forceinline__ bool sweep_block(int nbx, int nlines, int& bx, int& by) {
  const int b = blockIdx.x, t = b >> 3;
  bx = t % nbx;
  by = (t / nbx) * 8 + (b & 7);
  return by < nlines;
}

template <int DIR, int PF, int NQ, bool OTF, bool CONV = false, bool POST = false,
          bool DMAV = false>
__device__ __forceinline__ void sweep23_body(const DevGrid& g, const DevState& s,
                                             const SweepFields& F, int isp0, double dt, int bx,
                                             int by, double* xch, double* met, double* ring) {
  static_assert(!(CONV || POST) || NQ == 3, "the conversions come with all three fields");
  constexpr bool SHM = SWEEP_SHARE_METRICS && (CONV || POST);
  static_assert(!DMAV || (OTF && NQ == 3 && (CONV || POST)), "DMA: the fused sweeps only");
  constexpr bool DMA = DMAV && SWEEP_DMA > 0;
  constexpr int D = DMA ? SWEEP_DMA : 1;
  constexpr bool MDMA = DMA && SHM;           // metric factors by DMA as well
  constexpr int MS = sweep_met_slots<DMAV, PF>();
  static_assert(!MDMA || PF == 2, "one DMA instruction loads a factor's two steps");
  static_assert(!MDMA || 3 * PF >= D, "the factors are loaded at least as far ahead as the rows");
  const int lx = DIR == 2 ? g.lx2 : g.lx3;
  // lanes past lx1 redo column lx1's work: same inputs, same values written to the same cells
  const int i1 = min((int)(bx * 64 + threadIdx.x), g.lx1 - 1) + 1;
  const int isp = isp0 + threadIdx.y;
  const int other = by + 1;                          // i3 for DIR 2, i2 for DIR 3
  const long cs_ = DIR == 2 ? (long)g.L1 : (long)g.L1 * g.L2;   // cell stride along the line
  const long c0 = DIR == 2 ? gix(g, i1, 0, other) : gix(g, i1, other, 0);  // cell k -> c0+k*cs_
  const long istr = DIR == 2 ? (long)g.lx1 : (long)g.lx1 * g.lx2;          // iface stride
  const long if0 = DIR == 2 ? (long)(i1 - 1) + (long)g.lx1 * ((long)(g.lx2 + 1) * (other - 1)) - istr
                            : (long)(i1 - 1) + (long)g.lx1 * (other - 1) - istr;  // iface k -> if0+k*istr
  const double* hA = DIR == 2 ? g.h1x2i : g.h1x3i;   // metric factor that frank squares
  const double* hB = DIR == 2 ? g.h3x2i : g.h2x3i;
  const double* hI = DIR == 2 ? g.h2x2i : g.h3x3i;   // h1i of advec1D_MC_curv
  const double* dx = DIR == 2 ? g.dx2 : g.dx3;
  const double* dxi = DIR == 2 ? g.dx2i : g.dx3i;
  const long so = (long)(isp - 1) * g.N;
  const double* w = (DIR == 2 ? s.vs2 : s.vs3) + so + c0;
  const double* vi = (DIR == 2 ? s.vs2i : s.vs3i) +
                     (long)(isp - 1) * (DIR == 2 ? (long)g.lx1 * (g.lx2 + 1) * g.lx3
                                                 : (long)g.lx1 * g.lx2 * (g.lx3 + 1));
  double* fp[NQ];
  const double* lp[NQ];        // CONV: where a line-interior cell of the field is loaded from
  bool fr1[NQ];
  const bool e7 = F.elec_energy && isp == LSP;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    fp[q] = (e7 ? F.f[NQ - 1] : F.f[q]) + so + c0;
    fr1[q] = !e7 && F.frank[q] == 1;
    lp[q] = fp[q];
  }
  const double msp = ms(isp), gm1 = gammas(isp) - 1.0;
  if (CONV) {
    if (e7) {                                  // density and momentum: converted, not advected
      fp[0] = s.ns + so + c0;
      fp[1] = s.rhovs1 + so + c0;
    }
    lp[0] = fp[0];
    lp[1] = s.vs1 + so + c0;
    lp[2] = s.Ts + so + c0;
  }
  // CONV: turn the loaded slots of a line-interior cell into momentum and energy
  auto convert = [&](double* v) {
    if (CONV) {
      const double n = v[0];
      v[1] = (n * msp) * v[1];
      v[2] = ((n * kB) * v[2]) / gm1;
    }
  };
  // pipeline registers of the steps in flight
  double pf[NQ][PF], pv[PF], pI[PF], pA[PF], pB[PF], p1[PF], p2[PF], p3[PF];
  auto issue = [&](int k, int u) {
    const int kf = min(k + 2, lx + 2);
    const bool conv = CONV && kf <= lx;        // kf >= 2: inside the line up to lx
#pragma unroll
    for (int q = 0; q < NQ; ++q) pf[q][u] = (conv ? lp[q] : fp[q])[(long)kf * cs_];
    const int ki = clampi(k, 1, lx + 1);
    const long ik = if0 + (long)ki * istr;
    pv[u] = OTF ? w[(long)ki * cs_] : vi[ik];
    if (!SHM) {
      pI[u] = hI[ik]; pA[u] = hA[ik]; pB[u] = hB[ik];
      const long cm = c0 + (long)clampi(k - 1, 1, lx) * cs_;
      p1[u] = g.h1[cm]; p2[u] = g.h2[cm]; p3[u] = g.h3[cm];
    }
  };
  // DMA: this lane loads x1 cells pp, pp+1 (0-based in the block's 64) of a row; past the
  // line it loads the block's first pair again (never read).  The row's cell of lane i is
  // at offset i of the slot row; lanes past lx1 read column lx1's (rel).
  const bool hi = threadIdx.x >= 32;
  const int rel = i1 - 1 - bx * 64;
  const int pp0 = bx * 64 + 2 * (threadIdx.x & 31);
  const long dc = (long)(pp0 <= g.lx1 - 1 ? pp0 : bx * 64) - (i1 - 1);
  auto ring_slot = [&](int k) -> double* {
    return ring + (threadIdx.y * (D + 1) + k % (D + 1)) * 256;
  };
  auto dma_issue = [&](int k) {                // the rows of step k
    const int kf = min(k + 2, lx + 2);
    const bool conv = CONV && kf <= lx;
    const int ki = clampi(k, 1, lx + 1);
    const double* a = hi ? (conv ? lp[1 % NQ] : fp[1 % NQ]) : (conv ? lp[0] : fp[0]);
    const double* b = hi ? w + (long)(ki - kf) * cs_ : (conv ? lp[2 % NQ] : fp[2 % NQ]);
    double* sl = ring_slot(k);
    dma16(a + dc + (long)kf * cs_, sl);
    dma16(b + dc + (long)kf * cs_, sl + 128);
  };
  // SHM: this wave's metric factor (wave wy < 6 loads factor wy) at step k, and a group's
  // PF steps into registers or from registers into LDS
  const int wy = threadIdx.y;
  auto met_at = [&](int k) -> double {
    const int ki = clampi(k, 1, lx + 1);
    const long ik = if0 + (long)ki * istr;
    const long cm = c0 + (long)clampi(k - 1, 1, lx) * cs_;
    if (wy == 0) return hI[ik];
    if (wy == 1) return hA[ik];
    if (wy == 2) return hB[ik];
    if (wy == 3) return g.h1[cm];
    if (wy == 4) return g.h2[cm];
    return g.h3[cm];
  };
  auto met_load = [&](int G, double* r) {            // group G: steps 2 + G PF ..
#pragma unroll
    for (int u = 0; u < PF; ++u) r[u] = wy < 6 ? met_at(2 + G * PF + u) : 0.0;
  };
  auto met_store = [&](int G, const double* r) {
    if (wy < 6) {
#pragma unroll
      for (int u = 0; u < PF; ++u) met[(((G & 1) * PF + u) * 6 + wy) * 64 + threadIdx.x] = r[u];
    }
  };
  auto met_get = [&](int G, int u, int m) -> double {
    return met[(((G & 1) * PF + u) * 6 + m) * 64 + threadIdx.x];
  };
  // MDMA: the factors by DMA into a ring of MS group slots [slot][factor][step][64], three
  // groups ahead; wave wy < 6 loads factor wy of a group's two steps with one instruction
  // (lanes 0-31 the first step's row, lanes 32-63 the second's)
  auto met_ptr = [&](int k) -> const double* {
    const int ki = clampi(k, 1, lx + 1);
    const long ik = if0 + (long)ki * istr;
    const long cm = c0 + (long)clampi(k - 1, 1, lx) * cs_;
    if (wy == 0) return hI + ik;
    if (wy == 1) return hA + ik;
    if (wy == 2) return hB + ik;
    if (wy == 3) return g.h1 + cm;
    if (wy == 4) return g.h2 + cm;
    return g.h3 + cm;
  };
  auto met_dma = [&](int G) {
    if (wy < 6) dma16(met_ptr(2 + G * PF + (hi ? 1 : 0)) + dc, met + ((G % MS) * 6 + wy) * PF * 64);
  };
  // before the barrier that ends group G, this wave's load of group G+1's factors has
  // landed: it was followed by the loads and stores of the steps of groups G-2 .. G (or,
  // for G < 2, by the ring's prologue loads and steps 0 and 1)
  auto met_wait = [&](int G) {
    if (wy >= 6) return;
    if (G < 2) wait_vmcnt<2 * D + 4>();
    else if (POST && !e7) wait_vmcnt<3 * PF * (2 + NQ + 2)>();
    else wait_vmcnt<3 * PF * (2 + NQ)>();
  };
  if (MDMA) {
    met_dma(0);
    met_dma(1);
    met_dma(2);
  }
  double fa[NQ], fb[NQ], fc[NQ], rsp[NQ], slp[NQ], php[NQ];
  double dxcur = dx[0];
  double wprev = w[0];         // vs(0) for the on-the-fly interface velocity
  double h2p0 = 0.0, h2p1 = 0.0;   // ha2i(k-1) for frank 0 / frank 1
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    fa[q] = fp[q][-cs_];       // f(-1)
    fb[q] = fp[q][0];          // f(0)
    fc[q] = lp[q][cs_];        // f(1)
  }
  if (lx >= 1) convert(fc);
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    rsp[q] = (fb[q] - fa[q]) / dxcur;
    slp[q] = php[q] = 0.0;
  }
  if (DMA) {
#pragma unroll
    for (int k = 0; k < D; ++k) dma_issue(k);
  } else {
#pragma unroll
    for (int u = 0; u < PF; ++u) issue(u, u);
  }
  // POST: per-wave constants, the group parity being written, and the electron wave's
  // updated energies of the current and the previous group
  const double gm1kB = (gammas(isp) - 1.0) / kB, qsp = qs(isp);
  int xbuf = 0;
  double eout[PF], eprev[PF];
  const int lane = threadIdx.x;
  auto xslot = [&](int b, int slot, int sp, int j) -> double& {
    return xch[(((b * PF + slot) * (LSP - 1) + sp) * 2 + j) * 64 + lane];
  };
  // electron density, velocity and temperature of the cells k0-1 .. k0+cnt-2 of a group
  // the ion waves have written into LDS parity b (electron energies in e)
  auto finalize = [&](int b, int k0, int cnt, const double* e) {
#pragma unroll
    for (int slot = 0; slot < PF; ++slot) {
      if (slot >= cnt) break;
      double cf = 0.0, sum = 0.0;
#pragma unroll
      for (int sp = 0; sp < LSP - 1; ++sp) cf = cf + xslot(b, slot, sp, 1);
#pragma unroll
      for (int sp = 0; sp < LSP - 1; ++sp) sum = sum + xslot(b, slot, sp, 0);
      const long c = c0 + (long)(k0 + slot - 1) * cs_ + so;
      s.ns[c] = sum;
      s.vs1[c] = -(((1.0 / fmax_f(sum, mindensdiv)) / qs(LSP)) * cf);
      s.Ts[c] = fmax_f((gm1kB * e[slot]) / fmax_f(sum, mindensdiv), 100.0);
    }
  };
  // mode 0: step 0 (slopes only), 1: step 1 (flux, no store), 2: steps 2..lx+1
  // slot: the step's position in its group; G: the group (SHM reads its metric factors)
  auto step = [&](int k, int u, int mode, int slot = 0, int G = 0) {
    const double dxk = dxcur, dxk1 = ld_uniform(dx + k + 1);
    dxcur = dxk1;
    double wv, nxt[NQ];
    if (DMA) {                                 // step k's rows have landed in its slot
      if (k <= D) wait_vmcnt<sweep_dma_after<D>(false, 0)>();
      else if (POST && !e7) wait_vmcnt<sweep_dma_after<D>(true, NQ + 2)>();
      else wait_vmcnt<sweep_dma_after<D>(true, NQ)>();
      ring_read4(ring_slot(k) + rel, nxt[0], nxt[1 % NQ], nxt[2 % NQ], wv);
    } else {
      wv = pv[u];
#pragma unroll
      for (int q = 0; q < NQ; ++q) nxt[q] = pf[q][u];
    }
    double hIv, a, b, h1, h2, h3;
    if (!SHM) {
      hIv = pI[u]; a = pA[u]; b = pB[u]; h1 = p1[u]; h2 = p2[u]; h3 = p3[u];
    } else if (mode == 2 && MDMA) {
      double m6[6];
      lds_read6(met + ((G % MS) * 6 * PF + slot) * 64 + rel, m6);
      hIv = m6[0]; a = m6[1]; b = m6[2]; h1 = m6[3]; h2 = m6[4]; h3 = m6[5];
    } else if (mode == 2) {
      hIv = met_get(G, slot, 0); a = met_get(G, slot, 1); b = met_get(G, slot, 2);
      h1 = met_get(G, slot, 3); h2 = met_get(G, slot, 4); h3 = met_get(G, slot, 5);
    } else {                                   // step 1 (step 0 uses none)
      const long ik = if0 + (long)clampi(k, 1, lx + 1) * istr;
      hIv = hI[ik]; a = hA[ik]; b = hB[ik];
      h1 = h2 = h3 = 1.0;
    }
    if (!DMA) issue(k + PF, u);                // (after this step's slot u was read)
    if (CONV && min(k + 2, lx + 2) <= lx) convert(nxt);
    double v = 0.0, vq = 0.0, h2f0 = 0.0, h2f1 = 0.0, ha1f0 = 1.0, ha1f1 = 1.0, dxim = 1.0;
    double newv[NQ];
    if (mode >= 1) {
      if (OTF) { v = 0.5 * (wprev + wv); wprev = wv; }
      else v = wv;
      vq = v / hIv;
      h2f0 = a * b;
      h2f1 = (a * a) * b;
    }
    if (mode == 2) {
      ha1f0 = (h1 * h2) * h3;
      ha1f1 = ((h1 * h1) * h2) * h3;
      dxim = ld_uniform(dxi + k - 2);
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const double rs = (fc[q] - fb[q]) / dxk1;
      const double csl = (fc[q] - fa[q]) / (dxk + dxk1);
      const double sl = minmod(csl, minmod(2.0 * rsp[q], 2.0 * rs));
      if (mode >= 1) {
        const double phn = fb[q] * v - ((0.5 * v) * (dxk + vq * dt)) * sl;
        const double phv = fa[q] * v + ((0.5 * v) * (dxk - vq * dt)) * slp[q];
        const double ph = v < 0.0 ? phn : phv;
        if (mode == 2) {
          const double h2k = fr1[q] ? h2f1 : h2f0;
          const double h2m = fr1[q] ? h2p1 : h2p0;
          const double ha1 = fr1[q] ? ha1f1 : ha1f0;
          const double out = fa[q] - ((dt * (h2k * ph - h2m * php[q])) / dxim) / ha1;
          // CONV: the electron wave's density and momentum go back unchanged
          fp[q][(long)(k - 1) * cs_] = (CONV && q < 2 && e7) ? fa[q] : out;
          if (POST) newv[q] = out;
        }
        php[q] = ph;
      }
      slp[q] = sl;
      rsp[q] = rs;
      fa[q] = fb[q];
      fb[q] = fc[q];
      fc[q] = nxt[q];
    }
    if (mode >= 1) { h2p0 = h2f0; h2p1 = h2f1; }
    if (POST && mode == 2) {
      if (e7) {
        eout[slot] = newv[NQ - 1];
      } else {                                  // ion: density, momentum, energy of cell k-1
        const double n = newv[0];
        const double vv = newv[1] / (msp * fmax_f(n, mindensdiv));
        const long c = c0 + (long)(k - 1) * cs_ + so;
        s.vs1[c] = vv;
        s.Ts[c] = fmax_f((gm1kB * newv[2]) / fmax_f(n, mindensdiv), 100.0);
        xslot(xbuf, slot, isp - 1, 0) = n;
        xslot(xbuf, slot, isp - 1, 1) = (n * qsp) * vv;
      }
    }
    if (DMA) dma_issue(k + D);                // into the slot step k - 1 read
  };
  // POST: LDS stores of one group are visible to the electron wave after this barrier; no
  // fence on global memory, so loads issued ahead stay in flight
  auto lds_barrier = [&]() {
    if (DMA) {
      // the fences would also wait for the ring's loads: this wave's LDS accesses are
      // complete (lgkmcnt) before the barrier, and nothing moves across it
      __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    } else {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
      __builtin_amdgcn_s_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    }
  };
  double rmA[PF], rmB[PF];                    // SHM: factors of two groups ahead
  if (MDMA) {
    // (the ring's prologue loads are issued after these: see met_wait)
  } else if (SHM) {
    met_load(0, rmA);
    met_store(0, rmA);
    met_load(1, rmB);
    lds_barrier();
  }
  step(0, 0, 0);
  step(1, 1 % PF, 1);
  if (MDMA) {                                 // group 0's factors published
    wait_vmcnt<2 * D + 4>();
    lds_barrier();
  }
  int kb = 2;
  int groups = 0;
  // group G (steps kb..kb+PF-1): SHM loads the factors of group G+2 into rl and stores
  // those of group G+1 from rw before the barrier (MDMA: loads those of group G+3 into
  // their slot and waits for its load of group G+1's before the barrier)
  auto group = [&](double* rl, double* rw) {
    if (MDMA) met_dma(groups + 3);
    else if (SHM) met_load(groups + 2, rl);
#pragma unroll
    for (int u = 0; u < PF; ++u) step(kb + u, (2 + u) % PF, 2, u, groups);
    if (MDMA) met_wait(groups);
    else if (SHM) met_store(groups + 1, rw);
    if (POST) {
      // the electron wave finishes the previous group (its LDS parity is not written in
      // this group), then the ion waves may write that parity again after the barrier
      if (e7 && groups > 0) finalize(xbuf ^ 1, kb - PF, PF, eprev);
#pragma unroll
      for (int u = 0; u < PF; ++u) eprev[u] = eout[u];
      lds_barrier();
      xbuf ^= 1;
    } else if (SHM) {
      lds_barrier();
    } else {
      // keep the species waves of the block in step so the metric factors they all read
      // are still cached when the slower waves reach them (no memory fence: nothing is
      // exchanged between the waves)
      __builtin_amdgcn_s_barrier();
    }
    ++groups;
    kb += PF;
  };
  while (kb + PF - 1 <= lx + 1) {              // two groups a turn: rmA and rmB alternate
    group(rmA, rmB);
    if (kb + PF - 1 > lx + 1) break;
    group(rmB, rmA);
  }
  int tail = 0;
#pragma unroll
  for (int u = 0; u < PF - 1; ++u) {
    if (kb + u > lx + 1) break;
    step(kb + u, (2 + u) % PF, 2, u, groups);
    ++tail;
  }
  if (POST) {
    if (e7 && groups > 0) finalize(xbuf ^ 1, kb - PF, PF, eprev);
    lds_barrier();
    if (e7 && tail > 0) finalize(xbuf, kb, tail, eout);
  }
  if (DMA) wait_vmcnt<0>();                   // no load still writing the ring on exit
}

// (a 1-D grid of sweep_blocks(nbx, lines) blocks, see sweep_block)
template <int DIR, int PF, int NQ, bool OTF, bool CONV = false, bool POST = false,
          bool DMAV = false>
__global__ __launch_bounds__(448) SWEEP23_WAVES void k_sweep23(DevGrid g, DevState s,
                                                               SweepFields F, int isp0, double dt) {
  __shared__ double xch[POST ? sweep_xch_doubles<PF>() : 1];
  __shared__ double met[(CONV || POST) ? sweep_met_doubles<PF, DMAV>() : 1];
  __shared__ double ring[DMAV && SWEEP_DMA > 0 ? sweep_ring_doubles<SWEEP_DMA>() : 1];
  int bx, by;
  if (!sweep_block((g.lx1 + 63) / 64, DIR == 2 ? g.lx3 : g.lx2, bx, by)) return;
  sweep23_body<DIR, PF, NQ, OTF, CONV, POST, DMAV>(g, s, F, isp0, dt, bx, by, xch, met, ring);
}

// Ts ghost cells: floor at 100 K (the reference leaves them undefined, see the
// diffusion kernel).  One thread per ghost cell.
__device__ __forceinline__ void floor_ghost_body(const DevGrid& g, double* Ts, int isp, long t) {
  double* T = Ts + (long)(isp - 1) * g.N;
  const long c = surface_cell(g, t);
  if (c >= 0) T[c] = fmax_f(T[c], 100.0);
}
__global__ void k_floor_ghosts(DevGrid g, double* Ts) {
  floor_ghost_body(g, Ts, blockIdx.y + 1, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------- batched ensemble kernels
// Persistent grids over the work units of every member (BatchTable): unit t of kernel k
// belongs to the member m with pre[m] <= t < pre[m + 1] and is that member's unit
// t - pre[m], decoded as the member's own launch numbers it.

// Batched kernels copy the member's grid and state by value, as a member's own launch
// receives them: hipcc (ROCm 7.2) has been seen to drop the last arm of the x2/x3 fills'
// four-way choice of field (q == 0 ? .. : q == 1 ? .. : q == 2 ? .. : rhoes) when the
// struct is read through a reference to memory, so that q = 3 wrote density.

// the thread kernels: ghost fills, boundary interface velocities, surface conversions.
// Every member's unit count is padded to whole waves, so the member is the same for all
// lanes of a wave and is read into scalar registers (readfirstlane): the bodies then see
// uniform grid and state, as in the members' own launches.  The x2 and x3 fills take the
// field from blockIdx.y, also uniform, as their own launches do.
template <int K>
__global__ __launch_bounds__(TPB) void kb_threads(const BatchMember* M, const long* pre, int n,
                                                  long total) {
  const long* p = pre + (long)K * (n + 1);
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x) {
    const int m = __builtin_amdgcn_readfirstlane(batch_member(p, n, t));
    const DevGrid g = M[m].g;                        // by value, as a member's own launch
    const DevState s = M[m].s;                       // receives them
    const long u = t - p[m];
    if (u >= batch_thread_units(g, K)) continue;     // padding of the member's last wave
    if (K == BK_BC1) {
      const long cols = (long)g.L2 * g.L3;
      bc_x1_body<true>(g, s, (int)(u / cols) + 1, u % cols);
    } else if (K == BK_BC2 || K == BK_BC3) {
      const long cols = K == BK_BC2 ? (long)g.L1 * g.L3 : (long)g.L1 * g.L2;
      const int q = blockIdx.y;
      if (K == BK_BC2) bc_x2_body<true>(g, s, (int)(u / cols) + 1, q, u % cols);
      else bc_x3_body<true>(g, s, (int)(u / cols) + 1, q, u % cols);
    } else if (K == BK_V1B) {
      const long cols = (long)g.lx2 * g.lx3;
      save_v1b_body(g, s, (int)(u / cols) + 1, u % cols);
    } else if (K == BK_POST) {
      const long c = surface_cell(g, u);
      if (c >= 0) post_convert_cell(g, s, c);
    } else if (K == BK_FLOOR) {
      const long cells = surface_cells(g);
      floor_ghost_body(g, s.Ts, (int)(u / cells) + 1, u % cells);
    }
  }
}

// the x2/x3 sweeps of all fields (x3 with the conversions folded in, x2 with rhov12v1 and
// rhoe2T); blocks of 64 x 7
template <int DIR, int PF>
__global__ __launch_bounds__(448) void kb_sweep23(const BatchMember* M, const long* pre, int n,
                                                  long total) {
  __shared__ double xch[sweep_xch_doubles<PF>()];
  __shared__ double met[sweep_met_doubles<PF>()];
  __shared__ double ring[1];                       // (register prefetch)
  constexpr int K = DIR == 3 ? BK_SW3 : BK_SW2;
  const long* p = pre + (long)K * (n + 1);
  for (long t = blockIdx.x; t < total; t += gridDim.x) {
    const int m = batch_member(p, n, t);
    const DevGrid g = M[m].g;                        // by value, as a member's own launch
    const DevState s = M[m].s;                       // receives them
    // the member's units in the XCD-aware order of its own launch (sweep_block): the
    // member's first unit and the grid are multiples of 8, so unit u runs on XCD u % 8
    const long u = t - p[m];
    const int nbx = (g.lx1 + 63) / 64, nl = DIR == 2 ? g.lx3 : g.lx2;
    const int bx = (int)((u >> 3) % nbx), by = (int)((u >> 3) / nbx) * 8 + (int)(u & 7);
    if (by >= nl) continue;                        // (the whole block)
    SweepFields F{};
    F.f[0] = s.ns; F.frank[0] = 0;
    F.f[1] = s.rhovs1; F.frank[1] = 1;
    F.f[2] = s.rhoes; F.frank[2] = 0;
    F.elec_energy = 1;
    sweep23_body<DIR, PF, 3, true, DIR == 3, DIR == 2>(g, s, F, 1, M[m].dt, bx, by, xch, met,
                                                      ring);
    __syncthreads();                               // the next unit reuses the LDS
  }
}

// the x1 sweep: blocks of 64 x 7, the electron wave sweeping its energy only (the one-field
// body), as a member's own launch
__global__ __launch_bounds__(448) void kb_sweep1(const BatchMember* M, const long* pre, int n,
                                                 long total) {
  const long* p = pre + (long)BK_SW1 * (n + 1);
  for (long t = blockIdx.x; t < total; t += gridDim.x) {
    const int m = batch_member(p, n, t);
    const DevGrid g = M[m].g;                        // by value, as a member's own launch
    const DevState s = M[m].s;                       // receives them
    const long u = t - p[m];
    SweepFields F{};
    if (threadIdx.y == LSP - 1) {
      F.f[0] = s.rhoes; F.frank[0] = 0;
      sweep1_body<1, true>(g, s, F, 1, M[m].dt, (int)(u % g.lx2), (int)(u / g.lx2));
    } else {
      F.f[0] = s.ns; F.f[1] = s.rhovs1; F.f[2] = s.rhoes;   // x1: every field of rank 0
      sweep1_body<3, true>(g, s, F, 1, M[m].dt, (int)(u % g.lx2), (int)(u / g.lx2));
    }
  }
}

}  // namespace

// a member's units of a kernel: threads (padded to whole waves; the x2 and x3 fills per
// field, the field being blockIdx.y) or blocks
long batch_units(const DevGrid& g, int kernel) {
  switch (kernel) {
    case BK_BC1: case BK_BC2: case BK_BC3: case BK_V1B: case BK_POST: case BK_FLOOR:
      return (batch_thread_units(g, kernel) + 63) / 64 * 64;
    case BK_SW3: return sweep_blocks((g.lx1 + 63) / 64, g.lx2);
    case BK_SW1: return (long)g.lx2 * g.lx3;
    case BK_SW2: return sweep_blocks((g.lx1 + 63) / 64, g.lx3);
    case BK_DIFF: return (((long)g.lx2 * g.lx3 + 63) / 64 + 7) / 8 * 8 * LSP;
  }
  return 0;
}

// The advection part of a batched step for every member, in the fused step's order:
// ghost fill (with v12rhov1, T2rhoe), x1 boundary interface velocities, x3 sweep
// (converting), x1 sweep, x2 sweep (with rhov12v1, rhoe2T), the surface conversions.
void launch_batch_advection(const BatchTable& t, int sweep_pf, hipStream_t st) {
  const dim3 thr(TPB), s7(64, LSP);
  auto tb = [&](long total) { return (unsigned)std::max(1L, std::min<long>(total, t.blocks_sweep)); };
  auto th = [&](long total) {
    return (unsigned)std::max(1L, std::min<long>((total + TPB - 1) / TPB, t.blocks_thread));
  };
  hipLaunchKernelGGL(kb_threads<BK_BC1>, dim3(th(t.total[BK_BC1])), thr, 0, st, t.m, t.pre, t.n,
                     t.total[BK_BC1]);
  hipLaunchKernelGGL(kb_threads<BK_BC2>, dim3(th(t.total[BK_BC2]), 4), thr, 0, st, t.m, t.pre,
                     t.n, t.total[BK_BC2]);
  hipLaunchKernelGGL(kb_threads<BK_BC3>, dim3(th(t.total[BK_BC3]), 4), thr, 0, st, t.m, t.pre,
                     t.n, t.total[BK_BC3]);
  hipLaunchKernelGGL(kb_threads<BK_V1B>, dim3(th(t.total[BK_V1B])), thr, 0, st, t.m, t.pre, t.n,
                     t.total[BK_V1B]);
  if (sweep_pf == 1)
    hipLaunchKernelGGL((kb_sweep23<3, 1>), dim3(tb(t.total[BK_SW3])), s7, 0, st, t.m, t.pre, t.n,
                       t.total[BK_SW3]);
  else
    hipLaunchKernelGGL((kb_sweep23<3, 2>), dim3(tb(t.total[BK_SW3])), s7, 0, st, t.m, t.pre, t.n,
                       t.total[BK_SW3]);
  hipLaunchKernelGGL(kb_sweep1, dim3(tb(t.total[BK_SW1])), s7, 0, st, t.m, t.pre, t.n,
                     t.total[BK_SW1]);
  if (sweep_pf == 1)
    hipLaunchKernelGGL((kb_sweep23<2, 1>), dim3(tb(t.total[BK_SW2])), s7, 0, st, t.m, t.pre, t.n,
                       t.total[BK_SW2]);
  else
    hipLaunchKernelGGL((kb_sweep23<2, 2>), dim3(tb(t.total[BK_SW2])), s7, 0, st, t.m, t.pre, t.n,
                       t.total[BK_SW2]);
  hipLaunchKernelGGL(kb_threads<BK_POST>, dim3(th(t.total[BK_POST])), thr, 0, st, t.m, t.pre, t.n,
                     t.total[BK_POST]);
}

void launch_batch_floor(const BatchTable& t, hipStream_t st) {
  const long total = t.total[BK_FLOOR];
  const unsigned b = (unsigned)std::max(1L, std::min<long>((total + TPB - 1) / TPB, t.blocks_thread));
  hipLaunchKernelGGL(kb_threads<BK_FLOOR>, dim3(b), dim3(TPB), 0, st, t.m, t.pre, t.n, total);
}

// ---------------------------------------------------------------- launchers
void launch_v12rhov1_T2rhoe(const DevGrid& g, const DevState& s, bool mom, bool energy,
                            hipStream_t st) {
  timing_begin("pre_convert", st);
  hipLaunchKernelGGL(k_pre_convert, dim3(blocks_for(g.N), LSP), dim3(TPB), 0, st, g, s,
                     (int)mom, (int)energy);
  timing_end("pre_convert", st);
}

void launch_rhov12v1(const DevGrid& g, const DevState& s, hipStream_t st) {
  timing_begin("rhov12v1", st);
  hipLaunchKernelGGL(k_rhov12v1, dim3(blocks_for(g.N)), dim3(TPB), 0, st, g, s);
  timing_end("rhov12v1", st);
}

void launch_post_convert(const DevGrid& g, const DevState& s, hipStream_t st) {
  timing_begin("post_convert", st);
  hipLaunchKernelGGL(k_post_convert, dim3(blocks_for(g.N)), dim3(TPB), 0, st, g, s);
  timing_end("post_convert", st);
}

void launch_rhoe2T(const DevGrid& g, const DevState& s, hipStream_t st) {
  timing_begin("rhoe2T", st);
  hipLaunchKernelGGL(k_rhoe2T, dim3(blocks_for(g.N), LSP), dim3(TPB), 0, st, g, s);
  timing_end("rhoe2T", st);
}

void launch_set_global_boundaries(const DevGrid& g, const DevState& s, int lsp,
                                  hipStream_t st) {
  timing_begin("ghost_fill", st);
  hipLaunchKernelGGL(k_bc_x1<false>, dim3(blocks_for((long)g.L2 * g.L3), lsp), dim3(TPB), 0, st,
                     g, s);
  hipLaunchKernelGGL(k_bc_x2<false>, dim3(blocks_for((long)g.L1 * g.L3), lsp, 4), dim3(TPB), 0,
                     st, g, s);
  if (!g.flagper)
    hipLaunchKernelGGL(k_bc_x3<false>, dim3(blocks_for((long)g.L1 * g.L2), lsp, 4), dim3(TPB), 0,
                       st, g, s);
  timing_end("ghost_fill", st);
}

bool preconvert_fusable(const DevGrid& g) {
  if (g.flagper || g.lx3 <= 1) return false;   // the x3 sweep converts the interior cells
  for (int k = 0; k < 6; ++k)
    if (!g.isglobal[k]) return false;
  return true;
}

void launch_set_global_boundaries_preconvert(const DevGrid& g, const DevState& s,
                                             hipStream_t st) {
  timing_begin("ghost_fill", st);
  hipLaunchKernelGGL(k_bc_x1<true>, dim3(blocks_for((long)g.L2 * g.L3), LSP), dim3(TPB), 0, st,
                     g, s);
  hipLaunchKernelGGL(k_bc_x2<true>, dim3(blocks_for((long)g.L1 * g.L3), LSP, 4), dim3(TPB), 0,
                     st, g, s);
  hipLaunchKernelGGL(k_bc_x3<true>, dim3(blocks_for((long)g.L1 * g.L2), LSP, 4), dim3(TPB), 0, st,
                     g, s);
  timing_end("ghost_fill", st);
}

void launch_halo_periodic(const DevGrid& g, double* const* fields, int nfields, int lhalo,
                          int nspec, hipStream_t st) {
  if (!g.flagper || nfields <= 0) return;
  FieldList fl;
  for (int i = 0; i < 8; ++i) fl.p[i] = i < nfields ? fields[i] : nullptr;
  timing_begin("halo_periodic", st);
  hipLaunchKernelGGL(k_halo_periodic, dim3(blocks_for((long)g.L1 * g.lx2), nspec, nfields),
                     dim3(TPB), 0, st, g, fl, lhalo);
  timing_end("halo_periodic", st);
}

void launch_interface_vels(const DevGrid& g, const DevState& s, int lsp, hipStream_t st) {
  timing_begin("interface_vels", st);
  hipLaunchKernelGGL(k_vs1i, dim3(blocks_for((long)(g.lx1 + 1) * g.lx2 * g.lx3), lsp), dim3(TPB),
                     0, st, g, s);
  hipLaunchKernelGGL(k_vs2i, dim3(blocks_for((long)g.lx1 * (g.lx2 + 1) * g.lx3), lsp), dim3(TPB),
                     0, st, g, s);
  hipLaunchKernelGGL(k_vs3i, dim3(blocks_for((long)g.lx1 * g.lx2 * (g.lx3 + 1)), lsp), dim3(TPB),
                     0, st, g, s);
  timing_end("interface_vels", st);
}

void launch_save_v1_boundaries(const DevGrid& g, const DevState& s, hipStream_t st) {
  timing_begin("save_v1b", st);
  hipLaunchKernelGGL(k_save_v1b, dim3(blocks_for((long)g.lx2 * g.lx3), LSP), dim3(TPB), 0, st, g,
                     s);
  timing_end("save_v1b", st);
}

// The DMA ring loads 16-byte pairs of cells: every pair must be 16-byte aligned, so rows of
// an even length (lx1 + 4) and fields at 16-byte aligned addresses; otherwise the sweeps
// prefetch through registers.
static bool sweep_dma_ok(const DevGrid& g, std::initializer_list<const double*> fields) {
  if (SWEEP_DMA <= 0 || g.L1 % 2 != 0) return false;
  for (const double* f : fields)
    if (reinterpret_cast<uintptr_t>(f) % 16 != 0) return false;
  return true;
}

namespace {
template <int NQ, int PFM = SWEEP_PF>
void sweep_launch(const DevGrid& g, const DevState& s, int dir, const SweepFields& F, int isp0,
                  int nsp, double dt, bool otf, hipStream_t st) {
  const dim3 blk(64, nsp);
  const dim3 grd = dir == 1 ? dim3(g.lx2, g.lx3)
                            : dim3(sweep_blocks((g.lx1 + 63) / 64, dir == 2 ? g.lx3 : g.lx2));
  if (dir == 1) {
    if (otf) hipLaunchKernelGGL((k_sweep1<NQ, true>), grd, blk, 0, st, g, s, F, isp0, dt);
    else hipLaunchKernelGGL((k_sweep1<NQ, false>), grd, blk, 0, st, g, s, F, isp0, dt);
  } else if (dir == 2) {
    constexpr int PF = NQ == 1 ? SWEEP_PF1 : PFM;
    if (otf) hipLaunchKernelGGL((k_sweep23<2, PF, NQ, true>), grd, blk, 0, st, g, s, F, isp0, dt);
    else hipLaunchKernelGGL((k_sweep23<2, PF, NQ, false>), grd, blk, 0, st, g, s, F, isp0, dt);
  } else {
    constexpr int PF = NQ == 1 ? SWEEP_PF1 : PFM;
    if (otf) hipLaunchKernelGGL((k_sweep23<3, PF, NQ, true>), grd, blk, 0, st, g, s, F, isp0, dt);
    else hipLaunchKernelGGL((k_sweep23<3, PF, NQ, false>), grd, blk, 0, st, g, s, F, isp0, dt);
  }
}
}  // namespace

void launch_sweep(const DevGrid& g, const DevState& s, int dir, int fields, double dt,
                  bool onthefly, hipStream_t st, const Tuning& tune) {
  const int lx = dir == 1 ? g.lx1 : dir == 2 ? g.lx2 : g.lx3;
  if (lx <= 1) return;                       // advec1D_MC_curv leaves lx = 1 lines alone
  SweepFields A{};
  int na = 0;
  if (fields & F_MASS) { A.f[na] = s.ns; A.frank[na++] = 0; }
  if (fields & F_MOMENTUM) { A.f[na] = s.rhovs1; A.frank[na++] = dir == 1 ? 0 : 1; }
  if (fields & F_ENERGY) { A.f[na] = s.rhoes; A.frank[na++] = 0; }
  if (na == 0) return;
  const char* name = dir == 1 ? "sweep1" : dir == 2 ? "sweep2" : "sweep3";
  timing_begin(name, st);
  if (na == 1 && fields == F_ENERGY) {       // energy alone: all seven species in one launch
    sweep_launch<1>(g, s, dir, A, 1, LSP, dt, onthefly, st);
  } else if (na == 3 && dir == 1 && SWEEP1_E7_IN_BLOCK) {
    // x1, all fields: the electron wave joins the block and sweeps its energy only
    A.elec_energy = 1;
    if (onthefly && SWEEP1_RCP && sweep1_lm_bytes(g.lx1, 2) <= SWEEP1_LM_MAX) {
      hipLaunchKernelGGL((k_sweep1<3, true, 2>), dim3(g.lx2, g.lx3), dim3(64, LSP),
                         sweep1_lm_bytes(g.lx1, 2), st, g, s, A, 1, dt);
    } else if (onthefly && sweep1_lm_bytes(g.lx1) <= SWEEP1_LM_MAX) {
      hipLaunchKernelGGL((k_sweep1<3, true, 1>), dim3(g.lx2, g.lx3), dim3(64, LSP),
                         sweep1_lm_bytes(g.lx1), st, g, s, A, 1, dt);
    } else {
      sweep_launch<3>(g, s, dir, A, 1, LSP, dt, onthefly, st);
    }
  } else if (na == 3 && dir != 1 && SWEEP_E7_IN_BLOCK) {
    // x2/x3, all fields: the electron wave joins the block and does energy only
    A.elec_energy = 1;
    if (tune.sweep_pf == 1) sweep_launch<3, 1>(g, s, dir, A, 1, LSP, dt, onthefly, st);
    else sweep_launch<3, 2>(g, s, dir, A, 1, LSP, dt, onthefly, st);
  } else {                                   // fewer fields: ions, then the electrons
    if (na == 1) sweep_launch<1>(g, s, dir, A, 1, LSP - 1, dt, onthefly, st);
    else if (na == 2) sweep_launch<2>(g, s, dir, A, 1, LSP - 1, dt, onthefly, st);
    else sweep_launch<3>(g, s, dir, A, 1, LSP - 1, dt, onthefly, st);
    if (fields & F_ENERGY) {                 // electrons: energy only
      SweepFields E{};
      E.f[0] = s.rhoes;
      E.frank[0] = 0;
      sweep_launch<1>(g, s, dir, E, LSP, 1, dt, onthefly, st);
    }
  }
  timing_end(name, st);
}

// The fused step's x3 sweep of all fields, with v12rhov1 and T2rhoe folded into its loads
// (after launch_set_global_boundaries_preconvert; interface velocities on the fly).
void launch_sweep3_preconvert(const DevGrid& g, const DevState& s, double dt, hipStream_t st,
                              const Tuning& tune) {
  SweepFields A{};
  A.f[0] = s.ns; A.frank[0] = 0;
  A.f[1] = s.rhovs1; A.frank[1] = 1;
  A.f[2] = s.rhoes; A.frank[2] = 0;
  A.elec_energy = 1;
  const dim3 blk(64, LSP), grd(sweep_blocks((g.lx1 + 63) / 64, g.lx2));
  const bool dma = sweep_dma_ok(g, {s.ns, s.rhovs1, s.rhoes, s.vs1, s.Ts, s.vs3, g.h1, g.h2,
                                    g.h3, g.h1x3i, g.h2x3i, g.h3x3i});
  timing_begin("sweep3", st);
  if (tune.sweep_pf == 1) {
    hipLaunchKernelGGL((k_sweep23<3, 1, 3, true, true>), grd, blk, 0, st, g, s, A, 1, dt);
  } else {
    if (dma) hipLaunchKernelGGL((k_sweep23<3, 2, 3, true, true, false, true>), grd, blk, 0, st, g, s, A, 1, dt);
    else hipLaunchKernelGGL((k_sweep23<3, 2, 3, true, true>), grd, blk, 0, st, g, s, A, 1, dt);
  }
  timing_end("sweep3", st);
}

// The fused step's x2 sweep of all fields with rhov12v1 and rhoe2T of every cell it
// updates, then the conversions of the cells outside the interior box.
void launch_sweep2_postconvert(const DevGrid& g, const DevState& s, double dt, hipStream_t st,
                               const Tuning& tune) {
  SweepFields A{};
  A.f[0] = s.ns; A.frank[0] = 0;
  A.f[1] = s.rhovs1; A.frank[1] = 1;
  A.f[2] = s.rhoes; A.frank[2] = 0;
  A.elec_energy = 1;
  const dim3 blk(64, LSP), grd(sweep_blocks((g.lx1 + 63) / 64, g.lx3));
  const bool dma = sweep_dma_ok(g, {s.ns, s.rhovs1, s.rhoes, s.vs1, s.Ts, s.vs2, g.h1, g.h2,
                                    g.h3, g.h1x2i, g.h2x2i, g.h3x2i});
  timing_begin("sweep2", st);
  if (tune.sweep_pf == 1) {
    hipLaunchKernelGGL((k_sweep23<2, 1, 3, true, false, true>), grd, blk, 0, st, g, s, A, 1, dt);
  } else {
    if (dma) hipLaunchKernelGGL((k_sweep23<2, 2, 3, true, false, true, true>), grd, blk, 0, st, g, s, A, 1, dt);
    else hipLaunchKernelGGL((k_sweep23<2, 2, 3, true, false, true>), grd, blk, 0, st, g, s, A, 1, dt);
  }
  timing_end("sweep2", st);
  timing_begin("post_convert", st);
  hipLaunchKernelGGL(k_post_convert_surface, dim3(blocks_for(surface_cells(g))), dim3(TPB), 0, st,
                     g, s);
  timing_end("post_convert", st);
}

bool postconvert_fusable(const DevGrid& g) { return g.lx2 > 1; }

void launch_floor_ghosts(const DevGrid& g, double* Ts, hipStream_t st) {
  hipLaunchKernelGGL(k_floor_ghosts, dim3(blocks_for(surface_cells(g)), LSP), dim3(TPB), 0, st, g,
                     Ts);
}

}  // namespace gemhip
