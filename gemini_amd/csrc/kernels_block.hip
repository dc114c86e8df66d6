// kernels_block.hip -- the rest of fluid_adv's transport block on gfx950: null-cell and
// ghost clean-up, artificial viscosity, the compression substep's velocity boundaries and
// the compression substep itself.
//
// Reference: src/multifluid/multifluid.f90:237-302 (VNRicht_artvisc, compression),
// :553-688 (clean_param), src/numerical/calculus/div.f90:118-206 (div3D_curv_23),
// src/sources/sources_mpi.f90:55-112 (RK2_global_boundary_allspec), with the order of
// src/libgemini.cpp:186-210.  Same arithmetic order as the reference; built with
// -ffp-contract=off like every kernel of the library.
#include "gemhip_internal.h"

namespace gemhip {

namespace {

constexpr int TPB = 256;
constexpr double xicon = 3.0;                                 // multifluid.f90:43

inline unsigned grid_exact(long n) { return (unsigned)((n + TPB - 1) / TPB); }
#ifndef BLOCK_GRID_CAP
// blocks of the grid-stride cell kernels: one thread per cell at the production size (a cap
// of 8192 blocks made the block step 0.6 ms slower at 512x256x192)
#define BLOCK_GRID_CAP 1048576
#endif
inline unsigned grid_capped(long n) {
  const unsigned b = grid_exact(n);
  return b > BLOCK_GRID_CAP ? BLOCK_GRID_CAP : b;
}

__device__ __forceinline__ void cell_of(const DevGrid& g, long c, int& i1, int& i2, int& i3) {
  i1 = (int)(c % g.L1) - 1;
  i2 = (int)((c / g.L1) % g.L2) - 1;
  i3 = (int)(c / ((long)g.L1 * g.L2)) - 1;
}

__device__ __forceinline__ bool interior(const DevGrid& g, int i1, int i2, int i3) {
  return i1 >= 1 && i1 <= g.lx1 && i2 >= 1 && i2 <= g.lx2 && i3 >= 1 && i3 <= g.lx3;
}

__device__ __forceinline__ bool is_null(const DevGrid& g, const int* nul, int i1, int i2, int i3) {
  return nul && nul[(i1 - 1) + (long)g.lx1 * ((i2 - 1) + (long)g.lx2 * (i3 - 1))] != 0;
}

// clean_param(1): ion densities floored at mindens, electrons = sum of ions, null cells
// and every ghost cell set to fill values.  One thread per ghosted cell, all species.
__global__ void k_clean_density(DevGrid g, double* ns, const int* nul) {
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < g.N;
       c += (long)gridDim.x * blockDim.x) {
    int i1, i2, i3;
    cell_of(g, c, i1, i2, i3);
    if (!interior(g, i1, i2, i3)) {
      for (int isp = 0; isp < LSP; ++isp) ns[isp * g.N + c] = mindensdiv;
      continue;
    }
    if (is_null(g, nul, i1, i2, i3)) {
      for (int isp = 0; isp < LSP; ++isp) ns[isp * g.N + c] = isp == 0 ? mindensnull * 1e-2 : mindensnull;
      continue;
    }
    double s = 0.0;
    for (int isp = 0; isp < LSP - 1; ++isp) {
      const double v = fmax_f(ns[isp * g.N + c], mindens);
      ns[isp * g.N + c] = v;
      s = s + v;
    }
    ns[(LSP - 1) * g.N + c] = s;
  }
}

// clean_param(3): temperature floor, null cells and ghost cells at 100 K
__global__ void k_clean_temperature(DevGrid g, double* Ts, const int* nul) {
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < g.N;
       c += (long)gridDim.x * blockDim.x) {
    int i1, i2, i3;
    cell_of(g, c, i1, i2, i3);
    const bool fill = !interior(g, i1, i2, i3) || is_null(g, nul, i1, i2, i3);
    for (int isp = 0; isp < LSP; ++isp) {
      double* p = Ts + isp * g.N + c;
      *p = fill ? 100.0 : fmax_f(*p, 100.0);
    }
  }
}

// clean_param(2): zero velocity in null cells, then on dipole grids copy the neighbour
// into the border cells of the null region of each field line (multifluid.f90:623-667).
// One thread per (x2, x3) column and species.
__global__ void k_clean_velocity(DevGrid g, double* vs1, const int* nul) {
  const long ncol = (long)g.lx2 * g.lx3;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncol * LSP) return;
  const int isp = (int)(t / ncol);
  const long col = t % ncol;
  const int i2 = (int)(col % g.lx2) + 1, i3 = (int)(col / g.lx2) + 1;
  double* p = vs1 + isp * g.N + gix(g, 0, i2, i3);     // p[i1]
  const int lx1 = g.lx1;
  for (int i1 = 1; i1 <= lx1; ++i1)
    if (is_null(g, nul, i1, i2, i3)) p[i1] = 0.0;
  if (g.gridflag == 0) {
    int beg = 1;
    while (!is_null(g, nul, beg, i2, i3) && beg < lx1) ++beg;
    int end = beg;
    while (is_null(g, nul, end, i2, i3) && end < lx1) ++end;
    if (beg != lx1) p[beg] = p[beg + 1];
    if (end != lx1) p[end] = p[end - 1];
  } else if (g.gridflag == 1) {
    int end = 1;
    while (!is_null(g, nul, end, i2, i3) && end < lx1) ++end;
    if (end != lx1) p[end] = p[end - 1];
  }
}

// VNRicht_artvisc: Q = ns ms 0.25 xicon^2 min(dv1, 0)^2 for the ions, 0 for electrons,
// dv1 the difference of the x1 interface velocities 0.5 (vs1(i-1) + vs1(i)) around the
// cell.  One thread per interior cell, all species.
__global__ void k_artvisc(DevGrid g, const double* ns, const double* vs1, double* Q) {
  const long ni = (long)g.lx1 * g.lx2 * g.lx3;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < ni;
       k += (long)gridDim.x * blockDim.x) {
    const int i1 = (int)(k % g.lx1) + 1;
    const long r = k / g.lx1;
    const int i2 = (int)(r % g.lx2) + 1, i3 = (int)(r / g.lx2) + 1;
    const long c = gix(g, i1, i2, i3);
    for (int isp = 1; isp < LSP; ++isp) {
      const double* u = vs1 + (isp - 1) * g.N;
      const double vl = 0.5 * (u[c - 1] + u[c]);
      const double vr = 0.5 * (u[c] + u[c + 1]);
      const double m = fmin_f(vr - vl, 0.0);
      Q[(isp - 1) * ni + k] = (((ns[(isp - 1) * g.N + c] * ms(isp)) * 0.25) * (xicon * xicon)) * (m * m);
    }
    Q[(LSP - 1) * ni + k] = 0.0;
  }
}

// RK2_global_boundary: zero-order hold of vs1 into its x1 ghosts 0 and lx1+1 (every x2,
// x3), then of vs1, vs2, vs3 into the x2 ghosts 0, lx2+1 (every x1, x3) and, unless
// periodic, the x3 ghosts 0, lx3+1 (every x1, x2).  Three launches keep that order.
__global__ void k_rk2_bc_x1(DevGrid g, double* vs1) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.L2 * g.L3) return;
  const int i2 = (int)(t % g.L2) - 1, i3 = (int)(t / g.L2) - 1;
  double* p = vs1 + blockIdx.y * g.N + gix(g, 0, i2, i3);
  p[0] = p[1];
  p[g.lx1 + 1] = p[g.lx1];
}

__global__ void k_rk2_bc_x2(DevGrid g, double* vs1, double* vs2, double* vs3) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.L1 * g.L3) return;
  const int i1 = (int)(t % g.L1) - 1, i3 = (int)(t / g.L1) - 1;
  double* f = (blockIdx.z == 0 ? vs1 : blockIdx.z == 1 ? vs2 : vs3) + blockIdx.y * g.N;
  f[gix(g, i1, 0, i3)] = f[gix(g, i1, 1, i3)];
  f[gix(g, i1, g.lx2 + 1, i3)] = f[gix(g, i1, g.lx2, i3)];
}

__global__ void k_rk2_bc_x3(DevGrid g, double* vs1, double* vs2, double* vs3) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.L1 * g.L2) return;
  const int i1 = (int)(t % g.L1) - 1, i2 = (int)(t / g.L1) - 1;
  double* f = (blockIdx.z == 0 ? vs1 : blockIdx.z == 1 ? vs2 : vs3) + blockIdx.y * g.N;
  f[gix(g, i1, i2, 0)] = f[gix(g, i1, i2, 1)];
  f[gix(g, i1, i2, g.lx3 + 1)] = f[gix(g, i1, i2, g.lx3)];
}

// compression: div(vs) by centred differences with the metric factors (div3D_curv_23 on
// the one-ghost extended grid, of which the interior is used), then the two-stage update
//   e_half = e - dt/2 ((gamma-1) e + Q) div,   e <- e - dt ((gamma-1) e_half + Q) div.
// One thread per interior cell, all species.
__global__ void k_compression(DevGrid g, double dt, const double* vs1, const double* vs2,
                              const double* vs3, const double* Q, double* rhoes) {
  const long ni = (long)g.lx1 * g.lx2 * g.lx3;
  const long s2 = g.L1, s3 = (long)g.L1 * g.L2;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < ni;
       k += (long)gridDim.x * blockDim.x) {
    const int i1 = (int)(k % g.lx1) + 1;
    const long r = k / g.lx1;
    const int i2 = (int)(r % g.lx2) + 1, i3 = (int)(r / g.lx2) + 1;
    const long c = gix(g, i1, i2, i3);
    const double* h1 = g.h1;
    const double* h2 = g.h2;
    const double* h3 = g.h3;
    const double m1a = h2[c + 1] * h3[c + 1], m1b = h2[c - 1] * h3[c - 1];
    const double d1 = g.dx1[i1 + 1] + g.dx1[i1];
    double m2a = 0.0, m2b = 0.0, d2 = 1.0, m3a = 0.0, m3b = 0.0, d3 = 1.0;
    if (g.lx2 > 1) { m2a = h1[c + s2] * h3[c + s2]; m2b = h1[c - s2] * h3[c - s2]; d2 = g.dx2[i2 + 1] + g.dx2[i2]; }
    if (g.lx3 > 1) { m3a = h1[c + s3] * h2[c + s3]; m3b = h1[c - s3] * h2[c - s3]; d3 = g.dx3[i3 + 1] + g.dx3[i3]; }
    const double hc = (h1[c] * h2[c]) * h3[c];
    for (int isp = 1; isp <= LSP; ++isp) {
      const long o = (isp - 1) * g.N;
      double dv = (m1a * vs1[o + c + 1] - m1b * vs1[o + c - 1]) / d1;
      if (g.lx2 > 1) dv = dv + (m2a * vs2[o + c + s2] - m2b * vs2[o + c - s2]) / d2;
      if (g.lx3 > 1) dv = dv + (m3a * vs3[o + c + s3] - m3b * vs3[o + c - s3]) / d3;
      dv = dv / hc;
      const double gm1 = gammas(isp) - 1.0;
      const double q = Q[(isp - 1) * ni + k];
      const double e0 = rhoes[o + c];
      const double half = e0 - ((dt / 2.0) * (e0 * gm1 + q)) * dv;
      rhoes[o + c] = e0 - (dt * (half * gm1 + q)) * dv;
    }
  }
}

// rhov12v1 (multifluid.f90:218-233) then clean_param(1), fused per cell
// (src/libgemini.cpp:186-187): velocities from the advected momentum densities with the
// densities as the sweeps left them, then the density clean-up.
__global__ void k_rhov12v1_clean_density(DevGrid g, double* ns, const double* rhovs1,
                                         double* vs1, const int* nul) {
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < g.N;
       c += (long)gridDim.x * blockDim.x) {
    double n[LSP - 1];
    double chrgflux = 0.0, sum = 0.0;
#pragma unroll
    for (int isp = 1; isp <= LSP - 1; ++isp) {
      const long o = (isp - 1) * g.N + c;
      n[isp - 1] = ns[o];
      const double v = rhovs1[o] / (ms(isp) * fmax_f(n[isp - 1], mindensdiv));
      vs1[o] = v;
      chrgflux = chrgflux + (n[isp - 1] * qs(isp)) * v;
    }
#pragma unroll
    for (int isp = 0; isp < LSP - 1; ++isp) sum = sum + n[isp];
    vs1[(LSP - 1) * g.N + c] = -(((1.0 / fmax_f(sum, mindensdiv)) / qs(LSP)) * chrgflux);
    int i1, i2, i3;
    cell_of(g, c, i1, i2, i3);
    if (!interior(g, i1, i2, i3)) {
      for (int isp = 0; isp < LSP; ++isp) ns[isp * g.N + c] = mindensdiv;
    } else if (is_null(g, nul, i1, i2, i3)) {
      for (int isp = 0; isp < LSP; ++isp) ns[isp * g.N + c] = isp == 0 ? mindensnull * 1e-2 : mindensnull;
    } else {
      double s = 0.0;
#pragma unroll
      for (int isp = 0; isp < LSP - 1; ++isp) {
        const double v = fmax_f(n[isp], mindens);
        ns[isp * g.N + c] = v;
        s = s + v;
      }
      ns[(LSP - 1) * g.N + c] = s;
    }
  }
}

// The tail of the compression substep fused per cell (src/libgemini.cpp:203-205):
// compression of the interior energy density, rhoe2T over every cell, clean_param(3).
// Ghost cells end at 100 K whatever rhoe2T gives them, and keep their energy density.
__global__ void k_compression_rhoe2T_clean(DevGrid g, double dt, const double* ns,
                                           const double* vs1, const double* vs2,
                                           const double* vs3, const double* Q, double* rhoes,
                                           double* Ts, const int* nul) {
  const long ni = (long)g.lx1 * g.lx2 * g.lx3;
  const long s2 = g.L1, s3 = (long)g.L1 * g.L2;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < g.N;
       c += (long)gridDim.x * blockDim.x) {
    int i1, i2, i3;
    cell_of(g, c, i1, i2, i3);
    if (!interior(g, i1, i2, i3)) {
      for (int isp = 0; isp < LSP; ++isp) Ts[isp * g.N + c] = 100.0;
      continue;
    }
    const bool isnull = is_null(g, nul, i1, i2, i3);
    const long k = (i1 - 1) + (long)g.lx1 * ((i2 - 1) + (long)g.lx2 * (i3 - 1));
    const double* h1 = g.h1;
    const double* h2 = g.h2;
    const double* h3 = g.h3;
    const double m1a = h2[c + 1] * h3[c + 1], m1b = h2[c - 1] * h3[c - 1];
    const double d1 = g.dx1[i1 + 1] + g.dx1[i1];
    double m2a = 0.0, m2b = 0.0, d2 = 1.0, m3a = 0.0, m3b = 0.0, d3 = 1.0;
    if (g.lx2 > 1) { m2a = h1[c + s2] * h3[c + s2]; m2b = h1[c - s2] * h3[c - s2]; d2 = g.dx2[i2 + 1] + g.dx2[i2]; }
    if (g.lx3 > 1) { m3a = h1[c + s3] * h2[c + s3]; m3b = h1[c - s3] * h2[c - s3]; d3 = g.dx3[i3 + 1] + g.dx3[i3]; }
    const double hc = (h1[c] * h2[c]) * h3[c];
    for (int isp = 1; isp <= LSP; ++isp) {
      const long o = (isp - 1) * g.N;
      double dv = (m1a * vs1[o + c + 1] - m1b * vs1[o + c - 1]) / d1;
      if (g.lx2 > 1) dv = dv + (m2a * vs2[o + c + s2] - m2b * vs2[o + c - s2]) / d2;
      if (g.lx3 > 1) dv = dv + (m3a * vs3[o + c + s3] - m3b * vs3[o + c - s3]) / d3;
      dv = dv / hc;
      const double gm1 = gammas(isp) - 1.0;
      const double q = Q[(isp - 1) * ni + k];
      const double e0 = rhoes[o + c];
      const double half = e0 - ((dt / 2.0) * (e0 * gm1 + q)) * dv;
      const double e = e0 - (dt * (half * gm1 + q)) * dv;
      rhoes[o + c] = e;
      const double t = fmax_f(((gm1 / kB) * e) / fmax_f(ns[o + c], mindensdiv), 100.0);
      Ts[o + c] = isnull ? 100.0 : fmax_f(t, 100.0);
    }
  }
}

// The end of the block fused per cell (src/libgemini.cpp:209-210): clean_param(3), then
// T2rhoe over every cell.
__global__ void k_clean_T2rhoe(DevGrid g, const double* ns, double* Ts, double* rhoes,
                               const int* nul) {
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < g.N;
       c += (long)gridDim.x * blockDim.x) {
    int i1, i2, i3;
    cell_of(g, c, i1, i2, i3);
    const bool fill = !interior(g, i1, i2, i3) || is_null(g, nul, i1, i2, i3);
    for (int isp = 1; isp <= LSP; ++isp) {
      const long o = (isp - 1) * g.N + c;
      const double t = fill ? 100.0 : fmax_f(Ts[o], 100.0);
      Ts[o] = t;
      rhoes[o] = ((ns[o] * kB) * t) / (gammas(isp) - 1.0);
    }
  }
}

// dt_calc's grid rates (temporal_mpi.f90:115-131): per species the largest
// (vsnd + |vs1|)/dl1, |vs2|/dl2, |vs3|/dl3 over the interior, dl_k = dxki h_k
// (meshobj.f90:436-442).  Maxima are exact, so the grid-stride partial maxima, the
// 64-lane shuffle tree and the final pass give the reference's values in any order.
constexpr int NRATE = 3 * LSP;

__global__ __launch_bounds__(TPB) void k_gridrate(DevGrid g, const double* Ts, const double* vs1,
                                                  const double* vs2, const double* vs3,
                                                  double* part) {
  double r[NRATE];
#pragma unroll
  for (int k = 0; k < NRATE; ++k) r[k] = 0.0;
  const long ni = (long)g.lx1 * g.lx2 * g.lx3;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < ni;
       k += (long)gridDim.x * blockDim.x) {
    const int i1 = (int)(k % g.lx1) + 1;
    const long rr = k / g.lx1;
    const int i2 = (int)(rr % g.lx2) + 1, i3 = (int)(rr / g.lx2) + 1;
    const long c = gix(g, i1, i2, i3);
    const double dl1 = g.dx1i[i1 - 1] * g.h1[c], dl2 = g.dx2i[i2 - 1] * g.h2[c],
                 dl3 = g.dx3i[i3 - 1] * g.h3[c];
    const double Te = Ts[(LSP - 1) * g.N + c];
#pragma unroll
    for (int isp = 1; isp <= LSP; ++isp) {
      const long o = (isp - 1) * g.N + c;
      const double vsnd = isp < LSP
          ? sqrt((kB * Ts[o]) / ms(isp) + (((5.0 / 3.0) * kB) * Te) / ms(isp)) : 0.0;
      r[3 * (isp - 1)] = fmax_f((vsnd + fabs(vs1[o])) / dl1, r[3 * (isp - 1)]);
      r[3 * (isp - 1) + 1] = fmax_f(fabs(vs2[o]) / dl2, r[3 * (isp - 1) + 1]);
      r[3 * (isp - 1) + 2] = fmax_f(fabs(vs3[o]) / dl3, r[3 * (isp - 1) + 2]);
    }
  }
  __shared__ double red[TPB / 64][NRATE];
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
#pragma unroll
  for (int k = 0; k < NRATE; ++k) {
    double v = r[k];
    for (int off = 32; off >= 1; off >>= 1) v = fmax_f(__shfl_xor(v, off), v);
    if (lane == 0) red[wave][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < NRATE) {
    double m = red[0][threadIdx.x];
    for (int w = 1; w < TPB / 64; ++w) m = fmax_f(red[w][threadIdx.x], m);
    part[(long)blockIdx.x * NRATE + threadIdx.x] = m;
  }
}

// one block per rate: the block's threads stride over the partial maxima, then a wave
// shuffle tree and the block's four waves (a lone thread per rate walking the partials
// took a third as long as the whole reduction kernel)
__global__ __launch_bounds__(TPB) void k_gridrate_final(const double* part, int nblocks,
                                                        double* out) {
  const int k = blockIdx.x;
  double m = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) m = fmax_f(part[(long)b * NRATE + k], m);
  for (int off = 32; off >= 1; off >>= 1) m = fmax_f(__shfl_xor(m, off), m);
  __shared__ double red[TPB / 64];
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int w = 1; w < TPB / 64; ++w) r = fmax_f(red[w], r);
    out[k] = r;
  }
}

}  // namespace

int gridrate_blocks(const DevGrid& g) {
  const long ni = (long)g.lx1 * g.lx2 * g.lx3;
  const long b = (ni + TPB - 1) / TPB;
  return (int)(b > 1024 ? 1024 : b);
}

void launch_gridrates(const DevGrid& g, const DevState& s, double* work, hipStream_t st) {
  const int nb = gridrate_blocks(g);
  timing_begin("gridrate", st);
  hipLaunchKernelGGL(k_gridrate, dim3(nb), dim3(TPB), 0, st, g, s.Ts, s.vs1, s.vs2, s.vs3, work);
  hipLaunchKernelGGL(k_gridrate_final, dim3(NRATE), dim3(TPB), 0, st, work, nb,
                     work + (long)nb * NRATE);
  timing_end("gridrate", st);
}

void launch_compression_rhoe2T_clean(const DevGrid& g, const DevState& s, const double* Q,
                                     double dt, const int* nul, hipStream_t st) {
  timing_begin("compression", st);
  hipLaunchKernelGGL(k_compression_rhoe2T_clean,
                     dim3(grid_capped(g.N)), dim3(TPB), 0, st, g,
                     dt, s.ns, s.vs1, s.vs2, s.vs3, Q, s.rhoes, s.Ts, nul);
  timing_end("compression", st);
}

void launch_rhov12v1_clean_density(const DevGrid& g, const DevState& s, const int* nul,
                                   hipStream_t st) {
  timing_begin("rhov12v1", st);
  hipLaunchKernelGGL(k_rhov12v1_clean_density,
                     dim3(grid_capped(g.N)), dim3(TPB), 0, st, g,
                     s.ns, s.rhovs1, s.vs1, nul);
  timing_end("rhov12v1", st);
}

void launch_clean_T2rhoe(const DevGrid& g, const DevState& s, const int* nul, hipStream_t st) {
  timing_begin("clean_param", st);
  hipLaunchKernelGGL(k_clean_T2rhoe, dim3(grid_capped(g.N)),
                     dim3(TPB), 0, st, g, s.ns, s.Ts, s.rhoes, nul);
  timing_end("clean_param", st);
}

void launch_clean_param(const DevGrid& g, const DevState& s, int iparm, const int* nul,
                        hipStream_t st) {
  timing_begin("clean_param", st);
  const unsigned cells = grid_capped(g.N);
  if (iparm == 1) {
    hipLaunchKernelGGL(k_clean_density, dim3(cells), dim3(TPB), 0, st, g, s.ns, nul);
  } else if (iparm == 3) {
    hipLaunchKernelGGL(k_clean_temperature, dim3(cells), dim3(TPB), 0, st, g, s.Ts, nul);
  } else if (nul) {           // without null cells the velocity clean-up changes nothing
    hipLaunchKernelGGL(k_clean_velocity, dim3(grid_exact((long)g.lx2 * g.lx3 * LSP)), dim3(TPB),
                       0, st, g, s.vs1, nul);
  }
  timing_end("clean_param", st);
}

void launch_artvisc(const DevGrid& g, const DevState& s, double* Q, hipStream_t st) {
  const long ni = (long)g.lx1 * g.lx2 * g.lx3;
  timing_begin("artvisc", st);
  hipLaunchKernelGGL(k_artvisc, dim3(grid_capped(ni)), dim3(TPB),
                     0, st, g, s.ns, s.vs1, Q);
  timing_end("artvisc", st);
}

void launch_rk2_global_boundary(const DevGrid& g, const DevState& s, hipStream_t st) {
  timing_begin("rk2_boundary", st);
  hipLaunchKernelGGL(k_rk2_bc_x1, dim3(grid_exact((long)g.L2 * g.L3), LSP), dim3(TPB), 0, st, g,
                     s.vs1);
  hipLaunchKernelGGL(k_rk2_bc_x2, dim3(grid_exact((long)g.L1 * g.L3), LSP, 3), dim3(TPB), 0, st,
                     g, s.vs1, s.vs2, s.vs3);
  if (!g.flagper)
    hipLaunchKernelGGL(k_rk2_bc_x3, dim3(grid_exact((long)g.L1 * g.L2), LSP, 3), dim3(TPB), 0, st,
                       g, s.vs1, s.vs2, s.vs3);
  timing_end("rk2_boundary", st);
}

void launch_compression(const DevGrid& g, const DevState& s, const double* Q, double dt,
                        hipStream_t st) {
  const long ni = (long)g.lx1 * g.lx2 * g.lx3;
  timing_begin("compression", st);
  hipLaunchKernelGGL(k_compression, dim3(grid_capped(ni)),
                     dim3(TPB), 0, st, g, dt, s.vs1, s.vs2, s.vs3, Q, s.rhoes);
  timing_end("compression", st);
}

}  // namespace gemhip
