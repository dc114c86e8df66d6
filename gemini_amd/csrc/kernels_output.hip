// kernels_output.hip -- plasma output packing (io_nompi.f90:183-258).
//
// One pass over the interior cells of the HBM state writes every real32 field dataset that
// plasma_output_nompi puts in the file, back to back in dataset order, so a single
// device-to-host copy of about a quarter of the double-precision state (no ghost cells, 4
// bytes a value) feeds the HDF5 writer.  The ion averages of io_nompi.f90:199-213 are
// formed here in double precision, in the reference's order: sum over the six ions of
// ns*v accumulated from zero, then divided by the electron density; every value is then
// rounded to the nearest real32 (Fortran's real()).
//
// The patch writer declares J1, J2, J3 as dimension(1:,1:,1:) (io_nompi.f90:192) and is
// handed the ghosted electrovars fields (libgemini.f90:549-551), so its J(1:lx1,1:lx2,1:lx3)
// is the ghosted block's ghost-origin corner: cell (i1, i2, i3) of its file holds J at
// ghosted indices (i1-2, i2-2, i3-2).  The root writer's J (plasma_output.f90:46) keeps
// the ghosted bounds, so its file holds the interior.  `jshift` is the offset between the
// two (0 for the root writer).
//
// Per interior cell the full output (flagoutput 1) reads 36 doubles (ns, vs1, Ts of the 7
// species, vs2 and vs3 of the 6 ions, J1-J3) and writes 26 floats: 392 bytes; flagoutput 2
// reads 35 doubles and writes 9 floats (316 bytes); flagoutput 3 reads and writes the
// electron density only (12 bytes).  The kernel is HBM bound.
#include <hip/hip_runtime.h>

#include "gemhip_internal.h"

namespace gemhip {
namespace {

constexpr int TPB = 256;

// the datasets of one interior cell (i1, line), line = (i2 - 1) + lx2 (i3 - 1)
template <int FLAG>
__device__ __forceinline__ void pack_cell(const DevGrid& g, const double* __restrict__ fluid,
                                          const double* __restrict__ J, float* __restrict__ out,
                                          int i1, int line, long jshift) {
  const int i2 = line % g.lx2;
  const int i3 = line / g.lx2;
  const long n = (long)g.lx1 * g.lx2 * g.lx3;
  const long k = i1 + (long)g.lx1 * line;
  const long N = g.N, B = LSP * N;
  const long c = gix(g, i1 + 1, i2 + 1, i3 + 1);
  const double* ns = fluid;
  const double* vs1 = fluid + B;
  const double* vs2 = fluid + 2 * B;
  const double* vs3 = fluid + 3 * B;
  const double* Ts = fluid + 4 * B;
  const double ne = ns[(LSP - 1) * N + c];
  if (FLAG == 3) {
    out[k] = (float)ne;
  } else {
    double a2 = 0.0, a3 = 0.0;
#pragma unroll
    for (int isp = 0; isp < LSP - 1; ++isp) {
      const double d = ns[isp * N + c];
      a2 = a2 + d * vs2[isp * N + c];
      a3 = a3 + d * vs3[isp * N + c];
    }
    const long cj = c - jshift;
    const float J1 = (float)J[cj];
    const float J2 = (float)J[N + cj];
    const float J3 = (float)J[2 * N + cj];
    if (FLAG == 2) {
      double a1 = 0.0, aT = 0.0;
#pragma unroll
      for (int isp = 0; isp < LSP - 1; ++isp) {
        const double d = ns[isp * N + c];
        a1 = a1 + d * vs1[isp * N + c];
        aT = aT + d * Ts[isp * N + c];
      }
      out[k] = (float)ne;
      out[n + k] = (float)(a1 / ne);
      out[2 * n + k] = (float)(aT / ne);
      out[3 * n + k] = (float)Ts[(LSP - 1) * N + c];
      out[4 * n + k] = J1;
      out[5 * n + k] = J2;
      out[6 * n + k] = J3;
      out[7 * n + k] = (float)(a2 / ne);
      out[8 * n + k] = (float)(a3 / ne);
    } else {
#pragma unroll
      for (int isp = 0; isp < LSP; ++isp) {
        out[isp * n + k] = (float)ns[isp * N + c];
        out[(LSP + isp) * n + k] = (float)vs1[isp * N + c];
        out[(2 * LSP + isp) * n + k] = (float)Ts[isp * N + c];
      }
      out[3 * LSP * n + k] = J1;
      out[(3 * LSP + 1) * n + k] = J2;
      out[(3 * LSP + 2) * n + k] = J3;
      out[(3 * LSP + 3) * n + k] = (float)(a2 / ne);
      out[(3 * LSP + 4) * n + k] = (float)(a3 / ne);
    }
  }
  // the (lx2, lx3) potential slab at the top of the field lines: x1 = 1 on an inverted
  // grid, x1 = lx1 otherwise (io_nompi.f90:260-266)
  const int top = g.gridflag == 1 ? 0 : g.lx1 - 1;
  if (i1 == top) {
    const long base = FLAG == 2 ? 9 * n : FLAG == 3 ? n : (3L * LSP + 5) * n;
    out[base + i2 + (long)g.lx2 * i3] = (float)J[3 * N + c];
  }
}

#ifndef OUTPUT_CELLS_PER_THREAD
// two cells a thread, a block apart: 1-2 % faster than one for the full and ISR-like
// outputs and 18 % for the electron density alone at 512x256x192
#define OUTPUT_CELLS_PER_THREAD 2
#endif
constexpr int CPT = OUTPUT_CELLS_PER_THREAD;

// blocks in memory order: block b covers part b % nb of x1 line b / nb (nb blocks a line,
// CPT cells a thread, a block's cells interleaved so each load instruction stays
// coalesced), 32-bit index arithmetic only
template <int FLAG>
__global__ __launch_bounds__(TPB) void k_output_pack(DevGrid g, const double* __restrict__ fluid,
                                                      const double* __restrict__ J,
                                                      float* __restrict__ out, long jshift) {
  const int per = TPB * CPT;
  const int nb = (g.lx1 + per - 1) / per;
  const int line = (int)(blockIdx.x / (unsigned)nb);
  const int i0 = (int)(blockIdx.x % (unsigned)nb) * per + threadIdx.x;
#pragma unroll
  for (int u = 0; u < CPT; ++u) {
    const int i1 = i0 + u * TPB;
    if (i1 < g.lx1) pack_cell<FLAG>(g, fluid, J, out, i1, line, jshift);
  }
}

}  // namespace

long output_count(const DevGrid& g, int flagoutput) {
  const long n = (long)g.lx1 * g.lx2 * g.lx3, m = (long)g.lx2 * g.lx3;
  if (flagoutput == 2) return 9 * n + m;
  if (flagoutput == 3) return n + m;
  return (3L * LSP + 5) * n + m;
}

void launch_output_pack(const DevGrid& g, const double* fluid, const double* J,
                        int flagoutput, bool patch, float* out, hipStream_t st) {
  const long jshift = patch ? 2 + 2L * g.L1 + 2L * g.L1 * g.L2 : 0;
  const dim3 grid((unsigned)(g.lx2 * g.lx3) * (unsigned)((g.lx1 + TPB * CPT - 1) / (TPB * CPT)));
  timing_begin("output_pack", st);
  if (flagoutput == 2)
    k_output_pack<2><<<grid, TPB, 0, st>>>(g, fluid, J, out, jshift);
  else if (flagoutput == 3)
    k_output_pack<3><<<grid, TPB, 0, st>>>(g, fluid, J, out, jshift);
  else
    k_output_pack<1><<<grid, TPB, 0, st>>>(g, fluid, J, out, jshift);
  timing_end("output_pack", st);
}

}  // namespace gemhip
