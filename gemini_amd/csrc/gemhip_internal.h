// gemhip_internal.h -- device-side views of the transport state and the kernel
// launchers shared by the kernel files and the C ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include "gemhip_math.h"

namespace gemhip {

// Geometry in HBM, same layouts as the reference's curvmesh arrays (Fortran order,
// x1 fastest; ghosted arrays start at index -1).
struct DevGrid {
  int lx1, lx2, lx3;
  int L1, L2, L3;          // lxN + 4
  long N;                  // L1*L2*L3, one ghosted field
  int gridflag, flagper;
  int isglobal[6];         // x1min x1max x2min x2max x3min x3max
  const double *h1, *h2, *h3;
  const double *h1x1i, *h2x1i, *h3x1i;
  const double *h1x2i, *h2x2i, *h3x2i;
  const double *h1x3i, *h2x3i, *h3x3i;
  const double *dx1, *dx1i, *dx2, *dx2i, *dx3, *dx3i;
  // h1, h2, h3 of the interior rows in the energy diffusion's column-group order: factor m
  // of row r (1..lx1) of column gcol at hcol[m hstride + (cg lx1 + r-1) 64 + lane], (cg,
  // lane) the column's group and lane (hcol_index in diffusion_cell.h)
  const double* hcol;
  long hstride;
};

// State pointers in HBM.  ns..Ts point into one fluidvars block of 35 fields,
// rhovs1/rhoes into a 14-field auxiliary block; each species field is N long.
struct DevState {
  double *ns, *vs1, *vs2, *vs3, *Ts;
  double *rhovs1, *rhoes;
  const double *J1;
  double *vs1i, *vs2i, *vs3i;   // materialised interface velocities (intvars%vs*i)
  double *v1b;                  // v1i(1,:,:,:) and v1i(lx1+1,:,:,:): (2, lx2, lx3, 7)
  const double *nn, *Tn;        // nn(lx1,lx2,lx3,1:3), Tn(lx1,lx2,lx3)
  // parameter-fitting extension, null unless registered: electron conductivity scale and
  // downward top heat flux [W/m^2] per (x2,x3) column, lx2*lx3 each
  const double *kappa, *qtop;
  // null unless an adjoint tape is being recorded: the row interchanges of the electron
  // columns' two band factorisations (TR, BDF2), one bit per elimination step, 32 steps a
  // word, [column][stage][word] (fitting extension kernels only)
  int* pivrec;
};

__host__ __device__ inline long gix(const DevGrid& g, int i1, int i2, int i3) {
  return (long)(i1 + 1) + (long)g.L1 * ((long)(i2 + 1) + (long)g.L2 * (long)(i3 + 1));
}

// advec.f90:171-196: x1 interface velocity k (1..lx1+1) of line (i2,i3) from the cell
// velocities u(-1..lx1+2) of that line.
__device__ __forceinline__ double v1i_at(const DevGrid& g, const double* u, int k) {
  const int lx1 = g.lx1;
  if (k >= 2 && k <= lx1) return 0.5 * (u[k - 1] + u[k]);
  if (k == 1) {
    if (g.isglobal[0]) {
      if (g.gridflag == 1) return fmin_f(0.5 * (u[1] + u[2]), 0.0);
      return u[1];
    }
    return 0.5 * (u[0] + u[1]);
  }
  if (g.isglobal[1]) {
    if (g.gridflag == 2) return fmax_f(0.5 * (u[lx1 - 1] + u[lx1]), 0.0);
    return u[lx1];
  }
  return 0.5 * (u[lx1] + u[lx1 + 1]);
}

// Kernel variants a context may select (compiled alongside each other; an autotuner picks
// per grid size): the x2/x3 sweep pipeline depth for three-field launches (1 or 2) and
// the TRBDF2 chunk height (4 or 8 rows).
struct Tuning {
  int sweep_pf = 2;
  int diff_chunk = 8;
};

// field-selection bits for the sweeps (multifluid.f90:49-126)
enum : int { F_MASS = 1, F_MOMENTUM = 2, F_ENERGY = 4, F_ALL = 7 };

// ---- launchers (kernels_advec.hip) ----
void launch_v12rhov1_T2rhoe(const DevGrid& g, const DevState& s, bool mom, bool energy,
                            hipStream_t st);
void launch_rhov12v1(const DevGrid& g, const DevState& s, hipStream_t st);
void launch_rhoe2T(const DevGrid& g, const DevState& s, hipStream_t st);
void launch_post_convert(const DevGrid& g, const DevState& s, hipStream_t st);  // both above
void launch_set_global_boundaries(const DevGrid& g, const DevState& s, int lsp, hipStream_t st);
// the fused step's form of v12rhov1 + T2rhoe + ghost fill + x3 sweep: the conversions are
// folded into the ghost fill and the x3 sweep (all sides global, x3 not periodic)
bool preconvert_fusable(const DevGrid& g);
void launch_set_global_boundaries_preconvert(const DevGrid& g, const DevState& s, hipStream_t st);
void launch_sweep3_preconvert(const DevGrid& g, const DevState& s, double dt, hipStream_t st,
                              const Tuning& tune);
// the fused step's x2 sweep with rhov12v1 + rhoe2T folded in (x2 not a single cell)
bool postconvert_fusable(const DevGrid& g);
void launch_sweep2_postconvert(const DevGrid& g, const DevState& s, double dt, hipStream_t st,
                               const Tuning& tune);
void launch_halo_periodic(const DevGrid& g, double* const* fields, int nfields, int lhalo,
                          int nspec, hipStream_t st);
void launch_interface_vels(const DevGrid& g, const DevState& s, int lsp, hipStream_t st);
void launch_save_v1_boundaries(const DevGrid& g, const DevState& s, hipStream_t st);
// dir 1,2,3; fields = F_* bits; onthefly: interface velocities from the cell velocities
// (bit-identical to the materialised ones) instead of intvars%vs*i.  An on-the-fly x1
// sweep of all fields reads the x1-boundary interface velocities from s.v1b, so the
// caller must have run launch_save_v1_boundaries on the same vs1 (the step does, before
// its x3 sweep; no sweep writes vs1).
void launch_sweep(const DevGrid& g, const DevState& s, int dir, int fields, double dt,
                  bool onthefly, hipStream_t st, const Tuning& tune = Tuning{});

// ---- launchers (kernels_diffusion.hip) ----
size_t diffusion_scratch_bytes(const DevGrid& g);
long diffusion_hcol_doubles(const DevGrid& g);              // size of DevGrid::hcol
void launch_fill_hcol(const DevGrid& g, double* hcol, hipStream_t st);

// ---- batched ensemble steps (kernels_advec.hip, kernels_diffusion.hip) ----
// One member of an ensemble advanced by batched kernels: every kernel of the fused
// transport step runs once for all members, as a persistent grid whose blocks take the
// members' work units in turn (the same per-cell arithmetic as the member's own step).
struct BatchMember {
  DevGrid g;
  DevState s;
  double dt;
  double Teinf;
  double* scratch;          // the member's TRBDF2 work space
};
// the batched kernels of a step, in order; units are threads or blocks as noted
enum BatchKernel : int {
  BK_BC1,    // x1 ghost fill: threads, L2 L3 lsp
  BK_BC2,    // x2 ghost fill: threads, L1 L3 lsp (times 4 fields in blockIdx.y)
  BK_BC3,    // x3 ghost fill: threads, L1 L2 lsp (times 4 fields in blockIdx.y)
  BK_V1B,    // x1 boundary interface velocities: threads, lx2 lx3 lsp
  BK_SW3,    // x3 sweep with the conversions: blocks (64 x 7), sweep_blocks(ceil(lx1/64), lx2)
  BK_SW1,    // x1 sweep (the electron wave: energy only): blocks (64 x 7), lx2 lx3
  BK_SW2,    // x2 sweep with rhov12v1 + rhoe2T: blocks (64 x 7), sweep_blocks(.., lx3)
  BK_POST,   // rhov12v1 + rhoe2T outside the interior box: threads, surface cells
  BK_DIFF,   // TRBDF2: blocks (64), 8 ceil(ceil(lx2 lx3/64)/8) lsp (a multiple of 8)
  BK_FLOOR,  // Ts ghost floor: threads, surface cells lsp
  BK_COUNT
};
long batch_units(const DevGrid& g, int kernel);
__host__ __device__ inline long surface_cells(const DevGrid& g) {
  return 4L * g.lx2 * g.lx3 + (4L * g.L2 + 4L * g.lx3) * g.L1;
}
// threads of a member's own launch of a thread kernel (the x2/x3 fills: per field)
__host__ __device__ inline long batch_thread_units(const DevGrid& g, int kernel) {
  switch (kernel) {
    case BK_BC1: return (long)g.L2 * g.L3 * LSP;
    case BK_BC2: return (long)g.L1 * g.L3 * LSP;
    case BK_BC3: return (long)g.L1 * g.L2 * LSP;
    case BK_V1B: return (long)g.lx2 * g.lx3 * LSP;
    case BK_POST: return surface_cells(g);
    case BK_FLOOR: return surface_cells(g) * LSP;
  }
  return 0;
}
// the member whose units of a kernel include unit t: the last m with pre[m] <= t
__device__ __forceinline__ int batch_member(const long* pre, int n, long t) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
// members (HBM, n entries) and the prefix sums of each kernel's units (HBM,
// [BK_COUNT][n + 1]); totals on the host; grids: blocks of the persistent launches
struct BatchTable {
  const BatchMember* m;
  const long* pre;
  int n;
  long total[BK_COUNT];
  int blocks_thread;        // persistent grid of the thread kernels (256-thread blocks)
  int blocks_sweep;         // of the sweep kernels (one block per CU resident)
  int blocks_diff;          // of the TRBDF2 kernel
};
void launch_batch_advection(const BatchTable& t, int sweep_pf, hipStream_t st);
void launch_batch_diffusion(const BatchTable& t, int diff_chunk, hipStream_t st);
void launch_batch_floor(const BatchTable& t, hipStream_t st);
// floor_ghosts: then Ts of the ghost cells floored at 100 K (a caller whose ghost
// temperatures are floored already skips it: the solve writes interior rows only)
void launch_energy_diffusion(const DevGrid& g, const DevState& s, double dt, int diffsolvetype,
                             double Teinf, double* scratch, hipStream_t st,
                             const Tuning& tune = Tuning{}, bool floor_ghosts = true);

// ---- the rest of fluid_adv's transport block (kernels_block.hip) ----
// iparm 1 density, 2 velocity (vs1), 3 temperature; nul: null-cell mask (lx1*lx2*lx3,
// nonzero = null) or null for none.  Q is (lx1, lx2, lx3, 7).
void launch_clean_param(const DevGrid& g, const DevState& s, int iparm, const int* nul,
                        hipStream_t st);
void launch_artvisc(const DevGrid& g, const DevState& s, double* Q, hipStream_t st);
void launch_rk2_global_boundary(const DevGrid& g, const DevState& s, hipStream_t st);
void launch_compression(const DevGrid& g, const DevState& s, const double* Q, double dt,
                        hipStream_t st);
// the same passes fused per cell for the block step: compression + rhoe2T + clean_param(3),
// and clean_param(3) + T2rhoe
void launch_compression_rhoe2T_clean(const DevGrid& g, const DevState& s, const double* Q,
                                     double dt, const int* nul, hipStream_t st);
void launch_clean_T2rhoe(const DevGrid& g, const DevState& s, const int* nul, hipStream_t st);
void launch_rhov12v1_clean_density(const DevGrid& g, const DevState& s, const int* nul,
                                   hipStream_t st);

// dt_calc's 21 grid rates (species-major, x1 x2 x3): work holds gridrate_blocks(g) * 21
// partial maxima followed by the 21 results.
int gridrate_blocks(const DevGrid& g);
void launch_gridrates(const DevGrid& g, const DevState& s, double* work, hipStream_t st);

// ---- adjoint of the electron energy path (kernels_adjoint.hip) ----
// What one recorded forward step keeps for the adjoint (electron fields only): densities at
// the step start and at the diffusion, the energy density entering each sweep, the
// temperature entering the diffusion and the three interface-velocity arrays.
struct AdjTape {
  const double *NE0, *ES3, *ES1, *ES2, *NE1, *TD;   // ghosted fields, N each
  const double *V1, *V2, *V3;                       // (lx1+1,lx2,lx3), (lx1,lx2+1,lx3), (lx1,lx2,lx3+1)
  int* PIV;                                         // DevState::pivrec of the recorded step
};
// ints of DevState::pivrec for a grid
long pivot_record_ints(const DevGrid& g);
size_t adjoint_scratch_bytes(const DevGrid& g);
void launch_tape_vels(const DevGrid& g, const DevState& s, double* v1, double* v2, double* v3,
                      hipStream_t st);
// lam (ghosted field): dL/dTe after the step in, dL/dTe before it out; gk, gq (lx2*lx3,
// either may be null) accumulate dL/dkappa and dL/dq; work is a second field.
void launch_adjoint_step(const DevGrid& g, const DevState& s, const AdjTape& t, double dt,
                         double Teinf, double* lam, double* work, double* gk, double* gq,
                         double* scratch, hipStream_t st);

// ---- the source/loss substep (kernels_sources.hip) ----
// The reference's intvars%atmos, ionization rates and the mesh's altitude and gravity, in
// HBM: nn (lx1,lx2,lx3,6) O, N2, O2, H, N, NO; vn1-3, Qeionize, g1 (lx1,lx2,lx3); Prionize
// (lx1,lx2,lx3,6); alt and E1 ghosted fields; Q the artificial viscosity (lx1,lx2,lx3,7).
// The neutral temperature is DevState::Tn.
struct DevSources {
  const double *nn, *vn1, *vn2, *vn3, *Prionize, *Qeionize, *g1, *alt, *E1, *Q;
};
void launch_source_loss_energy(const DevGrid& g, const DevState& s, const DevSources& a,
                               double dt, hipStream_t st);
void launch_source_loss_momentum(const DevGrid& g, const DevState& s, const DevSources& a,
                                 double dt, hipStream_t st);
// clean: also clean_param 3 and 1 of the sources block (null cells from clean_nul)
void launch_source_loss_mass(const DevGrid& g, const DevState& s, const DevSources& a,
                             double dt, hipStream_t st, const int* clean_nul = nullptr,
                             bool clean = false);
void launch_zero(double* p, long n, hipStream_t st);

// ---- plasma output (kernels_output.hip) ----
// Number of real32 values of the field datasets plasma_output_nompi writes for flagoutput
// (io_nompi.f90:230-266), and the kernel that forms them from the HBM state: fluid is the
// 35-field fluidvars block, J the ghosted fields J1, J2, J3, Phi back to back; out receives
// the datasets in file order (Phiall last).
long output_count(const DevGrid& g, int flagoutput);
// patch: J1-J3 as the patch writer reads them (ghost-origin corner), else the interior
void launch_output_pack(const DevGrid& g, const double* fluid, const double* J, int flagoutput,
                        bool patch, float* out, hipStream_t st);

// kernel-timing hooks implemented by the API layer
void timing_begin(const char* name, hipStream_t st);
void timing_end(const char* name, hipStream_t st);

}  // namespace gemhip
