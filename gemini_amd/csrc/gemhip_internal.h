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
void launch_halo_periodic(const DevGrid& g, double* const* fields, int nfields, int lhalo,
                          int nspec, hipStream_t st);
void launch_interface_vels(const DevGrid& g, const DevState& s, int lsp, hipStream_t st);
void launch_save_v1_boundaries(const DevGrid& g, const DevState& s, hipStream_t st);
// dir 1,2,3; fields = F_* bits; onthefly: interface velocities from the cell velocities
// (bit-identical to the materialised ones) instead of intvars%vs*i.
void launch_sweep(const DevGrid& g, const DevState& s, int dir, int fields, double dt,
                  bool onthefly, hipStream_t st, const Tuning& tune = Tuning{});

// ---- launchers (kernels_diffusion.hip) ----
size_t diffusion_scratch_bytes(const DevGrid& g);
void launch_energy_diffusion(const DevGrid& g, const DevState& s, double dt, int diffsolvetype,
                             double Teinf, double* scratch, hipStream_t st,
                             const Tuning& tune = Tuning{});

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
};
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
void launch_source_loss_mass(const DevGrid& g, const DevState& s, const DevSources& a,
                             double dt, hipStream_t st);
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
