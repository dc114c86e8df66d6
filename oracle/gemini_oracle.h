/*
 * gemini_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement (plain C, IEEE double, no FMA contraction) of the GEMINI
 * multifluid-transport hot path: MC-limited finite-volume advection sweeps
 * and TRBDF2 / backward-Euler field-aligned energy diffusion with LAPACK
 * dgbsv semantics.  It is the parity checker for the HIP library in
 * gemini_amd/ and the CPU baseline timed by bench.py.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path never calls it.
 *
 * Array conventions follow the reference (Fortran column-major, x1 fastest):
 *   ghosted cell arrays  A(-1:lx1+2,-1:lx2+2,-1:lx3+2[,isp])
 *   x1-interface arrays  A(1:lx1+1,1:lx2,1:lx3[,isp])
 *   x2-interface arrays  A(1:lx1,1:lx2+1,1:lx3[,isp])
 *   x3-interface arrays  A(1:lx1,1:lx2,1:lx3+1[,isp])
 *   interior arrays      A(1:lx1,1:lx2,1:lx3[,k])
 *   dxN(0:lxN+2) and dxNi(1:lxN) one-dimensional
 * All species counts are the reference's lsp = 7.
 */
#ifndef GEMINI_ORACLE_H
#define GEMINI_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int lx1, lx2, lx3;
  int gridflag;          /* meshobj.f90:563-569: 0 closed dipole, 1 inverted, 2 standard */
  int flagper;           /* x3 periodic (meshobj.f90 flagper) */
  int isglobal[6];       /* grid.f90:448-507: x1min,x1max,x2min,x2max,x3min,x3max */
  const double *h1, *h2, *h3;                 /* ghosted cell-centred metric factors */
  const double *h1x1i, *h2x1i, *h3x1i;        /* x1 interfaces */
  const double *h1x2i, *h2x2i, *h3x2i;        /* x2 interfaces */
  const double *h1x3i, *h2x3i, *h3x3i;        /* x3 interfaces */
  const double *dx1, *dx1i, *dx2, *dx2i, *dx3, *dx3i;
} orc_grid;

/* constants (phys_consts.in.f90) exported so tests can read what the oracle uses */
void orc_constants(double *ms7, double *qs7, double *gammas7, double *kB, double *elchrg,
                   double *thermal_coeff6);

double orc_minmod(double a, double b);
void orc_advec1D_MC_curv(int lx, const double *f, const double *v1i, double dt,
                         const double *dx1, const double *dx1i, const double *ha1,
                         const double *ha2i, const double *h1i, double *out);

void orc_v12rhov1(const orc_grid *g, const double *ns, const double *vs1, double *rhovs1);
void orc_T2rhoe(const orc_grid *g, const double *ns, const double *Ts, double *rhoes);
void orc_rhov12v1(const orc_grid *g, double *ns, const double *rhovs1, double *vs1);
void orc_rhoe2T(const orc_grid *g, const double *ns, const double *rhoes, double *Ts);

void orc_set_global_boundaries_allspec(const orc_grid *g, double *ns, double *rhovs1,
                                       const double *vs1, double *vs2, double *vs3,
                                       double *rhoes, const double *vs1i, int lsp);
void orc_halo_periodic(const orc_grid *g, double *param, int nspec, int lhalo);
void orc_interface_vels_allspec(const orc_grid *g, const double *vs1, const double *vs2,
                                const double *vs3, double *vs1i, double *vs2i, double *vs3i,
                                int lsp);
/* dir = 1,2,3; frank = 0 scalar, 1 first component of a rank-1 tensor */
void orc_sweep_allspec(const orc_grid *g, int dir, double *fs, const double *vsi, double dt,
                       int frank, int lsp);

void orc_thermal_conduct(const orc_grid *g, int isp, const double *Ts, const double *ns,
                         const double *nn, const double *J1, double *lambda, double *beta);
int  orc_gbsv(int n, int kl, double *ab, double *b, int *ipiv);
/* number of row interchanges made by orc_gbsv since the last reset */
void orc_reset_swap_count(void);
long orc_swap_count(void);
void orc_TRBDF21D(int lx1, const double *Ts, const double *A, const double *B,
                  const double *C, const double *D, const double *E, double Tsminx1,
                  double Tsmaxx1, double dt, const double *dx1, const double *dx1i,
                  double *out);
void orc_backEuler1D(int lx1, const double *Ts, const double *A, const double *B,
                     const double *C, const double *D, const double *E, double Tsminx1,
                     double Tsmaxx1, double dt, const double *dx1, const double *dx1i,
                     double *out, double *coeffs, double *rhs);
int  orc_energy_diffusion(const orc_grid *g, double dt, const double *ns, double *Ts,
                          const double *J1, const double *nn, const double *Tn,
                          int diffsolvetype, double Teinf);

/* One transport step in the reference's fluid_adv order (see the .c file). */
int orc_transport_step(const orc_grid *g, double dt, double *fluidvars, double *fluidauxvars,
                       const double *electrovars, double *vs1i, double *vs2i, double *vs3i,
                       const double *nn, const double *Tn, int diffsolvetype, double Teinf);

/* The rest of the fluid_adv transport block (multifluid.f90:237-302, 553-688,
   sources_mpi.f90:55-112).  nullpts(lx1,lx2,lx3): nonzero marks a null cell (NULL: none).
   orc_clean_param returns -1 for a paramflag other than 1, 2, 3 (the reference's error
   stop).  Q is (lx1,lx2,lx3,7).  Pinned in tests/test_block.py; the dipole border copy of
   clean_param 2 (multifluid.f90:613-657) by every null mask of a 6-cell field line. */
int orc_clean_param(const orc_grid *g, int iparm, double *param, const int *nullpts);
void orc_VNRicht_artvisc(const orc_grid *g, const double *ns, const double *vs1, double *Q);
void orc_RK2_global_boundary_allspec(const orc_grid *g, double *vs1, double *vs2, double *vs3);
void orc_compression(const orc_grid *g, double dt, const double *vs1, const double *vs2,
                     const double *vs3, const double *Q, double *rhoes);
int orc_fluid_adv_block(const orc_grid *g, double dt, double *fluidvars, double *fluidauxvars,
                        const double *electrovars, double *vs1i, double *vs2i, double *vs3i,
                        const double *nn, const double *Tn, int diffsolvetype, double Teinf,
                        const int *nullpts, double *Q);

/* Time step (temporal_mpi.f90:27-165, libgemini_mpi.f90:450-483) on one process.
   rates[3*(isp-1)+k]: species isp's largest grid rate along x(k+1). */
double orc_dt_calc(const orc_grid *g, const double *Ts, const double *vs1, const double *vs2,
                   const double *vs3, double tcfl, double *rates);
void orc_dt_select(const orc_grid *g, const double *Ts, const double *vs1, const double *vs2,
                   const double *vs3, double tcfl, int flagglow, double t, double tout,
                   double tglowout, int it, double *dt);

/* Extension for parameter fitting (no counterpart in the reference): kappa(lx2*lx3)
   scales the electron thermal conductivity of each (x2,x3) column; qtop(lx2*lx3), for
   gridflag 2 only, replaces the electron Dirichlet top row by the Neumann row
   (T(lx1)-T(lx1-1))/dx1(lx1) = qtop*h1(lx1)/lambda(lx1), a downward heat flux qtop
   [W/m^2] through the top face.  NULL leaves the reference behaviour.  TRBDF2 only:
   returns -2 for diffsolvetype 1 with either array, -3 for qtop on gridflag 0/1. */
void orc_TRBDF21D_flux(int lx1, const double *Ts, const double *A, const double *B,
                       const double *C, const double *D, const double *E, double Tsminx1,
                       double gtop, double dt, const double *dx1, const double *dx1i,
                       double *out);
int orc_energy_diffusion_ext(const orc_grid *g, double dt, const double *ns, double *Ts,
                             const double *J1, const double *nn, const double *Tn,
                             int diffsolvetype, double Teinf, const double *kappa,
                             const double *qtop);
int orc_transport_step_ext(const orc_grid *g, double dt, double *fluidvars,
                           double *fluidauxvars, const double *electrovars, double *vs1i,
                           double *vs2i, double *vs3i, const double *nn, const double *Tn,
                           int diffsolvetype, double Teinf, const double *kappa,
                           const double *qtop);

/* Source/loss step (multifluid.f90:130-175, 472-549, sources.f90, collisions.f90:58-159,
   calculus.in.f90:174-196).  nn(lx1,lx2,lx3,6): O, N2, O2, H, N, NO; Tn, vn1-3, Qeionize,
   g1 (lx1,lx2,lx3); Prionize (lx1,lx2,lx3,6); alt ghosted.  Pr, Lo (lx1,lx2,lx3,7). */
typedef struct {
  const double *nn, *Tn, *vn1, *vn2, *vn3, *Prionize, *Qeionize, *alt, *g1;
} orc_atmos;
double orc_ETD_uncoupled(double f, double P, double L, double dt);
double orc_maxwell_colln(int isp, int isp2, double nn2, double Tn, double Tsp);
void orc_coulomb_colln(int isp, int isp2, double ns2, double Ts1, double Ts2, double v1,
                       double v2, double *out);
void orc_grad3D1_interior(const orc_grid *g, const double *f, double *out);
void orc_srcs_energy(const orc_grid *g, const orc_atmos *a, const double *ns, const double *vs1,
                     const double *vs2, const double *vs3, const double *Ts, double *Pr,
                     double *Lo);
void orc_srcs_momentum(const orc_grid *g, const orc_atmos *a, const double *ns, const double *vs1,
                       const double *vs2, const double *vs3, const double *Ts, const double *E1,
                       const double *Q, double *Pr, double *Lo);
void orc_srcs_continuity(const orc_grid *g, const orc_atmos *a, const double *ns,
                         const double *vs1, const double *vs2, const double *vs3,
                         const double *Ts, double *Pr, double *Lo);
void orc_source_loss_energy(const orc_grid *g, const orc_atmos *a, double dt, const double *ns,
                            const double *vs1, const double *vs2, const double *vs3,
                            double *Ts, double *rhoes);
void orc_source_loss_momentum(const orc_grid *g, const orc_atmos *a, double dt, const double *ns,
                              double *vs1, const double *vs2, const double *vs3, const double *Ts,
                              const double *E1, const double *Q, double *rhovs1);
void orc_source_loss_mass(const orc_grid *g, const orc_atmos *a, double dt, double *ns,
                          const double *vs1, const double *vs2, const double *vs3,
                          const double *Ts);
int orc_sources_block(const orc_grid *g, const orc_atmos *a, double dt, double *fluidvars,
                      double *fluidauxvars, const double *electrovars, const double *vs1i,
                      const double *Q, const int *nullpts);

/* plasma output (io_nompi.f90:183-258, timeutils.f90:151-196): the file stem, the
   number of real32 values of the field datasets and their values in dataset order */
int orc_filestem(const int ymd[3], double UTsec, char out[22]);
int orc_dateinc(double dtsec, int ymd[3], double *UTsec);    /* timeutils.f90:79-107 */
long orc_output_count(int lx1, int lx2, int lx3, int flagoutput);
void orc_output_pack(int lx1, int lx2, int lx3, int gridflag, int flagoutput, int patch,
                     const double *fluidvars, const double *electrovars, float *out);

/* 0: x**2.5 by libm pow, as the reference (default); 1: correctly rounded x**2.5 */
void orc_set_pow_mode(int mode);
void orc_eval_pow25(const double *x, double *y, long n, int mode);
/* 0: Netlib reference BLAS arithmetic in the band solve (default); 1: fused multiply-add
   updates as optimised BLAS libraries evaluate them */
void orc_set_blas_fma(int mode);

/* threads used by the OpenMP loops (1 when built without OpenMP) */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
