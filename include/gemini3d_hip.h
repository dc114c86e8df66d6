/*
 * gemini3d_hip.h -- C ABI of the MI355X (gfx950) multifluid-transport library.
 *
 * Part 1 re-exports, with identical names and signatures, the hot-path entry
 * points of the reference's public C API (include/gemini3d.h), so the
 * reference's C++ driver fluid_adv (src/libgemini.cpp:143-234) or the Fortran
 * main (app/main.f90:258-383, through gemini_amd/fortran/gemhip_bind.f90)
 * compiles against this library unchanged.  As in the reference every
 * argument is passed by reference and sizes come from library state set once
 * (the reference keeps them in module globals, grid.f90:12; here
 * gemhip_set_grid_size).  The opaque void** arguments (cfgC, xC, intvars) are
 * Fortran objects C cannot read; this library takes their contents once
 * through the gemhip_register_* calls of part 2 and ignores the pointers,
 * except that *xtype must equal the registered mesh type.
 *
 * Errors: where the reference executes `error stop` (advec.f90:145,225,241,
 * 258,275; multifluid.f90:334-336; libgemini_c.f90:75-78) this library prints
 * the message and calls abort() by default; after gemhip_set_error_mode(1) it
 * records the message (gemhip_last_error) and returns instead.
 *
 * Data residency: in mode 0 (default) each part-1 call copies the host blocks
 * it reads to HBM, runs, copies back what it writes and synchronises, so the
 * caller's host arrays are current after every call, as in the reference.  In
 * mode 1 the HBM copy is authoritative between gemhip_state_push and
 * gemhip_state_pull and the calls only enqueue kernels on the library stream.
 */
#ifndef GEMINI3D_HIP_H
#define GEMINI3D_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- part 1: reference gemini3d.h hot-path entry points ---------------- */

/* gemini3d.h:64, libgemini_c.f90:445-455 -> multifluid.f90:206-214 */
void v12rhov1_C(double** fluidvars, double** fluidauxvars);
/* gemini3d.h:65, libgemini_c.f90:459-469 -> multifluid.f90:193-202 */
void T2rhoe_C(double** fluidvars, double** fluidauxvars);
/* gemini3d.h:83, libgemini_c.f90:797-807 -> multifluid.f90:179-189 */
void rhoe2T_C(double** fluidvars, double** fluidauxvars);
/* gemini3d.h:79, libgemini_c.f90:748-758 -> multifluid.f90:218-233 */
void rhov12v1_C(double** fluidvars, double** fluidauxvars);
/* gemini3d.h:66, libgemini_c.f90:473-488 -> advec.f90:155-229 */
void interface_vels_allspec_C(int* xtype, void** xC, double** fluidvars, void** intvars, int* lsp);
/* gemini3d.h:101, libgemini_c.f90:491-509 -> advec.f90:21-149 */
void set_global_boundaries_allspec_C(int* xtype, void** xC, double** fluidvars,
                                     double** fluidauxvars, void** intvars, int* lsp);
/* gemini3d.h:67-78, libgemini_c.f90:513-744 -> multifluid.f90:49-126 -> advec.f90:233-381 */
void sweep3_allparams_C(double** fluidvars, double** fluidauxvars, void** intvars, int* xtype,
                        void** xC, double* dt);
void sweep3_allspec_mass_C(double** fluidvars, double** fluidauxvars, void** intvars,
                           int* xtype, void** xC, double* dt);
void sweep3_allspec_momentum_C(double** fluidvars, double** fluidauxvars, void** intvars,
                               int* xtype, void** xC, double* dt);
void sweep3_allspec_energy_C(double** fluidvars, double** fluidauxvars, void** intvars,
                             int* xtype, void** xC, double* dt);
void sweep1_allparams_C(double** fluidvars, double** fluidauxvars, void** intvars, int* xtype,
                        void** xC, double* dt);
void sweep1_allspec_mass_C(double** fluidvars, double** fluidauxvars, void** intvars,
                           int* xtype, void** xC, double* dt);
void sweep1_allspec_momentum_C(double** fluidvars, double** fluidauxvars, void** intvars,
                               int* xtype, void** xC, double* dt);
void sweep1_allspec_energy_C(double** fluidvars, double** fluidauxvars, void** intvars,
                             int* xtype, void** xC, double* dt);
void sweep2_allparams_C(double** fluidvars, double** fluidauxvars, void** intvars, int* xtype,
                        void** xC, double* dt);
void sweep2_allspec_mass_C(double** fluidvars, double** fluidauxvars, void** intvars,
                           int* xtype, void** xC, double* dt);
void sweep2_allspec_momentum_C(double** fluidvars, double** fluidauxvars, void** intvars,
                               int* xtype, void** xC, double* dt);
void sweep2_allspec_energy_C(double** fluidvars, double** fluidauxvars, void** intvars,
                             int* xtype, void** xC, double* dt);
/* gemini3d.h:84, libgemini_c.f90:827-848 -> multifluid.f90:309-342 */
void energy_diffusion_C(void** cfgC, int* xtype, void** xC, double** fluidvars,
                        double** electrovars, void** intvars, double* dt);
/* gemini3d.h:135-137, libgemini_mpi_c.f90:414-458 -> multifluid_mpi.f90:12-41 and
   advec_mpi.f90:33-44.
   One process owns the whole grid: the exchange reduces to the periodic x3 wrap of
   mpihalo.f90:147-152 when the grid is periodic, and to nothing otherwise. */
void halo_interface_vels_allspec_C(int* xtype, void** xC, double** fluidvars, int* lsp);
void halo_allparams_C(int* xtype, void** xC, double** fluidvars, double** fluidauxvars);
void halo_fluidvars_C(int* xtype, void** xC, double** fluidvars, double** fluidauxvars);
/* The rest of fluid_adv's transport block (src/libgemini.cpp:187-210). */
/* gemini3d.h:80, libgemini_c.f90:811-823 -> libgemini.f90:1037-1056 -> multifluid.f90:553-688:
   iparm 1 densities, 2 vs1, 3 temperatures; null cells are those registered with
   gemhip_register_null_cells (the reference's x%nullpts, meshobj.f90:586-619). */
void clean_param_C(int* iparm, int* xtype, void** xC, double** fluidvars);
/* gemini3d.h:81, libgemini_c.f90:762-771 -> multifluid.f90:237-268; the library holds
   intvars%Q, which compression_C reads. */
void VNRicht_artvisc_C(double** fluidvars, void** intvars);
/* gemini3d.h:82, libgemini_c.f90:775-793 -> multifluid.f90:274-302 with div3D_curv_23
   (div.f90:118-206). */
void compression_C(double** fluidvars, double** fluidauxvars, void** intvars, int* xtype,
                   void** xC, double* dt);
/* gemini3d.h:138-139, libgemini_mpi_c.f90:461-490 -> sources_mpi.f90:14-112.  One process:
   the depth-1 periodic x3 wrap of vs1, vs2, vs3, and the zero-order-hold boundaries. */
void RK2_prep_mpi_allspec_C(int* xtype, void** xC, double** fluidvars);
void RK2_global_boundary_allspec_C(int* xtype, void** xC, double** fluidvars);
/* gemini3d.h:140, libgemini_mpi_c.f90:289-308 -> libgemini_mpi.f90:450-483 and
   temporal_mpi.f90:27-165 on one process: the CFL-limited dt from Ts, vs1, vs2, vs3 (an
   on-device max reduction), clamped to the next output time, floored at 1e-6 and, from
   the second call on, limited to twice the previous dt (*dt in and out).  cfg%tcfl and
   cfg%flagglow come from gemhip_register_cfl. */
void dt_select_C(void** cfgC, int* xtype, void** xC, double** fluidvars, double** fluidauxvars,
                 double* t, double* tout, double* tglowout, double* dt);

/* The sources substep (src/libgemini.cpp:216-227).  gemini3d.h:85-89, libgemini_c.f90:852-948
   -> libgemini.f90:1120-1195 -> multifluid.f90:130-175, 472-549, sources.f90:18-374,
   collisions.f90:58-159, calculus.in.f90:174-196.  The library holds what the reference
   keeps in intvars and the mesh: the neutral atmosphere (gemhip_register_atmosphere), the
   ionization and heating rates Prionize/Qeionize that impact_ionization_C and
   solar_ionization_C accumulate (gemhip_set_ionization_rates), the altitude and gravity
   (gemhip_register_gravity) and Q from VNRicht_artvisc_C; the momentum sources read
   electrovars field 1 (E1).  Energy updates rhoes and Ts, momentum rhovs1 and vs1, mass ns;
   allparams runs the three in that order (libgemini.f90:1120-1127). */
void source_loss_allparams_C(double** fluidvars, double** fluidauxvars, double** electrovars,
                             void** intvars, int* xtype, void** xC, double* dt);
void source_loss_mass_C(double** fluidvars, double** fluidauxvars, double** electrovars,
                        void** intvars, int* xtype, void** xC, double* dt);
void source_loss_momentum_C(double** fluidvars, double** fluidauxvars, double** electrovars,
                            void** intvars, int* xtype, void** xC, double* dt);
void source_loss_energy_C(double** fluidvars, double** fluidauxvars, double** electrovars,
                          void** intvars, int* xtype, void** xC, double* dt);
/* gemini3d.h:90, libgemini.f90:1202-1207: Prionize and Qeionize to zero. */
void clear_ionization_arrays_C(void** intvars);

/* gemini3d.h:95, libgemini_c.f90:1144-1150 -> timeutils.f90:79-107: advance ymd and UTsec
   by dt (0 <= dt <= 86400), carrying whole days into the date by the Gregorian rules. */
void dateinc_C(double* dt, int* ymd, double* UTsec);

/* gemini3d.h:96, libgemini_c.f90:267-284 -> io_nompi.f90:183-268: write
   <outdir>/<YYYYMMDD_SSSSS.ffffff>_<identifier>.h5 (timeutils.f90:133-196) holding
   /flagoutput, /time/ymd, /time/UThour, /x1lims../x3lims, /patchID and the real32 field
   datasets of cfg%flagoutput (outdir, flagoutput and the deflate level come from
   gemhip_register_output): 2 -> neall v1avgall Tavgall Teall J1all J2all J3all v2avgall
   v3avgall, 3 -> neall, otherwise nsall vs1all Tsall J1all J2all J3all v2avgall v3avgall;
   then Phiall, the (lx2,lx3) potential at x1 = 1 (gridflag 1) or x1 = lx1.  A kernel forms
   the datasets from the HBM state (from fluidvars in mode 0; in mode 1 fluidvars may be
   null), J1-J3 and Phi come from electrovars (host or device address). */
void plasma_output_nompi_C(void** cfgC, int* ymd, double* UTsec, double** fluidvars,
                           double** electrovars, int* identifier, double* x1lims,
                           double* x2lims, double* x3lims);

/* ---------------- part 2: library set-up (no counterpart in the reference ABI) ---------------- */
/* Array arguments of part 2 may be host or HBM addresses: copies use hipMemcpyDefault. */

/* Select the HIP device and create the library stream.  Returns 0 or an error code. */
int gemhip_init(int device);
/* Release every HBM buffer and the stream. */
void gemhip_finalize(void);
/* 0: abort() on error like the reference's `error stop`; 1: record and return. */
void gemhip_set_error_mode(int mode);
/* Message of the last error, or "" (reset by gemhip_clear_error). */
const char* gemhip_last_error(void);
void gemhip_clear_error(void);

/* Subgrid sizes, the counterpart of the grid-module globals lx1,lx2,lx3 (grid.f90:12). */
int gemhip_set_grid_size(int lx1, int lx2, int lx3);
/* Mesh contents the reference keeps in its curvmesh object (meshobj.f90:29-208):
   xtype (1 Cartesian, 2 dipole, libgemini_c.f90:69-74), gridflag, flagper, the six
   isglobal flags (grid.f90:448-507) and host pointers to h1,h2,h3 (lx1+4,lx2+4,lx3+4),
   hNx1i (lx1+1,lx2,lx3), hNx2i (lx1,lx2+1,lx3), hNx3i (lx1,lx2,lx3+1), dxN (lxN+3),
   dxNi (lxN).  The arrays are copied to HBM. */
int gemhip_register_grid(int xtype, int gridflag, int flagper, const int* isglobal6,
                         const double* h1, const double* h2, const double* h3,
                         const double* h1x1i, const double* h2x1i, const double* h3x1i,
                         const double* h1x2i, const double* h2x2i, const double* h3x2i,
                         const double* h1x3i, const double* h2x3i, const double* h3x3i,
                         const double* dx1, const double* dx1i, const double* dx2,
                         const double* dx2i, const double* dx3, const double* dx3i);
/* intvars%atmos%nn(lx1,lx2,lx3,1:3) (slices lx1*lx2*lx3 apart) and Tn(lx1,lx2,lx3). */
int gemhip_register_neutrals(const double* nn, const double* Tn);
/* cfg%diffsolvetype (1 backward Euler, 2 TRBDF2) and cfg%Teinf (config.f90:88). */
int gemhip_register_config(int diffsolvetype, double Teinf);

/* Residency mode (see above). */
int gemhip_set_residency(int mode);
/* Copy host blocks to HBM: fluidvars (...,35), the first 14 fields of fluidauxvars
   (rhovs1, rhoes) and field 4 of electrovars (J1).  Either auxiliary pointer may be
   NULL to leave that part untouched. */
int gemhip_state_push(const double* fluidvars, const double* fluidauxvars,
                      const double* electrovars);
/* Copy HBM state back: fluidvars (...,35) and the first 14 fields of fluidauxvars. */
int gemhip_state_pull(double* fluidvars, double* fluidauxvars);
/* Interface velocities held by the library (the reference's intvars%vs1i/vs2i/vs3i).
   Arrays are (lx1+1,lx2,lx3,7), (lx1,lx2+1,lx3,7), (lx1,lx2,lx3+1,7). */
int gemhip_get_interface_vels(double* vs1i, double* vs2i, double* vs3i);
int gemhip_set_interface_vels(const double* vs1i, const double* vs2i, const double* vs3i);

/* One transport step on the HBM state, in the order of src/libgemini.cpp:155-186 and
   :208: v12rhov1, T2rhoe, set_global_boundaries_allspec, halo_fluidvars,
   interface_vels_allspec, sweep3/sweep1/halo_allparams/sweep2 of mass, momentum and
   energy, rhov12v1, rhoe2T, energy_diffusion.  Kernels are fused where the result is
   bit-identical (interface velocities are formed inside the sweeps).  Requires
   residency mode 1.  Returns 0 or an error code. */
int gemhip_transport_step(double dt);

/* cfg%tcfl (config.f90:15; 0.9 until registered) and cfg%flagglow for dt_select_C; also
   restarts its iteration count. */
int gemhip_register_cfl(double tcfl, int flagglow);
/* dt_calc's grid rates of the current state: rates[3*(isp-1)+k] is species isp's largest
   grid rate along x(k+1) (host or HBM array of 21 doubles). */
int gemhip_grid_rates(double* rates);

/* Null cells (the reference's x%nullpts, computed by meshobj.f90:586-619 as alt < 80 km):
   lx1*lx2*lx3 ints, nonzero = null, host or HBM; NULL: none (the default). */
int gemhip_register_null_cells(const int* nullpts);

/* The whole transport block of src/libgemini.cpp:155-210 on the HBM state (mode 1): the
   advection of gemhip_transport_step, then rhov12v1, clean_param 1 and 2,
   VNRicht_artvisc, halo_fluidvars, RK2_global_boundary_allspec, compression, rhoe2T,
   clean_param 3, energy_diffusion, clean_param 3 and T2rhoe.  Returns 0 or an error code. */
int gemhip_fluid_adv_block(double dt);

/* Inputs of the source/loss substep.  intvars%atmos (neutral.f90): nn (lx1,lx2,lx3,6) with
   O, N2, O2, H, N, NO; Tn, vn1, vn2, vn3 (lx1,lx2,lx3).  Also replaces the neutrals of
   gemhip_register_neutrals (the first three densities and Tn), so diffusion and sources
   see one atmosphere. */
int gemhip_register_atmosphere(const double* nn, const double* Tn, const double* vn1,
                               const double* vn2, const double* vn3);
/* x%alt (lx1+4,lx2+4,lx3+4, ghosted) and x%g1 (lx1,lx2,lx3), meshobj.f90:135-143. */
int gemhip_register_gravity(const double* alt, const double* g1);
/* intvars%Prionize (lx1,lx2,lx3,6) and intvars%Qeionize (lx1,lx2,lx3): the totals the
   reference's ionization calls leave before the sources substep (libgemini.cpp:216-219). */
int gemhip_set_ionization_rates(const double* Prionize, const double* Qeionize);
/* The sources substep of src/libgemini.cpp:223-232 on the HBM state (mode 1), after
   gemhip_fluid_adv_block and the ionization rates: set_global_boundaries_allspec,
   source_loss_energy, _momentum, _mass, then clean_param 3, 2 and 1. */
int gemhip_sources_block(double dt);

/* HBM pointers of the state blocks (mode 1), for zero-copy callers. */
int gemhip_device_state(double** fluidvars, double** rhovs1_rhoes, double** J1);
/* The library's HIP stream (a hipStream_t) for event timing and ordering. */
void* gemhip_stream(void);
int gemhip_synchronize(void);

/* Per-kernel timing with HIP events on the library stream: enable, then read the
   accumulated milliseconds and launch counts by kernel name. */
int gemhip_kernel_timing(int enable);
int gemhip_kernel_time(const char* name, double* total_ms, long* launches);
int gemhip_kernel_timing_reset(void);

/* cfg%outdir and cfg%flagoutput (config.f90) for plasma_output_nompi_C, and the deflate
   level of its field datasets (0 stores them contiguous; the reference's phys_consts
   comp_lvl is 3).  The HDF5 C library is loaded when a file is first written or read:
   $GEMHIP_HDF5_LIB, else libhdf5.so on the loader path, else /opt/conda/lib/libhdf5.so. */
int gemhip_register_output(const char* outdir, int flagoutput, int comp_lvl);
/* The path plasma_output_nompi_C writes for a date, time and patch (host-only; no
   context needed).  Returns 0, or nonzero (message in gemhip_last_error) where
   timeutils.f90:182 stops or out is shorter than the path. */
int gemhip_output_filename(const char* outdir, const int* ymd, double UTsec, int identifier,
                           char* out, long len);
/* output_plasma (plasma_output.f90:23-31, 76-136 -> plasma_output_hdf5.f90:10-65), the
   file gemini.bin writes at every output time: <outdir>/<YYYYMMDD_SSSSS.ffffff>.h5 with
   /flagoutput, /time/ymd, /time/UThour and the datasets of plasma_output_nompi_C, except
   that flagoutput 2 names the electron temperature TEall.  One process holds the whole
   grid, so the root's gather is the identity.  Arguments as plasma_output_nompi_C's. */
int gemhip_output_plasma(const int* ymd, double UTsec, double** fluidvars, double** electrovars);
/* The name gemhip_output_plasma writes (host-only). */
int gemhip_output_filename_root(const char* outdir, const int* ymd, double UTsec, char* out,
                                long len);
/* get_simsize3 (reader_hdf5.f90:64-92): the full-grid sizes in a simsize file (/lx1 /lx2
   /lx3, else /lxs, else /lx).  Host-only; no context needed. */
int gemhip_read_simsize(const char* path, int* lx1, int* lx2, int* lx3);
/* get_grid3_coords_hdf5 (readgrid_hdf5.f90:11-39): the ghosted coordinates x1(lx1+4),
   x2(lx2+4), x3(lx3+4) and the grid centre /glonctr /glatctr (required, as there) of path,
   or of path/simgrid.h5 when path does not name a simgrid.h5 file.  Host-only. */
int gemhip_read_grid_coords(const char* path, int lx1, int lx2, int lx3, double* x1, double* x2,
                            double* x3, double* glonctr, double* glatctr);
/* getICs_hdf5 (plasma_input_hdf5.f90:93-144; io_nompi.f90:271-345) into the caller's host
   blocks, with the zeroing of plasma_input_hdf5.f90:78-86 (host-only; no context needed):
   the sizes in indatsize (/lx1 /lx2 /lx3, else /lxs, else /lx; reader_hdf5.f90:64-92) must
   be lx1, lx2, lx3; ns, vs1 and Ts of fluidvars are zeroed and their interiors read from
   /nsall /vs1all /Tsall (shape (lx1,lx2,lx3,7), any real type) of indatfile; if
   electrovars is given its potential field is zeroed and every interior x1 plane set to
   /Phiall ((lx2,lx3), or 1-D when lx2 or lx3 is 1; zero when absent).  With the state
   resident in HBM, gemhip_state_push afterwards. */
int gemhip_read_initial_conditions(const char* indatsize, const char* indatfile, int lx1,
                                   int lx2, int lx3, double* fluidvars, double* electrovars);

/* Evaluate on the device the scalar routines the kernels use, for checks against the
   host libm: op 0 x**2.5 as the conductivity uses it, 1 sqrt, 2 reciprocal 1/x,
   3 the device library pow(x, 2.5).  x and y are host arrays of n doubles; for op 4,
   the x1 sweep's division a/b through the reciprocal 1/b, x holds n pairs (a, b). */
int gemhip_eval_math(int op, const double* x, double* y, long n);

/* ---------------- part 3: contexts, tuning and ensembles (no counterpart in the reference) ---------------- */
/* The calls above act on the selected library context.  Handle 0 is the implicit context
   the reference API uses; gemhip_context_create makes another (then gemhip_init it), for
   ensembles of independent grids.  Selection is process-wide and not thread-safe. */
int gemhip_context_create(void);
int gemhip_context_select(int handle);
int gemhip_context_current(void);
/* Destroying a context that a captured ensemble graph uses is refused (returns 2) until
   the graph is released. */
int gemhip_context_destroy(int handle);

/* Kernel variants of the selected context: "sweep_pf" (1 or 2, x2/x3 sweep pipeline
   depth) and "diff_chunk" (4 or 8, TRBDF2 rows per chunk).  Results do not depend on
   them; only speed does. */
int gemhip_set_tuning(const char* knob, int value);
int gemhip_get_tuning(const char* knob);

/* Capture nsteps transport steps (gemhip_transport_step with dt[m]) of each of the n
   contexts handles[m] (residency mode 1, state pushed) into one HIP graph whose members
   run concurrently, each on its own context's stream.  Returns a graph id > 0, or -1. */
int gemhip_ensemble_capture(int n, const int* handles, const double* dt, int nsteps);
/* Launch a captured graph; with wait != 0, return after it has finished. */
int gemhip_ensemble_launch(int id, int wait);
int gemhip_ensemble_release(int id);

/* ---------------- part 4: parameter fitting and the discrete adjoint (no counterpart in the reference) ---------------- */
/* Extension of the electron energy diffusion for fitting: a per-column scale kappa on the
   electron thermal conductivity (collisions.f90:185-195 times kappa(i2,i3)) and, for
   gridflag 2, a downward electron heat flux q [W/m^2] through the top face, imposed as the
   Neumann row (T(lx1)-T(lx1-1))/dx1(lx1) = q h1(lx1)/lambda(lx1) that PDEparabolic.f90:321-326
   carries commented out, in both TRBDF2 stages.  Arrays are lx2*lx3 values, x2 fastest,
   host or HBM; NULL removes the parameter (the reference's behaviour).  TRBDF2 only. */
int gemhip_register_conductivity_scale(const double* kappa);
int gemhip_register_top_heat_flux(const double* qtop);

/* Restart checkpoint of the resident state: fluidvars (35 fields) and the x1 boundary
   interface velocities the next step's ghost fill reads.  Size in doubles; dst/src host
   or HBM. */
long gemhip_checkpoint_size(void);
int gemhip_checkpoint_save(double* dst);
int gemhip_checkpoint_load(const double* src);

/* Discrete adjoint of the electron energy path of one transport step (the only part of
   the state the fitting parameters reach).  gemhip_adjoint_record_step runs
   gemhip_transport_step and writes, at the HBM address tape (gemhip_adjoint_tape_size()
   doubles), the electron fields the adjoint needs.  gemhip_adjoint_step takes lam (HBM,
   one ghosted field) = dL/dTe after that step, overwrites it with dL/dTe before the step,
   and adds dL/dkappa and dL/dq per column into gkappa and gq (HBM, lx2*lx3 each; either
   may be NULL).  Limiter, upwind and floor branches are those of the recorded step.
   Needs diffsolvetype 2, and lx3 >= 2 on periodic grids.  With a conductivity scale or
   heat flux registered, the tape ends with the row interchanges of the electron
   columns' TR and BDF2 band factorisations (one bit per elimination step, 32 to an
   int, [column][stage][word]), so that two recorded runs can be compared branch by
   branch. */
long gemhip_adjoint_tape_size(void);
int gemhip_adjoint_record_step(double dt, double* tape);
int gemhip_adjoint_step(double dt, const double* tape, double* lam, double* gkappa, double* gq);

/* ---------------- part 5: allocation-free entry point (no counterpart in the reference) ---------------- */
/* For callers that cannot allocate device memory (code under HIP graph capture, tight
   serving loops, hosts that forbid hipMalloc): energy_diffusion_C's computation on HBM
   fields the caller owns -- Ts and ns the 7 ghosted species fields of
   fluidvars(:,:,:,29:35) and (:,:,:,1:7), J1 one ghosted field -- with the grid,
   neutrals and configuration of the selected context.  The TRBDF2 work space is a static
   device array of 256 MiB (one per GPU, held until the process ends); grids that need
   more (7 * ceil(lx2*lx3/64) * (4608*lx1 + 256*ceil(lx1/32) + 256) bytes, about 1000
   columns at lx1 = 512) are rejected with an error.  Calls no allocator; kernel timing
   must be off.  Enqueued on the context's stream (gemhip_synchronize waits).  Calls that
   overlap on one GPU share the work space and race.  Returns 0 or an error code. */
int gemhip_energy_diffusion_noalloc(double* Ts, const double* ns, const double* J1, double dt);

/* ---------------- part 6: batched ensembles (extension, no counterpart in the reference) ---------------- */
/* The transport step of many members (contexts in residency mode 1, state pushed) run by
   batched kernels: each kernel of the step is one persistent launch walking a table of
   every member's blocks.  Members must have global sides, a non-periodic x3, lx2 and
   lx3 above 1, TRBDF2 diffusion without the fitting extension, and one device.  Results
   are bit-identical to gemhip_transport_step on each member.  Steps run on the batch's
   stream: synchronize it (or launch graphs with wait = 1) before touching a member
   through its context, and do not reallocate a member's state (residency, grid) while
   the batch exists; members of a batch cannot be destroyed.  The time steps live in HBM,
   so gemhip_batch_set_dt also changes what captured graphs do.  Tuning knobs:
   "sweep_pf" (1, 2), "diff_chunk" (4, 8), "grid" (1, 2: persistent blocks per CU times
   this); they change speed only.  gemhip_batch_capture returns a graph id for
   gemhip_ensemble_launch (launched on the batch's stream), released with the batch at
   the latest. */
int gemhip_batch_create(int n, const int* handles);      /* batch id >= 1, or < 0 */
int gemhip_batch_set_dt(int batch, const double* dt);    /* one per member */
int gemhip_batch_set_tuning(int batch, const char* knob, int value);
int gemhip_batch_step(int batch);
int gemhip_batch_synchronize(int batch);
int gemhip_batch_capture(int batch, int nsteps);         /* graph id >= 1, or < 0 */
int gemhip_batch_release(int batch);

#ifdef __cplusplus
}
#endif
#endif
