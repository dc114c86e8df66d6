// A C++ host driving the transport hot path through include/gemini3d_hip.h in the order of
// the reference's fluid_adv (src/libgemini.cpp:155-208): the reference-named entry points
// take the address of each state-block pointer, sizes come from library state.  Test
// program: reads a case written by the test suite (case.txt, in.bin), runs nsteps steps
// and writes out.bin.  mode 0 calls the entry points one by one on host blocks; mode 1
// pushes the state once and runs gemhip_transport_step on HBM; modes 2 and 3 do the same
// for the whole transport block (src/libgemini.cpp:155-210, gemhip_fluid_adv_block);
// modes 4 and 5 follow the block with the sources substep (:223-232), call by call or
// fused (gemhip_sources_block).
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "gemini3d_hip.h"

namespace {

std::vector<double> take(FILE* f, size_t n) {
  std::vector<double> v(n);
  if (fread(v.data(), sizeof(double), n, f) != n) {
    fprintf(stderr, "capi_driver: short read\n");
    exit(2);
  }
  return v;
}

void check(int rc, const char* what) {
  if (rc != 0) {
    fprintf(stderr, "capi_driver: %s failed (%d): %s\n", what, rc, gemhip_last_error());
    exit(3);
  }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) { fprintf(stderr, "usage: capi_driver <case dir>\n"); return 1; }
  const std::string dir = argv[1];
  int lx1, lx2, lx3, xtype, gridflag, flagper, isg[6], solver, nsteps, mode;
  double dt, Teinf;
  FILE* c = fopen((dir + "/case.txt").c_str(), "r");
  if (!c || fscanf(c, "%d %d %d %d %d %d %d %d %d %d %d %d %d %d %d %lf %lf", &lx1, &lx2, &lx3,
                   &xtype, &gridflag, &flagper, &isg[0], &isg[1], &isg[2], &isg[3], &isg[4],
                   &isg[5], &solver, &nsteps, &mode, &dt, &Teinf) != 17) {
    fprintf(stderr, "capi_driver: bad case.txt\n");
    return 1;
  }
  fclose(c);
  const size_t L = (size_t)(lx1 + 4) * (lx2 + 4) * (lx3 + 4);
  const size_t n1 = (size_t)(lx1 + 1) * lx2 * lx3, n2 = (size_t)lx1 * (lx2 + 1) * lx3,
               n3 = (size_t)lx1 * lx2 * (lx3 + 1), ni = (size_t)lx1 * lx2 * lx3;
  FILE* f = fopen((dir + "/in.bin").c_str(), "rb");
  if (!f) { fprintf(stderr, "capi_driver: no in.bin\n"); return 1; }
  auto h1 = take(f, L), h2 = take(f, L), h3 = take(f, L);
  auto h1x1i = take(f, n1), h2x1i = take(f, n1), h3x1i = take(f, n1);
  auto h1x2i = take(f, n2), h2x2i = take(f, n2), h3x2i = take(f, n2);
  auto h1x3i = take(f, n3), h2x3i = take(f, n3), h3x3i = take(f, n3);
  auto dx1 = take(f, lx1 + 3), dx1i = take(f, lx1), dx2 = take(f, lx2 + 3), dx2i = take(f, lx2);
  auto dx3 = take(f, lx3 + 3), dx3i = take(f, lx3);
  auto nn = take(f, 3 * ni), Tn = take(f, ni);
  auto fluid = take(f, 35 * L), aux = take(f, 23 * L), electro = take(f, 7 * L);
  auto vs1i = take(f, 7 * n1), vs2i = take(f, 7 * n2), vs3i = take(f, 7 * n3);
  const bool sources = mode >= 4;
  std::vector<double> nn6, vn1, vn2, vn3, Prion, Qeion, alt, g1;
  if (sources) {
    nn6 = take(f, 6 * ni); vn1 = take(f, ni); vn2 = take(f, ni); vn3 = take(f, ni);
    Prion = take(f, 6 * ni); Qeion = take(f, ni); alt = take(f, L); g1 = take(f, ni);
  }
  fclose(f);

  check(gemhip_init(0), "gemhip_init");
  gemhip_set_error_mode(1);
  check(gemhip_set_grid_size(lx1, lx2, lx3), "gemhip_set_grid_size");
  check(gemhip_register_grid(xtype, gridflag, flagper, isg, h1.data(), h2.data(), h3.data(),
                             h1x1i.data(), h2x1i.data(), h3x1i.data(), h1x2i.data(),
                             h2x2i.data(), h3x2i.data(), h1x3i.data(), h2x3i.data(),
                             h3x3i.data(), dx1.data(), dx1i.data(), dx2.data(), dx2i.data(),
                             dx3.data(), dx3i.data()),
        "gemhip_register_grid");
  check(gemhip_register_neutrals(nn.data(), Tn.data()), "gemhip_register_neutrals");
  check(gemhip_register_config(solver, Teinf), "gemhip_register_config");
  check(gemhip_set_interface_vels(vs1i.data(), vs2i.data(), vs3i.data()), "set_interface_vels");
  if (sources) {
    check(gemhip_register_atmosphere(nn6.data(), Tn.data(), vn1.data(), vn2.data(), vn3.data()),
          "gemhip_register_atmosphere");
    check(gemhip_register_gravity(alt.data(), g1.data()), "gemhip_register_gravity");
    check(gemhip_set_ionization_rates(Prion.data(), Qeion.data()), "set_ionization_rates");
  }

  double* fluidvars = fluid.data();
  double* fluidauxvars = aux.data();
  double* electrovars = electro.data();
  void* xC = nullptr;       // the reference's curvmesh, cfg and intvars objects: their
  void* intvars = nullptr;  // contents were registered above
  void* cfgC = nullptr;
  int lsp = 7;
  int one = 1, two = 2, three = 3;
  if (mode == 1 || mode == 3 || mode == 5) {
    check(gemhip_set_residency(1), "set_residency");
    check(gemhip_state_push(fluidvars, fluidauxvars, electrovars), "state_push");
    for (int k = 0; k < nsteps; ++k) {
      check(mode == 1 ? gemhip_transport_step(dt) : gemhip_fluid_adv_block(dt), "fused step");
      if (mode == 5) check(gemhip_sources_block(dt), "sources block");
    }
    check(gemhip_state_pull(fluidvars, fluidauxvars), "state_pull");
  } else if (mode == 2 || mode == 4) {
    for (int k = 0; k < nsteps; ++k) {
      // src/libgemini.cpp:155-210, the whole transport block
      v12rhov1_C(&fluidvars, &fluidauxvars);
      T2rhoe_C(&fluidvars, &fluidauxvars);
      set_global_boundaries_allspec_C(&xtype, &xC, &fluidvars, &fluidauxvars, &intvars, &lsp);
      halo_fluidvars_C(&xtype, &xC, &fluidvars, &fluidauxvars);
      interface_vels_allspec_C(&xtype, &xC, &fluidvars, &intvars, &lsp);
      sweep3_allparams_C(&fluidvars, &fluidauxvars, &intvars, &xtype, &xC, &dt);
      sweep1_allparams_C(&fluidvars, &fluidauxvars, &intvars, &xtype, &xC, &dt);
      halo_allparams_C(&xtype, &xC, &fluidvars, &fluidauxvars);
      sweep2_allparams_C(&fluidvars, &fluidauxvars, &intvars, &xtype, &xC, &dt);
      rhov12v1_C(&fluidvars, &fluidauxvars);
      clean_param_C(&one, &xtype, &xC, &fluidvars);
      clean_param_C(&two, &xtype, &xC, &fluidvars);
      VNRicht_artvisc_C(&fluidvars, &intvars);
      halo_fluidvars_C(&xtype, &xC, &fluidvars, &fluidauxvars);
      RK2_global_boundary_allspec_C(&xtype, &xC, &fluidvars);
      compression_C(&fluidvars, &fluidauxvars, &intvars, &xtype, &xC, &dt);
      rhoe2T_C(&fluidvars, &fluidauxvars);
      clean_param_C(&three, &xtype, &xC, &fluidvars);
      energy_diffusion_C(&cfgC, &xtype, &xC, &fluidvars, &electrovars, &intvars, &dt);
      clean_param_C(&three, &xtype, &xC, &fluidvars);
      T2rhoe_C(&fluidvars, &fluidauxvars);
      if (mode == 4) {
        // src/libgemini.cpp:223-232, the sources substep
        set_global_boundaries_allspec_C(&xtype, &xC, &fluidvars, &fluidauxvars, &intvars, &lsp);
        source_loss_energy_C(&fluidvars, &fluidauxvars, &electrovars, &intvars, &xtype, &xC, &dt);
        source_loss_momentum_C(&fluidvars, &fluidauxvars, &electrovars, &intvars, &xtype, &xC, &dt);
        source_loss_mass_C(&fluidvars, &fluidauxvars, &electrovars, &intvars, &xtype, &xC, &dt);
        clean_param_C(&three, &xtype, &xC, &fluidvars);
        clean_param_C(&two, &xtype, &xC, &fluidvars);
        clean_param_C(&one, &xtype, &xC, &fluidvars);
      }
      if (gemhip_last_error()[0]) check(1, "step");
    }
    check(gemhip_get_interface_vels(vs1i.data(), vs2i.data(), vs3i.data()), "get_interface_vels");
  } else {
    for (int k = 0; k < nsteps; ++k) {
      // src/libgemini.cpp:155-186
      v12rhov1_C(&fluidvars, &fluidauxvars);
      T2rhoe_C(&fluidvars, &fluidauxvars);
      set_global_boundaries_allspec_C(&xtype, &xC, &fluidvars, &fluidauxvars, &intvars, &lsp);
      halo_fluidvars_C(&xtype, &xC, &fluidvars, &fluidauxvars);
      interface_vels_allspec_C(&xtype, &xC, &fluidvars, &intvars, &lsp);
      sweep3_allparams_C(&fluidvars, &fluidauxvars, &intvars, &xtype, &xC, &dt);
      sweep1_allparams_C(&fluidvars, &fluidauxvars, &intvars, &xtype, &xC, &dt);
      halo_allparams_C(&xtype, &xC, &fluidvars, &fluidauxvars);
      sweep2_allparams_C(&fluidvars, &fluidauxvars, &intvars, &xtype, &xC, &dt);
      rhov12v1_C(&fluidvars, &fluidauxvars);
      // multifluid.f90:179-189, then src/libgemini.cpp:208
      rhoe2T_C(&fluidvars, &fluidauxvars);
      energy_diffusion_C(&cfgC, &xtype, &xC, &fluidvars, &electrovars, &intvars, &dt);
      if (gemhip_last_error()[0]) check(1, "step");
    }
    check(gemhip_get_interface_vels(vs1i.data(), vs2i.data(), vs3i.data()), "get_interface_vels");
  }
  // the plasma output of the final state (gemini3d.h:96), flagoutput 1 + mode % 3, deflate 3
  check(gemhip_register_output(dir.c_str(), 1 + mode % 3, 3), "register_output");
  int ymd[3] = {2015, 4, 13}, ident = mode;
  double utsec = 12345.67891, l1[2] = {0.0, 1.0}, l2[2] = {2.0, 3.0}, l3[2] = {4.0, 5.0};
  dateinc_C(&dt, ymd, &utsec);                              // the clock after one step
  if (gemhip_last_error()[0]) check(1, "dateinc");
  plasma_output_nompi_C(&cfgC, ymd, &utsec, &fluidvars, &electrovars, &ident, l1, l2, l3);
  if (gemhip_last_error()[0]) check(1, "plasma output");
  check(gemhip_output_plasma(ymd, utsec, &fluidvars, &electrovars), "output_plasma");
  FILE* o = fopen((dir + "/out.bin").c_str(), "wb");
  fwrite(fluid.data(), sizeof(double), fluid.size(), o);
  fwrite(aux.data(), sizeof(double), 14 * L, o);
  fwrite(vs1i.data(), sizeof(double), vs1i.size(), o);
  fwrite(vs2i.data(), sizeof(double), vs2i.size(), o);
  fwrite(vs3i.data(), sizeof(double), vs3i.size(), o);
  fclose(o);
  gemhip_finalize();
  return 0;
}
