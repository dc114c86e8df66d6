!> A Fortran host driving the transport hot path through gemhip_bind, the way the
!> reference's main loop does (app/main.f90:280-355 call order, the C-API names of
!> src/libgemini.cpp:155-208).  Test program: reads a case written by the test suite
!> (case.txt, in.bin), runs nsteps steps and writes out.bin.
!>
!> mode 0: every step calls the reference entry points one by one on the host blocks
!>         (residency 0: each call copies to HBM, runs and copies back);
!> mode 1: the state is pushed once, gemhip_transport_step runs each step on HBM, and the
!>         state is pulled at the end.
!> modes 2 and 3: the same for the whole transport block (src/libgemini.cpp:155-210);
!> modes 4 and 5: the block followed by the sources substep (:223-232), call by call
!>         or fused (gemhip_sources_block).
!> Every mode then writes the final state with plasma_output_nompi_C and
!> gemhip_output_plasma (flagoutput 1 + mod(mode, 3)) into the case directory.
program fortran_driver
  use, intrinsic :: iso_c_binding
  use gemhip_bind
  implicit none
  integer(c_int) :: lx1, lx2, lx3, xtype, gridflag, flagper, isg(6), solver, nsteps, mode, lsp
  real(c_double) :: dt, Teinf
  real(c_double), allocatable, target :: h1(:,:,:), h2(:,:,:), h3(:,:,:)
  real(c_double), allocatable :: h1x1i(:,:,:), h2x1i(:,:,:), h3x1i(:,:,:)
  real(c_double), allocatable :: h1x2i(:,:,:), h2x2i(:,:,:), h3x2i(:,:,:)
  real(c_double), allocatable :: h1x3i(:,:,:), h2x3i(:,:,:), h3x3i(:,:,:)
  real(c_double), allocatable :: dx1(:), dx1i(:), dx2(:), dx2i(:), dx3(:), dx3i(:)
  real(c_double), allocatable :: nn(:,:,:,:), Tn(:,:,:)
  real(c_double), allocatable, target :: fluidvars(:,:,:,:), fluidauxvars(:,:,:,:), electrovars(:,:,:,:)
  real(c_double), allocatable :: vs1i(:,:,:,:), vs2i(:,:,:,:), vs3i(:,:,:,:)
  real(c_double), allocatable :: nn6(:,:,:,:), vn1(:,:,:), vn2(:,:,:), vn3(:,:,:)
  real(c_double), allocatable :: Prion(:,:,:,:), Qeion(:,:,:), alt(:,:,:), g1(:,:,:)
  type(c_ptr) :: fvp, fap, evp, xC, intvars, cfgC
  character(len=4096) :: dir
  integer :: u, k, rc, w
  integer(c_int) :: ymd(3), ident
  real(c_double) :: UTsec, x1lims(2), x2lims(2), x3lims(2)

  call get_command_argument(1, dir)
  open(newunit=u, file=trim(dir)//'/case.txt', status='old', action='read')
  read(u, *) lx1, lx2, lx3, xtype, gridflag, flagper, isg, solver, nsteps, mode, dt, Teinf
  close(u)
  allocate(h1(-1:lx1+2,-1:lx2+2,-1:lx3+2), h2(-1:lx1+2,-1:lx2+2,-1:lx3+2), h3(-1:lx1+2,-1:lx2+2,-1:lx3+2))
  allocate(h1x1i(lx1+1,lx2,lx3), h2x1i(lx1+1,lx2,lx3), h3x1i(lx1+1,lx2,lx3))
  allocate(h1x2i(lx1,lx2+1,lx3), h2x2i(lx1,lx2+1,lx3), h3x2i(lx1,lx2+1,lx3))
  allocate(h1x3i(lx1,lx2,lx3+1), h2x3i(lx1,lx2,lx3+1), h3x3i(lx1,lx2,lx3+1))
  allocate(dx1(0:lx1+2), dx1i(lx1), dx2(0:lx2+2), dx2i(lx2), dx3(0:lx3+2), dx3i(lx3))
  allocate(nn(lx1,lx2,lx3,3), Tn(lx1,lx2,lx3))
  allocate(fluidvars(-1:lx1+2,-1:lx2+2,-1:lx3+2,35), fluidauxvars(-1:lx1+2,-1:lx2+2,-1:lx3+2,23))
  allocate(electrovars(-1:lx1+2,-1:lx2+2,-1:lx3+2,7))
  allocate(vs1i(lx1+1,lx2,lx3,7), vs2i(lx1,lx2+1,lx3,7), vs3i(lx1,lx2,lx3+1,7))
  open(newunit=u, file=trim(dir)//'/in.bin', access='stream', form='unformatted', status='old')
  read(u) h1, h2, h3, h1x1i, h2x1i, h3x1i, h1x2i, h2x2i, h3x2i, h1x3i, h2x3i, h3x3i
  read(u) dx1, dx1i, dx2, dx2i, dx3, dx3i, nn, Tn, fluidvars, fluidauxvars, electrovars
  read(u) vs1i, vs2i, vs3i
  if (mode >= 4) then
    allocate(nn6(lx1,lx2,lx3,6), vn1(lx1,lx2,lx3), vn2(lx1,lx2,lx3), vn3(lx1,lx2,lx3))
    allocate(Prion(lx1,lx2,lx3,6), Qeion(lx1,lx2,lx3), alt(-1:lx1+2,-1:lx2+2,-1:lx3+2), g1(lx1,lx2,lx3))
    read(u) nn6, vn1, vn2, vn3, Prion, Qeion, alt, g1
  end if
  close(u)

  rc = gemhip_init(0_c_int)
  call gemhip_set_error_mode(1_c_int)
  rc = rc + gemhip_set_grid_size(lx1, lx2, lx3)
  rc = rc + gemhip_register_grid(xtype, gridflag, flagper, isg, h1, h2, h3, h1x1i, h2x1i, h3x1i, &
       h1x2i, h2x2i, h3x2i, h1x3i, h2x3i, h3x3i, dx1, dx1i, dx2, dx2i, dx3, dx3i)
  rc = rc + gemhip_register_neutrals(nn, Tn)
  rc = rc + gemhip_register_config(solver, Teinf)
  rc = rc + gemhip_set_interface_vels(vs1i, vs2i, vs3i)
  if (mode >= 4) then
    rc = rc + gemhip_register_atmosphere(nn6, Tn, vn1, vn2, vn3)
    rc = rc + gemhip_register_gravity(alt, g1)
    rc = rc + gemhip_set_ionization_rates(Prion, Qeion)
  end if
  if (rc /= 0) error stop 'set-up failed: '//gemhip_error_message()

  fvp = c_loc(fluidvars)
  fap = c_loc(fluidauxvars)
  evp = c_loc(electrovars)
  xC = c_null_ptr
  intvars = c_null_ptr
  cfgC = c_null_ptr
  lsp = 7
  if (mode == 1 .or. mode == 3 .or. mode == 5) then
    rc = gemhip_set_residency(1_c_int)
    rc = rc + gemhip_state_push(fluidvars, fluidauxvars, electrovars)
    do k = 1, nsteps
      if (mode == 1) then
        rc = rc + gemhip_transport_step(dt)
      else
        rc = rc + gemhip_fluid_adv_block(dt)
      end if
      if (mode == 5) rc = rc + gemhip_sources_block(dt)
    end do
    rc = rc + gemhip_state_pull(fluidvars, fluidauxvars)
    if (rc /= 0) error stop 'fused step failed: '//gemhip_error_message()
  else if (mode == 2 .or. mode == 4) then
    do k = 1, nsteps
      call v12rhov1_C(fvp, fap)
      call T2rhoe_C(fvp, fap)
      call set_global_boundaries_allspec_C(xtype, xC, fvp, fap, intvars, lsp)
      call halo_fluidvars_C(xtype, xC, fvp, fap)
      call interface_vels_allspec_C(xtype, xC, fvp, intvars, lsp)
      call sweep3_allparams_C(fvp, fap, intvars, xtype, xC, dt)
      call sweep1_allparams_C(fvp, fap, intvars, xtype, xC, dt)
      call halo_allparams_C(xtype, xC, fvp, fap)
      call sweep2_allparams_C(fvp, fap, intvars, xtype, xC, dt)
      call rhov12v1_C(fvp, fap)
      call clean_param_C(1_c_int, xtype, xC, fvp)
      call clean_param_C(2_c_int, xtype, xC, fvp)
      call VNRicht_artvisc_C(fvp, intvars)
      call halo_fluidvars_C(xtype, xC, fvp, fap)
      call RK2_global_boundary_allspec_C(xtype, xC, fvp)
      call compression_C(fvp, fap, intvars, xtype, xC, dt)
      call rhoe2T_C(fvp, fap)
      call clean_param_C(3_c_int, xtype, xC, fvp)
      call energy_diffusion_C(cfgC, xtype, xC, fvp, evp, intvars, dt)
      call clean_param_C(3_c_int, xtype, xC, fvp)
      call T2rhoe_C(fvp, fap)
      if (mode == 4) then
        call set_global_boundaries_allspec_C(xtype, xC, fvp, fap, intvars, lsp)
        call source_loss_energy_C(fvp, fap, evp, intvars, xtype, xC, dt)
        call source_loss_momentum_C(fvp, fap, evp, intvars, xtype, xC, dt)
        call source_loss_mass_C(fvp, fap, evp, intvars, xtype, xC, dt)
        call clean_param_C(3_c_int, xtype, xC, fvp)
        call clean_param_C(2_c_int, xtype, xC, fvp)
        call clean_param_C(1_c_int, xtype, xC, fvp)
      end if
      if (len(gemhip_error_message()) > 0) error stop 'step failed: '//gemhip_error_message()
    end do
  else
    do k = 1, nsteps
      call v12rhov1_C(fvp, fap)
      call T2rhoe_C(fvp, fap)
      call set_global_boundaries_allspec_C(xtype, xC, fvp, fap, intvars, lsp)
      call halo_fluidvars_C(xtype, xC, fvp, fap)
      call interface_vels_allspec_C(xtype, xC, fvp, intvars, lsp)
      call sweep3_allspec_mass_C(fvp, fap, intvars, xtype, xC, dt)
      call sweep3_allspec_momentum_C(fvp, fap, intvars, xtype, xC, dt)
      call sweep3_allspec_energy_C(fvp, fap, intvars, xtype, xC, dt)
      call sweep1_allspec_mass_C(fvp, fap, intvars, xtype, xC, dt)
      call sweep1_allspec_momentum_C(fvp, fap, intvars, xtype, xC, dt)
      call sweep1_allspec_energy_C(fvp, fap, intvars, xtype, xC, dt)
      call halo_allparams_C(xtype, xC, fvp, fap)
      call sweep2_allspec_mass_C(fvp, fap, intvars, xtype, xC, dt)
      call sweep2_allspec_momentum_C(fvp, fap, intvars, xtype, xC, dt)
      call sweep2_allspec_energy_C(fvp, fap, intvars, xtype, xC, dt)
      call rhov12v1_C(fvp, fap)
      call rhoe2T_C(fvp, fap)
      call energy_diffusion_C(cfgC, xtype, xC, fvp, evp, intvars, dt)
      if (len(gemhip_error_message()) > 0) error stop 'step failed: '//gemhip_error_message()
    end do
  end if
  if (mode == 0 .or. mode == 2 .or. mode == 4) rc = gemhip_get_interface_vels(vs1i, vs2i, vs3i)
  ! the plasma output of the final state (gemini3d.h:96), flagoutput 1 + mod(mode, 3), deflate 3
  rc = gemhip_register_output(trim(dir)//c_null_char, 1_c_int + mod(mode, 3_c_int), 3_c_int)
  ymd = [2015, 4, 13]
  UTsec = 12345.67891_c_double
  call dateinc_C(dt, ymd, UTsec)                           ! the clock after one step
  ident = mode
  x1lims = [0.0_c_double, 1.0_c_double]
  x2lims = [2.0_c_double, 3.0_c_double]
  x3lims = [4.0_c_double, 5.0_c_double]
  call plasma_output_nompi_C(cfgC, ymd, UTsec, fvp, evp, ident, x1lims, x2lims, x3lims)
  if (rc /= 0 .or. len(gemhip_error_message()) > 0) error stop 'output failed: '//gemhip_error_message()
  rc = gemhip_output_plasma(ymd, UTsec, fvp, evp)          ! gemini.bin's own file
  if (rc /= 0) error stop 'root output failed: '//gemhip_error_message()
  open(newunit=w, file=trim(dir)//'/out.bin', access='stream', form='unformatted', status='replace')
  write(w) fluidvars, fluidauxvars(:,:,:,1:14), vs1i, vs2i, vs3i
  close(w)
  call gemhip_finalize()
end program fortran_driver
