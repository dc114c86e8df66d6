!> gemhip_bind -- Fortran (iso_c_binding) interface to libgemhip.so, the MI355X
!> transport library declared in include/gemini3d_hip.h.
!>
!> Part 1 has the names and argument lists of the reference's C API for the transport
!> hot path (include/gemini3d.h:64-137, implemented there by src/libgemini_c.f90:445-848
!> and src/libgemini_mpi_c.f90:414-458): every argument by reference, the state blocks as
!> the address of a C pointer to the caller's Fortran-layout array.  A Fortran host that
!> already calls those *_C routines links this library instead; a host that calls the
!> Fortran procedures directly (app/main.f90:280-355) swaps each call for the *_C one
!> below after the one-time set-up of part 2.
module gemhip_bind
  use, intrinsic :: iso_c_binding, only: c_int, c_long, c_double, c_ptr, c_char, c_null_char
  implicit none
  private
  public :: v12rhov1_C, T2rhoe_C, rhoe2T_C, rhov12v1_C, interface_vels_allspec_C, &
            set_global_boundaries_allspec_C, sweep3_allparams_C, sweep3_allspec_mass_C, &
            sweep3_allspec_momentum_C, sweep3_allspec_energy_C, sweep1_allparams_C, &
            sweep1_allspec_mass_C, sweep1_allspec_momentum_C, sweep1_allspec_energy_C, &
            sweep2_allparams_C, sweep2_allspec_mass_C, sweep2_allspec_momentum_C, &
            sweep2_allspec_energy_C, energy_diffusion_C, halo_interface_vels_allspec_C, &
            halo_allparams_C, halo_fluidvars_C, clean_param_C, VNRicht_artvisc_C, &
            compression_C, RK2_prep_mpi_allspec_C, RK2_global_boundary_allspec_C, dt_select_C, &
            source_loss_allparams_C, source_loss_mass_C, source_loss_momentum_C, &
            source_loss_energy_C, clear_ionization_arrays_C, plasma_output_nompi_C, dateinc_C
  public :: gemhip_init, gemhip_finalize, gemhip_set_error_mode, gemhip_clear_error, &
            gemhip_set_grid_size, gemhip_register_grid, gemhip_register_neutrals, &
            gemhip_register_config, gemhip_set_residency, gemhip_state_push, &
            gemhip_state_pull, gemhip_get_interface_vels, gemhip_set_interface_vels, &
            gemhip_transport_step, gemhip_synchronize, gemhip_error_message, &
            gemhip_register_null_cells, gemhip_fluid_adv_block, gemhip_register_cfl, &
            gemhip_register_atmosphere, gemhip_register_gravity, gemhip_set_ionization_rates, &
            gemhip_sources_block, gemhip_register_output, gemhip_read_initial_conditions, &
            gemhip_output_plasma, gemhip_read_simsize, gemhip_read_grid_coords

  ! ------------------------------------------------ part 1: reference entry points
  interface
    subroutine v12rhov1_C(fluidvars, fluidauxvars) bind(C, name="v12rhov1_C")
      import :: c_ptr
      type(c_ptr) :: fluidvars, fluidauxvars
    end subroutine
    subroutine T2rhoe_C(fluidvars, fluidauxvars) bind(C, name="T2rhoe_C")
      import :: c_ptr
      type(c_ptr) :: fluidvars, fluidauxvars
    end subroutine
    subroutine rhoe2T_C(fluidvars, fluidauxvars) bind(C, name="rhoe2T_C")
      import :: c_ptr
      type(c_ptr) :: fluidvars, fluidauxvars
    end subroutine
    subroutine rhov12v1_C(fluidvars, fluidauxvars) bind(C, name="rhov12v1_C")
      import :: c_ptr
      type(c_ptr) :: fluidvars, fluidauxvars
    end subroutine
    subroutine interface_vels_allspec_C(xtype, xC, fluidvars, intvars, lsp) &
        bind(C, name="interface_vels_allspec_C")
      import :: c_ptr, c_int
      integer(c_int) :: xtype, lsp
      type(c_ptr) :: xC, fluidvars, intvars
    end subroutine
    subroutine set_global_boundaries_allspec_C(xtype, xC, fluidvars, fluidauxvars, intvars, lsp) &
        bind(C, name="set_global_boundaries_allspec_C")
      import :: c_ptr, c_int
      integer(c_int) :: xtype, lsp
      type(c_ptr) :: xC, fluidvars, fluidauxvars, intvars
    end subroutine
    subroutine sweep3_allparams_C(fluidvars, fluidauxvars, intvars, xtype, xC, dt) &
        bind(C, name="sweep3_allparams_C")
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: fluidvars, fluidauxvars, intvars, xC
      integer(c_int) :: xtype
      real(c_double) :: dt
    end subroutine
    subroutine sweep3_allspec_mass_C(fluidvars, fluidauxvars, intvars, xtype, xC, dt) &
        bind(C, name="sweep3_allspec_mass_C")
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: fluidvars, fluidauxvars, intvars, xC
      integer(c_int) :: xtype
      real(c_double) :: dt
    end subroutine
    subroutine sweep3_allspec_momentum_C(fluidvars, fluidauxvars, intvars, xtype, xC, dt) &
        bind(C, name="sweep3_allspec_momentum_C")
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: fluidvars, fluidauxvars, intvars, xC
      integer(c_int) :: xtype
      real(c_double) :: dt
    end subroutine
    subroutine sweep3_allspec_energy_C(fluidvars, fluidauxvars, intvars, xtype, xC, dt) &
        bind(C, name="sweep3_allspec_energy_C")
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: fluidvars, fluidauxvars, intvars, xC
      integer(c_int) :: xtype
      real(c_double) :: dt
    end subroutine
    subroutine sweep1_allparams_C(fluidvars, fluidauxvars, intvars, xtype, xC, dt) &
        bind(C, name="sweep1_allparams_C")
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: fluidvars, fluidauxvars, intvars, xC
      integer(c_int) :: xtype
      real(c_double) :: dt
    end subroutine
    subroutine sweep1_allspec_mass_C(fluidvars, fluidauxvars, intvars, xtype, xC, dt) &
        bind(C, name="sweep1_allspec_mass_C")
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: fluidvars, fluidauxvars, intvars, xC
      integer(c_int) :: xtype
      real(c_double) :: dt
    end subroutine
    subroutine sweep1_allspec_momentum_C(fluidvars, fluidauxvars, intvars, xtype, xC, dt) &
        bind(C, name="sweep1_allspec_momentum_C")
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: fluidvars, fluidauxvars, intvars, xC
      integer(c_int) :: xtype
      real(c_double) :: dt
    end subroutine
    subroutine sweep1_allspec_energy_C(fluidvars, fluidauxvars, intvars, xtype, xC, dt) &
        bind(C, name="sweep1_allspec_energy_C")
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: fluidvars, fluidauxvars, intvars, xC
      integer(c_int) :: xtype
      real(c_double) :: dt
    end subroutine
    subroutine sweep2_allparams_C(fluidvars, fluidauxvars, intvars, xtype, xC, dt) &
        bind(C, name="sweep2_allparams_C")
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: fluidvars, fluidauxvars, intvars, xC
      integer(c_int) :: xtype
      real(c_double) :: dt
    end subroutine
    subroutine sweep2_allspec_mass_C(fluidvars, fluidauxvars, intvars, xtype, xC, dt) &
        bind(C, name="sweep2_allspec_mass_C")
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: fluidvars, fluidauxvars, intvars, xC
      integer(c_int) :: xtype
      real(c_double) :: dt
    end subroutine
    subroutine sweep2_allspec_momentum_C(fluidvars, fluidauxvars, intvars, xtype, xC, dt) &
        bind(C, name="sweep2_allspec_momentum_C")
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: fluidvars, fluidauxvars, intvars, xC
      integer(c_int) :: xtype
      real(c_double) :: dt
    end subroutine
    subroutine sweep2_allspec_energy_C(fluidvars, fluidauxvars, intvars, xtype, xC, dt) &
        bind(C, name="sweep2_allspec_energy_C")
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: fluidvars, fluidauxvars, intvars, xC
      integer(c_int) :: xtype
      real(c_double) :: dt
    end subroutine
    subroutine energy_diffusion_C(cfgC, xtype, xC, fluidvars, electrovars, intvars, dt) &
        bind(C, name="energy_diffusion_C")
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: cfgC, xC, fluidvars, electrovars, intvars
      integer(c_int) :: xtype
      real(c_double) :: dt
    end subroutine
    subroutine halo_interface_vels_allspec_C(xtype, xC, fluidvars, lsp) &
        bind(C, name="halo_interface_vels_allspec_C")
      import :: c_ptr, c_int
      integer(c_int) :: xtype, lsp
      type(c_ptr) :: xC, fluidvars
    end subroutine
    subroutine halo_allparams_C(xtype, xC, fluidvars, fluidauxvars) bind(C, name="halo_allparams_C")
      import :: c_ptr, c_int
      integer(c_int) :: xtype
      type(c_ptr) :: xC, fluidvars, fluidauxvars
    end subroutine
    subroutine halo_fluidvars_C(xtype, xC, fluidvars, fluidauxvars) bind(C, name="halo_fluidvars_C")
      import :: c_ptr, c_int
      integer(c_int) :: xtype
      type(c_ptr) :: xC, fluidvars, fluidauxvars
    end subroutine
    subroutine clean_param_C(iparm, xtype, xC, fluidvars) bind(C, name="clean_param_C")
      import :: c_ptr, c_int
      integer(c_int) :: iparm, xtype
      type(c_ptr) :: xC, fluidvars
    end subroutine
    subroutine VNRicht_artvisc_C(fluidvars, intvars) bind(C, name="VNRicht_artvisc_C")
      import :: c_ptr
      type(c_ptr) :: fluidvars, intvars
    end subroutine
    subroutine compression_C(fluidvars, fluidauxvars, intvars, xtype, xC, dt) &
        bind(C, name="compression_C")
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: fluidvars, fluidauxvars, intvars, xC
      integer(c_int) :: xtype
      real(c_double) :: dt
    end subroutine
    subroutine RK2_prep_mpi_allspec_C(xtype, xC, fluidvars) bind(C, name="RK2_prep_mpi_allspec_C")
      import :: c_ptr, c_int
      integer(c_int) :: xtype
      type(c_ptr) :: xC, fluidvars
    end subroutine
    subroutine RK2_global_boundary_allspec_C(xtype, xC, fluidvars) &
        bind(C, name="RK2_global_boundary_allspec_C")
      import :: c_ptr, c_int
      integer(c_int) :: xtype
      type(c_ptr) :: xC, fluidvars
    end subroutine
    subroutine dt_select_C(cfgC, xtype, xC, fluidvars, fluidauxvars, t, tout, tglowout, dt) &
        bind(C, name="dt_select_C")
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: cfgC, xC, fluidvars, fluidauxvars
      integer(c_int) :: xtype
      real(c_double) :: t, tout, tglowout, dt
    end subroutine
    subroutine source_loss_allparams_C(fluidvars, fluidauxvars, electrovars, intvars, xtype, xC, dt) &
        bind(C, name="source_loss_allparams_C")
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: fluidvars, fluidauxvars, electrovars, intvars, xC
      integer(c_int) :: xtype
      real(c_double) :: dt
    end subroutine
    subroutine source_loss_mass_C(fluidvars, fluidauxvars, electrovars, intvars, xtype, xC, dt) &
        bind(C, name="source_loss_mass_C")
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: fluidvars, fluidauxvars, electrovars, intvars, xC
      integer(c_int) :: xtype
      real(c_double) :: dt
    end subroutine
    subroutine source_loss_momentum_C(fluidvars, fluidauxvars, electrovars, intvars, xtype, xC, dt) &
        bind(C, name="source_loss_momentum_C")
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: fluidvars, fluidauxvars, electrovars, intvars, xC
      integer(c_int) :: xtype
      real(c_double) :: dt
    end subroutine
    subroutine source_loss_energy_C(fluidvars, fluidauxvars, electrovars, intvars, xtype, xC, dt) &
        bind(C, name="source_loss_energy_C")
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: fluidvars, fluidauxvars, electrovars, intvars, xC
      integer(c_int) :: xtype
      real(c_double) :: dt
    end subroutine
    subroutine clear_ionization_arrays_C(intvars) bind(C, name="clear_ionization_arrays_C")
      import :: c_ptr
      type(c_ptr) :: intvars
    end subroutine
    subroutine dateinc_C(dt, ymd, UTsec) bind(C, name="dateinc_C")
      import :: c_int, c_double
      real(c_double) :: dt, UTsec
      integer(c_int) :: ymd(3)
    end subroutine
    subroutine plasma_output_nompi_C(cfgC, ymd, UTsec, fluidvars, electrovars, identifier, &
                                     x1lims, x2lims, x3lims) bind(C, name="plasma_output_nompi_C")
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: cfgC, fluidvars, electrovars
      integer(c_int) :: ymd(3), identifier
      real(c_double) :: UTsec, x1lims(2), x2lims(2), x3lims(2)
    end subroutine
  end interface

  ! ------------------------------------------------ part 2: library set-up
  interface
    integer(c_int) function gemhip_init(device) bind(C, name="gemhip_init")
      import :: c_int
      integer(c_int), value :: device
    end function
    subroutine gemhip_finalize() bind(C, name="gemhip_finalize")
    end subroutine
    subroutine gemhip_set_error_mode(mode) bind(C, name="gemhip_set_error_mode")
      import :: c_int
      integer(c_int), value :: mode
    end subroutine
    subroutine gemhip_clear_error() bind(C, name="gemhip_clear_error")
    end subroutine
    type(c_ptr) function gemhip_last_error_c() bind(C, name="gemhip_last_error")
      import :: c_ptr
    end function
    integer(c_int) function gemhip_set_grid_size(lx1, lx2, lx3) bind(C, name="gemhip_set_grid_size")
      import :: c_int
      integer(c_int), value :: lx1, lx2, lx3
    end function
    integer(c_int) function gemhip_register_grid(xtype, gridflag, flagper, isglobal6, &
        h1, h2, h3, h1x1i, h2x1i, h3x1i, h1x2i, h2x2i, h3x2i, h1x3i, h2x3i, h3x3i, &
        dx1, dx1i, dx2, dx2i, dx3, dx3i) bind(C, name="gemhip_register_grid")
      import :: c_int, c_double
      integer(c_int), value :: xtype, gridflag, flagper
      integer(c_int), intent(in) :: isglobal6(6)
      real(c_double), intent(in) :: h1(*), h2(*), h3(*), h1x1i(*), h2x1i(*), h3x1i(*), &
          h1x2i(*), h2x2i(*), h3x2i(*), h1x3i(*), h2x3i(*), h3x3i(*), &
          dx1(*), dx1i(*), dx2(*), dx2i(*), dx3(*), dx3i(*)
    end function
    integer(c_int) function gemhip_register_neutrals(nn, Tn) bind(C, name="gemhip_register_neutrals")
      import :: c_int, c_double
      real(c_double), intent(in) :: nn(*), Tn(*)
    end function
    integer(c_int) function gemhip_register_config(diffsolvetype, Teinf) &
        bind(C, name="gemhip_register_config")
      import :: c_int, c_double
      integer(c_int), value :: diffsolvetype
      real(c_double), value :: Teinf
    end function
    integer(c_int) function gemhip_set_residency(mode) bind(C, name="gemhip_set_residency")
      import :: c_int
      integer(c_int), value :: mode
    end function
    integer(c_int) function gemhip_state_push(fluidvars, fluidauxvars, electrovars) &
        bind(C, name="gemhip_state_push")
      import :: c_int, c_double
      real(c_double), intent(in) :: fluidvars(*), fluidauxvars(*), electrovars(*)
    end function
    integer(c_int) function gemhip_state_pull(fluidvars, fluidauxvars) bind(C, name="gemhip_state_pull")
      import :: c_int, c_double
      real(c_double) :: fluidvars(*), fluidauxvars(*)
    end function
    integer(c_int) function gemhip_get_interface_vels(vs1i, vs2i, vs3i) &
        bind(C, name="gemhip_get_interface_vels")
      import :: c_int, c_double
      real(c_double) :: vs1i(*), vs2i(*), vs3i(*)
    end function
    integer(c_int) function gemhip_set_interface_vels(vs1i, vs2i, vs3i) &
        bind(C, name="gemhip_set_interface_vels")
      import :: c_int, c_double
      real(c_double), intent(in) :: vs1i(*), vs2i(*), vs3i(*)
    end function
    integer(c_int) function gemhip_register_null_cells(nullpts) &
        bind(C, name="gemhip_register_null_cells")
      import :: c_int
      integer(c_int), intent(in) :: nullpts(*)
    end function
    integer(c_int) function gemhip_register_cfl(tcfl, flagglow) bind(C, name="gemhip_register_cfl")
      import :: c_int, c_double
      real(c_double), value :: tcfl
      integer(c_int), value :: flagglow
    end function
    integer(c_int) function gemhip_fluid_adv_block(dt) bind(C, name="gemhip_fluid_adv_block")
      import :: c_int, c_double
      real(c_double), value :: dt
    end function
    integer(c_int) function gemhip_sources_block(dt) bind(C, name="gemhip_sources_block")
      import :: c_int, c_double
      real(c_double), value :: dt
    end function
    integer(c_int) function gemhip_register_atmosphere(nn, Tn, vn1, vn2, vn3) &
        bind(C, name="gemhip_register_atmosphere")
      import :: c_int, c_double
      real(c_double), intent(in) :: nn(*), Tn(*), vn1(*), vn2(*), vn3(*)
    end function
    integer(c_int) function gemhip_register_gravity(alt, g1) bind(C, name="gemhip_register_gravity")
      import :: c_int, c_double
      real(c_double), intent(in) :: alt(*), g1(*)
    end function
    integer(c_int) function gemhip_set_ionization_rates(Prionize, Qeionize) &
        bind(C, name="gemhip_set_ionization_rates")
      import :: c_int, c_double
      real(c_double), intent(in) :: Prionize(*), Qeionize(*)
    end function
    integer(c_int) function gemhip_transport_step(dt) bind(C, name="gemhip_transport_step")
      import :: c_int, c_double
      real(c_double), value :: dt
    end function
    integer(c_int) function gemhip_synchronize() bind(C, name="gemhip_synchronize")
      import :: c_int
    end function
    ! character arguments are C strings: trim(path) // c_null_char
    integer(c_int) function gemhip_register_output(outdir, flagoutput, comp_lvl) &
        bind(C, name="gemhip_register_output")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: outdir(*)
      integer(c_int), value :: flagoutput, comp_lvl
    end function
    integer(c_int) function gemhip_read_initial_conditions(indatsize, indatfile, lx1, lx2, lx3, &
        fluidvars, electrovars) bind(C, name="gemhip_read_initial_conditions")
      import :: c_int, c_char, c_double
      character(kind=c_char), intent(in) :: indatsize(*), indatfile(*)
      integer(c_int), value :: lx1, lx2, lx3
      real(c_double) :: fluidvars(*), electrovars(*)
    end function
    integer(c_int) function gemhip_output_plasma(ymd, UTsec, fluidvars, electrovars) &
        bind(C, name="gemhip_output_plasma")
      import :: c_int, c_double, c_ptr
      integer(c_int), intent(in) :: ymd(3)
      real(c_double), value :: UTsec
      type(c_ptr) :: fluidvars, electrovars
    end function
    integer(c_int) function gemhip_read_simsize(path, lx1, lx2, lx3) bind(C, name="gemhip_read_simsize")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), intent(out) :: lx1, lx2, lx3
    end function
    integer(c_int) function gemhip_read_grid_coords(path, lx1, lx2, lx3, x1, x2, x3, glonctr, &
        glatctr) bind(C, name="gemhip_read_grid_coords")
      import :: c_int, c_char, c_double
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: lx1, lx2, lx3
      real(c_double), intent(out) :: x1(*), x2(*), x3(*), glonctr, glatctr
    end function
    integer(c_long) function c_strlen(s) bind(C, name="strlen")
      import :: c_long, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  !> The library's last error message as a Fortran string ("" when none), for hosts
  !> that turn a failed call into `error stop` as the reference does.
  function gemhip_error_message() result(msg)
    use, intrinsic :: iso_c_binding, only: c_f_pointer
    character(:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = gemhip_last_error_c()
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function

end module gemhip_bind
