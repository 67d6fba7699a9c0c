! iso_c_binding interface of libocfd_hip.so (include/ocfd_hip.h) for the
! reference's Fortran driver: the type ocfd_cfg mirrors the C struct field by
! field, and each interface binds one C entry point under its C name.
! Build:  flang -c ocfd_hip.f90   (any Fortran 2003 compiler)
! Link:   -L<repo>/opencfd-comb_amd -locfd_hip -Wl,-rpath,<repo>/opencfd-comb_amd
module ocfd_hip
  use iso_c_binding
  implicit none
  type, bind(C) :: ocfd_cfg
    integer(c_int32_t) :: nx, ny, nz, lap, n_spec
    integer(c_int32_t) :: periodic(3)
    integer(c_int32_t) :: scheme_invis, scheme_vis, flux_splitting, if_viscous, if_scheme_character
    integer(c_int32_t) :: scheme_boundary(6), ghost_cell(6)
    integer(c_int32_t) :: ibc, if_mass_force, if_real_gas, flag_vis, flag_cond, flag_diff
    integer(c_int32_t) :: wall_catalysis, pad0
    real(c_double)     :: ud7l_diss, hybrid_para(16), amu0, pr, sc, mass_force(3), bc_para(32)
  end type
  interface
    integer(c_int32_t) function ocfd_init(cfg) bind(C, name="ocfd_init")
      import; type(ocfd_cfg), intent(in) :: cfg
    end function
    integer(c_int32_t) function ocfd_set_mesh(axx, ayy, azz, span) bind(C, name="ocfd_set_mesh")
      import; real(c_double), intent(in) :: axx(*), ayy(*), azz(*), span(9)
    end function
    integer(c_int32_t) function ocfd_upload_field(id, a) bind(C, name="ocfd_upload_field")
      import; integer(c_int32_t), value :: id; real(c_double), intent(in) :: a(*)
    end function
    integer(c_int32_t) function ocfd_download_field(id, a) bind(C, name="ocfd_download_field")
      import; integer(c_int32_t), value :: id; real(c_double) :: a(*)
    end function
    integer(c_int32_t) function ocfd_flow_time_advance(dt1) bind(C, name="ocfd_flow_time_advance")
      import; real(c_double), value :: dt1
    end function
    integer(c_int32_t) function ocfd_set_bc_time(ht) bind(C, name="ocfd_set_bc_time")
      import; real(c_double), value :: ht
    end function
    integer(c_int32_t) function ocfd_set_species(n_tct, tct, tct_max, mi, coef) bind(C, name="ocfd_set_species")
      import; integer(c_int32_t), value :: n_tct; real(c_double), value :: tct_max
      real(c_double), intent(in) :: tct(*), mi(*), coef(*)
    end function
    integer(c_int32_t) function ocfd_set_transport(mu_an, k_an, d_an) bind(C, name="ocfd_set_transport")
      import; real(c_double), intent(in) :: mu_an(*), k_an(*), d_an(*)
    end function
    integer(c_int32_t) function ocfd_set_bc_jet(inlet2d) bind(C, name="ocfd_set_bc_jet")
      import; real(c_double), intent(in) :: inlet2d(*)
    end function
    integer(c_int32_t) function ocfd_ns_step(dt, nsub, mode, reaction) bind(C, name="ocfd_ns_step")
      import; real(c_double), value :: dt; integer(c_int32_t), value :: nsub, mode, reaction
    end function
    integer(c_int32_t) function ocfd_chemical_time_advance(dt2, mode) bind(C, name="ocfd_chemical_time_advance")
      import; real(c_double), value :: dt2; integer(c_int32_t), value :: mode
    end function
    integer(c_int32_t) function ocfd_comput_t_p() bind(C, name="ocfd_comput_T_P")
      import
    end function
    integer(c_int32_t) function ocfd_chem_set_mechanism(ns, nr, rereac, af_type, third_body, fc_troe, &
        arrhenius, vf, vr, eff, n_tct, tct, tct_max, mi, coef) bind(C, name="ocfd_chem_set_mechanism")
      import; integer(c_int32_t), value :: ns, nr, rereac, n_tct; real(c_double), value :: tct_max
      integer(c_int32_t), intent(in) :: af_type(*), third_body(*), vf(*), vr(*)
      real(c_double), intent(in) :: fc_troe(*), arrhenius(*), eff(*), tct(*), mi(*), coef(*)
    end function
    integer(c_int32_t) function ocfd_cfl_dt(cfl, dt) bind(C, name="ocfd_cfl_dt")
      import; real(c_double), value :: cfl; real(c_double) :: dt
    end function
    integer(c_int32_t) function ocfd_flux1d_batch(side, scheme, npen, n, v, hh, scm, bound_index, &
        scheme_boundary) bind(C, name="ocfd_flux1d_batch")
      import; integer(c_int32_t), value :: side, scheme, n; integer(c_int64_t), value :: npen
      type(c_ptr), value :: v, hh, scm          ! device arrays (hipMalloc'd)
      integer(c_int32_t), intent(in) :: bound_index(2), scheme_boundary(2)
    end function
    integer(c_int32_t) function ocfd_synchronize() bind(C, name="ocfd_synchronize")
      import
    end function
    integer(c_int32_t) function ocfd_finalize() bind(C, name="ocfd_finalize")
      import
    end function
    integer(c_int32_t) function ocfd_set_option(key, val) bind(C, name="ocfd_set_option")
      import; integer(c_int32_t), value :: key, val
    end function
    integer(c_int32_t) function ocfd_set_bc_boundarylayer(inlet, upper, nonref, pertb) &
        bind(C, name="ocfd_set_bc_boundarylayer")
      import; real(c_double), intent(in) :: inlet(*), upper(*), pertb(*); integer(c_int32_t), intent(in) :: nonref(*)
    end function
    integer(c_int32_t) function ocfd_save_stage0() bind(C, name="ocfd_save_stage0")
      import
    end function
    integer(c_int32_t) function ocfd_stage_begin(krk) bind(C, name="ocfd_stage_begin")
      import; integer(c_int32_t), value :: krk
    end function
    integer(c_int32_t) function ocfd_du_inviscous() bind(C, name="ocfd_du_inviscous")
      import
    end function
    integer(c_int32_t) function ocfd_du_viscous() bind(C, name="ocfd_du_viscous")
      import
    end function
    integer(c_int32_t) function ocfd_time_adv_rk3(krk, dt) bind(C, name="ocfd_time_adv_rk3")
      import; integer(c_int32_t), value :: krk; real(c_double), value :: dt
    end function
    integer(c_int32_t) function ocfd_comput_flow_variables() bind(C, name="ocfd_comput_flow_variables")
      import
    end function
    integer(c_int32_t) function ocfd_bc() bind(C, name="ocfd_bc")
      import
    end function
    integer(c_int32_t) function ocfd_filtering(istep, nfiltering, filter) bind(C, name="ocfd_filtering")
      import; integer(c_int32_t), value :: istep, nfiltering; real(c_double), intent(in) :: filter(100, 10)
    end function
    integer(c_int32_t) function ocfd_du_scaler_invis(m) bind(C, name="ocfd_du_scaler_invis")
      import; integer(c_int32_t), value :: m
    end function
    integer(c_int32_t) function ocfd_du_scaler_viscous(m) bind(C, name="ocfd_du_scaler_viscous")
      import; integer(c_int32_t), value :: m
    end function
    integer(c_int32_t) function ocfd_scalar_time_adv_rk3(krk, m, dt) bind(C, name="ocfd_scalar_time_adv_rk3")
      import; integer(c_int32_t), value :: krk, m; real(c_double), value :: dt
    end function
    integer(c_int32_t) function ocfd_chem_set_reverse_rates(rev) bind(C, name="ocfd_chem_set_reverse_rates")
      import; real(c_double), intent(in) :: rev(*)
    end function
    type(c_ptr) function ocfd_last_error() bind(C, name="ocfd_last_error")
      import
    end function
  end interface
contains
  ! the text of ocfd_last_error() as a Fortran string
  function ocfd_error_text() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: n
    type(c_ptr) :: cp
    cp = ocfd_last_error()
    msg = ""
    if (.not. c_associated(cp)) return
    call c_f_pointer(cp, p, [4096])
    n = 0
    do while (n < 4096)
      if (p(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(p(1:n), msg)
  end function
  subroutine ocfd_check(rc)            ! the reference's print + stop on errors
    integer(c_int32_t), intent(in) :: rc
    if (rc /= 0) then
      print*, "ocfd_hip error: ", ocfd_error_text(); stop
    endif
  end subroutine
end module
