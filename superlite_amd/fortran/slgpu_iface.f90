! iso_c_binding interfaces to libslgpu.so (include/slgpu.h), for a SuperLite host.
!
! A patched SuperLite keeps its driver (superlite.f90) and module state and replaces the
! bodies of physical_opacity + opacity_planckmean + opacity_rossmean, leakage_opacity +
! emission_probability (+ sourceenergy), sourcenumbers/vacancies/interior_source/
! source_transformdirection, particle_advance and fluxtally by calls below.
module slgpu_iface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: SLGPU_OK = 0, SLGPU_EINVAL = 1, SLGPU_EORDER = 2, &
       SLGPU_EHIP = 3, SLGPU_ETRANSPORT = 4, SLGPU_ENOMEM = 5
  integer(c_int), parameter :: SLGPU_FP64 = 0, SLGPU_MIXED = 1
  integer(c_int), parameter :: SLGPU_NOBB = 1, SLGPU_NOBF = 2, SLGPU_NOFF = 4, SLGPU_NOTHOMSON = 8

  type, bind(C) :: slgpu_config
    integer(c_int) :: nx, ng, nep, nepg
    real(c_double) :: tauddmc, taulump
    integer(c_int) :: puretran, errorfatal, precision
    integer(c_int32_t) :: seed
    integer(c_int64_t) :: npartmax, max_steps
    integer(c_int) :: ampfact, thm_aniso
  end type

  type, bind(C) :: slgpu_gas
    type(c_ptr) :: temp, natom, vol, mass, nelec, natomfr, oc, ginv
  end type

  interface
    integer(c_int) function slgpu_create(cfg, ctx) bind(C, name="slgpu_create")
      import :: c_int, c_ptr, slgpu_config
      type(slgpu_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function slgpu_destroy(ctx) bind(C, name="slgpu_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function slgpu_last_error_message() bind(C, name="slgpu_last_error_message")
      import :: c_ptr
    end function

    integer(c_int) function slgpu_set_grid(ctx, xarr, vxarr, divv, vol, rcell) &
         bind(C, name="slgpu_set_grid")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: xarr(*), vxarr(*), divv(*), vol(*), rcell(*)
    end function

    integer(c_int) function slgpu_set_groups(ctx, grp_wl) bind(C, name="slgpu_set_groups")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: grp_wl(*)
    end function

    integer(c_int) function slgpu_set_lines(ctx, nline, wl0, gxs, chilw, iz, ii) &
         bind(C, name="slgpu_set_lines")
      import :: c_int, c_ptr, c_float, c_int16_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nline
      real(c_float), intent(in) :: wl0(*), gxs(*), chilw(*)
      integer(c_int16_t), intent(in) :: iz(*), ii(*)
    end function

    integer(c_int) function slgpu_set_bf_tables(ctx, ph1, ph2) bind(C, name="slgpu_set_bf_tables")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: ctx
      real(c_float), intent(in) :: ph1(*), ph2(*)
    end function

    integer(c_int) function slgpu_set_ff_table(ctx, gff) bind(C, name="slgpu_set_ff_table")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: gff(*)
    end function

    integer(c_int) function slgpu_set_ions(ctx, nion, ion_iz, ion_ii) bind(C, name="slgpu_set_ions")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nion
      integer(c_int), intent(in) :: ion_iz(*), ion_ii(*)
    end function

    integer(c_int) function slgpu_opacity_build(ctx, gas, flags, cap64, cap32, sig, capgrey, &
         capross, emiss) bind(C, name="slgpu_opacity_build")
      import :: c_int, c_ptr, slgpu_gas
      type(c_ptr), value :: ctx
      type(slgpu_gas), intent(in) :: gas
      integer(c_int), value :: flags
      type(c_ptr), value :: cap64, cap32, sig, capgrey, capross, emiss
    end function

    integer(c_int) function slgpu_opacity_build_noalloc(ctx, gas, flags, cap64, cap32, sig, &
         capgrey, capross, emiss) bind(C, name="slgpu_opacity_build_noalloc")
      import :: c_int, c_ptr, slgpu_gas
      type(c_ptr), value :: ctx
      type(slgpu_gas), intent(in) :: gas
      integer(c_int), value :: flags
      type(c_ptr), value :: cap64, cap32, sig, capgrey, capross, emiss
    end function

    integer(c_int) function slgpu_set_nlte_data(ctx, nlev, nlin, ncoll, glev, llw, lup, wl0, f, &
         gl, gu, aul, bul, blu, llw_coll, lup_coll, coll_c, dele, lev_pi, pi_c, rr_c, dele_pi) &
         bind(C, name="slgpu_set_nlte_data")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nlev, nlin, ncoll
      real(c_double), intent(in) :: glev(*), wl0(*), f(*), gl(*), gu(*), aul(*), bul(*), blu(*)
      integer(c_int), intent(in) :: llw(*), lup(*), llw_coll(*), lup_coll(*), lev_pi(*)
      real(c_double), intent(in) :: coll_c(*), dele(*), pi_c(*), rr_c(*), dele_pi(*)
    end function

    integer(c_int) function slgpu_opacity_build_nlte(ctx, gas, radtemp, ndens, flags, cap64, &
         cap32, capemit, sig, capgrey, capemitgrey, capross, emiss, npop) &
         bind(C, name="slgpu_opacity_build_nlte")
      import :: c_int, c_ptr, c_double, slgpu_gas
      type(c_ptr), value :: ctx
      type(slgpu_gas), intent(in) :: gas
      real(c_double), intent(in) :: radtemp(*), ndens(*)
      integer(c_int), value :: flags
      type(c_ptr), value :: cap64, cap32, capemit, sig, capgrey, capemitgrey, capross, emiss, npop
    end function

    ! nlte_ndens(i,1:2,1) of eos_update.f:52-57 from the last slgpu_eos (2*nc, H I then H II)
    integer(c_int) function slgpu_eos_nlte_ndens(ctx, ndens) bind(C, name="slgpu_eos_nlte_ndens")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ndens(*)
    end function

    ! nlte_rad_est (nltemod.f:231-260): call after slgpu_transport with grd_radtemp (in/out);
    ! jrad (ng*nc) may be c_null_ptr
    integer(c_int) function slgpu_nlte_rad_est(ctx, radtemp, jrad) bind(C, name="slgpu_nlte_rad_est")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: radtemp(*)
      type(c_ptr), value :: jrad
    end function

    integer(c_int) function slgpu_ddmc_prep(ctx, radtemp, lbol, opaclump, emitprob, emit, &
         sthermal) bind(C, name="slgpu_ddmc_prep")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: radtemp(*)
      real(c_double), value :: lbol
      type(c_ptr), value :: opaclump, emitprob, emit, sthermal
    end function

    integer(c_int) function slgpu_sources(ctx, nstot, iteration, nvol, nnew) &
         bind(C, name="slgpu_sources")
      import :: c_int, c_ptr, c_int64_t, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nstot
      integer(c_int32_t), value :: iteration
      type(c_ptr), value :: nvol
      integer(c_int64_t), intent(out) :: nnew
    end function

    integer(c_int) function slgpu_transport(ctx, iteration, tally, jrad, eamp, totals, counters) &
         bind(C, name="slgpu_transport")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: iteration
      type(c_ptr), value :: tally, jrad, eamp, totals, counters
    end function

    integer(c_int) function slgpu_fluxtally(ctx, luminos, lumdev, lumnum) &
         bind(C, name="slgpu_fluxtally")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: luminos, lumdev, lumnum
    end function

    integer(c_int) function slgpu_get_particles(ctx, n, stat, x, mu, wl, e, ig, nstep, err) &
         bind(C, name="slgpu_get_particles")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      type(c_ptr), value :: stat, x, mu, wl, e, ig, nstep, err
    end function

    ! differentiation form: adjoint of the last slgpu_opacity_build (any pointer may be
    ! c_null_ptr): gradients in, gradients w.r.t. T, natom, n_e, natomfr, oc out
    integer(c_int) function slgpu_opacity_vjp(ctx, grad_cap64, grad_capgrey, grad_capross, &
         grad_temp, grad_natom, grad_nelec, grad_natomfr, grad_oc) &
         bind(C, name="slgpu_opacity_vjp")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: grad_cap64, grad_capgrey, grad_capross
      type(c_ptr), value :: grad_temp, grad_natom, grad_nelec, grad_natomfr, grad_oc
    end function

    ! the same with the three input gradients as device arrays on the context's GPU
    integer(c_int) function slgpu_opacity_vjp_dev(ctx, d_grad_cap64, d_grad_capgrey, &
         d_grad_capross, grad_temp, grad_natom, grad_nelec, grad_natomfr, grad_oc) &
         bind(C, name="slgpu_opacity_vjp_dev")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: d_grad_cap64, d_grad_capgrey, d_grad_capross
      type(c_ptr), value :: grad_temp, grad_natom, grad_nelec, grad_natomfr, grad_oc
    end function

    ! device copies of the last build's cap64, capgrey, capross
    integer(c_int) function slgpu_opacity_outputs_dev(ctx, d_cap64, d_capgrey, d_capross) &
         bind(C, name="slgpu_opacity_outputs_dev")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: d_cap64, d_capgrey, d_capross
    end function

    integer(c_int) function slgpu_set_unstaged(ctx, on) bind(C, name="slgpu_set_unstaged")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function

    integer(c_int) function slgpu_sample_groups(ctx, table, direct, n, r, cell, group) &
         bind(C, name="slgpu_sample_groups")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: table, direct
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: r(*)
      integer(c_int), intent(in) :: cell(*)
      integer(c_int), intent(out) :: group(*)
    end function

    integer(c_int) function slgpu_set_eos_data(ctx, ni, e, q, g1, amass) &
         bind(C, name="slgpu_set_eos_data")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), intent(in) :: ni(*)
      real(c_double), intent(in) :: e(*), q(*), g1(*), amass(*)
    end function

    integer(c_int) function slgpu_eos(ctx, temp, mass, massfr, acc, nelec, natom, natomfr, &
         oc, ginv, niter) bind(C, name="slgpu_eos")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: temp(*), mass(*), massfr(*)
      real(c_double), value :: acc
      real(c_double), intent(inout) :: nelec(*)
      type(c_ptr), value :: natom, natomfr, oc, ginv, niter
    end function

    integer(c_int) function slgpu_eos_vjp(ctx, grad_natom, grad_natomfr, grad_nelec, grad_oc, &
         grad_temp, grad_mass, grad_massfr) bind(C, name="slgpu_eos_vjp")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: grad_natom, grad_natomfr, grad_nelec, grad_oc
      type(c_ptr), value :: grad_temp, grad_mass, grad_massfr
    end function

    integer(c_int) function slgpu_timers(ctx, ms, n, reset) bind(C, name="slgpu_timers")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ms(*)
      integer(c_int), value :: n, reset
    end function

    type(c_ptr) function slgpu_stream(ctx) bind(C, name="slgpu_stream")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function slgpu_peak_rates(exp32, exp64) bind(C, name="slgpu_peak_rates")
      import :: c_int, c_double
      real(c_double), intent(out) :: exp32, exp64
    end function
  end interface

contains

  !-- Fortran string of the last error
  function slgpu_error_string() result(s)
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: p(:)
    integer :: n
    type(c_ptr) :: cp
    cp = slgpu_last_error_message()
    call c_f_pointer(cp, p, [4096])
    n = 0
    do while (n < 4096)
      if (p(n+1) == c_null_char) exit
      n = n + 1
    enddo
    allocate(character(len=n) :: s)
    s = transfer(p(1:n), s)
  end function

end module slgpu_iface
