! Test-infrastructure harness around the reference SuperLite hot path (compiled from the
! reference's own sources by oracle/Makefile into oracle/_ref/libslref.so).
!
! Every routine below only (a) copies caller arrays into the reference's module state,
! (b) calls the reference's own subroutines in the order superlite.f90:148-219 uses, and
! (c) copies module state back out.  No physics is restated here.
!
! Used only by tests/ (and by tests/golden/make_golden.py) as a checker of the oracle.
module slref_harness
  use iso_c_binding
  implicit none
contains

  !-- load the reference's atomic tables from the current directory
  !   (ions_read_data superlite.f90:69, bfxs_read_data :79, ffxs_read_data :81)
  subroutine slref_read_tables() bind(C, name="slref_read_tables")
    use gasmod, only: gas_nelem
    use ionsmod, only: ions_read_data
    use bfxsmod, only: bfxs_read_data
    use ffxsmod, only: ffxs_read_data
    use timingmod, only: timingmod_init
    use countersmod, only: countersmod_init
    call timingmod_init
    call countersmod_init
    call ions_read_data(gas_nelem)
    call bfxs_read_data
    call ffxs_read_data
  end subroutine

  subroutine slref_get_bf(ph1, ph2) bind(C, name="slref_get_bf")
    use bfxsmod, only: bf_ph1, bf_ph2
    real(c_float), intent(out) :: ph1(6,7,30,30), ph2(7,30,30)
    ph1 = bf_ph1
    ph2 = bf_ph2
  end subroutine

  subroutine slref_get_ff(gff) bind(C, name="slref_get_ff")
    use ffxsmod, only: ff_gff
    real(c_double), intent(out) :: gff(81,41)
    gff = ff_gff
  end subroutine

  subroutine slref_get_ni(ni) bind(C, name="slref_get_ni")
    use ionsmod, only: ion_el
    integer(c_int), intent(out) :: ni(30)
    integer :: iz
    do iz = 1, 30
      ni(iz) = ion_el(iz)%ni
    enddo
  end subroutine

  !-- grid, groups, flux grid, transport pointers, particle store, rng
  !   (mirrors superlite.f90:87-121 without file input)
  subroutine slref_setup(nx, ng, wlmin, wlmax, xleft, vxleft, n2s, nomp, &
       tauddmc, taulump, puretran, errorfatal, lbol, ampfact, thm_aniso) bind(C, name="slref_setup")
    use inputparmod
    use inputstrmod, only: str_xleft, str_yleft, str_zleft, str_vxleft, str_vyleft, &
         str_vzleft, str_idcell, str_nc
    use physconstmod, only: pc_pi2
    use groupmod, only: groupmod_init
    use gridmod
    use gasmod, only: gasmod_init, gas_nelem, gas_rcell
    use fluxmod, only: fluxgrid_setup
    use transportmod, only: transportmod_init
    use sourcemod, only: sourcemod_init
    use ionsmod, only: ions_alloc_grndlev
    use particlemod, only: particle_alloc
    use randommod, only: rnd_init, rnd_state, rnd_states
!$  use omp_lib, only: omp_set_num_threads
    integer(c_int), value :: nx, ng, n2s, nomp, puretran, errorfatal, ampfact, thm_aniso
    real(c_double), value :: wlmin, wlmax, tauddmc, taulump, lbol
    real(c_double), intent(in) :: xleft(nx+1), vxleft(nx+1)
    integer :: i
!$  call omp_set_num_threads(nomp)
    in_ndim = [nx, 1, 1]
    in_grp_ng = ng
    in_grp_ngs = 1
    in_grp_wlmin = wlmin
    in_grp_wlmax = wlmax
    in_src_n2s = n2s
    in_src_ns = 0
    in_nomp = nomp
    in_trn_tauddmc = tauddmc
    in_taulump = taulump
    in_puretran = puretran /= 0
    in_trn_errorfatal = errorfatal /= 0
    in_trn_noamp = ampfact == 0
    in_thm_aniso = thm_aniso /= 0
    in_L_bol = lbol
    in_flx_ndim = [0, 1, 1]
    call provide_inputpars(1)
    allocate(str_xleft(nx+1), str_yleft(2), str_zleft(2))
    allocate(str_vxleft(nx+1), str_vyleft(2), str_vzleft(2))
    allocate(str_idcell(nx))
    str_xleft = xleft
    str_vxleft = vxleft
    str_yleft = [-1d0, 1d0]
    str_zleft = [0d0, pc_pi2]
    str_vyleft = 0d0
    str_vzleft = 0d0
    str_nc = nx
    forall(i=1:nx) str_idcell(i) = i
    call groupmod_init
    call fluxgrid_setup
    call gridmod_init(.false., ng, nx, 1, nx)
    call grid_setup
    call gasmod_init(.false., 1, nx, ng)
    call transportmod_init(11)
    call sourcemod_init(1)
    call ions_alloc_grndlev(gas_nelem, nx)
    call particle_alloc(.false.)
    call rnd_init(nomp, 0)
    rnd_state = rnd_states(1)
    call grid_volume(11)
    gas_rcell = grd_rcell
  end subroutine

  !-- group edges supplied by the caller (so both sides use bit-identical grp_wl);
  !   grp_wlinv = 1/grp_wl as groupmod.f:50-51, flux grid = transport grid (fluxmod.f:132-136)
  subroutine slref_put_groups(ng, wl) bind(C, name="slref_put_groups")
    use groupmod, only: grp_wl, grp_wlinv
    use fluxmod, only: flx_wl
    integer(c_int), value :: ng
    real(c_double), intent(in) :: wl(ng+1)
    grp_wl = wl
    grp_wlinv = 1d0/grp_wl
    flx_wl = grp_wl
  end subroutine

  subroutine slref_get_grid(vol, rcell, divv, nep, nepg) bind(C, name="slref_get_grid")
    use gridmod
    real(c_double), intent(out) :: vol(grd_ncell), rcell(grd_ncell), divv(grd_ncell)
    integer(c_int), intent(out) :: nep, nepg
    vol = grd_vol
    rcell = grd_rcell
    divv = grd_divv
    nep = grd_nep
    nepg = grd_nepg
  end subroutine

  !-- gas state from mass, temperature, mass fractions (gas_setup.f:129-200 massfr2natomfr)
  subroutine slref_set_gas(nc, mass, temp, massfr) bind(C, name="slref_set_gas")
    use gasmod
    use gridmod, only: grd_vol, grd_radtemp
    integer(c_int), value :: nc
    real(c_double), intent(in) :: mass(nc), temp(nc), massfr(30, nc)
    real*8 :: mf(0:gas_nelem, nc)
    gas_mass = mass
    gas_temp = temp
    gas_radtemp = temp
    grd_radtemp = temp
    gas_vol = grd_vol
    gas_rho = gas_mass/gas_vol
    mf = 0d0
    mf(1:30, :) = massfr
    call massfr2natomfr(mf)
  end subroutine

  !-- LTE EOS (eos_update.f:23-45) and readback of its results
  subroutine slref_eos() bind(C, name="slref_eos")
    call eos_update(.false.)
  end subroutine

  subroutine slref_get_gas(nc, natom, natomfr, nelec, vol) bind(C, name="slref_get_gas")
    use gasmod
    integer(c_int), value :: nc
    real(c_double), intent(out) :: natom(nc), natomfr(30, nc), nelec(nc), vol(nc)
    natom = gas_natom
    natomfr = gas_natomfr(1:30, :)
    nelec = gas_nelec
    vol = gas_vol
  end subroutine

  !-- the LTE EOS data the host holds after its first eos_update: per element the number of
  !   ionisation stages ni (incl. the bare nucleus), per stage the ionisation energy e [erg],
  !   the partition function q frozen at the first call (ionsmod.f:208-215) and the ground
  !   level's statistical weight glev(1); and the element masses [amu] (elemdatamod.f)
  subroutine slref_get_eos_data(ni, e, q, g1, am) bind(C, name="slref_get_eos_data")
    use ionsmod, only: ion_el
    use elemdatamod, only: elem_data
    integer(c_int), intent(out) :: ni(30)
    real(c_double), intent(out) :: e(31, 30), q(31, 30), g1(31, 30), am(30)
    integer :: iz, ii
    e = 0d0
    q = 0d0
    g1 = 0d0
    do iz = 1, 30
      ni(iz) = ion_el(iz)%ni
      am(iz) = elem_data(iz)%m
      do ii = 1, ion_el(iz)%ni
        e(ii, iz) = ion_el(iz)%i(ii)%e
        q(ii, iz) = ion_el(iz)%i(ii)%q
        if (allocated(ion_el(iz)%i(ii)%glev)) then
          if (size(ion_el(iz)%i(ii)%glev) > 0) g1(ii, iz) = ion_el(iz)%i(ii)%glev(1)
        endif
      enddo
    enddo
  end subroutine

  !-- ground-level populations of the listed ions, (nion, nc) ion-fastest
  subroutine slref_get_pops(nc, nion, ion_iz, ion_ii, oc, ginv) bind(C, name="slref_get_pops")
    use ionsmod, only: ion_grndlev
    integer(c_int), value :: nc, nion
    integer(c_int), intent(in) :: ion_iz(nion), ion_ii(nion)
    real(c_double), intent(out) :: oc(nion, nc), ginv(nion, nc)
    integer :: i, k
    do i = 1, nc
      do k = 1, nion
        oc(k, i) = ion_grndlev(ion_iz(k), i)%oc(ion_ii(k))
        ginv(k, i) = ion_grndlev(ion_iz(k), i)%ginv(ion_ii(k))
      enddo
    enddo
  end subroutine

  !-- overwrite gas state directly (used for small hand-built cases)
  subroutine slref_put_gas(nc, mass, temp, natom, natomfr, nelec, nion, ion_iz, ion_ii, oc, ginv) &
       bind(C, name="slref_put_gas")
    use gasmod
    use gridmod, only: grd_vol, grd_radtemp
    use ionsmod, only: ion_grndlev
    integer(c_int), value :: nc, nion
    real(c_double), intent(in) :: mass(nc), temp(nc), natom(nc), natomfr(30, nc), nelec(nc)
    integer(c_int), intent(in) :: ion_iz(nion), ion_ii(nion)
    real(c_double), intent(in) :: oc(nion, nc), ginv(nion, nc)
    integer :: i, k, iz
    gas_mass = mass
    gas_temp = temp
    gas_radtemp = temp
    grd_radtemp = temp
    gas_vol = grd_vol
    gas_rho = gas_mass/gas_vol
    gas_natom = natom
    gas_natomfr = 0d0
    gas_natomfr(1:30, :) = natomfr
    gas_nelec = nelec
    do i = 1, nc
      do iz = 1, 30
        ion_grndlev(iz, i)%oc = 0d0
        ion_grndlev(iz, i)%ginv = 1d0
      enddo
      do k = 1, nion
        ion_grndlev(ion_iz(k), i)%oc(ion_ii(k)) = oc(k, i)
        ion_grndlev(ion_iz(k), i)%ginv(ion_ii(k)) = ginv(k, i)
      enddo
    enddo
  end subroutine

  !-- line list already lambda-sorted (bypasses read_bbxs_data's file reader and sort)
  subroutine slref_set_lines(n, wl0, gxs, chilw, iz, ii) bind(C, name="slref_set_lines")
    use bbxsmod, only: bb_xs, bb_nline
    integer(c_int), value :: n
    real(c_float), intent(in) :: wl0(n), gxs(n), chilw(n)
    integer(c_short), intent(in) :: iz(n), ii(n)
    integer :: l
    if(allocated(bb_xs)) deallocate(bb_xs)
    allocate(bb_xs(n))
    do l = 1, n
      bb_xs(l)%wl0 = wl0(l)
      bb_xs(l)%gxs = gxs(l)
      bb_xs(l)%chilw = chilw(l)
      bb_xs(l)%iz = iz(l)
      bb_xs(l)%ii = ii(l)
    enddo
    bb_nline = n
  end subroutine

  !-- physical_opacity + means (gas_update.f:95-100); flags bit0 no bb, bit1 no bf, bit2 no ff, bit3 no thomson
  subroutine slref_opacity(flags) bind(C, name="slref_opacity")
    use inputparmod, only: in_nobbopac, in_nobfopac, in_noffopac, in_nothmson
    integer(c_int), value :: flags
    in_nobbopac = iand(flags, 1) /= 0
    in_nobfopac = iand(flags, 2) /= 0
    in_noffopac = iand(flags, 4) /= 0
    in_nothmson = iand(flags, 8) /= 0
    call physical_opacity
    call opacity_planckmean
    call opacity_rossmean
  end subroutine

  subroutine slref_get_opacity(nc, ng, cap, sig, capgrey, capross, emiss) bind(C, name="slref_get_opacity")
    use gasmod
    integer(c_int), value :: nc, ng
    real(c_float), intent(out) :: cap(ng, nc), emiss(ng, nc)
    real(c_double), intent(out) :: sig(nc), capgrey(nc), capross(nc)
    cap = gas_cap
    emiss = gas_emiss
    sig = gas_sig
    capgrey = gas_capgrey
    capross = gas_capross
  end subroutine

  !-- overwrite the FP32 group opacity and means (decoupled tests of the DDMC prep)
  subroutine slref_put_opacity(nc, ng, cap, sig, capgrey, capross) bind(C, name="slref_put_opacity")
    use gasmod
    integer(c_int), value :: nc, ng
    real(c_float), intent(in) :: cap(ng, nc)
    real(c_double), intent(in) :: sig(nc), capgrey(nc), capross(nc)
    gas_cap = cap
    gas_sig = sig
    gas_capgrey = capgrey
    gas_capross = capross
  end subroutine

  !-- bcast_nonpermanent (serial copies), sourceenergy, leakage_opacity, emission_probability
  !   (superlite.f90:163-178)
  subroutine slref_prep() bind(C, name="slref_prep")
    use gasmod, only: gas_emit, gas_capemit, gas_capemitgrey, gas_jrad
    use inputparmod, only: in_nlte
    use mpimod, only: bcast_nonpermanent
    gas_emit = 0d0
    if(.not.in_nlte) then
      gas_capemit = 0.
      gas_capemitgrey = 0d0
    endif
    gas_jrad = 0d0
    call bcast_nonpermanent
    call sourceenergy
    call leakage_opacity
    call emission_probability
  end subroutine

  !-- in_nlte and the NLTE emission opacity gas_capemit / gas_capemitgrey that the next
  !   slref_prep broadcasts (flag 0: back to LTE)
  subroutine slref_set_nlte(flag, nc, ng, capemit, capemitgrey) bind(C, name="slref_set_nlte")
    use gasmod, only: gas_capemit, gas_capemitgrey
    use inputparmod, only: in_nlte
    integer(c_int), value :: flag, nc, ng
    real(c_float), intent(in) :: capemit(ng, nc)
    real(c_double), intent(in) :: capemitgrey(nc)
    in_nlte = flag /= 0
    if(in_nlte) then
      gas_capemit = capemit
      gas_capemitgrey = capemitgrey
    endif
  end subroutine

  !-- nlte_rad_est (superlite.f90:208) on a given grd_tally and grd_radtemp; the module's
  !   grd_tally, grd_radtemp and grd_jrad are restored afterwards
  subroutine slref_nlte_rad_est(nc, ng, tally, radtemp, jrad) bind(C, name="slref_nlte_rad_est")
    use gridmod, only: grd_tally, grd_radtemp, grd_jrad
    use nltemod, only: nlte_rad_est
    integer(c_int), value :: nc, ng
    real(c_double), intent(in) :: tally(nc)
    real(c_double), intent(inout) :: radtemp(nc)
    real(c_double), intent(out) :: jrad(ng, nc)
    real(c_double) :: t0(nc), r0(nc), j0(ng, nc)
    t0 = grd_tally
    r0 = grd_radtemp
    j0 = grd_jrad
    grd_tally = tally
    grd_radtemp = radtemp
    call nlte_rad_est
    radtemp = grd_radtemp
    jrad = grd_jrad
    grd_tally = t0
    grd_radtemp = r0
    grd_jrad = j0
  end subroutine

  subroutine slref_get_prep(nc, nep, opaclump, emitprob, emit) bind(C, name="slref_get_prep")
    use gridmod
    integer(c_int), value :: nc, nep
    real(c_double), intent(out) :: opaclump(11, nc), emitprob(nep, nc), emit(nc)
    opaclump = grd_opaclump
    emitprob = grd_emitprob
    emit = grd_emit
  end subroutine

  !-- sourcenumbers (superlite.f90:182-183) for a given grd_emit and particle count, one rank
  !   with every slot vacant; returns grd_nvol and src_nnew
  subroutine slref_sourcenumbers(nc, emit, ns, nvol, nnew) bind(C, name="slref_sourcenumbers")
    use gridmod, only: grd_emit, grd_nvol
    use sourcemod, only: src_ns, src_nvacantall, src_nnew
    integer(c_int), value :: nc, ns
    real(c_double), intent(in) :: emit(nc)
    integer(c_int), intent(out) :: nvol(nc), nnew
    grd_emit = emit
    src_ns = ns
    src_nvacantall(1) = int(ns, 8)
    call sourcenumbers
    nvol = grd_nvol
    nnew = src_nnew
  end subroutine

  function slref_emitgroup(r, ic) result(ig) bind(C, name="slref_emitgroup")
    use gridmod, only: emitgroup
    real(c_double), value :: r
    integer(c_int), value :: ic
    integer(c_int) :: ig
    ig = emitgroup(r, ic)
  end function

  function slref_binsrch(x, arr, n, widerange) result(i) bind(C, name="slref_binsrch")
    use miscmod, only: binsrch
    real(c_double), value :: x
    integer(c_int), value :: n, widerange
    real(c_double), intent(in) :: arr(n)
    integer(c_int) :: i
    i = binsrch(x, arr, n, widerange /= 0)
  end function

  function slref_bfxs(nz, ne, e) result(xs) bind(C, name="slref_bfxs")
    use bfxsmod, only: bfxs
    integer(c_int), value :: nz, ne
    real(c_double), value :: e
    real(c_double) :: xs
    xs = bfxs(nz, ne, e)
  end function

  subroutine slref_specintv(tempinv, n, ss, offset) bind(C, name="slref_specintv")
    use groupmod, only: specintv
    real(c_double), value :: tempinv
    integer(c_int), value :: n, offset
    real(c_double), intent(out) :: ss(n)
    call specintv(tempinv, n, ss, offset)
  end subroutine

  subroutine slref_spectintv(tempinv, n, ss) bind(C, name="slref_spectintv")
    use groupmod, only: spectintv
    real(c_double), value :: tempinv
    integer(c_int), value :: n
    real(c_double), intent(out) :: ss(n)
    call spectintv(tempinv, n, ss)
  end subroutine

  !-- one Monte Carlo iteration after the prep (superlite.f90:182-219): sources, transport, flux tally.
  !   out_tot = [eout, evelo, sflux, sthermal]; out_ct = [nnew, nactive, nstepimc, nstepddmc,
  !   nstepmax, nmethswap, nflux, 0]
  subroutine slref_iteration(nc, ng, tally, jrad, eamp, luminos, lumdev, lumnum, out_tot, out_ct) &
       bind(C, name="slref_iteration")
    use gridmod
    use sourcemod
    use particlemod, only: prt_isvacant
    use fluxmod
    use totalsmod
    use countersmod
    use timingmod
    integer(c_int), value :: nc, ng
    real(c_double), intent(out) :: tally(nc), jrad(ng, nc), eamp(nc), luminos(ng), lumdev(ng)
    integer(c_int), intent(out) :: lumnum(ng)
    real(c_double), intent(out) :: out_tot(4)
    integer(c_long_long), intent(out) :: out_ct(8)
    tot_eout = 0d0
    tot_evelo = 0d0
    tot_sflux = 0d0
    src_nvacantall(1) = count(prt_isvacant)
    call sourcenumbers
    if(src_nnew > 0) then
      allocate(src_ivacant(src_nnew))
      call vacancies
      call boundary_source
      call interior_source
      call source_transformdirection
      deallocate(src_ivacant)
    endif
    call particle_advance
    call fluxtally
    tally = grd_tally
    jrad = grd_jrad
    eamp = grd_eamp
    luminos = flx_luminos(:, 1, 1)
    lumdev = flx_lumdev(:, 1, 1)
    lumnum = flx_lumnum(:, 1, 1)
    out_tot = [tot_eout, tot_evelo, tot_sflux, tot_sthermal]
    out_ct = [int(src_nnew, 8), ct_npactive(1), ct_npstepimc(1), ct_npstepddmc(1), &
         ct_npstepmax(1), ct_npmethswap(1), ct_npflux(1), 0_8]
  end subroutine

  !-- the reference's namelist reader (inputparmod.f:216-231) on ./input.par, and the values
  !   of the parameters the hot path consumes
  subroutine slref_read_par(ndim, flxndim, ivals, rvals, lvals) bind(C, name="slref_read_par")
    use inputparmod
    integer(c_int), intent(out) :: ndim(3), flxndim(3), ivals(6)
    real(c_double), intent(out) :: rvals(9)
    integer(c_int), intent(out) :: lvals(10)
    call read_inputpars
    ndim = in_ndim
    flxndim = in_flx_ndim
    ivals = [in_grp_ng, in_src_n2s, in_src_ns, in_niter, in_nomp, in_grd_igeom]
    rvals = [in_grp_wlmin, in_grp_wlmax, in_flx_wlmin, in_flx_wlmax, in_L_bol, &
         in_trn_tauddmc, in_taulump, in_srcepwr, in_L_inner]
    lvals = merge(1, 0, [in_puretran, in_trn_errorfatal, in_trn_noamp, in_thm_aniso, &
         in_nobbopac, in_nobfopac, in_noffopac, in_nothmson, in_nlte, in_noeos])
  end subroutine

  !-- the reference's structure-file reader (inputstrmod.f:85-274) on ./input.str, 1D spherical
  subroutine slref_read_str(nx, nabund, flags) bind(C, name="slref_read_str")
    use inputstrmod
    integer(c_int), value :: nx
    integer(c_int), intent(out) :: nabund, flags(4)
    call read_inputstr(11, [nx, 1, 1], 1)
    nabund = str_nabund
    flags = merge(1, 0, [str_lvol, str_lradtemp, str_llum, str_lye])
  end subroutine

  subroutine slref_get_str(nx, nabund, xleft, vxleft, mass, temp, massfr, iabund, vol, radtemp, &
       lum, ye) bind(C, name="slref_get_str")
    use inputstrmod
    integer(c_int), value :: nx, nabund
    real(c_double), intent(out) :: xleft(nx+1), vxleft(nx+1), mass(nx), temp(nx), &
         massfr(nabund, nx), vol(nx), radtemp(nx), lum(nx), ye(nx)
    integer(c_int), intent(out) :: iabund(nabund)
    xleft = str_xleft
    vxleft = str_vxleft
    mass = str_mass(:, 1, 1)
    temp = str_temp(:, 1, 1)
    massfr = str_massfr(:, :, 1, 1)
    iabund = str_iabund
    vol = 0d0
    radtemp = 0d0
    lum = 0d0
    ye = 0d0
    if (str_lvol) vol = str_vol(:, 1, 1)
    if (str_lradtemp) radtemp = str_radtemp(:, 1, 1)
    if (str_llum) lum = str_lum(:, 1, 1)
    if (str_lye) ye = str_ye(:, 1, 1)
  end subroutine

  !-- the reference's writers on given values, into the current directory:
  !   output_flux (OUTPUT/output_flux.f:3-66) and output_grid (OUTPUT/output_grid.f:3-170),
  !   one iteration per call (the first call writes the headers and the grid files)
  subroutine slref_write_outputs(nc, ng, luminos, lumdev, lumnum, nvol, tempinv, radtemp, &
       capgrey, capross, sig, tally, vol, tot, opacdump) bind(C, name="slref_write_outputs")
    use fluxmod, only: flx_luminos, flx_lumdev, flx_lumnum
    use gridmod
    use totalsmod
    use inputparmod, only: in_io_opacdump, in_io_nogriddump
    integer(c_int), value :: nc, ng, opacdump
    real(c_double), intent(in) :: luminos(ng), lumdev(ng), tempinv(nc), radtemp(nc), capgrey(nc), &
         capross(nc), sig(nc), tally(nc), vol(nc), tot(4)
    integer(c_int), intent(in) :: lumnum(ng), nvol(nc)
    flx_luminos(:, 1, 1) = luminos
    flx_lumdev(:, 1, 1) = lumdev
    flx_lumnum(:, 1, 1) = lumnum
    grd_nvol = nvol
    grd_tempinv = tempinv
    grd_radtemp = radtemp
    grd_capgrey = capgrey
    grd_capemitgrey = 0d0
    grd_capross = capross
    grd_sig = sig
    grd_tally = tally
    grd_vol = vol
    tot_eout = tot(1)
    tot_evelo = tot(2)
    tot_sflux = tot(3)
    tot_sthermal = tot(4)
    in_io_nogriddump = .false.
    in_io_opacdump = merge('one ', 'off ', opacdump /= 0)
    call output_flux
    call output_grid
  end subroutine

end module slref_harness
