! A Fortran host of libslgpu.so through superlite_amd/fortran/slgpu_iface.f90, the binding a
! patched SuperLite uses (superlite.f90:87-131 set-up, then one iteration of :157-219):
! create a context, upload mesh, groups, ions, atomic tables and the line list, then run
! the opacity build, the DDMC preparation, the sources, the transport and the flux tally.
!
! Usage: slgpu_e2e <input file> <output file>
! The input file (stream, native byte order) holds the workload the test wrote; the output
! file receives every result of the iteration in the order written below.
program slgpu_e2e
  use iso_c_binding
  use slgpu_iface
  implicit none
  character(len=4096) :: fin, fout
  integer :: u, nx, ng, nep, nepg, nion, nline, nstot, iter, seed, prec, puretran, errfatal
  integer(c_int) :: hdr(12), rc
  real(c_double) :: tauddmc, taulump, lbol
  real(c_double), allocatable, target :: xarr(:), vxarr(:), divv(:), vol(:), rcell(:), wl(:)
  integer(c_int), allocatable :: ion_iz(:), ion_ii(:)
  real(c_float), allocatable :: ph1(:), ph2(:), wl0(:), gxs(:), chilw(:)
  real(c_double), allocatable :: gff(:)
  integer(c_int16_t), allocatable :: liz(:), lii(:)
  real(c_double), allocatable, target :: temp(:), natom(:), gvol(:), mass(:), nelec(:), &
       natomfr(:), oc(:), ginv(:), radtemp(:)
  real(c_float), allocatable, target :: cap(:), emiss(:)
  real(c_double), allocatable, target :: sig(:), capgrey(:), capross(:), opaclump(:), &
       emitprob(:), emit(:), tally(:), jrad(:), eamp(:), luminos(:), lumdev(:)
  integer(c_int), allocatable, target :: nvol(:), lumnum(:)
  real(c_double), target :: totals(7), sthermal
  integer(c_int64_t), target :: counters(10), nnew
  type(c_ptr) :: ctx
  type(slgpu_config) :: cfg
  type(slgpu_gas) :: gas

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  nx = hdr(1); ng = hdr(2); nep = hdr(3); nepg = hdr(4); nion = hdr(5); nline = hdr(6)
  nstot = hdr(7); iter = hdr(8); seed = hdr(9); prec = hdr(10); puretran = hdr(11)
  errfatal = hdr(12)
  read(u) tauddmc, taulump, lbol
  allocate(xarr(nx+1), vxarr(nx+1), divv(nx), vol(nx), rcell(nx), wl(ng+1))
  allocate(ion_iz(nion), ion_ii(nion), ph1(6*7*30*30), ph2(7*30*30), gff(81*41))
  allocate(wl0(nline), gxs(nline), chilw(nline), liz(nline), lii(nline))
  allocate(temp(nx), natom(nx), gvol(nx), mass(nx), nelec(nx), natomfr(30*nx), &
       oc(max(nion,1)*nx), ginv(max(nion,1)*nx), radtemp(nx))
  read(u) xarr, vxarr, divv, vol, rcell, wl, ion_iz, ion_ii, ph1, ph2, gff
  read(u) wl0, gxs, chilw, liz, lii
  read(u) temp, natom, gvol, mass, nelec, natomfr, oc, ginv, radtemp
  close(u)

  cfg = slgpu_config(nx, ng, nep, nepg, tauddmc, taulump, puretran, errfatal, prec, &
       int(seed, c_int32_t), int(nstot, c_int64_t), 0_c_int64_t, 0, 0)
  call check(slgpu_create(cfg, ctx), 'slgpu_create')
  call check(slgpu_set_grid(ctx, xarr, vxarr, divv, vol, rcell), 'slgpu_set_grid')
  call check(slgpu_set_groups(ctx, wl), 'slgpu_set_groups')
  call check(slgpu_set_ions(ctx, int(nion, c_int), ion_iz, ion_ii), 'slgpu_set_ions')
  call check(slgpu_set_bf_tables(ctx, ph1, ph2), 'slgpu_set_bf_tables')
  call check(slgpu_set_ff_table(ctx, gff), 'slgpu_set_ff_table')
  call check(slgpu_set_lines(ctx, int(nline, c_int64_t), wl0, gxs, chilw, liz, lii), 'slgpu_set_lines')

  allocate(cap(ng*nx), emiss(ng*nx), sig(nx), capgrey(nx), capross(nx), opaclump(11*nx), &
       emitprob(nep*nx), emit(nx), nvol(nx), tally(nx), jrad(ng*nx), eamp(nx), &
       luminos(ng), lumdev(ng), lumnum(ng))
  gas = slgpu_gas(c_loc(temp), c_loc(natom), c_loc(gvol), c_loc(mass), c_loc(nelec), &
       c_loc(natomfr), c_loc(oc), c_loc(ginv))
  ! physical_opacity + opacity_planckmean + opacity_rossmean
  call check(slgpu_opacity_build(ctx, gas, 0_c_int, c_null_ptr, c_loc(cap), c_loc(sig), &
       c_loc(capgrey), c_loc(capross), c_loc(emiss)), 'slgpu_opacity_build')
  ! sourceenergy + leakage_opacity + emission_probability
  call check(slgpu_ddmc_prep(ctx, radtemp, lbol, c_loc(opaclump), c_loc(emitprob), &
       c_loc(emit), c_loc(sthermal)), 'slgpu_ddmc_prep')
  ! sourcenumbers .. source_transformdirection, particle_advance, fluxtally
  call check(slgpu_sources(ctx, int(nstot, c_int64_t), int(iter, c_int32_t), c_loc(nvol), nnew), &
       'slgpu_sources')
  call check(slgpu_transport(ctx, int(iter, c_int32_t), c_loc(tally), c_loc(jrad), c_loc(eamp), &
       c_loc(totals), c_loc(counters)), 'slgpu_transport')
  call check(slgpu_fluxtally(ctx, c_loc(luminos), c_loc(lumdev), c_loc(lumnum)), 'slgpu_fluxtally')
  call check(slgpu_destroy(ctx), 'slgpu_destroy')

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) cap, emiss, sig, capgrey, capross, opaclump, emitprob, emit, sthermal, nvol, nnew
  write(u) tally, jrad, eamp, totals, counters, luminos, lumdev, lumnum
  close(u)
  print '(a, i0, a, i0)', 'slgpu_e2e ok: particles ', nnew, ', particle-steps ', &
       counters(2) + counters(3)

contains

  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(*), intent(in) :: what
    if (code /= SLGPU_OK) then
      print '(a, a, i0, a, a)', what, ' failed: ', code, ' ', slgpu_error_string()
      stop 1
    endif
  end subroutine

end program slgpu_e2e
