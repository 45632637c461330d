!  ramgpu_driver.f90 -- a Fortran host driving the MI355X split step through
!  ramgpu_mod, the way the reference's ram_run would: module-style arrays in
!  Fortran order, one context, per-species set-up, then nsteps ram_run calls
!  and the block after ram_run's species loop (ANISCH of every species, FLUX).
!
!  Usage: ramgpu_driver <input.bin> <output.bin>
!  input.bin (stream, native endian): int32 ns,nr,nt,ne,npa,nsteps; then the
!  scalars dts,betalim,fraccfl,mdr,dphi,conf1,conf2; int32 charge(ns),
!  cex(ns); real64 rmas(ns), rlz(nr+1), ekev,ebnd,we,de(ne), grel,grbnd,v(ns,ne),
!  mu,wmu,dmu(npa), upa(nr), fnhs,fnis,bounhs,bounis,hdns,didt,dibndt(nr+1,nt,npa),
!  bnes,dbdt,vt,eir,eip(nr+1,nt); int32 mgnp(nr,nt); real64 fgeos(ns,nt,ne,npa),
!  f2(ns,nr,nt,ne,npa), ffactor(ns,nr,ne,npa), ernh,epp(ns,ne).
!  output.bin: f2 after the steps and ANISCH, setrc,lsdr,lscha,lsatm(ns),
!  ppert,ppart,edent,rnhtt(ns,nr,nt), flux(ns,nr,nt,ne,npa).
program ramgpu_driver
  use iso_c_binding
  use ramgpu_mod
  implicit none
  type(c_ptr) :: ctx
  integer(c_int) :: ns, nr, nt, ne, npa, nsteps, rc, s
  real(c_double) :: sc(7)
  integer(c_int), allocatable :: charge(:), cex(:), mgnp(:,:)
  real(c_double), allocatable :: rmas(:), rlz(:), ekev(:), ebnd(:), we(:), de(:)
  real(c_double), allocatable :: grel(:,:), grbnd(:,:), v(:,:), mu(:), wmu(:), dmu(:), upa(:)
  real(c_double), allocatable :: fnhs(:,:,:), fnis(:,:,:), bounhs(:,:,:), bounis(:,:,:), &
       hdns(:,:,:), didt(:,:,:), dibndt(:,:,:)
  real(c_double), allocatable :: bnes(:,:), dbdt(:,:), vt(:,:), eir(:,:), eip(:,:)
  real(c_double), allocatable :: fgeos(:,:,:,:), f2(:,:,:,:,:)
  real(c_double), allocatable :: setrc(:), lsdr(:), lscha(:), lsatm(:)
  real(c_double), allocatable :: ffactor(:,:,:,:), ernh(:,:), epp(:,:), flux(:,:,:,:,:)
  real(c_double), allocatable :: ppert(:,:,:), ppart(:,:,:), edent(:,:,:), rnhtt(:,:,:)
  type(ramgpu_species_out), allocatable :: out(:)
  character(len=512) :: fin, fout
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access="stream", form="unformatted", status="old")
  read(u) ns, nr, nt, ne, npa, nsteps
  read(u) sc
  allocate(charge(ns), cex(ns), rmas(ns), rlz(nr+1), ekev(ne), ebnd(ne), we(ne), de(ne), &
       grel(ns,ne), grbnd(ns,ne), v(ns,ne), mu(npa), wmu(npa), dmu(npa), upa(nr))
  allocate(fnhs(nr+1,nt,npa), fnis(nr+1,nt,npa), bounhs(nr+1,nt,npa), bounis(nr+1,nt,npa), &
       hdns(nr+1,nt,npa), didt(nr+1,nt,npa), dibndt(nr+1,nt,npa))
  allocate(bnes(nr+1,nt), dbdt(nr+1,nt), vt(nr+1,nt), eir(nr+1,nt), eip(nr+1,nt), mgnp(nr,nt))
  allocate(fgeos(ns,nt,ne,npa), f2(ns,nr,nt,ne,npa), setrc(ns), lsdr(ns), lscha(ns), lsatm(ns))
  allocate(out(ns))
  allocate(ffactor(ns,nr,ne,npa), ernh(ns,ne), epp(ns,ne), flux(ns,nr,nt,ne,npa))
  allocate(ppert(ns,nr,nt), ppart(ns,nr,nt), edent(ns,nr,nt), rnhtt(ns,nr,nt))
  ppert = 0; ppart = 0; edent = 0; rnhtt = 0
  read(u) charge, cex, rmas, rlz, ekev, ebnd, we, de, grel, grbnd, v, mu, wmu, dmu, upa
  read(u) fnhs, fnis, bounhs, bounis, hdns, didt, dibndt, bnes, dbdt, vt, eir, eip, mgnp
  read(u) fgeos, f2
  read(u) ffactor, ernh, epp
  close(u)

  rc = ramgpu_create(ctx, ns, nr, nt, ne, npa, 0_c_int)
  call check(rc, "create")
  do s = 1, ns
     call check(ramgpu_set_species(ctx, s, charge(s), cex(s), rmas(s)), "set_species")
  end do
  call check(ramgpu_set_params(ctx, sc(1), sc(2), sc(3)), "set_params")
  call check(ramgpu_set_grid(ctx, sc(4), sc(5), sc(6), sc(7), rlz, ekev, ebnd, we, de, grel, &
       grbnd, v, mu, wmu, dmu, upa), "set_grid")
  call check(ramgpu_set_geometry(ctx, fnhs, fnis, bounhs, bounis, hdns, didt, dibndt, bnes, &
       dbdt, eir, eip, mgnp), "set_geometry")
  call check(ramgpu_set_efield(ctx, vt), "set_efield")
  call check(ramgpu_set_boundary(ctx, fgeos), "set_boundary")
  call check(ramgpu_put_f2(ctx, f2), "put_f2")
  call check(ramgpu_ram_run(ctx, nsteps, out), "ram_run")
  ! src/ModRamRun.f90:207-222: ANISCH of every species (FLUX left out), then FLUX
  call check(ramgpu_set_moment_factors(ctx, ffactor, ernh, epp), "set_moment_factors")
  call check(ramgpu_moments(ctx, ppert, ppart, edent, rnhtt), "moments")
  call check(ramgpu_flux(ctx, flux), "flux")
  call check(ramgpu_get_f2(ctx, f2), "get_f2")
  call check(ramgpu_get_energy_state(ctx, setrc, lsdr, lscha, lsatm), "get_energy_state")
  call check(ramgpu_destroy(ctx), "destroy")

  open(newunit=u, file=trim(fout), access="stream", form="unformatted", status="replace")
  write(u) f2, setrc, lsdr, lscha, lsatm, ppert, ppart, edent, rnhtt, flux
  close(u)

contains
  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= RAMGPU_OK) then
       write(0, '(a,a,a,i0,a,a)') 'ramgpu_', what, ' failed (', code, '): ', ramgpu_error()
       stop 1
    end if
  end subroutine check
end program ramgpu_driver
