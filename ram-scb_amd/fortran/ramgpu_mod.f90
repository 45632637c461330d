!  ramgpu_mod.f90 -- ISO_C_BINDING interface to the C ABI in include/ramgpu.h.
!
!  A Fortran host (the reference's ram_run, src/ModRamRun.f90:16-259) uses
!  this module to drive the MI355X split step.  Arrays are passed as they are
!  stored in the reference's modules (Fortran order), sizes by value, one
!  opaque context pointer per problem.  Every function returns the C status
!  (0 = RAMGPU_OK); ramgpu_error() returns the last message as a Fortran
!  string.
module ramgpu_mod
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: RAMGPU_OK = 0, RAMGPU_EINVAL = 1, RAMGPU_ENOMEM = 2, &
       RAMGPU_EHIP = 3, RAMGPU_ESTATE = 4
  integer(c_int), parameter :: RAMGPU_CEX_NONE = 0, RAMGPU_CEX_H = 1, RAMGPU_CEX_HE = 2, &
       RAMGPU_CEX_O = 3

  type, bind(C) :: ramgpu_species_out
     real(c_double) :: setrc, lsdr, lscha, lsatm
     real(c_double) :: dtdrift_r, dtdrift_p, dtdrift_e, dtdrift_mu
     real(c_double) :: lswae, lscoe, lscsc
  end type ramgpu_species_out

  interface
     function ramgpu_last_error() bind(C, name="ramgpu_last_error") result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function ramgpu_last_error

     integer(c_int) function ramgpu_create(ctx, ns, nr, nt, ne, npa, device) &
          bind(C, name="ramgpu_create")
       import :: c_int, c_ptr
       type(c_ptr), intent(out) :: ctx
       integer(c_int), value :: ns, nr, nt, ne, npa, device
     end function ramgpu_create

     integer(c_int) function ramgpu_destroy(ctx) bind(C, name="ramgpu_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function ramgpu_destroy

     integer(c_int) function ramgpu_set_species(ctx, s, charge, cex_kind, rmas) &
          bind(C, name="ramgpu_set_species")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: s, charge, cex_kind
       real(c_double), value :: rmas
     end function ramgpu_set_species

     integer(c_int) function ramgpu_set_params(ctx, dts, betalim, fraccfl) &
          bind(C, name="ramgpu_set_params")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: dts, betalim, fraccfl
     end function ramgpu_set_params

     integer(c_int) function ramgpu_set_grid(ctx, mdr, dphi, conf1, conf2, rlz, ekev, ebnd, &
          we, de, grel, grbnd, v, mu, wmu, dmu, upa) bind(C, name="ramgpu_set_grid")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: mdr, dphi, conf1, conf2
       real(c_double), intent(in) :: rlz(*), ekev(*), ebnd(*), we(*), de(*), grel(*), &
            grbnd(*), v(*), mu(*), wmu(*), dmu(*), upa(*)
     end function ramgpu_set_grid

     integer(c_int) function ramgpu_set_geometry(ctx, fnhs, fnis, bounhs, bounis, hdns, &
          didt, dibndt, bnes, dbdt, eir, eip, outside_mgnp) bind(C, name="ramgpu_set_geometry")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: fnhs(*), fnis(*), bounhs(*), bounis(*), hdns(*), &
            didt(*), dibndt(*), bnes(*), dbdt(*), eir(*), eip(*)
       integer(c_int), intent(in) :: outside_mgnp(*)
     end function ramgpu_set_geometry

     integer(c_int) function ramgpu_get_geometry(ctx, fnhs, fnis, bounhs, bounis, hdns, &
          didt, dibndt, bnes, dbdt, eir, eip, outside_mgnp) bind(C, name="ramgpu_get_geometry")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: fnhs(*), fnis(*), bounhs(*), bounis(*), hdns(*), &
            didt(*), dibndt(*), bnes(*), dbdt(*), eir(*), eip(*)
       integer(c_int), intent(out) :: outside_mgnp(*)
     end function ramgpu_get_geometry

     ! computehI (src/ModRamScb.f90:144-643) from xRAM, yRAM, zRAM, bRAM on
     ! the RAM grid; scale_at is ScaleAt (zeros: no row beyond the SCB domain)
     integer(c_int) function ramgpu_computehI(ctx, nthe, ieq, x, y, z, b, chi, pa, pabn, lz, &
          outside_mgnp, scale_at, bnormal, smooth) bind(C, name="ramgpu_computehI")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: nthe, ieq, smooth
       real(c_double), intent(in) :: x(*), y(*), z(*), b(*), chi(*), pa(*), pabn(*), lz(*)
       integer(c_int), intent(in) :: outside_mgnp(*), scale_at(*)
       real(c_double), value :: bnormal
     end function ramgpu_computehI

     integer(c_int) function ramgpu_set_efield(ctx, vt) bind(C, name="ramgpu_set_efield")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: vt(*)
     end function ramgpu_set_efield

     integer(c_int) function ramgpu_set_kp(ctx, kp, lz, phi) bind(C, name="ramgpu_set_kp")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: kp
       real(c_double), intent(in) :: lz(*), phi(*)
     end function ramgpu_set_kp

     integer(c_int) function ramgpu_set_boundary(ctx, fgeos) bind(C, name="ramgpu_set_boundary")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: fgeos(*)
     end function ramgpu_set_boundary

     integer(c_int) function ramgpu_set_boundary_scale(ctx, scale) &
          bind(C, name="ramgpu_set_boundary_scale")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: scale(*)
     end function ramgpu_set_boundary_scale

     integer(c_int) function ramgpu_put_f2(ctx, f2) bind(C, name="ramgpu_put_f2")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: f2(*)
     end function ramgpu_put_f2

     integer(c_int) function ramgpu_get_f2(ctx, f2) bind(C, name="ramgpu_get_f2")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: f2(*)
     end function ramgpu_get_f2

     integer(c_int) function ramgpu_set_energy_state(ctx, setrc, lsdr, lscha, lsatm) &
          bind(C, name="ramgpu_set_energy_state")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: setrc(*), lsdr(*), lscha(*), lsatm(*)
     end function ramgpu_set_energy_state

     integer(c_int) function ramgpu_get_energy_state(ctx, setrc, lsdr, lscha, lsatm) &
          bind(C, name="ramgpu_get_energy_state")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: setrc(*), lsdr(*), lscha(*), lsatm(*)
     end function ramgpu_get_energy_state

     integer(c_int) function ramgpu_driftpara(ctx, s) bind(C, name="ramgpu_driftpara")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: s
     end function ramgpu_driftpara

     integer(c_int) function ramgpu_driftr(ctx, s, dtdrift) bind(C, name="ramgpu_driftr")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: s
       real(c_double), intent(out) :: dtdrift
     end function ramgpu_driftr

     integer(c_int) function ramgpu_driftp(ctx, s, dtdrift) bind(C, name="ramgpu_driftp")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: s
       real(c_double), intent(out) :: dtdrift
     end function ramgpu_driftp

     integer(c_int) function ramgpu_drifte(ctx, s, dtdrift) bind(C, name="ramgpu_drifte")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: s
       real(c_double), intent(out) :: dtdrift
     end function ramgpu_drifte

     integer(c_int) function ramgpu_driftmu(ctx, s, dtdrift) bind(C, name="ramgpu_driftmu")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: s
       real(c_double), intent(out) :: dtdrift
     end function ramgpu_driftmu

     integer(c_int) function ramgpu_driftend(ctx) bind(C, name="ramgpu_driftend")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function ramgpu_driftend

     integer(c_int) function ramgpu_cepara(ctx, s) bind(C, name="ramgpu_cepara")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: s
     end function ramgpu_cepara

     integer(c_int) function ramgpu_charexchange(ctx, s) bind(C, name="ramgpu_charexchange")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: s
     end function ramgpu_charexchange

     integer(c_int) function ramgpu_atmol(ctx, s) bind(C, name="ramgpu_atmol")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: s
     end function ramgpu_atmol

     integer(c_int) function ramgpu_wavelo(ctx, s) bind(C, name="ramgpu_wavelo")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: s
     end function ramgpu_wavelo

     ! WAVELO's tables of species s: WALOS1, WALOS2, WALOS3 (NR,NE)
     integer(c_int) function ramgpu_set_wavelo(ctx, s, walos1, walos2, walos3) &
          bind(C, name="ramgpu_set_wavelo")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: s
       real(c_double), intent(in) :: walos1(*), walos2(*), walos3(*)
     end function ramgpu_set_wavelo

     ! KP, Kpmax12, LZ(NR+1), NECR(NR,NT) (with use_plasmasphere /= 0)
     integer(c_int) function ramgpu_set_wave_params(ctx, kp, kpmax12, lz, necr, use_pp) &
          bind(C, name="ramgpu_set_wave_params")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: kp, kpmax12
       real(c_double), intent(in) :: lz(*), necr(*)
       integer(c_int), value :: use_pp
     end function ramgpu_set_wave_params

     integer(c_int) function ramgpu_get_loss_state(ctx, lswae, lscoe, lscsc) &
          bind(C, name="ramgpu_get_loss_state")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: lswae(*), lscoe(*), lscsc(*)
     end function ramgpu_get_loss_state

     integer(c_int) function ramgpu_sumrc(ctx, s, setrc, elorc) bind(C, name="ramgpu_sumrc")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: s
       real(c_double), intent(out) :: setrc, elorc
     end function ramgpu_sumrc

     integer(c_int) function ramgpu_set_moment_factors(ctx, ffactor, ernh, epp) &
          bind(C, name="ramgpu_set_moment_factors")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: ffactor(*), ernh(*), epp(*)
     end function ramgpu_set_moment_factors

     integer(c_int) function ramgpu_anisch(ctx, s, ppert, ppart, edent, rnhtt) bind(C, name="ramgpu_anisch")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: s
       real(c_double), intent(inout) :: ppert(*), ppart(*), edent(*), rnhtt(*)
     end function ramgpu_anisch

     integer(c_int) function ramgpu_flux(ctx, flux) bind(C, name="ramgpu_flux")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: flux(*)
     end function ramgpu_flux

     integer(c_int) function ramgpu_moments(ctx, ppert, ppart, edent, rnhtt, flux) bind(C, name="ramgpu_moments")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: ppert(*), ppart(*), edent(*), rnhtt(*)
       real(c_double), intent(out), optional :: flux(*)  ! absent: NULL, no FLUX
     end function ramgpu_moments

     ! Coulomb collisions (src/ModRamCoul.f90): DoUseCoulomb with VBND(NS,NE),
     ! NECR(NR,NT), TimeRamElapsed, COULPARA/COULEN/COULMU(S)
     integer(c_int) function ramgpu_set_coulomb(ctx, use, vbnd) bind(C, name="ramgpu_set_coulomb")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: use
       real(c_double), intent(in) :: vbnd(*)
     end function ramgpu_set_coulomb

     integer(c_int) function ramgpu_set_necr(ctx, necr) bind(C, name="ramgpu_set_necr")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: necr(*)
     end function ramgpu_set_necr

     integer(c_int) function ramgpu_set_time(ctx, time_elapsed) bind(C, name="ramgpu_set_time")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: time_elapsed
     end function ramgpu_set_time

     integer(c_int) function ramgpu_set_told(ctx, told) bind(C, name="ramgpu_set_told")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: told
     end function ramgpu_set_told

     integer(c_int) function ramgpu_get_told(ctx, told) bind(C, name="ramgpu_get_told")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: told
     end function ramgpu_get_told

     integer(c_int) function ramgpu_get_time(ctx, time_elapsed) bind(C, name="ramgpu_get_time")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: time_elapsed
     end function ramgpu_get_time

     integer(c_int) function ramgpu_coulpara(ctx, s) bind(C, name="ramgpu_coulpara")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: s
     end function ramgpu_coulpara

     integer(c_int) function ramgpu_coulen(ctx, s) bind(C, name="ramgpu_coulen")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: s
     end function ramgpu_coulen

     integer(c_int) function ramgpu_coulmu(ctx, s) bind(C, name="ramgpu_coulmu")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: s
     end function ramgpu_coulmu

     ! COULE, COULI, ATA, GTA, CEDR, CIDR (NE,NPA) of species s
     integer(c_int) function ramgpu_get_coulomb_tables(ctx, s, coule, couli, ata, gta, cedr, cidr) &
          bind(C, name="ramgpu_get_coulomb_tables")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: s
       real(c_double), intent(out) :: coule(*), couli(*), ata(*), gta(*), cedr(*), cidr(*)
     end function ramgpu_get_coulomb_tables

     ! WPADIF (D = d1 + d2, (NR,NT,NE,NPA)) and FLCscatter (FLC_coef(S,...))
     integer(c_int) function ramgpu_set_wpadif(ctx, s, d1, d2, at_wpi, at_emic) &
          bind(C, name="ramgpu_set_wpadif")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: s, at_wpi, at_emic
       real(c_double), intent(in) :: d1(*), d2(*)
     end function ramgpu_set_wpadif

     integer(c_int) function ramgpu_wpadif(ctx, s) bind(C, name="ramgpu_wpadif")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: s
     end function ramgpu_wpadif

     integer(c_int) function ramgpu_set_flc(ctx, s, coef, dt_bc) bind(C, name="ramgpu_set_flc")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: s
       real(c_double), intent(in) :: coef(*)
       real(c_double), value :: dt_bc
     end function ramgpu_set_flc

     integer(c_int) function ramgpu_flcscatter(ctx, s) bind(C, name="ramgpu_flcscatter")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: s
     end function ramgpu_flcscatter

     integer(c_int) function ramgpu_ram_run(ctx, nsteps, out) bind(C, name="ramgpu_ram_run")
       import :: c_int, c_ptr, ramgpu_species_out
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps
       type(ramgpu_species_out), intent(out) :: out(*)
     end function ramgpu_ram_run
  end interface

contains

  function ramgpu_error() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: n
    type(c_ptr) :: cp
    cp = ramgpu_last_error()
    call c_f_pointer(cp, p, [4096])
    n = 0
    do while (n < 4096)
       if (p(n + 1) == c_null_char) exit
       n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(p(1:n), msg)
  end function ramgpu_error

end module ramgpu_mod
