!  ModRamDrift_gpu.f90 -- drop-in replacement of the reference's
!  src/ModRamDrift.f90 (module ModRamDrift) and the GPU bodies for CEPARA,
!  CHAREXCHANGE, ATMOL (src/ModRamLoss.f90:19,457,485) and SUMRC
!  (src/ModRamRun.f90:231).  Compiled instead of src/ModRamDrift.f90, it keeps
!  the public interface DRIFTPARA(S), DRIFTR(S), DRIFTP(S), DRIFTE(S),
!  DRIFTMU(S), DRIFTEND and reads the same module state (ModRamVariables,
!  ModRamTiming::DTs, ModRamParams::BetaLim), so ram_run compiles unchanged.
!
!  One device context per species (ns=1), so the reference's OpenMP loop
!  over species (src/ModRamRun.f90:64) may call in from several threads.
!  The species' F2 slice goes to the device in ramgpu_begin_species (called
!  from CEPARA, the first operator of a species in ram_run) and back in
!  DRIFTEND (the last); geometry, field, boundary and DTs are uploaded with
!  it, since the reference may change them between steps.
!
!  ram_run also runs host operators on F2 between these calls when they are
!  switched on (src/ModRamRun.f90:81-168): COULEN/COULMU (DoUseCoulomb),
!  WAVELO/WPADIF (species%WPI, on for electrons by default), WPADIF for EMIC
!  (species%EMIC and DoUseEMIC) and FLCScatter (species%FLC and DoUseFLC).
!  For such a species the host copy stays authoritative: every operator here
!  puts the species' F2 on the device before it runs and takes it back after
!  (host_sync), so those host operators see and keep every change.
module ModRamDrift

  use iso_c_binding
  use ModRamMain,      ONLY: Real8_
  use ModRamGrids,     ONLY: nR, nT, nE, nPA
  use ModRamVariables, ONLY: FracCFL, DtDriftR, DtDriftP, DtDriftE, DtDriftMu
  use ramgpu_mod

  implicit none

  type(c_ptr), save, allocatable :: gpu_ctx(:)
  integer, save :: current_s = 0
  logical, save :: host_sync = .false.
  !$OMP THREADPRIVATE(current_s, host_sync)

  contains

  subroutine gpu_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= RAMGPU_OK) then
       write(*,*) 'ramgpu '//what//' failed: '//ramgpu_error()
       call CON_stop('ModRamDrift (GPU): '//what)
    end if
  end subroutine gpu_check

  ! Host F2 of species S to the device before an operator, when host
  ! operators also update it this step.
  subroutine sync_in(S)
    use ModRamVariables, ONLY: F2
    integer, intent(in) :: S
    if (host_sync) call gpu_check(ramgpu_put_f2(gpu_ctx(S), F2(S:S,:,:,:,:)), 'put_f2')
  end subroutine sync_in

  ! ... and back after it.
  subroutine sync_out(S)
    use ModRamVariables, ONLY: F2
    integer, intent(in) :: S
    real(c_double), allocatable :: f(:,:,:,:,:)
    if (.not. host_sync) return
    allocate(f(1, nR, nT, nE, nPA))
    call gpu_check(ramgpu_get_f2(gpu_ctx(S), f), 'get_f2')
    F2(S:S,:,:,:,:) = f
  end subroutine sync_out

  ! Upload the state species S needs and its F2 slice (first call per species).
  subroutine ramgpu_begin_species(S)
    use ModRamConst,     ONLY: PI
    use ModRamGrids,     ONLY: nS
    use ModRamTiming,    ONLY: DTs
    use ModRamParams,    ONLY: BetaLim, DoUseCoulomb, DoUseEMIC, DoUseFLC
    use ModRamVariables, ONLY: F2, FNHS, FNIS, BOUNHS, BOUNIS, HDNS, dIdt, dIbndt, BNES, &
                               dBdt, VT, EIR, EIP, outsideMGNP, FGEOS, RLZ, EKEV, EBND, &
                               WE, DE, GREL, GRBND, V, RMAS, MU, WMU, DMU, UPA, MDR, &
                               CONF1, CONF2, species
    use ModRamGrids,     ONLY: dPhi
    integer, intent(in) :: S
    integer(c_int) :: cex
    type(c_ptr) :: ctx
    !$OMP CRITICAL (ramgpu_ctx_alloc)
    if (.not. allocated(gpu_ctx)) then
       allocate(gpu_ctx(nS))
       gpu_ctx = c_null_ptr
    end if
    if (.not. c_associated(gpu_ctx(S))) &
         call gpu_check(ramgpu_create(gpu_ctx(S), 1_c_int, int(nR, c_int), int(nT, c_int), &
                                      int(nE, c_int), int(nPA, c_int), 0_c_int), 'create')
    !$OMP END CRITICAL (ramgpu_ctx_alloc)
    ctx = gpu_ctx(S)
    current_s = S
    host_sync = species(S)%WPI .or. DoUseCoulomb .or. (species(S)%EMIC .and. DoUseEMIC) .or. &
                (species(S)%FLC .and. DoUseFLC)
    cex = RAMGPU_CEX_NONE                     ! CEPARA's species selection, src/ModRamLoss.f90:37-83
    if (species(S)%CEX) then
       select case (species(S)%s_name)
       case ("Hydrogen"); cex = RAMGPU_CEX_H
       case ("HeliumP1"); cex = RAMGPU_CEX_HE
       case ("OxygenP1"); cex = RAMGPU_CEX_O
       case default
          call CON_stop('ModRamDrift (GPU): file-based charge exchange is not on the GPU path')
       end select
    end if
    call gpu_check(ramgpu_set_species(ctx, 1_c_int, int(species(S)%s_charge, c_int), cex, &
                                      RMAS(S)), 'set_species')
    call gpu_check(ramgpu_set_params(ctx, DTs, BetaLim, FracCFL), 'set_params')
    call gpu_check(ramgpu_set_grid(ctx, MDR, dPhi, CONF1, CONF2, RLZ, EKEV, EBND, WE, DE, &
                                   GREL(S:S,:), GRBND(S:S,:), V(S:S,:), MU, WMU, DMU, UPA), &
                   'set_grid')
    call gpu_check(ramgpu_set_geometry(ctx, FNHS, FNIS, BOUNHS, BOUNIS, HDNS, dIdt, dIbndt, &
                                       BNES, dBdt, EIR, EIP, outsideMGNP), 'set_geometry')
    call gpu_check(ramgpu_set_efield(ctx, VT), 'set_efield')
    call gpu_check(ramgpu_set_boundary(ctx, FGEOS(S:S,:,:,:)), 'set_boundary')
    call gpu_check(ramgpu_put_f2(ctx, F2(S:S,:,:,:,:)), 'put_f2')
  end subroutine ramgpu_begin_species

  SUBROUTINE DRIFTEND
    ! F2 slice of the current species back to the host (the reference frees
    ! its drift scratch here, src/ModRamDrift.f90:23-30)
    use ModRamVariables, ONLY: F2
    real(c_double), allocatable :: f(:,:,:,:,:)
    allocate(f(1, nR, nT, nE, nPA))
    if (.not. host_sync) &  ! with host_sync the host copy is already current
         call gpu_check(ramgpu_get_f2(gpu_ctx(current_s), f), 'get_f2')
    if (host_sync) f = F2(current_s:current_s,:,:,:,:)
    F2(current_s:current_s,:,:,:,:) = f
    call gpu_check(ramgpu_driftend(gpu_ctx(current_s)), 'driftend')
  END SUBROUTINE DRIFTEND

  SUBROUTINE DRIFTPARA(S)
    integer, intent(in) :: S
    if (current_s /= S) call ramgpu_begin_species(S)
    call gpu_check(ramgpu_driftpara(gpu_ctx(S), 1_c_int), 'driftpara')
  END SUBROUTINE DRIFTPARA

  SUBROUTINE DRIFTR(S)
    integer, intent(in) :: S
    real(c_double) :: dt
    call sync_in(S)
    call gpu_check(ramgpu_driftr(gpu_ctx(S), 1_c_int, dt), 'driftr')
    call sync_out(S)
    DtDriftR(S) = dt
  END SUBROUTINE DRIFTR

  SUBROUTINE DRIFTP(S)
    integer, intent(in) :: S
    real(c_double) :: dt
    call sync_in(S)
    call gpu_check(ramgpu_driftp(gpu_ctx(S), 1_c_int, dt), 'driftp')
    call sync_out(S)
    DtDriftP(S) = dt
  END SUBROUTINE DRIFTP

  SUBROUTINE DRIFTE(S)
    integer, intent(in) :: S
    real(c_double) :: dt
    call sync_in(S)
    call gpu_check(ramgpu_drifte(gpu_ctx(S), 1_c_int, dt), 'drifte')
    call sync_out(S)
    DtDriftE(S) = dt
  END SUBROUTINE DRIFTE

  SUBROUTINE DRIFTMU(S)
    integer, intent(in) :: S
    real(c_double) :: dt
    call sync_in(S)
    call gpu_check(ramgpu_driftmu(gpu_ctx(S), 1_c_int, dt), 'driftmu')
    call sync_out(S)
    DtDriftMu(S) = dt
  END SUBROUTINE DRIFTMU

  ! --- GPU bodies for the loss operators and SUMRC: the reference's
  ! CEPARA/CHAREXCHANGE/ATMOL (src/ModRamLoss.f90) and SUMRC
  ! (src/ModRamRun.f90) each call the matching routine below as their body.
  subroutine gpu_cepara(S)
    integer, intent(in) :: S
    call ramgpu_begin_species(S)
    call gpu_check(ramgpu_cepara(gpu_ctx(S), 1_c_int), 'cepara')
  end subroutine gpu_cepara

  subroutine gpu_charexchange(S)
    integer, intent(in) :: S
    call sync_in(S)
    call gpu_check(ramgpu_charexchange(gpu_ctx(S), 1_c_int), 'charexchange')
    call sync_out(S)
  end subroutine gpu_charexchange

  subroutine gpu_atmol(S)
    integer, intent(in) :: S
    call sync_in(S)
    call gpu_check(ramgpu_atmol(gpu_ctx(S), 1_c_int), 'atmol')
    call sync_out(S)
  end subroutine gpu_atmol

  subroutine gpu_sumrc(S)
    use ModRamVariables, ONLY: SETRC, ELORC
    integer, intent(in) :: S
    real(c_double) :: setrc1(1), zero(1), setrc_new, elorc_new
    setrc1(1) = SETRC(S)
    zero(1) = 0.0_c_double
    call sync_in(S)
    call gpu_check(ramgpu_set_energy_state(gpu_ctx(S), setrc1, zero, zero, zero), 'energy')
    call gpu_check(ramgpu_sumrc(gpu_ctx(S), 1_c_int, setrc_new, elorc_new), 'sumrc')
    SETRC(S) = setrc_new
    ELORC(S) = elorc_new
  end subroutine gpu_sumrc

END MODULE ModRamDrift
