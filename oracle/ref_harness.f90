!  ref_harness.f90 -- TEST INFRASTRUCTURE ONLY.
!
!  A C-callable harness around the reference's own, unmodified drift
!  operators (src/ModRamDrift.f90), Coulomb operators (src/ModRamCoul.f90)
!  and dipole bounce functions (src/ModRamFunctions.f90:90-143), compiled
!  from the reference sources by
!  oracle/Makefile into oracle/_ref/.  The harness owns nothing but the copy
!  in/out of the reference's module state; it fills that state the way
!  ram_allocate does (src/ModRamInit.f90:67-138) and then calls the reference
!  subroutines by name.  Used only to pin the C restatement in
!  oracle/ram_oracle.c.
module ram_ref_harness
  use iso_c_binding
  use ModRamGrids,     only: nS, nR, nT, nE, nPA, dPhi
  use ModRamVariables, only: F2, FNHS, FNIS, BOUNHS, BOUNIS, HDNS, dIdt, dIbndt, &
                             BNES, dBdt, VT, EIR, EIP, outsideMGNP, FGEOS, RLZ, &
                             EKEV, EBND, WE, DE, GREL, GRBND, V, RMAS, MU, WMU, &
                             DMU, UPA, species, DtDriftR, DtDriftP, DtDriftE, &
                             DtDriftMu, MDR, CONF1, CONF2, FracCFL, VBND, &
                             COULE, COULI, ATA, GTA, GTAE, GTAI, CEDR, CIDR, NECR, XNE
  use ModRamTiming,    only: DTs, TimeRamElapsed
  use ModRamParams,    only: BetaLim
  use ModRamSpecies,   only: DefineSpecies
  use ModRamDrift,     only: DRIFTPARA, DRIFTR, DRIFTP, DRIFTE, DRIFTMU, DRIFTEND
  use ModRamCoul,      only: COULPARA, COULEN, COULMU
  use ModRamFunctions, only: funt, funi
  implicit none
  logical, save :: para_live = .false.
contains

  subroutine release()
    if (para_live) then
       call DRIFTEND
       para_live = .false.
    end if
    if (allocated(F2)) deallocate(F2, FNHS, FNIS, BOUNHS, BOUNIS, HDNS, dIdt, &
         dIbndt, BNES, dBdt, VT, EIR, EIP, outsideMGNP, FGEOS, RLZ, EKEV, EBND, &
         WE, DE, GREL, GRBND, V, RMAS, MU, WMU, DMU, UPA, species, DtDriftR, &
         DtDriftP, DtDriftE, DtDriftMu)
    if (allocated(COULE)) deallocate(VBND, COULE, COULI, ATA, GTA, GTAE, GTAI, &
         CEDR, CIDR, NECR, XNE)
  end subroutine release

  subroutine ref_setup(ns_, nr_, nt_, ne_, npa_, dts_, betalim_, fraccfl_, mdr_, &
       dphi_, conf1_, conf2_, charge_, rmas_, rlz_, ekev_, ebnd_, we_, de_, grel_, &
       grbnd_, v_, mu_, wmu_, dmu_, upa_, fnhs_, fnis_, bounhs_, bounis_, hdns_, &
       didt_, dibndt_, bnes_, dbdt_, vt_, eir_, eip_, mgnp_, fgeos_, f2_) &
       bind(C, name="ref_setup")
    integer(c_int), value :: ns_, nr_, nt_, ne_, npa_
    real(c_double), value :: dts_, betalim_, fraccfl_, mdr_, dphi_, conf1_, conf2_
    integer(c_int) :: charge_(ns_), mgnp_(nr_, nt_)
    real(c_double) :: rmas_(ns_), rlz_(nr_+1), ekev_(ne_), ebnd_(ne_), we_(ne_), &
         de_(ne_), grel_(ns_, ne_), grbnd_(ns_, ne_), v_(ns_, ne_), mu_(npa_), &
         wmu_(npa_), dmu_(npa_), upa_(nr_)
    real(c_double) :: fnhs_(nr_+1, nt_, npa_), fnis_(nr_+1, nt_, npa_), &
         bounhs_(nr_+1, nt_, npa_), bounis_(nr_+1, nt_, npa_), hdns_(nr_+1, nt_, npa_), &
         didt_(nr_+1, nt_, npa_), dibndt_(nr_+1, nt_, npa_)
    real(c_double) :: bnes_(nr_+1, nt_), dbdt_(nr_+1, nt_), vt_(nr_+1, nt_), &
         eir_(nr_+1, nt_), eip_(nr_+1, nt_)
    real(c_double) :: fgeos_(ns_, nt_, ne_, npa_), f2_(ns_, nr_, nt_, ne_, npa_)
    integer :: s

    call release()
    nS = ns_; nR = nr_; nT = nt_; nE = ne_; nPA = npa_
    allocate(F2(nS,nR,nT,nE,nPA), FNHS(nR+1,nT,nPA), FNIS(nR+1,nT,nPA), &
         BOUNHS(nR+1,nT,nPA), BOUNIS(nR+1,nT,nPA), HDNS(nR+1,nT,nPA), &
         dIdt(nR+1,nT,nPA), dIbndt(nR+1,nT,nPA), BNES(nR+1,nT), dBdt(nR+1,nT), &
         VT(nR+1,nT), EIR(nR+1,nT), EIP(nR+1,nT), outsideMGNP(nR,nT), &
         FGEOS(nS,nT,nE,nPA), RLZ(nR+1), EKEV(nE), EBND(nE), WE(nE), DE(nE), &
         GREL(nS,nE), GRBND(nS,nE), V(nS,nE), RMAS(nS), MU(nPA), WMU(nPA), &
         DMU(nPA), UPA(nR), species(nS), DtDriftR(nS), DtDriftP(nS), &
         DtDriftE(nS), DtDriftMu(nS))
    DTs = dts_; BetaLim = betalim_; FracCFL = fraccfl_
    MDR = mdr_; dPhi = dphi_; CONF1 = conf1_; CONF2 = conf2_
    do s = 1, nS
       species(s)%s_charge = charge_(s)
    end do
    RMAS = rmas_; RLZ = rlz_; EKEV = ekev_; EBND = ebnd_; WE = we_; DE = de_
    GREL = grel_; GRBND = grbnd_; V = v_; MU = mu_; WMU = wmu_; DMU = dmu_; UPA = upa_
    FNHS = fnhs_; FNIS = fnis_; BOUNHS = bounhs_; BOUNIS = bounis_; HDNS = hdns_
    dIdt = didt_; dIbndt = dibndt_; BNES = bnes_; dBdt = dbdt_; VT = vt_
    EIR = eir_; EIP = eip_; outsideMGNP = mgnp_; FGEOS = fgeos_; F2 = f2_
    DtDriftR = 0; DtDriftP = 0; DtDriftE = 0; DtDriftMu = 0
  end subroutine ref_setup

  ! the Coulomb operators' state, zeroed as ram_allocate does
  ! (src/ModRamInit.f90:97-151): VBND(NS,NE), NECR(NR,NT), the elapsed time
  subroutine ref_setup_coulomb(vbnd_, necr_, time_) bind(C, name="ref_setup_coulomb")
    real(c_double) :: vbnd_(nS, nE), necr_(nR, nT)
    real(c_double), value :: time_
    if (allocated(COULE)) deallocate(VBND, COULE, COULI, ATA, GTA, GTAE, GTAI, &
         CEDR, CIDR, NECR, XNE)
    allocate(VBND(nS,nE), COULE(nS,nE,nPa), COULI(nS,nE,nPa), ATA(nS,nE,nPa), &
         GTA(nS,nE,nPa), GTAE(nS,nE,nPa), GTAI(nS,nE,nPa), CEDR(nS,nE,nPa), &
         CIDR(nS,nE,nPa), NECR(nR,nT), XNE(nR,nT))
    COULE = 0.0; COULI = 0.0; ATA = 0.0; GTA = 0.0; GTAE = 0.0; GTAI = 0.0
    CEDR = 0.0; CIDR = 0.0; XNE = 0.0
    VBND = vbnd_; NECR = necr_; TimeRamElapsed = time_
    call DefineSpecies
  end subroutine ref_setup_coulomb

  ! COULE, COULI, ATA, GTA, CEDR, CIDR (NS,NE,NPA) after COULPARA
  subroutine ref_get_coulomb(coule_, couli_, ata_, gta_, cedr_, cidr_) &
       bind(C, name="ref_get_coulomb")
    real(c_double) :: coule_(nS,nE,nPa), couli_(nS,nE,nPa), ata_(nS,nE,nPa), &
         gta_(nS,nE,nPa), cedr_(nS,nE,nPa), cidr_(nS,nE,nPa)
    coule_ = COULE; couli_ = COULI; ata_ = ATA; gta_ = GTA; cedr_ = CEDR; cidr_ = CIDR
  end subroutine ref_get_coulomb

  subroutine ref_put_f2(f2_) bind(C, name="ref_put_f2")
    real(c_double) :: f2_(nS, nR, nT, nE, nPA)
    F2 = f2_
  end subroutine ref_put_f2

  subroutine ref_get_f2(f2_) bind(C, name="ref_get_f2")
    real(c_double) :: f2_(nS, nR, nT, nE, nPA)
    f2_ = F2
  end subroutine ref_get_f2

  subroutine ref_get_dtdrift(out) bind(C, name="ref_get_dtdrift")
    real(c_double) :: out(nS, 4)
    out(:, 1) = DtDriftR; out(:, 2) = DtDriftP
    out(:, 3) = DtDriftE; out(:, 4) = DtDriftMu
  end subroutine ref_get_dtdrift

  ! op: 0 DRIFTPARA, 1 DRIFTR, 2 DRIFTP, 3 DRIFTE, 4 DRIFTMU, 5 COULPARA,
  ! 6 COULEN, 7 COULMU
  subroutine ref_call(op, s) bind(C, name="ref_call")
    integer(c_int), value :: op, s
    select case (op)
    case (0)
       call DRIFTPARA(s)
       para_live = .true.
    case (1)
       call DRIFTR(s)
    case (2)
       call DRIFTP(s)
    case (3)
       call DRIFTE(s)
    case (4)
       call DRIFTMU(s)
    case (5)
       call COULPARA(s)
    case (6)
       call COULEN(s)
    case (7)
       call COULMU(s)
    end select
  end subroutine ref_call

  subroutine ref_release() bind(C, name="ref_release")
    call release()
  end subroutine ref_release

  real(c_double) function ref_funt(x) bind(C, name="ref_funt")
    real(c_double), value :: x
    ref_funt = funt(x)
  end function ref_funt

  real(c_double) function ref_funi(x) bind(C, name="ref_funi")
    real(c_double), value :: x
    ref_funi = funi(x)
  end function ref_funi

end module ram_ref_harness
