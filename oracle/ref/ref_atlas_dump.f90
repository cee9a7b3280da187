! ref_atlas_dump -- runs the reference ATLAS12 first-iteration set-up
! (READIN, populations, Doppler widths, KAPCONT) on a model deck with a
! given line store, then records what the opacity-sampling hot path reads
! and what the reference computes from it:
!   * LINOP1's XLINES,
!   * KAPP's outputs at every frequency of the grid,
!   * RUN_RT_PASS's frequency-integrated accumulators.
! Test infrastructure (built by oracle/ref/Makefile).
!
!   ATLAS12=<root> ref_atlas_dump <deck> <nustart> <nuhi> <lines.bin> <out.bin> [<seldir>]
!
! lines.bin: int32 count, then count LINEDATA records (4 x int32).
! With <seldir>, the line store is instead SELECTLINES' product from the
! raw lists in <seldir> (gfpred29dec2014.bin, lowobsat12.bin, hilines.bin,
! h2opokazatel.bin, h3plus.dat, lines.list naming a MOL2 TiO list and the
! ASCII diatomic lists); it is written to <out.bin>.sel as int32 count, the
! records, then WAVETAB(1:NWAVE), and lines.bin is not read.  The diatomic
! records read_diatomics_for_atlas returns for the same window go to
! <out.bin>.dia: int32 count, then the wavelengths, log gf, E_lo, log
! gamma_rad, log gamma_w (REAL(8) arrays) and the MOLCODEs (int32).
PROGRAM ref_atlas_dump
  USE mod_atlas_data
  USE mod_mklinelist, only: read_diatomics_for_atlas, diatomic_record_t
  IMPLICIT NONE
  CHARACTER(LEN=512) :: deck, fin, fout, arg, envval, seldir
  INTEGER :: envlen, envstat, u, nustart_in, nuhi_in, nl, nu, j, nel, i
  REAL(8) :: freq15_l, xmax_l
  ! TCORR(3) sections G2-K, TTAUP, ROSSTAB and CZ_CONSTRUCT probes
  INTEGER :: uf, idum_l, k, irt, nq, iq
  REAL(8) :: rx_l, taustd_l(kw), tn_l(kw), pr_l(kw), ab_l(kw), pt_l(kw), pg_l(kw), t1_l(kw)
  REAL(8) :: t0_l(kw), p0_l(kw), qt_l(kw * 12), qp_l(kw * 12), qr_l(kw * 12)
  ! CONVEC probe: the EOS evaluations it makes (EDENS, RHO at four states)
  INTEGER :: uc
  REAL(8) :: cv_e(kw, 4), cv_r(kw, 4)
  ! diatomic reader probe
  TYPE(diatomic_record_t), ALLOCATABLE :: drec_l(:)
  INTEGER :: nd_l

  CALL GET_COMMAND_ARGUMENT(1, deck)
  CALL GET_COMMAND_ARGUMENT(2, arg); READ(arg, *) nustart_in
  CALL GET_COMMAND_ARGUMENT(3, arg); READ(arg, *) nuhi_in
  CALL GET_COMMAND_ARGUMENT(4, fin)
  CALL GET_COMMAND_ARGUMENT(5, fout)
  seldir = ''
  IF (COMMAND_ARGUMENT_COUNT() .GE. 6) CALL GET_COMMAND_ARGUMENT(6, seldir)

  CALL GET_ENVIRONMENT_VARIABLE('ATLAS12', envval, envlen, envstat)
  DATADIR = TRIM(envval) // '/data/'
  CALL set_bc_data_dir(TRIM(DATADIR))
  CALL TABVOIGT(200.0D0, 2001)
  INPUTDATA = 5
  CALL BLOCKJ
  CALL BLOCKH
  CALL IONPOTS
  ITEMP = 0
  INPUT_MODEL_FILE = deck
  NUMITS = 1
  IFPNCH(1) = 0
  IFPRNT(1) = 1
  CALL READIN(1)

  ! Species masses: dominant isotope of each ion (src/atlas12c.f90:418-440)
  XNF = 0.0D0
  XNFP = 0.0D0
  DOPPLE = 0.0D0
  CALL ISOTOPES
  DO nel = 1, MION
    xmax_l = 0.0D0
    DO i = 1, 10
      IF (ISOTOPE(i, 2, nel) .GT. xmax_l) THEN
        AMASSISO(1, nel) = ISOTOPE(i, 1, nel)
        xmax_l = ISOTOPE(i, 2, nel)
      END IF
    END DO
    IF (xmax_l .LE. 0.0D0 .AND. ISOTOPE(1, 1, nel) .GT. 0.0) AMASSISO(1, nel) = ISOTOPE(1, 1, nel)
  END DO

  ! Frequency grid of the ATLAS12 driver (src/atlas12c.f90:451-491),
  ! log-spaced at 1e-4 dex from the given start index.
  NULO = 1
  NUHI = nuhi_in
  NUSTEP = 1
  NUMNU = NUHI
  DO nu = NULO, NUHI
    WAVESET(nu) = 10.0D0 ** (1.0D0 + 0.0001D0 * (nu + nustart_in - 1))
  END DO
  RCOSET(1) = (CLIGHT_NMS / WAVESET(1) - CLIGHT_NMS / WAVESET(2)) * 1.5D0
  DO nu = 2, NUMNU - 1
    RCOSET(nu) = (CLIGHT_NMS / WAVESET(nu - 1) - CLIGHT_NMS / WAVESET(nu + 1)) * 0.5D0
  END DO
  RCOSET(NUMNU) = (CLIGHT_NMS / WAVESET(NUMNU - 1) + CLIGHT_NMS / WAVESET(NUMNU)) * 0.25D0

  ! First-iteration state (src/atlas12c.f90:500-563, without line selection)
  ITER = 1
  ITEMP = ITEMP + ITER
  IF (IFPRES .NE. 0) THEN
    PZERO = PCON + PRADK0 + PTURB0
    DO j = 1, NRHOX
      CHARGESQ(j) = XNE(j) * 2.0D0
      IF (2.0D0 * XNE(j) - P(j) / TK(j) .GT. 0.0D0) &
        CHARGESQ(j) = CHARGESQ(j) + 2.0D0 * (2.0D0 * XNE(j) - P(j) / TK(j))
      PTOTAL(j) = GRAV * RHOX(j) + PZERO
    END DO
    IFEDNS = 0
    CALL COMPUTE_ONE_POP(0.D0, 1, XNE)
    CALL COMPUTE_ALL_POPS
  END IF
  IF (IFEDNS .EQ. 1) CALL ENERGY_DENSITY
  DO j = 1, NRHOX
    DO nel = 1, MION - 1
      IF (AMASSISO(1, nel) .LE. 0.0D0) CYCLE
      DOPPLE(j, nel) = SQRT(2.0D0 * TK(j) / AMASSISO(1, nel) / AMU + VTURB(j)**2) / CLIGHT
      XNFDOP(j, nel) = XNFP(j, nel) / DOPPLE(j, nel) / RHO(j)
    END DO
  END DO
  CALL KAPCONT

  IF (LEN_TRIM(seldir) .GT. 0) THEN
    ! The reference's own selection from the raw lists
    DATADIR = TRIM(seldir) // '/'
    CALL SELECTLINES
    nl = NLINES_STORED
    OPEN(NEWUNIT=u, FILE=TRIM(fout)//'.sel', ACCESS='STREAM', FORM='UNFORMATTED', &
         STATUS='REPLACE')
    WRITE(u) nl
    WRITE(u) ((LINEDATA(i, j), i = 1, 4), j = 1, nl)
    WRITE(u) WAVETAB(1:NWAVE)
    CLOSE(u)
    CALL read_diatomics_for_atlas(lines_list_path=TRIM(DATADIR)//'lines.list', &
                                  datadir=TRIM(DATADIR), wlbeg_nm=WAVETAB(1), &
                                  wlend_nm=WAVETAB(NWAVE), recs=drec_l, n_recs=nd_l)
    OPEN(NEWUNIT=u, FILE=TRIM(fout)//'.dia', ACCESS='STREAM', FORM='UNFORMATTED', &
         STATUS='REPLACE')
    WRITE(u) nd_l
    IF (nd_l .GT. 0) THEN
      WRITE(u) drec_l(1:nd_l)%wlvac_nm, drec_l(1:nd_l)%gflog_dex, drec_l(1:nd_l)%elo_cm, &
               drec_l(1:nd_l)%gammar_log, drec_l(1:nd_l)%gammaw_log, drec_l(1:nd_l)%molcode
    END IF
    CLOSE(u)
  ELSE
    ! The line store (SELECTLINES' product) from the caller's file
    OPEN(NEWUNIT=u, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
    READ(u) nl
    IF (ALLOCATED(LINEDATA)) DEALLOCATE(LINEDATA)
    ALLOCATE(LINEDATA(4, MAX(nl, 1)))
    READ(u) ((LINEDATA(i, j), i = 1, 4), j = 1, nl)
    CLOSE(u)
    NLINES_STORED = nl
  END IF

  OPEN(NEWUNIT=u, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(u) NRHOX, NUMNU, nl, MION, NWAVE
  WRITE(u) TEFF, FLUX
  WRITE(u) WAVESET(1:NUMNU), RCOSET(1:NUMNU)
  WRITE(u) T(1:NRHOX), HKT(1:NRHOX), HCKT(1:NRHOX), XNE(1:NRHOX), RHOX(1:NRHOX), &
           XNF(1:NRHOX, 1), XNF(1:NRHOX, 3), XNF(1:NRHOX, 841)
  WRITE(u) XNFDOP(1:NRHOX, 1:MION), DOPPLE(1:NRHOX, 1:MION)
  WRITE(u) TABCONT(1:NRHOX, 1:NWAVE + 1), IWAVETAB(1:NWAVE + 1)
  WRITE(u) H0TAB(1:2001), H1TAB(1:2001), H2TAB(1:2001)
  WRITE(u) COEFJ, COEFH

  ! LINOP1 on its own
  CALL LINOP1
  WRITE(u) XLINES(1:NRHOX, 1:NUMNU)

  ! KAPP at every frequency, with the frequency set-up of RUN_RT_PASS
  DO nu = NULO, NUHI
    WAVE = WAVESET(nu)
    FREQ = CLIGHT_NMS / WAVE
    WAVENO = 1.0D7 / WAVE
    FREQLG = LOG(FREQ)
    freq15_l = FREQ / 1.0D15
    EHVKT = EXP(-FREQ * HKT)
    STIM = 1.0D0 - EHVKT
    BNU = BNU_PREFAC * freq15_l**3 * EHVKT / STIM
    CALL KAPP
    WRITE(u) ACONT(1:NRHOX), SIGMAC(1:NRHOX), SCONT(1:NRHOX), SLINE(1:NRHOX), &
             ALINE(1:NRHOX), SIGMAL(1:NRHOX)
  END DO

  ! The reference's own frequency loop with its accumulators
  CALL RUN_RT_PASS(.FALSE.)
  WRITE(u) RJMINS(1:NRHOX), FLXRAD(1:NRHOX), RADEN(1:NRHOX), ACCRAD(1:NRHOX), &
           ABROSS(1:NRHOX), PRADK0, RDIAGJNU(1:NRHOX), HNU(1:NRHOX), TAUNU(1:NRHOX)
  CLOSE(u)

  ! With A12_TCORR_OUT set: the driver's end of iteration up to the
  ! temperature correction (src/atlas12c.f90:581-587: ROSS(3), the ROSSTAB
  ! store, RADIAP(3), COMPUTE_HEIGHT, CONVEC), the state TCORR(3) reads,
  ! and TCORR(3) itself with its per-iteration table on unit 66.  The file
  ! <A12_TCORR_OUT>.fin gets what the sections after the table (G2-K) read
  ! and what TCORR(3) leaves in the module arrays, TTAUP called on its own
  ! on the same state, ROSSTAB lookups in its three modes, and, on a deck
  ! that convects, CZ_CONSTRUCT on its own on inputs planted to engage it.
  CALL GET_ENVIRONMENT_VARIABLE('A12_TCORR_OUT', envval, envlen, envstat)
  IF (envstat .EQ. 0 .AND. envlen .GT. 0) THEN
    CALL ROSS(3, 0.0D0)
    rx_l = ROSSTAB(0.0D0, 0.0D0, 0.0D0)
    CALL RADIAP(3, 0.0D0)
    CALL COMPUTE_HEIGHT
    IF (IFPRES .EQ. 1 .AND. IFCONV .EQ. 1) THEN
      ! <A12_TCORR_OUT>.cnv: CONVEC's inputs, the four EOS evaluations it
      ! makes (EDENS and RHO at T(1.001), T(0.999), P(1.001), P(0.999),
      ! src/atlas12_modules.f90:10457-10535, taken here with the same
      ! sequence of perturbations from a snapshot that NEWTON_SAVE_STATE /
      ! NEWTON_RESTORE_STATE put back exactly), then what CONVEC leaves.
      CALL NEWTON_SAVE_STATE
      IFEDNS = 1
      TLOG = TLOG + 0.0009995003D0
      T    = T * 1.001D0
      TK   = TK * 1.001D0
      HKT  = HKT / 1.001D0
      HCKT = HCKT / 1.001D0
      TKEV = TKEV * 1.001D0
      ITEMP = ITEMP + 1
      CALL COMPUTE_ONE_POP(0.0D0, 1, XNE)
      cv_e(1:NRHOX, 1) = EDENS(1:NRHOX)
      cv_r(1:NRHOX, 1) = RHO(1:NRHOX)
      TLOG = TLOG - 0.0009995003D0 - 0.0010005003D0
      T    = T / 1.001D0 * 0.999D0
      TK   = TK / 1.001D0 * 0.999D0
      HKT  = HKT * 1.001D0 / 0.999D0
      HCKT = HCKT * 1.001D0 / 0.999D0
      TKEV = TKEV / 1.001D0 * 0.999D0
      ITEMP = ITEMP + 1
      CALL COMPUTE_ONE_POP(0.0D0, 1, XNE)
      cv_e(1:NRHOX, 2) = EDENS(1:NRHOX)
      cv_r(1:NRHOX, 2) = RHO(1:NRHOX)
      TLOG = TLOG + 0.0010005003D0
      T    = T / 0.999D0
      TK   = TK / 0.999D0
      HKT  = HKT * 0.999D0
      HCKT = HCKT * 0.999D0
      TKEV = TKEV / 0.999D0
      P    = P * 1.001D0
      ITEMP = ITEMP + 1
      CALL COMPUTE_ONE_POP(0.0D0, 1, XNE)
      cv_e(1:NRHOX, 3) = EDENS(1:NRHOX)
      cv_r(1:NRHOX, 3) = RHO(1:NRHOX)
      P = P / 1.001D0 * 0.999D0
      ITEMP = ITEMP + 1
      CALL COMPUTE_ONE_POP(0.0D0, 1, XNE)
      cv_e(1:NRHOX, 4) = EDENS(1:NRHOX)
      cv_r(1:NRHOX, 4) = RHO(1:NRHOX)
      CALL NEWTON_RESTORE_STATE
      OPEN(NEWUNIT=uc, FILE=TRIM(envval)//'.cnv', ACCESS='STREAM', FORM='UNFORMATTED', &
           STATUS='REPLACE')
      WRITE(uc) NRHOX, IFCORR, NCONV, MERGE(1, 0, USE_4POINT_MLT), MERGE(1, 0, USE_FLXCNV_SMOOTH)
      WRITE(uc) GRAV, MIXLTH, FLUX, OVERWT
      WRITE(uc) T(1:NRHOX), TK(1:NRHOX), HKT(1:NRHOX), HCKT(1:NRHOX), TKEV(1:NRHOX), &
                TLOG(1:NRHOX), P(1:NRHOX), PTOTAL(1:NRHOX), RHO(1:NRHOX), PRADK(1:NRHOX), &
                TAUROS(1:NRHOX), RHOX(1:NRHOX), ABROSS(1:NRHOX), VTURB(1:NRHOX), &
                HEIGHT(1:NRHOX), FLXRAD(1:NRHOX), FLXCNV1(1:NRHOX)
      WRITE(uc) cv_e(1:NRHOX, 1:4), cv_r(1:NRHOX, 1:4)
      WRITE(uc) ROSSTAB_NROSS
      WRITE(uc) ROSSTAB_ZEROT, ROSSTAB_SLOPET, ROSSTAB_ZEROP, ROSSTAB_SLOPEP
      WRITE(uc) ROSSTAB_TABT(1:ROSSTAB_NROSS), ROSSTAB_TABP(1:ROSSTAB_NROSS), &
                ROSSTAB_ROSS(1:ROSSTAB_NROSS)
      CALL CONVEC(.FALSE.)
      WRITE(uc) T(1:NRHOX), TK(1:NRHOX), HKT(1:NRHOX), HCKT(1:NRHOX), TKEV(1:NRHOX), &
                TLOG(1:NRHOX), P(1:NRHOX), ITEMP
      WRITE(uc) DLTDLP(1:NRHOX), HEATCP(1:NRHOX), VELSND(1:NRHOX), DLRDLT(1:NRHOX), &
                GRDADB(1:NRHOX), HSCALE(1:NRHOX), VCONV(1:NRHOX), FLXCNV(1:NRHOX), &
                FLXCNV0(1:NRHOX), FLXCNV1(1:NRHOX)
      ! Overshooting (OVERWT > 0; the reference's default is 0): CONVEC again
      ! from the same snapshot with OVERWT = 0.5, then once more with the
      ! default so that TCORR(3) below reads what the first call left.
      CALL NEWTON_RESTORE_STATE
      OVERWT = 0.5D0
      CALL CONVEC(.FALSE.)
      WRITE(uc) FLXCNV(1:NRHOX), FLXCNV0(1:NRHOX), FLXCNV1(1:NRHOX)
      CALL NEWTON_RESTORE_STATE
      OVERWT = 0.0D0
      CALL CONVEC(.FALSE.)
      CLOSE(uc)
    END IF
    OPEN(NEWUNIT=u, FILE=TRIM(envval), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
    WRITE(u) NRHOX, IFCONV, ITER
    WRITE(u) FLUX, TEFF, MIXLTH
    WRITE(u) TAUROS(1:NRHOX), RHOX(1:NRHOX), T(1:NRHOX), DLTDLP(1:NRHOX), ABROSS(1:NRHOX), &
             FLXCNV(1:NRHOX), FLXCNV0(1:NRHOX), GRDADB(1:NRHOX), PTOTAL(1:NRHOX), &
             RHO(1:NRHOX), DLRDLT(1:NRHOX), HEATCP(1:NRHOX), HSCALE(1:NRHOX), &
             FLXRAD(1:NRHOX), RJMINS(1:NRHOX)
    CLOSE(u)
    OPEN(NEWUNIT=uf, FILE=TRIM(envval)//'.fin', ACCESS='STREAM', FORM='UNFORMATTED', &
         STATUS='REPLACE')
    WRITE(uf) ROSSTAB_NROSS, IROSSTAB, J1SMOOTH, J2SMOOTH, NUMITS, ITEMP
    WRITE(uf) GRAV, TAU1LG, STEPLG, PRADK0, TFLOOR_ATM, WTJM1, WTJ, WTJP1
    WRITE(uf) ROSSTAB_ZEROT, ROSSTAB_SLOPET, ROSSTAB_ZEROP, ROSSTAB_SLOPEP
    WRITE(uf) ROSSTAB_TABT(1:ROSSTAB_NROSS), ROSSTAB_TABP(1:ROSSTAB_NROSS), &
              ROSSTAB_ROSS(1:ROSSTAB_NROSS)
    WRITE(uf) P(1:NRHOX), XNE(1:NRHOX), PRAD(1:NRHOX), PTURB(1:NRHOX), VTURB(1:NRHOX), &
              BMIN(1:NRHOX), ACCRAD(1:NRHOX), BHYD(1:NRHOX, 1:6)
    ! TTAUP on its own: T and PRAD mapped onto TAUSTD as section (H) maps them
    DO j = 1, NRHOX
      taustd_l(j) = 10.0D0**(TAU1LG + (j - 1) * STEPLG)
    END DO
    idum_l = MAP1(TAUROS, T, NRHOX, taustd_l, tn_l, NRHOX)
    idum_l = MAP1(TAUROS, PRAD, NRHOX, taustd_l, pr_l, NRHOX)
    CALL TTAUP(tn_l, taustd_l, ab_l, pt_l, pg_l, pr_l, PTURB, VTURB, GRAV, NRHOX)
    WRITE(uf) tn_l(1:NRHOX), pr_l(1:NRHOX), ab_l(1:NRHOX), pt_l(1:NRHOX), pg_l(1:NRHOX)
    OPEN(UNIT=66, FILE=TRIM(envval)//'.iter', STATUS='REPLACE')
    IFCORR = 1
    IFPRNT(ITER) = 1
    CALL TCORR(3, 0.0D0)
    CLOSE(66)
    ! What TCORR(3) leaves behind (sections I-K)
    WRITE(uf) RHOX(1:NRHOX), T(1:NRHOX), TK(1:NRHOX), HKT(1:NRHOX), HCKT(1:NRHOX), &
              TKEV(1:NRHOX), TLOG(1:NRHOX), P(1:NRHOX), XNE(1:NRHOX), ABROSS(1:NRHOX), &
              PRAD(1:NRHOX), PRADK(1:NRHOX), VTURB(1:NRHOX), BMIN(1:NRHOX), PTURB(1:NRHOX), &
              TAUROS(1:NRHOX), BHYD(1:NRHOX, 1:6), FLXERR(1:NRHOX)
    WRITE(uf) ITEMP, MERGE(1, 0, CZC_RUN_ENGAGED), CZC_RUN_JC
    ! ROSSTAB lookups: a second profile stored (the corrected structure), then
    ! queries around every layer of the first, in modes 1, 2 and 3
    t0_l(1:NRHOX) = tn_l(1:NRHOX)
    p0_l(1:NRHOX) = pg_l(1:NRHOX)
    rx_l = ROSSTAB(0.0D0, 0.0D0, 0.0D0)
    nq = 0
    DO j = 1, NRHOX
      DO k = 1, 12
        nq = nq + 1
        qt_l(nq) = t0_l(j) * (0.94D0 + 0.01D0 * k)
        qp_l(nq) = p0_l(j) * 10.0D0**(0.3D0 * MOD(k * 7, 5) - 0.6D0)
      END DO
    END DO
    WRITE(uf) ROSSTAB_NROSS, nq
    WRITE(uf) ROSSTAB_TABT(1:ROSSTAB_NROSS), ROSSTAB_TABP(1:ROSSTAB_NROSS), &
              ROSSTAB_ROSS(1:ROSSTAB_NROSS)
    WRITE(uf) qt_l(1:nq), qp_l(1:nq)
    DO irt = 1, 3
      IROSSTAB = irt
      DO iq = 1, nq
        qr_l(iq) = ROSSTAB(qt_l(iq), qp_l(iq), 0.0D0)
      END DO
      WRITE(uf) qr_l(1:nq)
    END DO
    IROSSTAB = 1
    ! CZ_CONSTRUCT on its own (IFCONV = 1), on planted inputs that engage
    ! it: convective flux 0.9 FLUX in the deepest 24 layers (they pass the
    ! efficiency gate; their nabla - nabla_ad is CONVEC's), and a flux
    ! error of 8, 4, then 0.5 % over three calls (engage, stay engaged,
    ! release), T1 = 0 on entry to each.
    IF (IFCONV .EQ. 1) THEN
      FLXCNV(NRHOX - 23:NRHOX) = 0.9D0 * FLUX
      WRITE(uf) 1
      WRITE(uf) FLXCNV(1:NRHOX), FLXRAD(1:NRHOX), DLTDLP(1:NRHOX), GRDADB(1:NRHOX), &
                T(1:NRHOX), RHOX(1:NRHOX), PTOTAL(1:NRHOX), RHO(1:NRHOX), HEATCP(1:NRHOX), &
                DLRDLT(1:NRHOX), P(1:NRHOX), VTURB(1:NRHOX), ABROSS(1:NRHOX), &
                HSCALE(1:NRHOX)
      WRITE(uf) GRAV, FLUX, MIXLTH, MERGE(1, 0, USE_4POINT_MLT)
      DO k = 1, 3
        ITER = k
        FLXERR(1:NRHOX) = 8.0D0 / 2.0D0**(2 * (k - 1) - MERGE(1, 0, k .EQ. 2))
        t1_l(1:NRHOX) = 0.0D0
        CALL CZ_CONSTRUCT(t1_l)
        WRITE(uf) FLXERR(1:NRHOX), t1_l(1:NRHOX), MERGE(1, 0, CZC_RUN_ENGAGED), CZC_RUN_JC
      END DO
    ELSE
      WRITE(uf) 0
    END IF
    CLOSE(uf)
  END IF

  ! With A12_KAPP_DUMP and A12_KAPP_WAVES set: KAPP's inputs (the depth
  ! state and the tables its sources load on first use), then, at every
  ! wavelength of A12_KAPP_WAVES (int32 count, then REAL(8) nm), the
  ! frequency set-up of RUN_RT_PASS, KAPP, and every per-source array.
  CALL GET_ENVIRONMENT_VARIABLE('A12_KAPP_DUMP', envval, envlen, envstat)
  IF (envstat .EQ. 0 .AND. envlen .GT. 0) CALL dump_kapp(TRIM(envval))

CONTAINS

  SUBROUTINE dump_kapp(path)
    CHARACTER(LEN=*), INTENT(IN) :: path
    CHARACTER(LEN=512) :: wpath
    INTEGER :: uo, uw, nw, iw, is, il, s, nt, k
    REAL(8), ALLOCATABLE :: waves(:)
    REAL(8) :: eqh2(kw)

    CALL GET_ENVIRONMENT_VARIABLE('A12_KAPP_WAVES', wpath)
    OPEN(NEWUNIT=uw, FILE=TRIM(wpath), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
    READ(uw) nw
    ALLOCATE(waves(nw))
    READ(uw) waves
    CLOSE(uw)
    DO k = 1, NRHOX
      eqh2(k) = EQUILH2(T(k))
    END DO

    OPEN(NEWUNIT=uo, FILE=path, ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
    WRITE(uo) NRHOX, MION, kw, nw
    WRITE(uo) HCK, HOVERK, KBOL_EV, THETA_COEFF, BNU_PREFAC, SAHA_PREFAC, FREQ_RYDH, RYD_ANG, &
              LN10, COEFF_FF, SIGMA_THOMSON, CLIGHT_ANGS
    WRITE(uo) T(1:NRHOX), TKEV(1:NRHOX), TLOG(1:NRHOX), HKT(1:NRHOX), HCKT(1:NRHOX), &
              XNE(1:NRHOX), RHO(1:NRHOX), BMIN(1:NRHOX), BHYD(1:NRHOX, 1:6), eqh2(1:NRHOX)
    WRITE(uo) XNF(1:NRHOX, 1:MION), XNFP(1:NRHOX, 1:MION)
    WRITE(uo) Q_HOLTSMARK, LOG_BETA_MIN, LOG_BETA_STEP
    ! TOPbase levels as INIT_MBF_TOPBASE left them
    WRITE(uo) N_TB_SPECIES
    DO is = 1, N_TB_SPECIES
      WRITE(uo) TB_SPECIES(is)%XNFP_INDEX, TB_SPECIES(is)%NLEVELS, TB_SPECIES(is)%LOWEST_THR_RY
      DO il = 1, TB_SPECIES(is)%NLEVELS
        WRITE(uo) TB_SPECIES(is)%LEVELS(il)%G_STAT, TB_SPECIES(is)%LEVELS(il)%NP, &
                  TB_SPECIES(is)%LEVELS(il)%E_BIND_RY, TB_SPECIES(is)%LEVELS(il)%E_EXC_RY, &
                  TB_SPECIES(is)%LEVELS(il)%LOG_E_MIN, TB_SPECIES(is)%LEVELS(il)%LOG_E_STRIDE
        WRITE(uo) TB_SPECIES(is)%LEVELS(il)%E_PHOTON_RY, TB_SPECIES(is)%LEVELS(il)%SIGMA_MB
      END DO
    END DO
    ! hotop.dat edges kept by INIT_MBF_HIGH_ION
    WRITE(uo) N_HIGH_ION_EDGES
    WRITE(uo) HI_EDGE_DATA(1:7, 1:N_HIGH_ION_EDGES)
    ! Fe I / Fe II (INIT_FELO)
    WRITE(uo) FELO_NLEV, FELO_XNFP_IDX, FELO_NUTH, FELO_G, FELO_E, FELO_IFANCY, FELO_NPTS, &
              FELO_X, FELO_LOGSIG, FELO_SEATON
    ! Stehle hydrogen Stark tables (INIT_STARK_TABLES)
    DO s = 1, NSTARK_SERIES
      k = 0
      IF (STEHLE_DATA(s)%loaded) k = 1
      WRITE(uo) k, STEHLE_DATA(s)%n_lower, STEHLE_DATA(s)%n_upper_min, &
                STEHLE_DATA(s)%n_upper_max, STEHLE_DATA(s)%n_dens, &
                STEHLE_DATA(s)%density_grid, STEHLE_DATA(s)%temp_grid, &
                STEHLE_DATA(s)%log_dalpha_grid
      IF (k .EQ. 1) THEN
        nt = STEHLE_DATA(s)%n_upper_max - STEHLE_DATA(s)%n_upper_min + 1
        WRITE(uo) STEHLE_DATA(s)%max_dens_idx(1:nt), STEHLE_DATA(s)%k_alpha(1:nt)
        WRITE(uo) STEHLE_DATA(s)%profiles
      END IF
    END DO

    DO iw = 1, nw
      WAVE = waves(iw)
      FREQ = CLIGHT_NMS / WAVE
      WAVENO = 1.0D7 / WAVE
      FREQLG = LOG(FREQ)
      freq15_l = FREQ / 1.0D15
      EHVKT = EXP(-FREQ * HKT)
      STIM = 1.0D0 - EHVKT
      BNU = BNU_PREFAC * freq15_l**3 * EHVKT / STIM
      CALL KAPP
      WRITE(uo) AHYD(1:NRHOX), SHYD(1:NRHOX), AH2P(1:NRHOX), AHMIN(1:NRHOX), SHMIN(1:NRHOX), &
                SIGH(1:NRHOX), AHE1(1:NRHOX), AHE2(1:NRHOX), AHEMIN(1:NRHOX), SIGHE(1:NRHOX), &
                AH2COLL(1:NRHOX), AH2MIN(1:NRHOX), ACONT_METAL(1:NRHOX), SIGEL(1:NRHOX), &
                SIGH2(1:NRHOX), AHLINE(1:NRHOX), SHLINE(1:NRHOX), &
                ACONT(1:NRHOX), SIGMAC(1:NRHOX), SCONT(1:NRHOX), SLINE(1:NRHOX), &
                ALINE(1:NRHOX), SIGMAL(1:NRHOX)
    END DO
    CLOSE(uo)
  END SUBROUTINE dump_kapp

END PROGRAM ref_atlas_dump
