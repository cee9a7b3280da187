! ref_synthe_dump -- runs the reference's SYNTHE set-up stages (READIN,
! run_xnfpelsyn, run_mklinelist, BLOCKJ/BLOCKH) on a model deck and writes
! the state that the SYNTHE hot loops (src/synthe.f90:476-718) read, so the
! oracle and the HIP library can be fed exactly what the reference's loops
! see.  Test infrastructure (built by oracle/ref/Makefile).
!
!     ATLAS12=<root> ref_synthe_dump <model> <wlbeg> <wlend> <resolu> <out.bin>
!
! With A12_CPLX_DUMP=<file>: the driver's per-depth set-up (src/synthe.f90:
! 374-551) and compute_line_opacity (the complex-profile lines of
! nlte_lines) at every depth, with the line records, the per-depth inputs
! and the CONTINUUM and BUFFER arrays written to <file>.
!
! Output (stream, little-endian): see tests/golden/make_golden.py.
PROGRAM ref_synthe_dump
  USE synthe_module
  USE mod_mklinelist, ONLY: run_mklinelist, lte_lines, nlines_lte, nlines_nlte, nlte_lines
  USE mod_parameters, ONLY: LINE_CUTOFF
  USE mod_atlas_data, ONLY: READIN, BLOCKJ, BLOCKH, set_bc_data_dir, DATADIR, &
      IFSYNTHE, TURBV_UNSET, TEFF, COEFJ, COEFH, &
      hkt_a => HKT, nrhox_a => NRHOX, rhox_a => RHOX, itemp_a => ITEMP
  IMPLICIT NONE
  CHARACTER(LEN=512) :: model_file, fout, arg, envval
  INTEGER :: envlen, envstat, u, n, i
  REAL(8) :: wlb, wle, res, rat, ratlg, wbeg
  INTEGER :: len_grid, ixb

  CALL GET_COMMAND_ARGUMENT(1, model_file)
  CALL GET_COMMAND_ARGUMENT(2, arg); READ(arg, *) wlb
  CALL GET_COMMAND_ARGUMENT(3, arg); READ(arg, *) wle
  CALL GET_COMMAND_ARGUMENT(4, arg); READ(arg, *) res
  CALL GET_COMMAND_ARGUMENT(5, fout)

  IFSYNTHE = 1
  CALL GET_ENVIRONMENT_VARIABLE('ATLAS12', envval, envlen, envstat)
  DATADIR = TRIM(envval) // '/data/'
  CALL set_bc_data_dir(TRIM(DATADIR))

  ! Grid as the SYNTHE driver defines it (src/synthe.f90:316-324).
  rat = 1.0D0 + 1.0D0 / res
  ratlg = LOG(rat)
  len_grid = INT(LOG(wle / wlb) / ratlg)
  ixb = INT(LOG(wlb) / ratlg)
  wbeg = EXP(DBLE(ixb) * ratlg)
  IF (wbeg .LT. wlb) THEN
    ixb = ixb + 1
    wbeg = EXP(DBLE(ixb) * ratlg)
  END IF

  itemp_a = 1
  OPEN(UNIT=5, FILE=TRIM(model_file), STATUS='OLD', ACTION='READ')
  CALL readin(20)
  CALL run_xnfpelsyn(TURBV_UNSET)
  CLOSE(UNIT=5)
  CALL run_mklinelist(wlb, wle, res, TEFF, TRIM(DATADIR) // 'lines.list', DATADIR)
  CALL BLOCKJ
  CALL BLOCKH

  n = nrhox_a
  OPEN(NEWUNIT=u, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(u) n, len_grid, ixb, nedge_m, numnu_m, mw, nlines_lte, nlines_nlte
  WRITE(u) wbeg, rat, res
  WRITE(u) (wledge_m(i), i = 1, nedge_m)
  WRITE(u) continall_m(1:numnu_m, 1:n)
  WRITE(u) contabs_m(1:numnu_m, 1:n)
  WRITE(u) contscat_m(1:numnu_m, 1:n)
  WRITE(u) xnfpel_m(1:6, 1:mw, 1:n)
  WRITE(u) dopple_m(1:6, 1:mw, 1:n)
  WRITE(u) xf_t(1:n), xf_xne(1:n), xf_rho(1:n), xf_xnfh(1:n), xf_xnfhe(1:n,1), &
           xf_xnfh2(1:n), xf_hckt(1:n), xf_hkt(1:n), hkt_a(1:n), rhox_a(1:n)
  WRITE(u) COEFJ, COEFH
  DO i = 1, nlines_lte
    WRITE(u) lte_lines(i)%nbuff, lte_lines(i)%cgf, lte_lines(i)%nelion, &
             lte_lines(i)%elo, lte_lines(i)%gamrf, lte_lines(i)%gamsf, &
             lte_lines(i)%gamwf, lte_lines(i)%code
  END DO
  CLOSE(u)

  CALL GET_ENVIRONMENT_VARIABLE('A12_CPLX_DUMP', envval, envlen, envstat)
  IF (envstat .EQ. 0 .AND. envlen .GT. 0) CALL dump_cplx(TRIM(envval))

CONTAINS

  SUBROUTINE dump_cplx(path)
    CHARACTER(LEN=*), INTENT(IN) :: path
    REAL(8) :: wledge(me), halfedge(me), deledge(me), qablog(nf), ablog(3, me)
    REAL(8) :: qxnfpel(mw6), qdopple(mw6), wave8
    INTEGER :: uo, j, k, iedge, nu, nb, nel, ii, nedge

    ! the grid (src/synthe.f90:316-324)
    WLBEG = wlb
    WLEND = wle
    RESOLU = res
    RATIO = rat
    RATIOLG = ratlg
    LENGTH = len_grid
    IXWLBEG = ixb
    WBEGIN = wbeg
    ! the depth structure (:374-404)
    NRHOX = nrhox_a
    nedge = nedge_m
    DO iedge = 1, nedge_m
      wledge(iedge) = ABS(wledge_m(iedge))
    END DO
    DO iedge = 2, nedge
      halfedge(iedge-1) = (wledge(iedge-1) + wledge(iedge)) * 0.5D0
      deledge(iedge-1) = (wledge(iedge) - wledge(iedge-1))**2 * 0.5D0
    END DO
    ITEMP = 1
    DO j = 1, NRHOX
      T(j) = REAL(xf_t(j))
      TKEV(j) = REAL(xf_tkev(j))
      TK(j) = REAL(xf_tk(j))
      HKT(j) = xf_hkt(j)
      TLOG(j) = REAL(xf_tlog(j))
      HCKT(j) = xf_hckt(j)
      P(j) = REAL(xf_p(j))
      XNE(j) = REAL(xf_xne(j))
      XNATOM(j) = REAL(xf_xnatom(j))
      RHO(j) = REAL(xf_rho(j))
      RHOX(j) = REAL(xf_rhox(j))
      VTURB(j) = REAL(xf_vturb(j))
      XNFH(j) = REAL(xf_xnfh(j))
      XNFHE(j,1) = REAL(xf_xnfhe(j,1))
      XNFHE(j,2) = REAL(xf_xnfhe(j,2))
      XNFH2(j) = REAL(xf_xnfh2(j))
    END DO

    OPEN(NEWUNIT=uo, FILE=path, ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
    WRITE(uo) NRHOX, LENGTH, nlines_nlte, mw6
    WRITE(uo) LINE_CUTOFF
    DO k = 1, nlines_nlte
      WRITE(uo) nlte_lines(k)%wlvac, nlte_lines(k)%elo, nlte_lines(k)%gf, nlte_lines(k)%nblo, &
                nlte_lines(k)%nbup, nlte_lines(k)%nelion, nlte_lines(k)%itype, nlte_lines(k)%ncon, &
                nlte_lines(k)%nelionx, nlte_lines(k)%gammar, nlte_lines(k)%gammas, &
                nlte_lines(k)%gammaw, nlte_lines(k)%nbuff, nlte_lines(k)%lim
    END DO
    WRITE(uo) T(1:NRHOX), XNE(1:NRHOX), RHO(1:NRHOX), HCKT(1:NRHOX), XNFHE(1:NRHOX,1), &
              XNFH2(1:NRHOX)

    ! per depth (:476-554)
    DO j = 1, NRHOX
      BUFFER(1:LENGTH) = 0.0
      DO nu = 1, numnu_m
        qablog(nu) = continall_m(nu, j)
      END DO
      nu = 0
      DO iedge = 1, nedge-1
        nu = nu + 1;  ablog(1,iedge) = qablog(nu)
        nu = nu + 1;  ablog(2,iedge) = qablog(nu)
        nu = nu + 1;  ablog(3,iedge) = qablog(nu)
      END DO
      iedge = 1
      DO nb = 1, LENGTH
        wave8 = WBEGIN * RATIO**(nb-1)
        DO WHILE (wave8 .GE. wledge(iedge+1) .AND. iedge .LT. nedge-1)
          iedge = iedge + 1
        END DO
        CONTINUUM(nb) = REAL( &
          ((wave8 - halfedge(iedge))*(wave8 - wledge(iedge+1))*ablog(1,iedge) + &
           (wledge(iedge) - wave8)*(wave8 - wledge(iedge+1))*2.0D0*ablog(2,iedge) + &
           (wave8 - wledge(iedge))*(wave8 - halfedge(iedge))*ablog(3,iedge)) / &
          deledge(iedge) )
        IF (CONTINUUM(nb) .LT. -15.0) CONTINUUM(nb) = -15.0
      END DO
      DO nb = 1, LENGTH
        CONTINUUM(nb) = 10.0**CONTINUUM(nb)
      END DO
      DO nel = 1, mw
        DO ii = 1, 6
          qxnfpel((nel-1)*6 + ii) = xnfpel_m(ii, nel, j)
          qdopple((nel-1)*6 + ii) = dopple_m(ii, nel, j)
        END DO
      END DO
      XNFPH(j,1) = REAL(qxnfpel(1))
      XNFPH(j,2) = REAL(qxnfpel(2))
      XNFPHE(j,1) = REAL(qxnfpel(7))
      XNFPHE(j,2) = REAL(qxnfpel(8))
      XNFPHE(j,3) = REAL(qxnfpel(9))
      DO ii = 1, mw6
        XNFPEL(ii) = 0.0D0
        IF (qxnfpel(ii) .LT. 1.0D25) XNFPEL(ii) = qxnfpel(ii)
      END DO
      DO ii = 1, mw6
        DOPPLE(ii) = qdopple(ii)
        XNFPEL(ii) = XNFPEL(ii) / RHO(j)
        IF (qdopple(ii) .GT. 0.0D0) THEN
          XNFDOP(ii) = qxnfpel(ii) / xf_rho(j) / qdopple(ii)
        ELSE
          XNFDOP(ii) = 0.0D0
        END IF
      END DO
      TXNXN(j) = REAL( (XNFH(j) + 0.42D0*XNFHE(j,1) + 0.85D0*XNFH2(j)) * &
                       (T(j)/10000.0D0)**0.3D0 )
      IF (nlines_nlte .GT. 0) CALL compute_line_opacity(j, nlines_nlte, LINE_CUTOFF, 0.0, 1)
      WRITE(uo) TXNXN(j), XNFPH(j,1), XNFPH(j,2)
      WRITE(uo) XNFPEL(1:mw6), DOPPLE(1:mw6), XNFDOP(1:mw6)
      WRITE(uo) CONTINUUM(1:LENGTH), BUFFER(1:LENGTH)
    END DO
    CLOSE(uo)
  END SUBROUTINE dump_cplx

END PROGRAM ref_synthe_dump
