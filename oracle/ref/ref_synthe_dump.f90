! ref_synthe_dump -- runs the reference's SYNTHE set-up stages (READIN,
! run_xnfpelsyn, run_mklinelist, BLOCKJ/BLOCKH) on a model deck and writes
! the state that the SYNTHE hot loops (src/synthe.f90:476-718) read, so the
! oracle and the HIP library can be fed exactly what the reference's loops
! see.  Test infrastructure (built by oracle/ref/Makefile).
!
!     ATLAS12=<root> ref_synthe_dump <model> <wlbeg> <wlend> <resolu> <out.bin>
!
! Output (stream, little-endian): see tests/golden/make_golden.py.
PROGRAM ref_synthe_dump
  USE synthe_module
  USE mod_mklinelist, ONLY: run_mklinelist, lte_lines, nlines_lte, nlines_nlte
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
END PROGRAM ref_synthe_dump
