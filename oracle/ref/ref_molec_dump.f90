! ref_molec_dump -- READIN on a model deck, the first-iteration populations
! (COMPUTE_ONE_POP, which runs NMOLEC), then NMOLEC(1) once more, and
! records what that last call's gas-phase Newton solve
! (src/atlas12_modules.f90:9078-9217) reads and writes:
!   * the molecule table (NEQUA, NUMMOL, LOCJ, KCOMPS, IDEQUA, XNMOLCODE),
!   * per depth P, TK, T, XABUND and the equilibrium constants EQUILJ as
!     NMOLEC forms them (:8897-8946): with XNE(J) at that point the
!     previous depth's converged value times P(J)/P(J-1), or the surface
!     guess at J = 1 (:8851-8882),
!   * the converged XNATOM, XNE and XNMOL.
! Test infrastructure (built by oracle/ref/Makefile).
!
!   ATLAS12=<root> ref_molec_dump <deck> <out.bin>
PROGRAM ref_molec_dump
  USE mod_atlas_data
  IMPLICIT NONE
  CHARACTER(LEN=512) :: deck, fout, envval
  INTEGER :: envlen, envstat, u, j, jmol, ncomp, ion, id, nlc
  REAL(8) :: xnconv(kw), eqj(maxmol, kw), frac(kw, 6), xtot, xn1

  CALL GET_COMMAND_ARGUMENT(1, deck)
  CALL GET_COMMAND_ARGUMENT(2, fout)
  CALL GET_ENVIRONMENT_VARIABLE('ATLAS12', envval, envlen, envstat)
  DATADIR = TRIM(envval) // '/data/'
  CALL set_bc_data_dir(TRIM(DATADIR))
  INPUTDATA = 5
  CALL IONPOTS
  ITEMP = 0
  INPUT_MODEL_FILE = deck
  NUMITS = 1
  IFPNCH(1) = 0
  IFPRNT(1) = 1
  CALL READIN(1)
  ITER = 1
  ITEMP = ITEMP + ITER
  PZERO = PCON + PRADK0 + PTURB0
  DO j = 1, NRHOX
    CHARGESQ(j) = XNE(j) * 2.0D0
    IF (2.0D0 * XNE(j) - P(j) / TK(j) .GT. 0.0D0) &
      CHARGESQ(j) = CHARGESQ(j) + 2.0D0 * (2.0D0 * XNE(j) - P(j) / TK(j))
    PTOTAL(j) = GRAV * RHOX(j) + PZERO
  END DO
  IFEDNS = 0
  CALL COMPUTE_ONE_POP(0.D0, 1, XNE)
  CALL NMOLEC(1)

  ! The equilibrium constants of that call, depth by depth, with XNE(J)
  ! set to what NMOLEC held when it formed them.
  xnconv(1:NRHOX) = XNE(1:NRHOX)
  eqj = 0.0D0
  DO j = 1, NRHOX
    IF (j .EQ. 1) THEN
      xtot = P(1) / TK(1)
      xn1 = xtot / 2.0D0
      IF (T(1) .LT. 4000.0D0) xn1 = xtot
      XNE(1) = xn1 / 10.0D0
    ELSE
      XNE(j) = xnconv(j - 1) * (P(j) / P(j - 1))
    END IF
    DO jmol = 1, NUMMOL
      ncomp = LOCJ(jmol + 1) - LOCJ(jmol)
      IF (EQUIL(1, jmol) .NE. 0.0D0) THEN
        ion = int((XNMOLCODE(jmol) - aint(XNMOLCODE(jmol))) * 100.0D0 + 0.5D0)
        IF (XNMOLCODE(jmol) .EQ. 101.0D0) THEN
          IF (T(j) .LE. 20000.0D0) eqj(jmol, j) = EQUILH2(T(j))
        ELSE IF (T(j) .LE. 10000.0D0) THEN
          eqj(jmol, j) = exp(EQUIL(1, jmol) / TKEV(j) - 1.5D0 * (ncomp - ion - ion - 1) * TLOG(j) &
                             + EQUIL(2, jmol) + EQUIL(3, jmol) / T(j) + EQUIL(4, jmol) * TLOG(j) &
                             + EQUIL(5, jmol) * T(j) + EQUIL(6, jmol) / (T(j) * T(j)))
        END IF
      ELSE IF (ncomp .EQ. 1) THEN
        eqj(jmol, j) = 1.0D0
      ELSE
        id = int(XNMOLCODE(jmol))
        ion = ncomp - 1
        CALL PFSAHA(j, id, ncomp, 12, frac)
        eqj(jmol, j) = frac(j, ncomp) / frac(j, 1) * XNE(j)**ion
      END IF
    END DO
    XNE(j) = xnconv(j)
  END DO

  nlc = LOCJ(NUMMOL + 1) - 1
  OPEN(NEWUNIT=u, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(u) NRHOX, NEQUA, NUMMOL, nlc
  WRITE(u) LOCJ(1:NUMMOL + 1), KCOMPS(1:nlc), IDEQUA(1:NEQUA)
  WRITE(u) XNMOLCODE(1:NUMMOL)
  WRITE(u) P(1:NRHOX), TK(1:NRHOX), T(1:NRHOX)
  WRITE(u) XABUND(1:NRHOX, 1:99)
  WRITE(u) eqj(1:NUMMOL, 1:NRHOX)
  WRITE(u) XNATOM(1:NRHOX), XNE(1:NRHOX), XNMOL(1:NRHOX, 1:NUMMOL)
  CLOSE(u)
END PROGRAM ref_molec_dump
