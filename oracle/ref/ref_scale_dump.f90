! ref_scale_dump -- runs the reference's READIN on a model deck, then
! SCALE_MODEL(TEFF_NEW, LOGG_NEW), and records the structure before and
! after.  Test infrastructure (built by oracle/ref/Makefile).
!
!   ATLAS12=<root> ref_scale_dump <deck> <teff_new> <logg_new> <out.bin>
PROGRAM ref_scale_dump
  USE mod_atlas_data
  IMPLICIT NONE
  CHARACTER(LEN=512) :: deck, fout, arg, envval
  INTEGER :: envlen, envstat, u, n
  REAL(8) :: teff_new, logg_new

  CALL GET_COMMAND_ARGUMENT(1, deck)
  CALL GET_COMMAND_ARGUMENT(2, arg); READ(arg, *) teff_new
  CALL GET_COMMAND_ARGUMENT(3, arg); READ(arg, *) logg_new
  CALL GET_COMMAND_ARGUMENT(4, fout)
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

  n = NRHOX
  OPEN(NEWUNIT=u, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(u) n, kw
  WRITE(u) TEFF, GRAV, PRADK0, PCON, PTURB0, TAU1LG, STEPLG, teff_new, logg_new
  WRITE(u) RHOX(1:n), T(1:n), P(1:n), XNE(1:n), ABROSS(1:n), PRAD(1:n), VTURB(1:n)
  CALL SCALE_MODEL(teff_new, logg_new)
  WRITE(u) NRHOX
  WRITE(u) TAUSTD(1:kw), TAUROS(1:kw)
  WRITE(u) RHOX(1:kw), T(1:kw), P(1:kw), XNE(1:kw), ABROSS(1:kw), PRAD(1:kw), VTURB(1:kw), &
           PRADK(1:kw), PTOTAL(1:kw)
  WRITE(u) TEFF, GRAV, GLOG, FLUX, PRADK0, PZERO
  CLOSE(u)
END PROGRAM ref_scale_dump
