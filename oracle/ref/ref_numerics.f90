! ref_numerics -- calls the reference's own numerical routines on input
! vectors supplied in a stream binary file and writes their outputs.
!
! Test infrastructure: built by oracle/ref/Makefile against the reference
! modules compiled from the reference's src/; its outputs become golden
! vectors under tests/golden/.  Usage:
!     ref_numerics <op> <in.bin> <out.bin> [datadir/]
! Every input file starts with int32 header words described per op below;
! arrays are little-endian, Fortran order.
PROGRAM ref_numerics
  USE mod_atlas_data
  USE synthe_module, ONLY: voigt_profile
  IMPLICIT NONE
  CHARACTER(LEN=32)  :: op
  CHARACTER(LEN=512) :: fin, fout, ddir
  INTEGER :: uin, uout, n, m, cnt, k, i, nold, nnew, iscat, isurf, mj
  REAL(4), ALLOCATABLE :: v4(:), a4(:), r4(:)
  REAL(8), ALLOCATABLE :: x(:), f(:), g(:), aa(:), bb(:), cc(:), am(:,:), xn(:)
  REAL(8) :: start, hsurf
  INTEGER, ALLOCATABLE :: ipiv(:), nn(:)

  CALL GET_COMMAND_ARGUMENT(1, op)
  CALL GET_COMMAND_ARGUMENT(2, fin)
  CALL GET_COMMAND_ARGUMENT(3, fout)
  OPEN(NEWUNIT=uin, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  OPEN(NEWUNIT=uout, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')

  SELECT CASE (TRIM(op))
  CASE ('voigt')          ! n ; v(n) a(n) float -> H(n) float
    READ(uin) n
    ALLOCATE(v4(n), a4(n), r4(n))
    READ(uin) v4, a4
    DO i = 1, n
      r4(i) = voigt_profile(v4(i), a4(i))
    END DO
    WRITE(uout) r4
  CASE ('solvit')         ! n cnt ; cnt x (A(n,n), b(n)) -> cnt x b(n)
    READ(uin) n, cnt
    ALLOCATE(am(n,n), f(n), ipiv(n))
    DO k = 1, cnt
      READ(uin) am, f
      CALL SOLVIT(am, n, f, ipiv)
      WRITE(uout) f
    END DO
  CASE ('map1')           ! nold nnew cnt ; cnt x (xold, fold, xnew) -> (fnew, int32 ret)
    READ(uin) nold, nnew, cnt
    ALLOCATE(x(nold), f(nold), xn(nnew), g(nnew))
    DO k = 1, cnt
      READ(uin) x, f, xn
      mj = MAP1(x, f, nold, xn, g, nnew)
      WRITE(uout) g, mj
    END DO
  CASE ('integ')          ! n cnt ; cnt x (x, f, start) -> fint
    READ(uin) n, cnt
    ALLOCATE(x(n), f(n), g(n))
    DO k = 1, cnt
      READ(uin) x, f, start
      CALL INTEG(x, f, g, n, start)
      WRITE(uout) g
    END DO
  CASE ('parcoe')         ! n cnt ; cnt x (f, x) -> a, b, c
    READ(uin) n, cnt
    ALLOCATE(x(n), f(n), aa(n), bb(n), cc(n))
    DO k = 1, cnt
      READ(uin) f, x
      CALL PARCOE(f, x, aa, bb, cc, n)
      WRITE(uout) aa, bb, cc
    END DO
  CASE ('deriv')          ! n cnt ; cnt x (x, f) -> dfdx
    READ(uin) n, cnt
    ALLOCATE(x(n), f(n), g(n))
    DO k = 1, cnt
      READ(uin) x, f
      CALL DERIV(x, f, g, n)
      WRITE(uout) g
    END DO
  CASE ('expi')           ! cnt ; nn(cnt) int32, x(cnt) -> E_n(x)
    READ(uin) cnt
    ALLOCATE(nn(cnt), x(cnt), g(cnt))
    READ(uin) nn, x
    DO k = 1, cnt
      g(k) = EXPI(nn(k), x(k))
    END DO
    WRITE(uout) g
  CASE ('josh')           ! n cnt ; cnt x (ifscat, ifsurf, ACONT ALINE SIGMAC SIGMAL
                          !   SCONT SLINE RHOX BNU) -> HNU1, KNU1, TAUNU SNU JNU HNU JMINS
    CALL GET_COMMAND_ARGUMENT(4, ddir)
    DATADIR = TRIM(ddir)
    CALL BLOCKJ
    CALL BLOCKH
    READ(uin) n, cnt
    NRHOX = n
    FREQ = 1.0D15
    DO k = 1, cnt
      READ(uin) iscat, isurf
      READ(uin) ACONT(1:n), ALINE(1:n), SIGMAC(1:n), SIGMAL(1:n), &
                SCONT(1:n), SLINE(1:n), RHOX(1:n), BNU(1:n)
      TAUNU = 0; SNU = 0; JNU = 0; HNU = 0; JMINS = 0; KNU = 0
      CALL JOSH(iscat, isurf)
      hsurf = HNU(1)
      WRITE(uout) hsurf, KNU(1), TAUNU(1:n), SNU(1:n), JNU(1:n), HNU(1:n), JMINS(1:n)
    END DO
  CASE ('coef')           ! -> COEFJ, COEFH as read by BLOCKJ/BLOCKH
    CALL GET_COMMAND_ARGUMENT(4, ddir)
    DATADIR = TRIM(ddir)
    CALL BLOCKJ
    CALL BLOCKH
    WRITE(uout) COEFJ, COEFH
  CASE DEFAULT
    WRITE(6,*) 'ref_numerics: unknown op ', TRIM(op)
    CALL EXIT(2)
  END SELECT
  CLOSE(uin)
  CLOSE(uout)
END PROGRAM ref_numerics
