! A Fortran host of the library: Do_FullChem's batch call
! (GeosCore/fullchem_mod.F90:528-1551 restructured into gather, one call,
! scatter; the set-up of Init_FullChem, :2899-2919) through the
! ISO_C_BINDING module include/gckpp_hip_mod.F90, on host arrays laid out
! as the gather loop would fill them: C(ncell,NSPEC), MET(ncell,NMET),
! PHOTOL(ncell,NPHOTOL), HET(ncell or 1,NHET), HSTART(ncell).
!
! Usage: fullchem_host <batch file> <result file>
! Batch file (stream, native byte order): ncell, max_cells, het columns
! (int64 each), solver, tail hand-off cells (0: none; int32 each), DT,
! ICNTRL(20) (int32), RCNTRL(20),
! ATOL(NVAR), RTOL(NVAR), MW(NSPEC), HENRY_K0(NSPEC), HENRY_CR(NSPEC), MET,
! PHOTOL, HET, HSTART, C (float64, column-major).
! Result file: the step's return code (int32), C, HNEW (float64), ISTATUS,
! IERR (int32).
PROGRAM fullchem_host
  USE ISO_C_BINDING
  USE gckpp_hip_mod
  IMPLICIT NONE
  INTEGER(C_INT32_T)              :: sz(8), icntrl(20), solver, handoff, rcode
  INTEGER(C_INT64_T)              :: ncell, max_cells, hcols, hstride
  INTEGER                         :: nspec, nvar, nmet, nhet, nphot, u, rc
  REAL(C_DOUBLE)                  :: dt, rcntrl(20)
  REAL(C_DOUBLE),     ALLOCATABLE :: atol(:), rtol(:), mw(:), hk0(:), hcr(:)
  REAL(C_DOUBLE),     ALLOCATABLE :: met(:,:), phot(:,:), het(:,:), hstart(:), c(:,:), hnew(:)
  INTEGER(C_INT32_T), ALLOCATABLE :: istatus(:,:), ierr(:)
  TYPE(C_PTR)                     :: ctx
  CHARACTER(LEN=4096)             :: fin, fout

  CALL GET_COMMAND_ARGUMENT(1, fin)
  CALL GET_COMMAND_ARGUMENT(2, fout)
  rc = gckpp_mech_sizes(GCKPP_FULLCHEM, sz)
  IF (rc /= 0) CALL die('gckpp_mech_sizes')
  nspec = sz(1); nvar = sz(2); nmet = sz(6); nhet = sz(7); nphot = sz(8)

  ! gathered batch (phase 1 of Do_FullChem, done by the caller)
  OPEN(NEWUNIT=u, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  READ(u) ncell, max_cells, hcols, solver, handoff, dt, icntrl, rcntrl
  ALLOCATE(atol(nvar), rtol(nvar), mw(nspec), hk0(nspec), hcr(nspec))
  ALLOCATE(met(ncell,nmet), phot(ncell,nphot), het(hcols,nhet), hstart(ncell), c(ncell,nspec))
  ALLOCATE(hnew(ncell), istatus(ncell,20), ierr(ncell))
  READ(u) atol, rtol, mw, hk0, hcr, met, phot, het, hstart, c
  CLOSE(u)

  ! Init_FullChem: context, tolerances, species constants, solver
  rc = gckpp_init(GCKPP_FULLCHEM, max_cells, ctx)
  IF (rc /= 0) CALL die('gckpp_init')
  rc = gckpp_set_tolerances(ctx, atol, rtol)
  IF (rc /= 0) CALL die('gckpp_set_tolerances')
  rc = gckpp_set_species_consts(ctx, mw, hk0, hcr)
  IF (rc /= 0) CALL die('gckpp_set_species_consts')
  rc = gckpp_set_solver(ctx, solver)
  IF (rc /= 0) CALL die('gckpp_set_solver')
  IF (handoff > 0) THEN
    rc = gckpp_set_persistent_handoff(ctx, INT(handoff, C_INT64_T))
    IF (rc /= 0) CALL die('gckpp_set_persistent_handoff')
  END IF

  ! phase 2: Update_RCONST, Integrate, retry, MAX(C,0) for every cell
  hstride = ncell
  IF (hcols == 1) hstride = 0
  rcode = gckpp_fullchem_step(ctx, ncell, dt, icntrl, rcntrl, met, phot, het, hstride, hstart, &
                              c, hnew, istatus, ierr)
  IF (rcode /= 0) CALL report('gckpp_fullchem_step')

  ! phase 3 (the scatter) is the caller's: hand the results back
  OPEN(NEWUNIT=u, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(u) rcode, c, hnew, istatus, ierr
  CLOSE(u)
  rc = gckpp_finalize(ctx)
  IF (rc /= 0) CALL die('gckpp_finalize')

CONTAINS

  SUBROUTINE report(what)
    CHARACTER(*), INTENT(IN) :: what
    CHARACTER(KIND=C_CHAR), POINTER :: msg(:)
    INTEGER :: k
    CALL C_F_POINTER(gckpp_last_error(), msg, [512])
    k = 0
    DO WHILE (k < 512)
      IF (msg(k + 1) == C_NULL_CHAR) EXIT
      k = k + 1
    END DO
    WRITE(*, '(A, ": ", 512A)') what, msg(1:k)
  END SUBROUTINE report

  SUBROUTINE die(what)
    CHARACTER(*), INTENT(IN) :: what
    CALL report(what)
    STOP 1
  END SUBROUTINE die

END PROGRAM fullchem_host
