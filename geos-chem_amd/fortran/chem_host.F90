! A Fortran host of the library for the Rosenbrock mechanisms: one
! chemistry step of a gathered grid through gckpp_chem_mod (Do_FullChem's
! or ChemMercury's protocol with the KppDiags statistics), Hstart from the
! restart file's KPPHvalue (GeosCore/hco_utilities_gc_mod.F90:2144-2177)
! and the next KPPHvalue written back as the restart variable, and the
! KPP-Standalone state file of one active cell
! (GeosCore/fullchem_mod.F90:1306-1321).
!
! Usage: chem_host <batch file> <restart in, or -> <restart out> <result file> [<state file>]
! Batch file (stream, native byte order): mech, solver, tail hand-off
! cells, active cell (1-based, 0: none) (int32 each), NX, NY, NZ,
! max_cells, het rows (int64 each), DT, ICNTRL(20) (int32), RCNTRL(20),
! ATOL(NVAR), RTOL(NVAR), MW(NSPEC), HENRY_K0(NSPEC), HENRY_CR(NSPEC),
! MET(ncell,NMET), PHOTOL(ncell,NPHOTOL), HET(rows,NHET), C(ncell,NSPEC)
! (float64, column-major; ncell = NX*NY*NZ, cells in (I,J,L) order), then
! the NSPEC species names (64 characters each).  With het rows = -1 the
! file holds HETIN(ncell,GCKPP_NHIN) -- the model's fields that
! fullchem_SetStateHet reads -- in place of HET, and the host fills HET and
! MET's K_MT rows on the GPU (gckpp_set_state_het, gckpp_sulfur_aq: the
! options nAeroType 14, gamma_HO2 0.2, LPSCCHEM on, het rates on) before
! the step.
! Result file: the step's return code (int32), the cells that failed twice
! (int64), C, KPPHvalue (float64), IERR (int32), the nine KppDiags fields
! (float32: IntCounts, JacCounts, TotSteps, AccSteps, RejSteps, LuDecomps,
! Substs, SmDecomps, Negatives0), whether the restart held KPPHvalue
! (int32) and its step-model tag (64 characters).
PROGRAM chem_host
  USE ISO_C_BINDING
  USE gckpp_hip_mod
  USE gckpp_chem_mod
  IMPLICIT NONE
  INTEGER(C_INT32_T)              :: sz(8), icntrl(20), mech, solver, handoff, active, found
  INTEGER(C_INT64_T)              :: nx, ny, nz, ncell, max_cells, hcols, hstride, nfail
  INTEGER                         :: nspec, nvar, nreact, nmet, nhet, nphot, u, rc, rcode, k
  REAL(C_DOUBLE)                  :: dt, rcntrl(20)
  REAL(C_DOUBLE),     ALLOCATABLE :: atol(:), rtol(:), mw(:), hk0(:), hcr(:), satol(:)
  REAL(C_DOUBLE),     ALLOCATABLE :: met(:,:), phot(:,:), het(:,:), kpph(:), kpph0(:), c(:,:)
  REAL(C_DOUBLE),     ALLOCATABLE :: cinit(:,:), rconst(:,:), rstatus(:,:), aout(:,:)
  REAL(C_DOUBLE),     ALLOCATABLE :: hetin(:,:)
  LOGICAL                         :: producers
  INTEGER(C_INT32_T), ALLOCATABLE :: ierr(:), istatus(:,:)
  CHARACTER(LEN=64),  ALLOCATABLE :: names(:)
  CHARACTER(KIND=C_CHAR)          :: tag(GCKPP_NAME_LEN)
  CHARACTER(LEN=64)               :: tagstr
  TYPE(gckpp_kppdiags)            :: diags
  TYPE(gckpp_kppsa_meta)          :: meta
  TYPE(C_PTR)                     :: ctx
  CHARACTER(LEN=4096)             :: fin, frin, frout, fout, fsa

  CALL GET_COMMAND_ARGUMENT(1, fin)
  CALL GET_COMMAND_ARGUMENT(2, frin)
  CALL GET_COMMAND_ARGUMENT(3, frout)
  CALL GET_COMMAND_ARGUMENT(4, fout)
  fsa = ''
  IF (COMMAND_ARGUMENT_COUNT() >= 5) CALL GET_COMMAND_ARGUMENT(5, fsa)

  OPEN(NEWUNIT=u, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  READ(u) mech, solver, handoff, active, nx, ny, nz, max_cells, hcols, dt, icntrl, rcntrl
  rc = gckpp_mech_sizes(mech, sz)
  IF (rc /= 0) CALL die('gckpp_mech_sizes')
  nspec = sz(1); nvar = sz(2); nreact = sz(4); nmet = sz(6); nhet = sz(7); nphot = sz(8)
  ncell = nx * ny * nz
  ALLOCATE(atol(nvar), rtol(nvar), mw(nspec), hk0(nspec), hcr(nspec), names(nspec))
  producers = hcols < 0
  IF (producers) THEN
    hcols = ncell
    ALLOCATE(hetin(ncell,GCKPP_NHIN))
  END IF
  ALLOCATE(met(ncell,nmet), phot(ncell,nphot), het(hcols,nhet), c(ncell,nspec), kpph(ncell))
  ALLOCATE(kpph0(ncell))
  ALLOCATE(ierr(ncell), cinit(ncell,nspec), rconst(ncell,nreact), istatus(ncell,20))
  ALLOCATE(rstatus(ncell,20))
  IF (producers) THEN
    READ(u) atol, rtol, mw, hk0, hcr, met, phot, hetin, c, names
  ELSE
    READ(u) atol, rtol, mw, hk0, hcr, met, phot, het, c, names
  END IF
  CLOSE(u)

  ! Init_FullChem / Init_Mercury: context, tolerances, species constants
  rc = gckpp_init(mech, max_cells, ctx)
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

  ! Set_Sulfur_Chem_Rates' sea-salt part and fullchem_SetStateHet on the GPU
  IF (producers) THEN
    rc = gckpp_sulfur_aq(ctx, ncell, hetin, c, met)
    IF (rc /= 0) CALL die('gckpp_sulfur_aq')
    rc = gckpp_set_state_het(ctx, ncell, hetin, c, 14_C_INT32_T, 0.2_C_DOUBLE, 1_C_INT32_T, &
                             0_C_INT32_T, het)
    IF (rc /= 0) CALL die('gckpp_set_state_het')
  END IF

  ! KPPHvalue from the restart file (zero without one, or without the variable)
  kpph = 0.0_C_DOUBLE
  found = 0
  tag = C_NULL_CHAR
  IF (TRIM(frin) /= '-') THEN
    rc = gckpp_kpph_restart_read(cstr(frin), nx, ny, nz, kpph, found, tag)
    IF (rc /= 0) CALL die('gckpp_kpph_restart_read')
  END IF
  kpph0 = kpph

  ! the step
  hstride = ncell
  IF (hcols == 1) hstride = 0
  CALL gckpp_chem_step(ctx, mech, ncell, nspec, nreact, dt, icntrl, rcntrl, met, phot, het, &
                       hstride, kpph, c, diags, ierr, nfail, rcode, c_init=cinit,           &
                       rconst_out=rconst, istatus_out=istatus, rstatus_out=rstatus)
  IF (rcode /= 0) CALL report('gckpp_chem_step')

  ! the next step's Hstart as the restart variable
  rc = gckpp_kpph_restart_write(cstr(frout), nx, ny, nz, kpph, cstr('chem_host'))
  IF (rc /= 0) CALL die('gckpp_kpph_restart_write')

  ! the active cell's KPP-Standalone file: its state before Integrate, the
  ! Hstart it was given, the last call's Hexit and step count
  IF (active > 0 .AND. LEN_TRIM(fsa) > 0) THEN
    IF (ierr(active) >= 0) THEN
      ALLOCATE(aout(1,nreact), satol(nspec))
      rc = gckpp_aout(ctx, 1_C_INT64_T, cinit(active,:), rconst(active,:), aout)
      IF (rc /= 0) CALL die('gckpp_aout')
      satol = atol(1)
      satol(1:nvar) = atol
      CALL fill(meta%location, 'ActiveCell')
      CALL fill(meta%timestamp, '2019/06/21 12:00')
      meta%lon = 0.0_C_DOUBLE; meta%lat = 0.0_C_DOUBLE
      meta%level = INT(1 + (active - 1) / (nx * ny), C_INT32_T)
      meta%temperature = met(active, 1); meta%pressure_hpa = met(active, 2)
      meta%airnumden = met(active, 3); meta%h2o_vmr = met(active, 4) / met(active, 3)
      meta%suncos = met(active, 5); meta%cloud_frac = 0.0_C_DOUBLE
      meta%h_init = kpph0(active)
      meta%h_exit = rstatus(active, 2)
      meta%dt = dt
      meta%nsteps = istatus(active, 3)
      meta%icntrl = icntrl
      meta%rcntrl = rcntrl
      meta%rcntrl(3) = kpph0(active)
      rc = gckpp_kppsa_write(mech, cstr(fsa), meta, names_c(), cinit(active,:), satol, &
                             rconst(active,:), aout(1,:))
      IF (rc /= 0) CALL die('gckpp_kppsa_write')
    END IF
  END IF

  tagstr = ''
  DO k = 1, GCKPP_NAME_LEN
    IF (tag(k) == C_NULL_CHAR) EXIT
    tagstr(k:k) = tag(k)
  END DO
  OPEN(NEWUNIT=u, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(u) INT(rcode, C_INT32_T), nfail, c, kpph, ierr
  WRITE(u) diags%IntCounts, diags%JacCounts, diags%TotSteps, diags%AccSteps, diags%RejSteps, &
           diags%LuDecomps, diags%Substs, diags%SmDecomps, diags%Negatives0
  WRITE(u) found, tagstr
  CLOSE(u)
  rc = gckpp_finalize(ctx)
  IF (rc /= 0) CALL die('gckpp_finalize')

CONTAINS

  ! NUL-terminated copy of a Fortran string
  FUNCTION cstr(s) RESULT(r)
    CHARACTER(*), INTENT(IN) :: s
    CHARACTER(KIND=C_CHAR)   :: r(LEN_TRIM(s) + 1)
    INTEGER :: i
    DO i = 1, LEN_TRIM(s)
      r(i) = s(i:i)
    END DO
    r(LEN_TRIM(s) + 1) = C_NULL_CHAR
  END FUNCTION cstr

  ! the species names as NSPEC blank-padded records of GCKPP_NAME_LEN
  FUNCTION names_c() RESULT(r)
    CHARACTER(KIND=C_CHAR) :: r(GCKPP_NAME_LEN * nspec)
    INTEGER :: n, i
    DO n = 1, nspec
      DO i = 1, GCKPP_NAME_LEN
        r((n - 1) * GCKPP_NAME_LEN + i) = names(n)(i:i)
      END DO
    END DO
  END FUNCTION names_c

  SUBROUTINE fill(dst, s)
    CHARACTER(KIND=C_CHAR), INTENT(OUT) :: dst(:)
    CHARACTER(*),           INTENT(IN)  :: s
    INTEGER :: i
    dst = C_NULL_CHAR
    DO i = 1, MIN(LEN(s), SIZE(dst) - 1)
      dst(i) = s(i:i)
    END DO
  END SUBROUTINE fill

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

END PROGRAM chem_host
