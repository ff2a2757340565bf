! The host side of a batched chemistry step, for the Rosenbrock mechanisms
! (fullchem, Hg), on top of the C ABI module include/gckpp_hip_mod.F90.
!
! The reference runs its per-cell protocol inside the OpenMP cell loop:
!   fullchem  GeosCore/fullchem_mod.F90:953-1350 -- Update_RCONST, Integrate
!             with RCNTRL(3) = KPPHvalue, KppDiags from the first call's
!             ISTATUS(1:8) and its negative count (:1079-1133), on IERR < 0
!             one retry from the original C with RCNTRL(3) = 0 whose
!             ISTATUS(1:7) are added (:1138-1203), a cell that fails twice
!             is skipped (its species and KPPHvalue keep their values,
!             Failed2x, :1204-1245), KPPHvalue = RSTATE(Nhnew) (:1294),
!             C = MAX(C, 0) (:1329-1350);
!   Hg        GeosCore/mercury_mod.F90:1106-1262 -- the same calls, with
!             Hg_UpdateKppDiags (:4622-4719): every ISTATUS(1:8) and the
!             negative count of each call accumulated over both calls.
! Here the cells are gathered first (by the caller), then one rates call,
! one Integrate call for the batch, one Integrate call for the batch of
! failed cells, and the per-cell bookkeeping in loops over the batch.  The
! rate constants of the retry are the first call's: Update_RCONST of the
! restored C gives the same values.  The caller zeroes the prod/loss
! family species in C before the call, as the reference's gather does.
! The carbon mechanism (one forward-Euler step, no retry) calls
! gckpp_integrate directly.
MODULE gckpp_chem_mod
  USE ISO_C_BINDING
  USE gckpp_hip_mod
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: gckpp_kppdiags_init, gckpp_chem_step

  ! State_Diag's KppDiags fields for the batch's cells (REAL(4), as in
  ! Headers/state_diag_mod.F90): integrator calls, Jacobians, total,
  ! accepted and rejected steps, LU decompositions, substitutions,
  ! singular decompositions, negative concentrations after Integrate
  TYPE, PUBLIC :: gckpp_kppdiags
    REAL(4), ALLOCATABLE :: IntCounts(:), JacCounts(:), TotSteps(:), AccSteps(:)
    REAL(4), ALLOCATABLE :: RejSteps(:), LuDecomps(:), Substs(:), SmDecomps(:)
    REAL(4), ALLOCATABLE :: Negatives0(:)
  END TYPE gckpp_kppdiags

CONTAINS

  SUBROUTINE gckpp_kppdiags_init(d, ncell)
    TYPE(gckpp_kppdiags), INTENT(INOUT) :: d
    INTEGER(C_INT64_T),   INTENT(IN)    :: ncell
    IF (ALLOCATED(d%IntCounts)) THEN
      DEALLOCATE(d%IntCounts, d%JacCounts, d%TotSteps, d%AccSteps, d%RejSteps, d%LuDecomps, &
                 d%Substs, d%SmDecomps, d%Negatives0)
    END IF
    ALLOCATE(d%IntCounts(ncell), d%JacCounts(ncell), d%TotSteps(ncell), d%AccSteps(ncell), &
             d%RejSteps(ncell), d%LuDecomps(ncell), d%Substs(ncell), d%SmDecomps(ncell),   &
             d%Negatives0(ncell))
    d%IntCounts = 0.0; d%JacCounts = 0.0; d%TotSteps = 0.0; d%AccSteps = 0.0
    d%RejSteps = 0.0; d%LuDecomps = 0.0; d%Substs = 0.0; d%SmDecomps = 0.0
    d%Negatives0 = 0.0
  END SUBROUTINE gckpp_kppdiags_init

  ! ISTATUS(1:n) of cell i (istatus(ncell,20)) added to the diagnostics
  SUBROUTINE add_istatus(d, i, ist, n)
    TYPE(gckpp_kppdiags), INTENT(INOUT) :: d
    INTEGER(C_INT64_T),   INTENT(IN)    :: i
    INTEGER(C_INT32_T),   INTENT(IN)    :: ist(20)
    INTEGER,              INTENT(IN)    :: n
    d%IntCounts(i) = d%IntCounts(i) + REAL(ist(1), 4)
    d%JacCounts(i) = d%JacCounts(i) + REAL(ist(2), 4)
    d%TotSteps(i)  = d%TotSteps(i)  + REAL(ist(3), 4)
    d%AccSteps(i)  = d%AccSteps(i)  + REAL(ist(4), 4)
    d%RejSteps(i)  = d%RejSteps(i)  + REAL(ist(5), 4)
    d%LuDecomps(i) = d%LuDecomps(i) + REAL(ist(6), 4)
    d%Substs(i)    = d%Substs(i)    + REAL(ist(7), 4)
    IF (n >= 8) d%SmDecomps(i) = d%SmDecomps(i) + REAL(ist(8), 4)
  END SUBROUTINE add_istatus

  ! One chemistry step of a gathered batch of ncell cells of mechanism mech
  ! (GCKPP_FULLCHEM or GCKPP_HG) through context ctx.
  !   met(ncell,NMET), photol(ncell,NPHOTOL), het(hstride or 1,NHET) inputs
  !   kpph(ncell)   KPPHvalue in (RCNTRL(3) of each cell), the next out
  !   c(ncell,NSPEC) species in KPP units, in/out
  !   diags         KppDiags of the batch (allocated here)
  !   ierr(ncell)   the last Integrate's IERR of each cell
  !   nfail         cells that failed twice (their c and kpph are unchanged)
  !   rc            0, or -1 when a call failed or some cell failed twice
  !                 (GC_FAILURE, as Do_FullChem after the loop, :1524-1534)
  ! Optional, what the KPP-Standalone writer takes after the step
  ! (GeosCore/fullchem_mod.F90:1306-1321): c_init(ncell,NSPEC) the
  ! concentrations before Integrate, rconst_out(ncell,NREACT), and the last
  ! Integrate call's istatus_out(ncell,20) and rstatus_out(ncell,20).
  SUBROUTINE gckpp_chem_step(ctx, mech, ncell, nspec, nreact, dt, icntrl, rcntrl, met, photol, &
                             het, hstride, kpph, c, diags, ierr, nfail, rc, c_init, rconst_out,   &
                             istatus_out, rstatus_out)
    TYPE(C_PTR),          INTENT(IN)    :: ctx
    INTEGER(C_INT),       INTENT(IN)    :: mech
    INTEGER(C_INT64_T),   INTENT(IN)    :: ncell, hstride
    INTEGER,              INTENT(IN)    :: nspec, nreact
    REAL(C_DOUBLE),       INTENT(IN)    :: dt
    INTEGER(C_INT32_T),   INTENT(IN)    :: icntrl(20)
    REAL(C_DOUBLE),       INTENT(IN)    :: rcntrl(20)
    REAL(C_DOUBLE),       INTENT(IN)    :: met(*), photol(*), het(*)
    REAL(C_DOUBLE),       INTENT(INOUT) :: kpph(ncell), c(ncell, nspec)
    TYPE(gckpp_kppdiags), INTENT(INOUT) :: diags
    INTEGER(C_INT32_T),   INTENT(OUT)   :: ierr(ncell)
    INTEGER(C_INT64_T),   INTENT(OUT)   :: nfail
    INTEGER,              INTENT(OUT)   :: rc
    REAL(C_DOUBLE),     OPTIONAL, INTENT(OUT) :: c_init(ncell, nspec), rconst_out(ncell, nreact)
    REAL(C_DOUBLE),     OPTIONAL, INTENT(OUT) :: rstatus_out(ncell, 20)
    INTEGER(C_INT32_T), OPTIONAL, INTENT(OUT) :: istatus_out(ncell, 20)

    REAL(C_DOUBLE),     ALLOCATABLE :: rconst(:,:), cbefore(:,:), rstatus(:,:)
    REAL(C_DOUBLE),     ALLOCATABLE :: c2(:,:), r2(:,:), rst2(:,:), h2(:)
    INTEGER(C_INT32_T), ALLOCATABLE :: istatus(:,:), ist2(:,:), ierr2(:)
    INTEGER(C_INT64_T), ALLOCATABLE :: idx(:)
    INTEGER(C_INT64_T) :: i, k, nf
    REAL(C_DOUBLE)     :: rc2(20)
    LOGICAL            :: hg

    rc = 0
    nfail = 0
    hg = mech == GCKPP_HG
    CALL gckpp_kppdiags_init(diags, ncell)
    ALLOCATE(rconst(ncell, nreact), cbefore(ncell, nspec), rstatus(ncell, 20), istatus(ncell, 20))

    ! Update_RCONST of every cell
    IF (gckpp_rates(ctx, ncell, met, photol, het, hstride, c, rconst) /= 0) THEN
      rc = -1
      RETURN
    END IF

    ! C_before_integrate; Integrate(0, DT) with RCNTRL(3) = KPPHvalue
    cbefore = c
    IF (gckpp_integrate(ctx, ncell, 0.0_C_DOUBLE, dt, icntrl, rcntrl, kpph, c, rconst, &
                        istatus, rstatus, ierr) /= 0) THEN
      rc = -1
      RETURN
    END IF
    DO i = 1, ncell
      CALL add_istatus(diags, i, istatus(i, :), 8)
      diags%Negatives0(i) = REAL(COUNT(c(i, :) < 0.0_C_DOUBLE), 4)
    END DO

    ! the retry: the failed cells from their original C with Hstart = 0
    nf = COUNT(ierr < 0)
    IF (nf > 0) THEN
      ALLOCATE(idx(nf), c2(nf, nspec), r2(nf, nreact), rst2(nf, 20), h2(nf), ist2(nf, 20), ierr2(nf))
      k = 0
      DO i = 1, ncell
        IF (ierr(i) < 0) THEN
          k = k + 1
          idx(k) = i
        END IF
      END DO
      DO k = 1, nf
        c2(k, :) = cbefore(idx(k), :)
        r2(k, :) = rconst(idx(k), :)
      END DO
      h2 = 0.0_C_DOUBLE
      rc2 = rcntrl
      rc2(3) = 0.0_C_DOUBLE
      IF (gckpp_integrate(ctx, nf, 0.0_C_DOUBLE, dt, icntrl, rc2, h2, c2, r2, ist2, rst2, &
                          ierr2) /= 0) THEN
        rc = -1
        RETURN
      END IF
      DO k = 1, nf
        i = idx(k)
        c(i, :) = c2(k, :)
        rstatus(i, :) = rst2(k, :)
        istatus(i, :) = ist2(k, :)
        ierr(i) = ierr2(k)
        IF (hg) THEN
          CALL add_istatus(diags, i, ist2(k, :), 8)
          diags%Negatives0(i) = diags%Negatives0(i) + REAL(COUNT(c2(k, :) < 0.0_C_DOUBLE), 4)
        ELSE
          CALL add_istatus(diags, i, ist2(k, :), 7)
        END IF
      END DO
    END IF

    ! write-back: skipped cells keep their values; the others take Hnew
    ! and the clipped concentrations
    DO i = 1, ncell
      IF (ierr(i) < 0) THEN
        nfail = nfail + 1
        c(i, :) = cbefore(i, :)
      ELSE
        kpph(i) = rstatus(i, 3)
        c(i, :) = MAX(c(i, :), 0.0_C_DOUBLE)
      END IF
    END DO
    IF (nfail > 0) rc = -1
    IF (PRESENT(c_init)) c_init = cbefore
    IF (PRESENT(rconst_out)) rconst_out = rconst
    IF (PRESENT(istatus_out)) istatus_out = istatus
    IF (PRESENT(rstatus_out)) rstatus_out = rstatus
  END SUBROUTINE gckpp_chem_step

END MODULE gckpp_chem_mod
