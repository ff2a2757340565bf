! Test infrastructure only (a checker, never the product): a C-callable
! driver around the reference's own KPP fullchem core, which oracle/Makefile
! compiles where it lies under the reference checkout.  Only tests/,
! __graft_entry__.smoke() and bench.py's cpu_baseline leg load the library.
!
! Every per-cell array is cell-fastest, matching the C ABI of the GPU library:
!   met(ncell,NMET), photol(ncell,200), het(ncell|1,NHET), conc(ncell,NSPEC),
!   rct_io(ncell,NREACT), istatus(ncell,20), rstatus(ncell,20), ierr(ncell)
!
! The loops mirror GeosCore/fullchem_mod.F90:528-1551 (Do_FullChem):
! OpenMP PARALLEL DO SCHEDULE(DYNAMIC,24) over cells, THREADPRIVATE KPP
! globals, Update_RCONST once per cell (:953), Integrate(0,DT,...) (:1033),
! one retry with Hstart=0 and C restored (:1138-1277), C=max(C,0) (:1329-1350).
MODULE gckpp_ref_driver
  USE ISO_C_BINDING
  USE gckpp_Precision
  USE gckpp_Parameters
  USE gckpp_Global
  USE gckpp_Rates,         ONLY : Update_RCONST
  USE gckpp_Function,      ONLY : Fun, Fun_SPLIT
#ifndef MECH_CARBON
  USE gckpp_Jacobian,      ONLY : Jac_SP
  USE gckpp_LinearAlgebra, ONLY : KppDecomp, KppSolve
#endif
  USE gckpp_Integrator,    ONLY : Integrate
  IMPLICIT NONE
  PRIVATE
  INTEGER, PARAMETER :: NMET = 32, NHET = 165, NPHOT = 200
  ! physical constants of Headers/physconstants.F90:26-90
  REAL(dp), PARAMETER :: BOLTZ    = 1.38064852e-23_dp
  REAL(dp), PARAMETER :: CONSVAP  = 6.1078e+03_dp / ( BOLTZ * 1e+7_dp )
  REAL(dp), PARAMETER :: RGASLATM = 8.2057e-2_dp
  REAL(dp), PARAMETER :: RSTARG   = 8.3144598_dp
  REAL(dp), PARAMETER :: AVO_C    = 6.022140857e+23_dp
  REAL(dp), PARAMETER :: PI_C     = 3.14159265358979323_dp
  PUBLIC :: ref_sizes, ref_init, ref_rates, ref_fun, ref_jac, ref_decomp, ref_solve
  PUBLIC :: ref_integrate, ref_fullchem_step

CONTAINS

  SUBROUTINE ref_sizes( sz ) BIND(C, name="ref_sizes")
    INTEGER(C_INT), INTENT(OUT) :: sz(8)
    sz = 0
    sz(1) = NSPEC; sz(2) = NVAR; sz(3) = NFIX; sz(4) = NREACT
    sz(5) = LU_NONZERO; sz(6) = NMET; sz(7) = NHET; sz(8) = NPHOT
  END SUBROUTINE ref_sizes

  ! Shared (non-THREADPRIVATE) inputs: tolerances and species constants.
  SUBROUTINE ref_init( atol_in, rtol_in, mw_in, hk0_in, hcr_in ) BIND(C, name="ref_init")
    REAL(C_DOUBLE), INTENT(IN) :: atol_in(NVAR), rtol_in(NVAR)
    REAL(C_DOUBLE), INTENT(IN) :: mw_in(NSPEC), hk0_in(NSPEC), hcr_in(NSPEC)
    ATOL(1:NVAR) = atol_in
    RTOL(1:NVAR) = rtol_in
    MW       = mw_in
    SR_MW    = SQRT( mw_in )
    HENRY_K0 = hk0_in
    HENRY_CR = hcr_in
  END SUBROUTINE ref_init

  ! Sets the per-cell THREADPRIVATE globals of cell i, as
  ! Set_Kpp_GridBox_Values does (GeosCore/fullchem_mod.F90:2089-2167).
  SUBROUTINE set_cell( i, ncell, met, phot_io, het, het_stride, conc )
    INTEGER(C_INT64_T), INTENT(IN) :: i, ncell, het_stride
    REAL(C_DOUBLE),     INTENT(IN) :: met(ncell,NMET), phot_io(ncell,NPHOT)
    REAL(C_DOUBLE),     INTENT(IN) :: het(*), conc(ncell,NSPEC)
    REAL(dp) :: MREC(NMET), HREC(NHET), CONSEXP, VPRESH2O
    INTEGER  :: k
    INTEGER(C_INT64_T) :: hrow, j
    MREC = met(i,:)
    IF ( het_stride == 0 ) THEN
       hrow = 1
       j = 1
    ELSE
       hrow = het_stride
       j = i
    ENDIF
    DO k = 1, NHET
       HREC(k) = het( j + (k-1)*hrow )
    ENDDO
#include "cell_fields.inc"
    State_Het%AVO = AVO_C
    State_Het%PI  = PI_C
    PHOTOL(1:NPHOT) = phot_io(i,:)
    INV_TEMP        = 1.0_dp   / TEMP
    TEMP_OVER_K300  = TEMP     / 300.0_dp
    K300_OVER_TEMP  = 300.0_dp / TEMP
    SR_TEMP         = SQRT( TEMP )
    FOUR_R_T        = 4.0_dp * CON_R    * TEMP
    FOUR_RGASLATM_T = 4.0_dp * RGASLATM * TEMP
    EIGHT_RSTARG_T  = 8.0_dp * RSTARG   * TEMP
    CONSEXP         = 17.2693882_dp * (TEMP - 273.16_dp) / (TEMP - 35.86_dp)
    VPRESH2O        = CONSVAP * EXP( CONSEXP ) / TEMP
    RELHUM          = ( H2O / VPRESH2O ) * 100_dp
    C(1:NSPEC) = conc(i,:)
  END SUBROUTINE set_cell

  ! Update_RCONST for every cell (gckpp_Rates.F90:408-1503).
  SUBROUTINE ref_rates( ncell, met, phot_io, het, het_stride, conc, rct_io ) &
       BIND(C, name="ref_rates")
    INTEGER(C_INT64_T), VALUE      :: ncell, het_stride
    REAL(C_DOUBLE),     INTENT(IN) :: met(ncell,NMET), phot_io(ncell,NPHOT), het(*)
    REAL(C_DOUBLE),     INTENT(IN) :: conc(ncell,NSPEC)
    REAL(C_DOUBLE),     INTENT(OUT):: rct_io(ncell,NREACT)
    INTEGER(C_INT64_T) :: i
    !$OMP PARALLEL DO DEFAULT(SHARED) PRIVATE(i) SCHEDULE(DYNAMIC,24)
    DO i = 1, ncell
       CALL set_cell( i, ncell, met, phot_io, het, het_stride, conc )
       RCONST = 0.0_dp
       CALL Update_RCONST()
       rct_io(i,:) = RCONST(1:NREACT)
    ENDDO
    !$OMP END PARALLEL DO
  END SUBROUTINE ref_rates

  ! One Fun_SPLIT evaluation per cell (gckpp_Function.F90:2172-5060).
  SUBROUTINE ref_fun( ncell, conc, rct_io, vdot ) BIND(C, name="ref_fun")
    INTEGER(C_INT64_T), VALUE      :: ncell
    REAL(C_DOUBLE),     INTENT(IN) :: conc(ncell,NSPEC), rct_io(ncell,NREACT)
    REAL(C_DOUBLE),     INTENT(OUT):: vdot(ncell,NVAR)
    REAL(dp) :: V(NVAR), F(NFIX), RCT(NREACT), Vd(NVAR), P(NVAR), D(NVAR)
    INTEGER(C_INT64_T) :: i
#if !defined(MECH_HG) && !defined(MECH_CARBON)
    DO_FUN = .TRUE.
#endif
    DO i = 1, ncell
       V = conc(i,1:NVAR); F = conc(i,NVAR+1:NSPEC); RCT = rct_io(i,:)
#ifdef MECH_CARBON
       CALL Fun( V, F, RCT, Vd )
#else
       CALL Fun_SPLIT( V, F, RCT, Vd, P, D )
#endif
       vdot(i,:) = Vd
    ENDDO
  END SUBROUTINE ref_fun

  ! Jac_SP per cell (gckpp_Jacobian.F90:48-20887), LU_NONZERO layout.
  SUBROUTINE ref_jac( ncell, conc, rct_io, jvs ) BIND(C, name="ref_jac")
    INTEGER(C_INT64_T), VALUE      :: ncell
    REAL(C_DOUBLE),     INTENT(IN) :: conc(ncell,NSPEC), rct_io(ncell,NREACT)
    REAL(C_DOUBLE),     INTENT(OUT):: jvs(ncell,LU_NONZERO)
    REAL(dp) :: V(NVAR), F(NFIX), RCT(NREACT), J(LU_NONZERO)
    INTEGER(C_INT64_T) :: i
#if !defined(MECH_HG) && !defined(MECH_CARBON)
    DO_JVS = .TRUE.
#endif
#ifndef MECH_CARBON
    DO i = 1, ncell
       V = conc(i,1:NVAR); F = conc(i,NVAR+1:NSPEC); RCT = rct_io(i,:)
       J = 0.0_dp
       CALL Jac_SP( V, F, RCT, J )
       jvs(i,:) = J
    ENDDO
#endif
  END SUBROUTINE ref_jac

  ! KppDecomp in place (gckpp_LinearAlgebra.F90:46-83); ier(i) = IER.
  SUBROUTINE ref_decomp( ncell, jvs, ier ) BIND(C, name="ref_decomp")
    INTEGER(C_INT64_T), VALUE         :: ncell
    REAL(C_DOUBLE),     INTENT(INOUT) :: jvs(ncell,LU_NONZERO)
    INTEGER(C_INT),     INTENT(OUT)   :: ier(ncell)
    REAL(dp) :: J(LU_NONZERO)
    INTEGER  :: e
    INTEGER(C_INT64_T) :: i
#ifndef MECH_CARBON
    DO i = 1, ncell
       J = jvs(i,:)
       CALL KppDecomp( J, e )
       jvs(i,:) = J
       ier(i) = e
    ENDDO
#endif
  END SUBROUTINE ref_decomp

  ! KppSolve in place (gckpp_LinearAlgebra.F90:644-2309).
  SUBROUTINE ref_solve( ncell, lu, x ) BIND(C, name="ref_solve")
    INTEGER(C_INT64_T), VALUE         :: ncell
    REAL(C_DOUBLE),     INTENT(IN)    :: lu(ncell,LU_NONZERO)
    REAL(C_DOUBLE),     INTENT(INOUT) :: x(ncell,NVAR)
    REAL(dp) :: J(LU_NONZERO), B(NVAR)
    INTEGER(C_INT64_T) :: i
#if !defined(MECH_HG) && !defined(MECH_CARBON)
    DO_SLV = .TRUE.
#endif
#ifndef MECH_CARBON
    DO i = 1, ncell
       J = lu(i,:); B = x(i,:)
       CALL KppSolve( J, B )
       x(i,:) = B
    ENDDO
#endif
  END SUBROUTINE ref_solve

  ! Integrate(TIN,TOUT,...) per cell with rates given (ICNTRL(15)=-1 keeps
  ! them frozen); RCNTRL(3) is replaced by hstart(i).  No retry, no clipping.
  SUBROUTINE ref_integrate( ncell, conc, rct_io, hstart, tin, tout, icntrl_u, &
                            rcntrl_u, istatus, rstatus, ierr )               &
       BIND(C, name="ref_integrate")
    INTEGER(C_INT64_T), VALUE         :: ncell
    REAL(C_DOUBLE),     INTENT(INOUT) :: conc(ncell,NSPEC)
    REAL(C_DOUBLE),     INTENT(IN)    :: rct_io(ncell,NREACT), hstart(ncell)
    REAL(C_DOUBLE),     VALUE         :: tin, tout
    INTEGER(C_INT),     INTENT(IN)    :: icntrl_u(20)
    REAL(C_DOUBLE),     INTENT(IN)    :: rcntrl_u(20)
    INTEGER(C_INT),     INTENT(OUT)   :: istatus(ncell,20), ierr(ncell)
    REAL(C_DOUBLE),     INTENT(OUT)   :: rstatus(ncell,20)
    INTEGER  :: IC(20), IS(20), IE
    REAL(dp) :: RC(20), RS(20)
    INTEGER(C_INT64_T) :: i
    !$OMP PARALLEL DO DEFAULT(SHARED) PRIVATE(i, IC, IS, IE, RC, RS) SCHEDULE(DYNAMIC,24)
    DO i = 1, ncell
       C(1:NSPEC) = conc(i,:)
       RCONST(1:NREACT) = rct_io(i,:)
       IC = icntrl_u
       RC = rcntrl_u
       RC(3) = hstart(i)
       IS = 0; RS = 0.0_dp; IE = 0
       CALL Integrate( tin, tout, IC, RC, IS, RS, IE )
       conc(i,:) = C(1:NSPEC)
       istatus(i,:) = IS
       rstatus(i,:) = RS
       ierr(i) = IE
    ENDDO
    !$OMP END PARALLEL DO
  END SUBROUTINE ref_integrate

  ! One chemistry step of Do_FullChem for every cell: Update_RCONST,
  ! Integrate(0,dt), retry once on failure with Hstart=0 and C restored,
  ! clip negatives.  hnew(i) = RSTATE(Nhnew) as stored in KPPHvalue (:1294).
  SUBROUTINE ref_fullchem_step( ncell, met, phot_io, het, het_stride, conc,   &
                                hstart, dtchem, icntrl_u, rcntrl_u, istatus,     &
                                hnew, ierr, nthreads )                       &
       BIND(C, name="ref_fullchem_step")
    INTEGER(C_INT64_T), VALUE         :: ncell, het_stride
    REAL(C_DOUBLE),     INTENT(IN)    :: met(ncell,NMET), phot_io(ncell,NPHOT), het(*)
    REAL(C_DOUBLE),     INTENT(INOUT) :: conc(ncell,NSPEC)
    REAL(C_DOUBLE),     INTENT(IN)    :: hstart(ncell)
    REAL(C_DOUBLE),     VALUE         :: dtchem
    INTEGER(C_INT),     INTENT(IN)    :: icntrl_u(20)
    REAL(C_DOUBLE),     INTENT(IN)    :: rcntrl_u(20)
    INTEGER(C_INT),     INTENT(OUT)   :: istatus(ncell,20), ierr(ncell)
    REAL(C_DOUBLE),     INTENT(OUT)   :: hnew(ncell)
    INTEGER(C_INT),     VALUE         :: nthreads
    INTEGER  :: IC(20), IS(20), IE
    REAL(dp) :: RC(20), RS(20), C_before(NSPEC)
    INTEGER(C_INT64_T) :: i
    IF ( nthreads > 0 ) CALL omp_set_num_threads_wrap( nthreads )
    !$OMP PARALLEL DO DEFAULT(SHARED) PRIVATE(i, IC, IS, IE, RC, RS, C_before) &
    !$OMP SCHEDULE(DYNAMIC,24)
    DO i = 1, ncell
       CALL set_cell( i, ncell, met, phot_io, het, het_stride, conc )
       CALL Update_RCONST()
       IC = icntrl_u
       RC = rcntrl_u
       RC(3) = hstart(i)
       C_before = C(1:NSPEC)
       IS = 0; RS = 0.0_dp; IE = 0
       CALL Integrate( 0.0_dp, dtchem, IC, RC, IS, RS, IE )
       IF ( IE < 0 ) THEN
          RC(3) = 0.0_dp
          C(1:NSPEC) = C_before
          CALL Update_RCONST()
          CALL Integrate( 0.0_dp, dtchem, IC, RC, IS, RS, IE )
       ENDIF
       hnew(i) = RS(3)
#ifdef MECH_CARBON
       ! carbon_ConvertMolecCm3ToKg stores C as it is (carbon_Funcs.F90:300-322)
       conc(i,:) = C(1:NSPEC)
#else
       conc(i,:) = MAX( C(1:NSPEC), 0.0_dp )
#endif
       istatus(i,:) = IS
       ierr(i) = IE
    ENDDO
    !$OMP END PARALLEL DO
  END SUBROUTINE ref_fullchem_step

#if !defined(MECH_HG) && !defined(MECH_CARBON)
  ! Cld_Params per cell (rateLawUtilFuncs.F90:281-417), vAir [cm3] given:
  ! out(:,1..6) = rLiq, rIce, ALiq, VLiq, AIce, VIce.
  SUBROUTINE ref_cld_params( n, ad, cldf, frland, frocean, qi, ql, t, vair, out ) &
       BIND(C, name="ref_cld_params")
    USE rateLawUtilFuncs, ONLY : Cld_Params
    INTEGER(C_INT64_T), VALUE      :: n
    REAL(C_DOUBLE),     INTENT(IN) :: ad(n), cldf(n), frland(n), frocean(n), qi(n), ql(n)
    REAL(C_DOUBLE),     INTENT(IN) :: t(n), vair(n)
    REAL(C_DOUBLE),     INTENT(OUT):: out(n,6)
    TYPE(HetState) :: Hs
    INTEGER(C_INT64_T) :: i
    DO i = 1, n
       Hs%vAir = vair(i)
       CALL Cld_Params( ad(i), cldf(i), frland(i), frocean(i), qi(i), ql(i), t(i), Hs )
       out(i,:) = (/ Hs%rLiq, Hs%rIce, Hs%ALiq, Hs%VLiq, Hs%AIce, Hs%VIce /)
    ENDDO
  END SUBROUTINE ref_cld_params

  ! Ars_L1k and kIIR1Ltd per element (rateLawUtilFuncs.F90:77-150), with the
  ! globals NUMDEN and SR_TEMP of each element's cell:
  ! out(:,1) = Ars_L1k(area, radius, gamma, srMw), out(:,2) =
  ! kIIR1Ltd(cgas, ceduct, out(:,1)).
  SUBROUTINE ref_uptake( n, area, radius, gamma, srmw, numden_in, temp_in, cgas, ceduct, out ) &
       BIND(C, name="ref_uptake")
    USE rateLawUtilFuncs, ONLY : Ars_L1k, kIIR1Ltd
    INTEGER(C_INT64_T), VALUE      :: n
    REAL(C_DOUBLE),     INTENT(IN) :: area(n), radius(n), gamma(n), srmw(n), numden_in(n)
    REAL(C_DOUBLE),     INTENT(IN) :: temp_in(n), cgas(n), ceduct(n)
    REAL(C_DOUBLE),     INTENT(OUT):: out(n,2)
    INTEGER(C_INT64_T) :: i
    DO i = 1, n
       NUMDEN  = numden_in(i)
       TEMP    = temp_in(i)
       SR_TEMP = SQRT( TEMP )
       out(i,1) = Ars_L1k( area(i), radius(i), gamma(i), srmw(i) )
       out(i,2) = kIIR1Ltd( cgas(i), ceduct(i), out(i,1) )
    ENDDO
  END SUBROUTINE ref_uptake

  ! SafeDiv per element (rateLawUtilFuncs.F90:459-478).
  SUBROUTINE ref_safediv( n, num, den, alt, out ) BIND(C, name="ref_safediv")
    USE rateLawUtilFuncs, ONLY : SafeDiv
    INTEGER(C_INT64_T), VALUE      :: n
    REAL(C_DOUBLE),     INTENT(IN) :: num(n), den(n), alt(n)
    REAL(C_DOUBLE),     INTENT(OUT):: out(n)
    INTEGER(C_INT64_T) :: i
    DO i = 1, n
       out(i) = SafeDiv( num(i), den(i), alt(i) )
    ENDDO
  END SUBROUTINE ref_safediv

  ! The liquid/gas ratio of Compute_L2G_Local (fullchem_HetStateFuncs.F90:
  ! 631-693) from henry_mod's CALC_KH and CALC_HEFF (GeosUtil/henry_mod.F90).
  SUBROUTINE ref_l2g( n, k0, cr, pka, tk, h2oliq, ph, out ) BIND(C, name="ref_l2g")
    USE Henry_Mod, ONLY : CALC_KH, CALC_HEFF
    INTEGER(C_INT64_T), VALUE      :: n
    REAL(C_DOUBLE),     INTENT(IN) :: k0(n), cr(n), pka(n), tk(n), h2oliq(n), ph(n)
    REAL(C_DOUBLE),     INTENT(OUT):: out(n)
    REAL*8  :: KH, HEFF
    INTEGER :: RC
    INTEGER(C_INT64_T) :: i
    DO i = 1, n
       CALL CALC_KH( k0(i), cr(i), tk(i), KH, RC )
       CALL CALC_HEFF( pka(i), ph(i), KH, HEFF, RC )
       out(i) = HEFF * h2oliq(i)
    ENDDO
  END SUBROUTINE ref_l2g
#endif

  SUBROUTINE omp_set_num_threads_wrap( n )
    INTEGER(C_INT), INTENT(IN) :: n
!$  CALL omp_set_num_threads( INT(n) )
  END SUBROUTINE omp_set_num_threads_wrap

END MODULE gckpp_ref_driver
