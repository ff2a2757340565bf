/* Plain-C restatement of the reference's per-cell KPP Rosenbrock solve.
 * TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it as a checker; the product never links it.
 *
 * Table-driven: the caller passes the mechanism tables (from
 * geos-chem_amd/mech/<mech>.json).  One cell at a time, arrays per cell
 * (species index fastest), as in the reference.
 */
#pragma once
#include <stdint.h>

typedef struct {
  int nvar, nfix, nreact, lu_nonzero;
  /* reactants: reaction r uses rx_spc[rx_ptr[r] .. rx_ptr[r+1]) with
   * multiplicity rx_mult[..]; species index >= nvar means fixed species */
  const int *rx_ptr, *rx_spc, *rx_mult;
  /* net stoichiometry: reaction r changes st_spc[st_ptr[r]..] by st_coef[..] */
  const int *st_ptr, *st_spc;
  const double *st_coef;
  /* sparse LU pattern (0-based KPP LU_IROW/LU_ICOL/LU_CROW/LU_DIAG) */
  const int *lu_irow, *lu_icol, *lu_crow, *lu_diag;
  /* dense map (row*nvar+col) -> LU index or -1 */
  const int *lu_map;
} orc_mech;

/* Fun_SPLIT: Vdot = P - D*V (KPP/fullchem/gckpp_Function.F90:2172-5060) */
void orc_fun(const orc_mech *m, const double *V, const double *F,
             const double *RCT, double *Vdot);
/* Jac_SP in LU_NONZERO layout (KPP/fullchem/gckpp_Jacobian.F90:48-20887) */
void orc_jac(const orc_mech *m, const double *V, const double *F,
             const double *RCT, double *JVS);
/* KppDecomp (KPP/fullchem/gckpp_LinearAlgebra.F90:46-83); returns IER */
int orc_decomp(const orc_mech *m, double *JVS);
/* KppSolve (KPP/fullchem/gckpp_LinearAlgebra.F90:644-2309) */
void orc_solve(const orc_mech *m, const double *JVS, double *X);

/* Integrate -> Rosenbrock -> ros_Integrator with Rodas3, for ncell cells.
 * conc[ncell][nspec] in/out (per cell, species fastest), rct[ncell][nreact],
 * hstart[ncell]; istatus[ncell][20], rstatus[ncell][20], ierr[ncell].
 * (KPP/fullchem/gckpp_Integrator.F90:80-162,165-532,578-786) */
void orc_integrate(const orc_mech *m, int64_t ncell, double *conc,
                   const double *rct, const double *hstart, double tin,
                   double tout, const double *atol, const double *rtol,
                   const int *icntrl, const double *rcntrl, int *istatus,
                   double *rstatus, int *ierr, int nthreads);

/* The carbon mechanism's Integrate: Y += Fun(Y)*(tout-tin), ICNTRL(16)
 * negative handling (KPP/carbon/gckpp_Integrator.F90:155-215).  conc
 * [ncell][nspec] in/out, rct [ncell][nreact], ierr [ncell]. */
void orc_euler(const orc_mech *m, int64_t ncell, double *conc, const double *rct,
               double tin, double tout, const int *icntrl, int *ierr);
