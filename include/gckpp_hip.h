/* gckpp_hip: MI355X-native KPP chemistry solve for GEOS-Chem (C ABI).
 *
 * Replaces, for a batch of grid cells at once, the per-cell calls that
 * GeosCore/fullchem_mod.F90:953,1033 (and mercury_mod.F90:1106,1132,
 * carbon_gases_mod.F90:532,539) make into the KPP-generated solver:
 *   Update_RCONST()                      KPP/fullchem/gckpp_Rates.F90:408-1503
 *   Integrate(TIN,TOUT,ICNTRL_U,RCNTRL_U,ISTATUS_U,RSTATUS_U,IERR_U)
 *                                        KPP/fullchem/gckpp_Integrator.F90:80-162
 * and the pieces Integrate calls (Fun_SPLIT, Jac_SP, KppDecomp, KppSolve).
 *
 * Layout: every per-cell array is field-major with cells fastest, i.e. the
 * Fortran array A(ncell, NFIELD) of the caller, element (cell, field) at
 * A[field*ncell + cell].  KPP species order (ind_* of gckpp_Parameters.F90).
 *
 * Entry points whose names end in _dev take device pointers and a HIP
 * stream and return once the work is queued; the others take host pointers
 * and are synchronous, like the Fortran calls they replace.
 *
 * Return value: 0 (GC_SUCCESS) or -1 (GC_FAILURE), Headers/errcode_mod.F90:29-30;
 * gckpp_last_error() describes the failure.  Per-cell integrator status codes
 * are KPP's IERR values (1 success, -1..-8 as in gckpp_Integrator.F90:538-575).
 *
 * Threading: one context per host thread; contexts are independent.  Not to
 * be called from inside an OpenMP parallel region.
 */
#ifndef GCKPP_HIP_H
#define GCKPP_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gckpp_ctx gckpp_ctx;

enum { GCKPP_FULLCHEM = 0, GCKPP_HG = 1, GCKPP_CARBON = 2 };
enum { GCKPP_SOLVER_PLAIN = 0, GCKPP_SOLVER_PERSISTENT = 1 };

/* Sizes of one mechanism: out[0..7] = NSPEC, NVAR, NFIX, NREACT, LU_NONZERO,
 * NMET, NHET, NPHOTOL (gckpp_Parameters.F90:33-58; per-cell record sizes of
 * cell_fields.h).  No GPU needed. */
int gckpp_mech_sizes(int mech_id, int32_t out[8]);

/* Creates a context for one mechanism on the current HIP device, with device
 * workspace for up to max_cells cells per launch (larger batches are solved
 * in chunks).  Replaces the module-level KPP state and Init_FullChem's
 * tolerance setup (GeosCore/fullchem_mod.F90:2899-2905); tolerances default
 * to ATOL=1e-2, RTOL=0.5e-2 until gckpp_set_tolerances. */
int gckpp_init(int mech_id, int64_t max_cells, gckpp_ctx **out);
int gckpp_finalize(gckpp_ctx *ctx);

/* How Integrate schedules cells (results are the same to roundoff):
 * GCKPP_SOLVER_PLAIN       one lane per cell, one launch per workspace chunk
 * GCKPP_SOLVER_PERSISTENT  resident waves that take the next cell from a
 *                          global counter as soon as a lane's cell is done
 *                          (no reference counterpart: the reference's OpenMP
 *                          DYNAMIC,24 schedule plays this role on the CPU,
 *                          GeosCore/fullchem_mod.F90:528-546) */
int gckpp_set_solver(gckpp_ctx *ctx, int solver);

/* ATOL(1:NVAR), RTOL(1:NVAR) (gckpp_Global.F90:78-80), host arrays. */
int gckpp_set_tolerances(gckpp_ctx *ctx, const double *atol, const double *rtol);

/* MW(NSPEC) [g/mol], HENRY_K0(NSPEC), HENRY_CR(NSPEC) used by the rate laws
 * (commonIncludeVars.H:7-12; set by Init_FullChem, fullchem_mod.F90:2911-2919). */
int gckpp_set_species_consts(gckpp_ctx *ctx, const double *mw,
                             const double *henry_k0, const double *henry_cr);

/* Update_RCONST for ncell cells (gckpp_Rates.F90:408-1503).
 * met[NMET][ncell], photol[NPHOTOL][ncell], het[NHET][het_stride] (het_stride
 * 0: one record shared by all cells, else ncell), conc[NSPEC][ncell] ->
 * rconst[NREACT][ncell]. */
int gckpp_rates_dev(gckpp_ctx *ctx, int64_t ncell, const double *met,
                    const double *photol, const double *het, int64_t het_stride,
                    const double *conc, double *rconst, void *stream);

/* Integrate(TIN,TOUT,ICNTRL_U,RCNTRL_U,ISTATUS_U,RSTATUS_U,IERR_U) for ncell
 * cells with frozen rates (ICNTRL(15)=-1).  icntrl/rcntrl: the 20 options
 * shared by all cells; hstart[ncell] replaces RCNTRL(3) per cell (0: start
 * at 1e-5 s); conc[NSPEC][ncell] in/out; rconst[NREACT][ncell];
 * istatus[20][ncell], rstatus[20][ncell], ierr[ncell] out.  Method: Rodas3
 * (ICNTRL(3) = 0 or 4). */
int gckpp_integrate_dev(gckpp_ctx *ctx, int64_t ncell, double tin, double tout,
                        const int32_t icntrl[20], const double rcntrl[20],
                        const double *hstart, double *conc, const double *rconst,
                        int32_t *istatus, double *rstatus, int32_t *ierr,
                        void *stream);
int gckpp_integrate(gckpp_ctx *ctx, int64_t ncell, double tin, double tout,
                    const int32_t icntrl[20], const double rcntrl[20],
                    const double *hstart, double *conc, const double *rconst,
                    int32_t *istatus, double *rstatus, int32_t *ierr);

/* One Do_FullChem chemistry step for every cell (GeosCore/fullchem_mod.F90:
 * 953-1350): Update_RCONST, Integrate(0,dt), one retry of failed cells with
 * Hstart=0 and the concentrations restored (:1138-1277), then C=max(C,0).
 * hnew[ncell] = RSTATUS(Nhnew), the next step's Hstart (KPPHvalue, :1294).
 * istatus[20][ncell], ierr[ncell]; returns -1 if any cell failed twice. */
int gckpp_fullchem_step_dev(gckpp_ctx *ctx, int64_t ncell, double dt,
                            const int32_t icntrl[20], const double rcntrl[20],
                            const double *met, const double *photol,
                            const double *het, int64_t het_stride,
                            const double *hstart, double *conc, double *hnew,
                            int32_t *istatus, int32_t *ierr, void *stream);

/* Components, for parity checks against the reference routines:
 * Fun_SPLIT (gckpp_Function.F90:2172-5060): vdot[NVAR][ncell]
 * Jac_SP    (gckpp_Jacobian.F90:48-20887):  jvs[LU_NONZERO][ncell]
 * KppDecomp (gckpp_LinearAlgebra.F90:46-83): jvs in place, ier[ncell]
 * KppSolve  (gckpp_LinearAlgebra.F90:644-2309): x[NVAR][ncell] in place */
int gckpp_fun_dev(gckpp_ctx *ctx, int64_t ncell, const double *conc,
                  const double *rconst, double *vdot, void *stream);
int gckpp_jac_dev(gckpp_ctx *ctx, int64_t ncell, const double *conc,
                  const double *rconst, double *jvs, void *stream);
int gckpp_decomp_dev(gckpp_ctx *ctx, int64_t ncell, double *jvs, int32_t *ier,
                     void *stream);
int gckpp_solve_dev(gckpp_ctx *ctx, int64_t ncell, const double *lu, double *x,
                    void *stream);

/* Device time [ms] of the parts of the last gckpp_fullchem_step_dev call on
 * this context, from HIP events on its stream: ms[0] Update_RCONST,
 * ms[1] Integrate, ms[2] retry of failed cells + clipping, ms[3] the whole
 * step.  The batch counterpart of the per-cell KppTime diagnostic
 * (GeosCore/fullchem_mod.F90:596-598,1298-1301). */
int gckpp_step_timings(gckpp_ctx *ctx, float ms[4]);

/* Message of the last failure on this thread ("" if none). */
const char *gckpp_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* GCKPP_HIP_H */
