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
/* The same with host arrays, synchronous; any ncell (the cells go through
 * the device in chunks of at most max_cells, staged in buffers the context
 * allocates at its first host-pointer call). */
int gckpp_rates(gckpp_ctx *ctx, int64_t ncell, const double *met, const double *photol,
                const double *het, int64_t het_stride, const double *conc, double *rconst);

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
/* gckpp_integrate_dev without any allocation (no reference counterpart:
 * for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops).  Same arguments and results (bitwise) as
 * gckpp_integrate_dev with the persistent solver's cell-per-wave variant;
 * any ncell.  Its scratch is a static device array of 28 MiB per device
 * (the work-stealing counter and up to 768 fullchem waves), reserved when
 * the library is loaded and held until the process ends; the call itself
 * makes no hipMalloc or other allocation.  Fails with a message when the
 * context's wave cap (gckpp_set_persistent_waves) asks for more waves than
 * the scratch holds.  All calls on one device share the scratch: calls
 * that overlap on one device (concurrent streams or threads) race.  The
 * cell order and step log of gckpp_set_cell_order / gckpp_set_step_log do
 * not apply. */
int gckpp_integrate_noalloc_dev(gckpp_ctx *ctx, int64_t ncell, double tin, double tout,
                                const int32_t icntrl[20], const double rcntrl[20],
                                const double *hstart, double *conc, const double *rconst,
                                int32_t *istatus, double *rstatus, int32_t *ierr,
                                void *stream);
/* The same with host arrays, synchronous; any ncell, in chunks of at most
 * max_cells through the staging buffers of gckpp_rates (allocated once);
 * istatus, rstatus and ierr may be null (not returned).  The cell order and
 * step log of gckpp_set_cell_order / gckpp_set_step_log do not apply. */
int gckpp_integrate(gckpp_ctx *ctx, int64_t ncell, double tin, double tout,
                    const int32_t icntrl[20], const double rcntrl[20],
                    const double *hstart, double *conc, const double *rconst,
                    int32_t *istatus, double *rstatus, int32_t *ierr);

/* One Do_FullChem chemistry step for every cell (GeosCore/fullchem_mod.F90:
 * 953-1350): Update_RCONST, Integrate(0,dt), one retry of failed cells with
 * Hstart=0 and the concentrations restored (:1138-1277), then C=max(C,0).
 * hnew[ncell] = RSTATUS(Nhnew), the next step's Hstart (KPPHvalue, :1294).
 * istatus[20][ncell], ierr[ncell] (IERR < 0: the cell failed twice).
 * Asynchronous like every _dev call: returns -1 only for a bad argument or
 * a launch failure; gckpp_step_failures gives the number of cells of the
 * last step that failed twice (the condition for which Do_FullChem returns
 * GC_FAILURE, :1255-1277). */
int gckpp_fullchem_step_dev(gckpp_ctx *ctx, int64_t ncell, double dt,
                            const int32_t icntrl[20], const double rcntrl[20],
                            const double *met, const double *photol,
                            const double *het, int64_t het_stride,
                            const double *hstart, double *conc, double *hnew,
                            int32_t *istatus, int32_t *ierr, void *stream);
/* Cells of the context's last gckpp_fullchem_step_dev that failed twice
 * (no reference counterpart: Do_FullChem stops at the first, :1255-1277).
 * Synchronises the device. */
int gckpp_step_failures(gckpp_ctx *ctx, int64_t *nfail);
/* The same with host arrays, synchronous -- the one call of the host's
 * Do_FullChem (the Fortran arrays C(ncell,NSPEC), MET(ncell,NMET), ...
 * passed as they are); any ncell, in chunks of at most max_cells.  hstart
 * may be null (cold start), hnew and istatus too (not returned).  Returns
 * -1 if any cell failed twice (after writing every output).  The cell
 * order and step log of gckpp_set_cell_order / gckpp_set_step_log do not
 * apply to this call. */
int gckpp_fullchem_step(gckpp_ctx *ctx, int64_t ncell, double dt,
                        const int32_t icntrl[20], const double rcntrl[20],
                        const double *met, const double *photol, const double *het,
                        int64_t het_stride, const double *hstart, double *conc,
                        double *hnew, int32_t *istatus, int32_t *ierr);

/* The model side of Do_FullChem on the device.  gckpp_set_species_map
 * (host arrays, once): map[NSPEC] the model species (0-based, of nmodel)
 * of each KPP species or -1 (State_Chm%Map_KppSpc, fullchem_mod.F90:
 * 817-821), mw_g[nmodel] molecular weights [g/mol], zero[NSPEC] (or null)
 * 1 for the prod/loss dummies that start each step at zero (:942-946),
 * alk_equiv 1 (fullchem, the default configuration: sulfur sea-salt
 * chemistry in KPP) to convert the alkalinity tracers SALAAL and SALCAL to
 * equivalents after the gather and back before the scatter
 * (fullchem_ConvertAlkToEquiv / fullchem_ConvertEquivToAlk,
 * KPP/fullchem/fullchem_SulfurChemFuncs.F90:67-105, called at
 * fullchem_mod.F90:2034-2038 and :1285-1287: C * (MW * 7e-5)).
 * gckpp_gather_dev: species[nmodel][ncell] in kg/kg dry air and
 * airden[ncell] [kg/m3] -> conc[NSPEC][ncell] in molec/cm3, the unit
 * conversion of Convert_Spc_Units (GeosUtil/unitconv_mod.F90:1287-1391,
 * Conc * AIRDEN * (AVO/MW_kg)/1e6) fused with the gather; unmapped KPP
 * species 0.  gckpp_scatter_dev: the mapped species back, MAX(C,0)
 * (:1329-1350) converted to kg/kg dry (:1407-1515, C * 1e6/(AVO/MW_kg) /
 * AIRDEN); the other model species untouched. */
int gckpp_set_species_map(gckpp_ctx *ctx, int32_t nmodel, const int32_t *map,
                          const double *mw_g, const int32_t *zero, int32_t alk_equiv);
int gckpp_gather_dev(gckpp_ctx *ctx, int64_t ncell, const double *species,
                     const double *airden, double *conc, void *stream);
int gckpp_scatter_dev(gckpp_ctx *ctx, int64_t ncell, const double *conc,
                      const double *airden, double *species, void *stream);

/* fullchem_SetStateHet (KPP/fullchem/fullchem_HetStateFuncs.F90:48-244, called
 * per cell at GeosCore/fullchem_mod.F90:877-888) for a batch: the HetState
 * record het[NHET][ncell] of the fullchem rate laws from the model's fields
 * hetin[NHIN][ncell] (State_Met/State_Chm values, layout HIN_* of
 * cell_fields.h) and the cell's KPP concentrations conc[NSPEC][ncell]
 * (molec/cm3, after the gather), with Input_Opt's nAeroType (0..14),
 * gamma_HO2, LPSCCHEM and TurnOffHetRates.  Includes Cld_Params
 * (rateLawUtilFuncs.F90:281-417), Get_Theta_Ice (:259-301) and Halide_Conc
 * (:316-616).  fullchem contexts only. */
int gckpp_set_state_het_dev(gckpp_ctx *ctx, int64_t ncell, const double *hetin,
                            const double *conc, int32_t naerotype, double gamma_ho2,
                            int32_t lpscchem, int32_t turn_off_het, double *het, void *stream);
/* fullchem_SulfurAqChem (KPP/fullchem/fullchem_SulfurChemFuncs.F90:120-305,
 * via Set_Sulfur_Chem_Rates, GeosCore/fullchem_mod.F90:1964-2074) for a
 * batch: K_MT(1:6) into the met record's rows MET_K_MT..+5 from hetin's
 * WetAeroArea and AeroRadi of aerosol types 11 and 12, the KPP vector conc
 * (alkalinity tracers in equivalents, as gckpp_gather_dev with alk_equiv
 * leaves them) and met's TEMP and NUMDEN.  Molecular weights from
 * gckpp_set_species_consts.  fullchem contexts only. */
int gckpp_sulfur_aq_dev(gckpp_ctx *ctx, int64_t ncell, const double *hetin, const double *conc,
                        double *met, void *stream);
/* The same two with host arrays (Fortran hetin(ncell,NHIN), conc(ncell,NSPEC),
 * het(ncell,NHET), met(ncell,NMET) as they are), synchronous, any ncell in
 * chunks of at most max_cells staged through the device; gckpp_sulfur_aq
 * writes only met's K_MT rows. */
int gckpp_set_state_het(gckpp_ctx *ctx, int64_t ncell, const double *hetin, const double *conc,
                        int32_t naerotype, double gamma_ho2, int32_t lpscchem,
                        int32_t turn_off_het, double *het);
int gckpp_sulfur_aq(gckpp_ctx *ctx, int64_t ncell, const double *hetin, const double *conc,
                    double *met);

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
/* Fun's reaction rates (the Aout argument of Fun, gckpp_Function.F90:51-2152):
 * aout[NREACT][ncell] = RCT(r) times the product of r's reactant
 * concentrations, conc[NSPEC][ncell]; any mechanism. */
int gckpp_aout_dev(gckpp_ctx *ctx, int64_t ncell, const double *conc,
                   const double *rconst, double *aout, void *stream);
/* The same with host arrays, synchronous; any ncell (chunks of at most
 * max_cells through the staging buffers of gckpp_rates). */
int gckpp_aout(gckpp_ctx *ctx, int64_t ncell, const double *conc, const double *rconst,
               double *aout);

/* KPP-Standalone chemical-state files, one cell per file
 * (GeosCore/kppsa_interface_mod.F90:569-843, KppSa_Write_Samples): a header
 * whose first line is its own line count, the cell's metadata and
 * integrator options, then "Name,Value,ATOL" for the NSPEC species, "R<n>,"
 * for the NREACT rate constants and "A<n>," for Fun's reaction rates, with
 * the reference's Fortran formats (es25.16e3, es10.2e2, f11.4, ...).  Host
 * code, no GPU.  The strings are NUL-terminated; location is the cell name
 * of the file ("Location:"), timestamp "YYYY/MM/DD hh:mm" as written by
 * TIMESTAMP_STRING. */
typedef struct {
  char location[256];
  char timestamp[64];
  double lon, lat;             /* degrees */
  int32_t level;               /* GEOS-Chem vertical level (1-based) */
  double pressure_hpa, temperature, airnumden, h2o_vmr, cloud_frac, suncos;
  double h_init, h_exit, dt;   /* KPPHvalue before Integrate, RSTATE(Nhexit), chem step [s] */
  int32_t nsteps;              /* ISTATUS(3) */
  int32_t icntrl[20];
  double rcntrl[20];
} gckpp_kppsa_meta;

/* Writes one cell's file.  names: NSPEC records of GCKPP_NAME_LEN bytes,
 * each ended by a NUL or padded with blanks (a Fortran CHARACTER(LEN=64)
 * array); an empty record, or a null names, writes "C<n>" as the reference
 * does for a KPP species with no species-database entry.  conc[NSPEC] (the
 * concentrations before Integrate), atol[NSPEC], rconst[NREACT],
 * aout[NREACT]. */
enum { GCKPP_NAME_LEN = 64 };
int gckpp_kppsa_write(int mech_id, const char *path, const gckpp_kppsa_meta *meta,
                      const char *names, const double *conc, const double *atol,
                      const double *rconst, const double *aout);
/* Reads one cell's file back (any of the outputs may be null; names, if
 * given, receives the NSPEC names as NUL-terminated records of
 * GCKPP_NAME_LEN bytes).  Fails with a message if the file's species or
 * reaction count is not the mechanism's, or a line does not parse. */
int gckpp_kppsa_read(int mech_id, const char *path, gckpp_kppsa_meta *meta, char *names,
                     double *conc, double *atol, double *rconst, double *aout);

/* The KPPHvalue restart variable: each cell's RSTATUS(Nhnew), the next
 * step's Hstart, kept across runs in the restart file's Chem_KPPHvalue
 * (registered at Headers/state_chm_mod.F90:1742-1753, read back with a
 * zero fallback at GeosCore/hco_utilities_gc_mod.F90:2144-2177).  The
 * writer makes a netCDF classic file (64-bit offset) with dimensions
 * time(1), lev(nz), lat(ny), lon(nx), the double variable
 * Chem_KPPHvalue(time,lev,lat,lon) [s] -- the Fortran array
 * KPPHvalue(nx,ny,nz) as it lies in memory -- and the global attribute
 * gckpp_step_model = model_tag (the step-size model the values came from,
 * "" for none).  Host code, no GPU. */
int gckpp_kpph_restart_write(const char *path, int64_t nx, int64_t ny, int64_t nz,
                             const double *kpph, const char *model_tag);
/* Reads Chem_KPPHvalue (double or float) from a netCDF classic file
 * (CDF-1 or CDF-2) into kpph[nx*ny*nz]; the variable must have that many
 * values.  Without the variable: *found = 0 and kpph = 0 (a cold start,
 * as the reference does), else *found = 1.  model_tag, if not null,
 * receives the file's gckpp_step_model attribute (GCKPP_NAME_LEN bytes,
 * NUL-terminated; empty when absent). */
int gckpp_kpph_restart_read(const char *path, int64_t nx, int64_t ny, int64_t nz,
                            double *kpph, int32_t *found, char *model_tag);

/* Device time [ms] of the parts of the last gckpp_fullchem_step_dev call on
 * this context, from HIP events on its stream: ms[0] Update_RCONST,
 * ms[1] Integrate, ms[2] retry of failed cells + clipping, ms[3] the whole
 * step.  The batch counterpart of the per-cell KppTime diagnostic
 * (GeosCore/fullchem_mod.F90:596-598,1298-1301). */
int gckpp_step_timings(gckpp_ctx *ctx, float ms[4]);

/* Persistent solver configuration (no reference counterpart; the OpenMP
 * schedule of the cell loop, GeosCore/fullchem_mod.F90:528-546, is the CPU
 * analogue).  variant in [0, gckpp_persistent_variants()).  Variants 0-6:
 * one lane per cell, state in HBM; (U, W, S) = terms per batched gather U,
 * minimum waves per SIMD W (the register budget), S = 1: the independent
 * sums (Fun, Jac, the U part of the LU, substitution in batches of
 * independent rows) as flat segment lists: (1,1,0) (8,1,0) (8,4,0) (8,4,1)
 * (16,4,0) (16,4,1) (8,5,0); waves_per_group 1, 2 or 4 (cells per
 * workgroup 64, 128, 256).  Variant 7: one wavefront per cell with the
 * cell's state in LDS (waves_per_group must be 1; U = W = 0, S = 2).
 * Variants 8, 9: one lane per cell with an operand cache of C = 16 / 32
 * values per lane in LDS for the LU and the substitutions, managed by a
 * schedule computed when the mechanism is compiled ((16,4,0), (16,2,0)).
 * All variants add the same terms in the same order.  get: out = {variant,
 * waves_per_group, waves per launch, U, W, S, C}. */
int gckpp_persistent_variants(void);
int gckpp_set_persistent_config(gckpp_ctx *ctx, int variant, int waves_per_group);
int gckpp_get_persistent_config(gckpp_ctx *ctx, int32_t out[7]);

/* Persistent solver extras (no reference counterpart).  order[ncell]
 * (device, a permutation of 0..ncell-1, or null): the order in which the
 * resident waves take cells, e.g. the costliest first.  hlog[2*n][ncell]
 * (device float, or null): the next persistent Integrate or step writes
 * (H, scaled error) of each cell's first n attempts to rows 2k, 2k+1 -- the
 * step histories the step-size MLP is trained on.  The pointers are used by
 * later calls until replaced. */
int gckpp_set_cell_order(gckpp_ctx *ctx, const int32_t *order);
int gckpp_set_step_log(gckpp_ctx *ctx, float *hlog, int nattempt);

/* Split persistent solve (no reference counterpart).  With nfirst > 0 and a
 * cell order set, the next persistent Integrate or step runs the first
 * nfirst cells of the order on the cell-per-wave kernel (variant 7) on a
 * second stream of the context, concurrently with the configured variant on
 * the rest of the order; the call's stream waits for both.  Meant for the
 * costliest cells of an uneven step (a warm step, cells ordered by
 * gckpp_order_by_h0_dev): one cell-per-wave attempt takes far less time than
 * a lane-per-cell attempt, so the longest chains of attempts stop setting
 * the step's length.  Results are bitwise those of one kernel.  0 turns it
 * off; ignored without a cell order, for carbon and when nfirst >= ncell.
 * The first call with nfirst > 0 allocates the cell-per-wave slots. */
int gckpp_set_persistent_split(gckpp_ctx *ctx, int64_t nfirst);

/* Cap on the waves of a persistent launch (no reference counterpart): at
 * most max_waves resident waves take cells (0: every wave that fits, the
 * default).  Fewer waves mean fewer concurrent slot blocks in HBM, which
 * can raise each wave's throughput; kept across gckpp_set_persistent_config
 * and the autotuner. */
int gckpp_set_persistent_waves(gckpp_ctx *ctx, int64_t max_waves);

/* Tail hand-off (no reference counterpart).  With max_cells > 0, a
 * persistent Integrate or step on a lane variant hands its last cells to the
 * cell-per-wave kernel: once at most max_cells cells of the lane launch are
 * unfinished, each of them stops at its next step boundary (before the
 * first attempt of an integration step) and saves its state -- T, H, the
 * ISTATUS counters, the last accepted step, Y -- and a cell-per-wave launch
 * that follows on the same stream continues them.  A lone cell's attempt
 * takes far less time on the cell-per-wave kernel than on a lane whose wave
 * is otherwise idle, so the last cells stop setting the step's length.
 * Results are bitwise those without the hand-off.  0 turns it off; not for
 * carbon.  A call with a larger max_cells than before allocates the list
 * ((15 + NVAR) doubles and one int per cell) and, the first time, the
 * cell-per-wave scratch. */
int gckpp_set_persistent_handoff(gckpp_ctx *ctx, int64_t max_cells);
/* Statistics of the context's last solve: n[0] cells handed over, n[1]
 * attempts the cell-per-wave launch made on them; ms[0] the lane launch's
 * duration, ms[1] the cell-per-wave launch's (HIP events on the call's
 * stream).  All zeros when the last solve ran no hand-off (plain solver,
 * hand-off off, carbon, the cell-per-wave variant) or before the first
 * solve.  Synchronises the device. */
int gckpp_persistent_handoff_stats(gckpp_ctx *ctx, int64_t n[2], float ms[2]);

/* Autotuner: integrates a copy of the given cells with every persistent
 * configuration, keeps the fastest for this context and writes the times
 * [ms] to ms[variant*3 + log2(waves_per_group)] (-1: does not fit).  conc is
 * not modified.  Synchronous on stream. */
int gckpp_autotune_dev(gckpp_ctx *ctx, int64_t ncell, double tin, double tout,
                       const int32_t icntrl[20], const double rcntrl[20],
                       const double *hstart, const double *conc,
                       const double *rconst, float *ms, void *stream);

/* Mixed-mechanism batch (no reference counterpart: GEOS-Chem runs one
 * mechanism per simulation, Do_FullChem / ChemMercury / Chem_Carbon_Gases,
 * GeosCore/fullchem_mod.F90:528, mercury_mod.F90:646, carbon_gases_mod.F90:450).
 * One chemistry step of the cells of up to three mechanisms is one
 * persistent launch taking fullchem cells first, then Hg, then carbon:
 * Update_RCONST of each part, Integrate (Rodas3; carbon: forward Euler), the
 * callers' retry with Hstart=0 from the original C (fullchem and Hg), the
 * clip C=max(C,0) (fullchem and Hg; carbon stores C as is).  A step can be
 * captured into a HIP graph and replayed: each replay advances conc, and
 * hstart (in: RCNTRL(3) per cell, out: RSTATUS(3)) carries the step size to
 * the next step. */
typedef struct gckpp_batch gckpp_batch;
typedef struct {
  int64_t ncell;              /* cells of this mechanism (0: none) */
  const double *met, *photol, *het;
  int64_t het_stride;         /* 0 or ncell */
  double *conc;               /* [NSPEC][ncell] in/out */
  double *hstart;             /* [ncell] in/out */
  int32_t *istatus;           /* [20][ncell] or null */
  int32_t *ierr;              /* [ncell] or null */
  int32_t icntrl[20];
  double rcntrl[20];
  const int32_t *order;       /* [ncell] order in which cells are taken, or null */
} gckpp_part;
/* contexts in the order fullchem, Hg, carbon; any may be null; each part's
 * ncell must not exceed its context's max_cells. */
int gckpp_batch_init(gckpp_ctx *fullchem, gckpp_ctx *hg, gckpp_ctx *carbon,
                     gckpp_batch **out);
int gckpp_batch_set_config(gckpp_batch *b, int variant, int waves_per_group);
/* Cap on the waves of the batch's persistent launch, as
 * gckpp_set_persistent_waves for a context (0: every resident wave); kept
 * across gckpp_batch_set_config.  A captured graph keeps the launch it was
 * captured with. */
int gckpp_batch_set_waves(gckpp_batch *b, int64_t max_waves);
/* The tail hand-off of gckpp_set_persistent_handoff for the batch's launch
 * (fullchem and Hg cells; carbon cells finish on their lanes), captured
 * with the step; gckpp_batch_handoff_count: cells the last step handed
 * over (synchronises the device).  A captured graph keeps the setting it
 * was captured with.  A max_cells larger than any before reallocates the
 * list: the call then waits for the device and destroys the captured
 * graph, and gckpp_batch_replay fails (-1, nothing captured) until the
 * next gckpp_batch_capture. */
int gckpp_batch_set_handoff(gckpp_batch *b, int64_t max_cells);
int gckpp_batch_handoff_count(gckpp_batch *b, int64_t *count);
int gckpp_batch_step_dev(gckpp_batch *b, const gckpp_part parts[3], double dt,
                         void *stream);
int gckpp_batch_capture(gckpp_batch *b, const gckpp_part parts[3], double dt);
int gckpp_batch_replay(gckpp_batch *b, void *stream);
int gckpp_batch_finalize(gckpp_batch *b);

/* Step-size MLP (no reference counterpart; it replaces Hstart -- the cold
 * start's 1e-5 s, gckpp_Integrator.F90:389-395, or the previous step's
 * KPPHvalue, GeosCore/fullchem_mod.F90:1016-1020 -- by a prediction, and the
 * integrator's step control still accepts or rejects every step).
 * 42 inputs: T, log10 P, log10 NUMDEN, SUNCOS, log10(J+1e-20) of the 12
 * PHOTOL slots jslot[] (0-based), log10(C+1) of the 24 species spc[]
 * (0-based KPP indices), ln(hprev) and 1 for cells with a previous step
 * size hprev > 0 (hprev[ncell] may be null: cold, 0 and 0); normalised by
 * mean/std (48 entries, the last 6 unused).  3 hidden layers of 128 GELU units; outputs (ln h0 - out_mu)/
 * out_sigma and 4 stiffness-class logits.  Parameters: GCKPP_MLP_NPARAM
 * floats, per layer W[in][out] row-major then bias, layer sizes 48x128,
 * 128x128, 128x128, 128x16 (outputs 5..15 unused).
 * predict: bf16 MFMA inference; h0[ncell] = exp(ln h0)*safety clipped to
 *   [hmin, hmax], cls[ncell] = argmax class (either may be null).
 * features: X[ncell][48] fp32 normalised inputs (for training).
 * forward: fp32 outputs out[nb][16] of X[nb][48].
 * grad / train_step: fp32 forward and backward of the loss
 *   wh*mean((o0 - yh)^2) + wc*mean(cross-entropy(o1..o4, ycls)) over nb
 *   rows; grad[GCKPP_MLP_NPARAM] (device) and loss[1] (device) may be null;
 *   train_step then applies one Adam step (b1 0.9, b2 0.999, eps 1e-8). */
#define GCKPP_MLP_NPARAM 41360
typedef struct gckpp_mlp gckpp_mlp;
int gckpp_mlp_init(int64_t max_batch, const int32_t jslot[12], const int32_t spc[24],
                   gckpp_mlp **out);
int gckpp_mlp_set_norm(gckpp_mlp *m, const float mean[48], const float stdev[48],
                       float out_mu, float out_sigma);
int gckpp_mlp_set_params(gckpp_mlp *m, const float *params);
int gckpp_mlp_get_params(gckpp_mlp *m, float *params);
int gckpp_mlp_features_dev(gckpp_mlp *m, int64_t ncell, const double *met,
                           const double *photol, const double *conc,
                           const double *hprev, float *X, void *stream);
int gckpp_mlp_predict_dev(gckpp_mlp *m, int64_t ncell, const double *met,
                          const double *photol, const double *conc,
                          const double *hprev, double safety, double hmin,
                          double hmax, double *h0, int32_t *cls, void *stream);
int gckpp_mlp_forward_dev(gckpp_mlp *m, int64_t nb, const float *X, float *out,
                          void *stream);
int gckpp_mlp_grad_dev(gckpp_mlp *m, int64_t nb, const float *X, const float *yh,
                       const int32_t *ycls, float wh, float wc, float *grad,
                       float *loss, void *stream);
int gckpp_mlp_train_step_dev(gckpp_mlp *m, int64_t nb, const float *X, const float *yh,
                             const int32_t *ycls, float wh, float wc, float lr,
                             float *loss, void *stream);
int gckpp_mlp_finalize(gckpp_mlp *m);
/* order[ncell] = the cells sorted by class, class 3 (costliest) first,
 * stable within a class: a queue order for gckpp_set_cell_order.
 * scratch: 4*ceil(ncell/256) int32 on the device. */
int gckpp_order_by_class_dev(int64_t ncell, const int32_t *cls, int32_t *order,
                             int32_t *scratch, void *stream);
/* Order of cells by predicted initial step size, smallest first (the
 * stiffest cells), stable: a counting sort over 32 quarter-decade bins of
 * log10(h0[ncell]) on [1e-5 s, 1200 s] (h0 <= 0: last).  scratch:
 * gckpp_order_by_h0_scratch(ncell) int32 (device). */
int gckpp_order_by_h0_dev(int64_t ncell, const double *h0, int32_t *order, int32_t *scratch,
                          void *stream);
int64_t gckpp_order_by_h0_scratch(int64_t ncell);

/* Message of the last failure on this thread ("" if none). */
const char *gckpp_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* GCKPP_HIP_H */
