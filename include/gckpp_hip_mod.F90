! Fortran ISO_C_BINDING interface to libgckpp_hip.so (include/gckpp_hip.h):
! the module a GEOS-Chem host binds to call the GPU chemistry solve in place
! of Update_RCONST / Integrate (KPP/fullchem/gckpp_Rates.F90:408,
! KPP/fullchem/gckpp_Integrator.F90:80).  Array arguments are the ABI's
! [field][ncell] layout, i.e. Fortran A(ncell, NFIELD).
MODULE gckpp_hip_mod
  USE ISO_C_BINDING
  IMPLICIT NONE
  PRIVATE
  INTEGER(C_INT), PARAMETER, PUBLIC :: GCKPP_FULLCHEM = 0, GCKPP_HG = 1, GCKPP_CARBON = 2
  INTEGER(C_INT), PARAMETER, PUBLIC :: GCKPP_SOLVER_PLAIN = 0, GCKPP_SOLVER_PERSISTENT = 1
  PUBLIC :: gckpp_init, gckpp_finalize, gckpp_set_tolerances, gckpp_set_species_consts
  PUBLIC :: gckpp_set_solver, gckpp_fullchem_step_dev, gckpp_integrate, gckpp_last_error
  PUBLIC :: gckpp_mech_sizes, gckpp_rates, gckpp_fullchem_step, gckpp_set_persistent_handoff
  PUBLIC :: gckpp_step_failures, gckpp_aout, gckpp_kppsa_write, gckpp_kppsa_read
  PUBLIC :: gckpp_kpph_restart_write, gckpp_kpph_restart_read
  PUBLIC :: gckpp_set_species_map, gckpp_gather_dev, gckpp_scatter_dev
  PUBLIC :: gckpp_set_state_het_dev, gckpp_sulfur_aq_dev, gckpp_set_state_het, gckpp_sulfur_aq
  INTEGER(C_INT), PARAMETER, PUBLIC :: GCKPP_NAME_LEN = 64
  ! fields of the model-field record hetin (HIN_* of geos-chem_amd/csrc/cell_fields.h)
  INTEGER(C_INT), PARAMETER, PUBLIC :: GCKPP_NHIN = 99

  ! gckpp_kppsa_meta: metadata and integrator fields of a KPP-Standalone
  ! state file (GeosCore/kppsa_interface_mod.F90:569-843); strings are
  ! NUL-terminated
  TYPE, BIND(C), PUBLIC :: gckpp_kppsa_meta
    CHARACTER(KIND=C_CHAR) :: location(256)
    CHARACTER(KIND=C_CHAR) :: timestamp(64)
    REAL(C_DOUBLE)         :: lon, lat
    INTEGER(C_INT32_T)     :: level
    REAL(C_DOUBLE)         :: pressure_hpa, temperature, airnumden, h2o_vmr, cloud_frac, suncos
    REAL(C_DOUBLE)         :: h_init, h_exit, dt
    INTEGER(C_INT32_T)     :: nsteps
    INTEGER(C_INT32_T)     :: icntrl(20)
    REAL(C_DOUBLE)         :: rcntrl(20)
  END TYPE gckpp_kppsa_meta

  INTERFACE
    ! out = NSPEC, NVAR, NFIX, NREACT, LU_NONZERO, NMET, NHET, NPHOTOL (no GPU needed)
    INTEGER(C_INT) FUNCTION gckpp_mech_sizes(mech_id, out) BIND(C, name="gckpp_mech_sizes")
      IMPORT :: C_INT, C_INT32_T
      INTEGER(C_INT),     VALUE       :: mech_id
      INTEGER(C_INT32_T), INTENT(OUT) :: out(8)
    END FUNCTION
    ! int gckpp_init(int mech_id, int64_t max_cells, gckpp_ctx **out)
    INTEGER(C_INT) FUNCTION gckpp_init(mech_id, max_cells, ctx) BIND(C, name="gckpp_init")
      IMPORT :: C_INT, C_INT64_T, C_PTR
      INTEGER(C_INT),     VALUE :: mech_id
      INTEGER(C_INT64_T), VALUE :: max_cells
      TYPE(C_PTR)               :: ctx            ! out
    END FUNCTION
    INTEGER(C_INT) FUNCTION gckpp_finalize(ctx) BIND(C, name="gckpp_finalize")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    ! ATOL(1:NVAR), RTOL(1:NVAR)
    INTEGER(C_INT) FUNCTION gckpp_set_tolerances(ctx, atol, rtol) BIND(C, name="gckpp_set_tolerances")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE         :: ctx
      REAL(C_DOUBLE), INTENT(IN) :: atol(*), rtol(*)
    END FUNCTION
    ! MW, HENRY_K0, HENRY_CR (1:NSPEC)
    INTEGER(C_INT) FUNCTION gckpp_set_species_consts(ctx, mw, hk0, hcr) &
         BIND(C, name="gckpp_set_species_consts")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE         :: ctx
      REAL(C_DOUBLE), INTENT(IN) :: mw(*), hk0(*), hcr(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION gckpp_set_solver(ctx, solver) BIND(C, name="gckpp_set_solver")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR),    VALUE :: ctx
      INTEGER(C_INT), VALUE :: solver
    END FUNCTION
    ! Persistent solver: the last max_cells unfinished cells of a step move
    ! to the cell-per-wave kernel at a step boundary (0: off)
    INTEGER(C_INT) FUNCTION gckpp_set_persistent_handoff(ctx, max_cells) &
         BIND(C, name="gckpp_set_persistent_handoff")
      IMPORT :: C_INT, C_INT64_T, C_PTR
      TYPE(C_PTR),        VALUE :: ctx
      INTEGER(C_INT64_T), VALUE :: max_cells
    END FUNCTION
    ! One chemistry step of ncell cells (device pointers, HIP stream);
    ! returns once queued (gckpp_step_failures: cells that failed twice)
    INTEGER(C_INT) FUNCTION gckpp_fullchem_step_dev(ctx, ncell, dt, icntrl, rcntrl, met, photol, &
         het, het_stride, hstart, conc, hnew, istatus, ierr, stream)                            &
         BIND(C, name="gckpp_fullchem_step_dev")
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_PTR, C_DOUBLE
      TYPE(C_PTR),        VALUE      :: ctx
      INTEGER(C_INT64_T), VALUE      :: ncell, het_stride
      REAL(C_DOUBLE),     VALUE      :: dt
      INTEGER(C_INT32_T), INTENT(IN) :: icntrl(20)
      REAL(C_DOUBLE),     INTENT(IN) :: rcntrl(20)
      TYPE(C_PTR),        VALUE      :: met, photol, het, hstart, conc, hnew, istatus, ierr, stream
    END FUNCTION
    ! Cells of the last gckpp_fullchem_step_dev that failed twice (synchronises)
    INTEGER(C_INT) FUNCTION gckpp_step_failures(ctx, nfail) BIND(C, name="gckpp_step_failures")
      IMPORT :: C_INT, C_INT64_T, C_PTR
      TYPE(C_PTR),        VALUE       :: ctx
      INTEGER(C_INT64_T), INTENT(OUT) :: nfail
    END FUNCTION
    ! Update_RCONST with host arrays (synchronous): MET(ncell,NMET),
    ! PHOTOL(ncell,NPHOTOL), HET(ncell or 1,NHET) (het_stride ncell or 0),
    ! C(ncell,NSPEC) -> RCONST(ncell,NREACT)
    INTEGER(C_INT) FUNCTION gckpp_rates(ctx, ncell, met, photol, het, het_stride, conc, rconst) &
         BIND(C, name="gckpp_rates")
      IMPORT :: C_INT, C_INT64_T, C_PTR, C_DOUBLE
      TYPE(C_PTR),        VALUE       :: ctx
      INTEGER(C_INT64_T), VALUE       :: ncell, het_stride
      REAL(C_DOUBLE),     INTENT(IN)  :: met(*), photol(*), het(*), conc(*)
      REAL(C_DOUBLE),     INTENT(OUT) :: rconst(*)
    END FUNCTION
    ! One chemistry step with host arrays (synchronous): Do_FullChem's batch
    ! call.  C(ncell,NSPEC) in/out, HSTART(ncell) = KPPHvalue in, HNEW(ncell)
    ! the next KPPHvalue, ISTATUS(ncell,20), IERR(ncell); any ncell (chunks of
    ! at most max_cells); returns -1 with "failed twice" when some cell's
    ! retry failed (its IERR < 0)
    INTEGER(C_INT) FUNCTION gckpp_fullchem_step(ctx, ncell, dt, icntrl, rcntrl, met, photol, het, &
         het_stride, hstart, conc, hnew, istatus, ierr) BIND(C, name="gckpp_fullchem_step")
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_PTR, C_DOUBLE
      TYPE(C_PTR),        VALUE         :: ctx
      INTEGER(C_INT64_T), VALUE         :: ncell, het_stride
      REAL(C_DOUBLE),     VALUE         :: dt
      INTEGER(C_INT32_T), INTENT(IN)    :: icntrl(20)
      REAL(C_DOUBLE),     INTENT(IN)    :: rcntrl(20)
      REAL(C_DOUBLE),     INTENT(IN)    :: met(*), photol(*), het(*), hstart(*)
      REAL(C_DOUBLE),     INTENT(INOUT) :: conc(*)
      REAL(C_DOUBLE),     INTENT(OUT)   :: hnew(*)
      INTEGER(C_INT32_T), INTENT(OUT)   :: istatus(*), ierr(*)
    END FUNCTION
    ! Host-pointer Integrate (synchronous): C(ncell,NSPEC), RCONST(ncell,NREACT), ...
    INTEGER(C_INT) FUNCTION gckpp_integrate(ctx, ncell, tin, tout, icntrl, rcntrl, hstart, conc, &
         rconst, istatus, rstatus, ierr) BIND(C, name="gckpp_integrate")
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_PTR, C_DOUBLE
      TYPE(C_PTR),        VALUE         :: ctx
      INTEGER(C_INT64_T), VALUE         :: ncell
      REAL(C_DOUBLE),     VALUE         :: tin, tout
      INTEGER(C_INT32_T), INTENT(IN)    :: icntrl(20)
      REAL(C_DOUBLE),     INTENT(IN)    :: rcntrl(20)
      REAL(C_DOUBLE),     INTENT(IN)    :: hstart(*)
      REAL(C_DOUBLE),     INTENT(INOUT) :: conc(*)
      REAL(C_DOUBLE),     INTENT(IN)    :: rconst(*)
      INTEGER(C_INT32_T), INTENT(OUT)   :: istatus(*), ierr(*)
      REAL(C_DOUBLE),     INTENT(OUT)   :: rstatus(*)
    END FUNCTION
    ! Fun's reaction rates A (host arrays, synchronous): C(ncell,NSPEC),
    ! RCONST(ncell,NREACT) -> AOUT(ncell,NREACT)
    INTEGER(C_INT) FUNCTION gckpp_aout(ctx, ncell, conc, rconst, aout) BIND(C, name="gckpp_aout")
      IMPORT :: C_INT, C_INT64_T, C_PTR, C_DOUBLE
      TYPE(C_PTR),        VALUE       :: ctx
      INTEGER(C_INT64_T), VALUE       :: ncell
      REAL(C_DOUBLE),     INTENT(IN)  :: conc(*), rconst(*)
      REAL(C_DOUBLE),     INTENT(OUT) :: aout(*)
    END FUNCTION
    ! One cell's KPP-Standalone state file (KppSa_Write_Samples); path
    ! NUL-terminated; names: NSPEC records of GCKPP_NAME_LEN characters
    ! (blank-padded), C(NSPEC), ATOL(NSPEC), RCONST(NREACT), AOUT(NREACT)
    INTEGER(C_INT) FUNCTION gckpp_kppsa_write(mech_id, path, meta, names, conc, atol, rconst, &
         aout) BIND(C, name="gckpp_kppsa_write")
      IMPORT :: C_INT, C_CHAR, C_DOUBLE, gckpp_kppsa_meta
      INTEGER(C_INT),         VALUE      :: mech_id
      CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*), names(*)
      TYPE(gckpp_kppsa_meta), INTENT(IN) :: meta
      REAL(C_DOUBLE),         INTENT(IN) :: conc(*), atol(*), rconst(*), aout(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION gckpp_kppsa_read(mech_id, path, meta, names, conc, atol, rconst, &
         aout) BIND(C, name="gckpp_kppsa_read")
      IMPORT :: C_INT, C_CHAR, C_DOUBLE, gckpp_kppsa_meta
      INTEGER(C_INT),         VALUE       :: mech_id
      CHARACTER(KIND=C_CHAR), INTENT(IN)  :: path(*)
      TYPE(gckpp_kppsa_meta), INTENT(OUT) :: meta
      CHARACTER(KIND=C_CHAR), INTENT(OUT) :: names(*)
      REAL(C_DOUBLE),         INTENT(OUT) :: conc(*), atol(*), rconst(*), aout(*)
    END FUNCTION
    ! KPPHvalue(nx,ny,nz) as the restart variable Chem_KPPHvalue (netCDF
    ! classic); path and model_tag NUL-terminated
    INTEGER(C_INT) FUNCTION gckpp_kpph_restart_write(path, nx, ny, nz, kpph, model_tag) &
         BIND(C, name="gckpp_kpph_restart_write")
      IMPORT :: C_INT, C_INT64_T, C_CHAR, C_DOUBLE
      CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*), model_tag(*)
      INTEGER(C_INT64_T),     VALUE      :: nx, ny, nz
      REAL(C_DOUBLE),         INTENT(IN) :: kpph(*)
    END FUNCTION
    ! found = 0 and KPPHvalue = 0 when the file has no Chem_KPPHvalue
    INTEGER(C_INT) FUNCTION gckpp_kpph_restart_read(path, nx, ny, nz, kpph, found, model_tag) &
         BIND(C, name="gckpp_kpph_restart_read")
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_CHAR, C_DOUBLE
      CHARACTER(KIND=C_CHAR), INTENT(IN)  :: path(*)
      INTEGER(C_INT64_T),     VALUE       :: nx, ny, nz
      REAL(C_DOUBLE),         INTENT(OUT) :: kpph(*)
      INTEGER(C_INT32_T),     INTENT(OUT) :: found
      CHARACTER(KIND=C_CHAR), INTENT(OUT) :: model_tag(*)
    END FUNCTION
    ! Map_KppSpc as 0-based model species (-1: none), molecular weights
    ! [g/mol] of the nmodel model species, prod/loss dummies (1) or null,
    ! alk_equiv 1: SALAAL/SALCAL to equivalents and back (fullchem)
    INTEGER(C_INT) FUNCTION gckpp_set_species_map(ctx, nmodel, map, mw_g, zero, alk_equiv) &
         BIND(C, name="gckpp_set_species_map")
      IMPORT :: C_INT, C_INT32_T, C_PTR, C_DOUBLE
      TYPE(C_PTR),        VALUE      :: ctx
      INTEGER(C_INT32_T), VALUE      :: nmodel, alk_equiv
      INTEGER(C_INT32_T), INTENT(IN) :: map(*)
      REAL(C_DOUBLE),     INTENT(IN) :: mw_g(*)
      TYPE(C_PTR),        VALUE      :: zero
    END FUNCTION
    ! device pointers: SPECIES(ncell,nmodel) kg/kg dry, AIRDEN(ncell) kg/m3,
    ! C(ncell,NSPEC) molec/cm3 (gather: C out; scatter: SPECIES out)
    INTEGER(C_INT) FUNCTION gckpp_gather_dev(ctx, ncell, species, airden, conc, stream) &
         BIND(C, name="gckpp_gather_dev")
      IMPORT :: C_INT, C_INT64_T, C_PTR
      TYPE(C_PTR),        VALUE :: ctx, species, airden, conc, stream
      INTEGER(C_INT64_T), VALUE :: ncell
    END FUNCTION
    INTEGER(C_INT) FUNCTION gckpp_scatter_dev(ctx, ncell, conc, airden, species, stream) &
         BIND(C, name="gckpp_scatter_dev")
      IMPORT :: C_INT, C_INT64_T, C_PTR
      TYPE(C_PTR),        VALUE :: ctx, conc, airden, species, stream
      INTEGER(C_INT64_T), VALUE :: ncell
    END FUNCTION
    ! fullchem_SetStateHet for a batch: het(ncell, NHET) from hetin(ncell, NHIN)
    ! and conc(ncell, NSPEC), device pointers
    INTEGER(C_INT) FUNCTION gckpp_set_state_het_dev(ctx, ncell, hetin, conc, naerotype, &
         gamma_ho2, lpscchem, turn_off_het, het, stream) BIND(C, name="gckpp_set_state_het_dev")
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR
      TYPE(C_PTR),        VALUE :: ctx, hetin, conc, het, stream
      INTEGER(C_INT64_T), VALUE :: ncell
      INTEGER(C_INT32_T), VALUE :: naerotype, lpscchem, turn_off_het
      REAL(C_DOUBLE),     VALUE :: gamma_ho2
    END FUNCTION
    ! fullchem_SulfurAqChem for a batch: K_MT(1:6) into met(ncell, NMET)
    INTEGER(C_INT) FUNCTION gckpp_sulfur_aq_dev(ctx, ncell, hetin, conc, met, stream) &
         BIND(C, name="gckpp_sulfur_aq_dev")
      IMPORT :: C_INT, C_INT64_T, C_PTR
      TYPE(C_PTR),        VALUE :: ctx, hetin, conc, met, stream
      INTEGER(C_INT64_T), VALUE :: ncell
    END FUNCTION
    ! The same two with host arrays (synchronous, chunked)
    INTEGER(C_INT) FUNCTION gckpp_set_state_het(ctx, ncell, hetin, conc, naerotype, gamma_ho2, &
         lpscchem, turn_off_het, het) BIND(C, name="gckpp_set_state_het")
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR
      TYPE(C_PTR),        VALUE       :: ctx
      INTEGER(C_INT64_T), VALUE       :: ncell
      REAL(C_DOUBLE),     INTENT(IN)  :: hetin(*), conc(*)
      INTEGER(C_INT32_T), VALUE       :: naerotype, lpscchem, turn_off_het
      REAL(C_DOUBLE),     VALUE       :: gamma_ho2
      REAL(C_DOUBLE),     INTENT(OUT) :: het(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION gckpp_sulfur_aq(ctx, ncell, hetin, conc, met) &
         BIND(C, name="gckpp_sulfur_aq")
      IMPORT :: C_INT, C_INT64_T, C_DOUBLE, C_PTR
      TYPE(C_PTR),        VALUE         :: ctx
      INTEGER(C_INT64_T), VALUE         :: ncell
      REAL(C_DOUBLE),     INTENT(IN)    :: hetin(*), conc(*)
      REAL(C_DOUBLE),     INTENT(INOUT) :: met(*)
    END FUNCTION
    TYPE(C_PTR) FUNCTION gckpp_last_error() BIND(C, name="gckpp_last_error")
      IMPORT :: C_PTR
    END FUNCTION
  END INTERFACE
END MODULE gckpp_hip_mod
