// Internal declarations shared by the HIP translation units of libgckpp_hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gckpp {

// Fortran MIN/MAX as the reference's compiler evaluates them (the first
// argument unless the second compares strictly smaller / larger), which fixes
// how NaN propagates through the step-size control.
__host__ __device__ inline double min_f(double a, double b) { return a < b ? a : b; }
__host__ __device__ inline double max_f(double a, double b) { return a > b ? a : b; }

enum { MECH_FULLCHEM = 0, MECH_HG = 1, MECH_CARBON = 2 };

struct MechInfo {
  int nspec, nvar, nfix, nreact, lu_nz, npart;
};

// Integrator settings shared by all cells, resolved on the host from
// ICNTRL/RCNTRL exactly as Rosenbrock() does (gckpp_Integrator.F90:165-532).
struct RosParams {
  double Tstart, Tend, Hmin, Hmax, Hstart_default, FacMin, FacMax, FacRej, FacSafe;
  double roundoff, DeltaMin, direction;
  int max_steps, vectol, autonomous, clip_negative;
  int neg_mode;  // ICNTRL(16) as given (forward Euler's negative handling)
};

struct IntegrateArgs {
  int64_t ncell;
  double *conc;          // [NSPEC][cstride], Y = rows 0..NVAR-1
  int64_t cstride;
  const double *rconst;  // [NREACT][rstride]
  int64_t rstride;
  const double *hstart;  // [ncell] or null
  const double *atol, *rtol;  // [NVAR] device
  double *work;          // per-cell workspace [wsize][wstride]
  int64_t wstride;
  int32_t *istatus;      // [20][sstride] or null
  double *rstatus;       // [20][sstride] or null
  int32_t *ierr;         // [ncell] or null
  int64_t sstride;
  const int32_t *only_failed;  // if set, solve only cells with only_failed[cell] < 0
  const double *restore_from;  // if set, solved cells first copy conc from here
  RosParams prm;
  // persistent solver only
  double *hnew;          // [ncell] RSTATUS(3) (next Hstart) or null
  int retry;             // a failed cell is integrated once more with Hstart = 0
  int clip_out;          // write back max(C, 0) for all NSPEC species
  const int32_t *order;  // queue position -> cell (a permutation), or null
  float *hlog;           // [2*hlog_n][ncell] (H, err) of the first attempts, or null
  int hlog_n;
};

// Tail hand-off of a persistent launch (cells of the Euler mechanism stay on
// their lanes).  The lane kernel, once at most `cap` of its cells are
// unfinished, stops each
// cell at its next step boundary (before the first attempt of a step, where
// the integrator state is T, H, the counters and Y) and saves that state here;
// a cell-per-wave launch after it (resume = 1) continues those cells.  Every
// cell handed off was one of the at most `cap` unfinished cells when the
// threshold was first met, so the list never overflows (an entry past `cap`
// is not used: that cell stays on its lane).
struct Handoff {
  unsigned int *done;    // lane kernel: cells finished or handed off
  unsigned int *count;   // cells handed off (entries used: min(count, cap))
  unsigned int *attempts;  // k_wave resume: attempts it made (a statistic)
  int32_t *cell;         // [cap] the handed-off cells
  double *rec;           // [HO_FIELDS + max NVAR][cap] their state, field-major
  int64_t cap;           // 0: no hand-off
  int resume;            // k_wave: take the cells of the list and resume them
};
constexpr int HO_FIELDS = 15;  // T, H, Hexit, Texit, HnewAcc, ISTATUS(1:8), tries, queue

// One launch of the persistent solver: up to three queues of cells (one per
// mechanism); global index c < qend[0] is cell c of queue 0, and so on.
struct PersistArgs {
  IntegrateArgs q[3];
  int64_t qend[3];
  double *work;          // [nblocks][slot_fields][64]
  int64_t slot_fields;
  unsigned int *counter; // work-stealing counter, or null (one cell per lane)
  Handoff ho;
};

struct RatesArgs {
  int64_t ncell;
  const double *met;     // [NMET][ncell]
  const double *photol;  // [NPHOTOL][ncell]
  const double *het;     // [NHET][het_stride]
  int64_t het_stride;    // 0 or ncell
  const double *conc;    // [NSPEC][ncell]
  double *rconst;        // [NREACT][ncell]
  const double *consts;  // [3][NSPEC]: MW, HENRY_K0, HENRY_CR
};

struct CompArgs {
  int64_t ncell;
  const double *in;
  int64_t istride;
  const double *rconst;
  int64_t rstride;
  double *out;
  int64_t ostride;
  double *work;
  int64_t wstride;
  int32_t *ier;
};

MechInfo mech_info(int mech);
// producers.hip: the gather of the KPP vector from the model's species
// (kg/kg dry) and the scatter back
// fullchem_SetStateHet per cell (producers.hip): hetin[NHIN][n], conc[NSPEC][n]
// of fullchem -> het[NHET][n].
void launch_set_state_het(int64_t n, const double *hetin, const double *conc, double *het,
                          int naerotype, double gamma_ho2, int lpscchem, int turn_off_het,
                          hipStream_t s);
// fullchem_SulfurAqChem per cell (producers.hip): K_MT(1:6) into met[MET_K_MT..]
// from hetin, conc and met's TEMP and NUMDEN; consts = the context's MW table.
void launch_sulfur_aq(int64_t n, const double *hetin, const double *conc, const double *consts,
                      double *met, hipStream_t s);
void launch_gather(int64_t n, int nspec, const int32_t *map, const double *cin, const double *eq,
                   const int32_t *zero, const double *species, const double *airden, double *conc,
                   hipStream_t s);
void launch_scatter(int64_t n, int nspec, const int32_t *map, const double *cout, const double *eq,
                    const double *conc, const double *airden, double *species, hipStream_t s);
int set_error(const char *msg);  // gckpp_last_error's message; returns -1
// doubles of per-cell workspace the integrator needs
inline int64_t integrate_work_per_cell(const MechInfo &m) {
  int64_t scr = m.nreact > m.npart ? m.nreact : m.npart;
  return (int64_t)(m.lu_nz > 0 ? m.lu_nz : 1) + 6 * (int64_t)m.nvar + scr;
}
void launch_integrate_mech(int mech, const IntegrateArgs &a, hipStream_t s);
// persistent work-stealing integrator (k_persist): queue sets and variants
enum { PQ_FULLCHEM = 0, PQ_HG = 1, PQ_CARBON = 2, PQ_MIXED = 3 };
constexpr int PERSIST_VARIANTS = 10;
constexpr int PERSIST_WAVE_VARIANT = 7;  // one wave per cell (wave_solver.hip)
int persist_variant_info(int variant, int *U, int *W);
int persist_variant_cache(int variant);  // operand-cache slots per lane (0: none)
int persist_variant_seg(int variant);  // 0 lane per cell, 1 with flat segments, 2 wave per cell
const void *wave_kernel(int qset);     // the cell-per-wave kernel (64 threads per workgroup)
int64_t wave_scratch_fields(int qset);  // its global scratch per wave, fields of 64 doubles
int64_t persist_slot_fields(int qset);                   // doubles per lane slot
int64_t persist_resident_waves(int qset, int variant, int wpb);  // 0 on error
int launch_persist(int qset, int variant, int wpb, const PersistArgs &pa, int64_t nwaves,
                   hipStream_t s);
void launch_zero_u32(unsigned int *p, int n, hipStream_t s);  // p[0..n) = 0, n <= 64
void launch_comp_mech(int mech, int which, const CompArgs &a, hipStream_t s);

}  // namespace gckpp
