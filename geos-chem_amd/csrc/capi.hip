// C ABI of libgckpp_hip (include/gckpp_hip.h): contexts, option resolution,
// workspace and chunking around the HIP kernels.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gckpp_hip.h"
#include "cell_fields.h"
#include "gckpp_internal.h"
#include "gen/species_fullchem.h"
#include "mlp.h"

namespace gckpp {
void launch_rates_mech(int mech, const RatesArgs &a, hipStream_t s);
bool rates_available(int mech);
}

using namespace gckpp;

static thread_local char g_err[512];

// Avogadro's number, Headers/physconstants.F90:32
constexpr double kAvogadro = 6.022140857e+23;

// Records the message of a failed call.  The HIP runtime's last-error state
// is consumed too: a failure reported here (a failed allocation, say) must
// not be read back later as the error of an unrelated kernel launch.
static int fail(const char *fmt, ...) {
  (void)hipGetLastError();
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return -1;
}

#define HIP_OK(x)                                                              \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess)                                                      \
      return fail("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
                  __LINE__);                                                   \
  } while (0)

struct gckpp_ctx {
  int mech;
  int device;
  MechInfo info;
  int64_t max_cells;
  int64_t wstride;     // cells per workspace chunk (multiple of 64)
  double *work;        // [integrate_work_per_cell][wstride]; persistent slots
  int64_t work_doubles;
  double *atol, *rtol; // [NVAR] device
  double *consts;      // [3][NSPEC] device: MW, HENRY_K0, HENRY_CR
  double *step_buf;    // fullchem_step: rconst [NREACT][wstride], conc backup
                       // [NSPEC][wstride], rstatus [20][wstride]
  int32_t *nfail;      // [1] device counter of cells that failed twice
  hipEvent_t ev[5];    // step timing: rates | backup | integrate | retry+finish
  int solver;          // GCKPP_SOLVER_PLAIN or GCKPP_SOLVER_PERSISTENT
  unsigned int *counter;  // work-stealing cell counter of the persistent solver
  const int32_t *order;  // persistent solver: queue order of the cells (device), or null
  float *hlog;         // persistent solver: step-history log (device), or null
  int hlog_n;
  int pvariant;        // persistent kernel variant (loads in flight x register budget)
  int pwpb;            // waves per workgroup of the persistent launch
  int64_t pwaves;      // waves of one persistent launch (multiple of pwpb)
  int64_t pwaves_max;  // caller's cap on pwaves (0: every resident wave)
  // scratch for the host-pointer and retry paths
  std::vector<double> h_atol, h_rtol;
  double *hs_d;        // host-pointer calls: device staging of one chunk (allocated on use)
  int32_t *hs_i;
  // split persistent solve (gckpp_set_persistent_split): the first `split`
  // cells of the order on the cell-per-wave kernel, on stream s2
  int64_t split;
  double *wwork;       // its slot blocks, wwaves waves
  int64_t wwaves;
  unsigned int *wcounter;
  hipStream_t s2;
  hipEvent_t ev_fork, ev_join;
  // tail hand-off (gckpp_set_persistent_handoff): at most ho_cap cells of a
  // lane-kernel launch move to a cell-per-wave launch after it
  int64_t ho_cap, ho_alloc;
  unsigned int *ho_ctr;  // [4]: cells done, cells handed off, the resume launch's counter,
                         // the resume launch's attempts
  hipEvent_t ev_ho[3];   // lane launch start | resume start | resume end
  int32_t *ho_cell;      // [ho_alloc]
  double *ho_rec;        // [HO_FIELDS + NVAR][ho_cap] (room for ho_alloc)
  double *hwork;         // scratch of the resume launch, hwaves waves of hw_fields*64
  int64_t hwaves, hw_fields;
  bool ho_timed;         // ev_ho recorded by a launch
  int64_t na_waves;      // resident waves of the cell-per-wave kernel (no-allocation entry)
  // gckpp_set_species_map: the model species of each KPP species (-1:
  // none), the unit-conversion factors, the alkalinity-equivalent factors
  // (1 for the other species), the prod/loss dummies (device)
  int32_t *sm_map, *sm_zero;
  double *sm_cin, *sm_cout, *sm_eq;
  int32_t sm_nmodel;
};

extern "C" const char *gckpp_last_error(void) { return g_err; }

namespace gckpp {
// The message of a failure detected outside this file (host-side file I/O);
// an empty message clears it.  Returns -1.
int set_error(const char *msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return -1;
}
}  // namespace gckpp

static bool mech_ok(int mech) {
  return mech == MECH_FULLCHEM || mech == MECH_HG || mech == MECH_CARBON;
}

extern "C" int gckpp_mech_sizes(int mech_id, int32_t out[8]) {
  if (!mech_ok(mech_id)) return fail("unknown or unsupported mechanism id %d", mech_id);
  MechInfo m = mech_info(mech_id);
  out[0] = m.nspec; out[1] = m.nvar; out[2] = m.nfix; out[3] = m.nreact;
  out[4] = m.lu_nz; out[5] = NMET; out[6] = NHET; out[7] = NPHOTOL;
  g_err[0] = 0;
  return 0;
}

// Persistent configuration and its launch size: every resident wave, at
// most one wave per 64 cells of the workspace, a whole number of workgroups.
static const int kDefaultPersistVariant = 4, kDefaultPersistWpb = 1;
// Waves a persistent launch needs for ncell cells (before the resident cap):
// a lane-per-cell wave takes 64 cells at a time, a wave-per-cell wave one.
static int64_t cells_to_waves(int variant, int64_t ncell) {
  return variant == PERSIST_WAVE_VARIANT ? ncell : (ncell + 63) / 64;
}
// Fields (64 doubles) of one wave's block of the persistent workspace: a
// lane-per-cell wave's 64 slots, or a cell-per-wave wave's scratch.
static int64_t launch_slot_fields(int qset, int variant) {
  return variant == PERSIST_WAVE_VARIANT ? wave_scratch_fields(qset) : persist_slot_fields(qset);
}
static int set_pconfig(gckpp_ctx *c, int variant, int wpb) {
  int64_t cap = c->wstride / 64 / wpb * wpb;
  if (variant == PERSIST_WAVE_VARIANT)   // waves whose scratch fits the workspace
    cap = c->work_doubles / (64 * wave_scratch_fields(c->mech)) / wpb * wpb;
  int64_t res = persist_resident_waves(c->mech, variant, wpb);  // PQ_* == MECH_* for one mechanism
  if (cap <= 0 || res <= 0) return -1;
  c->pvariant = variant;
  c->pwpb = wpb;
  c->pwaves = res < cap ? res / wpb * wpb : cap;
  if (c->pwaves_max > 0 && c->pwaves > c->pwaves_max) c->pwaves = c->pwaves_max / wpb * wpb;
  if (c->pwaves <= 0) c->pwaves = wpb;
  return 0;
}

// Scratch of gckpp_integrate_noalloc_dev: the work-stealing counter
// (the first 256 bytes) and the blocks of up to kNoAllocWaves waves of the
// cell-per-wave solver (each wave's rate constants, reaction products,
// Fcn0 and stage vectors: wave_scratch_fields(mech) fields of 64 doubles,
// 73 for fullchem, the largest).  768 waves are three fullchem cells per
// CU on 256 CUs; 28 MiB per device, held until the process ends.  Host
// code reaches each device's copy through the address that
// hipGetSymbolAddress gives with that device current (g_na_ptr), never
// through the array's name.
constexpr int64_t kNoAllocWaves = 768, kNoAllocFields = 73, kNoAllocHead = 32;
constexpr int64_t kNoAllocDoubles = kNoAllocHead + kNoAllocWaves * kNoAllocFields * 64;
static __device__ __align__(256) double scratch[kNoAllocDoubles];
constexpr int kMaxDevices = 64;
static std::mutex g_na_mu;
static double *g_na_ptr[kMaxDevices];

// Once per device, with the device current: the address of its scratch.
static int noalloc_setup(int dev) {
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (dev < 0 || dev >= kMaxDevices) return fail("gckpp_init: device %d beyond %d", dev, kMaxDevices);
  if (g_na_ptr[dev]) return 0;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail("gckpp_init: no address for the static scratch: %s", hipGetErrorString(e));
  g_na_ptr[dev] = (double *)p;
  return 0;
}

// A failed gckpp_init releases what it allocated before reporting.
#define INIT_OK(x)                                                             \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      gckpp_finalize(c);                                                       \
      return fail("gckpp_init: %s failed: %s", #x, hipGetErrorString(e_));     \
    }                                                                          \
  } while (0)

extern "C" int gckpp_init(int mech_id, int64_t max_cells, gckpp_ctx **out) {
  if (!out) return fail("gckpp_init: null output pointer");
  *out = nullptr;
  if (!mech_ok(mech_id)) return fail("gckpp_init: unknown or unsupported mechanism id %d", mech_id);
  if (max_cells <= 0) return fail("gckpp_init: max_cells must be positive");
  auto *c = new gckpp_ctx();  // value-initialised: every pointer and event null
  c->mech = mech_id;
  c->info = mech_info(mech_id);
  INIT_OK(hipGetDevice(&c->device));
  if (noalloc_setup(c->device) != 0) {
    const std::string why = g_err;
    gckpp_finalize(c);
    return fail("%s", why.c_str());
  }
  c->na_waves = persist_resident_waves(mech_id, PERSIST_WAVE_VARIANT, 1);
  c->max_cells = max_cells;
  c->wstride = (max_cells + 63) / 64 * 64;
  int64_t per_cell = integrate_work_per_cell(c->info);
  if (persist_slot_fields(mech_id) > per_cell) per_cell = persist_slot_fields(mech_id);
  size_t wbytes = sizeof(double) * (size_t)per_cell * (size_t)c->wstride;
  c->work_doubles = per_cell * c->wstride;
  hipError_t e = hipMalloc(&c->work, wbytes);
  if (e != hipSuccess) {
    c->work = nullptr;
    gckpp_finalize(c);
    return fail("gckpp_init: cannot allocate %.2f GB of workspace for %lld cells: %s",
                wbytes / 1e9, (long long)max_cells, hipGetErrorString(e));
  }
  INIT_OK(hipMalloc(&c->atol, sizeof(double) * c->info.nvar));
  INIT_OK(hipMalloc(&c->rtol, sizeof(double) * c->info.nvar));
  INIT_OK(hipMalloc(&c->consts, sizeof(double) * 3 * c->info.nspec));
  INIT_OK(hipMemset(c->consts, 0, sizeof(double) * 3 * c->info.nspec));
  {
    size_t sb = sizeof(double) * (size_t)(c->info.nreact + c->info.nspec + 20) * (size_t)c->wstride;
    hipError_t e2 = hipMalloc(&c->step_buf, sb);
    if (e2 != hipSuccess) {
      c->step_buf = nullptr;
      gckpp_finalize(c);
      return fail("gckpp_init: cannot allocate %.2f GB of step buffers: %s", sb / 1e9,
                  hipGetErrorString(e2));
    }
  }
  INIT_OK(hipMalloc(&c->nfail, sizeof(int32_t)));
  for (int k = 0; k < 5; ++k) INIT_OK(hipEventCreate(&c->ev[k]));
  INIT_OK(hipMalloc(&c->counter, sizeof(unsigned int)));
  c->solver = GCKPP_SOLVER_PLAIN;
  if (set_pconfig(c, kDefaultPersistVariant, kDefaultPersistWpb) != 0) {
    c->pvariant = kDefaultPersistVariant;
    c->pwpb = 1;
    c->pwaves = c->wstride / 64 < 1024 ? c->wstride / 64 : 1024;
  }
  // Init_FullChem defaults (GeosCore/fullchem_mod.F90:2899-2905)
  c->h_atol.assign(c->info.nvar, 1.0e-2);
  c->h_rtol.assign(c->info.nvar, 0.5e-2);
  INIT_OK(hipMemcpy(c->atol, c->h_atol.data(), sizeof(double) * c->info.nvar, hipMemcpyHostToDevice));
  INIT_OK(hipMemcpy(c->rtol, c->h_rtol.data(), sizeof(double) * c->info.nvar, hipMemcpyHostToDevice));
  *out = c;
  g_err[0] = 0;
  return 0;
}

extern "C" int gckpp_finalize(gckpp_ctx *c) {
  if (!c) return 0;
  bool ok = true;
  for (void *p : {(void *)c->work, (void *)c->atol, (void *)c->rtol, (void *)c->consts,
                  (void *)c->step_buf, (void *)c->nfail, (void *)c->counter, (void *)c->hs_d,
                  (void *)c->hs_i, (void *)c->wwork, (void *)c->wcounter, (void *)c->ho_ctr,
                  (void *)c->ho_cell, (void *)c->ho_rec, (void *)c->hwork, (void *)c->sm_map,
                  (void *)c->sm_zero, (void *)c->sm_cin, (void *)c->sm_cout, (void *)c->sm_eq})
    if (p) ok = hipFree(p) == hipSuccess && ok;
  if (c->s2) ok = hipStreamDestroy(c->s2) == hipSuccess && ok;
  if (c->ev_fork) ok = hipEventDestroy(c->ev_fork) == hipSuccess && ok;
  if (c->ev_join) ok = hipEventDestroy(c->ev_join) == hipSuccess && ok;
  for (int k = 0; k < 3; ++k)
    if (c->ev_ho[k]) ok = hipEventDestroy(c->ev_ho[k]) == hipSuccess && ok;
  for (int k = 0; k < 5; ++k)
    if (c->ev[k]) ok = hipEventDestroy(c->ev[k]) == hipSuccess && ok;
  delete c;
  return ok ? 0 : fail("gckpp_finalize: releasing device memory failed");
}

extern "C" int gckpp_set_tolerances(gckpp_ctx *c, const double *atol, const double *rtol) {
  if (!c || !atol || !rtol) return fail("gckpp_set_tolerances: null argument");
  c->h_atol.assign(atol, atol + c->info.nvar);
  c->h_rtol.assign(rtol, rtol + c->info.nvar);
  HIP_OK(hipMemcpy(c->atol, atol, sizeof(double) * c->info.nvar, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(c->rtol, rtol, sizeof(double) * c->info.nvar, hipMemcpyHostToDevice));
  return 0;
}

extern "C" int gckpp_set_species_consts(gckpp_ctx *c, const double *mw, const double *hk0,
                                        const double *hcr) {
  if (!c || !mw || !hk0 || !hcr) return fail("gckpp_set_species_consts: null argument");
  size_t n = c->info.nspec;
  HIP_OK(hipMemcpy(c->consts, mw, sizeof(double) * n, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(c->consts + n, hk0, sizeof(double) * n, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(c->consts + 2 * n, hcr, sizeof(double) * n, hipMemcpyHostToDevice));
  return 0;
}

// Rosenbrock() option handling (gckpp_Integrator.F90:80-162 merge rule,
// :165-532 defaults and checks).  Returns 0, or a KPP IERR (<0) that every
// cell gets, or -100 for an option this library does not implement.
static int resolve_options(const gckpp_ctx *c, double tin, double tout, const int32_t *icu,
                           const double *rcu, RosParams *p, char *why, size_t nwhy) {
  int ic[20] = {0};
  double rc[20] = {0};
  ic[14] = 5;
  for (int i = 0; i < 20; ++i) {
    if (icu && icu[i] != 0) ic[i] = icu[i];
    if (rcu && rcu[i] > 0) rc[i] = rcu[i];
  }
  memset(p, 0, sizeof *p);
  p->Tstart = tin;
  p->Tend = tout;
  p->neg_mode = ic[15];
  if (c->mech == MECH_CARBON) {
    // forward Euler (KPP/carbon/gckpp_Integrator.F90:93-150): no step options
    if (ic[14] != -1) {
      snprintf(why, nwhy, "ICNTRL(15)=%d: rates must be frozen during the step (-1)", ic[14]);
      return -100;
    }
    return 0;
  }
  p->autonomous = ic[0] != 0;
  p->vectol = ic[1] == 0;
  if (!(ic[2] == 0 || ic[2] == 4)) {
    if (ic[2] >= 1 && ic[2] <= 6) {
      snprintf(why, nwhy, "ICNTRL(3)=%d: only Rodas3 (0 or 4) is implemented", ic[2]);
      return -100;
    }
    return -2;
  }
  if (ic[14] != -1) {
    snprintf(why, nwhy, "ICNTRL(15)=%d: rates must be frozen during the step (-1)", ic[14]);
    return -100;
  }
  if (ic[11] == 1) {
    snprintf(why, nwhy, "ICNTRL(12)=1: the auto-reduce integrator is not implemented");
    return -100;
  }
  if (ic[3] == 0) p->max_steps = 200000;
  else if (ic[3] > 0) p->max_steps = ic[3];
  else return -1;
  p->roundoff = DBL_EPSILON;  // WLAMCH('E')
  p->DeltaMin = 1.0e-5;
  if (rc[0] == 0.0) p->Hmin = 0.0; else p->Hmin = rc[0];
  if (rc[1] == 0.0) p->Hmax = fabs(tout - tin);
  else p->Hmax = fmin(fabs(rc[1]), fabs(tout - tin));
  if (rc[2] == 0.0) p->Hstart_default = fmax(p->Hmin, p->DeltaMin);
  else p->Hstart_default = fmin(fabs(rc[2]), fabs(tout - tin));
  p->FacMin = rc[3] == 0.0 ? 0.2 : rc[3];
  p->FacMax = rc[4] == 0.0 ? 6.0 : rc[4];
  p->FacRej = rc[5] == 0.0 ? 0.1 : rc[5];
  p->FacSafe = rc[6] == 0.0 ? 0.9 : rc[6];
  int up = p->vectol ? c->info.nvar : 1;
  for (int i = 0; i < up; ++i)
    if (c->h_atol[i] <= 0.0 || c->h_rtol[i] <= 10.0 * p->roundoff || c->h_rtol[i] >= 1.0) return -5;
  p->direction = tout >= tin ? 1.0 : -1.0;
  p->clip_negative = ic[15] == 1;
  return 0;
}

__global__ void k_fill_status(int64_t n, int32_t *ierr, int code, int32_t *ist, double *rst) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (ierr) ierr[i] = code;
  for (int k = 0; k < 20; ++k) {
    if (ist) ist[(int64_t)k * n + i] = 0;
    if (rst) rst[(int64_t)k * n + i] = 0.0;
  }
}

// One queue of a persistent launch in step mode: the callers' retry for the
// Rosenbrock mechanisms, the write-back clip except for carbon (whose caller
// stores C as it is, carbon_Funcs.F90:300-322).
static void fill_persist_queue(const gckpp_ctx *c, IntegrateArgs &a, int64_t ncell,
                               const RosParams &p, double *conc, const double *rct,
                               const double *hstart, double *hnew, int32_t *istatus,
                               double *rstatus, int32_t *ierr) {
  a.ncell = ncell;
  a.conc = conc;
  a.cstride = ncell;
  a.rconst = rct;
  a.rstride = ncell;
  a.hstart = hstart;
  a.atol = c->atol;
  a.rtol = c->rtol;
  a.prm = p;
  a.istatus = istatus;
  a.rstatus = rstatus;
  a.ierr = ierr;
  a.sstride = ncell;
  a.hnew = hnew;
  a.retry = c->mech != MECH_CARBON;
  a.clip_out = c->mech != MECH_CARBON;
}

// The persistent solve of one queue (pa.q[0], ncell cells): one launch of
// the configured variant, or -- with a split set, a cell order and more
// cells than the split -- the first `split` cells of the order on the
// cell-per-wave kernel on the second stream, concurrently with the
// configured variant on the rest of the order; work later enqueued on s
// waits for both.  With a hand-off set (lane variants), the lane kernel's
// last unfinished cells move at a step boundary to a cell-per-wave launch
// that follows it on s.
static int launch_persist_queue(gckpp_ctx *c, PersistArgs &pa, int64_t ncell, hipStream_t s,
                                const char *who) {
  const bool can_split = c->split > 0 && c->wwork && pa.q[0].order && ncell > c->split &&
                         c->mech != MECH_CARBON && c->pvariant != PERSIST_WAVE_VARIANT;
  const int64_t k = can_split ? c->split : 0;
  if (k > 0) {
    PersistArgs pw = pa;
    pw.qend[0] = pw.qend[1] = pw.qend[2] = k;
    pw.work = c->wwork;
    pw.slot_fields = wave_scratch_fields(c->mech);
    pw.counter = c->wcounter;
    HIP_OK(hipEventRecord(c->ev_fork, s));
    HIP_OK(hipStreamWaitEvent(c->s2, c->ev_fork, 0));
    if (launch_persist(c->mech, PERSIST_WAVE_VARIANT, 1, pw, k < c->wwaves ? k : c->wwaves,
                       c->s2) != 0)
      return fail("%s: persistent launch failed: %s", who, hipGetErrorString(hipGetLastError()));
    HIP_OK(hipEventRecord(c->ev_join, c->s2));
    pa.q[0].order += k;
    pa.qend[0] = pa.qend[1] = pa.qend[2] = ncell - k;
  }
  const bool hand = c->ho_cap > 0 && c->ho_rec && c->mech != MECH_CARBON &&
                    c->pvariant != PERSIST_WAVE_VARIANT && ncell - k > 0;
  c->ho_timed = false;  // the statistics describe this solve from here on
  if (hand) {
    launch_zero_u32(c->ho_ctr, 4, s);
    pa.ho = Handoff{c->ho_ctr, c->ho_ctr + 1, c->ho_ctr + 3, c->ho_cell, c->ho_rec, c->ho_cap, 0};
    HIP_OK(hipEventRecord(c->ev_ho[0], s));
  }
  int64_t nw = cells_to_waves(c->pvariant, ncell - k);
  nw = (nw + c->pwpb - 1) / c->pwpb * c->pwpb;
  if (nw > c->pwaves) nw = c->pwaves;
  if (launch_persist(c->mech, c->pvariant, c->pwpb, pa, nw, s) != 0)
    return fail("%s: persistent launch failed: %s", who, hipGetErrorString(hipGetLastError()));
  if (hand) {
    HIP_OK(hipEventRecord(c->ev_ho[1], s));
    PersistArgs ph = pa;
    ph.ho.resume = 1;
    ph.work = c->hwork;
    ph.slot_fields = c->hw_fields;
    ph.counter = c->ho_ctr + 2;
    if (launch_persist(c->mech, PERSIST_WAVE_VARIANT, 1, ph, c->hwaves, s) != 0)
      return fail("%s: hand-off launch failed: %s", who, hipGetErrorString(hipGetLastError()));
    HIP_OK(hipEventRecord(c->ev_ho[2], s));
    c->ho_timed = true;
    pa.ho = Handoff{};
  }
  if (k > 0) HIP_OK(hipStreamWaitEvent(s, c->ev_join, 0));
  return 0;
}

static int integrate_impl(gckpp_ctx *c, int64_t ncell, double tin, double tout,
                          const int32_t icntrl[20], const double rcntrl[20],
                          const double *hstart, double *conc, const double *rconst,
                          int32_t *istatus, double *rstatus, int32_t *ierr,
                          const int32_t *only_failed, const double *restore_from,
                          void *stream) {
  if (!c) return fail("gckpp_integrate: null context");
  if (ncell < 0) return fail("gckpp_integrate: negative ncell");
  if (ncell == 0) return 0;
  if (!conc || !rconst) return fail("gckpp_integrate: null conc or rconst");
  hipStream_t s = (hipStream_t)stream;
  RosParams p;
  char why[256] = "";
  int code = resolve_options(c, tin, tout, icntrl, rcntrl, &p, why, sizeof why);
  if (code == -100) return fail("gckpp_integrate: %s", why);
  if (code < 0) {
    // every cell fails the same option check, as Rosenbrock() would
    c->ho_timed = false;
    hipLaunchKernelGGL(k_fill_status, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, s,
                       ncell, ierr, code, istatus, rstatus);
    HIP_OK(hipGetLastError());
    return 0;
  }
  if (c->solver == GCKPP_SOLVER_PERSISTENT && !only_failed && !restore_from) {
    // one launch of resident waves stealing cells; any ncell
    PersistArgs pa{};
    IntegrateArgs &a = pa.q[0];
    a.ncell = ncell;
    a.conc = conc;
    a.cstride = ncell;
    a.rconst = rconst;
    a.rstride = ncell;
    a.hstart = hstart;
    a.atol = c->atol;
    a.rtol = c->rtol;
    a.prm = p;
    a.istatus = istatus;
    a.rstatus = rstatus;
    a.ierr = ierr;
    a.sstride = ncell;
    a.order = c->order;
    a.hlog = c->hlog;
    a.hlog_n = c->hlog_n;
    pa.qend[0] = pa.qend[1] = pa.qend[2] = ncell;
    pa.work = c->work;
    pa.slot_fields = launch_slot_fields(c->mech, c->pvariant);
    pa.counter = c->counter;
    return launch_persist_queue(c, pa, ncell, s, "gckpp_integrate");
  }
  // chunks of at most wstride cells; outputs are indexed with the full ncell
  if (!only_failed) c->ho_timed = false;  // no hand-off in this solve
  for (int64_t c0 = 0; c0 < ncell; c0 += c->wstride) {
    int64_t n = ncell - c0 < c->wstride ? ncell - c0 : c->wstride;
    IntegrateArgs a{};
    a.ncell = n;
    a.conc = conc + c0;
    a.cstride = ncell;
    a.rconst = rconst + c0;
    a.rstride = ncell;
    a.hstart = hstart ? hstart + c0 : nullptr;
    a.atol = c->atol;
    a.rtol = c->rtol;
    a.work = c->work;
    a.wstride = c->wstride;
    a.prm = p;
    a.istatus = istatus ? istatus + c0 : nullptr;
    a.rstatus = rstatus ? rstatus + c0 : nullptr;
    a.ierr = ierr ? ierr + c0 : nullptr;
    a.sstride = ncell;
    a.only_failed = only_failed ? only_failed + c0 : nullptr;
    a.restore_from = restore_from ? restore_from + c0 : nullptr;
    launch_integrate_mech(c->mech, a, s);
    HIP_OK(hipGetLastError());
  }
  return 0;
}

extern "C" int gckpp_integrate_dev(gckpp_ctx *c, int64_t ncell, double tin, double tout,
                                   const int32_t icntrl[20], const double rcntrl[20],
                                   const double *hstart, double *conc, const double *rconst,
                                   int32_t *istatus, double *rstatus, int32_t *ierr,
                                   void *stream) {
  return integrate_impl(c, ncell, tin, tout, icntrl, rcntrl, hstart, conc, rconst, istatus,
                        rstatus, ierr, nullptr, nullptr, stream);
}

// Integrate without allocating: the cell-per-wave solver with its scratch
// in the static array above (see the header).  Everything it touches is
// the caller's arrays, the context's tolerances and that array.
extern "C" int gckpp_integrate_noalloc_dev(gckpp_ctx *c, int64_t ncell, double tin, double tout,
                                           const int32_t icntrl[20], const double rcntrl[20],
                                           const double *hstart, double *conc,
                                           const double *rconst, int32_t *istatus,
                                           double *rstatus, int32_t *ierr, void *stream) {
  if (!c) return fail("gckpp_integrate_noalloc_dev: null context");
  if (ncell < 0) return fail("gckpp_integrate_noalloc_dev: negative ncell");
  if (ncell == 0) return 0;
  if (!conc || !rconst) return fail("gckpp_integrate_noalloc_dev: null conc or rconst");
  double *ptr = (c->device >= 0 && c->device < kMaxDevices) ? g_na_ptr[c->device] : nullptr;
  if (!ptr) return fail("gckpp_integrate_noalloc_dev: no static scratch on device %d", c->device);
  const int64_t fields = wave_scratch_fields(c->mech);
  const int64_t cap = (kNoAllocDoubles - kNoAllocHead) / (64 * fields);
  if (cap < 1)
    return fail("gckpp_integrate_noalloc_dev: one wave needs %lld doubles of scratch, the static "
                "scratch holds %lld", (long long)(64 * fields), (long long)(kNoAllocDoubles - kNoAllocHead));
  if (c->pwaves_max > cap)
    return fail("gckpp_integrate_noalloc_dev: the context's wave cap (%lld waves, "
                "gckpp_set_persistent_waves) exceeds the %lld waves (%.1f MB) the static scratch "
                "holds for this mechanism", (long long)c->pwaves_max, (long long)cap,
                kNoAllocDoubles * 8.0 / 1e6);
  hipStream_t s = (hipStream_t)stream;
  RosParams p;
  char why[256] = "";
  const int code = resolve_options(c, tin, tout, icntrl, rcntrl, &p, why, sizeof why);
  if (code == -100) return fail("gckpp_integrate_noalloc_dev: %s", why);
  if (code < 0) {
    hipLaunchKernelGGL(k_fill_status, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, s,
                       ncell, ierr, code, istatus, rstatus);
    HIP_OK(hipGetLastError());
    return 0;
  }
  int64_t nw = c->na_waves > 0 ? c->na_waves : 1;
  if (nw > cap) nw = cap;
  if (c->pwaves_max > 0 && nw > c->pwaves_max) nw = c->pwaves_max;
  if (nw > ncell) nw = ncell;
  PersistArgs pa{};
  IntegrateArgs &a = pa.q[0];
  a.ncell = ncell;
  a.conc = conc;
  a.cstride = ncell;
  a.rconst = rconst;
  a.rstride = ncell;
  a.hstart = hstart;
  a.atol = c->atol;
  a.rtol = c->rtol;
  a.prm = p;
  a.istatus = istatus;
  a.rstatus = rstatus;
  a.ierr = ierr;
  a.sstride = ncell;
  pa.qend[0] = pa.qend[1] = pa.qend[2] = ncell;
  pa.work = ptr + kNoAllocHead;
  pa.slot_fields = fields;
  pa.counter = (unsigned int *)ptr;   // zeroed by a kernel in launch_persist
  if (launch_persist(c->mech, PERSIST_WAVE_VARIANT, 1, pa, nw, s) != 0)
    return fail("gckpp_integrate_noalloc_dev: launch failed: %s", hipGetErrorString(hipGetLastError()));
  return 0;
}

static int comp(gckpp_ctx *c, int which, int64_t ncell, const double *in, const double *rct,
                double *out, int32_t *ier, void *stream) {
  if (!c) return fail("null context");
  if (ncell <= 0) return 0;
  if (ncell > c->wstride) return fail("ncell=%lld exceeds max_cells", (long long)ncell);
  if (c->mech == MECH_CARBON && which != 0 && which != 4)
    return fail("the carbon mechanism has no sparse Jacobian (forward Euler)");
  CompArgs a{};
  a.ncell = ncell;
  a.in = in;
  a.istride = ncell;
  a.rconst = rct;
  a.rstride = ncell;
  a.out = out;
  a.ostride = ncell;
  a.work = c->work;
  a.wstride = c->wstride;
  a.ier = ier;
  launch_comp_mech(c->mech, which, a, (hipStream_t)stream);
  HIP_OK(hipGetLastError());
  return 0;
}

extern "C" int gckpp_fun_dev(gckpp_ctx *c, int64_t ncell, const double *conc, const double *rct,
                             double *vdot, void *stream) {
  return comp(c, 0, ncell, conc, rct, vdot, nullptr, stream);
}
extern "C" int gckpp_jac_dev(gckpp_ctx *c, int64_t ncell, const double *conc, const double *rct,
                             double *jvs, void *stream) {
  return comp(c, 1, ncell, conc, rct, jvs, nullptr, stream);
}
extern "C" int gckpp_decomp_dev(gckpp_ctx *c, int64_t ncell, double *jvs, int32_t *ier,
                                void *stream) {
  return comp(c, 2, ncell, nullptr, nullptr, jvs, ier, stream);
}
extern "C" int gckpp_solve_dev(gckpp_ctx *c, int64_t ncell, const double *lu, double *x,
                               void *stream) {
  return comp(c, 3, ncell, lu, nullptr, x, nullptr, stream);
}
extern "C" int gckpp_aout_dev(gckpp_ctx *c, int64_t ncell, const double *conc, const double *rct,
                              double *aout, void *stream) {
  return comp(c, 4, ncell, conc, rct, aout, nullptr, stream);
}

extern "C" int gckpp_rates_dev(gckpp_ctx *c, int64_t ncell, const double *met,
                               const double *photol, const double *het, int64_t het_stride,
                               const double *conc, double *rconst, void *stream) {
  if (!c) return fail("gckpp_rates: null context");
  if (ncell <= 0) return ncell == 0 ? 0 : fail("gckpp_rates: negative ncell");
  if (!met || !photol || !het || !conc || !rconst) return fail("gckpp_rates: null argument");
  if (het_stride != 0 && het_stride != ncell)
    return fail("gckpp_rates: het_stride must be 0 or ncell");
  if (!rates_available(c->mech))
    return fail("gckpp_rates: no rate program compiled for mechanism %d", c->mech);
  RatesArgs a{};
  a.ncell = ncell;
  a.met = met;
  a.photol = photol;
  a.het = het;
  a.het_stride = het_stride;
  a.conc = conc;
  a.rconst = rconst;
  a.consts = c->consts;
  launch_rates_mech(c->mech, a, (hipStream_t)stream);
  HIP_OK(hipGetLastError());
  return 0;
}

// C = max(C,0) for every species (fullchem_mod.F90:1329-1350), Hnew from
// RSTATUS(3), and one partial count of twice-failed cells per block.
__global__ __launch_bounds__(256) void k_step_finish(int64_t n, int nspec, int clip, double *conc,
                                                     const double *rst, double *hnew,
                                                     const int32_t *ierr, int32_t *partial) {
  __shared__ int32_t red[256];
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int bad = 0;
  if (i < n) {
    if (clip)
      for (int k = 0; k < nspec; ++k) {
        double v = conc[(int64_t)k * n + i];
        conc[(int64_t)k * n + i] = v > 0.0 ? v : 0.0;
      }
    if (hnew) hnew[i] = rst[2 * n + i];
    bad = ierr[i] < 0;
  }
  red[threadIdx.x] = bad;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Sums the per-block partial counts with one workgroup (no atomics).
__global__ __launch_bounds__(256) void k_sum_partials(int64_t nb, const int32_t *partial,
                                                      int32_t *out) {
  __shared__ int32_t red[256];
  int32_t acc = 0;
  for (int64_t b = threadIdx.x; b < nb; b += 256) acc += partial[b];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// The step of gckpp_fullchem_step_dev, queued on the stream; the number of
// cells that failed twice goes to the device counter c->nfail (not an error
// here: gckpp_step_failures reads it).
static int step_impl(gckpp_ctx *c, int64_t ncell, double dt, const int32_t icntrl[20],
                     const double rcntrl[20], const double *met, const double *photol,
                     const double *het, int64_t het_stride, const double *hstart, double *conc,
                     double *hnew, int32_t *istatus, int32_t *ierr, void *stream) {
  if (!c) return fail("gckpp_fullchem_step: null context");
  if (ncell <= 0) return ncell == 0 ? 0 : fail("gckpp_fullchem_step: negative ncell");
  if (ncell > c->wstride)
    return fail("gckpp_fullchem_step: ncell=%lld exceeds max_cells=%lld", (long long)ncell,
                (long long)c->max_cells);
  if (!ierr) return fail("gckpp_fullchem_step: ierr is required");
  hipStream_t s = (hipStream_t)stream;
  double *rct = c->step_buf;
  double *cbak = rct + (size_t)c->info.nreact * ncell;
  double *rst = cbak + (size_t)c->info.nspec * ncell;
  // Update_RCONST (fullchem_mod.F90:953)
  HIP_OK(hipEventRecord(c->ev[0], s));
  int rc = gckpp_rates_dev(c, ncell, met, photol, het, het_stride, conc, rct, stream);
  if (rc) return rc;
  HIP_OK(hipEventRecord(c->ev[1], s));
  if (c->solver == GCKPP_SOLVER_PERSISTENT) {
    // one launch: Integrate, the retry of failed cells from their original
    // C with Hstart = 0, the clip and KPPHvalue = RSTATUS(3), per cell
    RosParams p;
    char why[256] = "";
    int code = resolve_options(c, 0.0, dt, icntrl, rcntrl, &p, why, sizeof why);
    if (code == -100) return fail("gckpp_fullchem_step: %s", why);
    HIP_OK(hipEventRecord(c->ev[2], s));
    if (code < 0) {
      c->ho_timed = false;
      // every cell fails the same option check (twice, as the retry would):
      // IERR, zero ISTATUS and RSTATUS, then the clip and Hnew = RSTATUS(3)
      // below, as in the plain step
      hipLaunchKernelGGL(k_fill_status, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, s,
                         ncell, ierr, code, istatus, rst);
    } else {
      PersistArgs pa{};
      fill_persist_queue(c, pa.q[0], ncell, p, conc, rct, hstart, hnew, istatus, rst, ierr);
      pa.q[0].order = c->order;
      pa.q[0].hlog = c->hlog;
      pa.q[0].hlog_n = c->hlog_n;
      pa.qend[0] = pa.qend[1] = pa.qend[2] = ncell;
      pa.work = c->work;
      pa.slot_fields = launch_slot_fields(c->mech, c->pvariant);
      pa.counter = c->counter;
      int rcl = launch_persist_queue(c, pa, ncell, s, "gckpp_fullchem_step");
      if (rcl) return rcl;
    }
    HIP_OK(hipEventRecord(c->ev[3], s));
    int64_t nb = (ncell + 255) / 256;
    int32_t *partial = (int32_t *)cbak;
    const bool opt_err = code < 0;  // else the persistent kernel clipped and wrote Hnew
    hipLaunchKernelGGL(k_step_finish, dim3((unsigned)nb), dim3(256), 0, s, ncell,
                       c->info.nspec, opt_err && c->mech != MECH_CARBON ? 1 : 0, conc,
                       opt_err ? (const double *)rst : nullptr, opt_err ? hnew : nullptr, ierr,
                       partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, nb, partial, c->nfail);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(c->ev[4], s));
    return 0;
  }
  // C_before_integrate (:1026)
  HIP_OK(hipMemcpyAsync(cbak, conc, sizeof(double) * c->info.nspec * ncell,
                        hipMemcpyDeviceToDevice, s));
  // Integrate(0, DT, ...) (:1033)
  HIP_OK(hipEventRecord(c->ev[2], s));
  rc = integrate_impl(c, ncell, 0.0, dt, icntrl, rcntrl, hstart, conc, rct, istatus, rst, ierr,
                      nullptr, nullptr, stream);
  if (rc) return rc;
  HIP_OK(hipEventRecord(c->ev[3], s));
  // retry of failed cells: RCNTRL(3)=0, C restored, rates of the restored C
  // (identical to the first Update_RCONST) (:1138-1166)
  double rc2[20];
  for (int k = 0; k < 20; ++k) rc2[k] = rcntrl ? rcntrl[k] : 0.0;
  rc2[2] = 0.0;
  rc = integrate_impl(c, ncell, 0.0, dt, icntrl, rc2, nullptr, conc, rct, istatus, rst, ierr,
                      ierr, cbak, stream);
  if (rc) return rc;
  int64_t nb = (ncell + 255) / 256;
  int32_t *partial = (int32_t *)(rst + 3 * (size_t)ncell);  // rows 3.. of rstatus are free
  hipLaunchKernelGGL(k_step_finish, dim3((unsigned)nb), dim3(256), 0, s, ncell,
                     c->info.nspec, c->mech != MECH_CARBON, conc, rst, hnew, ierr, partial);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, nb, partial, c->nfail);
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventRecord(c->ev[4], s));
  return 0;
}

extern "C" int gckpp_fullchem_step_dev(gckpp_ctx *c, int64_t ncell, double dt,
                                       const int32_t icntrl[20], const double rcntrl[20],
                                       const double *met, const double *photol,
                                       const double *het, int64_t het_stride,
                                       const double *hstart, double *conc, double *hnew,
                                       int32_t *istatus, int32_t *ierr, void *stream) {
  return step_impl(c, ncell, dt, icntrl, rcntrl, met, photol, het, het_stride, hstart, conc,
                   hnew, istatus, ierr, stream);
}

extern "C" int gckpp_step_failures(gckpp_ctx *c, int64_t *nfail) {
  if (!c || !nfail) return fail("gckpp_step_failures: null argument");
  *nfail = 0;
  int32_t v = 0;
  HIP_OK(hipDeviceSynchronize());
  HIP_OK(hipMemcpy(&v, c->nfail, sizeof v, hipMemcpyDeviceToHost));
  *nfail = v;
  return 0;
}

// ---------------------------------------------------------------------------
// Host-pointer Update_RCONST and step: the caller's arrays stay in host
// memory ([field][ncell], Fortran A(ncell, NFIELD)); the cells go through the
// device in chunks of at most max_cells (a chunk of a [field][ncell] array is
// a pitched 2-D block, one hipMemcpy2D per array), staged in buffers the
// context allocates at the first such call.
static int host_stage(gckpp_ctx *c, const char *who) {
  if (c->hs_d) return 0;
  const size_t n = (size_t)c->wstride;
  const size_t nd = (size_t)(NMET + NPHOTOL + NHET + c->info.nspec + c->info.nreact + 22) * n;
  const size_t ni = (size_t)21 * n;
  hipError_t e = hipMalloc(&c->hs_d, sizeof(double) * nd);
  if (e == hipSuccess) e = hipMalloc(&c->hs_i, sizeof(int32_t) * ni);
  if (e != hipSuccess) {
    if (c->hs_d) (void)hipFree(c->hs_d);
    c->hs_d = nullptr;
    c->hs_i = nullptr;
    return fail("%s: cannot allocate %.2f GB of staging buffers: %s", who,
                (sizeof(double) * nd + sizeof(int32_t) * ni) / 1e9, hipGetErrorString(e));
  }
  return 0;
}
// rows x n elements of a [rows][ncell] host array from cell c0, to or from a
// [rows][n] device block.
template <class T>
static hipError_t stage_in(T *d, const T *h, int64_t rows, int64_t n, int64_t ncell, int64_t c0) {
  return hipMemcpy2D(d, sizeof(T) * n, h + c0, sizeof(T) * ncell, sizeof(T) * n, rows,
                     hipMemcpyHostToDevice);
}
template <class T>
static hipError_t stage_out(T *h, const T *d, int64_t rows, int64_t n, int64_t ncell, int64_t c0) {
  return hipMemcpy2D(h + c0, sizeof(T) * ncell, d, sizeof(T) * n, sizeof(T) * n, rows,
                     hipMemcpyDeviceToHost);
}
struct Stage {
  double *met, *photol, *het, *conc, *rconst, *hstart, *hnew, *rstatus;
  int32_t *istatus, *ierr;
};
static Stage stage_views(const gckpp_ctx *c, int64_t n) {
  Stage v;
  v.met = c->hs_d;
  v.photol = v.met + (size_t)NMET * n;
  v.het = v.photol + (size_t)NPHOTOL * n;
  v.conc = v.het + (size_t)NHET * n;
  v.rconst = v.conc + (size_t)c->info.nspec * n;
  v.hstart = v.rconst + (size_t)c->info.nreact * n;
  v.hnew = v.hstart + n;
  v.rstatus = v.hnew + n;
  v.istatus = c->hs_i;
  v.ierr = v.istatus + (size_t)20 * n;
  return v;
}
// The per-cell inputs of one chunk.
static hipError_t stage_inputs(const Stage &v, int64_t n, int64_t ncell, int64_t c0,
                               const double *met, const double *photol, const double *het,
                               int64_t het_stride, const double *conc, int nspec) {
  hipError_t e = stage_in(v.met, met, NMET, n, ncell, c0);
  if (e == hipSuccess) e = stage_in(v.photol, photol, NPHOTOL, n, ncell, c0);
  if (e == hipSuccess)
    e = het_stride ? stage_in(v.het, het, NHET, n, ncell, c0)
                   : hipMemcpy(v.het, het, sizeof(double) * NHET, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = stage_in(v.conc, conc, nspec, n, ncell, c0);
  return e;
}

extern "C" int gckpp_rates(gckpp_ctx *c, int64_t ncell, const double *met, const double *photol,
                           const double *het, int64_t het_stride, const double *conc,
                           double *rconst) {
  if (!c) return fail("gckpp_rates: null context");
  if (ncell <= 0) return ncell == 0 ? 0 : fail("gckpp_rates: negative ncell");
  if (!met || !photol || !het || !conc || !rconst) return fail("gckpp_rates: null argument");
  if (het_stride != 0 && het_stride != ncell)
    return fail("gckpp_rates: het_stride must be 0 or ncell");
  if (host_stage(c, "gckpp_rates")) return -1;
  for (int64_t c0 = 0; c0 < ncell; c0 += c->wstride) {
    const int64_t n = ncell - c0 < c->wstride ? ncell - c0 : c->wstride;
    const Stage v = stage_views(c, n);
    HIP_OK(stage_inputs(v, n, ncell, c0, met, photol, het, het_stride, conc, c->info.nspec));
    int rc = gckpp_rates_dev(c, n, v.met, v.photol, v.het, het_stride ? n : 0, v.conc, v.rconst,
                             nullptr);
    if (rc) return rc;
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(stage_out(rconst, v.rconst, c->info.nreact, n, ncell, c0));
  }
  return 0;
}

// ---------------------------------------------------------------------------
// The model side of the step: Map_KppSpc (fullchem_mod.F90:817-821), the
// unit conversions of Convert_Spc_Units (GeosUtil/unitconv_mod.F90:1287-1391,
// 1407-1515) and the prod/loss dummies zeroed before the rates (:942-946).
extern "C" int gckpp_set_species_map(gckpp_ctx *c, int32_t nmodel, const int32_t *map,
                                     const double *mw_g, const int32_t *zero, int32_t alk_equiv) {
  if (!c || !map || !mw_g) return fail("gckpp_set_species_map: null argument");
  if (nmodel <= 0) return fail("gckpp_set_species_map: nmodel must be positive");
  const int ns = c->info.nspec;
  std::vector<double> cin(ns, 0.0), cout(ns, 0.0), eq(ns, 1.0);
  if (alk_equiv && c->mech != MECH_FULLCHEM)
    return fail("gckpp_set_species_map: the alkalinity conversion is fullchem's (mechanism %d)",
                c->mech);
  std::vector<int32_t> zz(ns, 0);
  for (int k = 0; k < ns; ++k) {
    const int s = map[k];
    if (s < -1 || s >= nmodel)
      return fail("gckpp_set_species_map: KPP species %d maps to model species %d of %d", k, s, nmodel);
    if (s < 0) continue;
    if (!(mw_g[s] > 0.0))
      return fail("gckpp_set_species_map: model species %d has molecular weight %g", s, mw_g[s]);
    const double mw_kg = mw_g[s] * 1.0e-3;
    cin[k] = (kAvogadro / mw_kg) / 1.0e+6;
    cout[k] = 1.0e+6 / (kAvogadro / mw_kg);
    zz[k] = zero ? (zero[k] != 0) : 0;
    // fullchem_ConvertAlkToEquiv (KPP/fullchem/fullchem_SulfurChemFuncs.F90:
    // 67-78): C * (MW * 7e-5) for the sea-salt alkalinity tracers; KPP's MW
    // is the model species' molecular weight
    if (alk_equiv && (k == species_fullchem::SALAAL || k == species_fullchem::SALCAL))
      eq[k] = mw_g[s] * 7.0e-5;
  }
  if (!c->sm_map) {
    HIP_OK(hipMalloc(&c->sm_map, sizeof(int32_t) * ns));
    HIP_OK(hipMalloc(&c->sm_zero, sizeof(int32_t) * ns));
    HIP_OK(hipMalloc(&c->sm_cin, sizeof(double) * ns));
    HIP_OK(hipMalloc(&c->sm_cout, sizeof(double) * ns));
    HIP_OK(hipMalloc(&c->sm_eq, sizeof(double) * ns));
  }
  HIP_OK(hipMemcpy(c->sm_map, map, sizeof(int32_t) * ns, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(c->sm_zero, zz.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(c->sm_cin, cin.data(), sizeof(double) * ns, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(c->sm_cout, cout.data(), sizeof(double) * ns, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(c->sm_eq, eq.data(), sizeof(double) * ns, hipMemcpyHostToDevice));
  c->sm_nmodel = nmodel;
  return 0;
}

extern "C" int gckpp_gather_dev(gckpp_ctx *c, int64_t ncell, const double *species,
                                const double *airden, double *conc, void *stream) {
  if (!c) return fail("gckpp_gather_dev: null context");
  if (ncell < 0) return fail("gckpp_gather_dev: negative ncell");
  if (ncell == 0) return 0;
  if (!c->sm_map) return fail("gckpp_gather_dev: no species map (gckpp_set_species_map)");
  if (!species || !airden || !conc) return fail("gckpp_gather_dev: null argument");
  launch_gather(ncell, c->info.nspec, c->sm_map, c->sm_cin, c->sm_eq, c->sm_zero, species, airden,
                conc,
                (hipStream_t)stream);
  HIP_OK(hipGetLastError());
  return 0;
}

extern "C" int gckpp_scatter_dev(gckpp_ctx *c, int64_t ncell, const double *conc,
                                 const double *airden, double *species, void *stream) {
  if (!c) return fail("gckpp_scatter_dev: null context");
  if (ncell < 0) return fail("gckpp_scatter_dev: negative ncell");
  if (ncell == 0) return 0;
  if (!c->sm_map) return fail("gckpp_scatter_dev: no species map (gckpp_set_species_map)");
  if (!species || !airden || !conc) return fail("gckpp_scatter_dev: null argument");
  launch_scatter(ncell, c->info.nspec, c->sm_map, c->sm_cout, c->sm_eq, conc, airden, species,
                 (hipStream_t)stream);
  HIP_OK(hipGetLastError());
  return 0;
}

extern "C" int gckpp_set_state_het_dev(gckpp_ctx *c, int64_t ncell, const double *hetin,
                                       const double *conc, int32_t naerotype, double gamma_ho2,
                                       int32_t lpscchem, int32_t turn_off_het, double *het,
                                       void *stream) {
  if (!c) return fail("gckpp_set_state_het_dev: null context");
  if (c->mech != MECH_FULLCHEM)
    return fail("gckpp_set_state_het_dev: the HetState record is fullchem's (context mechanism %d)",
                c->mech);
  if (ncell < 0) return fail("gckpp_set_state_het_dev: negative ncell");
  if (naerotype < 0 || naerotype > 14)
    return fail("gckpp_set_state_het_dev: nAeroType %d outside 0..14", naerotype);
  if (ncell == 0) return 0;
  if (!hetin || !conc || !het) return fail("gckpp_set_state_het_dev: null argument");
  launch_set_state_het(ncell, hetin, conc, het, naerotype, gamma_ho2, lpscchem, turn_off_het,
                       (hipStream_t)stream);
  HIP_OK(hipGetLastError());
  return 0;
}

extern "C" int gckpp_sulfur_aq_dev(gckpp_ctx *c, int64_t ncell, const double *hetin,
                                   const double *conc, double *met, void *stream) {
  if (!c) return fail("gckpp_sulfur_aq_dev: null context");
  if (c->mech != MECH_FULLCHEM)
    return fail("gckpp_sulfur_aq_dev: K_MT is fullchem's (context mechanism %d)", c->mech);
  if (ncell < 0) return fail("gckpp_sulfur_aq_dev: negative ncell");
  if (ncell == 0) return 0;
  if (!hetin || !conc || !met) return fail("gckpp_sulfur_aq_dev: null argument");
  launch_sulfur_aq(ncell, hetin, conc, c->consts, met, (hipStream_t)stream);
  HIP_OK(hipGetLastError());
  return 0;
}

// Host-pointer producers, for a host that gathers the model's fields into
// host arrays (any ncell, in chunks of at most max_cells staged through the
// device, synchronous): the model-field record hetin rides in the staging
// rows of PHOTOL (NHIN <= NPHOTOL).
static_assert(NHIN <= NPHOTOL, "hetin staging");
static int set_het_args(gckpp_ctx *c, int64_t ncell, int32_t naerotype, const char *who) {
  if (!c) return fail("%s: null context", who);
  if (c->mech != MECH_FULLCHEM)
    return fail("%s: the HetState record is fullchem's (context mechanism %d)", who, c->mech);
  if (ncell < 0) return fail("%s: negative ncell", who);
  if (naerotype < 0 || naerotype > 14) return fail("%s: nAeroType %d outside 0..14", who, naerotype);
  return 0;
}

extern "C" int gckpp_set_state_het(gckpp_ctx *c, int64_t ncell, const double *hetin,
                                   const double *conc, int32_t naerotype, double gamma_ho2,
                                   int32_t lpscchem, int32_t turn_off_het, double *het) {
  if (set_het_args(c, ncell, naerotype, "gckpp_set_state_het")) return -1;
  if (ncell == 0) return 0;
  if (!hetin || !conc || !het) return fail("gckpp_set_state_het: null argument");
  if (host_stage(c, "gckpp_set_state_het")) return -1;
  for (int64_t c0 = 0; c0 < ncell; c0 += c->wstride) {
    const int64_t n = ncell - c0 < c->wstride ? ncell - c0 : c->wstride;
    const Stage v = stage_views(c, n);
    HIP_OK(stage_in(v.photol, hetin, NHIN, n, ncell, c0));
    HIP_OK(stage_in(v.conc, conc, c->info.nspec, n, ncell, c0));
    launch_set_state_het(n, v.photol, v.conc, v.het, naerotype, gamma_ho2, lpscchem, turn_off_het,
                         nullptr);
    HIP_OK(hipGetLastError());
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(stage_out(het, v.het, NHET, n, ncell, c0));
  }
  return 0;
}

extern "C" int gckpp_sulfur_aq(gckpp_ctx *c, int64_t ncell, const double *hetin,
                               const double *conc, double *met) {
  if (set_het_args(c, ncell, 0, "gckpp_sulfur_aq")) return -1;
  if (ncell == 0) return 0;
  if (!hetin || !conc || !met) return fail("gckpp_sulfur_aq: null argument");
  if (host_stage(c, "gckpp_sulfur_aq")) return -1;
  for (int64_t c0 = 0; c0 < ncell; c0 += c->wstride) {
    const int64_t n = ncell - c0 < c->wstride ? ncell - c0 : c->wstride;
    const Stage v = stage_views(c, n);
    HIP_OK(stage_in(v.photol, hetin, NHIN, n, ncell, c0));
    HIP_OK(stage_in(v.conc, conc, c->info.nspec, n, ncell, c0));
    HIP_OK(stage_in(v.met, met, NMET, n, ncell, c0));
    launch_sulfur_aq(n, v.photol, v.conc, c->consts, v.met, nullptr);
    HIP_OK(hipGetLastError());
    HIP_OK(hipDeviceSynchronize());
    // only the K_MT rows change
    HIP_OK(stage_out(met + (int64_t)MET_K_MT * ncell, v.met + (int64_t)MET_K_MT * n, 6, n, ncell,
                     c0));
  }
  return 0;
}

// Host-pointer Aout (Fun's reaction rates): per chunk, conc and rconst in,
// the rates computed in place of the staged rconst, copied out.
extern "C" int gckpp_aout(gckpp_ctx *c, int64_t ncell, const double *conc, const double *rconst,
                          double *aout) {
  if (!c) return fail("gckpp_aout: null context");
  if (ncell <= 0) return ncell == 0 ? 0 : fail("gckpp_aout: negative ncell");
  if (!conc || !rconst || !aout) return fail("gckpp_aout: null argument");
  if (host_stage(c, "gckpp_aout")) return -1;
  for (int64_t c0 = 0; c0 < ncell; c0 += c->wstride) {
    const int64_t n = ncell - c0 < c->wstride ? ncell - c0 : c->wstride;
    const Stage v = stage_views(c, n);
    HIP_OK(stage_in(v.conc, conc, c->info.nspec, n, ncell, c0));
    HIP_OK(stage_in(v.rconst, rconst, c->info.nreact, n, ncell, c0));
    int rc = comp(c, 4, n, v.conc, v.rconst, v.rconst, nullptr, nullptr);
    if (rc) return rc;
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(stage_out(aout, v.rconst, c->info.nreact, n, ncell, c0));
  }
  return 0;
}

// Host-pointer Integrate: per chunk of at most max_cells cells, copies in,
// solves, copies out (synchronous), through the context's staging buffers
// (allocated at its first host-pointer call and reused by later ones).  A
// cell order or step log set for device calls does not apply here.
extern "C" int gckpp_integrate(gckpp_ctx *c, int64_t ncell, double tin, double tout,
                               const int32_t icntrl[20], const double rcntrl[20],
                               const double *hstart, double *conc, const double *rconst,
                               int32_t *istatus, double *rstatus, int32_t *ierr) {
  if (!c) return fail("gckpp_integrate: null context");
  if (ncell <= 0) return ncell == 0 ? 0 : fail("gckpp_integrate: negative ncell");
  if (!conc || !rconst) return fail("gckpp_integrate: null conc or rconst");
  if (host_stage(c, "gckpp_integrate")) return -1;
  const int32_t *order = c->order;
  float *hlog = c->hlog;
  c->order = nullptr;
  c->hlog = nullptr;
  int rc = 0;
  for (int64_t c0 = 0; c0 < ncell && rc == 0; c0 += c->wstride) {
    const int64_t n = ncell - c0 < c->wstride ? ncell - c0 : c->wstride;
    const Stage v = stage_views(c, n);
    hipError_t e = stage_in(v.conc, conc, c->info.nspec, n, ncell, c0);
    if (e == hipSuccess) e = stage_in(v.rconst, rconst, c->info.nreact, n, ncell, c0);
    if (e == hipSuccess)
      e = hstart ? stage_in(v.hstart, hstart, 1, n, ncell, c0)
                 : hipMemset(v.hstart, 0, sizeof(double) * n);
    if (e != hipSuccess) {
      rc = fail("gckpp_integrate: %s", hipGetErrorString(e));
      break;
    }
    rc = gckpp_integrate_dev(c, n, tin, tout, icntrl, rcntrl, v.hstart, v.conc, v.rconst,
                             v.istatus, v.rstatus, v.ierr, nullptr);
    if (rc) break;
    e = hipDeviceSynchronize();
    if (e == hipSuccess) e = stage_out(conc, v.conc, c->info.nspec, n, ncell, c0);
    if (e == hipSuccess && istatus) e = stage_out(istatus, v.istatus, 20, n, ncell, c0);
    if (e == hipSuccess && rstatus) e = stage_out(rstatus, v.rstatus, 20, n, ncell, c0);
    if (e == hipSuccess && ierr) e = stage_out(ierr, v.ierr, 1, n, ncell, c0);
    if (e != hipSuccess) rc = fail("gckpp_integrate: %s", hipGetErrorString(e));
  }
  c->order = order;
  c->hlog = hlog;
  return rc;
}

static int host_step_chunks(gckpp_ctx *c, int64_t ncell, double dt, const int32_t icntrl[20],
                            const double rcntrl[20], const double *met, const double *photol,
                            const double *het, int64_t het_stride, const double *hstart,
                            double *conc, double *hnew, int32_t *istatus, int32_t *ierr) {
  int64_t failed = 0;
  for (int64_t c0 = 0; c0 < ncell; c0 += c->wstride) {
    const int64_t n = ncell - c0 < c->wstride ? ncell - c0 : c->wstride;
    const Stage v = stage_views(c, n);
    HIP_OK(stage_inputs(v, n, ncell, c0, met, photol, het, het_stride, conc, c->info.nspec));
    if (hstart) HIP_OK(stage_in(v.hstart, hstart, 1, n, ncell, c0));
    else HIP_OK(hipMemset(v.hstart, 0, sizeof(double) * n));
    int rc = step_impl(c, n, dt, icntrl, rcntrl, v.met, v.photol, v.het, het_stride ? n : 0,
                       v.hstart, v.conc, v.hnew, v.istatus, v.ierr, nullptr);
    if (rc) return rc;
    int32_t nfail = 0;
    HIP_OK(hipMemcpy(&nfail, c->nfail, sizeof nfail, hipMemcpyDeviceToHost));
    failed += nfail;
    HIP_OK(stage_out(conc, v.conc, c->info.nspec, n, ncell, c0));
    HIP_OK(stage_out(ierr, v.ierr, 1, n, ncell, c0));
    if (hnew) HIP_OK(stage_out(hnew, v.hnew, 1, n, ncell, c0));
    if (istatus) HIP_OK(stage_out(istatus, v.istatus, 20, n, ncell, c0));
  }
  if (failed > 0) return fail("gckpp_fullchem_step: %lld cell(s) failed twice", (long long)failed);
  return 0;
}

extern "C" int gckpp_fullchem_step(gckpp_ctx *c, int64_t ncell, double dt,
                                   const int32_t icntrl[20], const double rcntrl[20],
                                   const double *met, const double *photol, const double *het,
                                   int64_t het_stride, const double *hstart, double *conc,
                                   double *hnew, int32_t *istatus, int32_t *ierr) {
  if (!c) return fail("gckpp_fullchem_step: null context");
  if (ncell <= 0) return ncell == 0 ? 0 : fail("gckpp_fullchem_step: negative ncell");
  if (!met || !photol || !het || !conc || !ierr)
    return fail("gckpp_fullchem_step: null argument");
  if (het_stride != 0 && het_stride != ncell)
    return fail("gckpp_fullchem_step: het_stride must be 0 or ncell");
  if (host_stage(c, "gckpp_fullchem_step")) return -1;
  // a cell order or step log set for device calls indexes their batch, not
  // a chunk of this one: neither applies here
  const int32_t *order = c->order;
  float *hlog = c->hlog;
  c->order = nullptr;
  c->hlog = nullptr;
  const int rc = host_step_chunks(c, ncell, dt, icntrl, rcntrl, met, photol, het, het_stride,
                                  hstart, conc, hnew, istatus, ierr);
  c->order = order;
  c->hlog = hlog;
  return rc;
}

// Device time of the parts of the last gckpp_fullchem_step_dev call, from HIP
// events on its stream (the role of the reference's per-cell KppTime
// diagnostic, fullchem_mod.F90:596-598,1298-1301, for the whole batch):
// ms[0] Update_RCONST, ms[1] Integrate, ms[2] retry + clip, ms[3] total.
extern "C" int gckpp_step_timings(gckpp_ctx *c, float ms[4]) {
  if (!c || !ms) return fail("gckpp_step_timings: null argument");
  HIP_OK(hipEventElapsedTime(&ms[0], c->ev[0], c->ev[1]));
  HIP_OK(hipEventElapsedTime(&ms[1], c->ev[2], c->ev[3]));
  HIP_OK(hipEventElapsedTime(&ms[2], c->ev[3], c->ev[4]));
  HIP_OK(hipEventElapsedTime(&ms[3], c->ev[0], c->ev[4]));
  return 0;
}

extern "C" int gckpp_set_solver(gckpp_ctx *c, int solver) {
  if (!c) return fail("gckpp_set_solver: null context");
  if (solver != GCKPP_SOLVER_PLAIN && solver != GCKPP_SOLVER_PERSISTENT)
    return fail("gckpp_set_solver: unknown solver %d", solver);
  c->solver = solver;
  return 0;
}

// ---------------------------------------------------------------------------
// Persistent-kernel configuration and its autotuner.
extern "C" int gckpp_set_persistent_config(gckpp_ctx *c, int variant, int waves_per_group) {
  if (!c) return fail("gckpp_set_persistent_config: null context");
  if (variant < 0 || variant >= PERSIST_VARIANTS)
    return fail("gckpp_set_persistent_config: variant %d not in [0,%d)", variant, PERSIST_VARIANTS);
  if (waves_per_group != 1 && waves_per_group != 2 && waves_per_group != 4)
    return fail("gckpp_set_persistent_config: waves per workgroup must be 1, 2 or 4");
  if (set_pconfig(c, variant, waves_per_group) != 0)
    return fail("gckpp_set_persistent_config: variant %d with %d waves per workgroup does not fit",
                variant, waves_per_group);
  return 0;
}

extern "C" int gckpp_set_persistent_waves(gckpp_ctx *c, int64_t max_waves) {
  if (!c) return fail("gckpp_set_persistent_waves: null context");
  if (max_waves < 0) return fail("gckpp_set_persistent_waves: negative wave count");
  c->pwaves_max = max_waves;
  if (set_pconfig(c, c->pvariant, c->pwpb) != 0)
    return fail("gckpp_set_persistent_waves: configuration does not fit");
  return 0;
}

extern "C" int gckpp_set_cell_order(gckpp_ctx *c, const int32_t *order) {
  if (!c) return fail("gckpp_set_cell_order: null context");
  c->order = order;
  return 0;
}

extern "C" int gckpp_set_step_log(gckpp_ctx *c, float *hlog, int nattempt) {
  if (!c) return fail("gckpp_set_step_log: null context");
  if (hlog && nattempt <= 0) return fail("gckpp_set_step_log: nattempt must be positive");
  c->hlog = hlog;
  c->hlog_n = hlog ? nattempt : 0;
  return 0;
}

extern "C" int gckpp_get_persistent_config(gckpp_ctx *c, int32_t out[7]) {
  if (!c || !out) return fail("gckpp_get_persistent_config: null argument");
  int U = 0, W = 0;
  persist_variant_info(c->pvariant, &U, &W);
  out[0] = c->pvariant;
  out[1] = c->pwpb;
  out[2] = (int32_t)c->pwaves;
  out[3] = U;
  out[4] = W;
  out[5] = persist_variant_seg(c->pvariant);
  out[6] = persist_variant_cache(c->pvariant);
  return 0;
}

extern "C" int gckpp_persistent_variants(void) { return PERSIST_VARIANTS; }

// Times the persistent solver on the caller's cells for every variant and
// workgroup size, keeps the fastest, and leaves conc untouched (each trial
// integrates a copy).  ms[v*3 + g] receives the time of variant v with
// 2^g waves per workgroup (-1 where it does not fit).
extern "C" int gckpp_autotune_dev(gckpp_ctx *c, int64_t ncell, double tin, double tout,
                                  const int32_t icntrl[20], const double rcntrl[20],
                                  const double *hstart, const double *conc,
                                  const double *rconst, float *ms, void *stream) {
  if (!c) return fail("gckpp_autotune: null context");
  if (ncell <= 0 || ncell > c->wstride)
    return fail("gckpp_autotune: ncell must be in [1, max_cells]");
  if (!conc || !rconst) return fail("gckpp_autotune: null conc or rconst");
  hipStream_t s = (hipStream_t)stream;
  double *scratch = c->step_buf + (size_t)c->info.nreact * ncell;  // the conc backup rows
  const size_t cbytes = sizeof(double) * c->info.nspec * ncell;
  int32_t *ierr = (int32_t *)(scratch + (size_t)c->info.nspec * ncell);
  int keep_solver = c->solver, best_v = c->pvariant, best_g = c->pwpb;
  float best = -1.0f;
  c->solver = GCKPP_SOLVER_PERSISTENT;
  int rc = 0;
  for (int v = 0; v < PERSIST_VARIANTS && rc == 0; ++v) {
    for (int g = 0; g < 3 && rc == 0; ++g) {
      float t = -1.0f;
      if (set_pconfig(c, v, 1 << g) == 0) {
        // one untimed run of the first configuration warms caches and code
        for (int rep = (v == 0 && g == 0) ? 0 : 1; rep < 2 && rc == 0; ++rep) {
          HIP_OK(hipMemcpyAsync(scratch, conc, cbytes, hipMemcpyDeviceToDevice, s));
          HIP_OK(hipEventRecord(c->ev[0], s));
          rc = integrate_impl(c, ncell, tin, tout, icntrl, rcntrl, hstart, scratch, rconst,
                              nullptr, nullptr, ierr, nullptr, nullptr, stream);
          HIP_OK(hipEventRecord(c->ev[1], s));
          HIP_OK(hipEventSynchronize(c->ev[1]));
          HIP_OK(hipEventElapsedTime(&t, c->ev[0], c->ev[1]));
        }
        if (rc == 0 && (best < 0.0f || t < best)) {
          best = t;
          best_v = v;
          best_g = 1 << g;
        }
      }
      if (ms) ms[v * 3 + g] = t;
    }
  }
  c->solver = keep_solver;
  if (set_pconfig(c, best_v, best_g) != 0) set_pconfig(c, kDefaultPersistVariant, 1);
  return rc;
}

// ---------------------------------------------------------------------------
// Mixed-mechanism batch: cells of fullchem, Hg and carbon solved by one
// persistent launch per chemistry step (queues in that order, the heaviest
// first), with the whole step -- three Update_RCONST launches, the counter
// reset and the persistent solve with retry, clip and Hstart update --
// capturable as one graph and replayed step after step.
struct gckpp_batch {
  gckpp_ctx *ctx[3];      // fullchem, Hg, carbon (null: absent)
  int variant, wpb;
  int64_t nwaves, slot_fields;
  double *work;           // [cap_waves][slot_fields][64]
  int64_t cap_waves;
  int64_t max_waves;      // gckpp_batch_set_waves (0: every resident wave)
  unsigned int *counter;
  hipStream_t cap_stream; // stream the graph is captured on
  hipGraph_t graph;
  hipGraphExec_t exec;
  // tail hand-off (gckpp_batch_set_handoff), as in a context
  int64_t ho_cap, ho_alloc;
  unsigned int *ho_ctr;   // [4]
  int32_t *ho_cell;
  double *ho_rec;         // [HO_FIELDS + max NVAR][ho_cap]
  double *hwork;
  int64_t hwaves, hw_fields;
};

static int batch_config(gckpp_batch *b, int variant, int wpb, int64_t cap_waves) {
  int64_t res = persist_resident_waves(PQ_MIXED, variant, wpb);
  if (res <= 0) return -1;
  int64_t nw = res / wpb * wpb;
  if (cap_waves > 0 && nw > cap_waves) nw = cap_waves / wpb * wpb;
  if (b->max_waves > 0 && nw > b->max_waves) nw = b->max_waves / wpb * wpb;
  if (nw <= 0) return -1;
  b->variant = variant;
  b->wpb = wpb;
  b->nwaves = nw;
  return 0;
}

extern "C" int gckpp_batch_init(gckpp_ctx *fullchem, gckpp_ctx *hg, gckpp_ctx *carbon,
                                gckpp_batch **out) {
  if (!out) return fail("gckpp_batch_init: null output pointer");
  *out = nullptr;
  gckpp_ctx *cs[3] = {fullchem, hg, carbon};
  const int want[3] = {MECH_FULLCHEM, MECH_HG, MECH_CARBON};
  for (int q = 0; q < 3; ++q)
    if (cs[q] && cs[q]->mech != want[q])
      return fail("gckpp_batch_init: context %d is not of the expected mechanism", q);
  if (!fullchem && !hg && !carbon) return fail("gckpp_batch_init: no context given");
  auto *b = new gckpp_batch();
  for (int q = 0; q < 3; ++q) b->ctx[q] = cs[q];
  b->slot_fields = persist_slot_fields(PQ_MIXED);
  if (batch_config(b, kDefaultPersistVariant, 1, 0) != 0) {
    delete b;
    return fail("gckpp_batch_init: cannot size the persistent launch");
  }
  // workspace for the largest configuration (every resident wave of 1-wave groups)
  size_t wb = sizeof(double) * (size_t)b->nwaves * 64 * (size_t)b->slot_fields;
  for (int v = 0; v < PERSIST_VARIANTS; ++v) {
    int64_t r = persist_resident_waves(PQ_MIXED, v, 1);
    size_t x = sizeof(double) * (size_t)r * 64 * (size_t)b->slot_fields;
    if (x > wb) wb = x;
  }
  hipError_t e = hipMalloc(&b->work, wb);
  if (e != hipSuccess) {
    delete b;
    return fail("gckpp_batch_init: cannot allocate %.2f GB of slots: %s", wb / 1e9,
                hipGetErrorString(e));
  }
  b->cap_waves = (int64_t)(wb / (sizeof(double) * 64 * (size_t)b->slot_fields));
  e = hipMalloc(&b->counter, sizeof(unsigned int));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    gckpp_batch_finalize(b);  // releases what was allocated
    return fail("gckpp_batch_init: %s", hipGetErrorString(e));
  }
  *out = b;
  return 0;
}

// One part of a batch step: the cells of one mechanism (see gckpp_hip.h).
static int batch_enqueue(gckpp_batch *b, const gckpp_part parts[3], double dt, hipStream_t s) {
  PersistArgs pa{};
  int64_t end = 0;
  int nq = 0;
  for (int q = 0; q < 3; ++q) {
    const gckpp_part &P = parts[q];
    gckpp_ctx *c = b->ctx[q];
    int64_t n = (c && P.ncell > 0) ? P.ncell : 0;
    if (n > 0) {
      if (n > c->wstride)
        return fail("gckpp_batch: part %d has %lld cells, more than its context's max_cells", q,
                    (long long)n);
      if (!P.conc || !P.hstart || !P.met || !P.photol || !P.het)
        return fail("gckpp_batch: part %d: null argument", q);
      RosParams p;
      char why[256] = "";
      int code = resolve_options(c, 0.0, dt, P.icntrl, P.rcntrl, &p, why, sizeof why);
      if (code == -100) return fail("gckpp_batch: part %d: %s", q, why);
      if (code < 0) return fail("gckpp_batch: part %d: options give IERR %d", q, code);
      double *rct = c->step_buf;
      int rc = gckpp_rates_dev(c, n, P.met, P.photol, P.het, P.het_stride, P.conc, rct, s);
      if (rc) return rc;
      fill_persist_queue(c, pa.q[q], n, p, P.conc, rct, P.hstart, P.hstart, P.istatus, nullptr,
                         P.ierr);
      pa.q[q].order = P.order;
      ++nq;
    }
    end += n;
    pa.qend[q] = end;
  }
  if (end == 0) return 0;
  pa.work = b->work;
  pa.slot_fields = launch_slot_fields(PQ_MIXED, b->variant);
  pa.counter = b->counter;
  const bool hand = b->ho_cap > 0 && b->ho_rec && b->variant != PERSIST_WAVE_VARIANT;
  if (hand) {
    launch_zero_u32(b->ho_ctr, 4, s);
    pa.ho = Handoff{b->ho_ctr, b->ho_ctr + 1, b->ho_ctr + 3, b->ho_cell, b->ho_rec, b->ho_cap, 0};
  }
  int64_t nw = cells_to_waves(b->variant, end);
  nw = (nw + b->wpb - 1) / b->wpb * b->wpb;
  if (nw > b->nwaves) nw = b->nwaves;
  if (launch_persist(PQ_MIXED, b->variant, b->wpb, pa, nw, s) != 0)
    return fail("gckpp_batch: persistent launch failed: %s", hipGetErrorString(hipGetLastError()));
  if (hand) {
    PersistArgs ph = pa;
    ph.ho.resume = 1;
    ph.work = b->hwork;
    ph.slot_fields = b->hw_fields;
    ph.counter = b->ho_ctr + 2;
    if (launch_persist(PQ_MIXED, PERSIST_WAVE_VARIANT, 1, ph, b->hwaves, s) != 0)
      return fail("gckpp_batch: hand-off launch failed: %s", hipGetErrorString(hipGetLastError()));
  }
  return 0;
}

// Destroys the batch's captured graph (if any); replay then reports that
// nothing is captured.  Called before device buffers the graph's kernels
// point at are released.
static int batch_drop_graph(gckpp_batch *b) {
  bool ok = true;
  if (b->exec) ok = hipGraphExecDestroy(b->exec) == hipSuccess && ok;
  if (b->graph) ok = hipGraphDestroy(b->graph) == hipSuccess && ok;
  b->exec = nullptr;
  b->graph = nullptr;
  return ok ? 0 : -1;
}

extern "C" int gckpp_batch_set_handoff(gckpp_batch *b, int64_t max_cells) {
  if (!b) return fail("gckpp_batch_set_handoff: null batch");
  if (max_cells < 0) return fail("gckpp_batch_set_handoff: negative cell count");
  if (max_cells > INT32_MAX)
    return fail("gckpp_batch_set_handoff: %lld cells exceed the list's index range",
                (long long)max_cells);
  if (max_cells > b->ho_alloc) {
    // the list is reallocated: a captured graph would still launch kernels
    // on the old list, so it goes first (and in-flight work is waited for
    // before the old list is released)
    if (batch_drop_graph(b) != 0) return fail("gckpp_batch_set_handoff: releasing the graph failed");
    HIP_OK(hipDeviceSynchronize());
    int64_t nvar = 0;
    const int64_t hwf = wave_scratch_fields(PQ_MIXED);
    for (int q = 0; q < 3; ++q) {
      const MechInfo mi = mech_info(q);  // queue q holds mechanism q
      nvar = mi.nvar > nvar ? mi.nvar : nvar;
    }
    for (void *p : {(void *)b->ho_cell, (void *)b->ho_rec})
      if (p) (void)hipFree(p);
    b->ho_cell = nullptr;
    b->ho_rec = nullptr;
    b->ho_cap = b->ho_alloc = 0;
    hipError_t e = hipMalloc(&b->ho_cell, sizeof(int32_t) * (size_t)max_cells);
    if (e == hipSuccess)
      e = hipMalloc(&b->ho_rec, sizeof(double) * (size_t)(HO_FIELDS + nvar) * (size_t)max_cells);
    if (e == hipSuccess && !b->ho_ctr) e = hipMalloc(&b->ho_ctr, 4 * sizeof(unsigned int));
    if (e == hipSuccess && !b->hwork) {
      const int64_t res = persist_resident_waves(PQ_MIXED, PERSIST_WAVE_VARIANT, 1);
      if (res <= 0) e = hipErrorInvalidConfiguration;
      b->hw_fields = hwf;
      b->hwaves = res;
      if (e == hipSuccess)
        e = hipMalloc(&b->hwork, sizeof(double) * 64 * (size_t)hwf * (size_t)res);
    }
    if (e != hipSuccess) {
      for (void *p : {(void *)b->ho_cell, (void *)b->ho_rec})
        if (p) (void)hipFree(p);
      b->ho_cell = nullptr;
      b->ho_rec = nullptr;
      return fail("gckpp_batch_set_handoff: %s", hipGetErrorString(e));
    }
    b->ho_alloc = max_cells;
  }
  b->ho_cap = max_cells;
  return 0;
}

extern "C" int gckpp_batch_handoff_count(gckpp_batch *b, int64_t *count) {
  if (!b || !count) return fail("gckpp_batch_handoff_count: null argument");
  *count = 0;
  if (!b->ho_ctr) return 0;
  unsigned int n = 0;
  HIP_OK(hipDeviceSynchronize());
  HIP_OK(hipMemcpy(&n, b->ho_ctr + 1, sizeof n, hipMemcpyDeviceToHost));
  *count = (int64_t)n;
  return 0;
}

extern "C" int gckpp_batch_step_dev(gckpp_batch *b, const gckpp_part parts[3], double dt,
                                    void *stream) {
  if (!b || !parts) return fail("gckpp_batch_step: null argument");
  return batch_enqueue(b, parts, dt, (hipStream_t)stream);
}

extern "C" int gckpp_batch_capture(gckpp_batch *b, const gckpp_part parts[3], double dt) {
  if (!b || !parts) return fail("gckpp_batch_capture: null argument");
  if (batch_drop_graph(b) != 0) return fail("gckpp_batch_capture: releasing the old graph failed");
  HIP_OK(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
  int rc = batch_enqueue(b, parts, dt, b->cap_stream);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail("gckpp_batch_capture: %s", hipGetErrorString(e));
  b->graph = g;
  e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) return fail("gckpp_batch_capture: instantiate: %s", hipGetErrorString(e));
  return 0;
}

extern "C" int gckpp_batch_replay(gckpp_batch *b, void *stream) {
  if (!b) return fail("gckpp_batch_replay: null batch");
  if (!b->exec) return fail("gckpp_batch_replay: nothing captured");
  HIP_OK(hipGraphLaunch(b->exec, (hipStream_t)stream));
  return 0;
}

extern "C" int gckpp_batch_set_config(gckpp_batch *b, int variant, int waves_per_group) {
  if (!b) return fail("gckpp_batch_set_config: null batch");
  if (variant < 0 || variant >= PERSIST_VARIANTS ||
      (waves_per_group != 1 && waves_per_group != 2 && waves_per_group != 4))
    return fail("gckpp_batch_set_config: bad variant or waves per workgroup");
  if (batch_config(b, variant, waves_per_group, b->cap_waves) != 0)
    return fail("gckpp_batch_set_config: configuration does not fit");
  return 0;
}

extern "C" int gckpp_batch_set_waves(gckpp_batch *b, int64_t max_waves) {
  if (!b) return fail("gckpp_batch_set_waves: null batch");
  if (max_waves < 0) return fail("gckpp_batch_set_waves: negative wave count");
  const int64_t keep = b->max_waves;
  b->max_waves = max_waves;
  if (batch_config(b, b->variant, b->wpb, b->cap_waves) != 0) {
    b->max_waves = keep;
    return fail("gckpp_batch_set_waves: %lld waves do not hold one workgroup", (long long)max_waves);
  }
  return 0;
}

extern "C" int gckpp_batch_finalize(gckpp_batch *b) {
  if (!b) return 0;
  bool ok = batch_drop_graph(b) == 0;
  if (b->cap_stream) ok = hipStreamDestroy(b->cap_stream) == hipSuccess && ok;
  for (void *p : {(void *)b->work, (void *)b->counter, (void *)b->ho_ctr, (void *)b->ho_cell,
                  (void *)b->ho_rec, (void *)b->hwork})
    if (p) ok = hipFree(p) == hipSuccess && ok;
  delete b;
  return ok ? 0 : fail("gckpp_batch_finalize: releasing device resources failed");
}

// ---------------------------------------------------------------------------
// Step-size MLP (mlp.hip): inference for the integrator's Hstart and
// training on logged step histories.
struct gckpp_mlp {
  MlpDesc desc;
  int64_t max_batch;
  float *params, *grad, *adam_m, *adam_v;  // [MLP_NPARAM] each
  unsigned short *packed;                  // [MLP_NPACK] bf16
  bool packed_ok;
  int step;
  float *act, *dO, *dA, *lossp;            // training scratch for max_batch rows
};

extern "C" int gckpp_mlp_init(int64_t max_batch, const int32_t jslot[12], const int32_t spc[24],
                              gckpp_mlp **out) {
  if (!out || !jslot || !spc) return fail("gckpp_mlp_init: null argument");
  *out = nullptr;
  if (max_batch <= 0) return fail("gckpp_mlp_init: max_batch must be positive");
  auto *m = new gckpp_mlp();
  for (int k = 0; k < MLP_NJ; ++k) m->desc.jslot[k] = jslot[k];
  for (int k = 0; k < MLP_NS; ++k) m->desc.spc[k] = spc[k];
  for (int k = 0; k < MLP_F; ++k) {
    m->desc.mean[k] = 0.f;
    m->desc.rstd[k] = 1.f;
  }
  m->desc.out_mu = 0.f;
  m->desc.out_sigma = 1.f;
  m->max_batch = max_batch;
  size_t np = sizeof(float) * MLP_NPARAM;
  size_t na = sizeof(float) * (size_t)max_batch * (6 * MLP_H + MLP_O);
  if (hipMalloc(&m->params, np) != hipSuccess || hipMalloc(&m->grad, np) != hipSuccess ||
      hipMalloc(&m->adam_m, np) != hipSuccess || hipMalloc(&m->adam_v, np) != hipSuccess ||
      hipMalloc(&m->packed, sizeof(unsigned short) * MLP_NPACK) != hipSuccess ||
      hipMalloc(&m->act, na) != hipSuccess ||
      hipMalloc(&m->dO, sizeof(float) * (size_t)max_batch * MLP_O) != hipSuccess ||
      hipMalloc(&m->dA, sizeof(float) * (size_t)max_batch * MLP_H) != hipSuccess ||
      hipMalloc(&m->lossp, sizeof(float) * (size_t)((max_batch + 255) / 256)) != hipSuccess ||
      hipMemset(m->params, 0, np) != hipSuccess || hipMemset(m->adam_m, 0, np) != hipSuccess ||
      hipMemset(m->adam_v, 0, np) != hipSuccess) {
    gckpp_mlp_finalize(m);  // releases what was allocated (the rest is null)
    return fail("gckpp_mlp_init: device allocation failed");
  }
  m->packed_ok = false;
  *out = m;
  return 0;
}

extern "C" int gckpp_mlp_set_norm(gckpp_mlp *m, const float mean[48], const float stdev[48],
                                  float out_mu, float out_sigma) {
  if (!m || !mean || !stdev) return fail("gckpp_mlp_set_norm: null argument");
  for (int k = 0; k < MLP_F; ++k) {
    if (!(stdev[k] > 0.f)) return fail("gckpp_mlp_set_norm: std[%d] must be positive", k);
    m->desc.mean[k] = mean[k];
    m->desc.rstd[k] = 1.f / stdev[k];
  }
  if (!(out_sigma > 0.f)) return fail("gckpp_mlp_set_norm: out_sigma must be positive");
  m->desc.out_mu = out_mu;
  m->desc.out_sigma = out_sigma;
  return 0;
}

extern "C" int gckpp_mlp_set_params(gckpp_mlp *m, const float *params) {
  if (!m || !params) return fail("gckpp_mlp_set_params: null argument");
  HIP_OK(hipMemcpy(m->params, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
  size_t np = sizeof(float) * MLP_NPARAM;
  HIP_OK(hipMemset(m->adam_m, 0, np));
  HIP_OK(hipMemset(m->adam_v, 0, np));
  m->step = 0;
  m->packed_ok = false;
  return 0;
}

extern "C" int gckpp_mlp_get_params(gckpp_mlp *m, float *params) {
  if (!m || !params) return fail("gckpp_mlp_get_params: null argument");
  HIP_OK(hipMemcpy(params, m->params, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int gckpp_mlp_features_dev(gckpp_mlp *m, int64_t ncell, const double *met,
                                      const double *photol, const double *conc,
                                      const double *hprev, float *X, void *stream) {
  if (!m || !met || !photol || !conc || !X) return fail("gckpp_mlp_features: null argument");
  if (ncell <= 0) return ncell == 0 ? 0 : fail("gckpp_mlp_features: negative ncell");
  mlp_features(m->desc, ncell, met, photol, conc, hprev, X, (hipStream_t)stream);
  HIP_OK(hipGetLastError());
  return 0;
}

extern "C" int gckpp_mlp_predict_dev(gckpp_mlp *m, int64_t ncell, const double *met,
                                     const double *photol, const double *conc,
                                     const double *hprev, double safety, double hmin,
                                     double hmax, double *h0, int32_t *cls, void *stream) {
  if (!m || !met || !photol || !conc) return fail("gckpp_mlp_predict: null argument");
  if (ncell <= 0) return ncell == 0 ? 0 : fail("gckpp_mlp_predict: negative ncell");
  if (!(hmin > 0.0) || !(hmax >= hmin) || !(safety > 0.0))
    return fail("gckpp_mlp_predict: need 0 < hmin <= hmax and safety > 0");
  hipStream_t s = (hipStream_t)stream;
  if (!m->packed_ok) {
    mlp_pack(m->params, m->packed, s);
    m->packed_ok = true;
  }
  mlp_predict(m->desc, ncell, met, photol, conc, hprev, m->packed, m->params, safety, hmin, hmax,
              h0, cls, s);
  HIP_OK(hipGetLastError());
  return 0;
}

extern "C" int gckpp_mlp_forward_dev(gckpp_mlp *m, int64_t nb, const float *X, float *out,
                                     void *stream) {
  if (!m || !X || !out) return fail("gckpp_mlp_forward: null argument");
  if (nb <= 0 || nb > m->max_batch) return fail("gckpp_mlp_forward: nb must be in [1, max_batch]");
  hipStream_t s = (hipStream_t)stream;
  mlp_forward_f32(m->params, nb, X, m->act, s);
  HIP_OK(hipMemcpyAsync(out, m->act + (size_t)nb * 6 * MLP_H, sizeof(float) * nb * MLP_O,
                        hipMemcpyDeviceToDevice, s));
  return 0;
}

static int mlp_grad_impl(gckpp_mlp *m, int64_t nb, const float *X, const float *yh,
                         const int32_t *ycls, float wh, float wc, float *loss, hipStream_t s) {
  if (!m || !X || !yh || !ycls) return fail("gckpp_mlp_grad: null argument");
  if (nb <= 0 || nb > m->max_batch) return fail("gckpp_mlp_grad: nb must be in [1, max_batch]");
  mlp_grad_f32(m->params, nb, X, yh, ycls, wh, wc, m->act, m->dO, m->dA, m->grad, m->lossp, s);
  if (loss) mlp_sum(m->lossp, (nb + 255) / 256, loss, s);
  HIP_OK(hipGetLastError());
  return 0;
}

extern "C" int gckpp_mlp_grad_dev(gckpp_mlp *m, int64_t nb, const float *X, const float *yh,
                                  const int32_t *ycls, float wh, float wc, float *grad,
                                  float *loss, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  int rc = mlp_grad_impl(m, nb, X, yh, ycls, wh, wc, loss, s);
  if (rc) return rc;
  if (grad)
    HIP_OK(hipMemcpyAsync(grad, m->grad, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToDevice, s));
  return 0;
}

extern "C" int gckpp_mlp_train_step_dev(gckpp_mlp *m, int64_t nb, const float *X, const float *yh,
                                        const int32_t *ycls, float wh, float wc, float lr,
                                        float *loss, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  int rc = mlp_grad_impl(m, nb, X, yh, ycls, wh, wc, loss, s);
  if (rc) return rc;
  m->step += 1;
  mlp_adam(m->params, m->grad, m->adam_m, m->adam_v, lr, m->step, s);
  m->packed_ok = false;
  HIP_OK(hipGetLastError());
  return 0;
}

extern "C" int gckpp_mlp_finalize(gckpp_mlp *m) {
  if (!m) return 0;
  bool ok = true;
  for (void *p : {(void *)m->params, (void *)m->grad, (void *)m->adam_m, (void *)m->adam_v,
                  (void *)m->packed, (void *)m->act, (void *)m->dO, (void *)m->dA,
                  (void *)m->lossp})
    if (p) ok = hipFree(p) == hipSuccess && ok;
  delete m;
  return ok ? 0 : fail("gckpp_mlp_finalize: releasing device memory failed");
}

extern "C" int gckpp_order_by_class_dev(int64_t ncell, const int32_t *cls, int32_t *order,
                                        int32_t *scratch, void *stream) {
  if (ncell <= 0) return ncell == 0 ? 0 : fail("gckpp_order_by_class: negative ncell");
  if (!cls || !order || !scratch) return fail("gckpp_order_by_class: null argument");
  order_by_class(ncell, cls, order, scratch, (hipStream_t)stream);
  HIP_OK(hipGetLastError());
  return 0;
}

extern "C" int gckpp_order_by_h0_dev(int64_t ncell, const double *h0, int32_t *order,
                                     int32_t *scratch, void *stream) {
  if (ncell <= 0) return ncell == 0 ? 0 : fail("gckpp_order_by_h0: negative ncell");
  if (!h0 || !order || !scratch) return fail("gckpp_order_by_h0: null argument");
  order_by_h0(ncell, h0, order, scratch, (hipStream_t)stream);
  HIP_OK(hipGetLastError());
  return 0;
}

extern "C" int64_t gckpp_order_by_h0_scratch(int64_t ncell) {
  return ncell > 0 ? order_by_h0_scratch(ncell) : 0;
}

extern "C" int gckpp_set_persistent_split(gckpp_ctx *c, int64_t nfirst) {
  if (!c) return fail("gckpp_set_persistent_split: null context");
  if (nfirst < 0) return fail("gckpp_set_persistent_split: negative cell count");
  if (nfirst > 0 && c->mech == MECH_CARBON)
    return fail("gckpp_set_persistent_split: the carbon mechanism has no Rosenbrock solve");
  if (nfirst > 0 && !c->wwork) {
    const int64_t res = persist_resident_waves(c->mech, PERSIST_WAVE_VARIANT, 1);
    if (res <= 0) return fail("gckpp_set_persistent_split: cannot size the cell-per-wave launch");
    const size_t wb = sizeof(double) * (size_t)res * 64 * (size_t)wave_scratch_fields(c->mech);
    hipError_t e = hipMalloc(&c->wwork, wb);
    if (e == hipSuccess) e = hipMalloc(&c->wcounter, sizeof(unsigned int));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->s2, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
    if (e != hipSuccess) {
      if (c->wwork) (void)hipFree(c->wwork);
      if (c->wcounter) (void)hipFree(c->wcounter);
      if (c->s2) (void)hipStreamDestroy(c->s2);
      if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
      if (c->ev_join) (void)hipEventDestroy(c->ev_join);
      c->wwork = nullptr;
      c->wcounter = nullptr;
      c->s2 = nullptr;
      c->ev_fork = c->ev_join = nullptr;
      return fail("gckpp_set_persistent_split: %s (%.2f GB of slots)", hipGetErrorString(e),
                  wb / 1e9);
    }
    c->wwaves = res;
  }
  c->split = nfirst;
  return 0;
}

extern "C" int gckpp_set_persistent_handoff(gckpp_ctx *c, int64_t max_cells) {
  if (!c) return fail("gckpp_set_persistent_handoff: null context");
  if (max_cells < 0) return fail("gckpp_set_persistent_handoff: negative cell count");
  if (max_cells > 0 && c->mech == MECH_CARBON)
    return fail("gckpp_set_persistent_handoff: the carbon mechanism has no Rosenbrock solve");
  if (max_cells > INT32_MAX)
    return fail("gckpp_set_persistent_handoff: %lld cells exceed the list's index range",
                (long long)max_cells);
  if (max_cells > c->ho_alloc) {
    // (re)allocate the list for max_cells; the resume launch's scratch once
    const MechInfo mi = mech_info(c->mech);
    for (void *p : {(void *)c->ho_cell, (void *)c->ho_rec})
      if (p) (void)hipFree(p);
    c->ho_cell = nullptr;
    c->ho_rec = nullptr;
    c->ho_cap = c->ho_alloc = 0;
    hipError_t e = hipMalloc(&c->ho_cell, sizeof(int32_t) * (size_t)max_cells);
    if (e == hipSuccess)
      e = hipMalloc(&c->ho_rec, sizeof(double) * (size_t)(HO_FIELDS + mi.nvar) * (size_t)max_cells);
    if (e == hipSuccess && !c->ho_ctr) e = hipMalloc(&c->ho_ctr, 4 * sizeof(unsigned int));
    for (int k = 0; k < 3 && e == hipSuccess; ++k)
      if (!c->ev_ho[k]) e = hipEventCreate(&c->ev_ho[k]);
    if (e == hipSuccess && !c->hwork) {
      const int64_t res = persist_resident_waves(c->mech, PERSIST_WAVE_VARIANT, 1);
      if (res <= 0) e = hipErrorInvalidConfiguration;
      c->hw_fields = wave_scratch_fields(c->mech);
      c->hwaves = res;
      if (e == hipSuccess)
        e = hipMalloc(&c->hwork, sizeof(double) * 64 * (size_t)c->hw_fields * (size_t)res);
    }
    if (e != hipSuccess) {
      for (void *p : {(void *)c->ho_cell, (void *)c->ho_rec})
        if (p) (void)hipFree(p);
      c->ho_cell = nullptr;
      c->ho_rec = nullptr;
      return fail("gckpp_set_persistent_handoff: %s", hipGetErrorString(e));
    }
    c->ho_alloc = max_cells;
  }
  c->ho_cap = max_cells;
  return 0;
}

extern "C" int gckpp_persistent_handoff_stats(gckpp_ctx *c, int64_t n[2], float ms[2]) {
  if (!c || !n || !ms) return fail("gckpp_persistent_handoff_stats: null argument");
  n[0] = n[1] = 0;
  ms[0] = ms[1] = 0.0f;
  if (!c->ho_ctr || !c->ho_timed) return 0;
  unsigned int v[4];
  HIP_OK(hipDeviceSynchronize());
  HIP_OK(hipMemcpy(v, c->ho_ctr, sizeof v, hipMemcpyDeviceToHost));
  n[0] = (int64_t)v[1];
  n[1] = (int64_t)v[3];
  HIP_OK(hipEventElapsedTime(&ms[0], c->ev_ho[0], c->ev_ho[1]));
  HIP_OK(hipEventElapsedTime(&ms[1], c->ev_ho[1], c->ev_ho[2]));
  return 0;
}
