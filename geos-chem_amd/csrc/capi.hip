// C ABI of libgckpp_hip (include/gckpp_hip.h): contexts, option resolution,
// workspace and chunking around the HIP kernels.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/gckpp_hip.h"
#include "cell_fields.h"
#include "gckpp_internal.h"

namespace gckpp {
void launch_rates_mech(int mech, const RatesArgs &a, hipStream_t s);
bool rates_available(int mech);
}

using namespace gckpp;

static thread_local char g_err[512];

static int fail(const char *fmt, ...) {
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
  double *work;        // [integrate_work_per_cell][wstride]
  double *atol, *rtol; // [NVAR] device
  double *consts;      // [3][NSPEC] device: MW, HENRY_K0, HENRY_CR
  double *step_buf;    // fullchem_step: rconst [NREACT][wstride], conc backup
                       // [NSPEC][wstride], rstatus [20][wstride]
  int32_t *nfail;      // [1] device counter of cells that failed twice
  hipEvent_t ev[5];    // step timing: rates | backup | integrate | retry+finish
  int solver;          // GCKPP_SOLVER_PLAIN or GCKPP_SOLVER_PERSISTENT
  unsigned int *counter;  // work-stealing cell counter of the persistent solver
  int64_t pblocks;     // waves of one persistent launch
  // scratch for the host-pointer and retry paths
  std::vector<double> h_atol, h_rtol;
};

extern "C" const char *gckpp_last_error(void) { return g_err; }

static bool mech_ok(int mech) { return mech == MECH_FULLCHEM || mech == MECH_HG; }

extern "C" int gckpp_mech_sizes(int mech_id, int32_t out[8]) {
  if (!mech_ok(mech_id)) return fail("unknown or unsupported mechanism id %d", mech_id);
  MechInfo m = mech_info(mech_id);
  out[0] = m.nspec; out[1] = m.nvar; out[2] = m.nfix; out[3] = m.nreact;
  out[4] = m.lu_nz; out[5] = NMET; out[6] = NHET; out[7] = NPHOTOL;
  g_err[0] = 0;
  return 0;
}

extern "C" int gckpp_init(int mech_id, int64_t max_cells, gckpp_ctx **out) {
  if (!out) return fail("gckpp_init: null output pointer");
  *out = nullptr;
  if (!mech_ok(mech_id)) return fail("gckpp_init: unknown or unsupported mechanism id %d", mech_id);
  if (max_cells <= 0) return fail("gckpp_init: max_cells must be positive");
  auto *c = new gckpp_ctx();
  c->mech = mech_id;
  c->info = mech_info(mech_id);
  HIP_OK(hipGetDevice(&c->device));
  c->max_cells = max_cells;
  c->wstride = (max_cells + 63) / 64 * 64;
  int64_t per_cell = integrate_work_per_cell(c->info);
  if (persistent_slot_fields(mech_id) > per_cell) per_cell = persistent_slot_fields(mech_id);
  size_t wbytes = sizeof(double) * (size_t)per_cell * (size_t)c->wstride;
  hipError_t e = hipMalloc(&c->work, wbytes);
  if (e != hipSuccess) {
    delete c;
    return fail("gckpp_init: cannot allocate %.2f GB of workspace for %lld cells: %s",
                wbytes / 1e9, (long long)max_cells, hipGetErrorString(e));
  }
  HIP_OK(hipMalloc(&c->atol, sizeof(double) * c->info.nvar));
  HIP_OK(hipMalloc(&c->rtol, sizeof(double) * c->info.nvar));
  HIP_OK(hipMalloc(&c->consts, sizeof(double) * 3 * c->info.nspec));
  HIP_OK(hipMemset(c->consts, 0, sizeof(double) * 3 * c->info.nspec));
  {
    size_t sb = sizeof(double) * (size_t)(c->info.nreact + c->info.nspec + 20) * (size_t)c->wstride;
    hipError_t e2 = hipMalloc(&c->step_buf, sb);
    if (e2 != hipSuccess)
      return fail("gckpp_init: cannot allocate %.2f GB of step buffers: %s", sb / 1e9,
                  hipGetErrorString(e2));
  }
  HIP_OK(hipMalloc(&c->nfail, sizeof(int32_t)));
  for (int k = 0; k < 5; ++k) HIP_OK(hipEventCreate(&c->ev[k]));
  HIP_OK(hipMalloc(&c->counter, sizeof(unsigned int)));
  c->solver = GCKPP_SOLVER_PLAIN;
  c->pblocks = persistent_blocks(mech_id);
  if (c->pblocks <= 0) c->pblocks = 1024;
  if (c->pblocks > c->wstride / 64) c->pblocks = c->wstride / 64;
  // Init_FullChem defaults (GeosCore/fullchem_mod.F90:2899-2905)
  c->h_atol.assign(c->info.nvar, 1.0e-2);
  c->h_rtol.assign(c->info.nvar, 0.5e-2);
  HIP_OK(hipMemcpy(c->atol, c->h_atol.data(), sizeof(double) * c->info.nvar, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(c->rtol, c->h_rtol.data(), sizeof(double) * c->info.nvar, hipMemcpyHostToDevice));
  *out = c;
  g_err[0] = 0;
  return 0;
}

extern "C" int gckpp_finalize(gckpp_ctx *c) {
  if (!c) return 0;
  hipFree(c->work);
  hipFree(c->atol);
  hipFree(c->rtol);
  hipFree(c->consts);
  hipFree(c->step_buf);
  hipFree(c->nfail);
  for (int k = 0; k < 5; ++k) hipEventDestroy(c->ev[k]);
  hipFree(c->counter);
  delete c;
  return 0;
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
    hipLaunchKernelGGL(k_fill_status, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, s,
                       ncell, ierr, code, istatus, rstatus);
    HIP_OK(hipGetLastError());
    return 0;
  }
  if (c->solver == GCKPP_SOLVER_PERSISTENT && !only_failed) {
    // one launch of resident waves stealing cells; any ncell
    IntegrateArgs a{};
    a.ncell = ncell;
    a.conc = conc;
    a.cstride = ncell;
    a.rconst = rconst;
    a.rstride = ncell;
    a.hstart = hstart;
    a.atol = c->atol;
    a.rtol = c->rtol;
    a.work = c->work;
    a.prm = p;
    a.istatus = istatus;
    a.rstatus = rstatus;
    a.ierr = ierr;
    a.sstride = ncell;
    int64_t nb = (ncell + 63) / 64;
    if (nb > c->pblocks) nb = c->pblocks;
    launch_integrate_persistent_mech(c->mech, a, nb, c->counter, s);
    HIP_OK(hipGetLastError());
    return 0;
  }
  // chunks of at most wstride cells; outputs are indexed with the full ncell
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

static int comp(gckpp_ctx *c, int which, int64_t ncell, const double *in, const double *rct,
                double *out, int32_t *ier, void *stream) {
  if (!c) return fail("null context");
  if (ncell <= 0) return 0;
  if (ncell > c->wstride) return fail("ncell=%lld exceeds max_cells", (long long)ncell);
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

// Host-pointer Integrate: copies in, solves, copies out (synchronous).
extern "C" int gckpp_integrate(gckpp_ctx *c, int64_t ncell, double tin, double tout,
                               const int32_t icntrl[20], const double rcntrl[20],
                               const double *hstart, double *conc, const double *rconst,
                               int32_t *istatus, double *rstatus, int32_t *ierr) {
  if (!c) return fail("gckpp_integrate: null context");
  if (ncell <= 0) return ncell == 0 ? 0 : fail("gckpp_integrate: negative ncell");
  size_t nsp = c->info.nspec, nre = c->info.nreact, n = ncell;
  double *dc = nullptr, *dr = nullptr, *dh = nullptr, *drs = nullptr;
  int32_t *dis = nullptr, *die = nullptr;
  int rc = 0;
  if (hipMalloc(&dc, 8 * nsp * n) != hipSuccess || hipMalloc(&dr, 8 * nre * n) != hipSuccess ||
      hipMalloc(&dh, 8 * n) != hipSuccess || hipMalloc(&drs, 8 * 20 * n) != hipSuccess ||
      hipMalloc(&dis, 4 * 20 * n) != hipSuccess || hipMalloc(&die, 4 * n) != hipSuccess) {
    rc = fail("gckpp_integrate: device allocation failed");
  } else {
    hipMemcpy(dc, conc, 8 * nsp * n, hipMemcpyHostToDevice);
    hipMemcpy(dr, rconst, 8 * nre * n, hipMemcpyHostToDevice);
    if (hstart) hipMemcpy(dh, hstart, 8 * n, hipMemcpyHostToDevice);
    else hipMemset(dh, 0, 8 * n);
    rc = gckpp_integrate_dev(c, ncell, tin, tout, icntrl, rcntrl, dh, dc, dr, dis, drs, die, nullptr);
    if (rc == 0) {
      hipError_t e = hipDeviceSynchronize();
      if (e != hipSuccess) rc = fail("gckpp_integrate: %s", hipGetErrorString(e));
    }
    if (rc == 0) {
      hipMemcpy(conc, dc, 8 * nsp * n, hipMemcpyDeviceToHost);
      if (istatus) hipMemcpy(istatus, dis, 4 * 20 * n, hipMemcpyDeviceToHost);
      if (rstatus) hipMemcpy(rstatus, drs, 8 * 20 * n, hipMemcpyDeviceToHost);
      if (ierr) hipMemcpy(ierr, die, 4 * n, hipMemcpyDeviceToHost);
    }
  }
  hipFree(dc); hipFree(dr); hipFree(dh); hipFree(drs); hipFree(dis); hipFree(die);
  return rc;
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
__global__ __launch_bounds__(256) void k_step_finish(int64_t n, int nspec, double *conc,
                                                     const double *rst, double *hnew,
                                                     const int32_t *ierr, int32_t *partial) {
  __shared__ int32_t red[256];
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int bad = 0;
  if (i < n) {
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

extern "C" int gckpp_fullchem_step_dev(gckpp_ctx *c, int64_t ncell, double dt,
                                       const int32_t icntrl[20], const double rcntrl[20],
                                       const double *met, const double *photol,
                                       const double *het, int64_t het_stride,
                                       const double *hstart, double *conc, double *hnew,
                                       int32_t *istatus, int32_t *ierr, void *stream) {
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
                     c->info.nspec, conc, rst, hnew, ierr, partial);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, nb, partial, c->nfail);
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventRecord(c->ev[4], s));
  int32_t nfail = 0;
  HIP_OK(hipMemcpyAsync(&nfail, c->nfail, sizeof nfail, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  if (nfail > 0) return fail("gckpp_fullchem_step: %d cell(s) failed twice", nfail);
  return 0;
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
