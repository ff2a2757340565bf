// Rodas3 Rosenbrock solve of a batch of KPP cells on gfx950: one lane per
// cell, every per-cell array field-major with cells fastest so each table
// step is one coalesced 512-byte access per wavefront.  The mechanism tables
// (csrc/gen/mech_<name>.h) are walked with wavefront-uniform indices.
//
// Algorithm (restated, not translated) from the reference's KPP-3.2.0 code:
//   ros_Integrator time loop      KPP/fullchem/gckpp_Integrator.F90:578-786
//   Rodas3 tableau                KPP/fullchem/gckpp_Integrator.F90:2239-2303
//   ros_PrepareMatrix             KPP/fullchem/gckpp_Integrator.F90:1921-1999
//   ros_ErrorNorm                 KPP/fullchem/gckpp_Integrator.F90:1715-1745
//   Fun_SPLIT, Jac_SP             gckpp_Function.F90:2172-5060, gckpp_Jacobian.F90:48-20887
//   KppDecomp, KppSolve           gckpp_LinearAlgebra.F90:46-83, :644-2309
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "gckpp_internal.h"
#include "gen/mech_Hg.h"
#include "gen/mech_fullchem.h"

namespace gckpp {

// Per-lane view of a cells-fastest array: element (k, cell) at p[k*stride + cell].
struct Col {
  double *p;
  int64_t stride;
  __device__ double &operator[](int k) const { return p[(int64_t)k * stride]; }
};
struct CCol {
  const double *p;
  int64_t stride;
  __device__ double operator[](int k) const { return p[(int64_t)k * stride]; }
};

// ---------------------------------------------------------------------------
// Mechanism-generic device pieces.  M is one of the mech_<name> namespaces,
// wrapped in a traits struct so it can be a template argument.
#define GCKPP_MECH_TRAITS(NS)                                                  \
  struct Traits_##NS {                                                         \
    static constexpr int NVAR = mech_##NS::NVAR, NFIX = mech_##NS::NFIX;       \
    static constexpr int NSPEC = mech_##NS::NSPEC, NREACT = mech_##NS::NREACT; \
    static constexpr int LU_NZ = mech_##NS::LU_NZ, NPART = mech_##NS::NPART;   \
    static constexpr int RX_MAX = mech_##NS::RX_MAX;                           \
    static __device__ int rx_s(int i) { return mech_##NS::rx_s[i]; }           \
    static __device__ int fv_ptr(int i) { return mech_##NS::fv_ptr[i]; }       \
    static __device__ int fv_r(int i) { return mech_##NS::fv_r[i]; }           \
    static __device__ double fv_c(int i) { return mech_##NS::fv_c[i]; }        \
    static __device__ int pt_r(int i) { return mech_##NS::pt_r[i]; }           \
    static __device__ double pt_m(int i) { return mech_##NS::pt_m[i]; }        \
    static __device__ int pt_s(int i) { return mech_##NS::pt_s[i]; }           \
    static __device__ int jv_ptr(int i) { return mech_##NS::jv_ptr[i]; }       \
    static __device__ int jv_b(int i) { return mech_##NS::jv_b[i]; }           \
    static __device__ double jv_c(int i) { return mech_##NS::jv_c[i]; }        \
    static __device__ int row_e0(int i) { return mech_##NS::row_e0[i]; }       \
    static __device__ int le_kk(int i) { return mech_##NS::le_kk[i]; }         \
    static __device__ int le_dj(int i) { return mech_##NS::le_dj[i]; }         \
    static __device__ int le_p0(int i) { return mech_##NS::le_p0[i]; }         \
    static __device__ int le_p1(int i) { return mech_##NS::le_p1[i]; }         \
    static __device__ int lp_dst(int i) { return mech_##NS::lp_dst[i]; }       \
    static __device__ int lp_src(int i) { return mech_##NS::lp_src[i]; }       \
    static __device__ int lu_crow(int i) { return mech_##NS::lu_crow[i]; }     \
    static __device__ int lu_diag(int i) { return mech_##NS::lu_diag[i]; }     \
    static __device__ int lu_icol(int i) { return mech_##NS::lu_icol[i]; }     \
    static __device__ int cr_ptr(int i) { return mech_##NS::cr_ptr[i]; }       \
    static __device__ int cr_l(int i) { return mech_##NS::cr_l[i]; }           \
    static __device__ int cr_u(int i) { return mech_##NS::cr_u[i]; }           \
    static __device__ int cr_div(int i) { return mech_##NS::cr_div[i]; }       \
  };
GCKPP_MECH_TRAITS(fullchem)
GCKPP_MECH_TRAITS(Hg)

// Concentration of species s (variable from y, fixed from the conc array).
template <class M>
__device__ inline double spc(const CCol &y, const CCol &fix, int s) {
  return s < M::NVAR ? y[s] : fix[s];
}

// A(r) = RCT(r)*prod(reactants) into a[], then Vdot(i) = sum c*A(r).
template <class M>
__device__ void dev_fun(const CCol &y, const CCol &fix, const CCol &rct, Col a, Col vdot) {
  for (int r = 0; r < M::NREACT; ++r) {
    double v = rct[r];
#pragma unroll
    for (int q = 0; q < M::RX_MAX; ++q) {
      int s = M::rx_s(r * M::RX_MAX + q);
      if (s < 0) break;
      v *= spc<M>(y, fix, s);
    }
    a[r] = v;
  }
  for (int i = 0; i < M::NVAR; ++i) {
    double acc = 0.0;
    int t1 = M::fv_ptr(i + 1);
    for (int t = M::fv_ptr(i); t < t1; ++t) acc = fma(M::fv_c(t), a[M::fv_r(t)], acc);
    vdot[i] = acc;
  }
}

// JVS in LU layout: partials into b[], then each LU entry gathers its terms;
// out[k] = sign*JVS(k) (+ diag on the diagonal), so sign=-1, diag=1/(H*gamma)
// builds ros_PrepareMatrix's Ghimj directly.
template <class M>
__device__ void dev_jac(const CCol &y, const CCol &fix, const CCol &rct, Col b, Col out,
                        double sign, double diag) {
  for (int p = 0; p < M::NPART; ++p) {
    double v = M::pt_m(p) * rct[M::pt_r(p)];
#pragma unroll
    for (int q = 0; q < M::RX_MAX - 1; ++q) {
      int s = M::pt_s(p * (M::RX_MAX - 1) + q);
      if (s < 0) break;
      v *= spc<M>(y, fix, s);
    }
    b[p] = v;
  }
  for (int k = 0; k < M::LU_NZ; ++k) {
    double acc = 0.0;
    int t1 = M::jv_ptr(k + 1);
    for (int t = M::jv_ptr(k); t < t1; ++t) acc = fma(M::jv_c(t), b[M::jv_b(t)], acc);
    out[k] = sign * acc;
  }
  if (diag != 0.0)
    for (int i = 0; i < M::NVAR; ++i) out[M::lu_diag(i)] += diag;
}

// KppDecomp in place; returns IER (0 or the 1-based row whose diagonal was
// below TINY when its elimination started, as the reference checks it).
template <class M>
__device__ int dev_decomp(Col lu) {
  int ier = 0;
  for (int k = 0; k < M::NVAR; ++k) {
    if (ier == 0 && fabs(lu[M::lu_diag(k)]) < DBL_MIN) ier = k + 1;
    int e1 = M::row_e0(k + 1);
    for (int e = M::row_e0(k); e < e1; ++e) {
      int kk = M::le_kk(e);
      double a = lu[kk] / lu[M::le_dj(e)];
      lu[kk] = a;
      int p1 = M::le_p1(e);
      for (int p = M::le_p0(e); p < p1; ++p) {
        int d = M::lp_dst(p);
        lu[d] = fma(-a, lu[M::lp_src(p)], lu[d]);
      }
    }
  }
  return ier;
}

// KppDecomp in place, as one dot product per LU position in row-major order
// (row-Crout form): position (k,c) accumulates -L(k,j)*U(j,c) over row k's
// pivots j < min(c,k) in increasing j -- the same updates, in the same order,
// as the row-wise elimination -- then is divided by U(c,c) when c < k.  Each
// position is written once and its partial sum stays in a register.
template <class M>
__device__ int dev_decomp_crout(Col lu) {
  int ier = 0;
  int q = 0;
  for (int k = 0; k < M::NVAR; ++k) {
    if (ier == 0 && fabs(lu[M::lu_diag(k)]) < DBL_MIN) ier = k + 1;
    const int q1 = M::lu_crow(k + 1);
    for (; q < q1; ++q) {
      double acc = lu[q];
      const int t1 = M::cr_ptr(q + 1);
      for (int t = M::cr_ptr(q); t < t1; ++t) acc = fma(-lu[M::cr_l(t)], lu[M::cr_u(t)], acc);
      const int dv = M::cr_div(q);
      if (dv >= 0) acc = acc / lu[dv];
      lu[q] = acc;
    }
  }
  return ier;
}

// KppSolve in place: unit-lower forward then upper backward substitution.
template <class M>
__device__ void dev_solve(const CCol &lu, Col x) {
  for (int i = 0; i < M::NVAR; ++i) {
    int k0 = M::lu_crow(i), k1 = M::lu_diag(i);
    if (k0 == k1) continue;
    double acc = x[i];
    for (int k = k0; k < k1; ++k) acc = fma(-lu[k], x[M::lu_icol(k)], acc);
    x[i] = acc;
  }
  for (int i = M::NVAR - 1; i >= 0; --i) {
    int kd = M::lu_diag(i), k1 = M::lu_crow(i + 1);
    double acc = x[i];
    for (int k = kd + 1; k < k1; ++k) acc = fma(-lu[k], x[M::lu_icol(k)], acc);
    x[i] = acc / lu[kd];
  }
}

// ---------------------------------------------------------------------------
// Rodas3 (gckpp_Integrator.F90:2239-2303): stages s=1..4; A, C lower
// triangular row-wise; M, E solution/error weights; gamma(1)=0.5.
__constant__ double kRosA[6] = {0.0, 2.0, 0.0, 2.0, 0.0, 1.0};
__constant__ double kRosC[6] = {4.0, 1.0, -1.0, 1.0, -1.0, -(8.0 / 3.0)};
__constant__ double kRosM[4] = {2.0, 0.0, 1.0, 1.0};
__constant__ double kRosE[4] = {0.0, 0.0, 0.0, 1.0};
__constant__ int kRosNewF[4] = {1, 0, 1, 1};
constexpr double kRosGamma1 = 0.5;
constexpr double kRosELO = 3.0;

template <class M>
__global__ __launch_bounds__(64) void k_integrate(IntegrateArgs a) {
  const int64_t cell = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (cell >= a.ncell) return;
  if (a.only_failed && a.only_failed[cell] >= 0) return;
  if (a.restore_from)
    for (int k = 0; k < M::NSPEC; ++k)
      a.conc[(int64_t)k * a.cstride + cell] = a.restore_from[(int64_t)k * a.cstride + cell];
  const int N = M::NVAR;
  Col Y{a.conc + cell, a.cstride};
  CCol Yc{a.conc + cell, a.cstride};
  CCol fix{a.conc + cell, a.cstride};
  CCol rct{a.rconst + cell, a.rstride};
  const int64_t ws = a.wstride;
  double *w = a.work + cell;
  Col LU{w, ws};
  Col Fcn0{w + (int64_t)M::LU_NZ * ws, ws};
  Col K{w + (int64_t)(M::LU_NZ + N) * ws, ws};            // 4*N
  Col Ynew{w + (int64_t)(M::LU_NZ + 5 * N) * ws, ws};
  Col Ascr{w + (int64_t)(M::LU_NZ + 6 * N) * ws, ws};      // NREACT (Fun) / NPART (Jac)
  CCol Ynewc{Ynew.p, ws};
  auto Ks = [&](int s) { return Col{K.p + (int64_t)s * N * ws, ws}; };
  auto Ksc = [&](int s) { return CCol{K.p + (int64_t)s * N * ws, ws}; };

  const RosParams &P = a.prm;
  const double Tend = P.Tend;
  double hs = a.hstart ? a.hstart[cell] : 0.0;
  // RCNTRL(3) > 0 wins over the shared value (Integrate merges RCNTRL_U > 0)
  double Hstart = P.Hstart_default;
  if (hs > 0.0) Hstart = min_f(fabs(hs), fabs(Tend - P.Tstart));
  double T = P.Tstart;
  double H = min_f(max_f(fabs(P.Hmin), fabs(Hstart)), fabs(P.Hmax));
  if (fabs(H) <= 10.0 * P.roundoff) H = P.DeltaMin;
  const double dir = P.direction;
  H = dir * H;
  double Hnew = H, Hexit = 0.0, Texit = 0.0, HnewAcc = 0.0;
  int st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int ierr = 1;
  bool rejLast = false, rejMore = false, newStep = true;

  for (;;) {
    if (newStep) {
      bool more = (dir > 0) ? ((T - Tend) + P.roundoff <= 0.0) : ((Tend - T) + P.roundoff <= 0.0);
      if (!more) break;
      if (st[2] > P.max_steps) { ierr = -6; break; }
      if ((T + 0.1 * H) == T || H <= P.roundoff) { ierr = -7; break; }
      H = min_f(H, fabs(Tend - T));
      dev_fun<M>(Yc, fix, rct, Ascr, Fcn0);
      st[0] += P.autonomous ? 1 : 2;  // non-autonomous: dF/dT by finite difference (0 with frozen rates)
      st[1] += 1;
    }
    // ros_PrepareMatrix: Ghimj = 1/(H*gamma1) - Jac0, LU, halve H on failure
    int nconsec = 0;
    bool singular = true;
    while (singular) {
      dev_jac<M>(Yc, fix, rct, Ascr, LU, -1.0, 1.0 / (dir * H * kRosGamma1));
      int ising = dev_decomp<M>(LU);
      st[5] += 1;
      if (ising == 0) {
        singular = false;
      } else {
        st[7] += 1;
        ++nconsec;
        if (nconsec <= 5) H *= 0.5;
        else break;
      }
    }
    if (singular) { ierr = -8; break; }
    // stages
    for (int s = 0; s < 4; ++s) {
      Col Kc = Ks(s);
      if (s == 0) {
        for (int i = 0; i < N; ++i) Kc[i] = Fcn0[i];
      } else if (kRosNewF[s]) {
        for (int i = 0; i < N; ++i) {
          double v = Yc[i];
          for (int j = 0; j < s; ++j) {
            double aj = kRosA[s * (s - 1) / 2 + j];
            if (aj != 0.0) v = fma(aj, Ksc(j)[i], v);
          }
          Ynew[i] = v;
        }
        dev_fun<M>(Ynewc, fix, rct, Ascr, Kc);
        st[0] += 1;
      } else {
        for (int i = 0; i < N; ++i) Kc[i] = Fcn0[i];
      }
      for (int j = 0; j < s; ++j) {
        double hc = kRosC[s * (s - 1) / 2 + j] / (dir * H);
        if (hc != 0.0)
          for (int i = 0; i < N; ++i) Kc[i] = fma(hc, Ksc(j)[i], Kc[i]);
      }
      dev_solve<M>(CCol{LU.p, ws}, Kc);
      st[6] += 1;
    }
    // new solution, error estimate (E = (0,0,0,1): Yerr = K4), scaled norm
    double err = 0.0;
    for (int i = 0; i < N; ++i) {
      double y0 = Yc[i];
      double v = y0;
      for (int j = 0; j < 4; ++j)
        if (kRosM[j] != 0.0) v = fma(kRosM[j], Ksc(j)[i], v);
      Ynew[i] = v;
      double ye = 0.0;
      for (int j = 0; j < 4; ++j)
        if (kRosE[j] != 0.0) ye = fma(kRosE[j], Ksc(j)[i], ye);
      double ymax = max_f(fabs(y0), fabs(v));
      double sc = P.vectol ? a.atol[i] + a.rtol[i] * ymax : a.atol[0] + a.rtol[0] * ymax;
      double q = ye / sc;
      err = fma(q, q, err);
    }
    err = max_f(sqrt(err / N), 1.0e-10);
    double fac = min_f(P.FacMax, max_f(P.FacMin, P.FacSafe / pow(err, 1.0 / kRosELO)));
    Hnew = H * fac;
    st[2] += 1;
    if (err <= 1.0 || H <= P.Hmin) {
      st[3] += 1;
      if (P.clip_negative)
        for (int i = 0; i < N; ++i) Y[i] = max_f(Ynew[i], 0.0);
      else
        for (int i = 0; i < N; ++i) Y[i] = Ynew[i];
      T = T + dir * H;
      Hnew = max_f(P.Hmin, min_f(Hnew, P.Hmax));
      if (rejLast) Hnew = min_f(Hnew, H);
      Hexit = H;
      Texit = T;
      HnewAcc = Hnew;
      rejLast = false;
      rejMore = false;
      H = Hnew;
      newStep = true;
    } else {
      if (rejMore) Hnew = H * P.FacRej;
      rejMore = rejLast;
      rejLast = true;
      H = Hnew;
      if (st[3] >= 1) st[4] += 1;
      newStep = false;
    }
  }
  if (a.istatus)
    for (int k = 0; k < 20; ++k) a.istatus[(int64_t)k * a.sstride + cell] = k < 8 ? st[k] : 0;
  if (a.rstatus) {
    for (int k = 0; k < 20; ++k) a.rstatus[(int64_t)k * a.sstride + cell] = 0.0;
    a.rstatus[cell] = Texit;
    a.rstatus[a.sstride + cell] = Hexit;
    a.rstatus[2 * a.sstride + cell] = HnewAcc;
  }
  if (a.ierr) a.ierr[cell] = ierr;
}

// ---------------------------------------------------------------------------
// Persistent, work-stealing Rodas3 integrator.  Each workgroup is one wave;
// each lane owns a slot of the workspace block of its wave, laid out
// field-major with the 64 slots fastest ([field][64], one contiguous block per
// wave), and integrates one cell after another: when its cell reaches Tend
// (or fails) it writes the cell back and takes the next cell index from a
// global counter, so lanes of one wave never wait for a slow cell of another
// lane.  Without a counter each lane integrates exactly one cell (the cell of
// its global thread index).  The per-cell algorithm is that of k_integrate,
// with the LU factorisation in row-Crout order.
template <class M>
struct SlotLayout {
  static constexpr int SCR = M::NREACT > M::NPART ? M::NREACT : M::NPART;
  static constexpr int Y = 0;                       // NSPEC (variable + fixed)
  static constexpr int R = Y + M::NSPEC;            // NREACT rate constants
  static constexpr int F0 = R + M::NREACT;          // NVAR Fcn0
  static constexpr int K = F0 + M::NVAR;            // 4*NVAR stage vectors
  static constexpr int YN = K + 4 * M::NVAR;        // NVAR Ynew
  static constexpr int S = YN + M::NVAR;            // SCR reaction rates / partials
  static constexpr int LU = S + SCR;                // LU_NZ
  static constexpr int FIELDS = LU + M::LU_NZ;
};

template <class M>
__global__ __launch_bounds__(64) void k_integrate_p(IntegrateArgs a, unsigned int *counter) {
  using L = SlotLayout<M>;
  const int lane = threadIdx.x;
  const int N = M::NVAR;
  double *blk = a.work + (int64_t)blockIdx.x * (64 * (int64_t)L::FIELDS) + lane;
  auto F = [&](int off) { return Col{blk + (int64_t)off * 64, 64}; };
  auto Fc = [&](int off) { return CCol{blk + (int64_t)off * 64, 64}; };
  Col Y = F(L::Y), Fcn0 = F(L::F0), Ynew = F(L::YN), Ascr = F(L::S), LU = F(L::LU);
  CCol Yc = Fc(L::Y), rct = Fc(L::R), Ynewc = Fc(L::YN);
  auto Ks = [&](int s) { return F(L::K + s * N); };
  auto Ksc = [&](int s) { return Fc(L::K + s * N); };
  const RosParams &P = a.prm;
  const double Tend = P.Tend, dir = P.direction;

  int64_t cell = -1;
  bool first = true;
  double T = 0, H = 0, Hnew = 0, Hexit = 0, Texit = 0, HnewAcc = 0;
  int st[8];
  int ierr = 1;
  bool rejLast = false, rejMore = false, newStep = true;

  for (;;) {
    if (cell < 0) {
      // take the next cell
      int64_t c;
      if (counter) c = (int64_t)atomicAdd(counter, 1u);
      else c = first ? (int64_t)blockIdx.x * 64 + lane : a.ncell;
      first = false;
      if (c >= a.ncell) break;
      cell = c;
      for (int s = 0; s < M::NSPEC; ++s) Y[s] = a.conc[(int64_t)s * a.cstride + cell];
      for (int r = 0; r < M::NREACT; ++r) F(L::R)[r] = a.rconst[(int64_t)r * a.rstride + cell];
      double hs = a.hstart ? a.hstart[cell] : 0.0;
      double Hstart = P.Hstart_default;
      if (hs > 0.0) Hstart = min_f(fabs(hs), fabs(Tend - P.Tstart));
      T = P.Tstart;
      H = min_f(max_f(fabs(P.Hmin), fabs(Hstart)), fabs(P.Hmax));
      if (fabs(H) <= 10.0 * P.roundoff) H = P.DeltaMin;
      H = dir * H;
      Hnew = H;
      Hexit = Texit = HnewAcc = 0.0;
      for (int k = 0; k < 8; ++k) st[k] = 0;
      ierr = 1;
      rejLast = rejMore = false;
      newStep = true;
    }
    bool finished = false;
    if (newStep) {
      bool more = (dir > 0) ? ((T - Tend) + P.roundoff <= 0.0) : ((Tend - T) + P.roundoff <= 0.0);
      if (!more) finished = true;
      else if (st[2] > P.max_steps) { ierr = -6; finished = true; }
      else if ((T + 0.1 * H) == T || H <= P.roundoff) { ierr = -7; finished = true; }
    }
    if (!finished) {
      if (newStep) {
        H = min_f(H, fabs(Tend - T));
        dev_fun<M>(Yc, Yc, rct, Ascr, Fcn0);
        st[0] += P.autonomous ? 1 : 2;
        st[1] += 1;
      }
      int nconsec = 0;
      bool singular = true;
      while (singular) {
        dev_jac<M>(Yc, Yc, rct, Ascr, LU, -1.0, 1.0 / (dir * H * kRosGamma1));
        int ising = dev_decomp_crout<M>(LU);
        st[5] += 1;
        if (ising == 0) {
          singular = false;
        } else {
          st[7] += 1;
          ++nconsec;
          if (nconsec <= 5) H *= 0.5;
          else break;
        }
      }
      if (singular) {
        ierr = -8;
        finished = true;
      } else {
        for (int s = 0; s < 4; ++s) {
          Col Kc = Ks(s);
          if (s == 0 || !kRosNewF[s]) {
            for (int i = 0; i < N; ++i) Kc[i] = Fcn0[i];
          } else {
            for (int i = 0; i < N; ++i) {
              double v = Yc[i];
              for (int j = 0; j < s; ++j) {
                double aj = kRosA[s * (s - 1) / 2 + j];
                if (aj != 0.0) v = fma(aj, Ksc(j)[i], v);
              }
              Ynew[i] = v;
            }
            dev_fun<M>(Ynewc, Yc, rct, Ascr, Kc);
            st[0] += 1;
          }
          for (int j = 0; j < s; ++j) {
            double hc = kRosC[s * (s - 1) / 2 + j] / (dir * H);
            if (hc != 0.0)
              for (int i = 0; i < N; ++i) Kc[i] = fma(hc, Ksc(j)[i], Kc[i]);
          }
          dev_solve<M>(CCol{LU.p, 64}, Kc);
          st[6] += 1;
        }
        double err = 0.0;
        for (int i = 0; i < N; ++i) {
          double y0 = Yc[i];
          double v = y0;
          for (int j = 0; j < 4; ++j)
            if (kRosM[j] != 0.0) v = fma(kRosM[j], Ksc(j)[i], v);
          Ynew[i] = v;
          double ye = 0.0;
          for (int j = 0; j < 4; ++j)
            if (kRosE[j] != 0.0) ye = fma(kRosE[j], Ksc(j)[i], ye);
          double ymax = max_f(fabs(y0), fabs(v));
          double sc = P.vectol ? a.atol[i] + a.rtol[i] * ymax : a.atol[0] + a.rtol[0] * ymax;
          double q = ye / sc;
          err = fma(q, q, err);
        }
        err = max_f(sqrt(err / N), 1.0e-10);
        double fac = min_f(P.FacMax, max_f(P.FacMin, P.FacSafe / pow(err, 1.0 / kRosELO)));
        Hnew = H * fac;
        st[2] += 1;
        if (err <= 1.0 || H <= P.Hmin) {
          st[3] += 1;
          if (P.clip_negative)
            for (int i = 0; i < N; ++i) Y[i] = max_f(Ynew[i], 0.0);
          else
            for (int i = 0; i < N; ++i) Y[i] = Ynew[i];
          T = T + dir * H;
          Hnew = max_f(P.Hmin, min_f(Hnew, P.Hmax));
          if (rejLast) Hnew = min_f(Hnew, H);
          Hexit = H;
          Texit = T;
          HnewAcc = Hnew;
          rejLast = rejMore = false;
          H = Hnew;
          newStep = true;
        } else {
          if (rejMore) Hnew = H * P.FacRej;
          rejMore = rejLast;
          rejLast = true;
          H = Hnew;
          if (st[3] >= 1) st[4] += 1;
          newStep = false;
        }
      }
    }
    if (finished) {
      for (int s = 0; s < M::NVAR; ++s) a.conc[(int64_t)s * a.cstride + cell] = Y[s];
      if (a.istatus)
        for (int k = 0; k < 20; ++k) a.istatus[(int64_t)k * a.sstride + cell] = k < 8 ? st[k] : 0;
      if (a.rstatus) {
        for (int k = 3; k < 20; ++k) a.rstatus[(int64_t)k * a.sstride + cell] = 0.0;
        a.rstatus[cell] = Texit;
        a.rstatus[a.sstride + cell] = Hexit;
        a.rstatus[2 * a.sstride + cell] = HnewAcc;
      }
      if (a.ierr) a.ierr[cell] = ierr;
      cell = -1;
    }
  }
}

__global__ void k_zero_u32(unsigned int *p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = 0u;
}

// Component kernels (parity checks) ------------------------------------------
template <class M>
__global__ __launch_bounds__(64) void k_fun(CompArgs a) {
  const int64_t cell = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (cell >= a.ncell) return;
  CCol y{a.in + cell, a.istride};
  CCol r{a.rconst + cell, a.rstride};
  dev_fun<M>(y, y, r, Col{a.work + cell, a.wstride}, Col{a.out + cell, a.ostride});
}
template <class M>
__global__ __launch_bounds__(64) void k_jac(CompArgs a) {
  const int64_t cell = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (cell >= a.ncell) return;
  CCol y{a.in + cell, a.istride};
  CCol r{a.rconst + cell, a.rstride};
  dev_jac<M>(y, y, r, Col{a.work + cell, a.wstride}, Col{a.out + cell, a.ostride}, 1.0, 0.0);
}
template <class M>
__global__ __launch_bounds__(64) void k_decomp(CompArgs a) {
  const int64_t cell = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (cell >= a.ncell) return;
  int ier = dev_decomp<M>(Col{a.out + cell, a.ostride});
  if (a.ier) a.ier[cell] = ier;
}
template <class M>
__global__ __launch_bounds__(64) void k_solve(CompArgs a) {
  const int64_t cell = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (cell >= a.ncell) return;
  dev_solve<M>(CCol{a.in + cell, a.istride}, Col{a.out + cell, a.ostride});
}

// Host-side launchers ---------------------------------------------------------
template <class M>
static void launch_integrate(const IntegrateArgs &a, hipStream_t s) {
  dim3 grid((unsigned)((a.ncell + 63) / 64));
  hipLaunchKernelGGL(k_integrate<M>, grid, dim3(64), 0, s, a);
}
template <class M>
static void launch_comp(int which, const CompArgs &a, hipStream_t s) {
  dim3 grid((unsigned)((a.ncell + 63) / 64));
  switch (which) {
    case 0: hipLaunchKernelGGL(k_fun<M>, grid, dim3(64), 0, s, a); break;
    case 1: hipLaunchKernelGGL(k_jac<M>, grid, dim3(64), 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_decomp<M>, grid, dim3(64), 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_solve<M>, grid, dim3(64), 0, s, a); break;
  }
}

void launch_integrate_mech(int mech, const IntegrateArgs &a, hipStream_t s) {
  if (mech == MECH_FULLCHEM) launch_integrate<Traits_fullchem>(a, s);
  else if (mech == MECH_HG) launch_integrate<Traits_Hg>(a, s);
}

template <class M>
static int64_t persist_blocks_for() {
  int dev = 0, ncu = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_integrate_p<M>, 64, 0) != hipSuccess)
    return 0;
  return (int64_t)ncu * (per > 0 ? per : 1);
}
int64_t persistent_slot_fields(int mech) {
  if (mech == MECH_FULLCHEM) return SlotLayout<Traits_fullchem>::FIELDS;
  if (mech == MECH_HG) return SlotLayout<Traits_Hg>::FIELDS;
  return 0;
}
int64_t persistent_blocks(int mech) {
  if (mech == MECH_FULLCHEM) return persist_blocks_for<Traits_fullchem>();
  if (mech == MECH_HG) return persist_blocks_for<Traits_Hg>();
  return 0;
}
// nblocks waves of 64 lanes; counter null: one cell per lane (nblocks*64 >=
// ncell), else work stealing from *counter (reset to 0 here, by a kernel).
void launch_integrate_persistent_mech(int mech, const IntegrateArgs &a, int64_t nblocks,
                                      unsigned int *counter, hipStream_t s) {
  if (counter) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, counter);
  dim3 grid((unsigned)nblocks);
  if (mech == MECH_FULLCHEM)
    hipLaunchKernelGGL(k_integrate_p<Traits_fullchem>, grid, dim3(64), 0, s, a, counter);
  else if (mech == MECH_HG)
    hipLaunchKernelGGL(k_integrate_p<Traits_Hg>, grid, dim3(64), 0, s, a, counter);
}
void launch_comp_mech(int mech, int which, const CompArgs &a, hipStream_t s) {
  if (mech == MECH_FULLCHEM) launch_comp<Traits_fullchem>(which, a, s);
  else if (mech == MECH_HG) launch_comp<Traits_Hg>(which, a, s);
}

MechInfo mech_info(int mech) {
  MechInfo m{};
  if (mech == MECH_FULLCHEM) {
    using T = Traits_fullchem;
    m = {T::NSPEC, T::NVAR, T::NFIX, T::NREACT, T::LU_NZ, T::NPART};
  } else if (mech == MECH_HG) {
    using T = Traits_Hg;
    m = {T::NSPEC, T::NVAR, T::NFIX, T::NREACT, T::LU_NZ, T::NPART};
  }
  return m;
}

}  // namespace gckpp
