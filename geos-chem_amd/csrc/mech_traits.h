// Mechanism traits shared by the solver translation units: each generated
// mechanism namespace (gen/mech_<name>.h) wrapped in a struct so kernels can
// take it as a template argument, and the Rodas3 tableau.
#pragma once
#include <hip/hip_runtime.h>

#include "gen/mech_Hg.h"
#include "gen/mech_carbon.h"
#include "gen/mech_fullchem.h"

namespace gckpp {

// One program of the cell-per-wave solver (see mechgen/wave_sched.py):
// per block of four steps and lane a uint4 of term words, per block four
// flag bits (eight blocks a word), per round and lane a head word; nb
// blocks, a multiple of eight.
struct WProg {
  const uint4 *ops;
  const unsigned *flag;
  int nb;
};
enum { WP_FV = 0, WP_JV = 1, WP_LU = 2, WP_SF = 3, WP_SB = 4 };
// Mechanism-generic device pieces.  M is one of the mech_<name> namespaces,
// wrapped in a traits struct so it can be a template argument.
#define GCKPP_MECH_TRAITS(NS)                                                  \
  struct Traits_##NS {                                                         \
    static constexpr int NVAR = mech_##NS::NVAR, NFIX = mech_##NS::NFIX;       \
    static constexpr int NSPEC = mech_##NS::NSPEC, NREACT = mech_##NS::NREACT; \
    static constexpr int LU_NZ = mech_##NS::LU_NZ, NPART = mech_##NS::NPART;   \
    static constexpr int RX_MAX = mech_##NS::RX_MAX;                           \
    /* entries of the Jacobian-term and Crout-term tables */                   \
    static constexpr int NJV_TAB = sizeof(mech_##NS::jv_b) / sizeof(int);      \
    static constexpr int NCR_TAB = sizeof(mech_##NS::cr_l) / sizeof(int);      \
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
    static __device__ int cu_row(int i) { return mech_##NS::cu_row[i]; }       \
    static __device__ int cu_sp(int i) { return mech_##NS::cu_sp[i]; }         \
    static __device__ int cu_q(int i) { return mech_##NS::cu_q[i]; }           \
    static __device__ int cu_l(int i) { return mech_##NS::cu_l[i]; }           \
    static __device__ int cu_u(int i) { return mech_##NS::cu_u[i]; }           \
    static __device__ int sf_sp(int i) { return mech_##NS::sf_sp[i]; }         \
    static __device__ int sf_l(int i) { return mech_##NS::sf_l[i]; }           \
    static __device__ int sf_x(int i) { return mech_##NS::sf_x[i]; }           \
    static __device__ int sf_batch(int i) { return mech_##NS::sf_batch[i]; }   \
    static __device__ int sb_sp(int i) { return mech_##NS::sb_sp[i]; }         \
    static __device__ int sb_l(int i) { return mech_##NS::sb_l[i]; }           \
    static __device__ int sb_x(int i) { return mech_##NS::sb_x[i]; }           \
    static __device__ int sb_batch(int i) { return mech_##NS::sb_batch[i]; }   \
    static __device__ unsigned lc_lu(int s, int i) {                           \
      return s == 32 ? mech_##NS::lc32_lu[i] : mech_##NS::lc16_lu[i];          \
    }                                                                          \
    static __device__ unsigned lc_sv(int s, int i) {                           \
      return s == 32 ? mech_##NS::lc32_sv[i] : mech_##NS::lc16_sv[i];          \
    }                                                                          \
    static constexpr int NSF_BATCH = mech_##NS::NSF_BATCH;                     \
    static constexpr int NSB_BATCH = mech_##NS::NSB_BATCH;                     \
    static constexpr int WV_ZERO = mech_##NS::WV_ZERO, WV_X0 = mech_##NS::WV_X0; \
    static constexpr int WV_ONE = mech_##NS::WV_ONE, WV_SINK = mech_##NS::WV_SINK; \
    static constexpr int WV_Y0 = mech_##NS::WV_Y0, WV_SH = mech_##NS::WV_SH;   \
    static constexpr int WV_NCOEF = mech_##NS::WV_NCOEF;                       \
    static constexpr int WV_G_AB = mech_##NS::WV_G_AB, WV_NAB = mech_##NS::WV_NAB; \
    static constexpr int WV_G_F0 = mech_##NS::WV_G_F0, WV_G_K = mech_##NS::WV_G_K; \
    static constexpr int WV_G_SIZE = mech_##NS::WV_G_SIZE;                     \
    static __device__ const double *wv_coef() { return mech_##NS::wv_coef; }   \
    static __device__ WProg wprog(int p) {                                     \
      switch (p) {                                                             \
        case WP_FV: return WProg{(const uint4 *)mech_##NS::wv_fv_ops,         \
                                 mech_##NS::wv_fv_flag, \
                                 mech_##NS::WV_FV_NB};                         \
        case WP_JV: return WProg{(const uint4 *)mech_##NS::wv_jv_ops,         \
                                 mech_##NS::wv_jv_flag, \
                                 mech_##NS::WV_JV_NB};                         \
        case WP_LU: return WProg{(const uint4 *)mech_##NS::wv_lu_ops,         \
                                 mech_##NS::wv_lu_flag, \
                                 mech_##NS::WV_LU_NB};                         \
        case WP_SF: return WProg{(const uint4 *)mech_##NS::wv_sf_ops,         \
                                 mech_##NS::wv_sf_flag, \
                                 mech_##NS::WV_SF_NB};                         \
        default: return WProg{(const uint4 *)mech_##NS::wv_sb_ops,            \
                              mech_##NS::wv_sb_flag,    \
                              mech_##NS::WV_SB_NB};                            \
      }                                                                        \
    }                                                                          \
  };
GCKPP_MECH_TRAITS(fullchem)
GCKPP_MECH_TRAITS(Hg)
GCKPP_MECH_TRAITS(carbon)

// Concentration of species s (variable from y, fixed from the conc array).
template <class M, class CY, class CF>
__device__ inline double spc(const CY &y, const CF &fix, int s) {
  return s < M::NVAR ? (double)y[s] : (double)fix[s];
}

// Rodas3 (gckpp_Integrator.F90:2239-2303): stages s=1..4; A, C lower
// triangular row-wise; M, E solution/error weights; gamma(1)=0.5.
static __constant__ double kRosA[6] = {0.0, 2.0, 0.0, 2.0, 0.0, 1.0};
static __constant__ double kRosC[6] = {4.0, 1.0, -1.0, 1.0, -1.0, -(8.0 / 3.0)};
static __constant__ double kRosM[4] = {2.0, 0.0, 1.0, 1.0};
static __constant__ double kRosE[4] = {0.0, 0.0, 0.0, 1.0};
static __constant__ int kRosNewF[4] = {1, 0, 1, 1};
constexpr double kRosGamma1 = 0.5;
constexpr double kRosELO = 3.0;

}  // namespace gckpp
