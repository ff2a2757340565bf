// Update_RCONST on the GPU (KPP/fullchem/gckpp_Rates.F90:408-1503): one lane
// per cell evaluates the mechanism's generated rate program
// (gen/rates_<mech>.h) against its per-cell inputs, read field-major with
// cells fastest.  The derived temperature terms and RELHUM are formed as
// Set_Kpp_GridBox_Values does (GeosCore/fullchem_mod.F90:2089-2167).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "cell_fields.h"
#include "gckpp_internal.h"
#include "hetlaws.h"
#include "ratelaws.h"
#include "gen/rates_fullchem.h"
#include "hglaws.h"
#include "gen/rates_Hg.h"
#include "gen/rates_carbon.h"

#define NSPEC_FULLCHEM species_fullchem::NSPEC

namespace gckpp {

__device__ inline rl::RateEnv make_env(const RatesArgs &a, int64_t cell, int nspec) {
  using namespace rl;
  RateEnv E;
  const int64_t n = a.ncell;
  E.met = a.met + cell;
  E.mstride = n;
  E.photol = a.photol + cell;
  E.pstride = n;
  E.het = a.het + (a.het_stride ? cell : 0);
  E.hstride = a.het_stride ? a.het_stride : 1;
  E.conc_p = a.conc + cell;
  E.cstride = n;
  E.consts = a.consts;
  E.nspec = nspec;
  const double T = E.met[(int64_t)MET_TEMP * n];
  E.TEMP = T;
  E.PRESS = E.met[(int64_t)MET_PRESS * n];
  E.NUMDEN = E.met[(int64_t)MET_NUMDEN * n];
  E.H2O = E.met[(int64_t)MET_H2O * n];
  E.SUNCOS = E.met[(int64_t)MET_SUNCOS * n];
  E.H_PLUS_G = E.met[(int64_t)MET_H_PLUS * n];
  E.SRHOBr = E.met[(int64_t)MET_SRHOBr * n];
  E.SRHOCl = E.met[(int64_t)MET_SRHOCl * n];
  E.SRO3 = E.met[(int64_t)MET_SRO3 * n];
  E.TROP = E.met[(int64_t)MET_TROP * n];
  E.INV_TEMP = 1.0 / T;
  E.TEMP_OVER_K300 = T / 300.0;
  E.K300_OVER_TEMP = 300.0 / T;
  E.SR_TEMP = sqrt(T);
  E.FOUR_R_T = 4.0 * kConR * T;
  E.FOUR_RGASLATM_T = 4.0 * kRgasLatm * T;
  E.EIGHT_RSTARG_T = 8.0 * kRstarG * T;
  double consexp = 17.2693882 * (T - 273.16) / (T - 35.86);
  double vpresh2o = kConsvap * exp(consexp) / T;
  E.RELHUM = (E.H2O / vpresh2o) * 100.0;
  return E;
}

__global__ __launch_bounds__(64) void k_rates_fullchem(RatesArgs a) {
  const int64_t cell = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (cell >= a.ncell) return;
  rl::RateEnv E = make_env(a, cell, NSPEC_FULLCHEM);
  rates_fullchem(E, a.rconst + cell, a.ncell);
}

__global__ __launch_bounds__(64) void k_rates_Hg(RatesArgs a) {
  const int64_t cell = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (cell >= a.ncell) return;
  rl::RateEnv E = make_env(a, cell, species_Hg::NSPEC);
  rates_Hg(E, a.rconst + cell, a.ncell);
}

__global__ __launch_bounds__(64) void k_rates_carbon(RatesArgs a) {
  const int64_t cell = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (cell >= a.ncell) return;
  rl::RateEnv E = make_env(a, cell, species_carbon::NSPEC);
  rates_carbon(E, a.rconst + cell, a.ncell);
}

bool rates_available(int mech) {
  return mech == MECH_FULLCHEM || mech == MECH_HG || mech == MECH_CARBON;
}

void launch_rates_mech(int mech, const RatesArgs &a, hipStream_t s) {
  dim3 grid((unsigned)((a.ncell + 63) / 64));
  if (mech == MECH_FULLCHEM) hipLaunchKernelGGL(k_rates_fullchem, grid, dim3(64), 0, s, a);
  else if (mech == MECH_HG) hipLaunchKernelGGL(k_rates_Hg, grid, dim3(64), 0, s, a);
  else if (mech == MECH_CARBON) hipLaunchKernelGGL(k_rates_carbon, grid, dim3(64), 0, s, a);
}

}  // namespace gckpp
