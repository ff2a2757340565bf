// Heterogeneous rate laws of the Hg mechanism, restated from
// KPP/Hg/Hg_RateLawFuncs.F90:67-158: uptake of Hg(II) on cloud droplets with
// entrainment limitation, split into organic / inorganic aerosol fractions.
// The gas-phase laws (GCARR_*, GCJPLPR_abab) are the shared ones of
// ratelaws.h, whose formulas the Hg module repeats.
#pragma once
#include "ratelaws.h"

namespace gckpp {
namespace rl {

// CloudHet_Hg(H, srMw, gamma), Hg_RateLawFuncs.F90:67-118
GCKPP_HD double CloudHet_Hg(const RateEnv &E, double srMw, double gamma) {
  double cldfr = E.h(HET_cldFr), aliq = E.h(HET_aLiq), aice = E.h(HET_aIce);
  if (cldfr < 0.0001 || aliq + aice <= 0.0) return 0.0;
  double kic = Ars_L1k(E, aliq, E.h(HET_rLiq), gamma, srMw);
  double kEinv = safe_div(3600.0 * E.h(HET_clearFr), cldfr, 1e+30);
  double kIinv = safe_div(1.0, cldfr * kic, 1e+30);
  return safe_div(1.0, kEinv + kIinv, 0.0);
}
// Het_HgIIP_Org / _Inorg(H, gamma, srMw), Hg_RateLawFuncs.F90:121-158.  The
// reference passes its (gamma, srMw) to CloudHet_Hg in that order, i.e. into
// CloudHet_Hg's (srMw, gamma) slots; the same argument order is kept here so
// the rates agree with the reference's.
GCKPP_HD double Het_HgIIP_Org(const RateEnv &E, double gamma, double srMw) {
  if (!E.hb(HET_cloudBox)) return 0.0;
  return CloudHet_Hg(E, gamma, srMw) * E.h(HET_fracOrgAer);
}
GCKPP_HD double Het_HgIIP_Inorg(const RateEnv &E, double gamma, double srMw) {
  if (!E.hb(HET_cloudBox)) return 0.0;
  return CloudHet_Hg(E, gamma, srMw) * E.h(HET_fracInorgAer);
}

}  // namespace rl
}  // namespace gckpp
