// Per-cell producers on the device, upstream and downstream of the solve:
// Do_FullChem's gather of the KPP vector from the model's species and its
// scatter back (GeosCore/fullchem_mod.F90:817-821, 942-946, 1329-1350),
// fused with the unit conversions around the chemistry
// (GeosUtil/unitconv_mod.F90:1287-1391 kg/kg dry -> molec/cm3 and
// :1407-1515 back).  One thread per (species, cell); the model array
// species[nmodel][ncell] and the KPP array conc[NSPEC][ncell] are both
// cells-fastest, so every access is coalesced.
//
// fullchem_SetStateHet (KPP/fullchem/fullchem_HetStateFuncs.F90:48-244) on
// the device: the HetState record the fullchem rate laws read, filled from
// the model's per-cell fields (hetin, cell_fields.h HIN_*) and the cell's
// KPP concentrations, with the cloud parameters (rateLawUtilFuncs.F90:
// 281-417), the ice-surface coverages (fullchem_HetStateFuncs.F90:259-301)
// and the halide concentrations (:316-616, Henry's law from
// GeosUtil/henry_mod.F90:74-204).  One thread per cell.
#include <hip/hip_runtime.h>

#include "gckpp_internal.h"
#include "gen/species_fullchem.h"
#include "ratelaws.h"

namespace gckpp {
namespace {

// C(n) = 0, or the model species map[n] converted to molec/cm3 as
// ConvertSpc_KgKgDry_to_MND does it, Conc * AIRDEN * const with
// const = (AVO / MW_kg) / 1e6 (cin[n]), then times eq[n]: MW * 7e-5 for the
// sea-salt alkalinity tracers (fullchem_ConvertAlkToEquiv,
// KPP/fullchem/fullchem_SulfurChemFuncs.F90:67-78), 1 for the others;
// zero[n]: a prod/loss dummy that starts each step at zero
// (fullchem_mod.F90:942-946).
__global__ __launch_bounds__(256) void k_gather(int64_t n, const int32_t *map, const double *cin,
                                                const double *eq, const int32_t *zero,
                                                const double *species, const double *airden,
                                                double *conc) {
  const int64_t cell = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int k = blockIdx.y;
  if (cell >= n) return;
  const int s = map[k];
  double v = 0.0;
  if (s >= 0 && !(zero && zero[k])) {
    v = species[(int64_t)s * n + cell] * airden[cell] * cin[k];
    v = v * eq[k];
  }
  conc[(int64_t)k * n + cell] = v;
}

// The mapped species back into the model array: C / eq[n]
// (fullchem_ConvertEquivToAlk, fullchem_SulfurChemFuncs.F90:94-105, at
// fullchem_mod.F90:1285-1287), MAX(C, 0) (:1329-1350), then
// ConvertSpc_MND_to_KgKgDry, Conc * const / AIRDEN with
// const = 1e6 / (AVO / MW_kg) (cout[n]).
__global__ __launch_bounds__(256) void k_scatter(int64_t n, const int32_t *map, const double *cout,
                                                 const double *eq, const double *conc,
                                                 const double *airden, double *species) {
  const int64_t cell = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int k = blockIdx.y;
  if (cell >= n) return;
  const int s = map[k];
  if (s < 0) return;
  double v = conc[(int64_t)k * n + cell] / eq[k];
  v = v > 0.0 ? v : 0.0;
  species[(int64_t)s * n + cell] = v * cout[k] / airden[cell];
}

// ---- fullchem_SetStateHet --------------------------------------------------

using rl::fmax_f;
using rl::fmin_f;
using rl::safe_div;
constexpr double kAvoHet = 6.022140857e+23;  // Headers/physconstants.F90:32

// The options fullchem_SetStateHet takes from Input_Opt.
struct SetHetOpts {
  int naerotype;       // State_Chm%nAeroType
  double gamma_ho2;    // Input_Opt%gamma_HO2
  int lpscchem;        // Input_Opt%LPSCCHEM
  int turn_off_het;    // Input_Opt%TurnOffHetRates
};

// Henry's law liquid/gas ratio (fullchem_HetStateFuncs.F90:631-693 via
// CALC_KH, henry_mod.F90:74-113, and CALC_HEFF, :154-204).
__device__ double l2g_local(double k0, double cr, double pka, double tk, double h2oliq,
                            double ph) {
  const double tref = 298.15, rgas = 8.3144598, atm = 101.325;  // henry_mod.F90:52-58
  double kh = k0 * exp(cr * (1.0 / tk - 1.0 / tref)) * rgas * tk / atm;
  double heff = kh;
  if (ph > 0.0 && pka > -100.0) heff = heff * (1.0 + pow(10.0, ph - pka));
  return heff * h2oliq;
}

// Halide (or nitrate) in sea-salt water [mol/L] (:572-616).
__device__ double halide_ssa(double n_x, double area, double r_w) {
  const double v = (area * r_w / 3.0) * 1e-3;
  if (v <= 1.0e-20) return 0.0;
  return fmax_f((n_x / kAvoHet) / v, 0.0);
}

template <class Out>
__device__ void set_state_het_cell(const double *in, int64_t is, const double *c, int64_t cs,
                                   const SetHetOpts &o, Out h) {
  namespace S = species_fullchem;
  auto I = [&](int f) { return in[(int64_t)f * is]; };
  auto C = [&](int k) { return c[(int64_t)k * cs]; };
  const int na = o.naerotype;
  // meteorology (:110-114)
  const double cldf = I(HIN_CLDF);
  const double cldfr = fmin_f(fmax_f(cldf, 0.0), 1.0);
  h(HET_cldFr, cldfr);
  h(HET_clearFr, 1.0 - cldfr);
  h(HET_qIce, I(HIN_QI));
  h(HET_qLiq, I(HIN_QL));
  const double vair = I(HIN_AIRVOL) * 1.0e6;
  h(HET_vAir, vair);
  // aerosol (:117-129); entries past nAeroType are zero
  h(HET_NAEROTYPE, (double)na);
  const double aclarea = I(HIN_aClArea), aclradi = I(HIN_aClRadi);
  h(HET_aClArea, aclarea);
  h(HET_aClRadi, aclradi);
  h(HET_aClVol, aclarea * aclradi / 3.0);
  h(HET_aWater, I(HIN_IsorropAeroH2O));
  h(HET_aWater + 1, I(HIN_IsorropAeroH2O + 1));
  for (int k = 0; k < 14; ++k) {
    const bool on = k < na;
    const double xa = on ? I(HIN_AeroArea + k) : 0.0, xr = on ? I(HIN_AeroRadi + k) : 0.0;
    h(HET_xArea + k, xa);
    h(HET_xRadi + k, xr);
    h(HET_xVol + k, on ? xa * xr / 3.0 : 0.0);
    h(HET_wetArea + k, on ? I(HIN_WetAeroArea + k) : 0.0);
    h(HET_xH2O + k, on ? I(HIN_AeroH2O + k) * 1.0e-6 : 0.0);
  }
  h(HET_OMOC_POA, I(HIN_OMOC_POA));
  h(HET_OMOC_OPOA, I(HIN_OMOC_OPOA));
  // cloud S(IV) (:132-136)
  const double hso3 = I(HIN_HSO3_aq), so3 = I(HIN_SO3_aq);
  const double tso3 = hso3 + so3;
  h(HET_HSO3_aq, hso3);
  h(HET_SO3_aq, so3);
  h(HET_TSO3_aq, tso3);
  h(HET_frac_HSO3_aq, safe_div(hso3, tso3, 0.0));
  h(HET_frac_SO3_aq, safe_div(so3, tso3, 0.0));
  // ISORROPIA/HETP (:139-141), pH and alkalinity (:144-168)
  h(HET_HSO4_molal, I(HIN_IsorropBisulfate));
  h(HET_NO3_molal, I(HIN_IsorropNitrate));
  h(HET_SO4_molal, I(HIN_IsorropSulfate));
  h(HET_H_plus, I(HIN_IsorropHplus));
  const double phcld = I(HIN_pHCloud), phssa1 = I(HIN_IsorropAeropH);
  h(HET_pHCloud, phcld);
  h(HET_pHSSA, phssa1);
  h(HET_pHSSA + 1, I(HIN_IsorropAeropH + 1));
  const double hsul = pow(10.0, -1.0 * phssa1);
  h(HET_H_conc_Sul, hsul);
  h(HET_H_conc_LCl, pow(10.0, -1.0 * phcld));
  h(HET_H_conc_ICl, pow(10.0, -4.5));
  h(HET_H_conc_SSA, hsul);
  h(HET_H_conc_SSC, pow(10.0, -5.0));
  double fa = safe_div(I(HIN_SALAAL), I(HIN_SALA), 0.0);
  fa = fmax_f(fmin_f(fa, 1.0), 0.0);
  double fc = safe_div(I(HIN_SALCAL), I(HIN_SALC), 0.0);
  fc = fmax_f(fmin_f(fc, 1.0), 0.0);
  h(HET_f_Alk_SSA, fa);
  h(HET_f_Acid_SSA, 1.0 - fa);
  h(HET_f_Alk_SSC, fc);
  h(HET_f_Acid_SSC, 1.0 - fc);
  const bool ssa_alk = fabs(fa) > 0.01, ssc_alk = fabs(fc) > 0.01;
  h(HET_SSA_is_Alk, ssa_alk ? 1.0 : 0.0);
  h(HET_SSA_is_Acid, ssa_alk ? 0.0 : 1.0);
  h(HET_SSC_is_Alk, ssc_alk ? 1.0 : 0.0);
  h(HET_SSC_is_Acid, ssc_alk ? 0.0 : 1.0);
  h(HET_gamma_HO2, o.gamma_ho2);
  h(HET_fupdateHOBr, I(HIN_fupdateHOBr));
  h(HET_fupdateHOCl, I(HIN_fupdateHOCl));
  // HSO3-/SO2 and SO3--/SO2 in cloud water (:182-209)
  const double ql = I(HIN_QL), airden = I(HIN_AIRDEN), so2 = C(S::SO2);
  double hso3m = safe_div(hso3 * 1.0e-3 * kAvoHet * ql * airden * 1.0e-3, cldf, 0.0);
  h(HET_HSO3m, safe_div(hso3m, so2, 0.0));
  double so3mm = safe_div(so3 * 1.0e-3 * kAvoHet * ql * airden * 1.0e-3, cldf, 0.0);
  h(HET_SO3mm, safe_div(so3mm, so2, 0.0));
  // Cld_Params (rateLawUtilFuncs.F90:281-417)
  const double qi = I(HIN_QI), t = I(HIN_T), ad = I(HIN_AD);
  double rliq, rice, aliq, vliq, aice, vice;
  if (ql + qi <= 0.0 || cldf <= 0.0) {
    rliq = 6.0e-4;
    rice = 38.5e-4;
    aliq = vliq = aice = vice = 0.0;
  } else {
    rliq = I(HIN_FRLAND) > I(HIN_FROCEAN) ? 6.0e-4 : 10.0e-4;
    vliq = ql * ad / 0.001 / vair;
    vice = qi * ad / 0.91e-3 / vair;
    aliq = 3.0 * vliq / rliq;
    double alpha, beta;
    if (t < 202.0) {
      alpha = 83.3;
      beta = 0.0184;
    } else if (t < 217.0) {
      alpha = 9.1744e+4;
      beta = 0.117;
    } else {
      alpha = 308.4;
      beta = 0.0152;
    }
    rice = 0.5 * alpha * exp(beta * (t - 273.15)) / 1e+4;
    aice = 3.0 * vice / rice * 2.25;
  }
  h(HET_rLiq, rliq);
  h(HET_rIce, rice);
  h(HET_aLiq, aliq);
  h(HET_vLiq, vliq);
  h(HET_aIce, aice);
  h(HET_vIce, vice);
  // Get_Theta_Ice (:259-301); TEMP is the cell's temperature
  {
    const double hno3 = C(S::HNO3), hcl = C(S::HCl), hbr = C(S::HBr);
    double klin = 7.5e-5 * exp(4585.0 / t);
    const double k1 = klin / 2.7e+14;
    klin = 1.3e-5 * exp(4600.0 / t);
    const double k2 = klin / 3.0e+14;
    const double den = 1.0 + k1 * hno3 + k2 * hcl;
    h(HET_HNO3_theta, k1 * hno3 / den);
    h(HET_HCl_theta, k2 * hcl / den);
    h(HET_HBr_theta, fmin_f(4.14e-10 * pow(hbr, 0.88), 1.0));
  }
  // Halide_Conc (:316-472) with Get_Halide_CldConc (:487-558)
  {
    const double cbr = C(S::HBr), cbra = C(S::BrSALA), cbrc = C(S::BrSALC);
    const double ccl = C(S::HCl), ccla = C(S::SALACL), cclc = C(S::SALCCL);
    const double hbr = cbr + (cbra * 0.7) + cbrc;
    const double hcl = ccl + (ccla * 0.7) + cclc;
    double br_conc = 0.0, cl_conc = 0.0;
    const double vtot = safe_div(vliq, cldfr, 0.0);
    if (!(vtot < 1.0e-20)) {
      double l2g = l2g_local(1.0, 9000.0, -6.3, t, vtot, phcld);
      double fl = l2g / (1.0 + l2g);
      cl_conc = fl * hcl / (vtot * kAvoHet * 1.0e-3);
      l2g = l2g_local(7.5e-1, 10200.0, -9.0, t, vtot, phcld);
      fl = l2g / (1.0 + l2g);
      br_conc = fl * hbr / (vtot * kAvoHet * 1.0e-3);
    }
    double brg = 0.0, bra = 0.0, brc = 0.0, clg = 0.0, cla = 0.0, clc = 0.0;
    double den = cbr + (cbra * 0.7) + cbrc;
    if (den > 0.0) {
      brg = (br_conc * cbr) / den;
      bra = (br_conc * cbra * 0.7) / den;
      brc = (br_conc * cbrc) / den;
    }
    den = ccl + (ccla * 0.7) + cclc;
    if (den > 0.0) {
      clg = (cl_conc * ccl) / den;
      cla = (cl_conc * ccla * 0.7) / den;
      clc = (cl_conc * cclc) / den;
    }
    const double brcld = bra + brc + brg, clcld = cla + clc + clg;
    h(HET_Br_conc_CldG, brg);
    h(HET_Br_conc_CldA, bra);
    h(HET_Br_conc_CldC, brc);
    h(HET_Cl_conc_CldG, clg);
    h(HET_Cl_conc_CldA, cla);
    h(HET_Cl_conc_CldC, clc);
    h(HET_Br_conc_Cld, brcld);
    h(HET_Cl_conc_Cld, clcld);
    const bool bpos = brcld > 0.0, cpos = clcld > 0.0;
    h(HET_frac_Br_CldA, bpos ? bra / brcld : 0.0);
    h(HET_frac_Br_CldC, bpos ? brc / brcld : 0.0);
    h(HET_frac_Br_CldG, bpos ? brg / brcld : 0.0);
    h(HET_frac_Cl_CldA, cpos ? cla / clcld : 0.0);
    h(HET_frac_Cl_CldC, cpos ? clc / clcld : 0.0);
    h(HET_frac_Cl_CldG, cpos ? clg / clcld : 0.0);
    const double a12 = na >= 12 ? I(HIN_AeroArea + 11) : 0.0;
    const double r12 = na >= 12 ? I(HIN_AeroRadi + 11) : 0.0;
    const double brssa = halide_ssa(cbra, aclarea, aclradi);
    const double brssc = halide_ssa(cbrc, a12, r12);
    const double clssa = halide_ssa(ccla, aclarea, aclradi);
    const double clssc = halide_ssa(cclc, a12, r12);
    h(HET_Br_conc_SSA, brssa);
    h(HET_Br_conc_SSC, brssc);
    h(HET_Cl_conc_SSA, clssa);
    h(HET_Cl_conc_SSC, clssc);
    h(HET_NIT_conc_SSA, halide_ssa(C(S::NIT), aclarea, aclradi));
    h(HET_NIT_conc_SSC, halide_ssa(C(S::NITs), a12, r12));
    h(HET_Br_over_Cl_Cld, cpos ? brcld / clcld : 0.0);
    h(HET_Br_over_Cl_SSA, clssa > 0.0 ? brssa / clssa : 0.0);
    h(HET_Br_over_Cl_SSC, clssc > 0.0 ? brssc / clssc : 0.0);
    const double sal = ccla + C(S::NIT) + C(S::SO4);
    h(HET_frac_SALACL, sal > 0.0 ? ccla / sal : 0.0);
  }
  // PSC and stratosphere flags (:224-241)
  for (int k = 0; k < 11; ++k) h(HET_KHETI_SLA + k, I(HIN_KHETI_SLA + k));
  const bool strat = I(HIN_InStratosphere) != 0.0;
  const bool psc = o.lpscchem != 0 && I(HIN_STATE_PSC) >= 2.0 && strat;
  h(HET_stratBox, strat ? 1.0 : 0.0);
  h(HET_pscBox, psc ? 1.0 : 0.0);
  h(HET_natSurface, psc && C(S::NIT) > 0.0 ? 1.0 : 0.0);
  h(HET_TurnOffHetRates, o.turn_off_het ? 1.0 : 0.0);
  // declared in HetState, set by no caller of the fullchem path
  h(HET_H2O, 0.0);
  h(HET_cloudBox, 0.0);
  h(HET_fracOrgAer, 0.0);
  h(HET_fracInorgAer, 0.0);
}

__global__ __launch_bounds__(256) void k_set_state_het(int64_t n, const double *hetin,
                                                       const double *conc, double *het,
                                                       SetHetOpts o) {
  const int64_t cell = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (cell >= n) return;
  set_state_het_cell(hetin + cell, n, conc + cell, n, o,
                     [&](int f, double v) { het[(int64_t)f * n + cell] = v; });
}

// ---- fullchem_SulfurAqChem --------------------------------------------------
// K_MT(1:6), the sea-salt alkalinity uptake rates of SO2, HCl and HNO3 on
// fine (AeroRadi/WetAeroArea type 11) and coarse (type 12) sea salt
// (KPP/fullchem/fullchem_SulfurChemFuncs.F90:120-305), from the KPP vector
// after fullchem_ConvertAlkToEquiv (SALAAL, SALCAL in equivalents) and
// Set_Kpp_GridBox_Values' NUMDEN and SR_TEMP (fullchem_mod.F90:2089-2167).
// The SO2 rates are per O3 molecule (the mechanism multiplies by O3).
__global__ __launch_bounds__(256) void k_sulfur_aq(int64_t n, const double *hetin,
                                                   const double *conc, const double *mw,
                                                   double *met) {
  namespace S = species_fullchem;
  const int64_t cell = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (cell >= n) return;
  auto I = [&](int f) { return hetin[(int64_t)f * n + cell]; };
  auto C = [&](int k) { return conc[(int64_t)k * n + cell]; };
  rl::RateEnv E{};
  E.TEMP = met[(int64_t)MET_TEMP * n + cell];
  E.NUMDEN = met[(int64_t)MET_NUMDEN * n + cell];
  E.SR_TEMP = sqrt(E.TEMP);
  const double salaal = C(S::SALAAL), salcal = C(S::SALCAL), o3 = C(S::O3);
  const bool aa = salaal > 0.1, ca = salcal > 0.1, o3ok = o3 > 1.0e+10;
  const double srso2 = sqrt(mw[S::SO2]), srhcl = sqrt(mw[S::HCl]), srhno3 = sqrt(mw[S::HNO3]);
  double k[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int f = 0; f < 2; ++f) {   // fine (type 11), coarse (type 12)
    const bool on = f == 0 ? aa : ca;
    if (!on) continue;
    const double alk = f == 0 ? salaal : salcal;
    const double area = I(HIN_WetAeroArea + 10 + f), radius = I(HIN_AeroRadi + 10 + f);
    if (o3ok) {
      const double kex = rl::Ars_L1k(E, area, radius, 0.11, srso2);
      k[3 * f] = rl::kIIR1Ltd(C(S::SO2), alk, kex) / o3;
    }
    double kex = rl::Ars_L1k(E, area, radius, 0.07, srhcl);
    k[3 * f + 1] = rl::kIIR1Ltd(C(S::HCl), alk, kex);
    kex = rl::Ars_L1k(E, area, radius, 0.5, srhno3);
    k[3 * f + 2] = rl::kIIR1Ltd(C(S::HNO3), alk, kex);
  }
  for (int j = 0; j < 6; ++j) met[(int64_t)(MET_K_MT + j) * n + cell] = k[j];
}

}  // namespace

void launch_sulfur_aq(int64_t n, const double *hetin, const double *conc, const double *consts,
                      double *met, hipStream_t s) {
  hipLaunchKernelGGL(k_sulfur_aq, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, hetin,
                     conc, consts, met);
}

void launch_set_state_het(int64_t n, const double *hetin, const double *conc, double *het,
                          int naerotype, double gamma_ho2, int lpscchem, int turn_off_het,
                          hipStream_t s) {
  SetHetOpts o{naerotype, gamma_ho2, lpscchem, turn_off_het};
  hipLaunchKernelGGL(k_set_state_het, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n,
                     hetin, conc, het, o);
}

void launch_gather(int64_t n, int nspec, const int32_t *map, const double *cin, const double *eq,
                   const int32_t *zero, const double *species, const double *airden, double *conc,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_gather, dim3((unsigned)((n + 255) / 256), (unsigned)nspec), dim3(256), 0, s,
                     n, map, cin, eq, zero, species, airden, conc);
}
void launch_scatter(int64_t n, int nspec, const int32_t *map, const double *cout, const double *eq,
                    const double *conc, const double *airden, double *species, hipStream_t s) {
  hipLaunchKernelGGL(k_scatter, dim3((unsigned)((n + 255) / 256), (unsigned)nspec), dim3(256), 0, s,
                     n, map, cout, eq, conc, airden, species);
}

}  // namespace gckpp
