// Rate-coefficient laws of the fullchem mechanism, evaluated per cell on the
// GPU by the generated rate program (gen/rates_<mech>.h).  Restated from the
// reference's rate-law library; each group cites the routines it follows:
//   KPP/fullchem/rateLawUtilFuncs.F90       Arrhenius, uptake helpers, safe math
//   KPP/fullchem/fullchem_RateLawFuncs.F90  gas-phase (:107-790) and
//                                           heterogeneous (:795-3460) laws
// Per-cell globals (TEMP, NUMDEN, ..., KPP/fullchem/commonIncludeVars.H) come
// from RateEnv, which the rates kernel fills as Set_Kpp_GridBox_Values does
// (GeosCore/fullchem_mod.F90:2089-2167).  FP64 throughout.
#pragma once
#include <cmath>
#include <cstdint>

#include "cell_fields.h"

#ifndef GCKPP_HD
#define GCKPP_HD __host__ __device__ inline
#endif

namespace gckpp {
namespace rl {

// Headers/physconstants.F90:26-90; CON_R (commonIncludeVars.H:7)
constexpr double kBoltz = 1.38064852e-23;
constexpr double kConsvap = 6.1078e+03 / (kBoltz * 1e+7);
constexpr double kRgasLatm = 8.2057e-2;
constexpr double kRstarG = 8.3144598;
constexpr double kConR = 0.083144598;
constexpr double kAvo = 6.022140857e+23;
constexpr double kPi = 3.14159265358979323;
// fullchem_RateLawFuncs.F90:62-68; rateLawUtilFuncs.F90:30-31
constexpr double kCritRH = 35.0;
constexpr double kConAtmBar = 1.0 / 1.01325;
constexpr double kInvT298 = 1.0 / 298.15;
constexpr double kHetMinLife = 1.e-3;
constexpr double kHetMinRate = 1.0 / kHetMinLife;

// aerosol types in HetState arrays (fullchem_RateLawFuncs.F90:29-42), 0-based
enum { DU1 = 0, DU2, DU3, DU4, DU5, DU6, DU7, SUL, BKC, ORC, SSA, SSC, SLA, IIC };
// KHETI_SLA slots (fullchem_RateLawFuncs.F90:49-59), 0-based
enum {
  N2O5_plus_H2O = 0, N2O5_plus_HCl, ClNO3_plus_H2O, ClNO3_plus_HCl, ClNO3_plus_HBr,
  BrNO3_plus_H2O, BrNO3_plus_HCl, HOCl_plus_HCl, HOCl_plus_HBr, HOBr_plus_HCl, HOBr_plus_HBr
};

// Everything a rate law of one cell may read.
struct RateEnv {
  // Set_Kpp_GridBox_Values (fullchem_mod.F90:2089-2167)
  double TEMP, INV_TEMP, TEMP_OVER_K300, K300_OVER_TEMP, SR_TEMP;
  double FOUR_R_T, FOUR_RGASLATM_T, EIGHT_RSTARG_T;
  double NUMDEN, H2O, PRESS, RELHUM, SUNCOS;
  double SRHOBr, SRHOCl, SRO3, H_PLUS_G, TROP;
  const double *met, *photol, *het, *conc_p, *consts;
  int64_t mstride, pstride, hstride, cstride, nspec;
  GCKPP_HD double PHOTOL(int i) const { return photol[(int64_t)i * pstride]; }
  GCKPP_HD double K_MT(int i) const { return met[(int64_t)(MET_K_MT + i) * mstride]; }
  GCKPP_HD double K_CLD(int i) const { return met[(int64_t)(MET_K_CLD + i) * mstride]; }
  GCKPP_HD double K_TROP(int i) const { return met[(int64_t)(MET_K_TROP + i) * mstride]; }
  GCKPP_HD double K_STRAT(int i) const { return met[(int64_t)(MET_K_STRAT + i) * mstride]; }
  GCKPP_HD double h(int f) const { return het[(int64_t)f * hstride]; }
  GCKPP_HD bool hb(int f) const { return het[(int64_t)f * hstride] != 0.0; }
  GCKPP_HD double conc(int s) const { return conc_p[(int64_t)s * cstride]; }
  GCKPP_HD double mw(int s) const { return consts[s]; }
  GCKPP_HD double srmw(int s) const { return sqrt(consts[s]); }
  GCKPP_HD double henry_k0(int s) const { return consts[nspec + s]; }
  GCKPP_HD double henry_cr(int s) const { return consts[2 * nspec + s]; }
  // HetState shorthands
  GCKPP_HD double xArea(int t) const { return h(HET_xArea + t); }
  GCKPP_HD double xRadi(int t) const { return h(HET_xRadi + t); }
  GCKPP_HD double xVol(int t) const { return h(HET_xVol + t); }
  GCKPP_HD double xH2O(int t) const { return h(HET_xH2O + t); }
  GCKPP_HD double kheti(int i) const { return h(HET_KHETI_SLA + i); }
  GCKPP_HD bool strat() const { return hb(HET_stratBox); }
  GCKPP_HD bool nat() const { return hb(HET_natSurface); }
  GCKPP_HD bool off() const { return hb(HET_TurnOffHetRates); }
};

// ---- Fortran intrinsic semantics the reference relies on --------------------
// MIN/MAX as the reference's compiler evaluates them: the first argument
// unless the second compares strictly smaller (larger), so a NaN first
// argument yields the second and a NaN second argument is returned.
GCKPP_HD double fmin_f(double a, double b) { return a < b ? a : b; }
GCKPP_HD double fmax_f(double a, double b) { return a > b ? a : b; }
// EXPONENT(x): x = f*2^e with f in [0.5,1); 0 for 0, HUGE(0) for NaN/Inf.
GCKPP_HD int fexponent(double x) {
  if (x == 0.0) return 0;
  if (!__builtin_isfinite(x)) return 2147483647;
  int e = 0;
  frexp(x, &e);
  return e;
}
// EXPONENT(a) - EXPONENT(b) in default (32-bit, wrapping) integers
GCKPP_HD int expdiff(double a, double b) {
  return (int)(int32_t)((uint32_t)fexponent(a) - (uint32_t)fexponent(b));
}

// ---- safe math (rateLawUtilFuncs.F90:459-520) ------------------------------
GCKPP_HD double safe_div(double num, double den, double alt) {
  int ed = expdiff(num, den);
  if (ed > 1023 || den == 0.0) return alt;
  if (ed < -1020) return 0.0;
  return num / den;
}
GCKPP_HD bool is_safe_div(double num, double den) {
  int ed = expdiff(num, den);
  return !(ed < -1020 || ed > 1023 || den == 0.0);
}
GCKPP_HD double pow_int(double x, int n) {  // Fortran x**n, integer n
  bool neg = n < 0;
  if (neg) n = -n;
  double r = 1.0, b = x;
  while (n) {
    if (n & 1) r *= b;
    b *= b;
    n >>= 1;
  }
  return neg ? 1.0 / r : r;
}

// ---- Arrhenius forms (rateLawUtilFuncs.F90:41-72) --------------------------
GCKPP_HD double GCARR_ab(const RateEnv &E, double a0, double b0) {
  return a0 * pow(E.K300_OVER_TEMP, b0);
}
GCKPP_HD double GCARR_ac(const RateEnv &E, double a0, double c0) { return a0 * exp(c0 / E.TEMP); }
GCKPP_HD double GCARR_abc(const RateEnv &E, double a0, double b0, double c0) {
  return a0 * exp(c0 / E.TEMP) * pow(E.K300_OVER_TEMP, b0);
}

// ---- gas-phase laws (fullchem_RateLawFuncs.F90:107-790) --------------------
GCKPP_HD double ARRPLUS_ade(const RateEnv &E, double a0, double d0, double e0) {
  return fmax_f(a0 * (d0 + E.TEMP * e0), 0.0);
}
GCKPP_HD double ARRPLUS_abde(const RateEnv &E, double a0, double b0, double d0, double e0) {
  return fmax_f(a0 * (d0 + E.TEMP * e0) * exp(-b0 / E.TEMP), 0.0);
}
GCKPP_HD double TUNPLUS_abcde(const RateEnv &E, double a0, double b0, double c0, double d0,
                              double e0) {
  double k = a0 * (d0 + E.TEMP * e0);
  k = k * exp(b0 / E.TEMP) * exp(c0 / pow_int(E.TEMP, 3));
  return fmax_f(k, 0.0);
}
// isoprene + OH branching (GC_ISO1 / GC_ISO2): fraction k2 of the nitrate-
// forming channel, applied as (1-k2) or k2
GCKPP_HD double iso_frac(const RateEnv &E, double c0, double d0, double e0, double f0, double g0) {
  double k0 = d0 * exp(e0 / E.TEMP) * exp(1.0e8 / pow_int(E.TEMP, 3));
  double k1 = f0 * exp(g0 / E.TEMP);
  return c0 * k0 / (k0 + k1);
}
GCKPP_HD double GC_ISO1(const RateEnv &E, double a0, double b0, double c0, double d0, double e0,
                        double f0, double g0) {
  return a0 * exp(b0 / E.TEMP) * (1.0 - iso_frac(E, c0, d0, e0, f0, g0));
}
GCKPP_HD double GC_ISO2(const RateEnv &E, double a0, double b0, double c0, double d0, double e0,
                        double f0, double g0) {
  return a0 * exp(b0 / E.TEMP) * iso_frac(E, c0, d0, e0, f0, g0);
}
GCKPP_HD double GC_EPO_a(const RateEnv &E, double a1, double e1, double m1) {
  return a1 * exp(e1 / E.TEMP) * (1.0 / (m1 * E.NUMDEN + 1.0));
}
// Troe-type PAN falloff with broadening cf
GCKPP_HD double pan_falloff(double k0, double k1, double cf) {
  double kr = k0 / k1;
  double nc = 0.75 - 1.27 * log10(cf);
  double q = log10(kr) / nc;
  double f = pow(10.0, log10(cf) / (1.0 + q * q));
  return k0 * k1 * f / (k0 + k1);
}
GCKPP_HD double GC_PAN_abab(const RateEnv &E, double a0, double b0, double a1, double b1,
                            double cf) {
  double k0 = a0 * exp(b0 / E.TEMP) * E.NUMDEN;
  double k1 = a1 * exp(b1 / E.TEMP);
  return pan_falloff(k0, k1, cf);
}
GCKPP_HD double GC_PAN_acac(const RateEnv &E, double a0, double c0, double a1, double c1,
                            double cf) {
  double k0 = a0 * pow(E.TEMP_OVER_K300, c0) * E.NUMDEN;
  double k1 = a1 * pow(E.TEMP_OVER_K300, c1);
  return pan_falloff(k0, k1, cf);
}
// alkyl-nitrate branching (GC_NIT takes the nitrate fraction k2/(k2+c0),
// GC_ALK the complementary c0/(k2+c0))
GCKPP_HD double nit_k2(const RateEnv &E, double n) {
  double k0 = 2.0E-22 * exp(n) * E.NUMDEN;
  double k1 = k0 / (4.3E-1 * pow_int(E.TEMP / 298.0, -8));
  double lk = log10(k1);
  return (k0 / (1.0 + k1)) * pow(4.1E-1, 1.0 / (1.0 + lk * lk));
}
GCKPP_HD double GC_NIT(const RateEnv &E, double a0, double b0, double c0, double n, double x0,
                       double y0) {
  double k2 = nit_k2(E, n);
  double k = a0 * (x0 - E.TEMP * y0) * exp(b0 / E.TEMP) * (k2 / (k2 + c0));
  return fmax_f(k, 0.0);
}
GCKPP_HD double GC_ALK(const RateEnv &E, double a0, double b0, double c0, double n, double x0,
                       double y0) {
  double k2 = nit_k2(E, n);
  double k = a0 * (x0 - E.TEMP * y0) * exp(b0 / E.TEMP) * (c0 / (k2 + c0));
  return fmax_f(k, 0.0);
}
GCKPP_HD double GC_HO2HO2_acac(const RateEnv &E, double a0, double c0, double a1, double c1) {
  double k0 = a0 * exp(c0 / E.TEMP), k1 = a1 * exp(c1 / E.TEMP);
  return (k0 + k1 * E.NUMDEN) * (1.0 + 1.4E-21 * E.H2O * exp(2200.0 / E.TEMP));
}
GCKPP_HD double GC_TBRANCH_1_acac(const RateEnv &E, double a0, double c0, double a1, double c1) {
  return a0 * exp(c0 / E.TEMP) / (1.0 + a1 * exp(c1 / E.TEMP));
}
GCKPP_HD double GC_RO2HO2_aca(const RateEnv &E, double a0, double c0, double a1) {
  return a0 * exp(c0 / E.TEMP) * (1.0 - exp(-0.245 * a1));
}
GCKPP_HD double GC_DMSOH_acac(const RateEnv &E, double a0, double c0, double a1, double c1) {
  double k0 = a0 * exp(c0 / E.TEMP), k1 = a1 * exp(c1 / E.TEMP);
  return (k0 * E.NUMDEN * 0.2095e0) / (1.0 + k1 * 0.2095e0);
}
GCKPP_HD double GC_GLYXNO3_ac(const RateEnv &E, double a0, double c0) {
  double o2 = E.NUMDEN * 0.2095;
  return a0 * exp(c0 / E.TEMP) * (o2 + 3.5E+18) / (2.0 * o2 + 3.5E+18);
}
GCKPP_HD double glyc_frac(const RateEnv &E) {
  return fmax_f(1.0 - 11.0729 * exp((-1.0 / 73.0) * E.TEMP), 0.0);
}
GCKPP_HD double GC_GLYCOH_A_a(const RateEnv &E, double a0) { return a0 * glyc_frac(E); }
GCKPP_HD double GC_GLYCOH_B_a(const RateEnv &E, double a0) { return a0 * (1.0 - glyc_frac(E)); }
GCKPP_HD double hac_frac(const RateEnv &E) {
  return fmax_f(1.0 - 23.7 * exp((-1.0 / 60.0) * E.TEMP), 0.0);
}
GCKPP_HD double GC_HACOH_A_ac(const RateEnv &E, double a0, double c0) {
  return a0 * exp(c0 / E.TEMP) * hac_frac(E);
}
GCKPP_HD double GC_HACOH_B_ac(const RateEnv &E, double a0, double c0) {
  return a0 * exp(c0 / E.TEMP) * (1.0 - hac_frac(E));
}
GCKPP_HD double GC_RO2NO_A1_ac(const RateEnv &E, double a0, double c0) {
  return a0 * exp(c0 / E.TEMP) * 3.0e-4;
}
GCKPP_HD double GC_RO2NO_B1_ac(const RateEnv &E, double a0, double c0) {
  return a0 * exp(c0 / E.TEMP) * (1.0 - 3.0e-4);
}
// RO2 + NO nitrate yield (Arey-type falloff in carbon number a1)
GCKPP_HD double ro2no_yield(const RateEnv &E, double a1) {
  double xxyn = 1.94e-22 * exp(0.97 * a1) * E.NUMDEN;
  double yyyn = 0.826 * pow(E.K300_OVER_TEMP, 8.1);
  double aaa = log10(xxyn / yyyn);
  double zzyn = 1.0 / (1.0 + aaa * aaa);
  double rarb = (xxyn / (1.0 + xxyn / yyyn)) * pow(0.411, zzyn);
  return rarb / (1.0 + rarb);
}
GCKPP_HD double GC_RO2NO_A2_aca(const RateEnv &E, double a0, double c0, double a1) {
  return a0 * exp(c0 / E.TEMP) * ro2no_yield(E, a1);
}
GCKPP_HD double GC_RO2NO_B2_aca(const RateEnv &E, double a0, double c0, double a1) {
  return a0 * exp(c0 / E.TEMP) * (1.0 - ro2no_yield(E, a1));
}
// JPL termolecular falloff: low-pressure rlow (already x [M]), high rhigh
GCKPP_HD double jpl_falloff(double rlow, double rhigh, double fv) {
  double xyrat = rlow / rhigh;
  double blog = log10(xyrat);
  double fexp = 1.0 / (1.0 + blog * blog);
  return rlow * pow(fv, fexp) / (1.0 + xyrat);
}
GCKPP_HD double GCJPLPR_aa(const RateEnv &E, double a1, double a2, double fv) {
  return jpl_falloff(a1 * E.NUMDEN, a2, fv);
}
GCKPP_HD double GCJPLPR_aba(const RateEnv &E, double a1, double b1, double a2, double fv) {
  return jpl_falloff(a1 * pow(E.K300_OVER_TEMP, b1) * E.NUMDEN, a2, fv);
}
GCKPP_HD double GCJPLPR_abab(const RateEnv &E, double a1, double b1, double a2, double b2,
                             double fv) {
  return jpl_falloff(a1 * pow(E.K300_OVER_TEMP, b1) * E.NUMDEN, a2 * pow(E.K300_OVER_TEMP, b2), fv);
}
GCKPP_HD double GCJPLPR_abcabc(const RateEnv &E, double a1, double b1, double c1, double a2,
                               double b2, double c2, double fv) {
  double rlow = a1 * pow(E.K300_OVER_TEMP, b1) * exp(c1 / E.TEMP) * E.NUMDEN;
  double rhigh = a2 * pow(E.K300_OVER_TEMP, b2) * exp(c2 / E.TEMP);
  return jpl_falloff(rlow, rhigh, fv);
}
GCKPP_HD double GCJPLEQ_acabab(const RateEnv &E, double a0, double c0, double a1, double b1,
                               double a2, double b2, double fv) {
  return GCJPLPR_abab(E, a1, b1, a2, b2, fv) / (a0 * exp(c0 / E.TEMP));
}
// chemical-activation channel: k3 * (1 - k_falloff/k_high)
GCKPP_HD double GCJPLAC_ababac(const RateEnv &E, double a1, double b1, double a2, double b2,
                               double a3, double c3, double fv) {
  double rhigh = a2 * pow(E.K300_OVER_TEMP, b2);
  double k1 = jpl_falloff(a1 * pow(E.K300_OVER_TEMP, b1) * E.NUMDEN, rhigh, fv);
  return a3 * exp(c3 / E.TEMP) * (1.0 - k1 / rhigh);
}

// ---- uptake helpers (rateLawUtilFuncs.F90:77-140,423-455) ------------------
// First-order uptake rate on an aerosol of given area [cm2/cm3], radius [cm]
// and reaction probability gamma (gas diffusion + free-molecular limits).
GCKPP_HD double Ars_L1k(const RateEnv &E, double area, double radius, double gamma, double srMw) {
  if (gamma < 1.0e-30 || radius < 1.0e-30) return 0.0;
  double dfkg = (9.45E+17 / E.NUMDEN) * E.SR_TEMP * sqrt(3.472E-2 + 1.0 / (srMw * srMw));
  return area / ((radius / dfkg) + 2.749064E-4 * srMw / (gamma * E.SR_TEMP));
}
// Second-order rate from a first-order uptake rate, limited so that neither
// reactant's lifetime falls below HET_MIN_LIFE.
GCKPP_HD double kIIR1Ltd(double concGas, double concEduct, double kISource) {
  if (concEduct < 1.0) return 0.0;
  if (!is_safe_div(concGas * kISource, concEduct)) return 0.0;
  double kIGas = kISource;
  double kIEduct = kIGas * concGas / concEduct;
  double kII = kIGas / concEduct;
  if (kIGas > 0.0) {
    double lifeA = safe_div(1.0, kIGas, 0.0);
    double lifeB = safe_div(1.0, kIEduct, 0.0);
    if (lifeA < lifeB && lifeA < kHetMinLife) kII = safe_div(kHetMinRate, concEduct, 0.0);
    else if (lifeB < kHetMinLife) kII = safe_div(kHetMinRate, concGas, 0.0);
  }
  return kII;
}
// Grid-average uptake on cloud liquid and ice with entrainment limitation
// (cloudy fraction exchanging with clear air on a 1 h timescale).
GCKPP_HD double CloudHet(const RateEnv &E, double srMw, double gamLiq, double gamIce,
                         double brLiq, double brIce) {
  double cldfr = E.h(HET_cldFr), aliq = E.h(HET_aLiq), aice = E.h(HET_aIce);
  if (cldfr < 0.0001 || aliq + aice <= 0.0) return 0.0;
  double kI = 0.0, kIb = 0.0;
  if (brLiq > 0.0) {
    double area = safe_div(aliq, cldfr, 0.0);
    if (area > 0.0) {
      double t = Ars_L1k(E, area, E.h(HET_rLiq), gamLiq, srMw);
      kI += t;
      kIb += t * brLiq;
    }
  }
  if (brIce > 0.0) {
    double area = safe_div(aice, cldfr, 0.0);
    if (area > 0.0) {
      double t = Ars_L1k(E, area, E.h(HET_rIce), gamIce, srMw);
      kI += t;
      kIb += t * brIce;
    }
  }
  double branch = safe_div(kIb, kI, 0.0);
  if (!(branch > 0.0)) return 0.0;
  double kk = kI * 3600.0;
  double ff = fmin_f(safe_div(cldfr, E.h(HET_clearFr), 1.0e+30), 1.0e+30);
  double xx = (ff - kk - 1.0) / 2.0 +
              sqrt(1.0 + ff * ff + kk * kk + 2.0 * ff + 2.0 * kk - 2.0 * ff * kk) / 2.0;
  xx = fmax_f(xx, 0.0);
  return kI / (1.0 + safe_div(1.0, xx, 1.0e+30)) * branch;
}
GCKPP_HD double coth_(double x) {
  double y = exp(-2.0 * x);
  return (1.0 + y) / (1.0 - y);
}
// reacto-diffusive correction for a particle of radius r, diffuso-reactive length l
GCKPP_HD double ReactoDiff_Corr(double radius, double l) {
  double x = radius / l;
  if (x > 1000.0) return 1.0;
  if (x < 0.1) return x / 3.0;
  return coth_(x) - 1.0 / x;
}
// mean molecular speed [cm/s] for species s
GCKPP_HD double cavg_of(const RateEnv &E, int s) {
  return sqrt(E.EIGHT_RSTARG_T / (kPi * (E.mw(s) * 1.0e-3))) * 100.0;
}
// resistance-sum reaction probability for bulk reaction with first-order
// rate k_tot in a particle: 1/(1/alpha + 1/(gb*corr))
GCKPP_HD double bulk_gamma(const RateEnv &E, double inv_alpha, double H_X, double D_l, double k_tot,
                           double cavg, double radius) {
  double l_r = sqrt(D_l / k_tot);
  double gb = E.FOUR_R_T * H_X * l_r * k_tot / cavg;
  gb = gb * ReactoDiff_Corr(radius, l_r);
  return 1.0 / (inv_alpha + 1.0 / gb);
}

}  // namespace rl
}  // namespace gckpp
