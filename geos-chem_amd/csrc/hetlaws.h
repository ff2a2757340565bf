// Heterogeneous (aerosol / cloud surface) rate laws of fullchem, restated
// from KPP/fullchem/fullchem_RateLawFuncs.F90:795-3460.  Each law is a sum of
// first-order uptake rates (Ars_L1k) over the relevant surfaces -- clear-sky
// aerosol types, stratospheric liquid aerosol (KHETI_SLA), ice, and cloud
// liquid/ice through CloudHet -- followed by the kIIR1Ltd conversion to a
// second-order rate limited by both reactants.  Branching between products
// uses the same reaction-probability parameterisations as the reference.
#pragma once
#include "gen/species_fullchem.h"
#include "ratelaws.h"

namespace gckpp {
namespace rl {
namespace sp = species_fullchem;

// ---- shared gamma parameterisations ----------------------------------------

// ClNO2 + Cl-/Br- in liquid (Gam_ClNO2, :884-935): gamma, Cl and Br branches
struct Gam3 {
  double gamma, b1, b2;
};
GCKPP_HD Gam3 gam_ClNO2(const RateEnv &E, double radius, double pH, double C_Cl, double C_Br) {
  const double inv_ab = 1.0 / 0.01, D_l = 1.0e-5;
  double cavg = cavg_of(E, sp::ClNO2);
  double H_X = 4.5e-2 * kConAtmBar;
  double k_Cl = 1.0e+7 * C_Cl;
  if (pH >= 2.0) k_Cl = 0.0;
  double k_Br = (1.01e-1 / (H_X * H_X * D_l)) * C_Br;
  double k_tot = k_Cl + k_Br;
  Gam3 g{0.0, 0.0, 0.0};
  if (k_tot > 0.0) {
    g.gamma = bulk_gamma(E, inv_ab, H_X, D_l, k_tot, cavg, radius);
    g.b1 = k_Cl / k_tot;
    g.b2 = k_Br / k_tot;
  }
  return g;
}
// ClNO3 on aqueous aerosol with Br- (Gam_ClNO3_Aer): gamma, Br branch
struct Gam2 {
  double gamma, b;
};
GCKPP_HD Gam2 gam_ClNO3_aer(const RateEnv &E, double C_Br) {
  const double inv_ab = 1.0 / 0.108, K_0 = pow(1.2e+5, 2.0), D_l = 5.0e-6;
  double cavg = cavg_of(E, sp::ClNO3);
  double k_Br = 1.0e+12 * C_Br;
  double k_tot = K_0 + k_Br;
  double gb_tot = E.FOUR_R_T * sqrt(k_tot * D_l) / cavg;
  return Gam2{1.0 / (inv_ab + 1.0 / gb_tot), k_Br / k_tot};
}
// ClNO3 on ice (Gam_ClNO3_Ice): gamma and HCl / HBr / H2O branches
struct Gam4 {
  double gamma, hcl, hbr, h2o;
};
GCKPP_HD Gam4 gam_ClNO3_ice(const RateEnv &E) {
  double g1 = 0.24 * E.h(HET_HCl_theta);
  double g2 = 0.56 * E.h(HET_HBr_theta);
  double cavg = cavg_of(E, sp::ClNO3);
  double H2Os = 1e+15 - (3.0 * 2.7e+14 * E.h(HET_HNO3_theta));
  double kks = 4.0 * 5.2e-17 * exp(2032.0 / E.TEMP);
  double g3 = 1.0 / ((1.0 / 0.5) + cavg / (kks * H2Os));
  double g = g1 + g2 + g3;
  return Gam4{g, g1 / g, g2 / g, g3 / g};
}
// HOBr uptake: Henry's law constant of HOBr at TEMP
GCKPP_HD double henry_HOBr(const RateEnv &E) {
  return (E.henry_k0(sp::HOBr) * kConAtmBar) * exp(E.henry_cr(sp::HOBr) * (1.0 / E.TEMP - kInvT298));
}
GCKPP_HD double henry_HOCl(const RateEnv &E) {
  return (E.henry_k0(sp::HOCl) * kConAtmBar) * exp(E.henry_cr(sp::HOCl) * (E.INV_TEMP - kInvT298));
}
// HOBr in acidic aerosol (Gam_HOBr_Aer)
GCKPP_HD double gam_HOBr_aer(const RateEnv &E, double radius, double C_Hp, double C_Clm,
                             double C_Brm) {
  const double inv_ab = 1.0 / 0.6, D_l = 1.4e-5;
  double H_X = henry_HOBr(E);
  double cavg = cavg_of(E, sp::HOBr);
  double C_Hp1 = fmax_f(fmin_f(C_Hp, 1.0e-6), 1.0e-9);
  double C_Hp2 = fmax_f(fmin_f(C_Hp, 1.0e-2), 1.0e-6);
  double k_tot = 2.3e+10 * C_Clm * C_Hp1 + 1.6e+10 * C_Brm * C_Hp2;
  if (!(k_tot > 0.0)) return 0.0;
  return bulk_gamma(E, inv_ab, H_X, D_l, k_tot, cavg, radius);
}
// HOBr in cloud water (Gam_HOBr_Cld): gamma and the four pathway rates
struct GamHOBrCld {
  double gamma, k_tot, k_Cl, k_Br, k_HSO3, k_SO3;
};
GCKPP_HD GamHOBrCld gam_HOBr_cld(const RateEnv &E) {
  const double inv_ab = 1.0 / 0.6, D_l = 1.4e-5;
  double H_X = henry_HOBr(E);
  double cavg = cavg_of(E, sp::HOBr);
  double hl = E.h(HET_H_conc_LCl);
  double C_Hp1 = fmax_f(fmin_f(hl, 1.0e-6), 1.0e-9);
  double C_Hp2 = fmax_f(fmin_f(hl, 1.0e-2), 1.0e-6);
  GamHOBrCld g;
  g.k_Cl = 2.3e+10 * E.h(HET_Cl_conc_Cld) * C_Hp1;
  g.k_Br = 1.6e+10 * E.h(HET_Br_conc_Cld) * C_Hp2;
  g.k_HSO3 = 2.6e+7 * E.h(HET_HSO3_aq);
  g.k_SO3 = 5.0e+9 * E.h(HET_SO3_aq);
  g.k_tot = g.k_Cl + g.k_Br + g.k_HSO3 + g.k_SO3;
  g.gamma = 0.0;
  if (g.k_tot > 0.0) g.gamma = bulk_gamma(E, inv_ab, H_X, D_l, g.k_tot, cavg, E.h(HET_rLiq));
  return g;
}
// HOBr on ice: gamma and HCl / HBr branches (Gam_HOBr_Ice)
GCKPP_HD Gam3 gam_HOBr_ice(const RateEnv &E) {
  double gCl = E.h(HET_HCl_theta) * 0.25;
  double gBr = E.h(HET_HBr_theta) * 4.8e-4 * exp(1240.0 / E.TEMP);
  double g = gCl + gBr;
  Gam3 r{g, 0.0, 0.0};
  if (g > 0.0) {
    r.b1 = gCl / g;
    r.b2 = gBr / g;
  }
  return r;
}
// Br2 yield of HOBr + Br-/Cl- as a function of Br-/Cl- (Br2_Yield)
GCKPP_HD double br2_yield(double r) {
  if (!(r > 0.0)) return 0.0;
  return fmax_f(fmin_f(0.41 * log10(r) + 2.25, 0.9), 0.0);
}
// HOCl in cloud (Gam_HOCl_Cld): gamma, Cl- and S(IV) branches
GCKPP_HD Gam3 gam_HOCl_cld(const RateEnv &E) {
  const double inv_ab = 1.0 / 0.8, D_l = 2.0e-5;
  double k_Cl = 1.5e+4 * E.h(HET_H_conc_LCl) * E.h(HET_Cl_conc_Cld);
  double k_SO3 = 2.8e+5 * E.h(HET_TSO3_aq);
  double k_tot = k_Cl + k_SO3;
  Gam3 g{0.0, 0.0, 0.0};
  if (k_tot > 0.0) {
    g.gamma = bulk_gamma(E, inv_ab, henry_HOCl(E), D_l, k_tot, cavg_of(E, sp::HOCl), E.h(HET_rLiq));
    g.b1 = k_Cl / k_tot;
    g.b2 = k_SO3 / k_tot;
  }
  return g;
}
// HOCl in acidic aerosol (Gam_HOCl_Aer)
GCKPP_HD double gam_HOCl_aer(const RateEnv &E, double radius, double C_Hp, double C_Cl) {
  const double inv_ab = 1.0 / 0.8, D_l = 2.0e-5, K_TER = 1.5e+4;
  if (!(C_Cl > 0.0)) return 0.0;
  double cavg = cavg_of(E, sp::HOCl);
  double H_X = henry_HOCl(E);
  double l_r = sqrt(D_l / (K_TER * C_Hp * C_Cl));
  double gb = E.FOUR_R_T * H_X * l_r * K_TER * C_Hp * C_Cl / cavg;
  gb = gb * ReactoDiff_Corr(radius, l_r);
  return 1.0 / (inv_ab + 1.0 / gb);
}
// O3 + Br- (bulk + surface Langmuir term; Gamma_O3_Br)
GCKPP_HD double gam_O3_Br(const RateEnv &E, double radius, double C_Br) {
  if (!(C_Br > 0.0)) return 0.0;
  const double K0_O3 = 1.1e-2 * kConAtmBar;
  double H_X = K0_O3 * exp(2300.0 * (E.INV_TEMP - kInvT298));
  double cavg = cavg_of(E, sp::O3);
  const double Nmax = 3.0e+14, KLangC = 1.0e-13, k_s = 1.0e-16;
  double C_Br_surf = fmin_f(3.41e+14 * C_Br, Nmax);
  double gs = (4.0 * k_s * C_Br_surf * KLangC * Nmax) / (cavg * (1.0 + KLangC * E.conc(sp::O3)));
  double k_b = 6.3e+8 * exp(-4.45e+3 / E.TEMP);
  const double D_l = 8.9e-6;
  double l_r = sqrt(D_l / (k_b * C_Br));
  double gb = E.FOUR_R_T * H_X * l_r * k_b * C_Br / cavg;
  gb = gb * ReactoDiff_Corr(radius, l_r);
  return gb + gs;
}
// NO3 on sea salt with Cl- (Gam_NO3)
GCKPP_HD double gam_NO3(const RateEnv &E, double aArea, double aRadi, double aWater, double C_X) {
  const double inv_ab = 1.0 / 1.3e-2;
  double vol = aArea * aRadi * 1.0e-3 / 3.0;
  double waterC = aWater / 18.0e+12 / vol;
  double cavg = sqrt(E.EIGHT_RSTARG_T / (kPi * (E.mw(sp::NO3) * 1.0e-3))) * 1.0e2;
  double k_tot = (2.76e+6 * C_X) + (23.0 * waterC);
  if (!(k_tot > 0.0)) return 0.0;
  double H_X = 0.6 * kConAtmBar;
  double l_r = sqrt(1.0e-5 / k_tot);
  double gb = E.FOUR_R_T * H_X * l_r * k_tot / cavg;
  gb = gb * ReactoDiff_Corr(aRadi, l_r);
  return 1.0 / (inv_ab + 1.0 / gb);
}
// N2O5 on mixed inorganic/organic particles (N2O5_InorgOrg) and the ClNO2
// yield (ClNO2_BT).  Returns gamma, ClNO2 yield, wet radius, total area.
struct N2O5Out {
  double gamma, y_clno2, rp, area;
};
GCKPP_HD double clno2_bt(double Cl, double H2O) {
  if (H2O < 0.1) return Cl > 1e-3 ? 1.0 : 0.0;
  return 1.0 / (1.0 + (1.0 / 4.5e+2) * safe_div(H2O, Cl, 1.0e+30));
}
GCKPP_HD N2O5Out n2o5_inorg_org(const RateEnv &E, double volInorg, double volOrg, double H2Oinorg,
                                double H2Oorg, double Rcore, double NIT, double Cl) {
  const double KH = 5.1e+1, k3k2b = 4.0e-2, beta = 1.15e+6, delta = 1.3e-1;
  const double Haq = 5e+3, Daq = 1e-9;
  double volTotal = volInorg + volOrg;
  double H2Ototal = H2Oinorg + H2Oorg;
  double volRatioDry = safe_div(fmax_f(volInorg - H2Oinorg, 0.0), fmax_f(volTotal - H2Ototal, 0.0), 0.0);
  double Rp = safe_div(Rcore, pow(volRatioDry, 1.0 / 3.0), Rcore);
  double l = Rp - Rcore;
  double speed = sqrt(E.EIGHT_RSTARG_T / (kPi * (E.mw(sp::N2O5) * 1.0e-3)));
  double M_H2O = H2Ototal / 18e+0 / volTotal * 1000.0;
  double M_NIT = NIT / volTotal / kAvo * 1000.0;
  double M_Cl = Cl / volTotal / kAvo * 1000.0;
  double OCratio = (((E.h(HET_OMOC_POA) + E.h(HET_OMOC_OPOA)) / 2.0) - 1.17) / 1.29;
  double eps = 1.5e-1 * OCratio + 1.6e-3 * E.RELHUM;
  double gamma_coat = 0.0;
  if (!(l <= 0.0e+0))
    gamma_coat = (E.FOUR_RGASLATM_T * 1.0e-3 * eps * Haq * Daq * Rcore / 100.0) /
                 (speed * l / 100.0 * Rp / 100.0);
  double areaTotal = 3.0 * volTotal / Rp;
  double gamma_core;
  if (M_H2O < 0.1) {
    gamma_core = 0.005;
  } else {
    double sp100 = speed * 1e+2;
    double A = fmin_f(((4.0 * volTotal) / (sp100 * areaTotal)) * KH, 3.2e-8);
    double k2f = (delta * M_H2O < 1e-2) ? beta * (delta * M_H2O) : beta * (1e+0 - exp(-delta * M_H2O));
    gamma_core = A * k2f * (1.0 - 1.0 / (1.0 + safe_div(k3k2b * M_H2O, M_NIT, 1.0e+30)));
  }
  double gamma;
  if (gamma_coat <= 0.0) gamma = gamma_core;
  else if (gamma_core <= 0.0) gamma = 0.0;
  else gamma = 1.0 / ((1.0 / gamma_core) + (1.0 / gamma_coat));
  return N2O5Out{gamma, clno2_bt(M_Cl, M_H2O), Rp, areaTotal};
}
// EpoxUptkGamma: acid-catalysed epoxide uptake on sulfate
GCKPP_HD double epox_gamma(const RateEnv &E, double srMw) {
  const double DIFF = 1.0e-1, MACOEFF = 1.0e-1, K_HPLUS = 3.6e-2, K_NUC = 2.0e-4;
  const double K_HSO4 = 7.3e-4, K_HYDRO = 0.0e+0, HSTAR_EPOX = 1.7e+7;
  double valTmp = 0.0;
  double aerVol = (E.xArea(SUL) * E.xRadi(SUL)) / 3.0;
  double xmms = sqrt((2.117e+8 * E.TEMP) / (srMw * srMw));
  double hp = E.h(HET_H_plus);
  double kPart = (K_HPLUS * hp) + (K_NUC * hp * (E.h(HET_NO3_molal) + E.h(HET_SO4_molal))) +
                 (K_HSO4 * E.h(HET_HSO4_molal)) + (K_HYDRO);
  double val1 = (E.xRadi(SUL) * xmms) / (4.0 * DIFF);
  double val2 = (1.0 / MACOEFF);
  if (E.xArea(SUL) > 0.0 && xmms > 0.0)
    valTmp = (E.FOUR_RGASLATM_T * aerVol * HSTAR_EPOX * kPart) / (E.xArea(SUL) * xmms);
  double val3 = valTmp > 0.0 ? 1.0 / valTmp : 0.0;
  double gamma = 0.0;
  if (kPart >= 1.e-8) gamma = 1.0 / (val1 + val2 + val3);
  if (gamma < 0.0) gamma = 0.0;
  return gamma;
}

// ---- the heterogeneous rate laws --------------------------------------------
// sum of Ars_L1k over aerosol types [t0, t1] with common gamma, area scaled
GCKPP_HD double uptake_types(const RateEnv &E, int t0, int t1, double gamma, double srMw,
                             double scale) {
  double k = 0.0;
  for (int t = t0; t <= t1; ++t) k = k + Ars_L1k(E, scale * E.xArea(t), E.xRadi(t), gamma, srMw);
  return k;
}

GCKPP_HD double BrNO3uptkByH2O(const RateEnv &E) {
  double gamLiq = 0.0021 * E.TEMP - 0.561, gamIce = 5.3e-4 * exp(1100.0 / E.TEMP);
  double srMw = E.srmw(sp::BrNO3), cf = E.h(HET_clearFr);
  double k = 0.0;
  k = k + Ars_L1k(E, cf * E.xArea(SUL), E.xRadi(SUL), gamLiq, srMw);
  k = k + Ars_L1k(E, cf * E.xArea(SSA), E.xRadi(SSA), gamLiq, srMw);
  k = k + Ars_L1k(E, cf * E.xArea(SSC), E.xRadi(SSC), gamLiq, srMw);
  k = k + E.xArea(SLA) * E.kheti(BrNO3_plus_H2O);
  k = k + Ars_L1k(E, cf * E.xArea(IIC), E.xRadi(IIC), E.nat() ? 0.001 : 0.3, srMw);
  k = k + CloudHet(E, srMw, gamLiq, gamIce, 1.0, 1.0);
  return kIIR1Ltd(E.conc(sp::BrNO3), E.conc(sp::H2O), k);
}
GCKPP_HD double BrNO3uptkByHCl(const RateEnv &E) {
  double srMw = E.srmw(sp::BrNO3), k = 0.0;
  if (E.strat()) {
    k = k + Ars_L1k(E, E.xArea(SUL), E.xRadi(SUL), 0.9, srMw);
    k = k + E.xArea(SLA) * E.kheti(BrNO3_plus_HCl);
    k = k + Ars_L1k(E, E.xArea(IIC), E.xRadi(IIC), 0.3, srMw);
  }
  k = kIIR1Ltd(E.conc(sp::BrNO3), E.conc(sp::HCl), k);
  return E.off() ? 0.0 : k;
}

// ClNO2 + halide: cloud part with branch frac_field (skipped in the
// stratosphere), optional fine/coarse sea-salt part, limited by `educt`.
GCKPP_HD double clno2_uptake(const RateEnv &E, bool chloride, int frac_field, int ss, int educt) {
  double srMw = E.srmw(sp::ClNO2), k = 0.0;
  if (!E.strat()) {
    Gam3 g = gam_ClNO2(E, E.h(HET_rLiq), E.h(HET_pHCloud), E.h(HET_Cl_conc_Cld), E.h(HET_Br_conc_Cld));
    double branch = (chloride ? g.b1 : g.b2) * E.h(frac_field);
    k = k + CloudHet(E, srMw, g.gamma, 0.0, branch, 0.0);
  }
  if (ss == 1) {  // fine sea salt
    Gam3 g = gam_ClNO2(E, E.h(HET_aClRadi), E.h(HET_pHSSA), E.h(HET_Cl_conc_SSA), E.h(HET_Br_conc_SSA));
    double area = E.h(HET_clearFr) * E.h(HET_aClArea);
    k = k + Ars_L1k(E, area, E.h(HET_aClRadi), g.gamma, srMw) * (chloride ? g.b1 : g.b2);
  } else if (ss == 2) {  // coarse sea salt
    Gam3 g = gam_ClNO2(E, E.xRadi(SSC), E.h(HET_pHSSA + 1), E.h(HET_Cl_conc_SSC), E.h(HET_Br_conc_SSC));
    double area = E.h(HET_clearFr) * E.xArea(SSC);
    k = k + Ars_L1k(E, area, E.xRadi(SSC), g.gamma, srMw) * (chloride ? g.b1 : g.b2);
  }
  return kIIR1Ltd(E.conc(sp::ClNO2), E.conc(educt), k);
}
GCKPP_HD double ClNO2uptkByBrSALA(const RateEnv &E) {
  return clno2_uptake(E, false, HET_frac_Br_CldA, 1, sp::BrSALA);
}
GCKPP_HD double ClNO2uptkByBrSALC(const RateEnv &E) {
  return clno2_uptake(E, false, HET_frac_Br_CldC, 2, sp::BrSALC);
}
GCKPP_HD double ClNO2uptkByHBr(const RateEnv &E) {
  return clno2_uptake(E, false, HET_frac_Br_CldG, 0, sp::HBr);
}
GCKPP_HD double ClNO2uptkBySALACL(const RateEnv &E) {
  return clno2_uptake(E, true, HET_frac_Cl_CldA, 1, sp::SALACL);
}
GCKPP_HD double ClNO2uptkBySALCCL(const RateEnv &E) {
  return clno2_uptake(E, true, HET_frac_Cl_CldC, 0, sp::SALCCL);
}
GCKPP_HD double ClNO2uptkByHCl(const RateEnv &E) {
  return clno2_uptake(E, true, HET_frac_Cl_CldG, 0, sp::HCl);
}

GCKPP_HD double ClNO3uptkByH2O(const RateEnv &E) {
  double srMw = E.srmw(sp::ClNO3), k = 0.0;
  Gam2 a = gam_ClNO3_aer(E, E.h(HET_Br_conc_SSA));
  double area = E.h(HET_clearFr) * E.h(HET_aClArea);
  double branchLiq = (1.0 - a.b) * (1.0 - E.h(HET_frac_SALACL));
  k = k + Ars_L1k(E, area, E.h(HET_aClRadi), a.gamma, srMw) * branchLiq;
  k = k + E.xArea(SLA) * E.kheti(ClNO3_plus_H2O);
  k = k + Ars_L1k(E, E.xArea(IIC), E.xRadi(IIC), E.nat() ? 0.004 : 0.3, srMw);
  if (!E.strat()) {
    Gam2 c = gam_ClNO3_aer(E, E.h(HET_Br_conc_Cld));
    Gam4 i = gam_ClNO3_ice(E);
    k = k + CloudHet(E, srMw, c.gamma, i.gamma, 1.0 - c.b, i.h2o);
  }
  return kIIR1Ltd(E.conc(sp::ClNO3), E.conc(sp::H2O), k);
}
GCKPP_HD double ClNO3uptkByHCl(const RateEnv &E) {
  double srMw = E.srmw(sp::ClNO3), k = 0.0;
  if (E.strat()) {
    k = k + Ars_L1k(E, E.xArea(SUL), E.xRadi(SUL), 0.1e-4, srMw);
    k = k + E.xArea(SLA) * E.kheti(ClNO3_plus_HCl);
    k = k + Ars_L1k(E, E.xArea(IIC), E.xRadi(IIC), E.nat() ? 0.2 : 0.3, srMw);
  } else {
    Gam4 i = gam_ClNO3_ice(E);
    k = k + CloudHet(E, srMw, 0.0, i.gamma, 0.0, i.hcl);
  }
  return kIIR1Ltd(E.conc(sp::ClNO3), E.conc(sp::HCl), k);
}
GCKPP_HD double ClNO3uptkByHBr(const RateEnv &E) {
  double srMw = E.srmw(sp::ClNO3), k = 0.0;
  if (E.strat()) {
    k = k + E.xArea(SLA) * E.kheti(ClNO3_plus_HBr);
    k = k + Ars_L1k(E, E.xArea(IIC), E.xRadi(IIC), 0.3, srMw);
  } else {
    Gam2 c = gam_ClNO3_aer(E, E.h(HET_Br_conc_Cld));
    Gam4 i = gam_ClNO3_ice(E);
    k = CloudHet(E, srMw, c.gamma, i.gamma, c.b * E.h(HET_frac_Br_CldG), i.hbr);
  }
  k = kIIR1Ltd(E.conc(sp::ClNO3), E.conc(sp::HBr), k);
  return E.off() ? 0.0 : k;
}
GCKPP_HD double clno3_by_brss(const RateEnv &E, int frac_field, bool fine, int educt) {
  double srMw = E.srmw(sp::ClNO3), k = 0.0;
  if (!E.strat()) {
    Gam2 c = gam_ClNO3_aer(E, E.h(HET_Br_conc_Cld));
    k = k + CloudHet(E, srMw, c.gamma, 0.0, c.b * E.h(frac_field), 0.0);
  }
  Gam2 a = gam_ClNO3_aer(E, E.h(fine ? HET_Br_conc_SSA : HET_Br_conc_SSC));
  double area = E.h(HET_clearFr) * (fine ? E.h(HET_aClArea) : E.xArea(SSC));
  double rad = fine ? E.h(HET_aClRadi) : E.xRadi(SSC);
  k = k + Ars_L1k(E, area, rad, a.gamma, srMw) * a.b;
  k = kIIR1Ltd(E.conc(sp::ClNO3), E.conc(educt), k);
  return E.off() ? 0.0 : k;
}
GCKPP_HD double ClNO3uptkByBrSALA(const RateEnv &E) {
  return clno3_by_brss(E, HET_frac_Br_CldA, true, sp::BrSALA);
}
GCKPP_HD double ClNO3uptkByBrSALC(const RateEnv &E) {
  return clno3_by_brss(E, HET_frac_Br_CldC, false, sp::BrSALC);
}
GCKPP_HD double ClNO3uptkBySALACL(const RateEnv &E) {
  if (E.strat()) return 0.0;
  Gam2 a = gam_ClNO3_aer(E, E.h(HET_Br_conc_SSA));
  double area = E.h(HET_clearFr) * E.h(HET_aClArea);
  double k = Ars_L1k(E, area, E.h(HET_aClRadi), a.gamma, E.srmw(sp::ClNO3)) *
             ((1.0 - a.b) * E.h(HET_frac_SALACL));
  return kIIR1Ltd(E.conc(sp::ClNO3), E.conc(sp::SALACL), k);
}
GCKPP_HD double ClNO3uptkBySALCCL(const RateEnv &E) {
  if (E.strat()) return 0.0;
  Gam2 a = gam_ClNO3_aer(E, E.h(HET_Br_conc_SSC));
  double area = E.h(HET_clearFr) * E.xArea(SSC);
  double k = Ars_L1k(E, area, E.xRadi(SSC), a.gamma, E.srmw(sp::ClNO3)) * (1.0 - a.b);
  return kIIR1Ltd(E.conc(sp::ClNO3), E.conc(sp::SALCCL), k);
}

GCKPP_HD double hbr_by_ss(const RateEnv &E, bool fine) {
  if (E.strat()) return 0.0;
  double area = E.h(HET_clearFr) * (fine ? E.h(HET_aClArea) : E.xArea(SSC));
  double gamma = 1.3e-8 * exp(4290.0 / E.TEMP);
  return Ars_L1k(E, area, fine ? E.h(HET_aClRadi) : E.xRadi(SSC), gamma, E.srmw(sp::HBr));
}
GCKPP_HD double HBrUptkBySALA(const RateEnv &E) { return hbr_by_ss(E, true); }
GCKPP_HD double HBrUptkBySALC(const RateEnv &E) { return hbr_by_ss(E, false); }

GCKPP_HD double HO2uptk1stOrd(const RateEnv &E) {
  return uptake_types(E, DU1, SSC, E.h(HET_gamma_HO2), E.srmw(sp::HO2), 1.0);
}

// HOBr + {HBr, HCl}: strat surfaces or tropospheric cloud with Br2/BrCl yield
GCKPP_HD double hobr_by_hx(const RateEnv &E, bool hbr) {
  double srMw = E.srmw(sp::HOBr), k = 0.0;
  if (E.strat()) {
    k = k + Ars_L1k(E, E.xArea(SUL), E.xRadi(SUL), hbr ? 0.25 : 0.2, srMw);
    k = k + E.xArea(SLA) * E.kheti(hbr ? HOBr_plus_HBr : HOBr_plus_HCl);
    double gi = hbr ? (E.nat() ? 0.001 : 0.3) : (E.nat() ? 0.1 : 0.3);
    k = k + Ars_L1k(E, E.xArea(IIC), E.xRadi(IIC), gi, srMw);
  } else {
    GamHOBrCld c = gam_HOBr_cld(E);
    double b0 = (c.k_Cl + c.k_Br) / c.k_tot;
    double r = E.h(HET_Br_over_Cl_Cld);
    double branch = b0 * (hbr ? 0.9 : 0.1);
    if (r <= 5.0e-4) branch = b0 * (hbr ? br2_yield(r) : (1.0 - br2_yield(r)));
    double brLiq = branch * E.h(hbr ? HET_frac_Br_CldG : HET_frac_Cl_CldG);
    Gam3 i = gam_HOBr_ice(E);
    k = k + CloudHet(E, srMw, c.gamma, i.gamma, brLiq, hbr ? i.b2 : i.b1);
  }
  return kIIR1Ltd(E.conc(sp::HOBr), E.conc(hbr ? sp::HBr : sp::HCl), k);
}
GCKPP_HD double HOBrUptkByHBr(const RateEnv &E) { return hobr_by_hx(E, true); }
GCKPP_HD double HOBrUptkByHCl(const RateEnv &E) { return hobr_by_hx(E, false); }
// HOBr + sea-salt Br- (bromide=true) or Cl- (false), fine or coarse
GCKPP_HD double hobr_by_ss(const RateEnv &E, bool bromide, bool fine, int educt) {
  double srMw = E.srmw(sp::HOBr), k = 0.0;
  if (!E.strat()) {
    GamHOBrCld c = gam_HOBr_cld(E);
    double b0 = (c.k_Cl + c.k_Br) / c.k_tot;
    double r = E.h(HET_Br_over_Cl_Cld);
    double branch = b0 * (bromide ? 0.9 : 0.1);
    if (r <= 5.0e-4) branch = b0 * (bromide ? br2_yield(r) : (1.0 - br2_yield(r)));
    int ff = bromide ? (fine ? HET_frac_Br_CldA : HET_frac_Br_CldC)
                     : (fine ? HET_frac_Cl_CldA : HET_frac_Cl_CldC);
    k = k + CloudHet(E, srMw, c.gamma, 0.0, branch * E.h(ff), 0.0);
  }
  if (E.hb(fine ? HET_SSA_is_Acid : HET_SSC_is_Acid)) {
    double rad = fine ? E.h(HET_aClRadi) : E.xRadi(SSC);
    double g = gam_HOBr_aer(E, rad, E.h(fine ? HET_H_conc_SSA : HET_H_conc_SSC),
                            E.h(fine ? HET_Cl_conc_SSA : HET_Cl_conc_SSC),
                            E.h(fine ? HET_Br_conc_SSA : HET_Br_conc_SSC));
    double r = E.h(fine ? HET_Br_over_Cl_SSA : HET_Br_over_Cl_SSC);
    double branch = bromide ? 0.9 : 0.1;
    if (r <= 5.0e-4) branch = bromide ? br2_yield(r) : 1.0 - br2_yield(r);
    double area = E.h(HET_clearFr) * (fine ? E.h(HET_aClArea) : E.xArea(SSC)) *
                  E.h(fine ? HET_f_Acid_SSA : HET_f_Acid_SSC);
    k = k + Ars_L1k(E, area, rad, g, srMw) * branch;
  }
  return kIIR1Ltd(E.conc(sp::HOBr), E.conc(educt), k);
}
GCKPP_HD double HOBrUptkByBrSALA(const RateEnv &E) { return hobr_by_ss(E, true, true, sp::BrSALA); }
GCKPP_HD double HOBrUptkByBrSALC(const RateEnv &E) { return hobr_by_ss(E, true, false, sp::BrSALC); }
GCKPP_HD double HOBrUptkBySALACL(const RateEnv &E) { return hobr_by_ss(E, false, true, sp::SALACL); }
GCKPP_HD double HOBrUptkBySALCCL(const RateEnv &E) { return hobr_by_ss(E, false, false, sp::SALCCL); }
GCKPP_HD double hobr_by_s4(const RateEnv &E, bool hso3) {
  double srMw = E.srmw(sp::HOBr), k = 0.0;
  if (!E.strat()) {
    GamHOBrCld c = gam_HOBr_cld(E);
    k = k + CloudHet(E, srMw, c.gamma, 0.0, (hso3 ? c.k_HSO3 : c.k_SO3) / c.k_tot, 0.0);
  }
  return kIIR1Ltd(E.conc(sp::HOBr), E.conc(sp::SO2), k);
}
GCKPP_HD double HOBrUptkByHSO3m(const RateEnv &E) { return hobr_by_s4(E, true); }
GCKPP_HD double HOBrUptkBySO3mm(const RateEnv &E) { return hobr_by_s4(E, false); }

GCKPP_HD double HOClUptkByHCl(const RateEnv &E) {
  double srMw = E.srmw(sp::HOCl), k = 0.0;
  if (E.strat()) {
    k = k + Ars_L1k(E, E.xArea(SUL), E.xRadi(SUL), 0.8, srMw);
    k = k + E.xArea(SLA) * E.kheti(HOCl_plus_HCl);
    k = k + Ars_L1k(E, E.xArea(IIC), E.xRadi(IIC), E.nat() ? 0.1 : 0.2, srMw);
    return kIIR1Ltd(E.conc(sp::HOCl), E.conc(sp::HCl), k);
  }
  Gam3 c = gam_HOCl_cld(E);
  k = k + CloudHet(E, srMw, c.gamma, 0.22 * E.h(HET_HCl_theta), c.b1 * E.h(HET_frac_Cl_CldG), 1.0);
  return kIIR1Ltd(E.conc(sp::HOCl), E.conc(sp::HCl), k);
}
GCKPP_HD double HOClUptkByHBr(const RateEnv &E) {
  double srMw = E.srmw(sp::HOCl), k = 0.0;
  if (E.strat()) {
    k = k + Ars_L1k(E, E.xArea(SUL), E.xRadi(SUL), 0.8, srMw);
    k = k + E.xArea(SLA) * E.kheti(HOCl_plus_HBr);
    k = k + Ars_L1k(E, E.xArea(IIC), E.xRadi(IIC), 0.3, srMw);
  }
  k = kIIR1Ltd(E.conc(sp::HOCl), E.conc(sp::HBr), k);
  return E.off() ? 0.0 : k;
}
GCKPP_HD double hocl_by_ss(const RateEnv &E, bool fine, int educt) {
  double srMw = E.srmw(sp::HOCl), k = 0.0;
  if (!E.strat()) {
    Gam3 c = gam_HOCl_cld(E);
    k = k + CloudHet(E, srMw, c.gamma, 0.0, c.b1 * E.h(fine ? HET_frac_Cl_CldA : HET_frac_Cl_CldC), 0.0);
  }
  if (E.hb(fine ? HET_SSA_is_Acid : HET_SSC_is_Acid)) {
    double rad = fine ? E.h(HET_aClRadi) : E.xRadi(SSC);
    double g = gam_HOCl_aer(E, rad, E.h(fine ? HET_H_conc_SSA : HET_H_conc_SSC),
                            E.h(fine ? HET_Cl_conc_SSA : HET_Cl_conc_SSC));
    double area = E.h(HET_clearFr) * (fine ? E.h(HET_aClArea) : E.xArea(SSC)) *
                  E.h(fine ? HET_f_Acid_SSA : HET_f_Acid_SSC);
    k = k + Ars_L1k(E, area, rad, g, srMw);
  }
  return kIIR1Ltd(E.conc(sp::HOCl), E.conc(educt), k);
}
GCKPP_HD double HOClUptkBySALACL(const RateEnv &E) { return hocl_by_ss(E, true, sp::SALACL); }
GCKPP_HD double HOClUptkBySALCCL(const RateEnv &E) { return hocl_by_ss(E, false, sp::SALCCL); }
GCKPP_HD double hocl_by_s4(const RateEnv &E, bool hso3) {
  double srMw = E.srmw(sp::HOCl), k = 0.0;
  if (!E.strat()) {
    Gam3 c = gam_HOCl_cld(E);
    double br = c.b2 * E.h(hso3 ? HET_frac_HSO3_aq : HET_frac_SO3_aq);
    k = k + CloudHet(E, srMw, c.gamma, 0.0, br, 0.0);
  }
  return kIIR1Ltd(E.conc(sp::HOCl), E.conc(sp::SO2), k) * E.h(hso3 ? HET_HSO3m : HET_SO3mm);
}
GCKPP_HD double HOClUptkByHSO3m(const RateEnv &E) { return hocl_by_s4(E, true); }
GCKPP_HD double HOClUptkBySO3mm(const RateEnv &E) { return hocl_by_s4(E, false); }

// iodine uptake on sulfate / sea salt (first order, no reactant limit)
GCKPP_HD double IuptkBySulf1stOrd(const RateEnv &E, double srMw, double gamma) {
  double k = Ars_L1k(E, E.xArea(SUL), E.xRadi(SUL), gamma, srMw);
  return k + Ars_L1k(E, E.xArea(SLA), E.xRadi(SLA), gamma, srMw);
}
GCKPP_HD double IuptkBySALA1stOrd(const RateEnv &E, double srMw, double gamma) {
  return E.strat() ? 0.0 : Ars_L1k(E, E.xArea(SSA), E.xRadi(SSA), gamma, srMw);
}
GCKPP_HD double IuptkBySALC1stOrd(const RateEnv &E, double srMw, double gamma) {
  return E.strat() ? 0.0 : Ars_L1k(E, E.xArea(SSC), E.xRadi(SSC), gamma, srMw);
}
GCKPP_HD double IuptkByAlkSALA1stOrd(const RateEnv &E, double srMw, double gamma) {
  if (E.strat() || !E.hb(HET_SSA_is_Alk)) return 0.0;
  return Ars_L1k(E, E.h(HET_f_Alk_SSA) * E.xArea(SSA), E.xRadi(SSA), gamma, srMw);
}
GCKPP_HD double IuptkByAlkSALC1stOrd(const RateEnv &E, double srMw, double gamma) {
  if (E.strat() || !E.hb(HET_SSC_is_Alk)) return 0.0;
  return Ars_L1k(E, E.h(HET_f_Alk_SSC) * E.xArea(SSC), E.xRadi(SSC), gamma, srMw);
}
// iodine breakdown on acidic sea salt: fraction `share` to the Br or Cl educt
GCKPP_HD double i_brkdn(const RateEnv &E, double srMw, double conc, double gamma, bool fine,
                        double share, int educt) {
  if (E.strat() || !E.hb(fine ? HET_SSA_is_Acid : HET_SSC_is_Acid)) return 0.0;
  int t = fine ? SSA : SSC;
  double ssarea = E.h(fine ? HET_f_Acid_SSA : HET_f_Acid_SSC) * E.xArea(t);
  double k = share * Ars_L1k(E, ssarea, E.xRadi(t), gamma, srMw);
  return kIIR1Ltd(conc, E.conc(educt), k);
}
GCKPP_HD double IbrkdnByAcidBrSALA(const RateEnv &E, double srMw, double conc, double gamma) {
  return i_brkdn(E, srMw, conc, gamma, true, 0.15, sp::BrSALA);
}
GCKPP_HD double IbrkdnByAcidBrSALC(const RateEnv &E, double srMw, double conc, double gamma) {
  return i_brkdn(E, srMw, conc, gamma, false, 0.15, sp::BrSALC);
}
GCKPP_HD double IbrkdnByAcidSALACl(const RateEnv &E, double srMw, double conc, double gamma) {
  return i_brkdn(E, srMw, conc, gamma, true, 0.85, sp::SALACL);
}
GCKPP_HD double IbrkdnByAcidSALCCl(const RateEnv &E, double srMw, double conc, double gamma) {
  return i_brkdn(E, srMw, conc, gamma, false, 0.85, sp::SALCCL);
}
GCKPP_HD double IONO2uptkByH2O(const RateEnv &E) {
  double srMw = E.srmw(sp::IONO2), cf = E.h(HET_clearFr), k = 0.0;
  k = k + Ars_L1k(E, cf * E.xArea(SUL), E.xRadi(SUL), fmax_f(0.0021 * E.TEMP - 0.561, 0.0), srMw);
  k = k + E.xArea(SLA) * E.kheti(BrNO3_plus_H2O);
  k = k + Ars_L1k(E, cf * E.xArea(IIC), E.xRadi(IIC), E.nat() ? 0.001 : 0.3, srMw);
  k = k + CloudHet(E, srMw, 0.01, 0.01, 1.0, 1.0);
  return kIIR1Ltd(E.conc(sp::IONO2), E.conc(sp::H2O), k);
}

GCKPP_HD N2O5Out n2o5_fine(const RateEnv &E) {
  return n2o5_inorg_org(E, E.h(HET_aClVol), E.xVol(ORC), E.xH2O(SUL), E.xH2O(ORC),
                        E.h(HET_aClRadi), E.conc(sp::NIT), E.conc(sp::SALACL));
}
GCKPP_HD N2O5Out n2o5_coarse(const RateEnv &E) {
  return n2o5_inorg_org(E, E.xVol(SSC), 0.0, E.xH2O(SSC), 0.0, E.xRadi(SSC), E.conc(sp::NITs),
                        E.conc(sp::SALCCL));
}
GCKPP_HD double N2O5uptkByH2O(const RateEnv &E) {
  double srMw = E.srmw(sp::N2O5), cf = E.h(HET_clearFr);
  double k = uptake_types(E, DU1, DU7, 0.02, srMw, cf);
  N2O5Out f = n2o5_fine(E);
  double t = Ars_L1k(E, cf * f.area, f.rp, f.gamma, srMw);
  k = k + t - (t * f.y_clno2 * 0.25);
  k = k + Ars_L1k(E, cf * E.xArea(BKC), E.xRadi(BKC), 0.005, srMw);
  N2O5Out c = n2o5_coarse(E);
  t = Ars_L1k(E, cf * c.area, c.rp, c.gamma, srMw);
  k = k + t - (t * c.y_clno2);
  k = k + E.xArea(SLA) * E.kheti(N2O5_plus_H2O);
  k = k + Ars_L1k(E, cf * E.xArea(IIC), E.xRadi(IIC), E.nat() ? 4.0e-4 : 0.02, srMw);
  return kIIR1Ltd(E.conc(sp::N2O5), E.conc(sp::H2O), k);
}
GCKPP_HD double N2O5uptkBySALACl(const RateEnv &E) {
  if (E.strat()) return 0.0;
  N2O5Out f = n2o5_fine(E);
  double k = Ars_L1k(E, E.h(HET_clearFr) * f.area, f.rp, f.gamma, E.srmw(sp::N2O5));
  k = k * f.y_clno2 * 0.25;
  return kIIR1Ltd(E.conc(sp::N2O5), E.conc(sp::SALACL), k);
}
GCKPP_HD double N2O5uptkBySALCCl(const RateEnv &E) {
  if (E.strat()) return 0.0;
  N2O5Out c = n2o5_coarse(E);
  double k = Ars_L1k(E, E.h(HET_clearFr) * c.area, c.rp, c.gamma, E.srmw(sp::N2O5));
  k = k * c.y_clno2;
  return kIIR1Ltd(E.conc(sp::N2O5), E.conc(sp::SALCCL), k);
}
GCKPP_HD double N2O5uptkByCloud(const RateEnv &E) {
  double gamma = (0.03 / 0.019) * exp(-25.5265 + 9283.76 / E.TEMP - 851801.0 / pow_int(E.TEMP, 2));
  return CloudHet(E, E.srmw(sp::N2O5), gamma, 0.02, 1.0, 1.0);
}
GCKPP_HD double N2O5uptkByStratHCl(const RateEnv &E) {
  double k = 0.0;
  if (E.strat()) {
    k = k + (E.xArea(SLA) * E.kheti(N2O5_plus_HCl));
    k = k + Ars_L1k(E, E.xArea(IIC), E.xRadi(IIC), E.nat() ? 0.003 : 0.03, E.srmw(sp::N2O5));
  }
  return kIIR1Ltd(E.conc(sp::N2O5), E.conc(sp::HCl), k);
}
GCKPP_HD double NO2uptk1stOrdAndCloud(const RateEnv &E) {
  double srMw = E.srmw(sp::NO2);
  double k = uptake_types(E, DU1, DU7, 1.0e-8, srMw, 1.0);
  k = k + Ars_L1k(E, E.xArea(SUL), E.xRadi(SUL), 5e-6, srMw);
  k = k + Ars_L1k(E, E.xArea(BKC), E.xRadi(BKC), 1e-4, srMw);
  k = k + Ars_L1k(E, E.xArea(ORC), E.xRadi(ORC), 1e-6, srMw);
  double rh = E.RELHUM, g;
  if (rh < 40.0) g = 1.0e-8;
  else if (rh > 70.0) g = 1.0e-4;
  else g = 1.0e-8 + (1e-4 - 1e-8) * (rh - 40.0) / 30.0;
  k = k + Ars_L1k(E, E.xArea(SSA), E.xRadi(SSA), g, srMw);
  k = k + Ars_L1k(E, E.xArea(SSC), E.xRadi(SSC), g, srMw);
  k = k + Ars_L1k(E, E.xArea(SLA), E.xRadi(SLA), 1.0e-4, srMw);
  k = k + Ars_L1k(E, E.xArea(IIC), E.xRadi(IIC), 1.0e-4, srMw);
  return k + CloudHet(E, srMw, 1.0e-8, 0.0, 1.0, 0.0);
}
GCKPP_HD double NO3uptk1stOrdAndCloud(const RateEnv &E) {
  double srMw = E.srmw(sp::NO3);
  double k = uptake_types(E, DU1, DU7, 0.01, srMw, 1.0);
  k = k + Ars_L1k(E, E.xArea(BKC), E.xRadi(BKC), E.RELHUM < 50.0 ? 2.0e-4 : 1.0e-3, srMw);
  k = k + Ars_L1k(E, E.xArea(ORC), E.xRadi(ORC), 0.005, srMw);
  k = k + Ars_L1k(E, E.xArea(SLA), E.xRadi(SLA), 0.1, srMw);
  k = k + Ars_L1k(E, E.xArea(IIC), E.xRadi(IIC), 0.1, srMw);
  return k + CloudHet(E, srMw, 0.002, 0.001, 1.0, 1.0);
}
GCKPP_HD double no3_hydrolysis_ss(const RateEnv &E, bool fine) {
  double area = fine ? E.h(HET_aClArea) : E.xArea(SSC);
  double radi = fine ? E.h(HET_aClRadi) : E.xRadi(SSC);
  double water = E.h(HET_aWater + (fine ? 0 : 1));
  double conc = E.h(fine ? HET_Cl_conc_SSA : HET_Cl_conc_SSC);
  double gamma = gam_NO3(E, area, radi, water, conc) * 0.01;
  return Ars_L1k(E, E.h(HET_clearFr) * area, radi, gamma, E.srmw(sp::NO3));
}
GCKPP_HD double NO3hypsisClonSALA(const RateEnv &E) { return no3_hydrolysis_ss(E, true); }
GCKPP_HD double NO3hypsisClonSALC(const RateEnv &E) { return no3_hydrolysis_ss(E, false); }

GCKPP_HD double o3_br_cloud(const RateEnv &E, double br_branch) {
  if (E.strat()) return 0.0;
  double g = gam_O3_Br(E, E.h(HET_rLiq), E.h(HET_Br_conc_Cld));
  return CloudHet(E, E.srmw(sp::O3), g, 0.0, br_branch, 0.0);
}
GCKPP_HD double O3uptkByHBr(const RateEnv &E) {
  return kIIR1Ltd(E.conc(sp::O3), E.conc(sp::HBr), o3_br_cloud(E, E.h(HET_frac_Br_CldG)));
}
GCKPP_HD double o3_br_ss(const RateEnv &E, bool fine, int educt) {
  if (E.strat()) return 0.0;
  double k = o3_br_cloud(E, E.h(fine ? HET_frac_Br_CldA : HET_frac_Br_CldC));
  if (E.hb(fine ? HET_SSA_is_Acid : HET_SSC_is_Acid)) {
    double rad = fine ? E.h(HET_aClRadi) : E.xRadi(SSC);
    double area = E.h(HET_clearFr) * (fine ? E.h(HET_aClArea) : E.xArea(SSC)) *
                  E.h(fine ? HET_f_Acid_SSA : HET_f_Acid_SSC);
    double g = gam_O3_Br(E, rad, E.h(fine ? HET_Br_conc_SSA : HET_Br_conc_SSC));
    k = k + Ars_L1k(E, area, rad, g, E.srmw(sp::O3));
  }
  return kIIR1Ltd(E.conc(sp::O3), E.conc(educt), k);
}
GCKPP_HD double O3uptkByBrSALA(const RateEnv &E) { return o3_br_ss(E, true, sp::BrSALA); }
GCKPP_HD double O3uptkByBrSALC(const RateEnv &E) { return o3_br_ss(E, false, sp::BrSALC); }
GCKPP_HD double OHuptkBySALACl(const RateEnv &E) {
  if (E.strat()) return 0.0;
  double k = Ars_L1k(E, E.h(HET_aClArea), E.h(HET_aClRadi), 0.04 * E.h(HET_Cl_conc_SSA), E.srmw(sp::OH));
  return kIIR1Ltd(E.conc(sp::OH), E.conc(sp::SALACL), k);
}
GCKPP_HD double OHuptkBySALCCl(const RateEnv &E) {
  if (E.strat()) return 0.0;
  double k = Ars_L1k(E, E.xArea(SSC), E.xRadi(SSC), 0.04 * E.h(HET_Cl_conc_SSC), E.srmw(sp::OH));
  return kIIR1Ltd(E.conc(sp::OH), E.conc(sp::SALCCL), k);
}
GCKPP_HD double GLYXuptk1stOrd(const RateEnv &E, double srMw) {
  if (!(E.RELHUM >= kCritRH)) return 0.0;
  double gamma = E.SUNCOS > 0.0 ? 4.4e-3 : 8.0e-6;
  return Ars_L1k(E, E.xArea(SUL), E.xRadi(SUL), gamma, srMw);
}
GCKPP_HD double IEPOXuptk1stOrd(const RateEnv &E, double srMw, double doScale) {
  if (!(E.RELHUM >= kCritRH)) return 0.0;
  double gamma = epox_gamma(E, srMw);
  if (doScale != 0.0 && E.h(HET_H_plus) > 8.0e-5) gamma = gamma / 30.0;
  return Ars_L1k(E, E.xArea(SUL), E.xRadi(SUL), gamma, srMw);
}
GCKPP_HD double MGLYuptk1stOrd(const RateEnv &E, double srMw) {
  if (!(E.RELHUM >= kCritRH)) return 0.0;
  return Ars_L1k(E, E.xArea(SUL), E.xRadi(SUL), 3.6e-7, srMw);
}
GCKPP_HD double VOCuptk1stOrd(const RateEnv &E, double srMw, double gamma) {
  if (!(E.RELHUM >= kCritRH)) return 0.0;
  double k = 0.0;
  const int types[7] = {SUL, BKC, ORC, SSA, SSC, SLA, IIC};
  for (int i = 0; i < 7; ++i) k = k + Ars_L1k(E, E.xArea(types[i]), E.xRadi(types[i]), gamma, srMw);
  return k;
}

}  // namespace rl
}  // namespace gckpp
