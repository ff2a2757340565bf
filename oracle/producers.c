/* Test infrastructure only (tests/, never the product path): plain-C
 * restatements of the per-cell producers upstream of the fullchem rates:
 * fullchem_SetStateHet, the HetState record the rate laws read, computed
 * from the model's per-cell fields, and fullchem_SulfurAqChem, the K_MT
 * sea-salt uptake rates.
 *
 *   fullChem_SetStateHet   KPP/fullchem/fullchem_HetStateFuncs.F90:48-244
 *   Cld_Params             KPP/fullchem/rateLawUtilFuncs.F90:281-417
 *   SafeDiv                KPP/fullchem/rateLawUtilFuncs.F90:459-478
 *   Get_Theta_Ice          KPP/fullchem/fullchem_HetStateFuncs.F90:259-301
 *   Halide_Conc            :316-472
 *   Get_Halide_CldConc     :487-558
 *   Get_Halide_SsaConc     :572-616
 *   Compute_L2G_Local      :631-693, with CALC_KH and CALC_HEFF
 *                          (GeosUtil/henry_mod.F90:74-113, 154-204)
 *   fullchem_SulfurAqChem  KPP/fullchem/fullchem_SulfurChemFuncs.F90:120-305
 *   Ars_L1k, kIIR1Ltd      KPP/fullchem/rateLawUtilFuncs.F90:77-150
 *
 * Layout: hetin[NHIN][n] (HIN_* of cell_fields.h), conc[nspec][n], out
 * het[NHET][n]; sp[] gives the KPP indices of SO2 HNO3 HCl HBr BrSALA BrSALC
 * SALACL SALCCL NIT NITs SO4.  Fortran MIN/MAX are (a<b)?a:b and (a>b)?a:b
 * and EXPONENT is frexp's exponent (0 for 0, HUGE for NaN/Inf), as flang
 * evaluates them.  Pinned by tests/test_oracle_setstatehet.py and
 * tests/test_oracle_sulfur_aq.py. */
#include <math.h>
#include <stdint.h>
#include <limits.h>

#include "../geos-chem_amd/csrc/cell_fields.h"

enum { I_SO2, I_HNO3, I_HCL, I_HBR, I_BRSALA, I_BRSALC, I_SALACL, I_SALCCL, I_NIT, I_NITS, I_SO4 };

static const double AVO = 6.022140857e+23; /* Headers/physconstants.F90:32 */

static double fmin2(double a, double b) { return a < b ? a : b; }
static double fmax2(double a, double b) { return a > b ? a : b; }

static int f_exponent(double x) {
  int e = 0;
  if (x == 0.0) return 0;
  if (!isfinite(x)) return INT_MAX;
  frexp(x, &e);
  return e;
}

/* SafeDiv: alt when the quotient would overflow or den is 0, 0 when it would
 * underflow (the exponent difference in wrapping 32-bit integers). */
double orc_safediv(double num, double den, double alt) {
  int32_t ed = (int32_t)((uint32_t)f_exponent(num) - (uint32_t)f_exponent(den));
  if (ed > 1023 || den == 0.0) return alt;
  if (ed < -1020) return 0.0;
  return num / den;
}

/* Compute_L2G_Local: KH = K0 exp(CR (1/T - 1/Tref)) R T / ATM, the effective
 * constant KH (1 + 10^(pH - pKa)) when pH > 0 and pKa > -100, times the
 * liquid water content. */
double orc_l2g(double k0, double cr, double pka, double tk, double h2oliq, double ph) {
  const double TREF = 298.15, R = 8.3144598, ATM = 101.325; /* henry_mod.F90 */
  double kh = k0 * exp(cr * (1.0 / tk - 1.0 / TREF)) * R * tk / ATM;
  double heff = kh;
  if (ph > 0.0 && pka > -100.0) heff = heff * (1.0 + pow(10.0, ph - pka));
  return heff * h2oliq;
}

/* Cld_Params: out = rLiq, rIce, ALiq, VLiq, AIce, VIce. */
void orc_cld_params(double ad, double cldf, double frland, double frocean, double qi, double ql,
                    double t, double vair, double out[6]) {
  const double CLDR_CONT = 6.0e-4, CLDR_MARI = 10.0e-4, CLDR_ICE = 38.5e-4;
  const double DENS_LIQ = 0.001, DENS_ICE = 0.91e-3;
  double alpha, beta;
  if (ql + qi <= 0.0 || cldf <= 0.0) {
    out[0] = CLDR_CONT;
    out[1] = CLDR_ICE;
    out[2] = out[3] = out[4] = out[5] = 0.0;
    return;
  }
  out[0] = frland > frocean ? CLDR_CONT : CLDR_MARI;
  out[3] = ql * ad / DENS_LIQ / vair;
  out[5] = qi * ad / DENS_ICE / vair;
  out[2] = 3.0 * out[3] / out[0];
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
  out[1] = 0.5 * alpha * exp(beta * (t - 273.15)) / 1e+4;
  out[4] = 3.0 * out[5] / out[1] * 2.25;
}

static double ssa_conc(double n_x, double area, double r_w) {
  double v = (area * r_w / 3.0) * 1e-3;
  if (v <= 1.0e-20) return 0.0;
  return fmax2((n_x / AVO) / v, 0.0);
}

void orc_set_state_het(int64_t n, const double *hetin, const double *conc, const int32_t *sp,
                       int naero, double gamma_ho2, int lpscchem, int turnoff, double *het) {
  for (int64_t i = 0; i < n; ++i) {
#define IN(f) hetin[(int64_t)(f) * n + i]
#define C(k) conc[(int64_t)sp[k] * n + i]
#define H(f) het[(int64_t)(f) * n + i]
    double cldf = IN(HIN_CLDF), t = IN(HIN_T), cld[6];
    H(HET_cldFr) = fmin2(fmax2(cldf, 0.0), 1.0);
    H(HET_clearFr) = 1.0 - H(HET_cldFr);
    H(HET_qIce) = IN(HIN_QI);
    H(HET_qLiq) = IN(HIN_QL);
    H(HET_vAir) = IN(HIN_AIRVOL) * 1.0e6;
    H(HET_NAEROTYPE) = naero;
    H(HET_aClArea) = IN(HIN_aClArea);
    H(HET_aClRadi) = IN(HIN_aClRadi);
    H(HET_aClVol) = H(HET_aClArea) * H(HET_aClRadi) / 3.0;
    for (int k = 0; k < 2; ++k) H(HET_aWater + k) = IN(HIN_IsorropAeroH2O + k);
    for (int k = 0; k < 14; ++k) {
      H(HET_xArea + k) = H(HET_xRadi + k) = H(HET_xVol + k) = 0.0;
      H(HET_wetArea + k) = H(HET_xH2O + k) = 0.0;
    }
    for (int k = 0; k < naero; ++k) {
      H(HET_xArea + k) = IN(HIN_AeroArea + k);
      H(HET_xRadi + k) = IN(HIN_AeroRadi + k);
      H(HET_xVol + k) = H(HET_xArea + k) * H(HET_xRadi + k) / 3.0;
      H(HET_wetArea + k) = IN(HIN_WetAeroArea + k);
      H(HET_xH2O + k) = IN(HIN_AeroH2O + k) * 1.0e-6;
    }
    H(HET_OMOC_POA) = IN(HIN_OMOC_POA);
    H(HET_OMOC_OPOA) = IN(HIN_OMOC_OPOA);
    H(HET_HSO3_aq) = IN(HIN_HSO3_aq);
    H(HET_SO3_aq) = IN(HIN_SO3_aq);
    H(HET_TSO3_aq) = H(HET_HSO3_aq) + H(HET_SO3_aq);
    H(HET_frac_HSO3_aq) = orc_safediv(H(HET_HSO3_aq), H(HET_TSO3_aq), 0.0);
    H(HET_frac_SO3_aq) = orc_safediv(H(HET_SO3_aq), H(HET_TSO3_aq), 0.0);
    H(HET_HSO4_molal) = IN(HIN_IsorropBisulfate);
    H(HET_NO3_molal) = IN(HIN_IsorropNitrate);
    H(HET_SO4_molal) = IN(HIN_IsorropSulfate);
    H(HET_H_plus) = IN(HIN_IsorropHplus);
    H(HET_pHCloud) = IN(HIN_pHCloud);
    for (int k = 0; k < 2; ++k) H(HET_pHSSA + k) = IN(HIN_IsorropAeropH + k);
    H(HET_H_conc_Sul) = pow(10.0, -1.0 * H(HET_pHSSA));
    H(HET_H_conc_LCl) = pow(10.0, -1.0 * H(HET_pHCloud));
    H(HET_H_conc_ICl) = pow(10.0, -4.5);
    H(HET_H_conc_SSA) = H(HET_H_conc_Sul);
    H(HET_H_conc_SSC) = pow(10.0, -5.0);
    H(HET_f_Alk_SSA) = orc_safediv(IN(HIN_SALAAL), IN(HIN_SALA), 0.0);
    H(HET_f_Alk_SSA) = fmax2(fmin2(H(HET_f_Alk_SSA), 1.0), 0.0);
    H(HET_f_Acid_SSA) = 1.0 - H(HET_f_Alk_SSA);
    H(HET_f_Alk_SSC) = orc_safediv(IN(HIN_SALCAL), IN(HIN_SALC), 0.0);
    H(HET_f_Alk_SSC) = fmax2(fmin2(H(HET_f_Alk_SSC), 1.0), 0.0);
    H(HET_f_Acid_SSC) = 1.0 - H(HET_f_Alk_SSC);
    H(HET_SSA_is_Alk) = fabs(H(HET_f_Alk_SSA)) > 0.01;
    H(HET_SSA_is_Acid) = !(H(HET_SSA_is_Alk) != 0.0);
    H(HET_SSC_is_Alk) = fabs(H(HET_f_Alk_SSC)) > 0.01;
    H(HET_SSC_is_Acid) = !(H(HET_SSC_is_Alk) != 0.0);
    H(HET_gamma_HO2) = gamma_ho2;
    H(HET_fupdateHOBr) = IN(HIN_fupdateHOBr);
    H(HET_fupdateHOCl) = IN(HIN_fupdateHOCl);
    H(HET_HSO3m) = orc_safediv(IN(HIN_HSO3_aq) * 1.0e-3 * AVO * IN(HIN_QL) * IN(HIN_AIRDEN) * 1.0e-3,
                               cldf, 0.0);
    H(HET_HSO3m) = orc_safediv(H(HET_HSO3m), C(I_SO2), 0.0);
    H(HET_SO3mm) = orc_safediv(IN(HIN_SO3_aq) * 1.0e-3 * AVO * IN(HIN_QL) * IN(HIN_AIRDEN) * 1.0e-3,
                               cldf, 0.0);
    H(HET_SO3mm) = orc_safediv(H(HET_SO3mm), C(I_SO2), 0.0);
    orc_cld_params(IN(HIN_AD), cldf, IN(HIN_FRLAND), IN(HIN_FROCEAN), IN(HIN_QI), IN(HIN_QL), t,
                   H(HET_vAir), cld);
    H(HET_rLiq) = cld[0];
    H(HET_rIce) = cld[1];
    H(HET_aLiq) = cld[2];
    H(HET_vLiq) = cld[3];
    H(HET_aIce) = cld[4];
    H(HET_vIce) = cld[5];
    { /* Get_Theta_Ice */
      double klinc = 7.5e-5 * exp(4585.0 / t);
      double klangc1 = klinc / 2.7e+14;
      double klangc2, denom;
      klinc = 1.3e-5 * exp(4600.0 / t);
      klangc2 = klinc / 3.0e+14;
      denom = 1.0 + klangc1 * C(I_HNO3) + klangc2 * C(I_HCL);
      H(HET_HNO3_theta) = klangc1 * C(I_HNO3) / denom;
      H(HET_HCl_theta) = klangc2 * C(I_HCL) / denom;
      H(HET_HBr_theta) = 4.14e-10 * pow(C(I_HBR), 0.88);
      H(HET_HBr_theta) = fmin2(H(HET_HBr_theta), 1.0);
    }
    { /* Halide_Conc and Get_Halide_CldConc */
      double hbr = C(I_HBR) + (C(I_BRSALA) * 0.7) + C(I_BRSALC);
      double hcl = C(I_HCL) + (C(I_SALACL) * 0.7) + C(I_SALCCL);
      double br_conc = 0.0, cl_conc = 0.0, denom, a12, r12;
      double v_tot = orc_safediv(H(HET_vLiq), H(HET_cldFr), 0.0);
      H(HET_Br_conc_CldG) = H(HET_Br_conc_CldA) = H(HET_Br_conc_CldC) = 0.0;
      H(HET_Cl_conc_CldG) = H(HET_Cl_conc_CldA) = H(HET_Cl_conc_CldC) = 0.0;
      H(HET_Br_over_Cl_Cld) = H(HET_Br_over_Cl_SSA) = H(HET_Br_over_Cl_SSC) = 0.0;
      H(HET_frac_Br_CldA) = H(HET_frac_Br_CldC) = H(HET_frac_Br_CldG) = 0.0;
      H(HET_frac_Cl_CldA) = H(HET_frac_Cl_CldC) = H(HET_frac_Cl_CldG) = 0.0;
      if (!(v_tot < 1.0e-20)) {
        double l2g = orc_l2g(1.0, 9000.0, -6.3, t, v_tot, H(HET_pHCloud));
        double f_l = l2g / (1.0 + l2g);
        cl_conc = f_l * hcl / (v_tot * AVO * 1.0e-3);
        l2g = orc_l2g(7.5e-1, 10200.0, -9.0, t, v_tot, H(HET_pHCloud));
        f_l = l2g / (1.0 + l2g);
        br_conc = f_l * hbr / (v_tot * AVO * 1.0e-3);
      }
      denom = C(I_HBR) + (C(I_BRSALA) * 0.7) + C(I_BRSALC);
      if (denom > 0.0) {
        H(HET_Br_conc_CldG) = (br_conc * C(I_HBR)) / denom;
        H(HET_Br_conc_CldA) = (br_conc * C(I_BRSALA) * 0.7) / denom;
        H(HET_Br_conc_CldC) = (br_conc * C(I_BRSALC)) / denom;
      }
      denom = C(I_HCL) + (C(I_SALACL) * 0.7) + C(I_SALCCL);
      if (denom > 0.0) {
        H(HET_Cl_conc_CldG) = (cl_conc * C(I_HCL)) / denom;
        H(HET_Cl_conc_CldA) = (cl_conc * C(I_SALACL) * 0.7) / denom;
        H(HET_Cl_conc_CldC) = (cl_conc * C(I_SALCCL)) / denom;
      }
      H(HET_Br_conc_Cld) = H(HET_Br_conc_CldA) + H(HET_Br_conc_CldC) + H(HET_Br_conc_CldG);
      H(HET_Cl_conc_Cld) = H(HET_Cl_conc_CldA) + H(HET_Cl_conc_CldC) + H(HET_Cl_conc_CldG);
      if (H(HET_Br_conc_Cld) > 0.0) {
        H(HET_frac_Br_CldA) = H(HET_Br_conc_CldA) / H(HET_Br_conc_Cld);
        H(HET_frac_Br_CldC) = H(HET_Br_conc_CldC) / H(HET_Br_conc_Cld);
        H(HET_frac_Br_CldG) = H(HET_Br_conc_CldG) / H(HET_Br_conc_Cld);
      }
      if (H(HET_Cl_conc_Cld) > 0.0) {
        H(HET_frac_Cl_CldA) = H(HET_Cl_conc_CldA) / H(HET_Cl_conc_Cld);
        H(HET_frac_Cl_CldC) = H(HET_Cl_conc_CldC) / H(HET_Cl_conc_Cld);
        H(HET_frac_Cl_CldG) = H(HET_Cl_conc_CldG) / H(HET_Cl_conc_Cld);
      }
      /* xArea(12), xRadi(12): coarse sea salt (zero when nAeroType < 12) */
      a12 = H(HET_xArea + 11);
      r12 = H(HET_xRadi + 11);
      H(HET_Br_conc_SSA) = ssa_conc(C(I_BRSALA), H(HET_aClArea), H(HET_aClRadi));
      H(HET_Br_conc_SSC) = ssa_conc(C(I_BRSALC), a12, r12);
      H(HET_Cl_conc_SSA) = ssa_conc(C(I_SALACL), H(HET_aClArea), H(HET_aClRadi));
      H(HET_Cl_conc_SSC) = ssa_conc(C(I_SALCCL), a12, r12);
      H(HET_NIT_conc_SSA) = ssa_conc(C(I_NIT), H(HET_aClArea), H(HET_aClRadi));
      H(HET_NIT_conc_SSC) = ssa_conc(C(I_NITS), a12, r12);
      if (H(HET_Cl_conc_Cld) > 0.0) H(HET_Br_over_Cl_Cld) = H(HET_Br_conc_Cld) / H(HET_Cl_conc_Cld);
      if (H(HET_Cl_conc_SSA) > 0.0) H(HET_Br_over_Cl_SSA) = H(HET_Br_conc_SSA) / H(HET_Cl_conc_SSA);
      if (H(HET_Cl_conc_SSC) > 0.0) H(HET_Br_over_Cl_SSC) = H(HET_Br_conc_SSC) / H(HET_Cl_conc_SSC);
      if ((C(I_SALACL) + C(I_NIT) + C(I_SO4)) > 0.0)
        H(HET_frac_SALACL) = C(I_SALACL) / (C(I_SALACL) + C(I_NIT) + C(I_SO4));
      else
        H(HET_frac_SALACL) = 0.0;
    }
    for (int k = 0; k < 11; ++k) H(HET_KHETI_SLA + k) = IN(HIN_KHETI_SLA + k);
    H(HET_stratBox) = IN(HIN_InStratosphere) != 0.0;
    H(HET_pscBox) = lpscchem && IN(HIN_STATE_PSC) >= 2.0 && H(HET_stratBox) != 0.0;
    H(HET_natSurface) = H(HET_pscBox) != 0.0 && C(I_NIT) > 0.0;
    H(HET_TurnOffHetRates) = turnoff != 0;
    /* HetState fields no routine of the fullchem path sets */
    H(HET_H2O) = H(HET_cloudBox) = H(HET_fracOrgAer) = H(HET_fracInorgAer) = 0.0;
#undef IN
#undef C
#undef H
  }
}

/* ---- fullchem_SulfurAqChem ------------------------------------------------ */

/* Ars_L1k: first-order uptake rate [1/s] on aerosol of area [cm2/cm3] and
 * radius [cm], with the cell's NUMDEN [molec/cm3] and SR_TEMP = sqrt(T). */
double orc_ars_l1k(double area, double radius, double gamma, double srmw, double numden,
                   double sr_temp) {
  double dfkg;
  if (gamma < 1.0e-30 || radius < 1.0e-30) return 0.0;
  dfkg = (9.45E+17 / numden) * sr_temp * sqrt(3.472E-2 + 1.0 / (srmw * srmw));
  return area / ((radius / dfkg) + 2.749064E-4 * srmw / (gamma * sr_temp));
}

static int is_safediv(double num, double den) {
  int32_t ed = (int32_t)((uint32_t)f_exponent(num) - (uint32_t)f_exponent(den));
  return !(ed < -1020 || ed > 1023 || den == 0.0);
}

/* kIIR1Ltd: the second-order rate of a first-order uptake, limited so that
 * neither reactant's lifetime falls below 1e-3 s. */
double orc_kiir1ltd(double conc_gas, double conc_educt, double ki_source) {
  const double min_life = 1.e-3, min_rate = 1.0 / min_life;
  double ki_gas, ki_educt, kii, life_a, life_b;
  if (conc_educt < 1.0) return 0.0;
  if (!is_safediv(conc_gas * ki_source, conc_educt)) return 0.0;
  ki_gas = ki_source;
  ki_educt = ki_gas * conc_gas / conc_educt;
  kii = ki_gas / conc_educt;
  if (ki_gas > 0.0) {
    life_a = orc_safediv(1.0, ki_gas, 0.0);
    life_b = orc_safediv(1.0, ki_educt, 0.0);
    if (life_a < life_b && life_a < min_life)
      kii = orc_safediv(min_rate, conc_educt, 0.0);
    else if (life_b < min_life)
      kii = orc_safediv(min_rate, conc_gas, 0.0);
  }
  return kii;
}

/* met[NMET][n]: rows K_MT..K_MT+5 written from TEMP, NUMDEN, hetin's
 * WetAeroArea and AeroRadi of types 11 and 12 and conc (SALAAL, SALCAL in
 * equivalents); sp[] = SO2 HNO3 HCl O3 SALAAL SALCAL; srmw[] = sqrt(MW) of
 * SO2, HCl, HNO3. */
void orc_sulfur_aq(int64_t n, const double *hetin, const double *conc, const int32_t *sp,
                   const double *srmw, double *met) {
  enum { J_SO2, J_HNO3, J_HCL, J_O3, J_SALAAL, J_SALCAL };
  for (int64_t i = 0; i < n; ++i) {
    double t = met[(int64_t)MET_TEMP * n + i], numden = met[(int64_t)MET_NUMDEN * n + i];
    double srt = sqrt(t), kmt[6] = {0, 0, 0, 0, 0, 0}, k_ex;
#define C(k) conc[(int64_t)sp[k] * n + i]
#define IN(f) hetin[(int64_t)(f) * n + i]
    int aa = C(J_SALAAL) > 0.1, ca = C(J_SALCAL) > 0.1, o3 = C(J_O3) > 1.0e+10;
    double a11 = IN(HIN_WetAeroArea + 10), r11 = IN(HIN_AeroRadi + 10);
    double a12 = IN(HIN_WetAeroArea + 11), r12 = IN(HIN_AeroRadi + 11);
    if (aa && o3) {
      k_ex = orc_ars_l1k(a11, r11, 0.11, srmw[0], numden, srt);
      kmt[0] = orc_kiir1ltd(C(J_SO2), C(J_SALAAL), k_ex) / C(J_O3);
    }
    if (aa) {
      k_ex = orc_ars_l1k(a11, r11, 0.07, srmw[1], numden, srt);
      kmt[1] = orc_kiir1ltd(C(J_HCL), C(J_SALAAL), k_ex);
      k_ex = orc_ars_l1k(a11, r11, 0.5, srmw[2], numden, srt);
      kmt[2] = orc_kiir1ltd(C(J_HNO3), C(J_SALAAL), k_ex);
    }
    if (ca && o3) {
      k_ex = orc_ars_l1k(a12, r12, 0.11, srmw[0], numden, srt);
      kmt[3] = orc_kiir1ltd(C(J_SO2), C(J_SALCAL), k_ex) / C(J_O3);
    }
    if (ca) {
      k_ex = orc_ars_l1k(a12, r12, 0.07, srmw[1], numden, srt);
      kmt[4] = orc_kiir1ltd(C(J_HCL), C(J_SALCAL), k_ex);
      k_ex = orc_ars_l1k(a12, r12, 0.5, srmw[2], numden, srt);
      kmt[5] = orc_kiir1ltd(C(J_HNO3), C(J_SALCAL), k_ex);
    }
    for (int k = 0; k < 6; ++k) met[(int64_t)(MET_K_MT + k) * n + i] = kmt[k];
#undef C
#undef IN
  }
}
