/* Plain-C restatement of the reference's per-cell KPP Rosenbrock solve.
 * TEST INFRASTRUCTURE ONLY (a checker, never the product).  Each function
 * cites the reference routine whose algorithm it restates.  Checks that pin
 * it live in tests/test_oracle.py. */
#include "gckpp_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* KPP/fullchem/gckpp_Function.F90:2172-5060 (Fun_SPLIT).  The reference
 * forms P (production) and D (destruction per unit V) and returns P - D*V;
 * a reaction's rate is A = RCT * prod(reactant concentrations), a consumed
 * reactant with multiplicity k contributes k*A/V to D.  Restated here from
 * the reactant lists and the net stoichiometry of the mechanism tables. */
void orc_fun(const orc_mech *m, const double *V, const double *F,
             const double *RCT, double *Vdot) {
  int nv = m->nvar;
  double *P = calloc((size_t)nv, sizeof(double));
  double *D = calloc((size_t)nv, sizeof(double));
  for (int r = 0; r < m->nreact; ++r) {
    int b = m->rx_ptr[r], e = m->rx_ptr[r + 1];
    double a = RCT[r];
    for (int k = b; k < e; ++k) {
      int s = m->rx_spc[k];
      double c = s < nv ? V[s] : F[s - nv];
      for (int q = 0; q < m->rx_mult[k]; ++q) a *= c;
    }
    for (int k = m->st_ptr[r]; k < m->st_ptr[r + 1]; ++k) {
      int s = m->st_spc[k];
      double coef = m->st_coef[k];
      /* consumed reactant species of r go to D, everything else to P */
      int is_reactant = 0, mult = 0;
      for (int q = b; q < e; ++q)
        if (m->rx_spc[q] == s) { is_reactant = 1; mult = m->rx_mult[q]; }
      if (is_reactant && coef == -(double)mult) {
        /* D += mult * RCT * prod(others) * V^(mult-1) */
        double d = RCT[r];
        for (int q = b; q < e; ++q) {
          int t = m->rx_spc[q];
          double c = t < nv ? V[t] : F[t - nv];
          int pw = (t == s) ? m->rx_mult[q] - 1 : m->rx_mult[q];
          for (int z = 0; z < pw; ++z) d *= c;
        }
        D[s] += (double)mult * d;
      } else {
        P[s] += coef * a;
      }
    }
  }
  for (int i = 0; i < nv; ++i) Vdot[i] = P[i] - D[i] * V[i];
  free(P);
  free(D);
}

/* KPP/fullchem/gckpp_Jacobian.F90:48-20887 (Jac_SP): JVS(k) = dVdot(i)/dV(j)
 * for LU entry k = (i,j); entries of the LU pattern that are not structural
 * Jacobian nonzeros (fill-in) are zero.  dA/dV(j) = RCT * mult_j *
 * V(j)^(mult_j-1) * prod(other reactants). */
void orc_jac(const orc_mech *m, const double *V, const double *F,
             const double *RCT, double *JVS) {
  int nv = m->nvar;
  memset(JVS, 0, sizeof(double) * (size_t)m->lu_nonzero);
  for (int r = 0; r < m->nreact; ++r) {
    int b = m->rx_ptr[r], e = m->rx_ptr[r + 1];
    for (int kj = b; kj < e; ++kj) {
      int j = m->rx_spc[kj];
      if (j >= nv) continue; /* fixed species: no Jacobian column */
      double d = RCT[r] * (double)m->rx_mult[kj];
      for (int q = b; q < e; ++q) {
        int t = m->rx_spc[q];
        double c = t < nv ? V[t] : F[t - nv];
        int pw = (q == kj) ? m->rx_mult[q] - 1 : m->rx_mult[q];
        for (int z = 0; z < pw; ++z) d *= c;
      }
      for (int k = m->st_ptr[r]; k < m->st_ptr[r + 1]; ++k) {
        int i = m->st_spc[k];
        int idx = m->lu_map[(size_t)i * nv + j];
        if (idx >= 0) JVS[idx] += m->st_coef[k] * d;
      }
    }
  }
}

/* KPP/fullchem/gckpp_LinearAlgebra.F90:46-83 (KppDecomp): row-wise sparse
 * Doolittle LU without pivoting in the fixed pattern, dense row workspace W;
 * fails with IER=k when |pivot k| < TINY (smallest normal double). */
int orc_decomp(const orc_mech *m, double *JVS) {
  int nv = m->nvar;
  double *W = calloc((size_t)nv, sizeof(double));
  int ier = 0;
  for (int k = 0; k < nv && !ier; ++k) {
    if (fabs(JVS[m->lu_diag[k]]) < DBL_MIN) { ier = k + 1; break; }
    for (int kk = m->lu_crow[k]; kk < m->lu_crow[k + 1]; ++kk)
      W[m->lu_icol[kk]] = JVS[kk];
    for (int kk = m->lu_crow[k]; kk < m->lu_diag[k]; ++kk) {
      int j = m->lu_icol[kk];
      double a = -W[j] / JVS[m->lu_diag[j]];
      W[j] = -a;
      for (int jj = m->lu_diag[j] + 1; jj < m->lu_crow[j + 1]; ++jj)
        W[m->lu_icol[jj]] += a * JVS[jj];
    }
    for (int kk = m->lu_crow[k]; kk < m->lu_crow[k + 1]; ++kk)
      JVS[kk] = W[m->lu_icol[kk]];
  }
  free(W);
  return ier;
}

/* KPP/fullchem/gckpp_LinearAlgebra.F90:644-2309 (KppSolve): forward
 * substitution with the unit lower factor, then backward substitution with
 * the upper factor, dividing by the diagonal. */
void orc_solve(const orc_mech *m, const double *JVS, double *X) {
  int nv = m->nvar;
  for (int i = 0; i < nv; ++i) {
    double s = X[i];
    for (int kk = m->lu_crow[i]; kk < m->lu_diag[i]; ++kk)
      s -= JVS[kk] * X[m->lu_icol[kk]];
    X[i] = s;
  }
  for (int i = nv - 1; i >= 0; --i) {
    double s = X[i];
    for (int kk = m->lu_diag[i] + 1; kk < m->lu_crow[i + 1]; ++kk)
      s -= JVS[kk] * X[m->lu_icol[kk]];
    X[i] = s / JVS[m->lu_diag[i]];
  }
}

/* Rodas3 tableau, KPP/fullchem/gckpp_Integrator.F90:2239-2303 */
static const double RA[6] = {0.0, 2.0, 0.0, 2.0, 0.0, 1.0};
static const double RC_[6] = {4.0, 1.0, -1.0, 1.0, -1.0, -(8.0 / 3.0)};
static const int RNEWF[4] = {1, 0, 1, 1};
static const double RM[4] = {2.0, 0.0, 1.0, 1.0};
static const double RE[4] = {0.0, 0.0, 0.0, 1.0};
static const double RELO = 3.0;
static const double RGAMMA[4] = {0.5, 1.5, 0.0, 0.0};

/* Fortran MIN/MAX as the reference's compiler evaluates them: the first
 * argument unless the second compares strictly smaller / larger (decides how
 * NaN propagates through the step-size control). */
static double min_f(double a, double b) { return a < b ? a : b; }
static double max_f(double a, double b) { return a > b ? a : b; }

/* WAXPY skips alpha == 0 (gckpp_LinearAlgebra.F90:3762) */
static void waxpy(int n, double a, const double *x, double *y) {
  if (a == 0.0) return;
  for (int i = 0; i < n; ++i) y[i] += a * x[i];
}

/* ros_ErrorNorm, gckpp_Integrator.F90:1715-1745 */
static double err_norm(int n, const double *Y, const double *Ynew,
                       const double *Yerr, const double *atol,
                       const double *rtol, int vectol) {
  double err = 0.0;
  for (int i = 0; i < n; ++i) {
    double ymax = max_f(fabs(Y[i]), fabs(Ynew[i]));
    double sc = vectol ? atol[i] + rtol[i] * ymax : atol[0] + rtol[0] * ymax;
    double q = Yerr[i] / sc;
    err += q * q;
  }
  err = sqrt(err / n);
  return max_f(err, 1.0e-10);
}

/* Integrate/Rosenbrock/ros_Integrator for one cell.
 * gckpp_Integrator.F90:80-162 (option merge), :165-532 (defaults and
 * checks), :578-786 (time loop), :1921-1999 (PrepareMatrix). */
static void integrate_cell(const orc_mech *m, double *C, const double *RCT,
                           double hstart_in, double Tstart, double Tend,
                           const double *atol, const double *rtol,
                           const int *icntrl_u, const double *rcntrl_u,
                           int *ISTATUS, double *RSTATUS, int *IERR) {
  int N = m->nvar, nz = m->lu_nonzero;
  int ICNTRL[20];
  double RCNTRL[20];
  for (int i = 0; i < 20; ++i) {
    ICNTRL[i] = 0;
    RCNTRL[i] = 0.0;
  }
  ICNTRL[14] = 5;
  for (int i = 0; i < 20; ++i) {
    if (icntrl_u[i] != 0) ICNTRL[i] = icntrl_u[i];
    if (rcntrl_u[i] > 0) RCNTRL[i] = rcntrl_u[i];
  }
  if (hstart_in > 0) RCNTRL[2] = hstart_in; else RCNTRL[2] = 0.0;
  for (int i = 0; i < 20; ++i) { ISTATUS[i] = 0; RSTATUS[i] = 0.0; }
  *IERR = 0;
  const double *F = C + N;
  double *Y = C;
  int autonomous = ICNTRL[0] != 0;
  int vectol = ICNTRL[1] == 0;
  int uplim = vectol ? N : 1;
  if (!(ICNTRL[2] == 0 || ICNTRL[2] == 4)) { *IERR = -2; return; }
  int max_steps;
  if (ICNTRL[3] == 0) max_steps = 200000;
  else if (ICNTRL[3] > 0) max_steps = ICNTRL[3];
  else { *IERR = -1; return; }
  const double roundoff = DBL_EPSILON; /* WLAMCH('E'), LinearAlgebra:3861-3895 */
  const double DeltaMin = 1.0e-5;
  double Hmin, Hmax, Hstart, FacMin, FacMax, FacRej, FacSafe;
  if (RCNTRL[0] == 0.0) Hmin = 0.0; else if (RCNTRL[0] > 0) Hmin = RCNTRL[0]; else { *IERR = -3; return; }
  if (RCNTRL[1] == 0.0) Hmax = fabs(Tend - Tstart);
  else if (RCNTRL[1] > 0) Hmax = min_f(fabs(RCNTRL[1]), fabs(Tend - Tstart));
  else { *IERR = -3; return; }
  if (RCNTRL[2] == 0.0) Hstart = max_f(Hmin, DeltaMin);
  else if (RCNTRL[2] > 0) Hstart = min_f(fabs(RCNTRL[2]), fabs(Tend - Tstart));
  else { *IERR = -3; return; }
  FacMin = RCNTRL[3] == 0.0 ? 0.2 : RCNTRL[3];
  FacMax = RCNTRL[4] == 0.0 ? 6.0 : RCNTRL[4];
  FacRej = RCNTRL[5] == 0.0 ? 0.1 : RCNTRL[5];
  FacSafe = RCNTRL[6] == 0.0 ? 0.9 : RCNTRL[6];
  for (int i = 0; i < uplim; ++i)
    if (atol[i] <= 0.0 || rtol[i] <= 10.0 * roundoff || rtol[i] >= 1.0) { *IERR = -5; return; }
  (void)autonomous; /* ICNTRL(1)=1: no dF/dT term (Rodas3 gammas 3,4 are 0) */

  size_t vb = sizeof(double) * (size_t)N;
  double *Ynew = malloc(vb), *Fcn0 = malloc(vb), *Fcn = malloc(vb), *Yerr = malloc(vb);
  double *K = malloc(vb * 4);
  double *Jac0 = malloc(sizeof(double) * (size_t)nz), *Ghimj = malloc(sizeof(double) * (size_t)nz);

  double T = Tstart;
  double H = min_f(max_f(fabs(Hmin), fabs(Hstart)), fabs(Hmax));
  if (fabs(H) <= 10.0 * roundoff) H = DeltaMin;
  int dir = Tend >= Tstart ? 1 : -1;
  H = dir * H;
  int rejLast = 0, rejMore = 0;
  double Hnew = H;
  *IERR = 1;
  while ((dir > 0 && (T - Tend) + roundoff <= 0.0) || (dir < 0 && (Tend - T) + roundoff <= 0.0)) {
    if (ISTATUS[2] > max_steps) { *IERR = -6; goto done; }
    if ((T + 0.1 * H) == T || H <= roundoff) { *IERR = -7; goto done; }
    H = min_f(H, fabs(Tend - T));
    orc_fun(m, Y, F, RCT, Fcn0);
    ISTATUS[0]++;
    orc_jac(m, Y, F, RCT, Jac0);
    ISTATUS[1]++;
    for (;;) { /* UntilAccepted */
      /* ros_PrepareMatrix */
      int nconsec = 0;
      for (;;) {
        for (int k = 0; k < nz; ++k) Ghimj[k] = -Jac0[k];
        double ghinv = 1.0 / (dir * H * RGAMMA[0]);
        for (int i = 0; i < N; ++i) Ghimj[m->lu_diag[i]] += ghinv;
        int ising = orc_decomp(m, Ghimj);
        ISTATUS[5]++;
        if (ising == 0) break;
        ISTATUS[7]++;
        nconsec++;
        if (nconsec <= 5) H *= 0.5;
        else { *IERR = -8; goto done; }
      }
      for (int s = 0; s < 4; ++s) {
        double *Ks = K + (size_t)s * N;
        if (s == 0) {
          memcpy(Fcn, Fcn0, vb);
        } else if (RNEWF[s]) {
          memcpy(Ynew, Y, vb);
          for (int j = 0; j < s; ++j) waxpy(N, RA[s * (s - 1) / 2 + j], K + (size_t)j * N, Ynew);
          orc_fun(m, Ynew, F, RCT, Fcn);
          ISTATUS[0]++;
        }
        memcpy(Ks, Fcn, vb);
        for (int j = 0; j < s; ++j) {
          double hc = RC_[s * (s - 1) / 2 + j] / (dir * H);
          waxpy(N, hc, K + (size_t)j * N, Ks);
        }
        orc_solve(m, Ghimj, Ks);
        ISTATUS[6]++;
      }
      memcpy(Ynew, Y, vb);
      for (int j = 0; j < 4; ++j) waxpy(N, RM[j], K + (size_t)j * N, Ynew);
      memset(Yerr, 0, vb);
      for (int j = 0; j < 4; ++j) waxpy(N, RE[j], K + (size_t)j * N, Yerr);
      double err = err_norm(N, Y, Ynew, Yerr, atol, rtol, vectol);
      double fac = min_f(FacMax, max_f(FacMin, FacSafe / pow(err, 1.0 / RELO)));
      Hnew = H * fac;
      ISTATUS[2]++;
      if (err <= 1.0 || H <= Hmin) {
        ISTATUS[3]++;
        if (ICNTRL[15] == 1) {
          for (int i = 0; i < N; ++i) Y[i] = max_f(Ynew[i], 0.0);
        } else {
          memcpy(Y, Ynew, vb);
        }
        T = T + dir * H;
        Hnew = max_f(Hmin, min_f(Hnew, Hmax));
        if (rejLast) Hnew = min_f(Hnew, H);
        RSTATUS[1] = H;
        RSTATUS[2] = Hnew;
        RSTATUS[0] = T;
        rejLast = 0;
        rejMore = 0;
        H = Hnew;
        break;
      } else {
        if (rejMore) Hnew = H * FacRej;
        rejMore = rejLast;
        rejLast = 1;
        H = Hnew;
        if (ISTATUS[3] >= 1) ISTATUS[4]++;
      }
    }
  }
done:
  free(Ynew); free(Fcn0); free(Fcn); free(Yerr); free(K); free(Jac0); free(Ghimj);
}

void orc_integrate(const orc_mech *m, int64_t ncell, double *conc,
                   const double *rct, const double *hstart, double tin,
                   double tout, const double *atol, const double *rtol,
                   const int *icntrl, const double *rcntrl, int *istatus,
                   double *rstatus, int *ierr, int nthreads) {
  int nspec = m->nvar + m->nfix;
#pragma omp parallel for schedule(dynamic, 24) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t c = 0; c < ncell; ++c)
    integrate_cell(m, conc + c * nspec, rct + c * m->nreact, hstart[c], tin,
                   tout, atol, rtol, icntrl, rcntrl, istatus + c * 20,
                   rstatus + c * 20, ierr + c);
}

/* The carbon mechanism's Integrate: one forward-Euler step
 * (KPP/carbon/gckpp_Integrator.F90:93-150 and ForwardEuler :155-215).
 * The reference's Fun there is the aggregate form (A(r) then Vdot = sum of
 * stoichiometric terms), which orc_fun equals up to summation order. */
void orc_euler(const orc_mech *m, int64_t ncell, double *conc, const double *rct,
               double tin, double tout, const int *icntrl, int *ierr) {
  int nspec = m->nvar + m->nfix;
  int icu16 = icntrl ? icntrl[15] : 0;
  double *Ynew = malloc(sizeof(double) * m->nvar);
  double *dYdt = malloc(sizeof(double) * m->nvar);
  for (int64_t c = 0; c < ncell; ++c) {
    double *Y = conc + c * nspec;
    orc_fun(m, Y, Y + m->nvar, rct + c * m->nreact, dYdt);
    for (int i = 0; i < m->nvar; ++i) Ynew[i] = Y[i] + dYdt[i] * (tout - tin);
    int e = 1, stop = 0;
    if (icu16 > 0)
      for (int i = 0; i < m->nvar && !stop; ++i)
        if (Ynew[i] < 0.0) {
          if (icu16 == 1) Ynew[i] = 0.0;
          else if (icu16 == 2 || icu16 == 3) { e = -9; stop = 1; }
        }
    if (!stop) memcpy(Y, Ynew, sizeof(double) * m->nvar);
    ierr[c] = e;
  }
  free(Ynew);
  free(dYdt);
}
