/*
 * oracle/riemann.c -- CPU restatement of M2C's exact two-material Riemann solver (test
 * infrastructure only).
 *
 *   driver (secant on p*)         ExactRiemannSolverBase.cpp:59-342 (adaptive), :2199-2474 (fixed step)
 *   star-state finalisation       :347-514
 *   bracketing interval           :631-795
 *   feasible starting points      :959-1114 (acoustic theory :1064-1114)
 *   1-/3-wave connection          :1169-1580 (adaptive), :2476-2848 (fixed step)
 *   Hugoniot residual             ExactRiemannSolverBase.h:58-75
 *   embedded RK (Cash-Karp)       :1729-1852; classic RK4 :2850-2946
 *   Hugoniot root                 the WITHOUT_BOOST secant/bisection branch :1501-1543 (Boost's
 *                                 TOMS748 is not available; see the design notes)
 *   interface face                SpaceOperator.cpp:2486-2542 (mesh normal, clip, store, fluxes)
 * Deviation (documented): the adaptive step-size control's pow(x, 0.25) and pow(x, 0.2)
 * (ExactRiemannSolverBase.cpp:1286-1310) are evaluated with orc_pow_quarter / orc_pow_fifth below,
 * deterministic routines built from +,-,*,/ and sqrt that the GPU code evaluates with the same
 * operations; glibc's pow differs from them by a few ulps, which only perturbs the adaptive step.
 * Parity pins (tests/test_oracle_riemann.py): stiffened-gas closed form (Toro's solver with
 * p -> p + p_inf, Newton to 1e-14) for p*, u* to the solver tolerance; Sod star state; symmetric
 * problems give u* = 0; the trivial case returns the input.
 */
#include "oracle_internal.h"
#include "flux.h"
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>

/* ------------------------------------------------------------ deterministic powers ------ */

static double orc_bits_ldexp(double m, int e) {   /* m * 2^e for |e| < 1000 (two exact steps) */
  int e1 = e / 2, e2 = e - e1;
  uint64_t b1 = (uint64_t)(1023 + e1) << 52, b2 = (uint64_t)(1023 + e2) << 52;
  double s1, s2;
  memcpy(&s1, &b1, 8);
  memcpy(&s2, &b2, 8);
  return m * s1 * s2;
}

/* natural log of a positive normal double: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
 * log m = 2 atanh(s), s = (m-1)/(m+1), series to s^23 */
double orc_det_log(double x) {
  uint64_t b;
  memcpy(&b, &x, 8);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  b = (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
  double m;
  memcpy(&m, &b, 8);
  if (m > 1.4142135623730951) { m = m * 0.5; e = e + 1; }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  double t = 1.0 / 23.0;
  t = t * s2 + 1.0 / 21.0;
  t = t * s2 + 1.0 / 19.0;
  t = t * s2 + 1.0 / 17.0;
  t = t * s2 + 1.0 / 15.0;
  t = t * s2 + 1.0 / 13.0;
  t = t * s2 + 1.0 / 11.0;
  t = t * s2 + 1.0 / 9.0;
  t = t * s2 + 1.0 / 7.0;
  t = t * s2 + 1.0 / 5.0;
  t = t * s2 + 1.0 / 3.0;
  t = t * s2 + 1.0;
  return (double)e * 0.6931471805599453 + 2.0 * s * t;
}

/* exp(y) for |y| < 700: y = n ln2 + r, |r| <= ln2/2, Taylor series to r^14 */
double orc_det_exp(double y) {
  double nf = floor(y * 1.4426950408889634 + 0.5);
  int n = (int)nf;
  double r = y - nf * 0.6931471805599453;
  double t = 1.0 / 87178291200.0;          /* 1/14! */
  t = t * r + 1.0 / 6227020800.0;
  t = t * r + 1.0 / 479001600.0;
  t = t * r + 1.0 / 39916800.0;
  t = t * r + 1.0 / 3628800.0;
  t = t * r + 1.0 / 362880.0;
  t = t * r + 1.0 / 40320.0;
  t = t * r + 1.0 / 5040.0;
  t = t * r + 1.0 / 720.0;
  t = t * r + 1.0 / 120.0;
  t = t * r + 1.0 / 24.0;
  t = t * r + 1.0 / 6.0;
  t = t * r + 0.5;
  t = t * r + 1.0;
  t = t * r + 1.0;
  return orc_bits_ldexp(t, n);
}

double orc_pow_quarter(double x) { return sqrt(sqrt(x)); }

double orc_pow_fifth(double x) {
  if (!(x > 2.2250738585072014e-308) || !(x < 1e300)) return pow(x, 0.2);
  return orc_det_exp(0.2 * orc_det_log(x));
}

/* ------------------------------------------------------------ solver state -------------- */

typedef struct { double *p, *rho, *u; int n, cap; } orc_path;

typedef struct {
  const orc_eos *mat;
  orc_riemann_cfg cfg;
  int nonadaptive;
  orc_path path1, path3;
} orc_rs;

static void path_reset(orc_path *P, double p, double rho, double u) {
  P->n = 0;
  if (P->cap < 16) {
    P->cap = 512;
    P->p = (double *)realloc(P->p, sizeof(double) * P->cap);
    P->rho = (double *)realloc(P->rho, sizeof(double) * P->cap);
    P->u = (double *)realloc(P->u, sizeof(double) * P->cap);
  }
  P->p[0] = p; P->rho[0] = rho; P->u[0] = u; P->n = 1;
}

static void path_push(orc_path *P, double p, double rho, double u) {
  if (P->n == P->cap) {
    P->cap *= 2;
    P->p = (double *)realloc(P->p, sizeof(double) * P->cap);
    P->rho = (double *)realloc(P->rho, sizeof(double) * P->cap);
    P->u = (double *)realloc(P->u, sizeof(double) * P->cap);
  }
  P->p[P->n] = p; P->rho[P->n] = rho; P->u[P->n] = u; P->n++;
}

static double dmin(double a, double b) { return (b < a) ? b : a; }   /* std::min */
static double dmax(double a, double b) { return (a < b) ? b : a; }   /* std::max */

static double c2_of(const orc_eos *m, double rho, double e) { return orc_c2(m, rho, e); }

/* ------------------------------------------------------------ RK steps ------------------ */

/* Cash-Karp embedded RK step along the isentrope (ExactRiemannSolverBase.cpp:1729-1852) */
static int rk_adaptive(const orc_eos *m, int wave, double rho_0, double u_0, double p_0, double dp,
                       double *rho, double *u, double *p, double *xi, double *uErr, double *rhoErr) {
  dp = -dp;
  double e_0 = orc_internal_energy(m, rho_0, p_0);
  double c_0_square = c2_of(m, rho_0, e_0);
  if (rho_0 <= 0 || c_0_square < 0) return 0;
  double c_0 = sqrt(c_0_square);
  double rho_1 = rho_0 + 0.2 * dp / c_0_square;
  double p_1 = p_0 + 0.2 * dp;
  double e_1 = orc_internal_energy(m, rho_1, p_1);
  double c_1_square = c2_of(m, rho_1, e_1);
  if (rho_1 <= 0 || c_1_square < 0) return 0;
  double rho_2 = rho_0 + 3. / 40. * dp / c_0_square + 9. / 40. * dp / c_1_square;
  double p_2 = p_1 + 3. / 10. * dp;
  double e_2 = orc_internal_energy(m, rho_2, p_2);
  double c_2_square = c2_of(m, rho_2, e_2);
  if (rho_2 <= 0 || c_2_square < 0) return 0;
  double c_2 = sqrt(c_2_square);
  double rho_3 = rho_0 + 3. / 10. * dp / c_0_square - 9. / 10. * dp / c_1_square + 6. / 5. * dp / c_2_square;
  double p_3 = p_0 + 3. / 5. * dp;
  double e_3 = orc_internal_energy(m, rho_3, p_3);
  double c_3_square = c2_of(m, rho_3, e_3);
  if (rho_3 <= 0 || c_3_square < 0) return 0;
  double c_3 = sqrt(c_3_square);
  double rho_4 = rho_0 - 11. / 54. * dp / c_0_square + 2.5 * dp / c_1_square - 70. / 27. * dp / c_2_square
               + 35. / 27. * dp / c_3_square;
  double p_4 = p_0 + dp;
  double e_4 = orc_internal_energy(m, rho_4, p_4);
  double c_4_square = c2_of(m, rho_4, e_4);
  if (rho_4 <= 0 || c_4_square < 0) return 0;
  double c_4 = sqrt(c_4_square);
  double rho_5 = rho_0 + 1631. / 55296. * dp / c_0_square + 175. / 512. * dp / c_1_square
               + 575. / 13824. * dp / c_2_square + 44275. / 110592. * dp / c_3_square
               + 253. / 4096. * dp / c_4_square;
  double p_5 = p_0 + 7. / 8. * dp;
  double e_5 = orc_internal_energy(m, rho_5, p_5);
  double c_5_square = c2_of(m, rho_5, e_5);
  if (rho_5 <= 0 || c_5_square < 0) return 0;
  double c_5 = sqrt(c_5_square);
  double drho = (37. / 378. / c_0_square + 250. / 621. / c_2_square + 125. / 594. / c_3_square
               + 512. / 1771. / c_5_square) * dp;
  double drho_err = (2825. / 27648. / c_0_square + 18575. / 48384. / c_2_square + 13525. / 55296. / c_3_square
                   + 277. / 14336. / c_4_square + 0.25 / c_5_square) * dp;
  double du = (37. / 378. / c_0 / rho_0 + 250. / 621. / c_2 / rho_2 + 125. / 594. / c_3 / rho_3
             + 512. / 1771. / c_5 / rho_5) * dp;
  double du_err = (2825. / 27648. / c_0 / rho_0 + 18575. / 48384. / c_2 / rho_2 + 13525. / 55296. / c_3 / rho_3
                 + 277. / 14336. / c_4 / rho_4 + 0.25 / c_5 / rho_5) * dp;
  if (isnan(du) || isnan(du_err)) return 0;
  *rhoErr = fabs(drho_err - drho);
  *uErr = fabs(du_err - du);
  *rho = rho_0 + drho;
  *p = p_0 + dp;
  *u = (wave == 1) ? u_0 - du : u_0 + du;
  double e = orc_internal_energy(m, *rho, *p);
  double c = c2_of(m, *rho, e);
  if (*rho <= 0 || c < 0) return 0;
  c = sqrt(c);
  *xi = (wave == 1) ? *u - c : *u + c;
  return 1;
}

/* classic RK4 step (ExactRiemannSolverBase.cpp:2850-2946) */
static int rk_classic(const orc_eos *m, int wave, double rho_0, double u_0, double p_0, double dp,
                      double *rho, double *u, double *p, double *xi) {
  dp = -dp;
  double e_0 = orc_internal_energy(m, rho_0, p_0);
  double c_0_square = c2_of(m, rho_0, e_0);
  if (rho_0 <= 0 || c_0_square < 0) return 0;
  double c_0 = sqrt(c_0_square);
  double rho_1 = rho_0 + 0.5 * dp / c_0_square;
  double p_1 = p_0 + 0.5 * dp;
  double e_1 = orc_internal_energy(m, rho_1, p_1);
  double c_1_square = c2_of(m, rho_1, e_1);
  if (rho_1 <= 0 || c_1_square < 0) return 0;
  double c_1 = sqrt(c_1_square);
  double rho_2 = rho_0 + 0.5 * dp / c_1_square;
  double p_2 = p_1;
  double e_2 = orc_internal_energy(m, rho_2, p_2);
  double c_2_square = c2_of(m, rho_2, e_2);
  if (rho_2 <= 0 || c_2_square < 0) return 0;
  double c_2 = sqrt(c_2_square);
  double rho_3 = rho_0 + dp / c_2_square;
  double p_3 = p_0 + dp;
  double e_3 = orc_internal_energy(m, rho_3, p_3);
  double c_3_square = c2_of(m, rho_3, e_3);
  if (rho_3 <= 0 || c_3_square < 0) return 0;
  double c_3 = sqrt(c_3_square);
  double drho = 1.0 / 6.0 * dp * (1. / c_0_square + 2.0 * (1. / c_1_square + 1. / c_2_square) + 1. / c_3_square);
  double du = 1.0 / 6.0 * dp * (1. / c_0 / rho_0 + 2.0 * (1. / c_1 / rho_1 + 1. / c_2 / rho_2) + 1. / c_3 / rho_3);
  if (isnan(du)) return 0;
  *rho = rho_0 + drho;
  *p = p_0 + dp;
  *u = (wave == 1) ? u_0 - du : u_0 + du;
  double e = orc_internal_energy(m, *rho, *p);
  double c = c2_of(m, *rho, e);
  if (*rho <= 0 || c < 0) return 0;
  c = sqrt(c);
  *xi = (wave == 1) ? *u - c : *u + c;
  return 1;
}

/* ------------------------------------------------------------ wave connection ----------- */

/* Hugoniot residual (ExactRiemannSolverBase.h:58-75) */
typedef struct { const orc_eos *m; double e, ps, pavg, one_over_rho; } hugo_t;
static double hugo(const hugo_t *h, double rhos) {
  double es = orc_internal_energy(h->m, rhos, h->ps);
  return es - h->e + h->pavg * (1.0 / rhos - h->one_over_rho);
}

static int rho_u_star(orc_rs *S, int wave, orc_path *P, double rho, double u, double p, double ps, int id,
                      double rhos0, double rhos1, double *rhos_out, double *us_out, int *trans_rare,
                      double *Vrare) {
  const orc_eos *m = &S->mat[id];
  const orc_riemann_cfg *C = &S->cfg;
  *rhos_out = rho;
  *us_out = u;
  if (p > ps) {
    double rhos_0 = rho, us_0 = u, ps_0 = p, xi_0;
    double rhos_1 = rho, us_1 = u, ps_1 = p, xi_1 = 0.0;
    double dp, dp_min_adaption = 0.0, dp_target = 0.0;
    double e = orc_internal_energy(m, rho, p);
    if (!S->nonadaptive) {
      dp = (p - ps) / C->numSteps_rarefaction;
      dp_min_adaption = dp / 2.5;
    } else {
      dp_target = (p - ps) / C->numSteps_rarefaction;
      double dpdrho = orc_dpdrho(m, rho, e);
      double drho_op1 = (p - ps) / dpdrho / C->numSteps_rarefaction;
      double drho_op2 = rho / (2.5 * C->numSteps_rarefaction);
      double drho_target = dmin(drho_op1, drho_op2);
      dp = dmin(dp_target, drho_target * dpdrho);
    }
    double pressure_endpoint_tol = C->tol_rarefaction * dmax(1.0, fabs(p));
    double c = c2_of(m, rho, e);
    if (rho <= 0 || c < 0) return 0;
    c = sqrt(c);
    int index0 = 0;
    if (P->n > 1) {
      for (int j = P->n - 1; j >= 0; j--)
        if (P->p[j] > ps) { index0 = j; break; }
      ps_0 = P->p[index0];
      rhos_0 = P->rho[index0];
      us_0 = P->u[index0];
      if (!S->nonadaptive) {
        dp_min_adaption = (ps_0 - ps) / C->numSteps_rarefaction / 2.5;
        if (index0 != P->n - 1) dp = ps_0 - ps;
        else dp = dmin(P->p[index0 - 1] - P->p[index0], ps_0 - ps);
      } else {
        dp = dmin((p - ps) / C->numSteps_rarefaction, ps_0 - ps);
      }
    }
    double xi = (wave == 1) ? u - c : u + c;
    xi_0 = xi;
    int done = 0;
    double uErr = 0., rhoErr = 0.;
    int moreSteps = S->nonadaptive ? 5 : 3;
    for (int i = 0; i < C->numSteps_rarefaction * moreSteps; i++) {
      if (fabs(ps_1 - ps) <= pressure_endpoint_tol) {
        *rhos_out = rhos_1;
        *us_out = us_1;
        if (orc_check_state_rp(m, *rhos_out, ps)) return 0;
        done = 1;
        break;
      }
      int ok;
      if (!S->nonadaptive)
        ok = rk_adaptive(m, wave, rhos_0, us_0, ps_0, dp, &rhos_1, &us_1, &ps_1, &xi_1, &uErr, &rhoErr);
      else
        ok = rk_classic(m, wave, rhos_0, us_0, ps_0, dp, &rhos_1, &us_1, &ps_1, &xi_1);
      if (!ok) {
        dp = dp / 2.0;
        if (S->nonadaptive && dp == 0.0) break;
        continue;
      }
      if (ps_1 - ps < -pressure_endpoint_tol) {
        dp = ps_0 - ps;
        continue;
      }
      if (!S->nonadaptive) {
        double tiny = 1.e-30;
        double errBar = dmax(1.0e-8, dmin(1.0e-5, 100.0 * C->tol_rarefaction));
        double uErrScaled = uErr / c + tiny;
        double rhoErrScaled = rhoErr / rho + tiny;
        if (isnan(rhoErrScaled) || isnan(uErrScaled)) return -1;   /* the reference exits */
        double dpTemp = dp;
        double safety = 0.9;
        if (rhoErrScaled >= errBar && dp > dp_min_adaption) {
          dpTemp = safety * dp * orc_pow_quarter(fabs(errBar / rhoErrScaled));
          dpTemp = dmax(dmax(dpTemp, 0.2 * dp), dp_min_adaption);
          dp = dmin(dpTemp, ps_0 - ps);
          continue;
        }
        if (uErrScaled >= errBar && dp > dp_min_adaption) {
          dpTemp = safety * dp * orc_pow_quarter(fabs(errBar / uErrScaled));
          dpTemp = dmax(dmax(dpTemp, 0.2 * dp), dp_min_adaption);
          dp = dmin(dpTemp, ps_0 - ps);
          continue;
        }
        if (ps_1 < P->p[P->n - 1]) path_push(P, ps_1, rhos_1, us_1);
        if (trans_rare && Vrare && xi_0 * xi_1 <= 0) {
          *trans_rare = 1;
          double w0 = fabs(xi_1), w1 = fabs(xi_0);
          double ww = w0 + w1;
          w0 /= ww;
          w1 /= ww;
          Vrare[0] = w0 * rhos_0 + w1 * rhos_1;
          Vrare[1] = w0 * us_0 + w1 * us_1;
          Vrare[2] = w0 * ps_0 + w1 * ps_1;
        }
        double dpTemp_rho = safety * dp * orc_pow_fifth(fabs(errBar / rhoErrScaled));
        double dpTemp_u = safety * dp * orc_pow_fifth(fabs(errBar / uErrScaled));
        dpTemp = dmin(dpTemp_rho, dpTemp_u);
        dpTemp = dmin(dpTemp, 10 * dp);
        dpTemp = dmax(dpTemp, dp_min_adaption);
        dp = dmin(dpTemp, ps_1 - ps);
      } else {
        if (ps_1 < P->p[P->n - 1]) path_push(P, ps_1, rhos_1, us_1);
        if (trans_rare && Vrare && xi_0 * xi_1 <= 0) {
          *trans_rare = 1;
          double w0 = fabs(xi_1), w1 = fabs(xi_0);
          double ww = w0 + w1;
          w0 /= ww;
          w1 /= ww;
          Vrare[0] = w0 * rhos_0 + w1 * rhos_1;
          Vrare[1] = w0 * us_0 + w1 * us_1;
          Vrare[2] = w0 * ps_0 + w1 * ps_1;
        }
        dp = dmin(dmin(dp_target, ps_1 - ps), 4.0 * dp);
      }
      rhos_0 = rhos_1;
      us_0 = us_1;
      ps_0 = ps_1;
      xi_0 = xi_1;
    }
    if (!done) return 0;
  } else {
    /* shock: bracket then secant/bisection on the Hugoniot residual (:1380-1580) */
    hugo_t H;
    H.m = m;
    H.e = orc_internal_energy(m, rho, p);
    H.ps = ps;
    H.pavg = 0.5 * (p + ps);
    H.one_over_rho = 1.0 / rho;
    double f0, f1;
    double drho = dmax(fabs(rhos0 - rhos1), 0.001 * rhos0);
    int found0 = 0, found1 = 0;
    if (dmin(rhos0, rhos1) >= rho) {
      f0 = hugo(&H, rhos0);
      f1 = hugo(&H, rhos1);
      if (f0 * f1 <= 0) {
        if (rhos0 > rhos1) {
          double t = rhos0; rhos0 = rhos1; rhos1 = t;
          t = f0; f0 = f1; f1 = t;
        }
        found0 = found1 = 1;
      } else {
        rhos0 = rhos1;
        f0 = f1;
      }
    } else {
      if (rhos1 > rhos0) {
        rhos0 = rhos1;
        f0 = hugo(&H, rhos0);
      }
      if (rhos0 < rho) {
        rhos0 = rho;
        f0 = hugo(&H, rhos0);
        found0 = 1;
      } else {
        f0 = hugo(&H, rhos0);
      }
    }
    if (!found0 || !found1) {
      int i = 0;
      double factor = 1.5;
      double tmp, ftmp;
      rhos1 = rhos0;
      f1 = f0;
      while (!found0) {
        if (++i >= C->maxIts_shock) return 0;
        tmp = rhos1;
        ftmp = f1;
        rhos1 = rhos0;
        f1 = f0;
        rhos0 = rhos1 - factor * drho;
        if (rhos0 <= rho) {
          rhos0 = rho;
          found0 = 1;
        }
        f0 = hugo(&H, rhos0);
        if (f0 * f1 <= 0) {
          found0 = found1 = 1;
        } else {
          rhos1 = tmp;
          f1 = ftmp;
          tmp = rhos0;
          ftmp = f0;
          rhos0 = rhos1;
          f0 = f1;
          rhos1 = rhos0 + factor * drho;
          f1 = hugo(&H, rhos1);
          if (f0 * f1 <= 0) {
            found0 = found1 = 1;
          } else {
            rhos0 = tmp;
            f0 = ftmp;
            drho = rhos1 - rhos0;
          }
        }
      }
      if (!found1) {
        i = 0;
        factor = 2.5;
        while (!found1) {
          if (++i >= C->maxIts_shock) return 0;
          rhos0 = rhos1;
          f0 = f1;
          rhos1 = rhos0 + factor * drho;
          f1 = hugo(&H, rhos1);
          if (f0 * f1 <= 0) found1 = 1;
          else drho = rhos1 - rhos0;
        }
      }
    }
    double sol_first, sol_second;
    double loc_tol_shock = C->tol_shock * dmin(rhos0, rhos1);
    if (f0 == 0.0) sol_first = sol_second = rhos0;
    else if (f1 == 0.0) sol_first = sol_second = rhos1;
    else {
      double rhos2 = rhos1, f2 = f1;
      int it;
      sol_first = sol_second = rhos1;
      for (it = 0; it < C->maxIts_shock; it++) {
        drho = rhos1 - rhos0;
        rhos2 = rhos2 - f2 * (rhos1 - rhos0) / (f1 - f0);
        if (rhos2 >= rhos1 || rhos2 <= rhos0) rhos2 = 0.5 * (rhos0 + rhos1);
        f2 = hugo(&H, rhos2);
        if (f2 == 0.0) { sol_first = sol_second = rhos2; break; }
        if (f2 * f0 < 0) { rhos1 = rhos2; f1 = f2; }
        else { rhos0 = rhos2; f0 = f2; }
        if (rhos1 - rhos0 < loc_tol_shock) { sol_first = rhos0; sol_second = rhos1; break; }
      }
      if (it == C->maxIts_shock) return 0;
    }
    double rhos = 0.5 * (sol_first + sol_second);
    double du = -(ps - p) * (1.0 / rhos - 1.0 / rho);
    if (du < 0) return 0;
    if (orc_check_state_rp(m, rhos, ps)) return 0;
    *rhos_out = rhos;
    *us_out = (wave == 1) ? u - sqrt(du) : u + sqrt(du);
  }
  return 1;
}

/* ------------------------------------------------------------ starting points ----------- */

static int feasible_acoustic(orc_rs *S, double rhol, double ul, double pl, double cl, int idl, double rhor,
                             double ur, double pr, double cr, int idr, double *p0, double *rhol0, double *rhor0,
                             double *ul0, double *ur0, double *p1, double *rhol1, double *rhor1, double *ul1,
                             double *ur1) {
  int found = 0;
  double Cl = rhol * cl, Cr = rhor * cr;
  *p0 = (Cr * pl + Cl * pr + Cl * Cr * (ul - ur)) / (Cl + Cr);
  if (rho_u_star(S, 1, &S->path1, rhol, ul, pl, *p0, idl, rhol, (*p0 > pl) ? rhol * 1.1 : rhol * 0.9, rhol0, ul0,
                 NULL, NULL) <= 0)
    return found;
  if (rho_u_star(S, 3, &S->path3, rhor, ur, pr, *p0, idr, rhor, (*p0 > pr) ? rhor * 1.1 : rhor * 0.9, rhor0, ur0,
                 NULL, NULL) <= 0)
    return found;
  found = 1;
  double Clbar = (*ul0 == ul) ? Cl : fabs(*p0 - pl) / fabs(*ul0 - ul);
  double Crbar = (*ur0 == ur) ? Cr : fabs(*p0 - pr) / fabs(*ur0 - ur);
  *p1 = (Crbar * pl + Clbar * pr + Clbar * Crbar * (ul - ur)) / (Clbar + Crbar);
  double tmp = dmax(fabs(*p0), fabs(*p1));
  if (fabs(*p1 - *p0) / tmp < 1.0e-8) *p1 = *p0 + 1.0e-8 * tmp;
  if (rho_u_star(S, 1, &S->path1, rhol, ul, pl, *p1, idl, rhol, *rhol0, rhol1, ul1, NULL, NULL) <= 0) return found;
  if (rho_u_star(S, 3, &S->path3, rhor, ur, pr, *p1, idr, rhor, *rhor0, rhor1, ur1, NULL, NULL) <= 0) return found;
  return 2;
}

#define SWAPD(a, b) do { double t_ = (a); (a) = (b); (b) = t_; } while (0)

static int feasible(orc_rs *S, double rhol, double ul, double pl, double cl, int idl, double rhor, double ur,
                    double pr, double cr, int idr, double *p0, double *rhol0, double *rhor0, double *ul0,
                    double *ur0, double *p1, double *rhol1, double *rhor1, double *ul1, double *ur1) {
  const orc_riemann_cfg *C = &S->cfg;
  int ok = 1;
  int found = feasible_acoustic(S, rhol, ul, pl, cl, idl, rhor, ur, pr, cr, idr, p0, rhol0, rhor0, ul0, ur0, p1,
                                rhol1, rhor1, ul1, ur1);
  if (found == 2) return 1;
  if (found != 1) {
    double dp = (pl != pr) ? fabs(pl - pr) : 0.5 * pl;
    for (int i = 0; i < C->maxIts_bracket; i++) {
      *p0 = dmin(pl, pr) + 0.01 * (i + 1) * (i + 1) * dmin(dp, dmin(fabs(pl), fabs(pr)));
      if (*p0 < C->min_pressure) *p0 = C->pressure_at_failure;
      ok = rho_u_star(S, 1, &S->path1, rhol, ul, pl, *p0, idl, rhol, (*p0 > pl) ? rhol * 1.1 : rhol * 0.9, rhol0,
                      ul0, NULL, NULL) > 0;
      ok = ok && rho_u_star(S, 3, &S->path3, rhor, ur, pr, *p0, idr, rhor, (*p0 > pr) ? rhor * 1.1 : rhor * 0.9,
                            rhor0, ur0, NULL, NULL) > 0;
      if (ok) break;
    }
    if (!ok) {
      for (int i = 0; i < C->maxIts_bracket; i++) {
        *p0 = dmin(pl, pr) - 0.01 * (i + 1) * (i + 1) * dmin(dp, dmin(fabs(pl), fabs(pr)));
        if (*p0 < C->min_pressure || i == (int)(C->maxIts_bracket / 2)) *p0 = C->pressure_at_failure;
        ok = rho_u_star(S, 1, &S->path1, rhol, ul, pl, *p0, idl, rhol, (*p0 > pl) ? rhol * 1.1 : rhol * 0.9,
                        rhol0, ul0, NULL, NULL) > 0;
        ok = ok && rho_u_star(S, 3, &S->path3, rhor, ur, pr, *p0, idr, rhor, (*p0 > pr) ? rhor * 1.1 : rhor * 0.9,
                              rhor0, ur0, NULL, NULL) > 0;
        if (ok) break;
      }
    }
    if (!ok) return 0;
  }
  {
    double dp = dmin(fabs(*p0 - pl), fabs(*p0 - pr));
    for (int i = 0; i < C->maxIts_bracket; i++) {
      *p1 = *p0 + 0.01 * (i + 1) * (i + 1) * dp;
      ok = rho_u_star(S, 1, &S->path1, rhol, ul, pl, *p1, idl, rhol, *rhol0, rhol1, ul1, NULL, NULL) > 0;
      ok = ok && rho_u_star(S, 3, &S->path3, rhor, ur, pr, *p1, idr, rhor, *rhor0, rhor1, ur1, NULL, NULL) > 0;
      if (ok) break;
    }
    if (!ok)
      for (int i = 0; i < C->maxIts_bracket; i++) {
        *p1 = *p0 - 0.01 * (i + 1) * (i + 1) * dp;
        if (*p1 < C->min_pressure) *p1 = C->pressure_at_failure * 1000.0;
        ok = rho_u_star(S, 1, &S->path1, rhol, ul, pl, *p1, idl, rhol, *rhol0, rhol1, ul1, NULL, NULL) > 0;
        ok = ok && rho_u_star(S, 3, &S->path3, rhor, ur, pr, *p1, idr, rhor, *rhor0, rhor1, ur1, NULL, NULL) > 0;
        if (ok) break;
      }
    if (!ok) return 0;
  }
  if (*p0 > *p1) {
    SWAPD(*p0, *p1); SWAPD(*rhol0, *rhol1); SWAPD(*rhor0, *rhor1); SWAPD(*ul0, *ul1); SWAPD(*ur0, *ur1);
  }
  return 1;
}

static int initial_interval(orc_rs *S, double rhol, double ul, double pl, double cl, int idl, double rhor,
                            double ur, double pr, double cr, int idr, double *p0, double *rhol0, double *rhor0,
                            double *ul0, double *ur0, double *p1, double *rhol1, double *rhor1, double *ul1,
                            double *ur1) {
  const orc_riemann_cfg *C = &S->cfg;
  if (!feasible(S, rhol, ul, pl, cl, idl, rhor, ur, pr, cr, idr, p0, rhol0, rhor0, ul0, ur0, p1, rhol1, rhor1, ul1,
                ur1)) {
    *p0 = *p1 = C->pressure_at_failure;
    return 0;
  }
  double p2 = 0, rhol2 = 0, rhor2 = 0, ul2 = 0, ur2 = 0, f0, f1;
  int i, ok = 1;
  double fmin, p_fmin, rhol_fmin, rhor_fmin, ul_fmin, ur_fmin;
  if (fabs(*ul0 - *ur0) < fabs(*ul1 - *ur1)) {
    fmin = fabs(*ul0 - *ur0);
    p_fmin = *p0; rhol_fmin = *rhol0; rhor_fmin = *rhor0; ul_fmin = *ul0; ur_fmin = *ur0;
  } else {
    fmin = fabs(*ul1 - *ur1);
    p_fmin = *p1; rhol_fmin = *rhol1; rhor_fmin = *rhor1; ul_fmin = *ul1; ur_fmin = *ur1;
  }
  for (i = 0; i < C->maxIts_bracket; i++) {
    f0 = *ul0 - *ur0;
    f1 = *ul1 - *ur1;
    if (f0 * f1 <= 0.0) return 1;
    if (fabs(f0 - f1) > 1e-9) {
      p2 = *p1 - f1 * (*p1 - *p0) / (f1 - f0);
      if (p2 < *p0) p2 -= 0.1 * (*p1 - *p0);
      else p2 += 0.1 * (*p1 - *p0);
    } else {
      p2 = 1.1 * *p1;
    }
    if (p2 < C->min_pressure || i == (int)(C->maxIts_bracket / 2)) p2 = 1.0e-8;
    ok = rho_u_star(S, 1, &S->path1, rhol, ul, pl, p2, idl, *rhol0, *rhol1, &rhol2, &ul2, NULL, NULL) > 0;
    ok = ok && rho_u_star(S, 3, &S->path3, rhor, ur, pr, p2, idr, *rhor0, *rhor1, &rhor2, &ur2, NULL, NULL) > 0;
    if (!ok) {
      for (int j = 0; j < C->maxIts_bracket / 2; j++) {
        if (p2 < *p0) p2 = *p0 - 0.5 * (*p0 - p2);
        else p2 = *p1 + 0.5 * (p2 - *p1);
        ok = rho_u_star(S, 1, &S->path1, rhol, ul, pl, p2, idl, *rhol0, *rhol1, &rhol2, &ul2, NULL, NULL) > 0;
        ok = ok && rho_u_star(S, 3, &S->path3, rhor, ur, pr, p2, idr, *rhor0, *rhor1, &rhor2, &ur2, NULL, NULL) > 0;
        if (ok) break;
      }
    }
    if (!ok) break;
    if (fabs(ul2 - ur2) < fmin) {
      fmin = fabs(ul2 - ur2);
      p_fmin = p2; rhol_fmin = rhol2; rhor_fmin = rhor2; ul_fmin = ul2; ur_fmin = ur2;
    }
    if (p2 < *p0) {
      *p1 = *p0; *rhol1 = *rhol0; *rhor1 = *rhor0; *ul1 = *ul0; *ur1 = *ur0;
      *p0 = p2; *rhol0 = rhol2; *rhor0 = rhor2; *ul0 = ul2; *ur0 = ur2;
    } else {
      *p0 = *p1; *rhol0 = *rhol1; *rhor0 = *rhor1; *ul0 = *ul1; *ur0 = *ur1;
      *p1 = p2; *rhol1 = rhol2; *rhor1 = rhor2; *ul1 = ul2; *ur1 = ur2;
    }
  }
  if (!ok || i == C->maxIts_bracket) {
    if (fmin < C->failure_threshold * fabs(ul - ur)) {
      *p0 = *p1 = p_fmin; *rhol0 = *rhol1 = rhol_fmin; *rhor0 = *rhor1 = rhor_fmin;
      *ul0 = *ul1 = ul_fmin; *ur0 = *ur1 = ur_fmin;
    } else {
      p2 = C->pressure_at_failure;
      ok = rho_u_star(S, 1, &S->path1, rhol, ul, pl, p2, idl, *rhol0, *rhol1, &rhol2, &ul2, NULL, NULL) > 0;
      ok = ok && rho_u_star(S, 3, &S->path3, rhor, ur, pr, p2, idr, *rhor0, *rhor1, &rhor2, &ur2, NULL, NULL) > 0;
      if (ok) {
        *p0 = *p1 = p2; *rhol0 = *rhol1 = rhol2; *rhor0 = *rhor1 = rhor2; *ul0 = *ul1 = ul2; *ur0 = *ur1 = ur2;
      } else {
        *p0 = *p1 = p_fmin; *rhol0 = *rhol1 = rhol_fmin; *rhor0 = *rhor1 = rhor_fmin;
        *ul0 = *ul1 = ul_fmin; *ur0 = *ur1 = ur_fmin;
      }
    }
    return 0;
  }
  return 1;
}

/* ------------------------------------------------------------ finalisation ------------- */

static void finalize(orc_rs *S, const double *dir, const double *Vm, const double *Vp, double rhol, double ul,
                     double pl, int idl, double rhor, double ur, double pr, int idr, double rhol2, double rhor2,
                     double u2, double p2, int trans_rare, const double *Vrare, double *Vs, int *id, double *Vsm,
                     double *Vsp) {
  double utanl[3] = {Vm[1] - ul * dir[0], Vm[2] - ul * dir[1], Vm[3] - ul * dir[2]};
  double utanr[3] = {Vp[1] - ur * dir[0], Vp[2] - ur * dir[1], Vp[3] - ur * dir[2]};
  *id = (u2 >= 0) ? idl : idr;
  Vs[0] = Vs[1] = Vs[2] = Vs[3] = Vs[4] = 0.0;
  if (trans_rare) {
    Vs[0] = Vrare[0];
    Vs[1] = Vrare[1] * dir[0];
    Vs[2] = Vrare[1] * dir[1];
    Vs[3] = Vrare[1] * dir[2];
    Vs[4] = Vrare[2];
  } else if (u2 >= 0) {
    int star = 0;
    if (pl >= p2) {
      const orc_eos *m = &S->mat[idl];
      double el2 = orc_internal_energy(m, rhol2, p2);
      double cl2 = sqrt(orc_c2(m, rhol2, el2));     /* the reference exits on c^2 < 0 */
      if (u2 - cl2 <= 0) star = 1;
    } else {
      double us = (rhol2 * u2 - rhol * ul) / (rhol2 - rhol);
      if (us <= 0) star = 1;
    }
    if (star) { Vs[0] = rhol2; Vs[1] = u2 * dir[0]; Vs[2] = u2 * dir[1]; Vs[3] = u2 * dir[2]; Vs[4] = p2; }
    else { Vs[0] = rhol; Vs[1] = ul * dir[0]; Vs[2] = ul * dir[1]; Vs[3] = ul * dir[2]; Vs[4] = pl; }
  } else {
    int star = 0;
    if (pr >= p2) {
      const orc_eos *m = &S->mat[idr];
      double er2 = orc_internal_energy(m, rhor2, p2);
      double cr2 = sqrt(orc_c2(m, rhor2, er2));
      if (u2 - cr2 >= 0) star = 1;
    } else {
      double us = (rhor2 * u2 - rhor * ur) / (rhor2 - rhor);
      if (us >= 0) star = 1;
    }
    if (star) { Vs[0] = rhor2; Vs[1] = u2 * dir[0]; Vs[2] = u2 * dir[1]; Vs[3] = u2 * dir[2]; Vs[4] = p2; }
    else { Vs[0] = rhor; Vs[1] = ur * dir[0]; Vs[2] = ur * dir[1]; Vs[3] = ur * dir[2]; Vs[4] = pr; }
  }
  if (u2 > 0) {
    for (int i = 1; i <= 3; i++) Vs[i] += utanl[i - 1];
  } else if (u2 < 0) {
    for (int i = 1; i <= 3; i++) Vs[i] += utanr[i - 1];
  } else {
    for (int i = 1; i <= 3; i++) Vs[i] += 0.5 * (utanl[i - 1] + utanr[i - 1]);
  }
  Vsm[0] = rhol2;
  Vsm[1] = utanl[0] + u2 * dir[0];
  Vsm[2] = utanl[1] + u2 * dir[1];
  Vsm[3] = utanl[2] + u2 * dir[2];
  Vsm[4] = p2;
  Vsp[0] = rhor2;
  Vsp[1] = utanr[0] + u2 * dir[0];
  Vsp[2] = utanr[1] + u2 * dir[1];
  Vsp[3] = utanr[2] + u2 * dir[2];
  Vsp[4] = p2;
}

/* ------------------------------------------------------------ driver -------------------- */

static int solve(orc_rs *S, const double *dir, const double *Vm, int idl, const double *Vp, int idr, double *Vs,
                 int *id, double *Vsm, double *Vsp) {
  const orc_riemann_cfg *C = &S->cfg;
  double rhol = Vm[0];
  double ul = Vm[1] * dir[0] + Vm[2] * dir[1] + Vm[3] * dir[2];
  double pl = Vm[4];
  double rhor = Vp[0];
  double ur = Vp[1] * dir[0] + Vp[2] * dir[1] + Vp[3] * dir[2];
  double pr = Vp[4];
  path_reset(&S->path1, pl, rhol, ul);
  path_reset(&S->path3, pr, rhor, ur);
  const orc_eos *ml = &S->mat[idl], *mr = &S->mat[idr];
  double el = orc_internal_energy(ml, rhol, pl);
  double cl = orc_c2(ml, rhol, el);
  if (rhol <= 0 || cl < 0) return -1;       /* the reference exits */
  cl = sqrt(cl);
  double er = orc_internal_energy(mr, rhor, pr);
  double cr = orc_c2(mr, rhor, er);
  if (rhor <= 0 || cr < 0) return -1;
  cr = sqrt(cr);
  double p0 = DBL_MIN, ul0 = 0.0, ur0 = 0.0, rhol0 = DBL_MIN, rhor0 = DBL_MIN;
  double p1 = DBL_MIN, ul1 = 0.0, ur1 = 0.0, rhol1 = DBL_MIN, rhor1 = DBL_MIN;
  double p2 = DBL_MIN, ul2 = 0.0, ur2 = 0.0, rhol2 = DBL_MIN, rhor2 = DBL_MIN;
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  int trans_rare = 0;
  double Vrare[3] = {0, 0, 0};
  if (ul == ur && pl == pr) {
    finalize(S, dir, Vm, Vp, rhol, ul, pl, idl, rhor, ur, pr, idr, rhol, rhor, ul, pl, trans_rare, Vrare, Vs, id,
             Vsm, Vsp);
    return 0;
  }
  int ok = initial_interval(S, rhol, ul, pl, cl, idl, rhor, ur, pr, cr, idr, &p0, &rhol0, &rhor0, &ul0, &ur0, &p1,
                            &rhol1, &rhor1, &ul1, &ur1);
  if (!ok) {
    ok = rho_u_star(S, 1, &S->path1, rhol, ul, pl, p1, idl, rhol0, rhol0 * 1.1, &rhol2, &ul2, &trans_rare, Vrare) > 0;
    ok = ok && rho_u_star(S, 3, &S->path3, rhor, ur, pr, p1, idr, rhor0, rhol0 * 1.1, &rhor2, &ur2, &trans_rare,
                          Vrare) > 0;
    if (!ok) {
      for (int i = 0; i < 5; i++) { Vsm[i] = Vm[i]; Vsp[i] = Vp[i]; }
      double u_avg = 0.5 * (ul + ur);
      if (u_avg >= 0) { *id = idl; for (int i = 0; i < 5; i++) Vs[i] = Vm[i]; }
      else { *id = idr; for (int i = 0; i < 5; i++) Vs[i] = Vp[i]; }
      return 1;   /* the adaptive driver then runs the fixed-step solver (handled by the caller) */
    }
    finalize(S, dir, Vm, Vp, rhol, ul, pl, idl, rhor, ur, pr, idr, rhol2, rhor2, 0.5 * (ul2 + ur2), p1, trans_rare,
             Vrare, Vs, id, Vsm, Vsp);
    return 1;
  }
  f0 = ul0 - ur0;
  f1 = ul1 - ur1;
  double denom = 0;
  int iter = 0;
  double err_p = 1.0, err_u = 1.0;
  p2 = p1;
  f2 = f1;
  for (iter = 0; iter < C->maxIts_main; iter++) {
    denom = f1 - f0;
    if (denom == 0) {
      p2 = 0.5 * (p0 + p1);
    } else {
      p2 = p2 - f2 * (p1 - p0) / denom;
      if (p2 <= p0 || p2 >= p1) p2 = 0.5 * (p0 + p1);
    }
    int broke = 0;
    for (;;) {   /* the reference's try_again label */
      ok = rho_u_star(S, 1, &S->path1, rhol, ul, pl, p2, idl, rhol0, rhol1, &rhol2, &ul2, &trans_rare, Vrare) > 0;
      if (!ok) {
        p2 = 0.5 * (p2 + p0);
        iter++;
        if (iter < C->maxIts_main) continue;
        broke = 1;
        break;
      }
      ok = rho_u_star(S, 3, &S->path3, rhor, ur, pr, p2, idr, rhor0, rhor1, &rhor2, &ur2, &trans_rare, Vrare) > 0;
      if (!ok) {
        p2 = 0.5 * (p2 + p1);
        iter++;
        if (iter < C->maxIts_main) continue;
        broke = 1;
        break;
      }
      break;
    }
    if (broke) break;
    f2 = ul2 - ur2;
    if (f0 * f2 < 0.0) { p1 = p2; f1 = f2; rhol1 = rhol2; rhor1 = rhor2; }
    else { p0 = p2; f0 = f2; rhol0 = rhol2; rhor0 = rhor2; }
    err_p = fabs(p1 - p0) / dmax(fabs(pl + 0.5 * rhol * ul * ul), fabs(pr + 0.5 * rhor * ur * ur));
    err_u = fabs(f2) / dmax(cl, cr);
    if ((err_p < C->tol_main && err_u < C->tol_main) || (err_p < C->tol_main * 1e-3) ||
        (err_u < C->tol_main * 1e-3))
      break;
    trans_rare = 0;
  }
  double u2 = 0.5 * (ul2 + ur2);
  finalize(S, dir, Vm, Vp, rhol, ul, pl, idl, rhor, ur, pr, idr, rhol2, rhor2, u2, p2, trans_rare, Vrare, Vs, id,
           Vsm, Vsp);
  if (iter == C->maxIts_main) return 1;
  return 0;
}

/* ExactRiemannSolverBase::ComputeRiemannSolution: adaptive first, fixed-step solver on failure
 * (ExactRiemannSolverBase.cpp:145-189, 318-331). Returns 0 on success, 1 if the fallback also
 * failed (the reference counts it as a Riemann error), -1 on an invalid input state. */
int orc_riemann_solve(const orc_eos *mat, const orc_riemann_cfg *cfg, const double *dir, const double *Vm, int idl,
                      const double *Vp, int idr, double *Vs, int *id, double *Vsm, double *Vsp) {
  orc_rs S;
  memset(&S, 0, sizeof(S));
  S.mat = mat;
  S.cfg = *cfg;
  S.nonadaptive = 0;
  int r = solve(&S, dir, Vm, idl, Vp, idr, Vs, id, Vsm, Vsp);
  if (r == 1) {
    S.nonadaptive = 1;
    r = solve(&S, dir, Vm, idl, Vp, idr, Vs, id, Vsm, Vsp);
  }
  free(S.path1.p); free(S.path1.rho); free(S.path1.u);
  free(S.path3.p); free(S.path3.rho); free(S.path3.u);
  return r;
}

/* ------------------------------------------------------------ one-sided (wall) problem -- */

/* FinalizeOneSidedSolution (ExactRiemannSolverBase.cpp:519-630): the state at the wall (xi = 0) and
 * the fluid-side star state. id = -1 when the wall recedes (u* < 0); Vs is then left as it was.
 * Returns -1 where the reference exits (rho* <= 0 or c*^2 < 0 behind a rarefaction). */
static int os_finalize(orc_rs *S, const double *dir, const double *Vm, double rhol, double ul, double pl, int idl,
                       double rhol2, double u2, double p2, int trans_rare, const double *Vrare, double *Vs, int *id,
                       double *Vsm) {
  double utanl[3] = {Vm[1] - ul * dir[0], Vm[2] - ul * dir[1], Vm[3] - ul * dir[2]};
  *id = (u2 >= 0) ? idl : -1;
  if (*id != -1) {
    Vs[0] = Vs[1] = Vs[2] = Vs[3] = Vs[4] = 0.0;
    if (trans_rare) {
      Vs[0] = Vrare[0];
      Vs[1] = Vrare[1] * dir[0];
      Vs[2] = Vrare[1] * dir[1];
      Vs[3] = Vrare[1] * dir[2];
      Vs[4] = Vrare[2];
    } else {
      int star = 0;
      if (pl >= p2) {
        const orc_eos *m = &S->mat[idl];
        double el2 = orc_internal_energy(m, rhol2, p2);
        double cl2 = orc_c2(m, rhol2, el2);
        if (rhol2 <= 0 || cl2 < 0) return -1;
        if (u2 - sqrt(cl2) <= 0) star = 1;
      } else {
        double us = (rhol2 * u2 - rhol * ul) / (rhol2 - rhol);
        if (us <= 0) star = 1;
      }
      if (star) { Vs[0] = rhol2; Vs[1] = u2 * dir[0]; Vs[2] = u2 * dir[1]; Vs[3] = u2 * dir[2]; Vs[4] = p2; }
      else { Vs[0] = rhol; Vs[1] = ul * dir[0]; Vs[2] = ul * dir[1]; Vs[3] = ul * dir[2]; Vs[4] = pl; }
    }
    for (int i = 1; i <= 3; i++) Vs[i] += utanl[i - 1];
  }
  Vsm[0] = rhol2;
  Vsm[1] = utanl[0] + u2 * dir[0];
  Vsm[2] = utanl[1] + u2 * dir[1];
  Vsm[3] = utanl[2] + u2 * dir[2];
  Vsm[4] = p2;
  return 0;
}

/* ComputeOneSidedRarefaction (ExactRiemannSolverBase.cpp:1586-1728): the 1-rarefaction from
 * (rho, u, p) integrated with the embedded RK step until its velocity reaches the wall's, us. The
 * step is sized by velocity (du_target) and cut back when it overshoots. */
static int os_rarefaction(orc_rs *S, double rho, double u, double p, double c, int id, double us, double *rhos,
                          double *ps, int *trans_rare, double *Vrare) {
  const orc_eos *m = &S->mat[id];
  const orc_riemann_cfg *C = &S->cfg;
  *rhos = rho;
  *ps = p;
  double du_target = (us - u) / C->numSteps_rarefaction;
  double du_max = du_target * 1.25;
  double dp_op1 = fabs(du_target * c * rho);
  double dp_op2 = (fabs(p) + 1.0e-6) / (C->numSteps_rarefaction * 2.5);
  double dp = dmin(dp_op1, dp_op2);
  double rhos_0 = rho, us_0 = u, ps_0 = p, xi_0;
  double rhos_1 = rho, us_1 = u, ps_1 = p, xi_1 = 0.0;
  double du;
  xi_0 = u - c;
  double loc_tol = C->tol_rarefaction * dmax(us - u, dmax(fabs(us), fabs(u)));
  double uErr = 0, rhoErr = 0;
  for (int i = 0; i < C->numSteps_rarefaction * 5; i++) {
    if (!rk_adaptive(m, 1, rhos_0, us_0, ps_0, dp, &rhos_1, &us_1, &ps_1, &xi_1, &uErr, &rhoErr)) {
      dp /= 2.0;
      continue;
    }
    du = us_1 - us_0;
    if (du > du_max) {
      dp = dp / du * du_target;
      continue;
    }
    if (us_1 - us > loc_tol) {
      if (du != 0) dp = dp / du * (us - us_0);
      else dp /= 2.0;
      continue;
    }
    if (xi_0 * xi_1 <= 0) {
      *trans_rare = 1;
      double w0 = fabs(xi_1), w1 = fabs(xi_0);
      double ww = w0 + w1;
      w0 /= ww;
      w1 /= ww;
      Vrare[0] = w0 * rhos_0 + w1 * rhos_1;
      Vrare[1] = w0 * us_0 + w1 * us_1;
      Vrare[2] = w0 * ps_0 + w1 * ps_1;
    }
    if (fabs(us - us_1) <= loc_tol) {
      *rhos = rhos_1;
      *ps = ps_1;
      return orc_check_state_rp(m, *rhos, *ps) ? 0 : 1;
    }
    dp = dmin(dp / du * dmin(du_target, us - us_1), dp * 5.0);
    rhos_0 = rhos_1;
    us_0 = us_1;
    ps_0 = ps_1;
    xi_0 = xi_1;
  }
  return 0;
}

/* FindInitialFeasiblePointsOneSidedByAcousticTheory (ExactRiemannSolverBase.cpp:1119-1168):
 * number of starting points found (0, 1 or 2) */
static int os_feasible_acoustic(orc_rs *S, double rhol, double ul, double pl, double cl, int idl, double ustar,
                                double *p0, double *rhol0, double *ul0, double *p1, double *rhol1, double *ul1) {
  double Cl = rhol * cl;
  *p0 = pl + Cl * (ul - ustar);
  if (rho_u_star(S, 1, &S->path1, rhol, ul, pl, *p0, idl, rhol, (*p0 > pl) ? rhol * 1.1 : rhol * 0.9, rhol0, ul0,
                 NULL, NULL) <= 0) {
    *p0 = 1.5 * pl;
    if (rho_u_star(S, 1, &S->path1, rhol, ul, pl, *p0, idl, rhol, (*p0 > pl) ? rhol * 1.1 : rhol * 0.9, rhol0, ul0,
                   NULL, NULL) <= 0)
      return 0;
  }
  double Clbar = (*ul0 == ul) ? Cl : fabs(*p0 - pl) / fabs(*ul0 - ul);
  *p1 = pl + Clbar * (ul - ustar);
  double tmp = dmax(fabs(*p0), fabs(*p1));
  if (fabs(*p1 - *p0) / tmp < 1.0e-8) *p1 = *p0 + 1.0e-8 * tmp;
  if (rho_u_star(S, 1, &S->path1, rhol, ul, pl, *p1, idl, rhol, *rhol0, rhol1, ul1, NULL, NULL) <= 0) return 1;
  return 2;
}

/* FindInitialFeasiblePointsOneSided (ExactRiemannSolverBase.cpp:2946-3021) */
static int os_feasible(orc_rs *S, double rhol, double ul, double pl, double cl, int idl, double ustar, double *p0,
                       double *rhol0, double *ul0, double *p1, double *rhol1, double *ul1) {
  const orc_riemann_cfg *C = &S->cfg;
  int ok = 1;
  int found = os_feasible_acoustic(S, rhol, ul, pl, cl, idl, ustar, p0, rhol0, ul0, p1, rhol1, ul1);
  if (found == 2) return 1;
  if (found != 1) {
    double dp = ul > ustar ? 0.5 * pl : -0.5 * pl;
    for (int i = 0; i < C->maxIts_bracket; i++) {
      *p0 = pl + 0.01 * (i + 1) * (i + 1) * dp;
      if (*p0 < C->min_pressure) *p0 = C->pressure_at_failure;
      ok = rho_u_star(S, 1, &S->path1, rhol, ul, pl, *p0, idl, rhol, (*p0 > pl) ? rhol * 1.1 : rhol * 0.9, rhol0,
                      ul0, NULL, NULL) > 0;
      if (ok) break;
    }
    if (!ok) return 0;
  }
  double dp = *p0 - pl;
  for (int i = 0; i < C->maxIts_bracket; i++) {
    *p1 = *p0 + 0.01 * (i + 1) * (i + 1) * dp;
    ok = rho_u_star(S, 1, &S->path1, rhol, ul, pl, *p1, idl, rhol, *rhol0, rhol1, ul1, NULL, NULL) > 0;
    if (ok) break;
  }
  if (!ok) return 0;
  if (*p0 > *p1) { SWAPD(*p0, *p1); SWAPD(*rhol0, *rhol1); SWAPD(*ul0, *ul1); }
  return 1;
}

/* FindInitialIntervalOneSided (ExactRiemannSolverBase.cpp:800-958): secant steps outward from
 * the starting points until f = u*(p) - ustar changes sign; on failure p0 = p1 = the best point
 * found (or the prescribed failure pressure) and 0 is returned */
static int os_interval(orc_rs *S, double rhol, double ul, double pl, double cl, int idl, double ustar, double *p0,
                       double *rhol0, double *ul0, double *p1, double *rhol1, double *ul1) {
  const orc_riemann_cfg *C = &S->cfg;
  if (!os_feasible(S, rhol, ul, pl, cl, idl, ustar, p0, rhol0, ul0, p1, rhol1, ul1)) {
    *p0 = *p1 = C->pressure_at_failure;
    return 0;
  }
  double p2 = 0, rhol2 = 0, ul2 = 0, f0, f1;
  int i, ok = 1;
  double fmin, p_fmin, rhol_fmin, ul_fmin;
  if (fabs(*ul0 - ustar) < fabs(*ul1 - ustar)) {
    fmin = fabs(*ul0 - ustar); p_fmin = *p0; rhol_fmin = *rhol0; ul_fmin = *ul0;
  } else {
    fmin = fabs(*ul1 - ustar); p_fmin = *p1; rhol_fmin = *rhol1; ul_fmin = *ul1;
  }
  for (i = 0; i < C->maxIts_bracket; i++) {
    f0 = *ul0 - ustar;
    f1 = *ul1 - ustar;
    if (f0 * f1 <= 0.0) return 1;
    if (fabs(f0 - f1) > 1e-9) {
      p2 = *p1 - f1 * (*p1 - *p0) / (f1 - f0);
      if (p2 < *p0) p2 -= 0.1 * (*p1 - *p0);
      else p2 += 0.1 * (*p1 - *p0);
    } else {
      p2 = 1.1 * *p1;
    }
    if (p2 < pl || i == (int)(C->maxIts_bracket / 2)) p2 = pl;
    ok = rho_u_star(S, 1, &S->path1, rhol, ul, pl, p2, idl, *rhol0, *rhol1, &rhol2, &ul2, NULL, NULL) > 0;
    if (!ok) {
      for (int j = 0; j < C->maxIts_bracket / 2; j++) {
        if (p2 < *p0) p2 = *p0 - 0.5 * (*p0 - p2);
        else p2 = *p1 + 0.5 * (p2 - *p1);
        if (p2 < pl) p2 = pl;
        ok = rho_u_star(S, 1, &S->path1, rhol, ul, pl, p2, idl, *rhol0, *rhol1, &rhol2, &ul2, NULL, NULL) > 0;
        if (ok) break;
      }
    }
    if (!ok) break;
    if (fabs(ul2 - ustar) < fmin) { fmin = fabs(ul2 - ustar); p_fmin = p2; rhol_fmin = rhol2; ul_fmin = ul2; }
    if (p2 < *p0) {
      *p1 = *p0; *rhol1 = *rhol0; *ul1 = *ul0;
      *p0 = p2; *rhol0 = rhol2; *ul0 = ul2;
    } else {
      *p0 = *p1; *rhol0 = *rhol1; *ul0 = *ul1;
      *p1 = p2; *rhol1 = rhol2; *ul1 = ul2;
    }
  }
  if (!ok || i == C->maxIts_bracket) {
    if (fmin < C->failure_threshold * fabs(ul - ustar)) {
      *p0 = *p1 = p_fmin; *rhol0 = *rhol1 = rhol_fmin; *ul0 = *ul1 = ul_fmin;
    } else {
      p2 = C->pressure_at_failure;
      ok = rho_u_star(S, 1, &S->path1, rhol, ul, pl, p2, idl, *rhol0, *rhol1, &rhol2, &ul2, NULL, NULL) > 0;
      if (ok) { *p0 = *p1 = p2; *rhol0 = *rhol1 = rhol2; *ul0 = *ul1 = ul2; }
      else { *p0 = *p1 = p_fmin; *rhol0 = *rhol1 = rhol_fmin; *ul0 = *ul1 = ul_fmin; }
    }
    return 0;
  }
  return 1;
}

/* the reference's answer when the one-sided solver gives up: the input state moving with the wall
 * (ExactRiemannSolverBase.cpp:1990-2008, 2040-2056) */
static void os_give_up(const double *dir, const double *Vm, int idl, double ul, double ustar, double *Vs, int *id,
                       double *Vsm) {
  double utanl[3] = {Vm[1] - ul * dir[0], Vm[2] - ul * dir[1], Vm[3] - ul * dir[2]};
  Vsm[0] = Vm[0];
  Vsm[1] = utanl[0] + ustar * dir[0];
  Vsm[2] = utanl[1] + ustar * dir[1];
  Vsm[3] = utanl[2] + ustar * dir[2];
  Vsm[4] = Vm[4];
  if (ustar >= 0) {
    *id = idl;
    for (int i = 0; i < 5; i++) Vs[i] = Vsm[i];
  } else {
    *id = -1;
  }
}

/* ExactRiemannSolverBase::ComputeOneSidedRiemannSolution (ExactRiemannSolverBase.cpp:1931-2198):
 * the half Riemann problem of the fluid state Vm (material idl) against a wall moving with velocity
 * Ustar, along the unit normal dir pointing from the fluid to the wall. The solver object M2C
 * creates (Main.cpp:142) is the adaptive one, so the 1-wave connection is the adaptive
 * ComputeRhoUStar, and there is no fixed-step retry. 0 = success, 1 = failure (the reference counts
 * a Riemann error), -1 = invalid state (the reference exits). */
static int os_solve(orc_rs *S, const double *dir, const double *Vm, int idl, const double *Ustar, double *Vs,
                    int *id, double *Vsm) {
  const orc_riemann_cfg *C = &S->cfg;
  double rhol = Vm[0];
  double ul = Vm[1] * dir[0] + Vm[2] * dir[1] + Vm[3] * dir[2];
  double pl = Vm[4];
  double ustar = Ustar[0] * dir[0] + Ustar[1] * dir[1] + Ustar[2] * dir[2];
  const orc_eos *ml = &S->mat[idl];
  double el = orc_internal_energy(ml, rhol, pl);
  double cl = orc_c2(ml, rhol, el);
  if (rhol <= 0 || cl < 0) return -1;
  cl = sqrt(cl);
  path_reset(&S->path1, pl, rhol, ul);
  double p0 = DBL_MIN, ul0 = 0.0, rhol0 = DBL_MIN;
  double p1 = DBL_MIN, ul1 = 0.0, rhol1 = DBL_MIN;
  double p2 = DBL_MIN, ul2 = 0.0, rhol2 = DBL_MIN;
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  int trans_rare = 0;
  double Vrare[3] = {0, 0, 0};
  if (fabs(ul - ustar) < 1.0e-20)
    return os_finalize(S, dir, Vm, rhol, ul, pl, idl, rhol, ul, pl, trans_rare, Vrare, Vs, id, Vsm);
  if (ul < ustar) {
    if (!os_rarefaction(S, rhol, ul, pl, cl, idl, ustar, &rhol2, &p2, &trans_rare, Vrare)) {
      os_give_up(dir, Vm, idl, ul, ustar, Vs, id, Vsm);
      return 1;
    }
    return os_finalize(S, dir, Vm, rhol, ul, pl, idl, rhol2, ustar, p2, trans_rare, Vrare, Vs, id, Vsm);
  }
  if (!os_interval(S, rhol, ul, pl, cl, idl, ustar, &p0, &rhol0, &ul0, &p1, &rhol1, &ul1)) {
    if (rho_u_star(S, 1, &S->path1, rhol, ul, pl, p1, idl, rhol0, rhol0 * 1.1, &rhol2, &ul2, &trans_rare, Vrare) <= 0) {
      os_give_up(dir, Vm, idl, ul, ustar, Vs, id, Vsm);
      return 1;
    }
    int r = os_finalize(S, dir, Vm, rhol, ul, pl, idl, rhol2, ustar, p1, trans_rare, Vrare, Vs, id, Vsm);
    return r < 0 ? r : 1;
  }
  f0 = ul0 - ustar;
  f1 = ul1 - ustar;
  double denom = 0;
  int iter = 0;
  double err_p = 1.0, err_u = 1.0;
  p2 = p1;
  f2 = f1;
  for (iter = 0; iter < C->maxIts_main; iter++) {
    denom = f1 - f0;
    if (denom == 0) {
      p2 = 0.5 * (p0 + p1);
    } else {
      p2 = p2 - f2 * (p1 - p0) / denom;
      if (p2 <= p0 || p2 >= p1) p2 = 0.5 * (p0 + p1);
    }
    int broke = 0;
    for (;;) {   /* the reference's try_again label */
      if (rho_u_star(S, 1, &S->path1, rhol, ul, pl, p2, idl, rhol0, rhol1, &rhol2, &ul2, &trans_rare, Vrare) > 0)
        break;
      p2 = 0.5 * (p2 + p0);
      iter++;
      if (iter < C->maxIts_main) continue;
      broke = 1;
      break;
    }
    if (broke) break;
    f2 = ul2 - ustar;
    if (f0 * f2 < 0.0) { p1 = p2; f1 = f2; rhol1 = rhol2; }
    else { p0 = p2; f0 = f2; rhol0 = rhol2; }
    err_p = fabs(p1 - p0) / fabs(pl + 0.5 * rhol * ul * ul);
    err_u = fabs(f2) / cl;
    if ((err_p < C->tol_main && err_u < C->tol_main) || (err_p < C->tol_main * 1e-3) || (err_u < C->tol_main * 1e-3))
      break;
    trans_rare = 0;
  }
  int r = os_finalize(S, dir, Vm, rhol, ul, pl, idl, rhol2, ustar, p2, trans_rare, Vrare, Vs, id, Vsm);
  if (r < 0) return r;
  return iter == C->maxIts_main ? 1 : 0;
}

int orc_riemann_onesided_solve(const orc_eos *mat, const orc_riemann_cfg *cfg, const double *dir, const double *Vm,
                               int idl, const double *Ustar, double *Vs, int *id, double *Vsm) {
  orc_rs S;
  memset(&S, 0, sizeof(S));
  S.mat = mat;
  S.cfg = *cfg;
  S.nonadaptive = 0;
  int r = os_solve(&S, dir, Vm, idl, Ustar, Vs, id, Vsm);
  free(S.path1.p); free(S.path1.rho); free(S.path1.u);
  return r;
}

void orc_riemann_begin_stage(orc_ctx *c) { (void)c; }
void orc_riemann_end_stage(orc_ctx *c) { (void)c; }

/* ------------------------------------------------------------ interface normal ---------- */

/* phi at (i,j,k): periodic axes wrapped first, then the ghosted array (its face ghosts hold the
 * level-set boundary rule) */
static double phi_get(const orc_ctx *c, const double *phi, int i, int j, int k) {
  const int *b = c->cfg.ls_bc;
  if (i < 0 && b[0] == 7) i += c->nx; else if (i >= c->nx && b[1] == 7) i -= c->nx;
  if (j < 0 && b[2] == 7) j += c->ny; else if (j >= c->ny && b[3] == 7) j -= c->ny;
  if (k < 0 && b[4] == 7) k += c->nz; else if (k >= c->nz && b[5] == 7) k -= c->nz;
  return phi[ORC_IDX(c, i, j, k)];
}

#define XCO(c, d, q) ((c)->xc[d][(q) + 1])

/* CentralDifferenceLocal (SpaceOperator.h:218-223) */
static double cdiff_local(double p0, double p1, double p2, double x0, double x1, double x2) {
  double c0 = -(x2 - x1) / ((x1 - x0) * (x2 - x0));
  double c1 = 1.0 / (x1 - x0) - 1.0 / (x2 - x1);
  double c2 = (x1 - x0) / ((x2 - x0) * (x2 - x1));
  return c0 * p0 + c1 * p1 + c2 * p2;
}

/* SpaceOperator::CalculateGradientAtCellInterface (SpaceOperator.cpp:3052-3190), written out for
 * the three face directions as there */
static void grad_at_face(const orc_ctx *c, const double *phi, int d, int i, int j, int k, double *dir) {
  int q[3] = {i, j, k}, m[3] = {i, j, k}, N[3] = {c->nx, c->ny, c->nz};
  m[d] -= 1;
  dir[0] = dir[1] = dir[2] = 0.0;
  dir[d] = (phi_get(c, phi, q[0], q[1], q[2]) - phi_get(c, phi, m[0], m[1], m[2])) / (XCO(c, d, q[d]) - XCO(c, d, m[d]));
  int ta[2];
  if (d == 0) { ta[0] = 1; ta[1] = 2; } else if (d == 1) { ta[0] = 0; ta[1] = 2; } else { ta[0] = 0; ta[1] = 1; }
  double t1[2], t2[2];
  for (int a = 0; a < 2; a++) {
    int t = ta[a], qt = q[t];
    int mp[3] = {m[0], m[1], m[2]}, mm[3] = {m[0], m[1], m[2]}, qp[3] = {q[0], q[1], q[2]}, qm[3] = {q[0], q[1], q[2]};
    mp[t] += 1; mm[t] -= 1; qp[t] += 1; qm[t] -= 1;
    double fm = phi_get(c, phi, m[0], m[1], m[2]), fq = phi_get(c, phi, q[0], q[1], q[2]);
    if (qt < 0) {
      t1[a] = (phi_get(c, phi, mp[0], mp[1], mp[2]) - fm) / (XCO(c, t, qt + 1) - XCO(c, t, qt));
      t2[a] = (phi_get(c, phi, qp[0], qp[1], qp[2]) - fq) / (XCO(c, t, qt + 1) - XCO(c, t, qt));
    } else if (qt >= N[t]) {
      t1[a] = (fm - phi_get(c, phi, mm[0], mm[1], mm[2])) / (XCO(c, t, qt) - XCO(c, t, qt - 1));
      t2[a] = (fq - phi_get(c, phi, qm[0], qm[1], qm[2])) / (XCO(c, t, qt) - XCO(c, t, qt - 1));
    } else {
      t1[a] = cdiff_local(phi_get(c, phi, mm[0], mm[1], mm[2]), fm, phi_get(c, phi, mp[0], mp[1], mp[2]),
                          XCO(c, t, qt - 1), XCO(c, t, qt), XCO(c, t, qt + 1));
      t2[a] = cdiff_local(phi_get(c, phi, qm[0], qm[1], qm[2]), fq, phi_get(c, phi, qp[0], qp[1], qp[2]),
                          XCO(c, t, qt - 1), XCO(c, t, qt), XCO(c, t, qt + 1));
    }
  }
  double x_half = XCO(c, d, q[d]) - 0.5 * ORC_W(c, d, q[d]);
  double dd = XCO(c, d, q[d]) - XCO(c, d, m[d]);
  double c1 = (XCO(c, d, q[d]) - x_half) / dd;
  double c2 = (x_half - XCO(c, d, m[d])) / dd;
  dir[ta[0]] = c1 * t1[0] + c2 * t2[0];
  dir[ta[1]] = c1 * t1[1] + c2 * t2[1];
  double nrm = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
  if (nrm != 0.0) {
    double s = 1.0 / nrm;
    dir[0] *= s; dir[1] *= s; dir[2] *= s;
  }
}

/* SpaceOperator::GetNormalForBimaterialRiemann + CalculateGradPhiAtCellInterface
 * (SpaceOperator.cpp:2966-3048) for one level set; myid: plus cell, nbid: minus cell */
void orc_interface_normal(const orc_ctx *c, const double *phi, int d, int i, int j, int k, int myid, int nbid,
                          double *dir) {
  dir[0] = dir[1] = dir[2] = 0.0;
  int mode = c->cfg.mp_normal;
  if (mode == 1 || !phi) { dir[d] = 1.0; return; }
  (void)myid;
  grad_at_face(c, phi, d, i, j, k, dir);
  if (nbid == 0) { dir[0] *= -1.0; dir[1] *= -1.0; dir[2] *= -1.0; }
  if (sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]) == 0.0) dir[d] = 1.0;
  if (dir[d] < 0.0) dir[d] *= -1.0;
  if (mode == 2) {
    dir[d] += 1.0;
    double s = 1.0 / sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    dir[0] *= s; dir[1] *= s; dir[2] *= s;
  }
}

/* SpaceOperator.cpp:2486-2542: the direction of GetNormalForBimaterialRiemann, the solve, clip,
 * store, fluxes. */
int orc_interface_flux(orc_ctx *c, int dir, const double *vm, int idm, const double *vp, int idp, long lo, long hi,
                       double *lf1, double *lf2) {
  const orc_config *g = &c->cfg;
  double n[3];
  {
    long r = hi;
    int ii = (int)(r % c->gx) - 1;
    r /= c->gx;
    int jj = (int)(r % c->gy) - 1;
    int kk = (int)(r / c->gy) - 1;
    orc_interface_normal(c, c->phi_n, dir, ii, jj, kk, idp, idm, n);
  }
  double Vmid[5], Vsm[5], Vsp[5];
  int midid = 0;
  int err = orc_riemann_solve(g->mat, &g->riemann, n, vm, idm, vp, idp, Vmid, &midid, Vsm, Vsp);
  orc_clip(&g->mat[idm], Vsm);
  orc_clip(&g->mat[idp], Vsp);
  if (c->store_riemann && err == 0) orc_rs_store(c, dir, hi, Vsm, idm, Vsp, idp);
  if (g->mp_flux == ORC_MPFLUX_EXACT) {
    orc_physical_flux(&g->mat[midid], dir, Vmid, lf1);
    memcpy(lf2, lf1, 5 * sizeof(double));
  } else {
    orc_numerical_flux(c, dir, vm, Vsm, idm, lf1);
    orc_numerical_flux(c, dir, Vsp, vp, idp, lf2);
  }
  (void)lo;
  return err != 0;
}

/* FluxFcnGodunov.h:37-60 */
void orc_flux_godunov(const orc_ctx *c, int dir, const double *Vm, const double *Vp, int id, double *F) {
  double n[3] = {0.0, 0.0, 0.0};
  n[dir] = 1.0;
  double Vmid[5] = {0, 0, 0, 0, 0}, Vsm[5], Vsp[5];   /* an invalid state (the reference exits) leaves Vmid */
  int midid = id;
  orc_riemann_solve(c->cfg.mat, &c->cfg.riemann, n, Vm, id, Vp, id, Vmid, &midid, Vsm, Vsp);
  orc_physical_flux(&c->cfg.mat[id], dir, Vmid, F);
}

/* ------------------------------------------------------------ wall-cut faces ------------ */

/* SpaceOperator::GetNormalForOneSidedRiemann (SpaceOperator.cpp:3506-3533): the direction of the
 * one-sided problem on a face normal to axis d, fb = +1 when the wall lies ahead of the fluid cell
 * (x increasing), -1 behind it. mode 0 = the wall's normal with its d component turned to point
 * along fb (the other components are kept), 1 = the mesh normal, 2 = their normalised sum. */
void orc_onesided_normal(int d, int fb, int mode, const double *nwall, double *dir) {
  dir[0] = dir[1] = dir[2] = 0.0;
  if (mode == 1) {
    dir[d] = (double)fb;
    return;
  }
  dir[0] = nwall[0]; dir[1] = nwall[1]; dir[2] = nwall[2];
  if (dir[d] * (double)fb < 0.0) dir[d] *= -1;
  if (mode == 2) {
    dir[d] += (double)fb;
    double s = 1.0 / sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    dir[0] *= s; dir[1] *= s; dir[2] *= s;
  }
}

/* The flux on the fluid side of a mesh face cut by an embedded wall (SpaceOperator.cpp:2438-2484 for
 * i-1/2 faces, :2569-2611 for j-1/2, :2696-2737 for k-1/2): the one-sided problem along the
 * GetNormalForOneSidedRiemann direction, the star state clipped (ClipDensityAndPressure), then the
 * numerical flux along the axis between the fluid state and the star state -- F(V, V*) when the
 * fluid cell is on the minus side (side = +1), F(V*, V) when it is on the plus side (side = -1).
 * Vstar receives the clipped star state. Returns the solver's status. */
int orc_wall_face_flux(const orc_ctx *c, int axis, int side, int mode, const double *nwall, const double *vwall,
                       const double *V, int id, double *F, double *Vstar) {
  const orc_config *g = &c->cfg;
  double dir[3], Vs[5] = {0, 0, 0, 0, 0};
  int sid = 0;
  orc_onesided_normal(axis, side, mode, nwall, dir);
  int err = orc_riemann_onesided_solve(g->mat, &g->riemann, dir, V, id, vwall, Vs, &sid, Vstar);
  if (err < 0) return err;
  orc_clip(&g->mat[id], Vstar);
  if (side > 0) orc_numerical_flux(c, axis, V, Vstar, id, F);
  else orc_numerical_flux(c, axis, Vstar, V, id, F);
  return err;
}
