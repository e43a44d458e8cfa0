// riemann_dev.hpp -- exact two-material Riemann solver for the interface-face kernels (gfx950).
//
// Device restatement of ExactRiemannSolverBase::ComputeRiemannSolution:
//   driver (secant on p*)        ExactRiemannSolverBase.cpp:59-342 (adaptive), :2199-2474 (fixed step)
//   star-state finalisation      :347-514
//   bracketing interval          :631-795
//   feasible starting points     :959-1114 (acoustic theory :1064-1114)
//   1-/3-wave connection         :1169-1580 (adaptive), :2476-2848 (fixed step)
//   Hugoniot residual            ExactRiemannSolverBase.h:58-75
//   embedded RK (Cash-Karp)      :1729-1852; classic RK4 :2850-2946
//   Hugoniot root                secant/bisection (the reference's WITHOUT_BOOST branch :1501-1543)
// One thread solves one face. The rarefaction path cache (integrationPath1/3,
// ExactRiemannSolverBase.h:28-29) lives in a per-thread slice of a scratch buffer, laid out
// entry-major with the thread index fastest so that a wavefront's accesses to the same entry are
// contiguous. A path that would outgrow its slice sets `overflow` (reported as an error).
// The adaptive step control's pow(x, 1/4) and pow(x, 1/5) (:1286-1310) use det_pow_quarter /
// det_pow_fifth, fixed sequences of +,-,*,/ and sqrt (atanh-series log, Taylor-series exp).
#pragma once
#include "physics.hpp"

namespace m2c {

struct RiemannCfg {
  int maxIts_main, maxIts_bracket, maxIts_shock, numSteps_rarefaction;
  double tol_main, tol_shock, tol_rarefaction, min_pressure, failure_threshold, pressure_at_failure;
};

// ---------------------------------------------------------------- deterministic powers

__device__ __forceinline__ double det_ldexp(double m, int e) {
  int e1 = e / 2, e2 = e - e1;
  double s1 = __longlong_as_double((long long)((unsigned long long)(1023 + e1) << 52));
  double s2 = __longlong_as_double((long long)((unsigned long long)(1023 + e2) << 52));
  return m * s1 * s2;
}

// log of a positive normal double: x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s)
__device__ inline double det_log(double x) {
  unsigned long long b = (unsigned long long)__double_as_longlong(x);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  b = (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
  double m = __longlong_as_double((long long)b);
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

// exp(y), |y| < 700: y = n ln2 + r, Taylor series of e^r to r^14
__device__ inline double det_exp(double y) {
  double nf = __builtin_floor(y * 1.4426950408889634 + 0.5);
  int n = (int)nf;
  double r = y - nf * 0.6931471805599453;
  double t = 1.0 / 87178291200.0;
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
  return det_ldexp(t, n);
}

__device__ __forceinline__ double det_pow_quarter(double x) { return dsqrt<double>(dsqrt<double>(x)); }

__device__ inline double det_pow_fifth(double x) {
  if (!(x > 2.2250738585072014e-308) || !(x < 1e300)) return pow(x, 0.2);
  return det_exp(0.2 * det_log(x));
}

// ---------------------------------------------------------------- solver state

struct RPath {
  double *base;   // p[i] at base[i*stride], rho[i] at base[(cap+i)*stride], u[i] at base[(2cap+i)*stride]
  int n, cap, stride;
  __device__ __forceinline__ double &P(int i) { return base[(long long)i * stride]; }
  __device__ __forceinline__ double &R(int i) { return base[(long long)(cap + i) * stride]; }
  __device__ __forceinline__ double &U(int i) { return base[(long long)(2 * cap + i) * stride]; }
};

struct RSolver {
  const DevEos *mat;
  RiemannCfg cfg;
  int nonadaptive;
  int overflow;
  RPath path1, path3;
};

__device__ __forceinline__ void path_reset(RPath &P, double p, double rho, double u) {
  P.n = 1;
  P.P(0) = p; P.R(0) = rho; P.U(0) = u;
}

__device__ __forceinline__ void path_push(RSolver &S, RPath &P, double p, double rho, double u) {
  if (P.n >= P.cap) { S.overflow = 1; return; }
  P.P(P.n) = p; P.R(P.n) = rho; P.U(P.n) = u;
  P.n++;
}

__device__ __forceinline__ double rs_e(const DevEos &m, double rho, double p) { return eos_e<double, -1>(m, rho, p); }
__device__ __forceinline__ double rs_c2(const DevEos &m, double rho, double e) { return eos_c2<double, -1>(m, rho, e); }
__device__ __forceinline__ bool dnan(double x) { return __builtin_isnan(x); }

// CheckState(rho, p): true = invalid (VarFcnBase.h:113-142, VarFcnSG.h:96-108)
__device__ __forceinline__ bool eos_bad_rp(const DevEos &m, double rho, double p) {
  if (m.type == EOS_DUMMY) return false;
  if (!dfinite(rho) || !dfinite(p)) return true;
  if (m.type == EOS_SG) return rho <= 0.0 || p + m.pstiff <= 0.0;
  if (m.type == EOS_NASG) return rho <= 0.0 || p + m.pc <= 0.0;
  if (rho <= 0.0) return true;
  double e = rs_e(m, rho, p);
  double c2 = eos_dpdrho<double, -1>(m, rho, e) + p / rho * eos_G<double, -1>(m, rho);
  return c2 <= 0;
}

// ---------------------------------------------------------------- RK steps along the isentrope

__device__ inline int rk_adaptive(const DevEos &m, int wave, double rho_0, double u_0, double p_0, double dp,
                                  double *rho, double *u, double *p, double *xi, double *uErr, double *rhoErr) {
  dp = -dp;
  double e_0 = rs_e(m, rho_0, p_0);
  double c_0_square = rs_c2(m, rho_0, e_0);
  if (rho_0 <= 0 || c_0_square < 0) return 0;
  double c_0 = dsqrt<double>(c_0_square);
  double rho_1 = rho_0 + 0.2 * dp / c_0_square;
  double p_1 = p_0 + 0.2 * dp;
  double e_1 = rs_e(m, rho_1, p_1);
  double c_1_square = rs_c2(m, rho_1, e_1);
  if (rho_1 <= 0 || c_1_square < 0) return 0;
  double rho_2 = rho_0 + 3. / 40. * dp / c_0_square + 9. / 40. * dp / c_1_square;
  double p_2 = p_1 + 3. / 10. * dp;
  double e_2 = rs_e(m, rho_2, p_2);
  double c_2_square = rs_c2(m, rho_2, e_2);
  if (rho_2 <= 0 || c_2_square < 0) return 0;
  double c_2 = dsqrt<double>(c_2_square);
  double rho_3 = rho_0 + 3. / 10. * dp / c_0_square - 9. / 10. * dp / c_1_square + 6. / 5. * dp / c_2_square;
  double p_3 = p_0 + 3. / 5. * dp;
  double e_3 = rs_e(m, rho_3, p_3);
  double c_3_square = rs_c2(m, rho_3, e_3);
  if (rho_3 <= 0 || c_3_square < 0) return 0;
  double c_3 = dsqrt<double>(c_3_square);
  double rho_4 = rho_0 - 11. / 54. * dp / c_0_square + 2.5 * dp / c_1_square - 70. / 27. * dp / c_2_square
               + 35. / 27. * dp / c_3_square;
  double p_4 = p_0 + dp;
  double e_4 = rs_e(m, rho_4, p_4);
  double c_4_square = rs_c2(m, rho_4, e_4);
  if (rho_4 <= 0 || c_4_square < 0) return 0;
  double c_4 = dsqrt<double>(c_4_square);
  double rho_5 = rho_0 + 1631. / 55296. * dp / c_0_square + 175. / 512. * dp / c_1_square
               + 575. / 13824. * dp / c_2_square + 44275. / 110592. * dp / c_3_square
               + 253. / 4096. * dp / c_4_square;
  double p_5 = p_0 + 7. / 8. * dp;
  double e_5 = rs_e(m, rho_5, p_5);
  double c_5_square = rs_c2(m, rho_5, e_5);
  if (rho_5 <= 0 || c_5_square < 0) return 0;
  double c_5 = dsqrt<double>(c_5_square);
  double drho = (37. / 378. / c_0_square + 250. / 621. / c_2_square + 125. / 594. / c_3_square
               + 512. / 1771. / c_5_square) * dp;
  double drho_err = (2825. / 27648. / c_0_square + 18575. / 48384. / c_2_square + 13525. / 55296. / c_3_square
                   + 277. / 14336. / c_4_square + 0.25 / c_5_square) * dp;
  double du = (37. / 378. / c_0 / rho_0 + 250. / 621. / c_2 / rho_2 + 125. / 594. / c_3 / rho_3
             + 512. / 1771. / c_5 / rho_5) * dp;
  double du_err = (2825. / 27648. / c_0 / rho_0 + 18575. / 48384. / c_2 / rho_2 + 13525. / 55296. / c_3 / rho_3
                 + 277. / 14336. / c_4 / rho_4 + 0.25 / c_5 / rho_5) * dp;
  if (dnan(du) || dnan(du_err)) return 0;
  *rhoErr = dabs(drho_err - drho);
  *uErr = dabs(du_err - du);
  *rho = rho_0 + drho;
  *p = p_0 + dp;
  *u = (wave == 1) ? u_0 - du : u_0 + du;
  double e = rs_e(m, *rho, *p);
  double c = rs_c2(m, *rho, e);
  if (*rho <= 0 || c < 0) return 0;
  c = dsqrt<double>(c);
  *xi = (wave == 1) ? *u - c : *u + c;
  return 1;
}

__device__ inline int rk_classic(const DevEos &m, int wave, double rho_0, double u_0, double p_0, double dp,
                                 double *rho, double *u, double *p, double *xi) {
  dp = -dp;
  double e_0 = rs_e(m, rho_0, p_0);
  double c_0_square = rs_c2(m, rho_0, e_0);
  if (rho_0 <= 0 || c_0_square < 0) return 0;
  double c_0 = dsqrt<double>(c_0_square);
  double rho_1 = rho_0 + 0.5 * dp / c_0_square;
  double p_1 = p_0 + 0.5 * dp;
  double e_1 = rs_e(m, rho_1, p_1);
  double c_1_square = rs_c2(m, rho_1, e_1);
  if (rho_1 <= 0 || c_1_square < 0) return 0;
  double c_1 = dsqrt<double>(c_1_square);
  double rho_2 = rho_0 + 0.5 * dp / c_1_square;
  double p_2 = p_1;
  double e_2 = rs_e(m, rho_2, p_2);
  double c_2_square = rs_c2(m, rho_2, e_2);
  if (rho_2 <= 0 || c_2_square < 0) return 0;
  double c_2 = dsqrt<double>(c_2_square);
  double rho_3 = rho_0 + dp / c_2_square;
  double p_3 = p_0 + dp;
  double e_3 = rs_e(m, rho_3, p_3);
  double c_3_square = rs_c2(m, rho_3, e_3);
  if (rho_3 <= 0 || c_3_square < 0) return 0;
  double c_3 = dsqrt<double>(c_3_square);
  double drho = 1.0 / 6.0 * dp * (1. / c_0_square + 2.0 * (1. / c_1_square + 1. / c_2_square) + 1. / c_3_square);
  double du = 1.0 / 6.0 * dp * (1. / c_0 / rho_0 + 2.0 * (1. / c_1 / rho_1 + 1. / c_2 / rho_2) + 1. / c_3 / rho_3);
  if (dnan(du)) return 0;
  *rho = rho_0 + drho;
  *p = p_0 + dp;
  *u = (wave == 1) ? u_0 - du : u_0 + du;
  double e = rs_e(m, *rho, *p);
  double c = rs_c2(m, *rho, e);
  if (*rho <= 0 || c < 0) return 0;
  c = dsqrt<double>(c);
  *xi = (wave == 1) ? *u - c : *u + c;
  return 1;
}

// ---------------------------------------------------------------- wave connection

struct Hugo { const DevEos *m; double e, ps, pavg, one_over_rho; };
__device__ __forceinline__ double hugo(const Hugo &h, double rhos) {
  double es = rs_e(*h.m, rhos, h.ps);
  return es - h.e + h.pavg * (1.0 / rhos - h.one_over_rho);
}

// state behind the 1-wave (wave 1) or 3-wave (wave 3) at pressure ps; returns 1 on success,
// 0 on failure, -1 on a NaN in the step control (the reference exits)
__device__ inline int rho_u_star(RSolver &S, int wave, RPath &P, double rho, double u, double p, double ps, int id,
                                 double rhos0, double rhos1, double *rhos_out, double *us_out, int *trans_rare,
                                 double *Vrare) {
  const DevEos &m = S.mat[id];
  const RiemannCfg &C = S.cfg;
  *rhos_out = rho;
  *us_out = u;
  if (p > ps) {
    double rhos_0 = rho, us_0 = u, ps_0 = p, xi_0;
    double rhos_1 = rho, us_1 = u, ps_1 = p, xi_1 = 0.0;
    double dp, dp_min_adaption = 0.0, dp_target = 0.0;
    double e = rs_e(m, rho, p);
    if (!S.nonadaptive) {
      dp = (p - ps) / C.numSteps_rarefaction;
      dp_min_adaption = dp / 2.5;
    } else {
      dp_target = (p - ps) / C.numSteps_rarefaction;
      double dpdrho = eos_dpdrho<double, -1>(m, rho, e);
      double drho_op1 = (p - ps) / dpdrho / C.numSteps_rarefaction;
      double drho_op2 = rho / (2.5 * C.numSteps_rarefaction);
      double drho_target = smin(drho_op1, drho_op2);
      dp = smin(dp_target, drho_target * dpdrho);
    }
    double pressure_endpoint_tol = C.tol_rarefaction * smax(1.0, dabs(p));
    double c = rs_c2(m, rho, e);
    if (rho <= 0 || c < 0) return 0;
    c = dsqrt<double>(c);
    int index0 = 0;
    if (P.n > 1) {
      for (int j = P.n - 1; j >= 0; j--)
        if (P.P(j) > ps) { index0 = j; break; }
      ps_0 = P.P(index0);
      rhos_0 = P.R(index0);
      us_0 = P.U(index0);
      if (!S.nonadaptive) {
        dp_min_adaption = (ps_0 - ps) / C.numSteps_rarefaction / 2.5;
        if (index0 != P.n - 1) dp = ps_0 - ps;
        else dp = smin(P.P(index0 - 1) - P.P(index0), ps_0 - ps);
      } else {
        dp = smin((p - ps) / C.numSteps_rarefaction, ps_0 - ps);
      }
    }
    double xi = (wave == 1) ? u - c : u + c;
    xi_0 = xi;
    int done = 0;
    double uErr = 0., rhoErr = 0.;
    int moreSteps = S.nonadaptive ? 5 : 3;
    for (int i = 0; i < C.numSteps_rarefaction * moreSteps; i++) {
      if (dabs(ps_1 - ps) <= pressure_endpoint_tol) {
        *rhos_out = rhos_1;
        *us_out = us_1;
        if (eos_bad_rp(m, *rhos_out, ps)) return 0;
        done = 1;
        break;
      }
      int ok;
      if (!S.nonadaptive)
        ok = rk_adaptive(m, wave, rhos_0, us_0, ps_0, dp, &rhos_1, &us_1, &ps_1, &xi_1, &uErr, &rhoErr);
      else
        ok = rk_classic(m, wave, rhos_0, us_0, ps_0, dp, &rhos_1, &us_1, &ps_1, &xi_1);
      if (!ok) {
        dp = dp / 2.0;
        if (S.nonadaptive && dp == 0.0) break;
        continue;
      }
      if (ps_1 - ps < -pressure_endpoint_tol) {
        dp = ps_0 - ps;
        continue;
      }
      if (!S.nonadaptive) {
        double tiny = 1.e-30;
        double errBar = smax(1.0e-8, smin(1.0e-5, 100.0 * C.tol_rarefaction));
        double uErrScaled = uErr / c + tiny;
        double rhoErrScaled = rhoErr / rho + tiny;
        if (dnan(rhoErrScaled) || dnan(uErrScaled)) return -1;
        double dpTemp = dp;
        double safety = 0.9;
        if (rhoErrScaled >= errBar && dp > dp_min_adaption) {
          dpTemp = safety * dp * det_pow_quarter(dabs(errBar / rhoErrScaled));
          dpTemp = smax(smax(dpTemp, 0.2 * dp), dp_min_adaption);
          dp = smin(dpTemp, ps_0 - ps);
          continue;
        }
        if (uErrScaled >= errBar && dp > dp_min_adaption) {
          dpTemp = safety * dp * det_pow_quarter(dabs(errBar / uErrScaled));
          dpTemp = smax(smax(dpTemp, 0.2 * dp), dp_min_adaption);
          dp = smin(dpTemp, ps_0 - ps);
          continue;
        }
        if (ps_1 < P.P(P.n - 1)) path_push(S, P, ps_1, rhos_1, us_1);
        if (trans_rare && Vrare && xi_0 * xi_1 <= 0) {
          *trans_rare = 1;
          double w0 = dabs(xi_1), w1 = dabs(xi_0);
          double ww = w0 + w1;
          w0 /= ww;
          w1 /= ww;
          Vrare[0] = w0 * rhos_0 + w1 * rhos_1;
          Vrare[1] = w0 * us_0 + w1 * us_1;
          Vrare[2] = w0 * ps_0 + w1 * ps_1;
        }
        double dpTemp_rho = safety * dp * det_pow_fifth(dabs(errBar / rhoErrScaled));
        double dpTemp_u = safety * dp * det_pow_fifth(dabs(errBar / uErrScaled));
        dpTemp = smin(dpTemp_rho, dpTemp_u);
        dpTemp = smin(dpTemp, 10 * dp);
        dpTemp = smax(dpTemp, dp_min_adaption);
        dp = smin(dpTemp, ps_1 - ps);
      } else {
        if (ps_1 < P.P(P.n - 1)) path_push(S, P, ps_1, rhos_1, us_1);
        if (trans_rare && Vrare && xi_0 * xi_1 <= 0) {
          *trans_rare = 1;
          double w0 = dabs(xi_1), w1 = dabs(xi_0);
          double ww = w0 + w1;
          w0 /= ww;
          w1 /= ww;
          Vrare[0] = w0 * rhos_0 + w1 * rhos_1;
          Vrare[1] = w0 * us_0 + w1 * us_1;
          Vrare[2] = w0 * ps_0 + w1 * ps_1;
        }
        dp = smin(smin(dp_target, ps_1 - ps), 4.0 * dp);
      }
      rhos_0 = rhos_1;
      us_0 = us_1;
      ps_0 = ps_1;
      xi_0 = xi_1;
    }
    if (!done) return 0;
  } else {
    // shock: bracket, then secant/bisection on the Hugoniot residual (:1380-1580)
    Hugo H;
    H.m = &m;
    H.e = rs_e(m, rho, p);
    H.ps = ps;
    H.pavg = 0.5 * (p + ps);
    H.one_over_rho = 1.0 / rho;
    double f0, f1;
    double drho = smax(dabs(rhos0 - rhos1), 0.001 * rhos0);
    int found0 = 0, found1 = 0;
    if (smin(rhos0, rhos1) >= rho) {
      f0 = hugo(H, rhos0);
      f1 = hugo(H, rhos1);
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
        f0 = hugo(H, rhos0);
      }
      if (rhos0 < rho) {
        rhos0 = rho;
        f0 = hugo(H, rhos0);
        found0 = 1;
      } else {
        f0 = hugo(H, rhos0);
      }
    }
    if (!found0 || !found1) {
      int i = 0;
      double factor = 1.5;
      double tmp, ftmp;
      rhos1 = rhos0;
      f1 = f0;
      while (!found0) {
        if (++i >= C.maxIts_shock) return 0;
        tmp = rhos1;
        ftmp = f1;
        rhos1 = rhos0;
        f1 = f0;
        rhos0 = rhos1 - factor * drho;
        if (rhos0 <= rho) {
          rhos0 = rho;
          found0 = 1;
        }
        f0 = hugo(H, rhos0);
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
          f1 = hugo(H, rhos1);
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
          if (++i >= C.maxIts_shock) return 0;
          rhos0 = rhos1;
          f0 = f1;
          rhos1 = rhos0 + factor * drho;
          f1 = hugo(H, rhos1);
          if (f0 * f1 <= 0) found1 = 1;
          else drho = rhos1 - rhos0;
        }
      }
    }
    double sol_first, sol_second;
    double loc_tol_shock = C.tol_shock * smin(rhos0, rhos1);
    if (f0 == 0.0) sol_first = sol_second = rhos0;
    else if (f1 == 0.0) sol_first = sol_second = rhos1;
    else {
      double rhos2 = rhos1, f2 = f1;
      int it;
      sol_first = sol_second = rhos1;
      for (it = 0; it < C.maxIts_shock; it++) {
        drho = rhos1 - rhos0;
        rhos2 = rhos2 - f2 * (rhos1 - rhos0) / (f1 - f0);
        if (rhos2 >= rhos1 || rhos2 <= rhos0) rhos2 = 0.5 * (rhos0 + rhos1);
        f2 = hugo(H, rhos2);
        if (f2 == 0.0) { sol_first = sol_second = rhos2; break; }
        if (f2 * f0 < 0) { rhos1 = rhos2; f1 = f2; }
        else { rhos0 = rhos2; f0 = f2; }
        if (rhos1 - rhos0 < loc_tol_shock) { sol_first = rhos0; sol_second = rhos1; break; }
      }
      if (it == C.maxIts_shock) return 0;
    }
    double rhos = 0.5 * (sol_first + sol_second);
    double du = -(ps - p) * (1.0 / rhos - 1.0 / rho);
    if (du < 0) return 0;
    if (eos_bad_rp(m, rhos, ps)) return 0;
    *rhos_out = rhos;
    *us_out = (wave == 1) ? u - dsqrt<double>(du) : u + dsqrt<double>(du);
  }
  return 1;
}

// both waves at one trial pressure (short-circuit: the 3-wave is skipped if the 1-wave fails)
__device__ __forceinline__ bool both_waves(RSolver &S, double rhol, double ul, double pl, int idl, double rhor,
                                           double ur, double pr, int idr, double ps, double a0, double a1, double b0,
                                           double b1, double *rl, double *uls, double *rr, double *urs) {
  bool ok = rho_u_star(S, 1, S.path1, rhol, ul, pl, ps, idl, a0, a1, rl, uls, nullptr, nullptr) > 0;
  ok = ok && rho_u_star(S, 3, S.path3, rhor, ur, pr, ps, idr, b0, b1, rr, urs, nullptr, nullptr) > 0;
  return ok;
}

// ---------------------------------------------------------------- starting points

struct Bracket { double p0, rhol0, rhor0, ul0, ur0, p1, rhol1, rhor1, ul1, ur1; };

__device__ inline int feasible_acoustic(RSolver &S, double rhol, double ul, double pl, double cl, int idl,
                                        double rhor, double ur, double pr, double cr, int idr, Bracket &B) {
  int found = 0;
  double Cl = rhol * cl, Cr = rhor * cr;
  B.p0 = (Cr * pl + Cl * pr + Cl * Cr * (ul - ur)) / (Cl + Cr);
  if (rho_u_star(S, 1, S.path1, rhol, ul, pl, B.p0, idl, rhol, (B.p0 > pl) ? rhol * 1.1 : rhol * 0.9, &B.rhol0,
                 &B.ul0, nullptr, nullptr) <= 0)
    return found;
  if (rho_u_star(S, 3, S.path3, rhor, ur, pr, B.p0, idr, rhor, (B.p0 > pr) ? rhor * 1.1 : rhor * 0.9, &B.rhor0,
                 &B.ur0, nullptr, nullptr) <= 0)
    return found;
  found = 1;
  double Clbar = (B.ul0 == ul) ? Cl : dabs(B.p0 - pl) / dabs(B.ul0 - ul);
  double Crbar = (B.ur0 == ur) ? Cr : dabs(B.p0 - pr) / dabs(B.ur0 - ur);
  B.p1 = (Crbar * pl + Clbar * pr + Clbar * Crbar * (ul - ur)) / (Clbar + Crbar);
  double tmp = smax(dabs(B.p0), dabs(B.p1));
  if (dabs(B.p1 - B.p0) / tmp < 1.0e-8) B.p1 = B.p0 + 1.0e-8 * tmp;
  if (rho_u_star(S, 1, S.path1, rhol, ul, pl, B.p1, idl, rhol, B.rhol0, &B.rhol1, &B.ul1, nullptr, nullptr) <= 0)
    return found;
  if (rho_u_star(S, 3, S.path3, rhor, ur, pr, B.p1, idr, rhor, B.rhor0, &B.rhor1, &B.ur1, nullptr, nullptr) <= 0)
    return found;
  return 2;
}

__device__ __forceinline__ void dswap(double &a, double &b) { double t = a; a = b; b = t; }

__device__ inline int feasible(RSolver &S, double rhol, double ul, double pl, double cl, int idl, double rhor,
                               double ur, double pr, double cr, int idr, Bracket &B) {
  const RiemannCfg &C = S.cfg;
  bool ok = true;
  int found = feasible_acoustic(S, rhol, ul, pl, cl, idl, rhor, ur, pr, cr, idr, B);
  if (found == 2) return 1;
  if (found != 1) {
    double dp = (pl != pr) ? dabs(pl - pr) : 0.5 * pl;
    for (int i = 0; i < C.maxIts_bracket; i++) {
      B.p0 = smin(pl, pr) + 0.01 * (i + 1) * (i + 1) * smin(dp, smin(dabs(pl), dabs(pr)));
      if (B.p0 < C.min_pressure) B.p0 = C.pressure_at_failure;
      ok = both_waves(S, rhol, ul, pl, idl, rhor, ur, pr, idr, B.p0, rhol, (B.p0 > pl) ? rhol * 1.1 : rhol * 0.9,
                      rhor, (B.p0 > pr) ? rhor * 1.1 : rhor * 0.9, &B.rhol0, &B.ul0, &B.rhor0, &B.ur0);
      if (ok) break;
    }
    if (!ok) {
      for (int i = 0; i < C.maxIts_bracket; i++) {
        B.p0 = smin(pl, pr) - 0.01 * (i + 1) * (i + 1) * smin(dp, smin(dabs(pl), dabs(pr)));
        if (B.p0 < C.min_pressure || i == (int)(C.maxIts_bracket / 2)) B.p0 = C.pressure_at_failure;
        ok = both_waves(S, rhol, ul, pl, idl, rhor, ur, pr, idr, B.p0, rhol, (B.p0 > pl) ? rhol * 1.1 : rhol * 0.9,
                        rhor, (B.p0 > pr) ? rhor * 1.1 : rhor * 0.9, &B.rhol0, &B.ul0, &B.rhor0, &B.ur0);
        if (ok) break;
      }
    }
    if (!ok) return 0;
  }
  {
    double dp = smin(dabs(B.p0 - pl), dabs(B.p0 - pr));
    for (int i = 0; i < C.maxIts_bracket; i++) {
      B.p1 = B.p0 + 0.01 * (i + 1) * (i + 1) * dp;
      ok = both_waves(S, rhol, ul, pl, idl, rhor, ur, pr, idr, B.p1, rhol, B.rhol0, rhor, B.rhor0, &B.rhol1, &B.ul1,
                      &B.rhor1, &B.ur1);
      if (ok) break;
    }
    if (!ok)
      for (int i = 0; i < C.maxIts_bracket; i++) {
        B.p1 = B.p0 - 0.01 * (i + 1) * (i + 1) * dp;
        if (B.p1 < C.min_pressure) B.p1 = C.pressure_at_failure * 1000.0;
        ok = both_waves(S, rhol, ul, pl, idl, rhor, ur, pr, idr, B.p1, rhol, B.rhol0, rhor, B.rhor0, &B.rhol1,
                        &B.ul1, &B.rhor1, &B.ur1);
        if (ok) break;
      }
    if (!ok) return 0;
  }
  if (B.p0 > B.p1) {
    dswap(B.p0, B.p1); dswap(B.rhol0, B.rhol1); dswap(B.rhor0, B.rhor1); dswap(B.ul0, B.ul1); dswap(B.ur0, B.ur1);
  }
  return 1;
}

__device__ inline int initial_interval(RSolver &S, double rhol, double ul, double pl, double cl, int idl,
                                       double rhor, double ur, double pr, double cr, int idr, Bracket &B) {
  const RiemannCfg &C = S.cfg;
  if (!feasible(S, rhol, ul, pl, cl, idl, rhor, ur, pr, cr, idr, B)) {
    B.p0 = B.p1 = C.pressure_at_failure;
    return 0;
  }
  double p2 = 0, rhol2 = 0, rhor2 = 0, ul2 = 0, ur2 = 0, f0, f1;
  int i;
  bool ok = true;
  double fmin, p_fmin, rhol_fmin, rhor_fmin, ul_fmin, ur_fmin;
  if (dabs(B.ul0 - B.ur0) < dabs(B.ul1 - B.ur1)) {
    fmin = dabs(B.ul0 - B.ur0);
    p_fmin = B.p0; rhol_fmin = B.rhol0; rhor_fmin = B.rhor0; ul_fmin = B.ul0; ur_fmin = B.ur0;
  } else {
    fmin = dabs(B.ul1 - B.ur1);
    p_fmin = B.p1; rhol_fmin = B.rhol1; rhor_fmin = B.rhor1; ul_fmin = B.ul1; ur_fmin = B.ur1;
  }
  for (i = 0; i < C.maxIts_bracket; i++) {
    f0 = B.ul0 - B.ur0;
    f1 = B.ul1 - B.ur1;
    if (f0 * f1 <= 0.0) return 1;
    if (dabs(f0 - f1) > 1e-9) {
      p2 = B.p1 - f1 * (B.p1 - B.p0) / (f1 - f0);
      if (p2 < B.p0) p2 -= 0.1 * (B.p1 - B.p0);
      else p2 += 0.1 * (B.p1 - B.p0);
    } else {
      p2 = 1.1 * B.p1;
    }
    if (p2 < C.min_pressure || i == (int)(C.maxIts_bracket / 2)) p2 = 1.0e-8;
    ok = both_waves(S, rhol, ul, pl, idl, rhor, ur, pr, idr, p2, B.rhol0, B.rhol1, B.rhor0, B.rhor1, &rhol2, &ul2,
                    &rhor2, &ur2);
    if (!ok) {
      for (int j = 0; j < C.maxIts_bracket / 2; j++) {
        if (p2 < B.p0) p2 = B.p0 - 0.5 * (B.p0 - p2);
        else p2 = B.p1 + 0.5 * (p2 - B.p1);
        ok = both_waves(S, rhol, ul, pl, idl, rhor, ur, pr, idr, p2, B.rhol0, B.rhol1, B.rhor0, B.rhor1, &rhol2,
                        &ul2, &rhor2, &ur2);
        if (ok) break;
      }
    }
    if (!ok) break;
    if (dabs(ul2 - ur2) < fmin) {
      fmin = dabs(ul2 - ur2);
      p_fmin = p2; rhol_fmin = rhol2; rhor_fmin = rhor2; ul_fmin = ul2; ur_fmin = ur2;
    }
    if (p2 < B.p0) {
      B.p1 = B.p0; B.rhol1 = B.rhol0; B.rhor1 = B.rhor0; B.ul1 = B.ul0; B.ur1 = B.ur0;
      B.p0 = p2; B.rhol0 = rhol2; B.rhor0 = rhor2; B.ul0 = ul2; B.ur0 = ur2;
    } else {
      B.p0 = B.p1; B.rhol0 = B.rhol1; B.rhor0 = B.rhor1; B.ul0 = B.ul1; B.ur0 = B.ur1;
      B.p1 = p2; B.rhol1 = rhol2; B.rhor1 = rhor2; B.ul1 = ul2; B.ur1 = ur2;
    }
  }
  if (!ok || i == C.maxIts_bracket) {
    if (fmin < C.failure_threshold * dabs(ul - ur)) {
      B.p0 = B.p1 = p_fmin; B.rhol0 = B.rhol1 = rhol_fmin; B.rhor0 = B.rhor1 = rhor_fmin;
      B.ul0 = B.ul1 = ul_fmin; B.ur0 = B.ur1 = ur_fmin;
    } else {
      p2 = C.pressure_at_failure;
      ok = both_waves(S, rhol, ul, pl, idl, rhor, ur, pr, idr, p2, B.rhol0, B.rhol1, B.rhor0, B.rhor1, &rhol2, &ul2,
                      &rhor2, &ur2);
      if (ok) {
        B.p0 = B.p1 = p2; B.rhol0 = B.rhol1 = rhol2; B.rhor0 = B.rhor1 = rhor2; B.ul0 = B.ul1 = ul2;
        B.ur0 = B.ur1 = ur2;
      } else {
        B.p0 = B.p1 = p_fmin; B.rhol0 = B.rhol1 = rhol_fmin; B.rhor0 = B.rhor1 = rhor_fmin;
        B.ul0 = B.ul1 = ul_fmin; B.ur0 = B.ur1 = ur_fmin;
      }
    }
    return 0;
  }
  return 1;
}

// ---------------------------------------------------------------- finalisation

__device__ inline void rs_finalize(RSolver &S, const double *dir, const double *Vm, const double *Vp, double rhol,
                                   double ul, double pl, int idl, double rhor, double ur, double pr, int idr,
                                   double rhol2, double rhor2, double u2, double p2, int trans_rare,
                                   const double *Vrare, double *Vs, int *id, double *Vsm, double *Vsp) {
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
      const DevEos &m = S.mat[idl];
      double el2 = rs_e(m, rhol2, p2);
      double cl2 = dsqrt<double>(rs_c2(m, rhol2, el2));
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
      const DevEos &m = S.mat[idr];
      double er2 = rs_e(m, rhor2, p2);
      double cr2 = dsqrt<double>(rs_c2(m, rhor2, er2));
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

// ---------------------------------------------------------------- driver

__device__ inline int rs_solve(RSolver &S, const double *dir, const double *Vm, int idl, const double *Vp, int idr,
                               double *Vs, int *id, double *Vsm, double *Vsp) {
  const RiemannCfg &C = S.cfg;
  double rhol = Vm[0];
  double ul = Vm[1] * dir[0] + Vm[2] * dir[1] + Vm[3] * dir[2];
  double pl = Vm[4];
  double rhor = Vp[0];
  double ur = Vp[1] * dir[0] + Vp[2] * dir[1] + Vp[3] * dir[2];
  double pr = Vp[4];
  path_reset(S.path1, pl, rhol, ul);
  path_reset(S.path3, pr, rhor, ur);
  const DevEos &ml = S.mat[idl], &mr = S.mat[idr];
  double el = rs_e(ml, rhol, pl);
  double cl = rs_c2(ml, rhol, el);
  if (rhol <= 0 || cl < 0) return -1;
  cl = dsqrt<double>(cl);
  double er = rs_e(mr, rhor, pr);
  double cr = rs_c2(mr, rhor, er);
  if (rhor <= 0 || cr < 0) return -1;
  cr = dsqrt<double>(cr);
  Bracket B;
  const double dmin_ = 2.2250738585072014e-308;
  B.p0 = dmin_; B.ul0 = 0.0; B.ur0 = 0.0; B.rhol0 = dmin_; B.rhor0 = dmin_;
  B.p1 = dmin_; B.ul1 = 0.0; B.ur1 = 0.0; B.rhol1 = dmin_; B.rhor1 = dmin_;
  double p2 = dmin_, ul2 = 0.0, ur2 = 0.0, rhol2 = dmin_, rhor2 = dmin_;
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  int trans_rare = 0;
  double Vrare[3] = {0, 0, 0};
  if (ul == ur && pl == pr) {
    rs_finalize(S, dir, Vm, Vp, rhol, ul, pl, idl, rhor, ur, pr, idr, rhol, rhor, ul, pl, trans_rare, Vrare, Vs, id,
                Vsm, Vsp);
    return 0;
  }
  int ok = initial_interval(S, rhol, ul, pl, cl, idl, rhor, ur, pr, cr, idr, B);
  if (!ok) {
    bool ok2 = rho_u_star(S, 1, S.path1, rhol, ul, pl, B.p1, idl, B.rhol0, B.rhol0 * 1.1, &rhol2, &ul2, &trans_rare,
                          Vrare) > 0;
    ok2 = ok2 && rho_u_star(S, 3, S.path3, rhor, ur, pr, B.p1, idr, B.rhor0, B.rhol0 * 1.1, &rhor2, &ur2,
                            &trans_rare, Vrare) > 0;
    if (!ok2) {
      for (int i = 0; i < 5; i++) { Vsm[i] = Vm[i]; Vsp[i] = Vp[i]; }
      double u_avg = 0.5 * (ul + ur);
      if (u_avg >= 0) { *id = idl; for (int i = 0; i < 5; i++) Vs[i] = Vm[i]; }
      else { *id = idr; for (int i = 0; i < 5; i++) Vs[i] = Vp[i]; }
      return 1;
    }
    rs_finalize(S, dir, Vm, Vp, rhol, ul, pl, idl, rhor, ur, pr, idr, rhol2, rhor2, 0.5 * (ul2 + ur2), B.p1,
                trans_rare, Vrare, Vs, id, Vsm, Vsp);
    return 1;
  }
  double p0 = B.p0, p1 = B.p1, rhol0 = B.rhol0, rhol1 = B.rhol1, rhor0 = B.rhor0, rhor1 = B.rhor1;
  f0 = B.ul0 - B.ur0;
  f1 = B.ul1 - B.ur1;
  double denom = 0;
  int iter = 0;
  double err_p = 1.0, err_u = 1.0;
  p2 = p1;
  f2 = f1;
  for (iter = 0; iter < C.maxIts_main; iter++) {
    denom = f1 - f0;
    if (denom == 0) {
      p2 = 0.5 * (p0 + p1);
    } else {
      p2 = p2 - f2 * (p1 - p0) / denom;
      if (p2 <= p0 || p2 >= p1) p2 = 0.5 * (p0 + p1);
    }
    bool broke = false;
    for (;;) {   // the reference's try_again label
      bool okl = rho_u_star(S, 1, S.path1, rhol, ul, pl, p2, idl, rhol0, rhol1, &rhol2, &ul2, &trans_rare, Vrare) > 0;
      if (!okl) {
        p2 = 0.5 * (p2 + p0);
        iter++;
        if (iter < C.maxIts_main) continue;
        broke = true;
        break;
      }
      bool okr = rho_u_star(S, 3, S.path3, rhor, ur, pr, p2, idr, rhor0, rhor1, &rhor2, &ur2, &trans_rare, Vrare) > 0;
      if (!okr) {
        p2 = 0.5 * (p2 + p1);
        iter++;
        if (iter < C.maxIts_main) continue;
        broke = true;
        break;
      }
      break;
    }
    if (broke) break;
    f2 = ul2 - ur2;
    if (f0 * f2 < 0.0) { p1 = p2; f1 = f2; rhol1 = rhol2; rhor1 = rhor2; }
    else { p0 = p2; f0 = f2; rhol0 = rhol2; rhor0 = rhor2; }
    err_p = dabs(p1 - p0) / smax(dabs(pl + 0.5 * rhol * ul * ul), dabs(pr + 0.5 * rhor * ur * ur));
    err_u = dabs(f2) / smax(cl, cr);
    if ((err_p < C.tol_main && err_u < C.tol_main) || (err_p < C.tol_main * 1e-3) || (err_u < C.tol_main * 1e-3))
      break;
    trans_rare = 0;
  }
  double u2 = 0.5 * (ul2 + ur2);
  rs_finalize(S, dir, Vm, Vp, rhol, ul, pl, idl, rhor, ur, pr, idr, rhol2, rhor2, u2, p2, trans_rare, Vrare, Vs, id,
              Vsm, Vsp);
  if (iter == C.maxIts_main) return 1;
  return 0;
}

// ComputeRiemannSolution: adaptive first, the fixed-step solver on failure
// (ExactRiemannSolverBase.cpp:145-189, 318-331). 0 = success, 1 = failure (counted as a Riemann
// error), -1 = invalid input state (the reference exits).
__device__ inline int riemann_solve(RSolver &S, const double *dir, const double *Vm, int idl, const double *Vp,
                                    int idr, double *Vs, int *id, double *Vsm, double *Vsp) {
  S.nonadaptive = 0;
  int r = rs_solve(S, dir, Vm, idl, Vp, idr, Vs, id, Vsm, Vsp);
  if (r == 1) {
    S.nonadaptive = 1;
    r = rs_solve(S, dir, Vm, idl, Vp, idr, Vs, id, Vsm, Vsp);
  }
  return r;
}


// FluxFcnGodunov::ComputeNumericalFluxAtCellInterface (FluxFcnGodunov.h:37-60): the exact problem
// between Vm and Vp of one material along the axis, then the physical flux of the state at xi = 0 in
// that material. The reference does not look at the solver's status; it is returned for the
// invalid-state report (-1, where the reference exits).
template <int DIR>
__device__ inline int godunov_flux(RSolver &S, const double *Vm, const double *Vp, int id, double *F) {
  double nrm[3] = {0.0, 0.0, 0.0};
  nrm[DIR] = 1.0;
  double Vmid[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, Vsm[5], Vsp[5];
  int midid = id;
  const int r = riemann_solve(S, nrm, Vm, id, Vp, id, Vmid, &midid, Vsm, Vsp);
  phys_flux<double, -1, DIR>(S.mat[id], Vmid, F);
  return r;
}

// ---------------------------------------------------------------- one-sided (wall) problem
//
// ExactRiemannSolverBase::ComputeOneSidedRiemannSolution (ExactRiemannSolverBase.cpp:1931-2198) and
// its helpers: FinalizeOneSidedSolution :519-630, FindInitialIntervalOneSided :800-958,
// FindInitialFeasiblePointsOneSidedByAcousticTheory :1119-1168, ComputeOneSidedRarefaction
// :1586-1728, FindInitialFeasiblePointsOneSided :2946-3012. The fluid is on the left of the wall,
// dir points from the fluid to the wall; only the 1-wave exists. M2C's solver object
// (Main.cpp:142) is the adaptive one and the one-sided solver has no fixed-step retry, so every
// wave connection here runs with nonadaptive = 0.

// the state at the wall and the fluid-side star state; id = -1 (Vs untouched) when u* < 0.
// -1 where the reference exits (rho* <= 0 or c*^2 < 0 behind a rarefaction)
__device__ inline int os_finalize(RSolver &S, const double *dir, const double *Vm, double rhol, double ul, double pl,
                                  int idl, double rhol2, double u2, double p2, int trans_rare, const double *Vrare,
                                  double *Vs, int *id, double *Vsm) {
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
        const DevEos &m = S.mat[idl];
        double el2 = rs_e(m, rhol2, p2);
        double cl2 = rs_c2(m, rhol2, el2);
        if (rhol2 <= 0 || cl2 < 0) return -1;
        if (u2 - dsqrt<double>(cl2) <= 0) star = 1;
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

// the 1-rarefaction integrated with the embedded RK step until its velocity reaches us; steps are
// sized by velocity and cut back on overshoot
__device__ inline int os_rarefaction(RSolver &S, double rho, double u, double p, double c, int id, double us,
                                     double *rhos, double *ps, int *trans_rare, double *Vrare) {
  const DevEos &m = S.mat[id];
  const RiemannCfg &C = S.cfg;
  *rhos = rho;
  *ps = p;
  double du_target = (us - u) / C.numSteps_rarefaction;
  double du_max = du_target * 1.25;
  double dp_op1 = dabs(du_target * c * rho);
  double dp_op2 = (dabs(p) + 1.0e-6) / (C.numSteps_rarefaction * 2.5);
  double dp = smin(dp_op1, dp_op2);
  double rhos_0 = rho, us_0 = u, ps_0 = p, xi_0 = u - c;
  double rhos_1 = rho, us_1 = u, ps_1 = p, xi_1 = 0.0;
  double du;
  double loc_tol = C.tol_rarefaction * smax(us - u, smax(dabs(us), dabs(u)));
  double uErr = 0, rhoErr = 0;
  for (int i = 0; i < C.numSteps_rarefaction * 5; i++) {
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
      double w0 = dabs(xi_1), w1 = dabs(xi_0);
      double ww = w0 + w1;
      w0 /= ww;
      w1 /= ww;
      Vrare[0] = w0 * rhos_0 + w1 * rhos_1;
      Vrare[1] = w0 * us_0 + w1 * us_1;
      Vrare[2] = w0 * ps_0 + w1 * ps_1;
    }
    if (dabs(us - us_1) <= loc_tol) {
      *rhos = rhos_1;
      *ps = ps_1;
      return eos_bad_rp(m, *rhos, *ps) ? 0 : 1;
    }
    dp = smin(dp / du * smin(du_target, us - us_1), dp * 5.0);
    rhos_0 = rhos_1;
    us_0 = us_1;
    ps_0 = ps_1;
    xi_0 = xi_1;
  }
  return 0;
}

struct OsBracket { double p0, rhol0, ul0, p1, rhol1, ul1; };

__device__ inline int os_feasible_acoustic(RSolver &S, double rhol, double ul, double pl, double cl, int idl,
                                           double ustar, OsBracket &B) {
  double Cl = rhol * cl;
  B.p0 = pl + Cl * (ul - ustar);
  if (rho_u_star(S, 1, S.path1, rhol, ul, pl, B.p0, idl, rhol, (B.p0 > pl) ? rhol * 1.1 : rhol * 0.9, &B.rhol0,
                 &B.ul0, nullptr, nullptr) <= 0) {
    B.p0 = 1.5 * pl;
    if (rho_u_star(S, 1, S.path1, rhol, ul, pl, B.p0, idl, rhol, (B.p0 > pl) ? rhol * 1.1 : rhol * 0.9, &B.rhol0,
                   &B.ul0, nullptr, nullptr) <= 0)
      return 0;
  }
  double Clbar = (B.ul0 == ul) ? Cl : dabs(B.p0 - pl) / dabs(B.ul0 - ul);
  B.p1 = pl + Clbar * (ul - ustar);
  double tmp = smax(dabs(B.p0), dabs(B.p1));
  if (dabs(B.p1 - B.p0) / tmp < 1.0e-8) B.p1 = B.p0 + 1.0e-8 * tmp;
  if (rho_u_star(S, 1, S.path1, rhol, ul, pl, B.p1, idl, rhol, B.rhol0, &B.rhol1, &B.ul1, nullptr, nullptr) <= 0)
    return 1;
  return 2;
}

__device__ inline int os_feasible(RSolver &S, double rhol, double ul, double pl, double cl, int idl, double ustar,
                                  OsBracket &B) {
  const RiemannCfg &C = S.cfg;
  bool ok = true;
  int found = os_feasible_acoustic(S, rhol, ul, pl, cl, idl, ustar, B);
  if (found == 2) return 1;
  if (found != 1) {
    double dp = ul > ustar ? 0.5 * pl : -0.5 * pl;
    for (int i = 0; i < C.maxIts_bracket; i++) {
      B.p0 = pl + 0.01 * (i + 1) * (i + 1) * dp;
      if (B.p0 < C.min_pressure) B.p0 = C.pressure_at_failure;
      ok = rho_u_star(S, 1, S.path1, rhol, ul, pl, B.p0, idl, rhol, (B.p0 > pl) ? rhol * 1.1 : rhol * 0.9, &B.rhol0,
                      &B.ul0, nullptr, nullptr) > 0;
      if (ok) break;
    }
    if (!ok) return 0;
  }
  double dp = B.p0 - pl;
  for (int i = 0; i < C.maxIts_bracket; i++) {
    B.p1 = B.p0 + 0.01 * (i + 1) * (i + 1) * dp;
    ok = rho_u_star(S, 1, S.path1, rhol, ul, pl, B.p1, idl, rhol, B.rhol0, &B.rhol1, &B.ul1, nullptr, nullptr) > 0;
    if (ok) break;
  }
  if (!ok) return 0;
  if (B.p0 > B.p1) { dswap(B.p0, B.p1); dswap(B.rhol0, B.rhol1); dswap(B.ul0, B.ul1); }
  return 1;
}

__device__ inline int os_interval(RSolver &S, double rhol, double ul, double pl, double cl, int idl, double ustar,
                                  OsBracket &B) {
  const RiemannCfg &C = S.cfg;
  if (!os_feasible(S, rhol, ul, pl, cl, idl, ustar, B)) {
    B.p0 = B.p1 = C.pressure_at_failure;
    return 0;
  }
  double p2 = 0, rhol2 = 0, ul2 = 0, f0, f1;
  int i;
  bool ok = true;
  double fmin, p_fmin, rhol_fmin, ul_fmin;
  if (dabs(B.ul0 - ustar) < dabs(B.ul1 - ustar)) {
    fmin = dabs(B.ul0 - ustar); p_fmin = B.p0; rhol_fmin = B.rhol0; ul_fmin = B.ul0;
  } else {
    fmin = dabs(B.ul1 - ustar); p_fmin = B.p1; rhol_fmin = B.rhol1; ul_fmin = B.ul1;
  }
  for (i = 0; i < C.maxIts_bracket; i++) {
    f0 = B.ul0 - ustar;
    f1 = B.ul1 - ustar;
    if (f0 * f1 <= 0.0) return 1;
    if (dabs(f0 - f1) > 1e-9) {
      p2 = B.p1 - f1 * (B.p1 - B.p0) / (f1 - f0);
      if (p2 < B.p0) p2 -= 0.1 * (B.p1 - B.p0);
      else p2 += 0.1 * (B.p1 - B.p0);
    } else {
      p2 = 1.1 * B.p1;
    }
    if (p2 < pl || i == (int)(C.maxIts_bracket / 2)) p2 = pl;
    ok = rho_u_star(S, 1, S.path1, rhol, ul, pl, p2, idl, B.rhol0, B.rhol1, &rhol2, &ul2, nullptr, nullptr) > 0;
    if (!ok) {
      for (int j = 0; j < C.maxIts_bracket / 2; j++) {
        if (p2 < B.p0) p2 = B.p0 - 0.5 * (B.p0 - p2);
        else p2 = B.p1 + 0.5 * (p2 - B.p1);
        if (p2 < pl) p2 = pl;
        ok = rho_u_star(S, 1, S.path1, rhol, ul, pl, p2, idl, B.rhol0, B.rhol1, &rhol2, &ul2, nullptr, nullptr) > 0;
        if (ok) break;
      }
    }
    if (!ok) break;
    if (dabs(ul2 - ustar) < fmin) { fmin = dabs(ul2 - ustar); p_fmin = p2; rhol_fmin = rhol2; ul_fmin = ul2; }
    if (p2 < B.p0) {
      B.p1 = B.p0; B.rhol1 = B.rhol0; B.ul1 = B.ul0;
      B.p0 = p2; B.rhol0 = rhol2; B.ul0 = ul2;
    } else {
      B.p0 = B.p1; B.rhol0 = B.rhol1; B.ul0 = B.ul1;
      B.p1 = p2; B.rhol1 = rhol2; B.ul1 = ul2;
    }
  }
  if (!ok || i == C.maxIts_bracket) {
    if (fmin < C.failure_threshold * dabs(ul - ustar)) {
      B.p0 = B.p1 = p_fmin; B.rhol0 = B.rhol1 = rhol_fmin; B.ul0 = B.ul1 = ul_fmin;
    } else {
      p2 = C.pressure_at_failure;
      ok = rho_u_star(S, 1, S.path1, rhol, ul, pl, p2, idl, B.rhol0, B.rhol1, &rhol2, &ul2, nullptr, nullptr) > 0;
      if (ok) { B.p0 = B.p1 = p2; B.rhol0 = B.rhol1 = rhol2; B.ul0 = B.ul1 = ul2; }
      else { B.p0 = B.p1 = p_fmin; B.rhol0 = B.rhol1 = rhol_fmin; B.ul0 = B.ul1 = ul_fmin; }
    }
    return 0;
  }
  return 1;
}

// the answer when the solver gives up: the input state moving with the wall
__device__ inline void os_give_up(const double *dir, const double *Vm, int idl, double ul, double ustar, double *Vs,
                                  int *id, double *Vsm) {
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

// 0 = success, 1 = failure (counted as a Riemann error), -1 = invalid state (the reference exits)
__device__ inline int onesided_solve(RSolver &S, const double *dir, const double *Vm, int idl, const double *Ustar,
                                     double *Vs, int *id, double *Vsm) {
  const RiemannCfg &C = S.cfg;
  S.nonadaptive = 0;
  double rhol = Vm[0];
  double ul = Vm[1] * dir[0] + Vm[2] * dir[1] + Vm[3] * dir[2];
  double pl = Vm[4];
  double ustar = Ustar[0] * dir[0] + Ustar[1] * dir[1] + Ustar[2] * dir[2];
  const DevEos &ml = S.mat[idl];
  double el = rs_e(ml, rhol, pl);
  double cl = rs_c2(ml, rhol, el);
  if (rhol <= 0 || cl < 0) return -1;
  cl = dsqrt<double>(cl);
  path_reset(S.path1, pl, rhol, ul);
  OsBracket B;
  const double dmin_ = 2.2250738585072014e-308;
  B.p0 = dmin_; B.ul0 = 0.0; B.rhol0 = dmin_;
  B.p1 = dmin_; B.ul1 = 0.0; B.rhol1 = dmin_;
  double p2 = dmin_, ul2 = 0.0, rhol2 = dmin_;
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  int trans_rare = 0;
  double Vrare[3] = {0, 0, 0};
  if (dabs(ul - ustar) < 1.0e-20)
    return os_finalize(S, dir, Vm, rhol, ul, pl, idl, rhol, ul, pl, trans_rare, Vrare, Vs, id, Vsm);
  if (ul < ustar) {
    if (!os_rarefaction(S, rhol, ul, pl, cl, idl, ustar, &rhol2, &p2, &trans_rare, Vrare)) {
      os_give_up(dir, Vm, idl, ul, ustar, Vs, id, Vsm);
      return 1;
    }
    return os_finalize(S, dir, Vm, rhol, ul, pl, idl, rhol2, ustar, p2, trans_rare, Vrare, Vs, id, Vsm);
  }
  if (!os_interval(S, rhol, ul, pl, cl, idl, ustar, B)) {
    if (rho_u_star(S, 1, S.path1, rhol, ul, pl, B.p1, idl, B.rhol0, B.rhol0 * 1.1, &rhol2, &ul2, &trans_rare,
                   Vrare) <= 0) {
      os_give_up(dir, Vm, idl, ul, ustar, Vs, id, Vsm);
      return 1;
    }
    int r = os_finalize(S, dir, Vm, rhol, ul, pl, idl, rhol2, ustar, B.p1, trans_rare, Vrare, Vs, id, Vsm);
    return r < 0 ? r : 1;
  }
  double p0 = B.p0, p1 = B.p1, rhol0 = B.rhol0, rhol1 = B.rhol1;
  f0 = B.ul0 - ustar;
  f1 = B.ul1 - ustar;
  double denom = 0;
  int iter = 0;
  double err_p = 1.0, err_u = 1.0;
  p2 = p1;
  f2 = f1;
  for (iter = 0; iter < C.maxIts_main; iter++) {
    denom = f1 - f0;
    if (denom == 0) {
      p2 = 0.5 * (p0 + p1);
    } else {
      p2 = p2 - f2 * (p1 - p0) / denom;
      if (p2 <= p0 || p2 >= p1) p2 = 0.5 * (p0 + p1);
    }
    bool broke = false;
    for (;;) {   // the reference's try_again label
      if (rho_u_star(S, 1, S.path1, rhol, ul, pl, p2, idl, rhol0, rhol1, &rhol2, &ul2, &trans_rare, Vrare) > 0) break;
      p2 = 0.5 * (p2 + p0);
      iter++;
      if (iter < C.maxIts_main) continue;
      broke = true;
      break;
    }
    if (broke) break;
    f2 = ul2 - ustar;
    if (f0 * f2 < 0.0) { p1 = p2; f1 = f2; rhol1 = rhol2; }
    else { p0 = p2; f0 = f2; rhol0 = rhol2; }
    err_p = dabs(p1 - p0) / dabs(pl + 0.5 * rhol * ul * ul);
    err_u = dabs(f2) / cl;
    if ((err_p < C.tol_main && err_u < C.tol_main) || (err_p < C.tol_main * 1e-3) || (err_u < C.tol_main * 1e-3))
      break;
    trans_rare = 0;
  }
  int r = os_finalize(S, dir, Vm, rhol, ul, pl, idl, rhol2, ustar, p2, trans_rare, Vrare, Vs, id, Vsm);
  if (r < 0) return r;
  return iter == C.maxIts_main ? 1 : 0;
}

// GetNormalForOneSidedRiemann (SpaceOperator.cpp:3506-3533): mode 0 the wall's normal with its d
// component turned along fb, 1 the mesh normal, 2 their normalised sum
__device__ inline void onesided_normal(int d, int fb, int mode, const double *nwall, double *dir) {
  dir[0] = dir[1] = dir[2] = 0.0;
  if (mode == 1) {
    dir[d] = (double)fb;
    return;
  }
  dir[0] = nwall[0]; dir[1] = nwall[1]; dir[2] = nwall[2];
  if (dir[d] * (double)fb < 0.0) dir[d] *= -1;
  if (mode == 2) {
    dir[d] += (double)fb;
    double s = 1.0 / dsqrt<double>(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    dir[0] *= s; dir[1] *= s; dir[2] *= s;
  }
}

}  // namespace m2c
