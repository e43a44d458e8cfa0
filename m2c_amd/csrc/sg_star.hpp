// sg_star.hpp -- closed-form star state of the stiffened-gas Riemann problem and the input/output
// maps of the learned interface solver (gfx950 device code).
//
// For a stiffened gas p = (gam-1) rho e - gam pinf, the shock (Hugoniot) and rarefaction (isentrope)
// branches of the wave curves have closed forms in the effective pressure P = p + pinf; they are
// the ideal-gas relations of Toro, "Riemann Solvers and Numerical Methods for Fluid Dynamics",
// 3rd ed., eqs. (4.6)-(4.7), (4.50)-(4.60), with p replaced by P on each side. The star pressure p*
// is the root of F(p) = f_L(p) + f_R(p) + (u_R - u_L) (Toro eq. 4.5); the velocity continuity
// residual |u*_L(p) - u*_R(p)| / max(c_L, c_R) = |F(p)| / max(c_L, c_R) is the acceptance test of a
// predicted p*.
//
// The same pointwise problem is solved iteratively, for any equation of state, by
// ExactRiemannSolverBase::ComputeRiemannSolution (ExactRiemannSolverBase.cpp:59-342), which is the
// fallback whenever a prediction fails the test.
#pragma once
#include <hip/hip_runtime.h>

namespace m2c {

// learned-solver layout: 10 inputs, 4 hidden layers of 256, 3 outputs
constexpr int MLP_IN = 10, MLP_H = 256, MLP_OUT = 3, MLP_NLAYER = 5;
constexpr int MLP_NPARAM = MLP_IN * MLP_H + MLP_H + 3 * (MLP_H * MLP_H + MLP_H) + MLP_H * MLP_OUT + MLP_OUT;

struct SgSide {
  double rho, u, p, gam, pinf;
  double P, c;   // effective pressure p + pinf and sound speed sqrt(gam P / rho)
};

__device__ __forceinline__ bool sg_side_init(SgSide &s) {
  s.P = s.p + s.pinf;
  if (!(s.rho > 0.0) || !(s.P > 0.0) || !(s.gam > 1.0)) return false;
  s.c = sqrt(s.gam * s.P / s.rho);
  return true;
}

// f_K(p) and df_K/dp of one side (Toro 4.6-4.7 in effective pressure); p + pinf > 0 required
__device__ __forceinline__ void sg_wave(const SgSide &s, double p, double &f, double &df) {
  const double Pp = p + s.pinf;
  if (p > s.p) {   // shock
    const double A = 2.0 / ((s.gam + 1.0) * s.rho), B = (s.gam - 1.0) / (s.gam + 1.0) * s.P;
    const double q = sqrt(A / (Pp + B));
    f = (p - s.p) * q;
    df = q * (1.0 - 0.5 * (p - s.p) / (Pp + B));
  } else {         // rarefaction
    const double r = Pp / s.P;
    const double rw = pow(r, (s.gam - 1.0) / (2.0 * s.gam));
    f = 2.0 * s.c / (s.gam - 1.0) * (rw - 1.0);
    df = rw / (r * s.rho * s.c);
  }
}

// density behind the wave of side s at star pressure p (Toro 4.50 / 4.53 in effective pressure)
__device__ __forceinline__ double sg_star_rho(const SgSide &s, double p) {
  const double r = (p + s.pinf) / s.P;
  if (p > s.p) {
    const double mu = (s.gam - 1.0) / (s.gam + 1.0);
    return s.rho * (r + mu) / (mu * r + 1.0);
  }
  return s.rho * pow(r, 1.0 / s.gam);
}

// The solution at xi = x/t = 0 (Toro, sec. 4.5: sampling of shocks, rarefaction heads/tails and the
// fans, eqs. 4.56 and 4.63 with p -> p + pinf): w = (rho, normal velocity, p).
__device__ inline void sg_sample0(const SgSide &L, const SgSide &R, double ps, double us, double rsl, double rsr,
                                  double *w) {
  if (us >= 0.0) {   // the contact moves right: xi = 0 lies on the left of it
    const double Ps = ps + L.pinf;
    if (ps > L.p) {
      const double SL = L.u - L.c * sqrt((L.gam + 1.0) / (2.0 * L.gam) * Ps / L.P + (L.gam - 1.0) / (2.0 * L.gam));
      if (SL >= 0.0) { w[0] = L.rho; w[1] = L.u; w[2] = L.p; }
      else { w[0] = rsl; w[1] = us; w[2] = ps; }
    } else if (L.u - L.c >= 0.0) {
      w[0] = L.rho; w[1] = L.u; w[2] = L.p;
    } else if (us - L.c * pow(Ps / L.P, (L.gam - 1.0) / (2.0 * L.gam)) <= 0.0) {
      w[0] = rsl; w[1] = us; w[2] = ps;
    } else {   // inside the left fan
      const double u = 2.0 / (L.gam + 1.0) * (L.c + 0.5 * (L.gam - 1.0) * L.u), q = u / L.c;
      w[0] = L.rho * pow(q, 2.0 / (L.gam - 1.0));
      w[1] = u;
      w[2] = L.P * pow(q, 2.0 * L.gam / (L.gam - 1.0)) - L.pinf;
    }
  } else {
    const double Ps = ps + R.pinf;
    if (ps > R.p) {
      const double SR = R.u + R.c * sqrt((R.gam + 1.0) / (2.0 * R.gam) * Ps / R.P + (R.gam - 1.0) / (2.0 * R.gam));
      if (SR <= 0.0) { w[0] = R.rho; w[1] = R.u; w[2] = R.p; }
      else { w[0] = rsr; w[1] = us; w[2] = ps; }
    } else if (R.u + R.c <= 0.0) {
      w[0] = R.rho; w[1] = R.u; w[2] = R.p;
    } else if (us + R.c * pow(Ps / R.P, (R.gam - 1.0) / (2.0 * R.gam)) >= 0.0) {
      w[0] = rsr; w[1] = us; w[2] = ps;
    } else {   // inside the right fan
      const double u = 2.0 / (R.gam + 1.0) * (-R.c + 0.5 * (R.gam - 1.0) * R.u), q = -u / R.c;
      w[0] = R.rho * pow(q, 2.0 / (R.gam - 1.0));
      w[1] = u;
      w[2] = R.P * pow(q, 2.0 * R.gam / (R.gam - 1.0)) - R.pinf;
    }
  }
}

// Newton iterations on F from p0: at most `nmax` updates until the continuity residual
// |F(p)| / max(cL, cR) <= tol, then one further update (quadratic convergence takes p* from the
// tolerance to near machine precision; a stiff side turns a small velocity residual into a large
// pressure error). Returns 1 with the polished p (or the one that met tol, should the polish not
// improve it), else 0. *iters receives the number of updates made.
__device__ inline int sg_newton(const SgSide &L, const SgSide &R, double p0, int nmax, double tol, double *pstar,
                                double *resid, int *iters) {
  const double pfloor = -(L.pinf < R.pinf ? L.pinf : R.pinf);   // p + min(pinf) > 0 on both sides
  const double cmax = L.c > R.c ? L.c : R.c, du = R.u - L.u;
  double p = p0;
  if (!(p > pfloor)) p = pfloor + 1e-6 * (L.P < R.P ? L.P : R.P);
  int it = 0;
  double pok = 0.0, rok = 0.0;
  bool met = false;
  for (;;) {
    double fl, dfl, fr, dfr;
    sg_wave(L, p, fl, dfl);
    sg_wave(R, p, fr, dfr);
    const double F = fl + fr + du;
    const double res = fabs(F) / cmax;
    if (met) {   // after the polishing update
      if (res <= rok) { pok = p; rok = res; }
      *pstar = pok; *resid = rok; *iters = it;
      return 1;
    }
    if (res <= tol) { met = true; pok = p; rok = res; }
    else if (it == nmax || !(res == res)) { *pstar = p; *resid = res; *iters = it; return 0; }
    double pn = p - F / (dfl + dfr);
    if (!(pn > pfloor)) pn = 0.5 * (p + pfloor);   // stay on the admissible side
    if (met && !(pn == pn)) { *pstar = pok; *resid = rok; *iters = it; return 1; }
    p = pn;
    it++;
  }
}

// ---------------------------------------------------------------- learned solver maps

struct SgScales {
  double Ps, pimin, cs, ubar;
};

// the 10 network inputs of one (left, right) pair; false when the states are not admissible
__device__ __forceinline__ bool sg_features(const SgSide &L, const SgSide &R, float *x, SgScales *sc) {
  const double cs = L.c + R.c;
  const double ZL = L.rho * L.c, ZR = R.rho * R.c;
  const double pimin = L.pinf < R.pinf ? L.pinf : R.pinf;
  const double Ps = sqrt(L.P * R.P);
  const double du = R.u - L.u;
  // linearised (acoustic) star pressure from the impedances Z = rho c
  double pac = (ZR * L.p + ZL * R.p - ZL * ZR * du) / (ZL + ZR);
  double qac = (pac + pimin) / Ps;
  if (!(qac > 1e-8)) qac = 1e-8;
  x[0] = (float)(log(R.rho / L.rho) * (1.0 / 7.0));
  x[1] = (float)(log(R.P / L.P) * (1.0 / 9.0));
  x[2] = (float)(du / cs);
  x[3] = (float)(log(R.c / L.c) * (1.0 / 5.0));
  x[4] = (float)log(L.gam);
  x[5] = (float)log(R.gam);
  x[6] = (float)(L.pinf / L.P);
  x[7] = (float)(R.pinf / R.P);
  x[8] = (float)(log(ZR / ZL) * (1.0 / 7.0));
  x[9] = (float)(log(qac) * 0.25);
  sc->Ps = Ps;
  sc->pimin = pimin;
  sc->cs = cs;
  sc->ubar = 0.5 * (L.u + R.u);
  return true;
}

// network outputs <-> star quantities: y0 = ln((p* + min pinf)/Ps)/4, y1 = (u* - ubar)/cs,
// y2 = ln(rho*_L / rho_L)/4
__device__ __forceinline__ double sg_decode_p(const SgScales &sc, float y0) { return exp(4.0 * (double)y0) * sc.Ps - sc.pimin; }

}  // namespace m2c
