// physics.hpp -- pointwise equations of state and numerical fluxes for the gfx950 kernels.
//
// Device restatement of:
//   VarFcnSG.h:64-68, VarFcnNASG.h:55-59, VarFcnMG.h:64-81, VarFcnJWL.h:49-53,135-140,
//   VarFcnDummy.h:28-55, VarFcnBase.h:113-294 (C2P/P2C, c^2, H, CheckState, clipping),
//   FluxFcnBase.h:83-123 (physical fluxes), FluxFcnHLLC.h:48-245, FluxFcnGenRoe.h:54-221,
//   FluxFcnLLF.h:46-184.
// Every expression keeps the reference's operand order; the library is compiled with
// -ffp-contract=off and without fast-math, so double-precision results are IEEE-754 identical
// to the reference's x86-64 build (divisions and square roots are correctly rounded on gfx950).
// Template parameters: T is the arithmetic type (double for the drop-in path, float for the
// mixed-precision flux path); EK fixes the equation of state at compile time (-1 = dispatch on
// the material's kind at run time); DIR is the face normal (0 x, 1 y, 2 z), a compile-time
// constant so that no state vector is ever indexed dynamically (which would place it in scratch).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace m2c {

enum { EOS_SG = 0, EOS_NASG = 1, EOS_MG = 2, EOS_JWL = 3, EOS_DUMMY = 5 };
enum { FLUX_ROE = 0, FLUX_LLF = 1, FLUX_HLLC = 2, FLUX_GODUNOV = 3 };
enum { LIM_NONE = 0, LIM_GENMINMOD = 1, LIM_VANALBADA = 2 };

// Material parameters with the derived constants the reference constructors precompute.
struct DevEos {
  int type;
  double rhomin, pmin, rhomax, pmax;
  double gam, pstiff, gam1, gam_pstiff;                 // SG
  double pc, b, q, gam_pc, invgam1;                     // NASG
  double rho0, c0, Gamma0, s, e0, r0c0c0, G0half, G0r0;  // MG
  double omega, A1, A2, R1r0, R2r0, oR1r0, oR2r0;        // JWL (rho0 shared with MG)
  double d_p0, d_rho0, d_e0;                             // dummy
};

// Temperature laws per material: 0 none (JWL, dummy), 1 T0 + (e - e0)/cv (SG, MG), 2 T0 + (e + p/rho
// - h0)/cp (SG, MG with cp > 0 and cv <= 0), 3 the SG law with a reference density (cv > 0 and
// rho0 > 0), 4 NASG (VarFcnSG.h:70-78, 126-142; VarFcnNASG.h:61, 112; VarFcnMG.h:113-121, 172-179).
struct TLaw {
  int method[8];
  double invcv[8], T0[8], e0[8], rho0[8], invcp[8], h0[8];
};

template <typename T> __device__ __forceinline__ T dsqrt(T x) { return __builtin_sqrt(x); }
template <> __device__ __forceinline__ float dsqrt<float>(float x) { return __builtin_sqrtf(x); }
template <typename T> __device__ __forceinline__ T dexp(T x) { return exp(x); }
template <typename T> __device__ __forceinline__ T dabs(T x) { return __builtin_fabs(x); }   // C fabs
template <> __device__ __forceinline__ float dabs<float>(float x) { return __builtin_fabsf(x); }
template <typename T> __device__ __forceinline__ bool dfinite(T x) { return __builtin_isfinite(x); }
// std::min / std::max semantics (first argument wins ties)
template <typename T> __device__ __forceinline__ T smin(T a, T b) { return (b < a) ? b : a; }
template <typename T> __device__ __forceinline__ T smax(T a, T b) { return (a < b) ? b : a; }

template <int EK> __device__ __forceinline__ int kind_of(const DevEos &m) { return EK >= 0 ? EK : m.type; }

template <typename T> __device__ __forceinline__ T jwl_fun(const DevEos &m, T rho) {
  return T(m.A1) * (T(1) - T(m.oR1r0) * rho) * dexp<T>(-T(m.R1r0) / rho)
       + T(m.A2) * (T(1) - T(m.oR2r0) * rho) * dexp<T>(-T(m.R2r0) / rho);
}

template <typename T, int EK> __device__ __forceinline__ T eos_p(const DevEos &m, T rho, T e) {
  switch (kind_of<EK>(m)) {
  case EOS_SG: return T(m.gam1) * rho * e - T(m.gam_pstiff);
  case EOS_NASG: return T(m.gam1) * (e - T(m.q)) / (T(1) / rho - T(m.b)) - T(m.gam_pc);
  case EOS_MG: {
    T eta = T(1) - T(m.rho0) / rho;
    return T(m.r0c0c0) * eta * (T(1) - T(m.G0half) * eta) / ((T(1) - T(m.s) * eta) * (T(1) - T(m.s) * eta))
         + T(m.G0r0) * (e - T(m.e0));
  }
  case EOS_JWL: return T(m.omega) * rho * e + jwl_fun<T>(m, rho);
  default: return T(m.d_p0);
  }
}

template <typename T, int EK> __device__ __forceinline__ T eos_e(const DevEos &m, T rho, T p) {
  switch (kind_of<EK>(m)) {
  case EOS_SG: return (p + T(m.gam_pstiff)) / (T(m.gam1) * rho);
  case EOS_NASG: return T(m.invgam1) * (p + T(m.gam_pc)) * (T(1) / rho - T(m.b)) + T(m.q);
  case EOS_MG: {
    T eta = T(1) - T(m.rho0) / rho;
    return (p - T(m.r0c0c0) * eta * (T(1) - T(m.G0half) * eta) / ((T(1) - T(m.s) * eta) * (T(1) - T(m.s) * eta)))
           / T(m.G0r0) + T(m.e0);
  }
  case EOS_JWL: return (p - jwl_fun<T>(m, rho)) / (T(m.omega) * rho);
  default: return T(m.d_e0);
  }
}

template <typename T, int EK> __device__ __forceinline__ T eos_dpdrho(const DevEos &m, T rho, T e) {
  switch (kind_of<EK>(m)) {
  case EOS_SG: return T(m.gam1) * e;
  case EOS_NASG: { T V = T(1) / rho; return T(m.gam1) * V * V * (e - T(m.q)) / ((V - T(m.b)) * (V - T(m.b))); }
  case EOS_MG: {
    T eta = T(1) - T(m.rho0) / rho;
    T S = T(1) - T(m.s) * eta;
    return T(m.r0c0c0) * (T(1) + (T(m.s) - T(m.Gamma0)) * eta) / (S * S * S) * T(m.rho0) / (rho * rho);
  }
  case EOS_JWL:
    return T(m.omega) * e
         + T(m.A1) * (-T(m.oR1r0) + T(m.R1r0) / (rho * rho) - T(m.oR1r0) * T(m.R1r0) / rho) * dexp<T>(-T(m.R1r0) / rho)
         + T(m.A2) * (-T(m.oR2r0) + T(m.R2r0) / (rho * rho) - T(m.oR2r0) * T(m.R2r0) / rho) * dexp<T>(-T(m.R2r0) / rho);
  default: return T(0);
  }
}

template <typename T, int EK> __device__ __forceinline__ T eos_G(const DevEos &m, T rho) {
  switch (kind_of<EK>(m)) {
  case EOS_SG: return T(m.gam1);
  case EOS_NASG: return T(m.gam1) / (T(1) - T(m.b) * rho);
  case EOS_MG: return T(m.G0r0) / rho;
  case EOS_JWL: return T(m.omega);
  default: return T(0);
  }
}

// c^2 (VarFcnBase.h:215-219): recomputes p from (rho, e); the dummy returns DBL_MIN.
template <typename T, int EK> __device__ __forceinline__ T eos_c2(const DevEos &m, T rho, T e) {
  if (kind_of<EK>(m) == EOS_DUMMY) return T(2.2250738585072014e-308);
  return eos_dpdrho<T, EK>(m, rho, e) + eos_p<T, EK>(m, rho, e) / rho * eos_G<T, EK>(m, rho);
}

// c^2 of a state whose pressure is known: in double the reference's form above (p recomputed
// from e); in float the state's own p, which avoids the p_inf cancellation of p(e) in single
// precision (stiffened gas: (gamma-1) rho e - gamma p_inf)
template <typename T, int EK> __device__ __forceinline__ T eos_c2_st(const DevEos &m, T rho, T e, T p) {
  if (sizeof(T) == sizeof(double)) return eos_c2<T, EK>(m, rho, e);
  if (kind_of<EK>(m) == EOS_DUMMY) return T(2.2250738585072014e-308);
  return eos_dpdrho<T, EK>(m, rho, e) + p / rho * eos_G<T, EK>(m, rho);
}

// H = e + |u|^2/2 + p/rho (VarFcnBase.h:250-254)
template <typename T, int EK> __device__ __forceinline__ T eos_H(const DevEos &m, const T *V) {
  if (kind_of<EK>(m) == EOS_DUMMY) return T(m.d_e0);
  T e = eos_e<T, EK>(m, V[0], V[4]);
  return e + T(0.5) * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3]) + V[4] / V[0];
}

// CheckState(V): true = invalid (VarFcnSG.h:96-108, VarFcnNASG.h:73-85, VarFcnBase.h:113-142)
template <typename T, int EK> __device__ __forceinline__ bool eos_bad(const DevEos &m, const T *V) {
  const int k = kind_of<EK>(m);
  if (k == EOS_DUMMY) return false;
  if (!dfinite(V[0]) || !dfinite(V[1]) || !dfinite(V[2]) || !dfinite(V[3]) || !dfinite(V[4])) return true;
  T rho = V[0], p = V[4];
  if (k == EOS_SG) return rho <= T(0) || p + T(m.pstiff) <= T(0);
  if (k == EOS_NASG) return rho <= T(0) || p + T(m.pc) <= T(0);
  if (rho <= T(0)) return true;
  T e = eos_e<T, EK>(m, rho, p);
  T c2 = eos_dpdrho<T, EK>(m, rho, e) + p / rho * eos_G<T, EK>(m, rho);
  return c2 <= T(0);
}

// ClipDensityAndPressure(V) without U (VarFcnBase.h:259-294); returns true if clipped.
template <typename T, int EK> __device__ __forceinline__ bool eos_clip(const DevEos &m, T *V) {
  if (kind_of<EK>(m) == EOS_DUMMY) {
    V[0] = T(m.d_rho0); V[1] = V[2] = V[3] = T(0); V[4] = T(m.d_p0);
    return false;
  }
  bool c = false;
  if (V[0] < T(m.rhomin)) { V[0] = T(m.rhomin); c = true; }
  if (V[4] < T(m.pmin)) { V[4] = T(m.pmin); c = true; }
  if (V[0] > T(m.rhomax)) { V[0] = T(m.rhomax); c = true; }
  if (V[4] > T(m.pmax)) { V[4] = T(m.pmax); c = true; }
  return c;
}

// Would ClipDensityAndPressure modify V? (no side effects)
template <typename T, int EK> __device__ __forceinline__ bool eos_would_clip(const DevEos &m, const T *V) {
  if (kind_of<EK>(m) == EOS_DUMMY) return false;
  return V[0] < T(m.rhomin) || V[4] < T(m.pmin) || V[0] > T(m.rhomax) || V[4] > T(m.pmax);
}

// VarFcnBase.h:169-181
template <int EK> __device__ __forceinline__ void eos_c2p(const DevEos &m, const double *U, double *V) {
  if (kind_of<EK>(m) == EOS_DUMMY) { V[0] = m.d_rho0; V[1] = V[2] = V[3] = 0.0; V[4] = m.d_p0; return; }
  V[0] = U[0];
  double invRho = 1.0 / U[0];
  V[1] = U[1] * invRho;
  V[2] = U[2] * invRho;
  V[3] = U[3] * invRho;
  double e = (U[4] - 0.5 * V[0] * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3])) * invRho;
  V[4] = eos_p<double, EK>(m, V[0], e);
}

// VarFcnBase.h:185-197
template <typename T, int EK> __device__ __forceinline__ void eos_p2c(const DevEos &m, const T *V, T *U) {
  if (kind_of<EK>(m) == EOS_DUMMY) {
    U[0] = T(m.d_rho0); U[1] = U[2] = U[3] = T(0); U[4] = T(m.d_rho0) * T(m.d_e0);
    return;
  }
  U[0] = V[0];
  U[1] = V[0] * V[1];
  U[2] = V[0] * V[2];
  U[3] = V[0] * V[3];
  T e = eos_e<T, EK>(m, V[0], V[4]);
  U[4] = V[0] * (e + T(0.5) * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3]));
}

// Physical flux normal to DIR (FluxFcnBase.h:83-123).
template <typename T, int EK, int DIR> __device__ __forceinline__ void phys_flux(const DevEos &m, const T *V, T *F) {
  T H = eos_H<T, EK>(m, V);
  T vn = V[DIR + 1];
  T rhou = V[0] * vn;
  F[0] = rhou;
  F[1] = rhou * V[1];
  F[2] = rhou * V[2];
  F[3] = rhou * V[3];
  F[DIR + 1] = rhou * vn + V[4];
  F[4] = rhou * H;
}

template <typename T> __device__ __forceinline__ T wavg(T w1, T f1, T w2, T f2) {
  return (w1 * f1 + w2 * f2) / (w1 + w2);
}

// Physical flux with the state's e, p/rho and |u|^2 already known (the same values
// FluxFcnBase.h:83-123 recomputes through ComputeTotalEnthalpyPerUnitMass).
template <typename T, int EK, int DIR>
__device__ __forceinline__ void phys_flux_k(const DevEos &m, const T *V, T e, T por, T vel2, T *F) {
  T H = (kind_of<EK>(m) == EOS_DUMMY) ? T(m.d_e0) : e + T(0.5) * vel2 + por;
  T vn = V[DIR + 1];
  T rhou = V[0] * vn;
  F[0] = rhou;
  F[1] = rhou * V[1];
  F[2] = rhou * V[2];
  F[3] = rhou * V[3];
  F[DIR + 1] = rhou * vn + V[4];
  F[4] = rhou * H;
}

// HLLC star-region flux, FluxFcnHLLC.h:159-192 (e, p/rho, |u|^2 of Vs passed in; P2C of
// VarFcnBase.h:185-197 inlined with the same operations)
template <typename T, int EK, int DIR>
__device__ __forceinline__ void hllc_fstar(const DevEos &m, const T *Vs, T e, T por, T vel2, T S, T Sstar, T *F) {
  T velo = Vs[DIR + 1];
  T rhostar = Vs[0] * (S - velo) / (S - Sstar);
  T U[5], Us[5];
  if (kind_of<EK>(m) == EOS_DUMMY) {
    U[0] = T(m.d_rho0); U[1] = U[2] = U[3] = T(0); U[4] = T(m.d_rho0) * T(m.d_e0);
  } else {
    U[0] = Vs[0];
    U[1] = Vs[0] * Vs[1];
    U[2] = Vs[0] * Vs[2];
    U[3] = Vs[0] * Vs[3];
    U[4] = Vs[0] * (e + T(0.5) * vel2);
  }
  Us[0] = T(1);
  Us[1] = Vs[1]; Us[2] = Vs[2]; Us[3] = Vs[3];
  Us[DIR + 1] = Sstar;
  Us[4] = U[4] / Vs[0] + (Sstar - velo) * (Sstar + Vs[4] / (Vs[0] * (S - velo)));
#pragma unroll
  for (int i = 0; i < 5; i++) Us[i] *= rhostar;
  phys_flux_k<T, EK, DIR>(m, Vs, e, por, vel2, F);
#pragma unroll
  for (int i = 0; i < 5; i++) F[i] += S * (Us[i] - U[i]);
}

// HLLC, FluxFcnHLLC.h:58-245 (Roe-averaged wave speeds of Hu, Adams & Iaccarino 2009).
template <typename T, int EK, int DIR>
__device__ __forceinline__ void flux_hllc(const DevEos &m, const T *Vm, const T *Vp, T *F) {
  const T eps = T(1e-10);
  T rho_hat = dsqrt<T>(Vm[0] * Vp[0]);
  T sqrt_rhom = dsqrt<T>(Vm[0]);
  T sqrt_rhop = dsqrt<T>(Vp[0]);
  T drho = Vp[0] - Vm[0];
  T dp = Vp[4] - Vm[4];
  T vm = Vm[DIR + 1], vp = Vp[DIR + 1];
  T diff = (vp - vm) / (sqrt_rhom + sqrt_rhop);
  T porm = Vm[4] / Vm[0], porp = Vp[4] / Vp[0];
  T p_over_rho_hat = wavg<T>(sqrt_rhom, porm, sqrt_rhop, porp) + T(0.5) * diff * diff;
  T em = eos_e<T, EK>(m, Vm[0], Vm[4]);
  T ep = eos_e<T, EK>(m, Vp[0], Vp[4]);
  T de = ep - em;
  T e_hat = wavg<T>(sqrt_rhom, em, sqrt_rhop, ep);
  T w_rho = drho * drho / (rho_hat * rho_hat);
  T w_e = de * de / (e_hat * e_hat);
  T denom = w_rho + w_e + eps;
  T dpdrho_avg = wavg<T>(sqrt_rhom, eos_dpdrho<T, EK>(m, Vm[0], em), sqrt_rhop, eos_dpdrho<T, EK>(m, Vp[0], ep));
  T G_avg = wavg<T>(sqrt_rhom, eos_G<T, EK>(m, Vm[0]), sqrt_rhop, eos_G<T, EK>(m, Vp[0]));
  T t1 = (w_e + eps) * dpdrho_avg;
  T t2 = (dp - G_avg * rho_hat * de) * drho / (rho_hat * rho_hat);
  T dpdrho_hat = (t1 + t2) / denom;
  T g1 = (w_rho + eps) * G_avg;
  T g2 = (dp - dpdrho_avg * drho) * de / (rho_hat * e_hat * e_hat);
  T G_hat = (g1 + g2) / denom;
  T c_hat2 = dpdrho_hat + G_hat * p_over_rho_hat;
  T c_hat = (c_hat2 <= T(0)) ? eps : dsqrt<T>(c_hat2);
  T cm = dsqrt<T>(eos_c2_st<T, EK>(m, Vm[0], em, Vm[4]));
  T cp = dsqrt<T>(eos_c2_st<T, EK>(m, Vp[0], ep, Vp[4]));
  T u_hat = wavg<T>(sqrt_rhom, vm, sqrt_rhop, vp);
  T Sm = smin<T>(vm - cm, u_hat - c_hat);
  T Sp = smax<T>(vp + cp, u_hat + c_hat);
  // the state whose flux (or star flux) is taken: minus side for 0 <= Sm, plus side for
  // 0 >= Sp, else by the sign of S* (a NaN S* takes the plus side, where the reference aborts)
  T Sstar = T(0);
  bool star = false, left;
  if (T(0) <= Sm) left = true;
  else if (T(0) >= Sp) left = false;
  else {
    Sstar = (Vp[4] - Vm[4] + Vm[0] * vm * (Sm - vm) - Vp[0] * vp * (Sp - vp))
          / (Vm[0] * (Sm - vm) - Vp[0] * (Sp - vp));
    star = true;
    left = Sstar >= T(0);
  }
  T Vs[5];
#pragma unroll
  for (int i = 0; i < 5; i++) Vs[i] = left ? Vm[i] : Vp[i];
  const T es = left ? em : ep, pors = left ? porm : porp;
  const T vel2 = Vs[1] * Vs[1] + Vs[2] * Vs[2] + Vs[3] * Vs[3];
  if (star) hllc_fstar<T, EK, DIR>(m, Vs, es, pors, vel2, left ? Sm : Sp, Sstar, F);
  else phys_flux_k<T, EK, DIR>(m, Vs, es, pors, vel2, F);
}

// Local Lax-Friedrichs / Rusanov, FluxFcnLLF.h:48-106 and :112-184 (two materials).
template <typename T, int EKM, int EKP, int DIR>
__device__ __forceinline__ void flux_llf2(const DevEos &mm, const DevEos &mp, const T *Vm, const T *Vp, T *F) {
  T em = eos_e<T, EKM>(mm, Vm[0], Vm[4]);
  T ep = eos_e<T, EKP>(mp, Vp[0], Vp[4]);
  T cm = dsqrt<T>(eos_c2_st<T, EKM>(mm, Vm[0], em, Vm[4]));
  T cp = dsqrt<T>(eos_c2_st<T, EKP>(mp, Vp[0], ep, Vp[4]));
  T a = smax<T>(dabs<T>(Vm[DIR + 1]) + cm, dabs<T>(Vp[DIR + 1]) + cp);
  T fm[5], fp[5], Um[5], Up[5];
  phys_flux<T, EKM, DIR>(mm, Vm, fm);
  phys_flux<T, EKP, DIR>(mp, Vp, fp);
  eos_p2c<T, EKM>(mm, Vm, Um);
  eos_p2c<T, EKP>(mp, Vp, Up);
#pragma unroll
  for (int i = 0; i < 5; i++) F[i] = T(0.5) * ((fm[i] + fp[i]) - a * (Up[i] - Um[i]));
}

// Generalised Roe with Harten's entropy fix, FluxFcnGenRoe.h:64-221.
template <typename T, int EK, int DIR>
__device__ __forceinline__ void flux_roe(const DevEos &m, const T *Vm, const T *Vp, T del, T *F) {
  const T eps = T(1e-10);
  T rho_hat = dsqrt<T>(Vm[0] * Vp[0]);
  T sqrt_rhom = dsqrt<T>(Vm[0]);
  T sqrt_rhop = dsqrt<T>(Vp[0]);
  T vel_hat[3];
  vel_hat[0] = wavg<T>(sqrt_rhom, Vm[1], sqrt_rhop, Vp[1]);
  vel_hat[1] = wavg<T>(sqrt_rhom, Vm[2], sqrt_rhop, Vp[2]);
  vel_hat[2] = wavg<T>(sqrt_rhom, Vm[3], sqrt_rhop, Vp[3]);
  T Hm = eos_H<T, EK>(m, Vm), Hp = eos_H<T, EK>(m, Vp);
  T H_hat = wavg<T>(sqrt_rhom, Hm, sqrt_rhop, Hp);
  T drho = Vp[0] - Vm[0];
  T dvel[3] = {Vp[1] - Vm[1], Vp[2] - Vm[2], Vp[3] - Vm[3]};
  T dp = Vp[4] - Vm[4];
  T diff = dvel[DIR] / (sqrt_rhom + sqrt_rhop);
  T p_over_rho_hat = wavg<T>(sqrt_rhom, Vm[4] / Vm[0], sqrt_rhop, Vp[4] / Vp[0]) + T(0.5) * diff * diff;
  T em = eos_e<T, EK>(m, Vm[0], Vm[4]);
  T ep = eos_e<T, EK>(m, Vp[0], Vp[4]);
  T de = ep - em;
  T e_hat = wavg<T>(sqrt_rhom, em, sqrt_rhop, ep);
  T w_rho = drho * drho / (rho_hat * rho_hat);
  T w_e = de * de / (e_hat * e_hat);
  T denom = w_rho + w_e + eps;
  T dpdrho_avg = wavg<T>(sqrt_rhom, eos_dpdrho<T, EK>(m, Vm[0], em), sqrt_rhop, eos_dpdrho<T, EK>(m, Vp[0], ep));
  T G_avg = wavg<T>(sqrt_rhom, eos_G<T, EK>(m, Vm[0]), sqrt_rhop, eos_G<T, EK>(m, Vp[0]));
  T t1 = (w_e + eps) * dpdrho_avg;
  T t2 = (dp - G_avg * rho_hat * de) * drho / (rho_hat * rho_hat);
  T dpdrho_hat = (t1 + t2) / denom;
  T g1 = (w_rho + eps) * G_avg;
  T g2 = (dp - dpdrho_avg * drho) * de / (rho_hat * e_hat * e_hat);
  T G_hat = (g1 + g2) / denom;
  T c2h = dpdrho_hat + G_hat * p_over_rho_hat;
  T ch;
  if (c2h <= T(0)) { ch = eps; c2h = ch * ch; }
  else ch = dsqrt<T>(c2h);

  T vn = vel_hat[DIR];
  T lam[5], a[5];
  lam[0] = vn - ch; lam[1] = vn; lam[2] = vn; lam[3] = vn; lam[4] = vn + ch;
  T u = vel_hat[0], v = vel_hat[1], w = vel_hat[2];
  T vel2 = u * u + v * v + w * w;
  T h2 = e_hat + T(0.5) * vel2 - dpdrho_hat / G_hat;
  T dn = dvel[DIR];
  a[0] = (dp - rho_hat * ch * dn) / (T(2) * c2h);
  a[4] = (dp + rho_hat * ch * dn) / (T(2) * c2h);
#pragma unroll
  for (int t = 0; t < 3; t++) a[t + 1] = (t == DIR) ? (-dp / c2h + drho) : rho_hat * dvel[t];

  T tol = lam[0];
#pragma unroll
  for (int i = 1; i < 5; i++) if (dabs<T>(lam[i]) > tol) tol = dabs<T>(lam[i]);
  tol *= del;
#pragma unroll
  for (int i = 0; i < 5; i++) if (dabs<T>(lam[i]) < tol) lam[i] = (lam[i] * lam[i] + tol * tol) / (T(2) * tol);

  // eigenvectors (FluxFcnBase.h:147-247): r0, r4 acoustic; the middle vector of index DIR+1 is
  // (1,u,v,w,h2); the two tangential ones are unit vectors in the tangential velocities
  T fm[5], fp[5];
  phys_flux<T, EK, DIR>(m, Vm, fm);
  phys_flux<T, EK, DIR>(m, Vp, fp);
#pragma unroll
  for (int i = 0; i < 5; i++) {
    T f = T(0.5) * (fm[i] + fp[i]);
#pragma unroll
    for (int p = 0; p < 5; p++) {
      T r;
      if (p == 0 || p == 4) {
        T c = (p == 0) ? -ch : ch;
        if (i == 0) r = T(1);
        else if (i == 4) r = H_hat + vn * c;
        else if (i == DIR + 1) r = vn + c;
        else r = vel_hat[i - 1];
      } else if (p == DIR + 1) {
        r = (i == 0) ? T(1) : (i == 4 ? h2 : vel_hat[i - 1]);
      } else {
        r = (i == p) ? T(1) : (i == 4 ? vel_hat[p - 1] : T(0));
      }
      f -= T(0.5) * dabs<T>(lam[p]) * a[p] * r;
    }
    F[i] = f;
  }
}


// VarFcn*::GetTemperature(rho, e) of material k (NaN without a law)
__device__ inline double tlaw_temperature(const TLaw &t, const DevEos &m, int k, double r, double ee) {
  switch (t.method[k]) {
  case 1: return t.T0[k] + t.invcv[k] * (ee - t.e0[k]);
  case 2: return t.T0[k] + t.invcp[k] * (ee + eos_p<double, -1>(m, r, ee) / r - t.h0[k]);
  case 3:
    return t.invcv[k] * (ee + m.pstiff / r) +
           pow(r / t.rho0[k], m.gam1) * (t.T0[k] - t.invcv[k] * (t.e0[k] + m.pstiff / t.rho0[k]));
  case 4: return t.invcv[k] * (ee - m.q - m.pc * (1.0 / r - m.b));
  default: return __builtin_nan("");
  }
}

}  // namespace m2c
