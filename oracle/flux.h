/*
 * oracle/flux.h -- CPU restatement of M2C's numerical flux functions (test infrastructure only).
 *
 * Physical fluxes          FluxFcnBase.h:83-123
 * max eigenvalues          FluxFcnBase.h:287-303
 * HLLC (Hu et al. 2009)    FluxFcnHLLC.h:48-245
 * generalised Roe          FluxFcnGenRoe.h:54-221 (Harten entropy fix :217-221)
 * local Lax-Friedrichs     FluxFcnLLF.h:46-106 (material-interface form :112-184)
 * Operand order follows the reference so that -ffp-contract=off reproduces its bits.
 * Parity pins (tests/test_oracle_flux.py): consistency F(V,V) = physical flux for every scheme,
 * x<->-x mirror symmetry, HLLC exact for an isolated stationary contact, Rusanov closed form.
 */
#ifndef M2C_ORACLE_FLUX_H
#define M2C_ORACLE_FLUX_H

#include "eos.h"

enum { ORC_FLUX_ROE = 0, ORC_FLUX_LLF = 1, ORC_FLUX_HLLC = 2, ORC_FLUX_GODUNOV = 3 };

/* FluxFcnBase.h:83-123, dir = 0 (F), 1 (G), 2 (H) */
static inline void orc_physical_flux(const orc_eos *m, int dir, const double *V, double *F) {
  double H = orc_total_enthalpy(m, V);
  double rhou = V[0] * V[dir + 1];
  F[0] = rhou;
  F[1] = rhou * V[1];
  F[2] = rhou * V[2];
  F[3] = rhou * V[3];
  F[dir + 1] = rhou * V[dir + 1] + V[4];
  F[4] = rhou * H;
}

static inline double orc_avg(double w1, double f1, double w2, double f2) {
  return (w1 * f1 + w2 * f2) / (w1 + w2);
}

/* FluxFcnHLLC.h:58-154 */
static inline void orc_hllc_wave_speeds(const orc_eos *m, int dir, const double *Vm, const double *Vp,
                                        double *Sm, double *Sp) {
  const double eps = 1e-10;
  double rho_hat = sqrt(Vm[0] * Vp[0]);
  double sqrt_rhom = sqrt(Vm[0]);
  double sqrt_rhop = sqrt(Vp[0]);
  double drho = Vp[0] - Vm[0];
  double dp = Vp[4] - Vm[4];
  double diffvelo = Vp[dir + 1] - Vm[dir + 1];
  double diff = diffvelo / (sqrt_rhom + sqrt_rhop);
  double p_over_rho_hat = orc_avg(sqrt_rhom, Vm[4] / Vm[0], sqrt_rhop, Vp[4] / Vp[0]) + 0.5 * diff * diff;
  double em = orc_internal_energy(m, Vm[0], Vm[4]);
  double ep = orc_internal_energy(m, Vp[0], Vp[4]);
  double de = ep - em;
  double e_hat = orc_avg(sqrt_rhom, em, sqrt_rhop, ep);
  double w_rho = drho * drho / (rho_hat * rho_hat);
  double w_e = de * de / (e_hat * e_hat);
  double denominator = w_rho + w_e + eps;
  double dpdrho_roeavg = orc_avg(sqrt_rhom, orc_dpdrho(m, Vm[0], em), sqrt_rhop, orc_dpdrho(m, Vp[0], ep));
  double Gamma_roeavg = orc_avg(sqrt_rhom, orc_big_gamma(m, Vm[0], em), sqrt_rhop, orc_big_gamma(m, Vp[0], ep));
  double t1 = (w_e + eps) * dpdrho_roeavg;
  double t2 = (dp - Gamma_roeavg * rho_hat * de) * drho / (rho_hat * rho_hat);
  double dpdrho_hat = (t1 + t2) / denominator;
  double g1 = (w_rho + eps) * Gamma_roeavg;
  double g2 = (dp - dpdrho_roeavg * drho) * de / (rho_hat * e_hat * e_hat);
  double Gamma_hat = (g1 + g2) / denominator;
  double c_hat_square = dpdrho_hat + Gamma_hat * p_over_rho_hat;
  double c_hat = (c_hat_square <= 0) ? eps : sqrt(c_hat_square);
  double cm = sqrt(orc_c2(m, Vm[0], em)); /* the reference aborts on c^2 < 0; sqrt gives NaN here */
  double cp = sqrt(orc_c2(m, Vp[0], ep));
  double u_hat = orc_avg(sqrt_rhom, Vm[dir + 1], sqrt_rhop, Vp[dir + 1]);
  double a = Vm[dir + 1] - cm, b = u_hat - c_hat;
  *Sm = (b < a) ? b : a;                       /* std::min(a,b) */
  a = Vp[dir + 1] + cp; b = u_hat + c_hat;
  *Sp = (a < b) ? b : a;                       /* std::max(a,b) */
}

/* FluxFcnHLLC.h:159-192 */
static inline void orc_hllc_fstar(const orc_eos *m, int dir, const double *V, double S, double Sstar, double *Fs) {
  double velo = V[dir + 1];
  double rhostar = V[0] * (S - velo) / (S - Sstar);
  double U[5], Us[5];
  orc_p2c(m, V, U);
  Us[0] = 1.0;
  Us[1] = (dir == 0) ? Sstar : V[1];
  Us[2] = (dir == 1) ? Sstar : V[2];
  Us[3] = (dir == 2) ? Sstar : V[3];
  Us[4] = U[4] / V[0] + (Sstar - velo) * (Sstar + V[4] / (V[0] * (S - velo)));
  for (int i = 0; i < 5; i++) Us[i] *= rhostar;
  orc_physical_flux(m, dir, V, Fs);
  for (int i = 0; i < 5; i++) Fs[i] += S * (Us[i] - U[i]);
}

/* FluxFcnHLLC.h:197-245 */
static inline void orc_flux_hllc(const orc_eos *m, int dir, const double *Vm, const double *Vp, double *F) {
  double Sm = 0.0, Sp = 0.0;
  orc_hllc_wave_speeds(m, dir, Vm, Vp, &Sm, &Sp);
  if (0.0 <= Sm) {
    orc_physical_flux(m, dir, Vm, F);
  } else if (0.0 >= Sp) {
    orc_physical_flux(m, dir, Vp, F);
  } else {
    double vm = Vm[dir + 1], vp = Vp[dir + 1];
    double Sstar = (Vp[4] - Vm[4] + Vm[0] * vm * (Sm - vm) - Vp[0] * vp * (Sp - vp))
                 / (Vm[0] * (Sm - vm) - Vp[0] * (Sp - vp));
    if (Sstar >= 0.0) orc_hllc_fstar(m, dir, Vm, Sm, Sstar, F);
    else orc_hllc_fstar(m, dir, Vp, Sp, Sstar, F);   /* NaN Sstar: the reference aborts */
  }
}

/* FluxFcnBase.h:147-247: eigenvectors with the normal direction rotated into r1/r5. */
static inline void orc_roe_eigvec(int dir, double u, double v, double w, double e, double c, double H,
                                  double Gamma, double dpdrho, double r[5][5]) {
  double vel2 = u * u + v * v + w * w;
  double vel[3] = {u, v, w};
  double vn = vel[dir];
  double h2 = e + 0.5 * vel2 - dpdrho / Gamma;
  for (int a = 0; a < 5; a++) for (int b = 0; b < 5; b++) r[a][b] = 0.0;
  r[0][0] = 1; r[0][1] = u; r[0][2] = v; r[0][3] = w; r[0][dir + 1] = vn - c; r[0][4] = H - vn * c;
  r[4][0] = 1; r[4][1] = u; r[4][2] = v; r[4][3] = w; r[4][dir + 1] = vn + c; r[4][4] = H + vn * c;
  /* the three middle (linearly degenerate) vectors */
  if (dir == 0) {
    r[1][0] = 1; r[1][1] = u; r[1][2] = v; r[1][3] = w; r[1][4] = h2;
    r[2][2] = 1; r[2][4] = v;
    r[3][3] = 1; r[3][4] = w;
  } else if (dir == 1) {
    r[1][1] = 1; r[1][4] = u;
    r[2][0] = 1; r[2][1] = u; r[2][2] = v; r[2][3] = w; r[2][4] = h2;
    r[3][3] = 1; r[3][4] = w;
  } else {
    r[1][1] = 1; r[1][4] = u;
    r[2][2] = 1; r[2][4] = v;
    r[3][0] = 1; r[3][1] = u; r[3][2] = v; r[3][3] = w; r[3][4] = h2;
  }
}

/* FluxFcnGenRoe.h:64-221 */
static inline void orc_flux_roe(const orc_eos *m, int dir, const double *Vm, const double *Vp, double del, double *F) {
  const double eps = 1e-10;
  double rho_hat = sqrt(Vm[0] * Vp[0]);
  double sqrt_rhom = sqrt(Vm[0]);
  double sqrt_rhop = sqrt(Vp[0]);
  double u_hat = orc_avg(sqrt_rhom, Vm[1], sqrt_rhop, Vp[1]);
  double v_hat = orc_avg(sqrt_rhom, Vm[2], sqrt_rhop, Vp[2]);
  double w_hat = orc_avg(sqrt_rhom, Vm[3], sqrt_rhop, Vp[3]);
  double Hm = orc_total_enthalpy(m, Vm);
  double Hp = orc_total_enthalpy(m, Vp);
  double H_hat = orc_avg(sqrt_rhom, Hm, sqrt_rhop, Hp);
  double drho = Vp[0] - Vm[0];
  double du = Vp[1] - Vm[1], dv = Vp[2] - Vm[2], dw = Vp[3] - Vm[3];
  double dp = Vp[4] - Vm[4];
  double dvel[3] = {du, dv, dw};
  double diff = dvel[dir] / (sqrt_rhom + sqrt_rhop);
  double p_over_rho_hat = orc_avg(sqrt_rhom, Vm[4] / Vm[0], sqrt_rhop, Vp[4] / Vp[0]) + 0.5 * diff * diff;
  double em = orc_internal_energy(m, Vm[0], Vm[4]);
  double ep = orc_internal_energy(m, Vp[0], Vp[4]);
  double de = ep - em;
  double e_hat = orc_avg(sqrt_rhom, em, sqrt_rhop, ep);
  double w_rho = drho * drho / (rho_hat * rho_hat);
  double w_e = de * de / (e_hat * e_hat);
  double denominator = w_rho + w_e + eps;
  double dpdrho_roeavg = orc_avg(sqrt_rhom, orc_dpdrho(m, Vm[0], em), sqrt_rhop, orc_dpdrho(m, Vp[0], ep));
  double Gamma_roeavg = orc_avg(sqrt_rhom, orc_big_gamma(m, Vm[0], em), sqrt_rhop, orc_big_gamma(m, Vp[0], ep));
  double t1 = (w_e + eps) * dpdrho_roeavg;
  double t2 = (dp - Gamma_roeavg * rho_hat * de) * drho / (rho_hat * rho_hat);
  double dpdrho_hat = (t1 + t2) / denominator;
  double g1 = (w_rho + eps) * Gamma_roeavg;
  double g2 = (dp - dpdrho_roeavg * drho) * de / (rho_hat * e_hat * e_hat);
  double Gamma_hat = (g1 + g2) / denominator;
  double c_hat_square = dpdrho_hat + Gamma_hat * p_over_rho_hat;
  double c_hat;
  if (c_hat_square <= 0) { c_hat = eps; c_hat_square = c_hat * c_hat; }
  else c_hat = sqrt(c_hat_square);

  double vel_hat[3] = {u_hat, v_hat, w_hat};
  double vn = vel_hat[dir];
  double lam[5], a[5], r[5][5];
  lam[0] = vn - c_hat; lam[1] = lam[2] = lam[3] = vn; lam[4] = vn + c_hat;
  orc_roe_eigvec(dir, u_hat, v_hat, w_hat, e_hat, c_hat, H_hat, Gamma_hat, dpdrho_hat, r);
  double dn = dvel[dir];
  a[0] = (dp - rho_hat * c_hat * dn) / (2 * c_hat_square);
  a[4] = (dp + rho_hat * c_hat * dn) / (2 * c_hat_square);
  if (dir == 0) { a[1] = -dp / c_hat_square + drho; a[2] = rho_hat * dv; a[3] = rho_hat * dw; }
  else if (dir == 1) { a[1] = rho_hat * du; a[2] = -dp / c_hat_square + drho; a[3] = rho_hat * dw; }
  else { a[1] = rho_hat * du; a[2] = rho_hat * dv; a[3] = -dp / c_hat_square + drho; }

  double tol = lam[0];
  for (int i = 1; i < 5; i++) if (fabs(lam[i]) > tol) tol = fabs(lam[i]);
  tol *= del;
  for (int i = 0; i < 5; i++) if (fabs(lam[i]) < tol) lam[i] = (lam[i] * lam[i] + tol * tol) / (2 * tol);

  double fm[5], fp[5];
  orc_physical_flux(m, dir, Vm, fm);
  orc_physical_flux(m, dir, Vp, fp);
  for (int i = 0; i < 5; i++) {
    F[i] = 0.5 * (fm[i] + fp[i]);
    for (int p = 0; p < 5; p++) F[i] -= 0.5 * fabs(lam[p]) * a[p] * r[p][i];
  }
}

/* FluxFcnLLF.h:48-106 (idm == idp) and :112-184 (material interface) */
static inline void orc_flux_llf2(const orc_eos *mm, const orc_eos *mp, int dir, const double *Vm,
                                 const double *Vp, double *F) {
  double em = orc_internal_energy(mm, Vm[0], Vm[4]);
  double ep = orc_internal_energy(mp, Vp[0], Vp[4]);
  double cm = sqrt(orc_c2(mm, Vm[0], em));
  double cp = sqrt(orc_c2(mp, Vp[0], ep));
  double a1 = fabs(Vm[dir + 1]) + cm, a2 = fabs(Vp[dir + 1]) + cp;
  double a = (a1 < a2) ? a2 : a1;
  double fm[5], fp[5], Um[5], Up[5];
  orc_physical_flux(mm, dir, Vm, fm);
  orc_physical_flux(mp, dir, Vp, fp);
  orc_p2c(mm, Vm, Um);
  orc_p2c(mp, Vp, Up);
  for (int i = 0; i < 5; i++) F[i] = 0.5 * ((fm[i] + fp[i]) - a * (Up[i] - Um[i]));
}

#endif
