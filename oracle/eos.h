/*
 * oracle/eos.h -- CPU restatement of M2C's equations of state (test infrastructure only).
 *
 * This file is part of the ORACLE: a plain-C restatement of the reference algorithm that the
 * tests, the smoke check and the CPU-baseline leg of bench.py use as the checker. Product code
 * never includes it.
 *
 * Every expression keeps the reference's operand order so that, compiled with
 * -ffp-contract=off, it reproduces the reference's IEEE-754 results bit for bit:
 *   stiffened gas           VarFcnSG.h:64-68, CheckState VarFcnSG.h:96-108
 *   Noble-Abel SG           VarFcnNASG.h:55-59, CheckState VarFcnNASG.h:73-85
 *   Mie-Gruneisen           VarFcnMG.h:64-81
 *   JWL                     VarFcnJWL.h:49-53,64-67,135-140
 *   dummy (inactive)        VarFcnDummy.h:28-55
 *   base-class formulas     VarFcnBase.h:169-197 (C2P/P2C), :215-219 (c^2), :250-254 (H),
 *                           :113-142 (generic CheckState), :259-294 (clipping)
 * Parity pins (tests/test_oracle_eos.py, tests/test_oracle_fv.py): NASG with b=q=0 equals SG, JWL
 * with A1=A2=0 is an ideal gas with gamma-1 = omega, MG at rho=rho0 is linear in e, MG on its
 * principal Hugoniot gives rho0 Us up, JWL against its textbook relative-volume form, the NASG sound
 * speed of Le Metayer & Saurel, dp/drho|e, Gamma and c^2 against central differences of p(rho,e)
 * (c^2 along the isentrope de = p/rho^2 drho), p(rho,e(rho,p)) = p, SG c^2 = gamma (p+p_inf)/rho,
 * CheckState on the edges of the valid region.
 */
#ifndef M2C_ORACLE_EOS_H
#define M2C_ORACLE_EOS_H

#include <math.h>
#include <float.h>

enum { ORC_EOS_SG = 0, ORC_EOS_NASG = 1, ORC_EOS_MG = 2, ORC_EOS_JWL = 3, ORC_EOS_DUMMY = 5 };

typedef struct {
  int type;
  double rhomin, pmin, rhomax, pmax, failsafe_density;
  /* SG */
  double gam, pstiff, gam1;
  /* NASG */
  double pc, b, q, gam_pc;
  /* MG */
  double rho0, c0, Gamma0, s, e0, rho0_c0_c0, Gamma0_over_2, Gamma0_rho0;
  /* JWL */
  double omega, A1, A2, R1, R2, R1rho0, R2rho0, omega_over_R1rho0, omega_over_R2rho0;
  /* dummy */
  double d_p0, d_rho0, d_e0;
  /* temperature law (VarFcnSG.h:126-142, VarFcnNASG.h:112, VarFcnMG.h:113-121): cv, T0, e0, rho0, cp, h0 */
  double t_cv, t_T0, t_e0, t_rho0, t_cp, t_h0;
  /* viscosity (ViscoFcn.h): 0 none, 1 constant (mu, bulk), 2 Sutherland (mu = mu0, su_T0, su_S, bulk) */
  int visc;
  double mu, bulk, su_T0, su_S;
  /* heat diffusion (HeatDiffusionFcn.h): constant diffusivity, 0 = none */
  double kappa;
} orc_eos;

/* Derived constants, computed as the reference constructors do
 * (VarFcnSG.h:124-130, VarFcnNASG.h:104-109, VarFcnMG.h:117-119, VarFcnJWL.h:95-98). */
static inline void orc_eos_finalize(orc_eos *m) {
  m->gam1 = m->gam - 1.0;
  m->gam_pc = m->gam * m->pc;
  m->rho0_c0_c0 = m->rho0 * m->c0 * m->c0;
  m->Gamma0_over_2 = 0.5 * m->Gamma0;
  m->Gamma0_rho0 = m->Gamma0 * m->rho0;
  m->R1rho0 = m->R1 * m->rho0;
  m->R2rho0 = m->R2 * m->rho0;
  m->omega_over_R1rho0 = (m->R1rho0 != 0.0) ? m->omega / m->R1rho0 : 0.0;
  m->omega_over_R2rho0 = (m->R2rho0 != 0.0) ? m->omega / m->R2rho0 : 0.0;
}

static inline double orc_jwl_fun(const orc_eos *m, double rho) {
  return m->A1 * (1.0 - m->omega_over_R1rho0 * rho) * exp(-m->R1rho0 / rho)
       + m->A2 * (1.0 - m->omega_over_R2rho0 * rho) * exp(-m->R2rho0 / rho);
}

static inline double orc_pressure(const orc_eos *m, double rho, double e) {
  switch (m->type) {
  case ORC_EOS_SG: return m->gam1 * rho * e - m->gam * m->pstiff;
  case ORC_EOS_NASG: return m->gam1 * (e - m->q) / (1.0 / rho - m->b) - m->gam_pc;
  case ORC_EOS_MG: {
    double eta = 1.0 - m->rho0 / rho;
    return m->rho0_c0_c0 * eta * (1.0 - m->Gamma0_over_2 * eta) / ((1.0 - m->s * eta) * (1.0 - m->s * eta))
         + m->Gamma0_rho0 * (e - m->e0);
  }
  case ORC_EOS_JWL: return m->omega * rho * e + orc_jwl_fun(m, rho);
  default: return m->d_p0;
  }
}

static inline double orc_internal_energy(const orc_eos *m, double rho, double p) {
  switch (m->type) {
  case ORC_EOS_SG: return (p + m->gam * m->pstiff) / (m->gam1 * rho);
  case ORC_EOS_NASG: return (1.0 / m->gam1) * (p + m->gam_pc) * (1.0 / rho - m->b) + m->q;
  case ORC_EOS_MG: {
    double eta = 1.0 - m->rho0 / rho;
    return (p - m->rho0_c0_c0 * eta * (1.0 - m->Gamma0_over_2 * eta) / ((1.0 - m->s * eta) * (1.0 - m->s * eta)))
           / m->Gamma0_rho0 + m->e0;
  }
  case ORC_EOS_JWL: return (p - orc_jwl_fun(m, rho)) / (m->omega * rho);
  default: return m->d_e0;
  }
}

/* VarFcn*::GetTemperature(rho, e): SG VarFcnSG.h:70-78 (three laws chosen as VarFcnSG.h:141-142),
 * NASG VarFcnNASG.h:61, MG VarFcnMG.h:172-179 (its e0 is the EOS reference energy); NaN for the
 * others, which have no temperature law. */
static inline double orc_temperature(const orc_eos *m, double rho, double e) {
  const double invcv = m->t_cv == 0.0 ? 0.0 : 1.0 / m->t_cv, invcp = m->t_cp == 0.0 ? 0.0 : 1.0 / m->t_cp;
  switch (m->type) {
  case ORC_EOS_SG:
    if (m->t_cv > 0 && m->t_rho0 > 0)
      return invcv * (e + m->pstiff / rho) +
             pow(rho / m->t_rho0, m->gam1) * (m->t_T0 - invcv * (m->t_e0 + m->pstiff / m->t_rho0));
    if (m->t_cp > 0 && m->t_cv <= 0.0) {
      double p = orc_pressure(m, rho, e);
      return m->t_T0 + invcp * (e + p / rho - m->t_h0);
    }
    return m->t_T0 + invcv * (e - m->t_e0);
  case ORC_EOS_NASG:
    return invcv * (e - m->q - m->pc * (1.0 / rho - m->b));
  case ORC_EOS_MG:
    if (m->t_cp > 0 && m->t_cv <= 0.0) {
      double p = orc_pressure(m, rho, e);
      return m->t_T0 + invcp * (e + p / rho - m->t_h0);
    }
    return m->t_T0 + invcv * (e - m->e0);
  default:
    return NAN;
  }
}

static inline double orc_dpdrho(const orc_eos *m, double rho, double e) {
  switch (m->type) {
  case ORC_EOS_SG: return m->gam1 * e;
  case ORC_EOS_NASG: { double V = 1.0 / rho; return m->gam1 * V * V * (e - m->q) / ((V - m->b) * (V - m->b)); }
  case ORC_EOS_MG: {
    double eta = 1.0 - m->rho0 / rho;
    double S = 1.0 - m->s * eta;
    return m->rho0_c0_c0 * (1.0 + (m->s - m->Gamma0) * eta) / (S * S * S) * m->rho0 / (rho * rho);
  }
  case ORC_EOS_JWL:
    return m->omega * e
         + m->A1 * (-m->omega_over_R1rho0 + m->R1rho0 / (rho * rho) - m->omega_over_R1rho0 * m->R1rho0 / rho) * exp(-m->R1rho0 / rho)
         + m->A2 * (-m->omega_over_R2rho0 + m->R2rho0 / (rho * rho) - m->omega_over_R2rho0 * m->R2rho0 / rho) * exp(-m->R2rho0 / rho);
  default: return 0.0;
  }
}

static inline double orc_big_gamma(const orc_eos *m, double rho, double e) {
  (void)e;
  switch (m->type) {
  case ORC_EOS_SG: return m->gam1;
  case ORC_EOS_NASG: return m->gam1 / (1.0 - m->b * rho);
  case ORC_EOS_MG: return m->Gamma0_rho0 / rho;
  case ORC_EOS_JWL: return m->omega;
  default: return 0.0;
  }
}

/* VarFcnBase.h:215-219; the dummy material returns DBL_MIN (VarFcnDummy.h:43). */
static inline double orc_c2(const orc_eos *m, double rho, double e) {
  if (m->type == ORC_EOS_DUMMY) return DBL_MIN;
  return orc_dpdrho(m, rho, e) + orc_pressure(m, rho, e) / rho * orc_big_gamma(m, rho, e);
}

/* VarFcnBase.h:250-254 */
static inline double orc_total_enthalpy(const orc_eos *m, const double *V) {
  if (m->type == ORC_EOS_DUMMY) return m->d_e0;
  double e = orc_internal_energy(m, V[0], V[4]);
  return e + 0.5 * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3]) + V[4] / V[0];
}

/* CheckState(rho,p): true = invalid. SG VarFcnSG.h:96-108, NASG VarFcnNASG.h:73-85,
 * others VarFcnBase.h:113-130. */
static inline int orc_check_state_rp(const orc_eos *m, double rho, double p) {
  if (m->type == ORC_EOS_DUMMY) return 0;
  if (!isfinite(rho) || !isfinite(p)) return 1;
  if (m->type == ORC_EOS_SG) return (rho <= 0.0 || p + m->pstiff <= 0.0);
  if (m->type == ORC_EOS_NASG) return (rho <= 0.0 || p + m->pc <= 0.0);
  if (rho <= 0.0) return 1;
  double e = orc_internal_energy(m, rho, p);
  double c2 = orc_dpdrho(m, rho, e) + p / rho * orc_big_gamma(m, rho, e);
  return c2 <= 0;
}

/* CheckState(V): VarFcnBase.h:133-142 */
static inline int orc_check_state(const orc_eos *m, const double *V) {
  if (m->type == ORC_EOS_DUMMY) return 0;
  if (!isfinite(V[0]) || !isfinite(V[1]) || !isfinite(V[2]) || !isfinite(V[3]) || !isfinite(V[4])) return 1;
  return orc_check_state_rp(m, V[0], V[4]);
}

/* VarFcnBase.h:169-181 */
static inline void orc_c2p(const orc_eos *m, const double *U, double *V) {
  if (m->type == ORC_EOS_DUMMY) { V[0] = m->d_rho0; V[1] = V[2] = V[3] = 0.0; V[4] = m->d_p0; return; }
  V[0] = U[0];
  double invRho = 1.0 / U[0];
  V[1] = U[1] * invRho;
  V[2] = U[2] * invRho;
  V[3] = U[3] * invRho;
  double e = (U[4] - 0.5 * V[0] * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3])) * invRho;
  V[4] = orc_pressure(m, V[0], e);
}

/* VarFcnBase.h:185-197 */
static inline void orc_p2c(const orc_eos *m, const double *V, double *U) {
  if (m->type == ORC_EOS_DUMMY) { U[0] = m->d_rho0; U[1] = U[2] = U[3] = 0.0; U[4] = m->d_rho0 * m->d_e0; return; }
  U[0] = V[0];
  U[1] = V[0] * V[1];
  U[2] = V[0] * V[2];
  U[3] = V[0] * V[3];
  double e = orc_internal_energy(m, V[0], V[4]);
  U[4] = V[0] * (e + 0.5 * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3]));
}

/* VarFcnBase.h:259-294 (without the optional U update). Returns 1 if clipped. */
static inline int orc_clip(const orc_eos *m, double *V) {
  if (m->type == ORC_EOS_DUMMY) { V[0] = m->d_rho0; V[1] = V[2] = V[3] = 0.0; V[4] = m->d_p0; return 0; }
  int clip = 0;
  if (V[0] < m->rhomin) { V[0] = m->rhomin; clip = 1; }
  if (V[4] < m->pmin) { V[4] = m->pmin; clip = 1; }
  if (V[0] > m->rhomax) { V[0] = m->rhomax; clip = 1; }
  if (V[4] > m->pmax) { V[4] = m->pmax; clip = 1; }
  return clip;
}

#endif
