/* oracle/oracle_internal.h -- shared internals of the CPU oracle (test infrastructure only). */
#ifndef M2C_ORACLE_INTERNAL_H
#define M2C_ORACLE_INTERNAL_H

#include "m2c_oracle.h"

/* one stored Riemann star state (RiemannSolutions.h:12-30) */
typedef struct { long key; double V[5]; int id; } orc_rs_entry;
typedef struct { orc_rs_entry *e; long n, cap; } orc_rs_list;

struct orc_ctx {
  orc_config cfg;
  int nx, ny, nz, gx, gy, gz;
  double dx, dy, dz, vol, area[3];
  double A[3], B[3];
  int K[3];
  double *face[6];          /* vl, vr, vb, vt, vk, vf (ghosted, AoS) */
  double *Un, *U1, *R, *V1, *V2;
  double *Fx, *Fy, *Fz;     /* per face: localflux1[5], localflux2[5] */
  int *ferr;
  long n_reverted, n_clipped, n_riemann_err;
  /* multiphase state (oracle/multiphase.c) */
  int store_riemann;        /* 1 during the first stage of a step */
  orc_rs_list rs[6];        /* left/right x, y, z (RiemannSolutions.h:14-19) */
  void *ls;                 /* level-set state */
};

#define ORC_IDX(c, i, j, k) ((((long)(k) + 1) * (c)->gy + ((j) + 1)) * (long)(c)->gx + ((i) + 1))

/* exact Riemann solver (oracle/riemann.c) */
int orc_interface_flux(orc_ctx *c, int dir, const double *vm, int idm, const double *vp, int idp,
                       long lo, long hi, double *lf1, double *lf2);
void orc_flux_godunov(const orc_ctx *c, int dir, const double *Vm, const double *Vp, int id, double *F);
void orc_riemann_begin_stage(orc_ctx *c);
void orc_riemann_end_stage(orc_ctx *c);

/* level set + end-of-step tasks (oracle/multiphase.c) */
int orc_step_all(orc_ctx *c, double *V, int *ID, double dt, int step);
void orc_ls_free(orc_ctx *c);

#endif
