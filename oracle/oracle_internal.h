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
  double dx, dy, dz;
  /* mesh per axis: cell widths with one ghost per side (the boundary cell's), index i + 1; cell
   * centres of cells 0..n-1; limiter coefficients per cell (Reconstructor.cpp:53-121) */
  double *w[3], *mx[3], *cA[3], *cB[3];
  int *cK[3];
  double *face[6];          /* vl, vr, vb, vt, vk, vf (ghosted, AoS) */
  double *Un, *U1, *R, *V1, *V2;
  double *Fx, *Fy, *Fz;     /* per face: localflux1[5], localflux2[5] */
  int *ferr;
  long n_reverted, n_clipped, n_riemann_err;
  /* multiphase state (oracle/multiphase.c) */
  int store_riemann;        /* 1 during the first stage of a step */
  int rs_stamp;             /* time step of the stored Riemann solutions */
  double *rsV[3];           /* per face (keyed by the plus cell): Vsm[5], Vsp[5] */
  int *rsId[3];             /* per face: idm, idp, stamp */
  double *phi_tmp, *phi1, *Rls, *phi0, *phibk, *rphi1, *sign, *Rre;
  int *tag, *IDn;
  double *xc[3];            /* cell-centre coordinates with one ghost on each side */
  double *fd_l[3], *fd_r[3], *fd_r1y;   /* FD3 coefficients per index along each axis */
  double reinit_cfl_state;  /* LevelSetReinitializer::cfl (persists, divided by 1.5 on retries) */
  double *Dv[3], *Dh[3];    /* viscous (plus[5], minus[5]) and heat fluxes per face (oracle/diffusion.c) */
  const double *phi_n;      /* level set at t_n for the interface normal (RiemannNormal LevelSet/Average) */
  long n_reinit_iters, n_unresolved;
  int steps_done;           /* time steps of earlier orc_run calls: the numbering continues */
  double t_done;            /* and their time */
  /* mesh faces cut by embedded walls (orc_set_embedded_faces): per direction and face key
   * ((k*(ny+1)+j)*(nx+1)+i, keyed by the plus cell) the wall record or -1; 12 doubles per record
   * (nwallf, vwallf, nwallb, vwallb) */
  int *eb_rec[3];
  double *eb_wall;
  int eb_n, eb_normal_mode, eb_recon;
};

#define ORC_W(c, d, i) ((c)->w[d][(i) + 1])

#define ORC_FKEY(c, i, j, k) ((((long)(k)) * ((c)->ny + 1) + (j)) * (long)((c)->nx + 1) + (i))

#define ORC_IDX(c, i, j, k) ((((long)(k) + 1) * (c)->gy + ((j) + 1)) * (long)(c)->gx + ((i) + 1))

/* building blocks shared by fv.c and multiphase.c */
void orc_p2c_all(orc_ctx *c, const double *V, const int *ID, double *U);
void orc_axpby(orc_ctx *c, double *y, double a, double b, const double *x);
long orc_c2p_clip(orc_ctx *c, const double *U, const int *ID, double *V);
void orc_rs_store(orc_ctx *c, int dir, long hi, const double *Vsm, int idm, const double *Vsp, int idp);
int orc_riemann_solve(const orc_eos *mat, const orc_riemann_cfg *cfg, const double *dir, const double *Vm, int idl,
                      const double *Vp, int idr, double *Vs, int *id, double *Vsm, double *Vsp);
int orc_riemann_onesided_solve(const orc_eos *mat, const orc_riemann_cfg *cfg, const double *dir, const double *Vm,
                               int idl, const double *Ustar, double *Vs, int *id, double *Vsm);
double orc_pow_quarter(double x);
double orc_pow_fifth(double x);

/* exact Riemann solver (oracle/riemann.c) */
int orc_interface_flux(orc_ctx *c, int dir, const double *vm, int idm, const double *vp, int idp,
                       long lo, long hi, double *lf1, double *lf2);
void orc_flux_godunov(const orc_ctx *c, int dir, const double *Vm, const double *Vp, int id, double *F);
void orc_riemann_begin_stage(orc_ctx *c);
void orc_riemann_end_stage(orc_ctx *c);

/* viscous and heat-diffusion face fluxes (oracle/diffusion.c) */
int orc_has_diffusion(const orc_ctx *c);
void orc_diffusion(orc_ctx *c, const double *V, const int *ID, double *Dv[3], double *Dh[3]);

/* level set + end-of-step tasks (oracle/multiphase.c) */
int orc_step_all(orc_ctx *c, double *V, int *ID, double *phi, double dt, int step);
void orc_ls_free(orc_ctx *c);
void orc_ls_init(orc_ctx *c);

#endif
