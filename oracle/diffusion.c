/*
 * oracle/diffusion.c -- viscous and heat-diffusion terms of the residual (test infrastructure only).
 *
 * Restates, loop for loop, ViscosityOperator::AddDiffusionFluxes (ViscosityOperator.cpp:200-404) and
 * HeatDiffusionOperator::AddDiffusionFluxes (HeatDiffusionOperator.cpp:64-139) with the gradients of
 * GradientCalculatorCentral (GradientCalculatorCentral.cpp:22-207 node derivatives with the
 * one-sided rows, :208-245 face derivatives, :249-424 coefficients, :430-494 central differences at
 * faces) and the face interpolation of InterpolatorLinear (Interpolator.cpp:32-130, :150-181). The
 * results are added to the residual sums in the order the reference accumulates them: after the
 * advective fluxes, the viscous loop, then the heat loop, each in (k, j, i) cell order with the
 * i-1/2, j-1/2, k-1/2 faces of a cell in turn (each face adds to its plus cell and subtracts from its
 * minus cell).
 *
 * Reproduced as written in the reference:
 *  - the k-1/2 face of two different materials evaluates the lower cell's stress with the upper
 *    cell's viscosity function (ViscosityOperator.cpp:386);
 *  - Sutherland's law evaluates the temperature at (rho, p) where (rho, e) is meant (ViscoFcn.h:255);
 *  - the heat flux of the j-1/2 and k-1/2 faces divides by the i-1/2 face's diffusivity sum, which on
 *    the north ghost row and the top ghost plane is the last one computed
 *    (HeatDiffusionOperator.cpp:109-127).
 * Needs at least two cells in every direction (with one, the reference reads unpopulated edge
 * ghosts).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

#define XC(c, d, q) ((c)->xc[d][(q) + 1])

static const double *vel(const orc_ctx *c, const double *V, int i, int j, int k) {
  return V + 5 * ORC_IDX(c, i, j, k);
}

/* linear interpolation weights at the face q-1/2 between cells q-1 and q (Interpolator.cpp:150-181):
 * value = wl * a[q-1] + wr * a[q] */
static void interp_w(const orc_ctx *c, int d, int q, double *wl, double *wr) {
  const double h[3] = {c->dx, c->dy, c->dz};
  double xm = XC(c, d, q) - 0.5 * h[d];   /* dxyz of real and ghost cells is h */
  double dd = XC(c, d, q) - XC(c, d, q - 1);
  *wl = (XC(c, d, q) - xm) / dd;
  *wr = (xm - XC(c, d, q - 1)) / dd;
}

/* d(comp)/dx_d at the node (i, j, k) (GradientCalculatorCentral.cpp:22-207, coefficients :249-424);
 * `ghostrow` says that one of the other two indices is a ghost index */
static double node_deriv(const orc_ctx *c, const double *V, int d, int i, int j, int k, int comp) {
  const int N[3] = {c->nx, c->ny, c->nz};
  int q = d == 0 ? i : (d == 1 ? j : k);
  int ghostrow;
  if (d == 0) ghostrow = (j == -1 || j == c->ny || k == -1 || k == c->nz);
  else if (d == 1) ghostrow = (i == -1 || i == c->nx || k == -1 || k == c->nz);
  else ghostrow = (i == -1 || i == c->nx || j == -1 || j == c->ny);
  int im[3] = {i, j, k}, ip[3] = {i, j, k};
  im[d] -= 1;
  ip[d] += 1;
  const double u = vel(c, V, i, j, k)[comp];
  if (q == 0 && ghostrow) {
    double coeff = 1.0 / (XC(c, d, q + 1) - XC(c, d, q));
    double c1 = -coeff, c2 = coeff;
    return c1 * u + c2 * vel(c, V, ip[0], ip[1], ip[2])[comp];
  }
  if (q == N[d] - 1 && ghostrow) {
    double coeff = 1.0 / (XC(c, d, q) - XC(c, d, q - 1));
    double c0 = -coeff, c1 = coeff;
    return c0 * vel(c, V, im[0], im[1], im[2])[comp] + c1 * u;
  }
  double d0 = XC(c, d, q) - XC(c, d, q - 1);
  double d1 = XC(c, d, q + 1) - XC(c, d, q);
  double d2 = d0 + d1;
  double c0 = -d1 / (d0 * d2), c1 = 1.0 / d0 - 1.0 / d1, c2 = d0 / (d1 * d2);
  return c0 * vel(c, V, im[0], im[1], im[2])[comp] + c1 * u + c2 * vel(c, V, ip[0], ip[1], ip[2])[comp];
}

/* velocity and its gradient at the face `dir` of cell (i, j, k) (the face between it and its lower
 * neighbour along dir): V3 (interpolated), D[a][m] = d u_m / d x_a */
static void face_grad(const orc_ctx *c, const double *V, int dir, int i, int j, int k, double V3[3], double D[3][3]) {
  int lo[3] = {i, j, k};
  lo[dir] -= 1;
  const int q = dir == 0 ? i : (dir == 1 ? j : k);
  double wl, wr;
  interp_w(c, dir, q, &wl, &wr);
  const double *a = vel(c, V, lo[0], lo[1], lo[2]), *b = vel(c, V, i, j, k);
  const double dd = XC(c, dir, q) - XC(c, dir, q - 1);
  for (int m = 0; m < 3; m++) V3[m] = wl * a[1 + m] + wr * b[1 + m];
  for (int ax = 0; ax < 3; ax++)
    for (int m = 0; m < 3; m++) {
      if (ax == dir) D[ax][m] = (b[1 + m] - a[1 + m]) / dd;   /* GradientCalculatorCentral.cpp:430-494 */
      else D[ax][m] = wl * node_deriv(c, V, ax, lo[0], lo[1], lo[2], 1 + m) + wr * node_deriv(c, V, ax, i, j, k, 1 + m);
    }
}

/* viscosity of material m at (rho, p) (ViscoFcn.h) */
static double mu_of(const orc_eos *m, double rho, double p) {
  if (m->visc == 2) {
    double T = orc_temperature(m, rho, p);   /* (rho, p): as ViscoFcn.h:255 calls it */
    return m->mu * pow(T / m->su_T0, 1.5) * (m->su_T0 + m->su_S) / (T + m->su_S);
  }
  return m->mu;
}

/* EvaluateViscousFluxFunction_F/G/H (ViscoFcn.h:61-110) with the Constant / Sutherland stresses */
static void visc_flux(const orc_eos *m, int dir, double D[3][3], const double V3[3], double rho, double p, double *F) {
  if (m->visc == 0) {
    F[0] = F[1] = F[2] = F[3] = F[4] = 0.0;
    return;
  }
  const double *dVdx = D[0], *dVdy = D[1], *dVdz = D[2];
  double dyna = mu_of(m, rho, p), bulk = m->bulk;
  double div = dVdx[0] + dVdy[1] + dVdz[2];
  double t0, t1, t2;
  if (dir == 0) {
    t0 = bulk * div + dyna * (2.0 * dVdx[0] - 2.0 / 3.0 * div);
    t1 = dyna * (dVdx[1] + dVdy[0]);
    t2 = dyna * (dVdx[2] + dVdz[0]);
  } else if (dir == 1) {
    t0 = dyna * (dVdx[1] + dVdy[0]);
    t1 = bulk * div + dyna * (2.0 * dVdy[1] - 2.0 / 3.0 * div);
    t2 = dyna * (dVdy[2] + dVdz[1]);
  } else {
    t0 = dyna * (dVdx[2] + dVdz[0]);
    t1 = dyna * (dVdy[2] + dVdz[1]);
    t2 = bulk * div + dyna * (2.0 * dVdz[2] - 2.0 / 3.0 * div);
  }
  F[0] = 0.0;
  F[1] = t0;
  F[2] = t1;
  F[3] = t2;
  F[4] = V3[0] * t0 + V3[1] * t1 + V3[2] * t2;
}

int orc_has_diffusion(const orc_ctx *c) {
  for (int m = 0; m < c->cfg.nmat; m++)
    if (c->cfg.mat[m].visc != 0 || c->cfg.mat[m].kappa != 0.0) return 1;
  return 0;
}

/* Per face (x: (nx+1)*ny*nz, y: nx*(ny+1)*nz, z: nx*ny*(nz+1), indexed as Fx/Fy/Fz): Dv[10]
 * (viscous flux times area for the plus cell, then for the minus cell) and Dh[1] (heat flux times
 * area, the same for both cells). Faces the reference skips hold zeros. */
void orc_diffusion(orc_ctx *c, const double *V, const int *ID, double *Dv[3], double *Dh[3]) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const long nf[3] = {(long)(nx + 1) * ny * nz, (long)nx * (ny + 1) * nz, (long)nx * ny * (nz + 1)};
  for (int d = 0; d < 3; d++) {
    memset(Dv[d], 0, sizeof(double) * 10 * nf[d]);
    memset(Dh[d], 0, sizeof(double) * nf[d]);
  }
  const orc_eos *mat = c->cfg.mat;
  const int inactive = c->cfg.inactive_id;
  const double dx = c->dx, dy = c->dy, dz = c->dz;
  int visc = 0, heat = 0;
  for (int m = 0; m < c->cfg.nmat; m++) {
    visc |= mat[m].visc != 0;
    heat |= mat[m].kappa != 0.0;
  }
  /* viscous loop: k0..kkmax-1, j0..jjmax-1, i0..iimax-1 (ViscosityOperator.cpp:259-393) */
  if (visc) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= nz; k++)
      for (int j = 0; j <= ny; j++)
        for (int i = 0; i <= nx; i++) {
          const long me = ORC_IDX(c, i, j, k);
          const int myid = ID[me];
          if (myid == inactive) continue;
          const double rho = V[5 * me], p = V[5 * me + 4];
          for (int dir = 0; dir < 3; dir++) {
            if (dir == 0 && (k == nz || j == ny)) continue;
            if (dir == 1 && (i == nx || k == nz)) continue;
            if (dir == 2 && (i == nx || j == ny)) continue;
            int lo[3] = {i, j, k};
            lo[dir] -= 1;
            const long nb = ORC_IDX(c, lo[0], lo[1], lo[2]);
            const int id2 = ID[nb];
            double V3[3], D[3][3], F[5], G[5];
            face_grad(c, V, dir, i, j, k, V3, D);
            const double area = dir == 0 ? dy * dz : (dir == 1 ? dx * dz : dx * dy);
            double *out;
            if (dir == 0) out = Dv[0] + 10 * (((long)k * ny + j) * (nx + 1) + i);
            else if (dir == 1) out = Dv[1] + 10 * (((long)k * (ny + 1) + j) * nx + i);
            else out = Dv[2] + 10 * (((long)k * ny + j) * nx + i);
            if (id2 == myid) {      /* one flux at the averaged state */
              visc_flux(&mat[myid], dir, D, V3, 0.5 * (rho + V[5 * nb]), 0.5 * (p + V[5 * nb + 4]), F);
              for (int q = 0; q < 5; q++) out[q] = out[5 + q] = F[q] * area;
            } else {
              visc_flux(&mat[myid], dir, D, V3, rho, p, F);
              for (int q = 0; q < 5; q++) out[q] = area * F[q];
              if (id2 != inactive) {
                const orc_eos *m2 = dir == 2 ? &mat[myid] : &mat[id2];
                visc_flux(m2, dir, D, V3, V[5 * nb], V[5 * nb + 4], G);
                for (int q = 0; q < 5; q++) out[5 + q] = area * G[q];
              }
            }
          }
        }
  }
  /* heat loop (HeatDiffusionOperator.cpp:68-131): temperatures on the ghosted box, then faces in
   * cell order with the diffusivity sum `denom` carried from cell to cell */
  if (heat) {
    const long n = (long)c->gx * c->gy * c->gz;
    double *T = (double *)malloc(sizeof(double) * n);
    for (long q = 0; q < n; q++) {
      const orc_eos *m = &mat[ID[q] >= 0 && ID[q] < c->cfg.nmat ? ID[q] : 0];
      double e = orc_internal_energy(m, V[5 * q], V[5 * q + 4]);
      T[q] = orc_temperature(m, V[5 * q], e);
    }
    double denom = 0.0;
    for (int k = 0; k <= nz; k++)
      for (int j = 0; j <= ny; j++)
        for (int i = 0; i <= nx; i++) {
          const long me = ORC_IDX(c, i, j, k);
          const int myid = ID[me];
          if (myid == inactive) continue;
          const double myk = mat[myid].kappa;
          if (k != nz && j != ny) {
            const long nb = ORC_IDX(c, i - 1, j, k);
            const double neighk = mat[ID[nb]].kappa;
            denom = myk + neighk;
            double dT = (T[me] - T[nb]) / (XC(c, 0, i) - XC(c, 0, i - 1));
            double flux = (denom == 0) ? 0.0 : 2.0 * myk * neighk / denom * dT;
            flux *= dy * dz;
            Dh[0][((long)k * ny + j) * (nx + 1) + i] = flux;
          }
          if (i != nx && k != nz) {
            const long nb = ORC_IDX(c, i, j - 1, k);
            const double neighk = mat[ID[nb]].kappa;
            double dT = (T[me] - T[nb]) / (XC(c, 1, j) - XC(c, 1, j - 1));
            double flux = (denom == 0) ? 0.0 : 2.0 * myk * neighk / denom * dT;
            flux *= dx * dz;
            Dh[1][((long)k * (ny + 1) + j) * nx + i] = flux;
          }
          if (i != nx && j != ny) {
            const long nb = ORC_IDX(c, i, j, k - 1);
            const double neighk = mat[ID[nb]].kappa;
            double dT = (T[me] - T[nb]) / (XC(c, 2, k) - XC(c, 2, k - 1));
            double flux = (denom == 0) ? 0.0 : 2.0 * myk * neighk / denom * dT;
            flux *= dx * dy;
            Dh[2][((long)k * ny + j) * nx + i] = flux;
          }
        }
    free(T);
  }
}
