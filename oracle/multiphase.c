/*
 * oracle/multiphase.c -- CPU restatement of M2C's level-set advection, reinitialisation and
 * two-material update, and of the time step that ties them to the flow (test infrastructure).
 *
 *   level-set ghost fill        LevelSetOperator.cpp:934-1033
 *   FD3 coefficients            GradientCalculatorFD3.cpp:300-469 (closed-form Lagrange weights of
 *                               the same 4-point stencils; the reference solves the 4x4 Taylor
 *                               system with Eigen's QR, which is absent, so the coefficients agree
 *                               with it to rounding only -- "parity unpinned" at the ulp level)
 *   FD3 derivatives             GradientCalculatorFD3.cpp:54-171 (including the right-biased y
 *                               branch that tests i < NY-1, :415)
 *   advection residual          LevelSetOperator.cpp:1066-1141
 *   reinitialisation            LevelSetReinitializer.cpp:74-172 (FIXED first layer), :297-430,
 *                               :594-615, :953-1035, :1242-1355, :1970-1999
 *   material ID update          MultiPhaseOperator.cpp:127-330, ghost IDs :796-819
 *   state update after motion   MultiPhaseOperator.cpp:857-1028, FixUnresolvedNodes :1156-1320
 *   Riemann-solution store      RiemannSolutions.h:12-30, SpaceOperator.cpp:2518-2523
 *   time step                   TimeIntegrator.cpp:150-226 (FE), :291-428 (RK2), :496-690 (RK3),
 *                               end-of-step tasks :695-797
 * Parity pins: tests/test_oracle_levelset.py (FD3 exact on cubics, signed distance of a plane is a
 * reinitialisation fixed point, sign-change ID rule, constant-velocity advection of a plane).
 */
#include "oracle_internal.h"
#include "flux.h"
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

static double dmin(double a, double b) { return (b < a) ? b : a; }
static double dmax(double a, double b) { return (a < b) ? b : a; }

/* ------------------------------------------------------------ set-up -------------------- */

/* closed-form derivative-at-0 weights of the 4-point interpolant through nodes d[0..3], d[c] = 0 */
static void fd_weights(const double *d, int cidx, double *w) {
  for (int k = 0; k < 4; k++) {
    if (k == cidx) continue;
    double num = 1.0, den = 1.0;
    for (int l = 0; l < 4; l++)
      if (l != k && l != cidx) num = num * (-d[l]);
    for (int m = 0; m < 4; m++)
      if (m != k) den = den * (d[k] - d[m]);
    w[k] = num / den;
  }
  double s = 0.0;
  for (int m = 0; m < 4; m++)
    if (m != cidx) s = s + 1.0 / d[m];
  w[cidx] = -s;
}

void orc_ls_init(orc_ctx *c) {
  long n = orc_ncells_ghosted(c);
  int N[3] = {c->nx, c->ny, c->nz};
  for (int d = 0; d < 3; d++) {
    c->rsV[d] = (double *)calloc((size_t)n * 10, sizeof(double));
    c->rsId[d] = (int *)calloc((size_t)n * 3, sizeof(int));
    for (long q = 0; q < n; q++) c->rsId[d][3 * q + 2] = -1;
    /* coordinates x[-1..N] (MeshGenerator.cpp:153-157, SpaceOperator.cpp:188-230) */
    double *x = (double *)calloc((size_t)N[d] + 2, sizeof(double));
    for (int i = 0; i < N[d]; i++) x[i + 1] = c->mx[d][i];
    x[0] = x[1] - ORC_W(c, d, 0);
    x[N[d] + 1] = x[N[d]] + ORC_W(c, d, N[d] - 1);
    c->xc[d] = x;
    /* FD3 coefficients per index along the axis */
    double *L = (double *)calloc((size_t)N[d] * 4, sizeof(double));
    double *R = (double *)calloc((size_t)N[d] * 4, sizeof(double));
    for (int i = 0; i < N[d]; i++) {
      double *xi = x + 1;   /* xi[-1..N] */
      if (i > 0 && i + 1 <= N[d]) {
        double dd[4] = {(i - 2 >= -1) ? xi[i - 2] - xi[i] : 0.0 - xi[i], xi[i - 1] - xi[i], 0.0, xi[i + 1] - xi[i]};
        fd_weights(dd, 2, L + 4 * i);
      } else {
        double dx = xi[i] - xi[i - 1];
        L[4 * i] = 0.0; L[4 * i + 1] = -1.0 / dx; L[4 * i + 2] = 1.0 / dx; L[4 * i + 3] = 0.0;
      }
      if (i < N[d] - 1) {
        double dd[4] = {xi[i - 1] - xi[i], 0.0, xi[i + 1] - xi[i], xi[i + 2] - xi[i]};
        fd_weights(dd, 1, R + 4 * i);
      } else {
        double dx = xi[i + 1] - xi[i];
        R[4 * i] = 0.0; R[4 * i + 1] = -1.0 / dx; R[4 * i + 2] = 1.0 / dx; R[4 * i + 3] = 0.0;
      }
    }
    c->fd_l[d] = L;
    c->fd_r[d] = R;
  }
  /* first-order right-biased y coefficients, used where the reference's branch tests i >= NY-1 */
  c->fd_r1y = (double *)calloc((size_t)c->ny * 4, sizeof(double));
  for (int j = 0; j < c->ny; j++) {
    double *yj = c->xc[1] + 1;
    double dy = yj[j + 1] - yj[j];
    c->fd_r1y[4 * j] = 0.0; c->fd_r1y[4 * j + 1] = -1.0 / dy; c->fd_r1y[4 * j + 2] = 1.0 / dy; c->fd_r1y[4 * j + 3] = 0.0;
  }
  c->phi_tmp = (double *)calloc((size_t)n, sizeof(double));
  c->phi1 = (double *)calloc((size_t)n, sizeof(double));
  c->Rls = (double *)calloc((size_t)n, sizeof(double));
  c->phi0 = (double *)calloc((size_t)n, sizeof(double));
  c->phibk = (double *)calloc((size_t)n, sizeof(double));
  c->rphi1 = (double *)calloc((size_t)n, sizeof(double));
  c->sign = (double *)calloc((size_t)n, sizeof(double));
  c->Rre = (double *)calloc((size_t)n, sizeof(double));
  c->tag = (int *)calloc((size_t)n, sizeof(int));
  c->IDn = (int *)calloc((size_t)n, sizeof(int));
  c->reinit_cfl_state = c->cfg.reinit_cfl;
}

void orc_ls_free(orc_ctx *c) {
  for (int d = 0; d < 3; d++) {
    free(c->rsV[d]); free(c->rsId[d]); free(c->xc[d]); free(c->fd_l[d]); free(c->fd_r[d]);
  }
  free(c->fd_r1y);
  free(c->phi_tmp); free(c->phi1); free(c->Rls); free(c->phi0); free(c->phibk); free(c->rphi1); free(c->sign);
  free(c->Rre); free(c->tag); free(c->IDn);
}

#define X(c, d, i) ((c)->xc[d][(i) + 1])

/* ------------------------------------------------------------ Riemann-solution store ---- */

/* left[(k,j,i)] = (Vsm, idm), right[(k,j,i-1)] = (Vsp, idp): one record per face, keyed by the
 * face's plus cell (SpaceOperator.cpp:2518-2523 and the y/z analogues) */
void orc_rs_store(orc_ctx *c, int dir, long hi, const double *Vsm, int idm, const double *Vsp, int idp) {
  double *r = c->rsV[dir] + 10 * hi;
  memcpy(r, Vsm, 5 * sizeof(double));
  memcpy(r + 5, Vsp, 5 * sizeof(double));
  int *q = c->rsId[dir] + 3 * hi;
  q[0] = idm; q[1] = idp; q[2] = c->rs_stamp;
}

/* ------------------------------------------------------------ level-set ghosts --------- */

void orc_apply_bc_phi(const orc_ctx *c, double *phi) {
  int N[3] = {c->nx, c->ny, c->nz};
  for (int side = 0; side < 6; side++) {
    int dim = side / 2, hi = side % 2, bc = c->cfg.ls_bc[side];
    int n0 = dim == 0 ? c->ny : c->nx, n1 = dim == 2 ? c->ny : c->nz;
    for (int b = 0; b < n1; b++)
      for (int a = 0; a < n0; a++) {
        int ijk[3];
        if (dim == 0) { ijk[1] = a; ijk[2] = b; }
        else if (dim == 1) { ijk[0] = a; ijk[2] = b; }
        else { ijk[0] = a; ijk[1] = b; }
        int g = hi ? N[dim] : -1;
        ijk[dim] = g;
        long gi = ORC_IDX(c, ijk[0], ijk[1], ijk[2]);
        int im[3] = {ijk[0], ijk[1], ijk[2]};
        if (bc == ORC_BC_PERIODIC) im[dim] = hi ? 0 : N[dim] - 1;
        else im[dim] = hi ? N[dim] - 1 : 0;
        long ii = ORC_IDX(c, im[0], im[1], im[2]);
        if (bc == 2 || bc == 3) {
          /* linear extrapolation from the two nearest interior nodes when the axis has two */
          if (N[dim] >= 2) {
            int s = hi ? -1 : 1;
            int i1[3] = {ijk[0], ijk[1], ijk[2]}, i2[3] = {ijk[0], ijk[1], ijk[2]};
            i1[dim] = g + s; i2[dim] = g + 2 * s;
            double r = X(c, dim, g), r1 = X(c, dim, g + s), r2 = X(c, dim, g + 2 * s);
            double f1 = phi[ORC_IDX(c, i1[0], i1[1], i1[2])], f2 = phi[ORC_IDX(c, i2[0], i2[1], i2[2])];
            phi[gi] = f1 + (f2 - f1) / (r2 - r1) * (r - r1);
          } else
            phi[gi] = phi[ii];
          if (bc == 3) {   /* NON_NEGATIVE: no deeper than the wall distance */
            double d2wall = 0.5 * fabs(X(c, dim, im[dim]) - X(c, dim, g));
            phi[gi] = dmax(-d2wall, phi[gi]);
          }
        } else if (bc == 1 || bc == ORC_BC_PERIODIC) {
          phi[gi] = phi[ii];
        }
      }
  }
}

/* ------------------------------------------------------------ advection residual ------- */

void orc_ls_residual(orc_ctx *c, const double *V, const double *phi, double *R) {
  int NX = c->nx, NY = c->ny, NZ = c->nz;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < NZ; k++)
    for (int j = 0; j < NY; j++)
      for (int i = 0; i < NX; i++) {
        long me = ORC_IDX(c, i, j, k);
        double p0 = phi[me];
        double a, b, cc, d, e, f;
        if (i - 2 >= -1) {
          const double *w = c->fd_l[0] + 4 * i;
          a = w[0] * phi[ORC_IDX(c, i - 2, j, k)] + w[1] * phi[ORC_IDX(c, i - 1, j, k)] + w[2] * p0 + w[3] * phi[ORC_IDX(c, i + 1, j, k)];
        } else
          a = (p0 - phi[ORC_IDX(c, i - 1, j, k)]) / (X(c, 0, i) - X(c, 0, i - 1));
        if (i + 2 <= NX) {
          const double *w = c->fd_r[0] + 4 * i;
          b = w[0] * phi[ORC_IDX(c, i - 1, j, k)] + w[1] * p0 + w[2] * phi[ORC_IDX(c, i + 1, j, k)] + w[3] * phi[ORC_IDX(c, i + 2, j, k)];
        } else
          b = (phi[ORC_IDX(c, i + 1, j, k)] - p0) / (X(c, 0, i + 1) - X(c, 0, i));
        if (j - 2 >= -1) {
          const double *w = c->fd_l[1] + 4 * j;
          cc = w[0] * phi[ORC_IDX(c, i, j - 2, k)] + w[1] * phi[ORC_IDX(c, i, j - 1, k)] + w[2] * p0 + w[3] * phi[ORC_IDX(c, i, j + 1, k)];
        } else
          cc = (p0 - phi[ORC_IDX(c, i, j - 1, k)]) / (X(c, 1, j) - X(c, 1, j - 1));
        if (j + 2 <= NY) {
          const double *w = (i < NY - 1 && j < NY - 1) ? c->fd_r[1] + 4 * j : c->fd_r1y + 4 * j;
          d = w[0] * phi[ORC_IDX(c, i, j - 1, k)] + w[1] * p0 + w[2] * phi[ORC_IDX(c, i, j + 1, k)] + w[3] * phi[ORC_IDX(c, i, j + 2, k)];
        } else
          d = (phi[ORC_IDX(c, i, j + 1, k)] - p0) / (X(c, 1, j + 1) - X(c, 1, j));
        if (k - 2 >= -1) {
          const double *w = c->fd_l[2] + 4 * k;
          e = w[0] * phi[ORC_IDX(c, i, j, k - 2)] + w[1] * phi[ORC_IDX(c, i, j, k - 1)] + w[2] * p0 + w[3] * phi[ORC_IDX(c, i, j, k + 1)];
        } else
          e = (p0 - phi[ORC_IDX(c, i, j, k - 1)]) / (X(c, 2, k) - X(c, 2, k - 1));
        if (k + 2 <= NZ) {
          const double *w = c->fd_r[2] + 4 * k;
          f = w[0] * phi[ORC_IDX(c, i, j, k - 1)] + w[1] * p0 + w[2] * phi[ORC_IDX(c, i, j, k + 1)] + w[3] * phi[ORC_IDX(c, i, j, k + 2)];
        } else
          f = (phi[ORC_IDX(c, i, j, k + 1)] - p0) / (X(c, 2, k + 1) - X(c, 2, k));
        const double *v = V + 5 * me;
        double r = (v[1] >= 0 ? a * v[1] : b * v[1]) + (v[2] >= 0 ? cc * v[2] : d * v[2]) + (v[3] >= 0 ? e * v[3] : f * v[3]);
        R[me] = r * -1.0;
      }
}

/* x = a x + b y over the interior (LevelSetOperator.cpp:152-197) */
static void ls_axpby(orc_ctx *c, double *x, double a, double b, const double *y) {
#pragma omp parallel for schedule(static)
  for (int k = 0; k < c->nz; k++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nx; i++) {
        long me = ORC_IDX(c, i, j, k);
        x[me] = a * x[me] + b * y[me];
      }
}

/* ------------------------------------------------------------ reinitialisation --------- */

static int outside2(const orc_ctx *c, int i, int j, int k) {
  int o = (i < 0 || i >= c->nx) + (j < 0 || j >= c->ny) + (k < 0 || k >= c->nz);
  return o >= 2;
}

static double reinit_residual(orc_ctx *c, const double *phi, double *R, double cfl) {
  double max_residual = 0.0;
#pragma omp parallel for reduction(max : max_residual) schedule(static)
  for (int k = 0; k < c->nz; k++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nx; i++) {
        long me = ORC_IDX(c, i, j, k);
        if (c->tag[me] != 0) { R[me] = 0.0; continue; }
        double dx = dmin(ORC_W(c, 0, i), dmin(ORC_W(c, 1, j), ORC_W(c, 2, k)));
        double dt = cfl * dx;
        double p = phi[me];
        double a = (p - phi[ORC_IDX(c, i - 1, j, k)]) / (X(c, 0, i) - X(c, 0, i - 1));
        double b = (phi[ORC_IDX(c, i + 1, j, k)] - p) / (X(c, 0, i + 1) - X(c, 0, i));
        double cc = (p - phi[ORC_IDX(c, i, j - 1, k)]) / (X(c, 1, j) - X(c, 1, j - 1));
        double d = (phi[ORC_IDX(c, i, j + 1, k)] - p) / (X(c, 1, j + 1) - X(c, 1, j));
        double e = (p - phi[ORC_IDX(c, i, j, k - 1)]) / (X(c, 2, k) - X(c, 2, k - 1));
        double f = (phi[ORC_IDX(c, i, j, k + 1)] - p) / (X(c, 2, k + 1) - X(c, 2, k));
        double ap = dmax(a, 0.0), am = dmin(a, 0.0), bp = dmax(b, 0.0), bm = dmin(b, 0.0);
        double cp = dmax(cc, 0.0), cm = dmin(cc, 0.0), dp = dmax(d, 0.0), dm = dmin(d, 0.0);
        double ep = dmax(e, 0.0), em = dmin(e, 0.0), fp = dmax(f, 0.0), fm = dmin(f, 0.0);
        double lr;
        if (p >= 0)
          lr = sqrt(dmax(ap * ap, bm * bm) + dmax(cp * cp, dm * dm) + dmax(ep * ep, fm * fm)) - 1.0;
        else
          lr = sqrt(dmax(am * am, bp * bp) + dmax(cm * cm, dp * dp) + dmax(em * em, fp * fp)) - 1.0;
        if (fabs(lr) <= 1.0) R[me] = -dt * c->sign[me] * lr;
        else R[me] = -dt * c->sign[me] * lr / fabs(lr);
        if (fabs(lr) > max_residual) max_residual = fabs(lr);
      }
  return max_residual;
}

int orc_ls_reinitialize(orc_ctx *c, double *phi) {
  const orc_config *g = &c->cfg;
  long n = orc_ncells_ghosted(c);
  /* first-layer tags over the ghosted box without the edge/corner ghosts (:297-381) */
  int detected = 0;
  for (int k = -1; k <= c->nz; k++)
    for (int j = -1; j <= c->ny; j++)
      for (int i = -1; i <= c->nx; i++) {
        long me = ORC_IDX(c, i, j, k);
        c->tag[me] = 0;
        if (outside2(c, i, j, k)) continue;
        double p = phi[me];
        int t = 0;
        if (i - 1 >= -1 && p * phi[ORC_IDX(c, i - 1, j, k)] <= 0) t = 1;
        else if (i + 1 <= c->nx && p * phi[ORC_IDX(c, i + 1, j, k)] <= 0) t = 1;
        else if (j - 1 >= -1 && p * phi[ORC_IDX(c, i, j - 1, k)] <= 0) t = 1;
        else if (j + 1 <= c->ny && p * phi[ORC_IDX(c, i, j + 1, k)] <= 0) t = 1;
        else if (k - 1 >= -1 && p * phi[ORC_IDX(c, i, j, k - 1)] <= 0) t = 1;
        else if (k + 1 <= c->nz && p * phi[ORC_IDX(c, i, j, k + 1)] <= 0) t = 1;
        c->tag[me] = t;
        if (t && i >= 0 && i < c->nx && j >= 0 && j < c->ny && k >= 0 && k < c->nz) detected = 1;
      }
  if (!detected) return 0;
  /* sign function (:594-615), eps = 1 */
  for (int k = -1; k <= c->nz; k++)   /* every cell of the ghosted box, with its own min width */
    for (int j = -1; j <= c->ny; j++)
      for (int i = -1; i <= c->nx; i++) {
        const long q = ORC_IDX(c, i, j, k);
        const double factor = 1.0 * dmin(ORC_W(c, 0, i), dmin(ORC_W(c, 1, j), ORC_W(c, 2, k)));
        c->sign[q] = phi[q] / sqrt(phi[q] * phi[q] + factor * factor);
      }
  for (long q = 0; q < n; q++) c->phibk[q] = 0.0 * c->phibk[q] + 1.0 * phi[q];
  for (;;) {
    double residual = 0.0, dphi_max = 0.0;
    double cfl = c->reinit_cfl_state;
    int iter;
    for (iter = 0; iter < g->reinit_maxits; iter++) {
      ls_axpby(c, c->phi0, 0.0, 1.0, phi);
      residual = reinit_residual(c, phi, c->Rre, cfl);
      if (residual < g->reinit_tol) break;
      ls_axpby(c, c->rphi1, 0.0, 1.0, phi);
      ls_axpby(c, c->rphi1, 1.0, 1.0, c->Rre);
      orc_apply_bc_phi(c, c->rphi1);
      reinit_residual(c, c->rphi1, c->Rre, cfl);
      ls_axpby(c, c->rphi1, 0.25, 0.75, phi);
      ls_axpby(c, c->rphi1, 1.0, 0.25, c->Rre);
      orc_apply_bc_phi(c, c->rphi1);
      reinit_residual(c, c->rphi1, c->Rre, cfl);
      ls_axpby(c, phi, 1.0 / 3.0, 2.0 / 3.0, c->rphi1);
      ls_axpby(c, phi, 1.0, 2.0 / 3.0, c->Rre);
      orc_apply_bc_phi(c, phi);
      double err = 0.0;
      for (int k = 0; k < c->nz; k++)
        for (int j = 0; j < c->ny; j++)
          for (int i = 0; i < c->nx; i++) {
            long me = ORC_IDX(c, i, j, k);
            double le = fabs(phi[me] - c->phi0[me]);
            if (phi[me] != 0) le /= fabs(phi[me]);
            err = dmax(err, le);
          }
      dphi_max = err;
      if (dphi_max < g->reinit_tol) break;
    }
    c->n_reinit_iters += iter;
    if (iter < g->reinit_maxits) return iter;
    if (dphi_max < 0.02) return iter;
    if (cfl < 0.02) return -1;                      /* the reference exits */
    c->reinit_cfl_state = cfl / 1.5;
    for (long q = 0; q < n; q++) phi[q] = 0.0 * phi[q] + 1.0 * c->phibk[q];
  }
}

/* ------------------------------------------------------------ material IDs ------------- */

typedef struct { int i, j, k; } i3;
static int cmp_i3(const void *a, const void *b) {    /* Int3 ordering (Vector3D.h:31-35) */
  const i3 *x = (const i3 *)a, *y = (const i3 *)b;
  if (x->i != y->i) return x->i < y->i ? -1 : 1;
  if (x->j != y->j) return x->j < y->j ? -1 : 1;
  if (x->k != y->k) return x->k < y->k ? -1 : 1;
  return 0;
}

void orc_update_id(orc_ctx *c, const double *phi0, const double *phi, int *ID) {
  const orc_config *g = &c->cfg;
  long n = orc_ncells_ghosted(c);
  int *tag = c->tag;
  for (long q = 0; q < n; q++) tag[q] = 0;
  i3 *rem = NULL;
  long nrem = 0, caprem = 0;
  for (int k = 0; k < c->nz; k++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nx; i++) {
        long me = ORC_IDX(c, i, j, k);
        if (ID[me] == g->inactive_id) continue;
        if (phi[me] * phi0[me] > 0.0) continue;
        if (phi[me] < 0) ID[me] = g->ls_matid;
        else {
          tag[me] = 1;
          if (nrem == caprem) { caprem = caprem ? 2 * caprem : 256; rem = (i3 *)realloc(rem, sizeof(i3) * caprem); }
          rem[nrem].i = i; rem[nrem].j = j; rem[nrem].k = k; nrem++;
        }
      }
  qsort(rem, (size_t)nrem, sizeof(i3), cmp_i3);
  char *gone = (char *)calloc((size_t)nrem + 1, 1);
  long left = nrem;
  for (int iter = 0; iter < 100 && left > 0; iter++) {
    int progress = 0;
    for (long r = 0; r < nrem; r++) {
      if (gone[r]) continue;
      int i = rem[r].i, j = rem[r].j, k = rem[r].k;
      int best = -1;
      for (int nk = k - 1; nk <= k + 1; nk++)
        for (int nj = j - 1; nj <= j + 1; nj++)
          for (int ni = i - 1; ni <= i + 1; ni++) {
            if (ni < 0 || ni >= c->nx || nj < 0 || nj >= c->ny || nk < 0 || nk >= c->nz) continue;
            long nb = ORC_IDX(c, ni, nj, nk);
            if (tag[nb] == 1) continue;
            int nid = ID[nb];
            if (nid == g->inactive_id) continue;
            if (nid == g->ls_matid) continue;
            if (best < 0 || nid < best) best = nid;      /* std::set<int>::begin() */
          }
      if (best < 0) continue;
      long me = ORC_IDX(c, i, j, k);
      ID[me] = best;
      tag[me] = 0;
      gone[r] = 1;
      left--;
      progress++;
    }
    if (!progress) {
      for (long r = 0; r < nrem; r++)
        if (!gone[r]) ID[ORC_IDX(c, rem[r].i, rem[r].j, rem[r].k)] = 0;
      break;
    }
  }
  free(gone);
  free(rem);
  orc_apply_bc_id(c, ID);
}

/* ------------------------------------------------------------ state update ------------- */

static double norm3(double a, double b, double cc) { return sqrt(a * a + b * b + cc * cc); }

/* LocalUpdateByRiemannSolutions (MultiPhaseOperator.cpp:916-1028) */
static int local_update(orc_ctx *c, int i, int j, int k, int id, double *V, int upwind) {
  long me = ORC_IDX(c, i, j, k);
  double *v = V + 5 * me;
  int counter = 0;
  double sum_weight = 0.0;
  /* (face direction, face plus-cell, which record: 0 = Vsm (left/bottom/back), 1 = Vsp,
   *  neighbour, velocity component tested, sign) */
  struct { int dir; long face; int rec; long nb; int comp; double sgn; } L[6] = {
    {0, ORC_IDX(c, i, j, k), 0, ORC_IDX(c, i - 1, j, k), 1, 1.0},
    {0, ORC_IDX(c, i + 1, j, k), 1, ORC_IDX(c, i + 1, j, k), 1, -1.0},
    {1, ORC_IDX(c, i, j, k), 0, ORC_IDX(c, i, j - 1, k), 2, 1.0},
    {1, ORC_IDX(c, i, j + 1, k), 1, ORC_IDX(c, i, j + 1, k), 2, -1.0},
    {2, ORC_IDX(c, i, j, k), 0, ORC_IDX(c, i, j, k - 1), 3, 1.0},
    {2, ORC_IDX(c, i, j, k + 1), 1, ORC_IDX(c, i, j, k + 1), 3, -1.0}};
  for (int s = 0; s < 6; s++) {
    const int *q = c->rsId[L[s].dir] + 3 * L[s].face;
    if (q[2] != c->rs_stamp) continue;
    int rid = q[L[s].rec];
    const double *vn = V + 5 * L[s].nb;
    if (rid != id) continue;
    if (upwind && !(L[s].sgn > 0 ? vn[L[s].comp] > 0 : vn[L[s].comp] < 0)) continue;
    double w;
    if (upwind) {
      double nrm = norm3(vn[1], vn[2], vn[3]);
      w = (L[s].sgn > 0 ? vn[L[s].comp] : -vn[L[s].comp]) / nrm;
    } else w = 1.0;
    sum_weight += w;
    const double *rs = c->rsV[L[s].dir] + 10 * L[s].face + 5 * L[s].rec;
    if (counter == 0) for (int p = 0; p < 5; p++) v[p] = w * rs[p];
    else for (int p = 0; p < 5; p++) v[p] += w * rs[p];
    counter++;
  }
  if (sum_weight > 0.0) {
    double inv = 1.0 / sum_weight;
    for (int p = 0; p < 5; p++) v[p] *= inv;
  }
  return counter;
}

/* FixUnresolvedNodes (MultiPhaseOperator.cpp:1156-1320) with the failsafe density (default On);
 * `reset` persists across nodes as in the reference. */
static void fix_unresolved(orc_ctx *c, const i3 *un, long nun, const int *ID, double *V) {
  int reset = 0;
  for (long u = 0; u < nun; u++) {
    int i = un[u].i, j = un[u].j, k = un[u].k;
    long me = ORC_IDX(c, i, j, k);
    double *v = V + 5 * me;
    double x0[3] = {X(c, 0, i), X(c, 1, j), X(c, 2, k)};
    double sum_weight = 0.0, sum_weight2 = 0.0, vtmp[5] = {0, 0, 0, 0, 0};
    for (int nk = k - 1; nk <= k + 1; nk++)
      for (int nj = j - 1; nj <= j + 1; nj++)
        for (int ni = i - 1; ni <= i + 1; ni++) {
          if (ni < 0 || ni >= c->nx || nj < 0 || nj >= c->ny || nk < 0 || nk >= c->nz) continue;
          long nb = ORC_IDX(c, ni, nj, nk);
          if (ID[nb] != ID[me]) continue;
          if (nb == me) continue;
          int nu = 0;
          for (long w2 = 0; w2 < nun; w2++)
            if (un[w2].i == ni && un[w2].j == nj && un[w2].k == nk) { nu = 1; break; }
          if (nu) continue;
          const double *vn = V + 5 * nb;
          double v1[3] = {vn[1], vn[2], vn[3]};
          double v1n = norm3(v1[0], v1[1], v1[2]);
          if (v1n != 0) { double iv = 1.0 / v1n; v1[0] *= iv; v1[1] *= iv; v1[2] *= iv; }
          double d[3] = {x0[0] - X(c, 0, ni), x0[1] - X(c, 1, nj), x0[2] - X(c, 2, nk)};
          double dn = norm3(d[0], d[1], d[2]);
          double idn = 1.0 / dn;
          d[0] *= idn; d[1] *= idn; d[2] *= idn;
          double dot = d[0] * v1[0] + d[1] * v1[1] + d[2] * v1[2];
          double weight = dmax(0.0, dot);
          if (weight > 0) {
            sum_weight += weight;
            if (reset) for (int p = 0; p < 5; p++) v[p] += weight * vn[p];
            else { for (int p = 0; p < 5; p++) v[p] = weight * vn[p]; reset = 1; }
          }
          for (int p = 0; p < 5; p++) vtmp[p] += 1.0 / dn * vn[p];
          sum_weight2 += 1.0 / dn;
        }
    if (sum_weight > 0) {
      double inv = 1.0 / sum_weight;
      for (int p = 0; p < 5; p++) v[p] *= inv;
      continue;
    }
    if (sum_weight2 > 0) {
      for (int p = 0; p < 5; p++) v[p] = vtmp[p] / sum_weight2;
      continue;
    }
    double density = 0.0;
    for (int layer = 1; layer <= 10; layer++) {
      for (int nk = k - layer; nk <= k + layer; nk++)
        for (int nj = j - layer; nj <= j + layer; nj++)
          for (int ni = i - layer; ni <= i + layer; ni++) {
            if (ni < 0 || ni >= c->nx || nj < 0 || nj >= c->ny || nk < 0 || nk >= c->nz) continue;
            long nb = ORC_IDX(c, ni, nj, nk);
            if (ID[nb] != ID[me]) continue;
            if (nb == me) continue;
            int nu = 0;
            for (long w2 = 0; w2 < nun; w2++)
              if (un[w2].i == ni && un[w2].j == nj && un[w2].k == nk) { nu = 1; break; }
            if (nu) continue;
            double d[3] = {X(c, 0, ni) - x0[0], X(c, 1, nj) - x0[1], X(c, 2, nk) - x0[2]};
            double dist = norm3(d[0], d[1], d[2]);
            sum_weight += 1.0 / dist;
            density += (1.0 / dist) * V[5 * nb];
          }
      if (sum_weight > 0) { v[0] = density / sum_weight; break; }
    }
    if (sum_weight == 0) v[0] = c->cfg.mat[ID[me]].failsafe_density;
  }
}

static void update_states(orc_ctx *c, const int *IDn, const int *ID, double *V) {
  i3 *un = NULL;
  long nun = 0, cap = 0;
  for (int k = 0; k < c->nz; k++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nx; i++) {
        long me = ORC_IDX(c, i, j, k);
        if (ID[me] == IDn[me]) continue;
        if (ID[me] == c->cfg.inactive_id) continue;
        int cnt = local_update(c, i, j, k, ID[me], V, 1);
        if (cnt == 0) cnt = local_update(c, i, j, k, ID[me], V, 0);
        if (cnt == 0) {
          if (nun == cap) { cap = cap ? 2 * cap : 64; un = (i3 *)realloc(un, sizeof(i3) * cap); }
          un[nun].i = i; un[nun].j = j; un[nun].k = k; nun++;
        }
      }
  c->n_unresolved += nun;
  if (nun) fix_unresolved(c, un, nun, ID, V);
  free(un);
}

/* The interface Riemann solutions of V stored as the first stage of time step `stamp` stores them
 * (SpaceOperator.cpp:2518-2523): the residual of V is computed with the store on. Returns the
 * residual's status. */
int orc_store_riemann(orc_ctx *c, const double *V, const int *ID, int stamp) {
  c->rs_stamp = stamp;
  c->store_riemann = 1;
  int st = orc_compute_residual(c, V, ID, c->R);
  c->store_riemann = 0;
  return st;
}

/* UpdateStateVariablesByRiemannSolutions + FixUnresolvedNodes (MultiPhaseOperator.cpp:857-911,
 * 1156-1320) from the solutions of the last orc_store_riemann call, for the material IDs IDn (before)
 * and ID (after). Returns the number of unresolved cells. */
long orc_update_states(orc_ctx *c, const int *IDn, const int *ID, double *V) {
  long before = c->n_unresolved;
  update_states(c, IDn, ID, V);
  return c->n_unresolved - before;
}

/* Multiphase counters since creation: reinitialisation pseudo-time iterations (retries included),
 * unresolved cells of the state update, and the reinitialisation CFL number now in force. */
void orc_mp_counters(const orc_ctx *c, long *reinit_iters, long *unresolved, double *reinit_cfl) {
  *reinit_iters = c->n_reinit_iters;
  *unresolved = c->n_unresolved;
  *reinit_cfl = c->reinit_cfl_state;
}

/* ------------------------------------------------------------ time step ---------------- */

void orc_set_levelset(orc_ctx *c, const double *phi) { c->phi_n = phi; }

int orc_advance(orc_ctx *c, double *V, const int *ID, double dt) {
  return orc_step_all(c, V, (int *)ID, NULL, dt, 0);
}

static int ls_stage(orc_ctx *c, const double *Vres, double *phi_io, double *phi_s, double a, double b, double cdt,
                    int first, double *dst) {
  /* dst = a*dst + b*phi (or copy), dst += cdt*R(Vres, phi_s) -- LevelSetOperator::AXPlusBY order */
  orc_ls_residual(c, Vres, phi_s, c->Rls);
  if (first) {
    if (dst != phi_io) ls_axpby(c, dst, 0.0, 1.0, phi_io);
  } else {
    ls_axpby(c, dst, a, b, phi_io);
  }
  ls_axpby(c, dst, 1.0, cdt, c->Rls);
  orc_apply_bc_phi(c, dst);
  return 0;
}

static int step_body(orc_ctx *c, double *V, int *ID, double *phi, double dt, int step, int ls);

int orc_step_all(orc_ctx *c, double *V, int *ID, double *phi, double dt, int step) {
  long n = orc_ncells_ghosted(c);
  int ls = c->cfg.ls_on && phi;
  if (ls) memcpy(c->phi_tmp, phi, (size_t)n * sizeof(double));
  const double *phi_saved = c->phi_n;
  c->phi_n = ls ? phi : NULL;   /* every stage's interfaces use Phi(n) (TimeIntegrator.cpp:375, 581, 637) */
  int rc = step_body(c, V, ID, phi, dt, step, ls);
  c->phi_n = phi_saved;
  return rc;
}

static int step_body(orc_ctx *c, double *V, int *ID, double *phi, double dt, int step, int ls) {
  long n = orc_ncells_ghosted(c);
  const orc_config *g = &c->cfg;
  double *R = c->R, *Un = c->Un, *U1 = c->U1, *V1 = c->V1, *V2 = c->V2;
  c->rs_stamp = step;
  c->store_riemann = 1;
  int st = orc_compute_residual(c, V, ID, R);
  c->store_riemann = 0;
  if (st < 0) return -1;
  orc_p2c_all(c, V, ID, Un);
  if (g->ts == ORC_TS_FE) {
    orc_axpby(c, Un, 1.0, dt, R);
    if (orc_c2p_clip(c, Un, ID, V) < 0) return -1;
    orc_apply_bc(c, V);
    if (ls) {
      orc_ls_residual(c, V, phi, c->Rls);
      ls_axpby(c, phi, 1.0, dt, c->Rls);
      orc_apply_bc_phi(c, phi);
    }
  } else {
    memset(U1, 0, (size_t)n * 5 * sizeof(double));
    orc_axpby(c, U1, 0.0, 1.0, Un);
    orc_axpby(c, U1, 1.0, dt, R);
    memcpy(V1, V, (size_t)n * 5 * sizeof(double));
    long cl = orc_c2p_clip(c, U1, ID, V1);
    if (cl < 0) return -1;
    if (cl) orc_p2c_all(c, V1, ID, U1);
    orc_apply_bc(c, V1);
    if (ls) ls_stage(c, V, phi, phi, 0, 0, dt, 1, c->phi1);           /* Phi1 = Phi + dt R(Phi; V_n) */
    st = orc_compute_residual(c, V1, ID, R);
    if (st < 0) return -1;
    if (g->ts == ORC_TS_RK2) {
      orc_axpby(c, U1, 0.5, 0.5, Un);
      orc_axpby(c, U1, 1.0, 0.5 * dt, R);
      if (orc_c2p_clip(c, U1, ID, V) < 0) return -1;
      orc_apply_bc(c, V);
      if (ls) {
        orc_ls_residual(c, V1, c->phi1, c->Rls);
        ls_axpby(c, phi, 0.5, 0.5, c->phi1);
        ls_axpby(c, phi, 1.0, 0.5 * dt, c->Rls);
        orc_apply_bc_phi(c, phi);
      }
    } else {
      orc_axpby(c, U1, 0.25, 0.75, Un);
      orc_axpby(c, U1, 1.0, 0.25 * dt, R);
      memcpy(V2, V1, (size_t)n * 5 * sizeof(double));
      cl = orc_c2p_clip(c, U1, ID, V2);
      if (cl < 0) return -1;
      if (cl) orc_p2c_all(c, V2, ID, U1);
      orc_apply_bc(c, V2);
      if (ls) {
        orc_ls_residual(c, V1, c->phi1, c->Rls);
        ls_axpby(c, c->phi1, 0.25, 0.75, phi);
        ls_axpby(c, c->phi1, 1.0, 0.25 * dt, c->Rls);
        orc_apply_bc_phi(c, c->phi1);
      }
      st = orc_compute_residual(c, V2, ID, R);
      if (st < 0) return -1;
      orc_axpby(c, U1, 2.0 / 3.0, 1.0 / 3.0, Un);
      orc_axpby(c, U1, 1.0, 2.0 / 3.0 * dt, R);
      if (orc_c2p_clip(c, U1, ID, V) < 0) return -1;
      orc_apply_bc(c, V);
      if (ls) {
        orc_ls_residual(c, V2, c->phi1, c->Rls);
        ls_axpby(c, phi, 1.0 / 3.0, 2.0 / 3.0, c->phi1);
        ls_axpby(c, phi, 1.0, 2.0 / 3.0 * dt, c->Rls);
        orc_apply_bc_phi(c, phi);
      }
    }
  }
  if (!ls) return 0;
  /* end-of-step tasks (TimeIntegrator.cpp:695-797) */
  memcpy(c->IDn, ID, (size_t)n * sizeof(int));
  orc_update_id(c, c->phi_tmp, phi, ID);
  if (g->reinit_freq > 0 && step % g->reinit_freq == 0)
    if (orc_ls_reinitialize(c, phi) < 0) return -1;
  update_states(c, c->IDn, ID, V);
  if (orc_c2p_clip(c, NULL, ID, V) < 0) return -1;
  orc_apply_bc(c, V);
  return 0;
}

int orc_run(orc_ctx *c, double *V, int *ID, double *phi, double tmax, int maxsteps, double *t_out, int *steps_out) {
  /* a later call continues the time and the step numbering (the reinitialisation schedule and the
   * Riemann-solution stamps count from the first step of the first call) */
  double t = c->t_done;
  int n = 0, rc = 0;
  while (t < tmax && n < maxsteps) {
    n++;
    const int step = c->steps_done + n;
    double dt;
    if (orc_time_step(c, V, ID, &dt) < 0) { rc = -1; break; }
    if (t + dt > tmax) dt = tmax - t;
    t += dt;
    if (orc_step_all(c, V, ID, phi, dt, step) < 0) { rc = -1; break; }
  }
  c->steps_done += n;
  c->t_done = t;
  *t_out = t;
  *steps_out = n;
  return rc;
}

int orc_riemann(const orc_ctx *c, const double *dir, const double *Vm, int idm, const double *Vp, int idp,
                double *Vs, int *id, double *Vsm, double *Vsp) {
  return orc_riemann_solve(c->cfg.mat, &c->cfg.riemann, dir, Vm, idm, Vp, idp, Vs, id, Vsm, Vsp);
}

int orc_riemann_onesided(const orc_ctx *c, const double *dir, const double *Vm, int idl, const double *Ustar,
                         double *Vs, int *id, double *Vsm) {
  return orc_riemann_onesided_solve(c->cfg.mat, &c->cfg.riemann, dir, Vm, idl, Ustar, Vs, id, Vsm);
}
