/*
 * oracle/fv.c -- CPU restatement of M2C's explicit finite-volume update (test infrastructure).
 *
 * Functions follow, in order:
 *   mesh + MUSCL coefficients   MeshGenerator.cpp:146-158, SpaceOperator.cpp:129-183,
 *                               Reconstructor.cpp:53-121 (A, B, k of Zeng 2016)
 *   ghost fill                  SpaceOperator.cpp:428-516 (ghost list), :1667-1745 (BCs),
 *                               MultiPhaseOperator.cpp:796-819 (ghost IDs)
 *   reconstruction              Reconstructor.cpp:332-904, limiters Reconstructor.h:95-110
 *   reconstructed-state check   SpaceOperator.cpp:3198-3347
 *   face loop + accumulation    SpaceOperator.cpp:2323-2856, residual :3351-3409
 *   time-step size              SpaceOperator.cpp:2128-2227
 *   Runge-Kutta stages          TimeIntegrator.cpp:150-226, 291-428, 496-690
 *   time loop                   Main.cpp:514-621
 * The residual sums each cell's six face contributions in the reference's loop order
 * (west, south, back subtracted, then east, north, front added) and then multiplies by
 * 1/(-vol) as Vec5D::operator/= does (Vector5D.h:174-185). The result does not depend on the
 * number of OpenMP threads.
 * The periodic boundary code is an extension with no reference counterpart; it is pinned only
 * by the tests' invariants (translation invariance, conservation).
 */
#include "m2c_oracle.h"
#include "flux.h"
#include "oracle_internal.h"
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

long orc_ncells_ghosted(const orc_ctx *c) { return (long)c->gx * c->gy * c->gz; }

static int slope_k(double A, double B) {          /* Reconstructor.cpp:103-121 */
  int k;
  double rhs;
  for (k = 2; k < 1000; k++) {
    rhs = 2.0 / (1.0 + pow((k - 1.0) / k, k - 1) / k) * (A < 1.0 ? A : 1.0);
    if (B <= rhs) break;
  }
  return k;
}

orc_ctx *orc_create(const orc_config *cfg) {
  orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
  c->cfg = *cfg;
  for (int m = 0; m < cfg->nmat; m++) orc_eos_finalize(&c->cfg.mat[m]);
  c->nx = cfg->nx; c->ny = cfg->ny; c->nz = cfg->nz;
  c->gx = c->nx + 2; c->gy = c->ny + 2; c->gz = c->nz + 2;
  c->dx = (cfg->x1 - cfg->x0) / cfg->nx;
  c->dy = (cfg->y1 - cfg->y0) / cfg->ny;
  c->dz = (cfg->z1 - cfg->z0) / cfg->nz;
  /* the mesh (MeshGenerator.cpp:12-131): widths and centres per axis; ghosts take the boundary
   * cell's width (SpaceOperator.cpp:219-300); limiter coefficients per cell from the widths of the
   * cell and its neighbours (Reconstructor.cpp:53-121) */
  {
    const double lo[3] = {cfg->x0, cfg->y0, cfg->z0}, hi[3] = {cfg->x1, cfg->y1, cfg->z1};
    const int N[3] = {cfg->nx, cfg->ny, cfg->nz};
    for (int a = 0; a < 3; a++) {
      const int n = N[a];
      c->w[a] = (double *)calloc((size_t)n + 2, sizeof(double));
      c->mx[a] = (double *)calloc((size_t)n, sizeof(double));
      c->cA[a] = (double *)calloc((size_t)n, sizeof(double));
      c->cB[a] = (double *)calloc((size_t)n, sizeof(double));
      c->cK[a] = (int *)calloc((size_t)n, sizeof(int));
      int m = 0;
      if (orc_mesh_1d(lo[a], hi[a], n, cfg->mesh_npts[a], &cfg->mesh_pts[a][0][0], n, c->mx[a], c->w[a] + 1, &m) ||
          m != n)
        return NULL;
      c->w[a][0] = c->w[a][1];
      c->w[a][n + 1] = c->w[a][n];
      for (int i = 0; i < n; i++) {
        const double wm = ORC_W(c, a, i - 1), w0 = ORC_W(c, a, i), wp = ORC_W(c, a, i + 1);
        c->cA[a][i] = (wm + w0) / (w0 + wp);
        c->cB[a][i] = 2.0 * w0 / (w0 + wp);
        c->cK[a][i] = (cfg->limiter == ORC_LIM_VANALBADA) ? slope_k(c->cA[a][i], c->cB[a][i]) : 0;
      }
    }
  }
  long n = orc_ncells_ghosted(c);
  for (int f = 0; f < 6; f++) c->face[f] = (double *)calloc(n * 5, sizeof(double));
  c->Un = (double *)calloc(n * 5, sizeof(double));
  c->U1 = (double *)calloc(n * 5, sizeof(double));
  c->R = (double *)calloc(n * 5, sizeof(double));
  c->V1 = (double *)calloc(n * 5, sizeof(double));
  c->V2 = (double *)calloc(n * 5, sizeof(double));
  long nfx = (long)(c->nx + 1) * c->ny * c->nz, nfy = (long)c->nx * (c->ny + 1) * c->nz,
       nfz = (long)c->nx * c->ny * (c->nz + 1);
  c->Fx = (double *)calloc(nfx * 10, sizeof(double));
  c->Fy = (double *)calloc(nfy * 10, sizeof(double));
  c->Fz = (double *)calloc(nfz * 10, sizeof(double));
  c->ferr = (int *)calloc((size_t)(c->nz + 1), sizeof(int));
#ifdef _OPENMP
  if (cfg->nthreads > 0) omp_set_num_threads(cfg->nthreads);
#endif
  orc_ls_init(c);
  return c;
}

void orc_destroy(orc_ctx *c) {
  if (!c) return;
  for (int f = 0; f < 6; f++) free(c->face[f]);
  free(c->Un); free(c->U1); free(c->R); free(c->V1); free(c->V2);
  free(c->Fx); free(c->Fy); free(c->Fz); free(c->ferr);
  for (int d = 0; d < 3; d++) { free(c->Dv[d]); free(c->Dh[d]); free(c->eb_rec[d]); }
  for (int d = 0; d < 3; d++) { free(c->w[d]); free(c->mx[d]); free(c->cA[d]); free(c->cB[d]); free(c->cK[d]); }
  free(c->eb_wall);
  orc_ls_free(c);
  free(c);
}

void orc_counters(const orc_ctx *c, long *rev, long *clip, long *rerr) {
  if (rev) *rev = c->n_reverted;
  if (clip) *clip = c->n_clipped;
  if (rerr) *rerr = c->n_riemann_err;
}

int orc_set_embedded_faces(orc_ctx *c, int n, const int *face, const double *wall, int normal_mode, int recon) {
  const long nf = (long)(c->nx + 1) * (c->ny + 1) * (c->nz + 1);
  for (int d = 0; d < 3; d++) { free(c->eb_rec[d]); c->eb_rec[d] = NULL; }
  free(c->eb_wall);
  c->eb_wall = NULL;
  c->eb_n = 0;
  if (n <= 0) return 0;
  if (normal_mode < 0 || normal_mode > 2 || recon < 0 || recon > 1) return -1;
  for (int d = 0; d < 3; d++) {
    c->eb_rec[d] = (int *)malloc(sizeof(int) * nf);
    for (long f = 0; f < nf; f++) c->eb_rec[d][f] = -1;
  }
  c->eb_wall = (double *)malloc(sizeof(double) * 12 * n);
  memcpy(c->eb_wall, wall, sizeof(double) * 12 * n);
  const int N[3] = {c->nx, c->ny, c->nz};
  for (int t = 0; t < n; t++) {
    const int d = face[4 * t], q[3] = {face[4 * t + 1], face[4 * t + 2], face[4 * t + 3]};
    int ok = d >= 0 && d <= 2;
    for (int a = 0; ok && a < 3; a++)
      ok = a == d ? (q[a] >= 1 && q[a] <= N[a] - 1) : (q[a] >= 0 && q[a] < N[a]);
    if (!ok || c->eb_rec[d][ORC_FKEY(c, q[0], q[1], q[2])] >= 0) {
      orc_set_embedded_faces(c, 0, NULL, NULL, 0, 0);
      return -1;
    }
    c->eb_rec[d][ORC_FKEY(c, q[0], q[1], q[2])] = t;
  }
  c->eb_n = n;
  c->eb_normal_mode = normal_mode;
  c->eb_recon = recon;
  return 0;
}

/* the wall record of face (d, plus cell i,j,k), or NULL */
static inline const double *eb_face(const orc_ctx *c, int d, int i, int j, int k) {
  if (!c->eb_n || i < 0 || j < 0 || k < 0 || i > c->nx || j > c->ny || k > c->nz) return NULL;
  const int r = c->eb_rec[d][ORC_FKEY(c, i, j, k)];
  return r >= 0 ? c->eb_wall + 12 * r : NULL;
}

void orc_numerical_flux(const orc_ctx *c, int dir, const double *Vm, const double *Vp, int id, double *F) {
  const orc_eos *m = &c->cfg.mat[id];
  switch (c->cfg.flux) {
  case ORC_FLUX_ROE: orc_flux_roe(m, dir, Vm, Vp, c->cfg.roe_delta, F); break;
  case ORC_FLUX_LLF: orc_flux_llf2(m, m, dir, Vm, Vp, F); break;
  case ORC_FLUX_GODUNOV: orc_flux_godunov(c, dir, Vm, Vp, id, F); break;
  default: orc_flux_hllc(m, dir, Vm, Vp, F); break;
  }
}

/* ------------------------------------------------------------------ ghost fill ---------- */

/* image index of a ghost at -1 / n (SpaceOperator.cpp:449-480), periodic wraps */
static inline int image_of(int i, int n, int periodic) {
  if (i < 0) return periodic ? n + i : -i - 1;
  if (i >= n) return periodic ? i - n : n + (i - n) - 1;
  return i;
}

void orc_apply_bc(const orc_ctx *c, double *V) {
  const orc_config *g = &c->cfg;
  int nx = c->nx, ny = c->ny, nz = c->nz;
  for (int side = 0; side < 6; side++) {
    int dim = side / 2, bc = g->bc[side];
    int n0 = dim == 0 ? ny : nx, n1 = dim == 2 ? ny : nz;
    for (int b = 0; b < n1; b++)
      for (int a = 0; a < n0; a++) {
        int i, j, k;
        int gidx = (side % 2 == 0) ? -1 : (dim == 0 ? nx : dim == 1 ? ny : nz);
        if (dim == 0) { i = gidx; j = a; k = b; }
        else if (dim == 1) { i = a; j = gidx; k = b; }
        else { i = a; j = b; k = gidx; }
        double *v = V + 5 * ORC_IDX(c, i, j, k);
        int ii = i, jj = j, kk = k;
        if (dim == 0) ii = image_of(i, nx, bc == ORC_BC_PERIODIC);
        if (dim == 1) jj = image_of(j, ny, bc == ORC_BC_PERIODIC);
        if (dim == 2) kk = image_of(k, nz, bc == ORC_BC_PERIODIC);
        const double *im = V + 5 * ORC_IDX(c, ii, jj, kk);
        switch (bc) {
        case ORC_BC_INLET: for (int p = 0; p < 5; p++) v[p] = g->inlet[p]; break;
        case ORC_BC_OUTLET: for (int p = 0; p < 5; p++) v[p] = g->outlet[p]; break;
        case ORC_BC_SLIPWALL: case ORC_BC_SYMMETRY:
          for (int p = 0; p < 5; p++) v[p] = im[p];
          v[dim + 1] *= -1.0;
          break;
        case ORC_BC_STICKWALL:
          v[0] = im[0]; v[1] = -1.0 * im[1]; v[2] = -1.0 * im[2]; v[3] = -1.0 * im[3]; v[4] = im[4];
          break;
        case ORC_BC_PERIODIC: for (int p = 0; p < 5; p++) v[p] = im[p]; break;
        default: break;
        }
      }
  }
}

void orc_apply_bc_id(const orc_ctx *c, int *ID) {
  int nx = c->nx, ny = c->ny, nz = c->nz;
  for (int side = 0; side < 6; side++) {
    int dim = side / 2, per = c->cfg.bc[side] == ORC_BC_PERIODIC;
    int n0 = dim == 0 ? ny : nx, n1 = dim == 2 ? ny : nz;
    for (int b = 0; b < n1; b++)
      for (int a = 0; a < n0; a++) {
        int i, j, k;
        int gidx = (side % 2 == 0) ? -1 : (dim == 0 ? nx : dim == 1 ? ny : nz);
        if (dim == 0) { i = gidx; j = a; k = b; }
        else if (dim == 1) { i = a; j = gidx; k = b; }
        else { i = a; j = b; k = gidx; }
        int ii = dim == 0 ? image_of(i, nx, per) : i;
        int jj = dim == 1 ? image_of(j, ny, per) : j;
        int kk = dim == 2 ? image_of(k, nz, per) : k;
        ID[ORC_IDX(c, i, j, k)] = ID[ORC_IDX(c, ii, jj, kk)];
      }
  }
}

/* ------------------------------------------------------------------ reconstruction ------ */

static inline double gen_minmod(double A, double B, double alpha, double dq0, double dq1) {
  /* Reconstructor.h:95-102 */
  if (dq0 < 0.0 && dq1 < 0.0) {
    double a = alpha * dq0, b = B / (A + 1.0) * (dq0 + dq1), cc = alpha * dq1;
    double m = (b < cc) ? cc : b;            /* std::max(b, cc) */
    return (a < m) ? m : a;                  /* std::max(a, m)  */
  } else if (dq0 > 0.0 && dq1 > 0.0) {
    double a = alpha * dq0, b = B / (A + 1.0) * (dq0 + dq1), cc = alpha * dq1;
    double m = (cc < b) ? cc : b;            /* std::min(b, cc) */
    return (m < a) ? m : a;                  /* std::min(a, m)  */
  }
  return 0.0;
}

static inline double van_albada(double A, double B, int k, double dq0, double dq1, int use_pow) {
  /* Reconstructor.h:104-110 */
  if (dq0 * dq1 <= 0) return 0.0;
  if (fabs(dq0 * dq1) < 1e-16) return 0.0;
  double d0k, d1k;
  if (use_pow) { d0k = pow(dq0, k); d1k = pow(dq1, k); }
  else {
    d0k = dq0; d1k = dq1;
    for (int t = 1; t < k; t++) { d0k *= dq0; d1k *= dq1; }
  }
  return B * (dq1 * d0k + dq0 * d1k) / (d0k + A * d1k);
}

/* Reconstruct one interior cell: Reconstructor.cpp:423-779 (primitive variables). */
static void recon_cell(orc_ctx *c, const double *V, const int *ID, int i, int j, int k, long *nrev) {
  const orc_config *g = &c->cfg;
  long me = ORC_IDX(c, i, j, k);
  const double *v = V + 5 * me;
  double *out[6];
  for (int f = 0; f < 6; f++) out[f] = c->face[f] + 5 * me;
  int myid = ID[me];
  /* SpaceOperator::TagNodesOutsideConRecDepth (SpaceOperator.cpp:2861-2920): |phi(t_n)| below the
   * depth selects constant reconstruction (Reconstructor.cpp:434-447) */
  const int conrec = g->conrec_depth > 0 && c->phi_n && fabs(c->phi_n[me]) < g->conrec_depth;
  if (g->rec_type == ORC_REC_CONSTANT || myid == g->inactive_id || conrec) {
    for (int f = 0; f < 6; f++) memcpy(out[f], v, 5 * sizeof(double));
    return;
  }
  long nb[6] = {ORC_IDX(c, i - 1, j, k), ORC_IDX(c, i + 1, j, k), ORC_IDX(c, i, j - 1, k),
                ORC_IDX(c, i, j + 1, k), ORC_IDX(c, i, j, k - 1), ORC_IDX(c, i, j, k + 1)};
  double sig[3][5];
  const int ijk[3] = {i, j, k}, nn[3] = {c->nx, c->ny, c->nz};
  for (int d = 0; d < 3; d++) {
    long lo = nb[2 * d], hi = nb[2 * d + 1];
    const int q = ijk[d] < 0 ? 0 : (ijk[d] >= nn[d] ? nn[d] - 1 : ijk[d]);
    const double cA = c->cA[d][q], cB = c->cB[d][q];
    const int cK = c->cK[d][q];
    for (int p = 0; p < 5; p++) {
      double dq0 = (ID[lo] == g->inactive_id) ? 0.0 : v[p] - V[5 * lo + p];
      double dq1 = (ID[hi] == g->inactive_id) ? 0.0 : V[5 * hi + p] - v[p];
      double s;
      switch (g->limiter) {
      case ORC_LIM_GENMINMOD: s = gen_minmod(cA, cB, g->gminmod_alpha, dq0, dq1); break;
      case ORC_LIM_VANALBADA: s = van_albada(cA, cB, cK, dq0, dq1, g->vanalbada_pow); break;
      default: s = 0.5 * (dq0 + dq1); break;
      }
      sig[d][p] = s;
    }
    if (g->slope_zero_near_interface && (myid != ID[lo] || myid != ID[hi]))
      for (int p = 0; p < 5; p++) sig[d][p] = 0.0;
    /* Reconstructor.cpp:656-664: also where the cell's low or high face along d is cut by an
     * embedded wall */
    if (g->slope_zero_near_interface && c->eb_n &&
        (eb_face(c, d, i, j, k) || eb_face(c, d, i + (d == 0), j + (d == 1), k + (d == 2))))
      for (int p = 0; p < 5; p++) sig[d][p] = 0.0;
  }
  for (int p = 0; p < 5; p++) {
    out[0][p] = v[p] - 0.5 * sig[0][p];
    out[1][p] = v[p] + 0.5 * sig[0][p];
    out[2][p] = v[p] - 0.5 * sig[1][p];
    out[3][p] = v[p] + 0.5 * sig[1][p];
    out[4][p] = v[p] - 0.5 * sig[2][p];
    out[5][p] = v[p] + 0.5 * sig[2][p];
  }
  /* Reconstructor.cpp:719-775: the first face that fails CheckState reverts to the cell value */
  const orc_eos *m = &g->mat[myid];
  for (int f = 0; f < 6; f++)
    if (orc_check_state(m, out[f])) {
      memcpy(out[f], v, 5 * sizeof(double));
      (*nrev)++;
      break;
    }
}

/* Ghost face states (Reconstructor.cpp:795-875); periodic faces copy the wrapped cell's face. */
static void recon_ghost_faces(orc_ctx *c, const double *V) {
  const orc_config *g = &c->cfg;
  int n[3] = {c->nx, c->ny, c->nz};
  for (int side = 0; side < 6; side++) {
    int dim = side / 2, hi = side % 2, bc = g->bc[side];
    int n0 = dim == 0 ? c->ny : c->nx, n1 = dim == 2 ? c->ny : c->nz;
    int gface = hi ? 2 * dim : 2 * dim + 1;     /* face of the ghost pointing into the domain */
    int iface = hi ? 2 * dim + 1 : 2 * dim;     /* matching face of the image cell */
    for (int b = 0; b < n1; b++)
      for (int a = 0; a < n0; a++) {
        int ijk[3], im[3];
        if (dim == 0) { ijk[1] = a; ijk[2] = b; }
        else if (dim == 1) { ijk[0] = a; ijk[2] = b; }
        else { ijk[0] = a; ijk[1] = b; }
        ijk[dim] = hi ? n[dim] : -1;
        im[0] = ijk[0]; im[1] = ijk[1]; im[2] = ijk[2];
        im[dim] = image_of(ijk[dim], n[dim], bc == ORC_BC_PERIODIC);
        long gi = ORC_IDX(c, ijk[0], ijk[1], ijk[2]), ii = ORC_IDX(c, im[0], im[1], im[2]);
        double *dst = c->face[gface] + 5 * gi;
        const double *src = c->face[iface] + 5 * ii;
        switch (bc) {
        case ORC_BC_INLET: case ORC_BC_OUTLET: case ORC_BC_OVERSET:
          memcpy(dst, V + 5 * gi, 5 * sizeof(double)); break;
        case ORC_BC_SYMMETRY: case ORC_BC_SLIPWALL:
          for (int p = 0; p < 5; p++) dst[p] = (p == dim + 1) ? -src[p] : src[p];
          break;
        case ORC_BC_STICKWALL:
          for (int p = 0; p < 5; p++) dst[p] = (p >= 1 && p <= 3) ? -src[p] : src[p];
          break;
        case ORC_BC_PERIODIC: {
          /* the periodic image of ghost -1 is cell n-1; its face toward the domain is its
           * high face (and vice versa) */
          const double *ps = c->face[gface] + 5 * ii;
          memcpy(dst, ps, 5 * sizeof(double));
          break;
        }
        default: break;
        }
      }
  }
}

/* SpaceOperator.cpp:3198-3347. Returns 1 on a hyperbolicity violation. */
static int check_reconstructed(orc_ctx *c, const double *V, const int *ID, long *nclip) {
  const orc_config *g = &c->cfg;
  int err = 0;
  long clipped = 0;
#pragma omp parallel for reduction(+ : clipped) reduction(| : err) schedule(static)
  for (int k = -1; k <= c->nz; k++)
    for (int j = -1; j <= c->ny; j++)
      for (int i = -1; i <= c->nx; i++) {
        int boundary = (k == -1 || k == c->nz) + (j == -1 || j == c->ny) + (i == -1 || i == c->nx);
        if (boundary >= 2) continue;
        long me = ORC_IDX(c, i, j, k);
        int myid = ID[me];
        if (myid == g->inactive_id) continue;
        const orc_eos *m = &g->mat[myid];
        const double *v = V + 5 * me;
        if (boundary == 0) {
          for (int f = 0; f < 6; f++) {
            double *fs = c->face[f] + 5 * me;
            if (orc_clip(m, fs)) { memcpy(fs, v, 5 * sizeof(double)); clipped++; }
          }
          for (int f = 0; f < 6; f++) err |= orc_check_state(m, c->face[f] + 5 * me);
        } else {
          int f;
          if (i == -1) f = 1; else if (i == c->nx) f = 0;
          else if (j == -1) f = 3; else if (j == c->ny) f = 2;
          else if (k == -1) f = 5; else f = 4;
          double *fs = c->face[f] + 5 * me;
          if (orc_clip(m, fs)) { memcpy(fs, v, 5 * sizeof(double)); clipped++; }
          err |= orc_check_state(m, fs);
        }
      }
  *nclip += clipped;
  return err;
}

/* ------------------------------------------------------------------ residual ------------ */

/* one face: localflux1 (to the minus cell) and localflux2 (to the plus cell); the number of
 * Riemann-solver errors */
static int face_flux(orc_ctx *c, int dir, const double *V, const int *ID, long lo, long hi, const double *wall,
                     double *lf1, double *lf2) {
  const orc_config *g = &c->cfg;
  int nbid = ID[lo], myid = ID[hi];
  const double *vm = c->face[2 * dir + 1] + 5 * lo;   /* vr / vt / vf of the minus cell */
  const double *vp = c->face[2 * dir] + 5 * hi;       /* vl / vb / vk of the plus cell  */
  if (wall) {
    /* a face cut by an embedded wall (SpaceOperator.cpp:2438-2484): each active side solves its
     * one-sided problem against the wall (forward intersection for the minus cell, backward for
     * the plus cell) and takes the flux between its state and the clipped star state. Where the
     * reference exits (a negative status) the side's flux is zero and an error is counted. */
    double Vs[5];
    int nerr = 0;
    memset(lf1, 0, 5 * sizeof(double));
    memset(lf2, 0, 5 * sizeof(double));
    if (nbid != g->inactive_id) {
      const double *am = c->eb_recon == 0 ? V + 5 * lo : vm;
      int e = orc_wall_face_flux(c, dir, 1, c->eb_normal_mode, wall, wall + 3, am, nbid, lf1, Vs);
      if (e < 0) memset(lf1, 0, 5 * sizeof(double));
      nerr += e != 0;
    }
    if (myid != g->inactive_id) {
      const double *ap = c->eb_recon == 0 ? V + 5 * hi : vp;
      int e = orc_wall_face_flux(c, dir, -1, c->eb_normal_mode, wall + 6, wall + 9, ap, myid, lf2, Vs);
      if (e < 0) memset(lf2, 0, 5 * sizeof(double));
      nerr += e != 0;
    }
    return nerr;
  }
  if (nbid != myid) {
    if (nbid != g->inactive_id && myid != g->inactive_id) {
      const double *am = (g->mp_recon == ORC_REC_CONSTANT) ? V + 5 * lo : vm;
      const double *ap = (g->mp_recon == ORC_REC_CONSTANT) ? V + 5 * hi : vp;
      if (g->mp_flux == ORC_MPFLUX_LLF) {
        orc_flux_llf2(&g->mat[nbid], &g->mat[myid], dir, am, ap, lf1);
        memcpy(lf2, lf1, 5 * sizeof(double));
        return 0;
      }
      return orc_interface_flux(c, dir, am, nbid, ap, myid, lo, hi, lf1, lf2) != 0;
    }
    memset(lf1, 0, 5 * sizeof(double));
    memset(lf2, 0, 5 * sizeof(double));
    return 0;
  }
  if (myid != g->inactive_id) {
    orc_numerical_flux(c, dir, vm, vp, myid, lf1);
    memcpy(lf2, lf1, 5 * sizeof(double));
  } else {
    memset(lf1, 0, 5 * sizeof(double));
    memset(lf2, 0, 5 * sizeof(double));
  }
  return 0;
}

int orc_compute_residual(orc_ctx *c, const double *V, const int *ID, double *R) {
  int nx = c->nx, ny = c->ny, nz = c->nz;
  long nrev = 0;
  orc_riemann_begin_stage(c);
#pragma omp parallel for reduction(+ : nrev) schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) recon_cell(c, V, ID, i, j, k, &nrev);
  c->n_reverted += nrev;
  recon_ghost_faces(c, V);
  if (check_reconstructed(c, V, ID, &c->n_clipped)) return -1;

  long rerr = 0;
#pragma omp parallel for reduction(+ : rerr) schedule(static)
  for (int k = 0; k <= nz; k++) {
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i <= nx; i++) {
        if (j < ny && k < nz) {      /* x face i-1/2 */
          double *F = c->Fx + 10 * (((long)k * ny + j) * (nx + 1) + i);
          rerr += face_flux(c, 0, V, ID, ORC_IDX(c, i - 1, j, k), ORC_IDX(c, i, j, k), eb_face(c, 0, i, j, k),
                            F, F + 5);
        }
        if (i < nx && k < nz) {      /* y face j-1/2 */
          double *F = c->Fy + 10 * (((long)k * (ny + 1) + j) * nx + i);
          rerr += face_flux(c, 1, V, ID, ORC_IDX(c, i, j - 1, k), ORC_IDX(c, i, j, k), eb_face(c, 1, i, j, k),
                            F, F + 5);
        }
        if (i < nx && j < ny) {      /* z face k-1/2 */
          double *F = c->Fz + 10 * (((long)k * ny + j) * nx + i);
          rerr += face_flux(c, 2, V, ID, ORC_IDX(c, i, j, k - 1), ORC_IDX(c, i, j, k), eb_face(c, 2, i, j, k),
                            F, F + 5);
        }
      }
  }
  c->n_riemann_err += rerr;
  orc_riemann_end_stage(c);

  /* viscous and heat-diffusion terms (SpaceOperator.cpp:3367-3371), added after the advective sum */
  const int diff = orc_has_diffusion(c);
  if (diff) {
    const long nf[3] = {(long)(nx + 1) * ny * nz, (long)nx * (ny + 1) * nz, (long)nx * ny * (nz + 1)};
    for (int d = 0; d < 3; d++)
      if (!c->Dv[d]) {
        c->Dv[d] = (double *)calloc((size_t)nf[d] * 10, sizeof(double));
        c->Dh[d] = (double *)calloc((size_t)nf[d], sizeof(double));
      }
    orc_diffusion(c, V, ID, c->Dv, c->Dh);
  }

#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        /* face areas and volume of the cell (SpaceOperator.cpp:160-177, 2554) */
        const double wx = ORC_W(c, 0, i), wy = ORC_W(c, 1, j), wz = ORC_W(c, 2, k);
        const double ax = wy * wz, ay = wx * wz, az = wx * wy;
        const double inv = 1.0 / (-(wx * wy * wz));
        const long fxw = ((long)k * ny + j) * (nx + 1) + i, fys = ((long)k * (ny + 1) + j) * nx + i,
                   fzb = ((long)k * ny + j) * nx + i;
        const long fxe = fxw + 1, fyn = fys + nx, fzt = fzb + (long)nx * ny;
        const double *Fw = c->Fx + 10 * fxw + 5;
        const double *Fe = c->Fx + 10 * fxe;
        const double *Fs = c->Fy + 10 * fys + 5;
        const double *Fn = c->Fy + 10 * fyn;
        const double *Fb = c->Fz + 10 * fzb + 5;
        const double *Ft = c->Fz + 10 * fzt;
        double *r = R + 5 * ORC_IDX(c, i, j, k);
        for (int p = 0; p < 5; p++) {
          double s = 0.0;
          s -= Fw[p] * ax;
          s -= Fs[p] * ay;
          s -= Fb[p] * az;
          s += Fe[p] * ax;
          s += Fn[p] * ay;
          s += Ft[p] * az;
          if (diff) {
            /* ViscosityOperator.cpp:259-393: own faces added, then the east, north, front faces
             * (visited from the neighbouring cells) subtracted */
            s += c->Dv[0][10 * fxw + p];
            s += c->Dv[1][10 * fys + p];
            s += c->Dv[2][10 * fzb + p];
            s -= c->Dv[0][10 * fxe + 5 + p];
            s -= c->Dv[1][10 * fyn + 5 + p];
            s -= c->Dv[2][10 * fzt + 5 + p];
            if (p == 4) {            /* HeatDiffusionOperator.cpp:96-131 */
              s += c->Dh[0][fxw];
              s += c->Dh[1][fys];
              s += c->Dh[2][fzb];
              s -= c->Dh[0][fxe];
              s -= c->Dh[1][fyn];
              s -= c->Dh[2][fzt];
            }
          }
          r[p] = s * inv;
        }
      }
  return (int)rerr;
}

/* ------------------------------------------------------------------ time step size ------ */

int orc_time_step(const orc_ctx *c, const double *V, const int *ID, double *dt) {
  const orc_config *g = &c->cfg;
  double mn = DBL_MAX;
  int bad = 0;
#pragma omp parallel for reduction(min : mn) reduction(| : bad) schedule(static)
  for (int k = 0; k < c->nz; k++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nx; i++) {
        long me = ORC_IDX(c, i, j, k);
        int id = ID[me];
        if (id == g->inactive_id) continue;
        const orc_eos *m = &g->mat[id];
        const double *v = V + 5 * me;
        double c2 = orc_c2(m, v[0], orc_internal_energy(m, v[0], v[4]));
        if (c2 < 0) { bad = 1; continue; }
        double cs = sqrt(c2);
        double lf = fabs(v[1]) + cs, lg = fabs(v[2]) + cs, lh = fabs(v[3]) + cs;
        double a = ORC_W(c, 0, i) / lf, b = ORC_W(c, 1, j) / lg, cc = ORC_W(c, 2, k) / lh;
        double m2 = (cc < b) ? cc : b;
        double m1 = (m2 < a) ? m2 : a;
        mn = (m1 < mn) ? m1 : mn;
      }
  *dt = g->cfl * mn;
  return bad ? -1 : 0;
}

/* ------------------------------------------------------------------ RK stages ----------- */

void orc_p2c_all(orc_ctx *c, const double *V, const int *ID, double *U) {
#pragma omp parallel for schedule(static)
  for (int k = 0; k < c->nz; k++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nx; i++) {
        long me = ORC_IDX(c, i, j, k);
        orc_p2c(&c->cfg.mat[ID[me]], V + 5 * me, U + 5 * me);
      }
}

/* y = a*y + b*x over the interior (SpaceVariable.cpp:377-402) */
void orc_axpby(orc_ctx *c, double *y, double a, double b, const double *x) {
#pragma omp parallel for schedule(static)
  for (int k = 0; k < c->nz; k++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nx; i++) {
        long me = 5 * ORC_IDX(c, i, j, k);
        for (int p = 0; p < 5; p++) y[me + p] = a * y[me + p] + b * x[me + p];
      }
}

/* C2P over the interior + ClipDensityAndPressure with CheckState (SpaceOperator.cpp:553-649).
 * Returns the number of clipped cells, or -1 on a hyperbolicity violation. */
long orc_c2p_clip(orc_ctx *c, const double *U, const int *ID, double *V) {
  long nclip = 0;
  int bad = 0;
#pragma omp parallel for reduction(+ : nclip) reduction(| : bad) schedule(static)
  for (int k = 0; k < c->nz; k++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nx; i++) {
        long me = ORC_IDX(c, i, j, k);
        const orc_eos *m = &c->cfg.mat[ID[me]];
        if (U) orc_c2p(m, U + 5 * me, V + 5 * me);   /* U == NULL: clip V in place */
        nclip += orc_clip(m, V + 5 * me);
        bad |= orc_check_state(m, V + 5 * me);
      }
  c->n_clipped += nclip;
  return bad ? -1 : nclip;
}


/* ------------------------------------------------------------------ pointwise EOS ------- */

void orc_eos_p2c(const orc_ctx *c, int id, const double *V, double *U) { orc_p2c(&c->cfg.mat[id], V, U); }
void orc_eos_c2p(const orc_ctx *c, int id, const double *U, double *V) { orc_c2p(&c->cfg.mat[id], U, V); }
double orc_eos_c2(const orc_ctx *c, int id, double rho, double p) {
  const orc_eos *m = &c->cfg.mat[id];
  return orc_c2(m, rho, orc_internal_energy(m, rho, p));
}

/* Pointwise EOS of material id at (rho, e) and (rho, p): out[0] p(rho,e), out[1] e(rho,p),
 * out[2] dp/drho|e (rho,e), out[3] Gamma(rho,e), out[4] c^2(rho,e), out[5] CheckState(rho,p)
 * (VarFcnBase.h:62-164 through the per-EOS closed forms of eos.h). */
void orc_eos_point(const orc_ctx *c, int id, double rho, double e, double p, double *out) {
  const orc_eos *m = &c->cfg.mat[id];
  out[0] = orc_pressure(m, rho, e);
  out[1] = orc_internal_energy(m, rho, p);
  out[2] = orc_dpdrho(m, rho, e);
  out[3] = orc_big_gamma(m, rho, e);
  out[4] = orc_c2(m, rho, e);
  out[5] = (double)orc_check_state_rp(m, rho, p);
}

/* FluxFcnLLF::ComputeNumericalFluxAtMaterialInterface (FluxFcnLLF.h:112-185) of materials idm | idp. */
void orc_material_interface_flux(const orc_ctx *c, int dir, const double *Vm, int idm, const double *Vp, int idp,
                                 double *F) {
  orc_flux_llf2(&c->cfg.mat[idm], &c->cfg.mat[idp], dir, Vm, Vp, F);
}
