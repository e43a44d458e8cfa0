/* oracle/mesh.c -- the cells of one mesh axis (CPU oracle: test infrastructure only).
 *
 * Restates MeshGenerator::ComputeMesh1DUniform (MeshGenerator.cpp:141-153) and
 * MeshGenerator::ComputeMesh1DNonUniform (MeshGenerator.cpp:163-259): between control points
 * (coordinate, cell width) the width varies linearly, h(x) = a x + b; the cell count is the integral
 * of dx / h rounded down, N; each cell then covers the same share 1/kappa = Nreal/N of that
 * integral, so its right boundary solves int dx/h = 1/kappa in closed form, x' = (x + b/a)
 * exp(a/kappa) - b/a (x' = x + b/kappa where h is constant), restarted with the remainder at each
 * control point. Centres: x0 + dx_0/2, then x_{i+1} = x_i + (dx_i + dx_{i+1})/2. */
#include <math.h>
#include <stdlib.h>

#include "m2c_oracle.h"

static int cmp_pair(const void *a, const void *b) {   /* lexicographic (coordinate, width), as std::sort */
  const double *p = (const double *)a, *q = (const double *)b;
  if (p[0] < q[0]) return -1;
  if (p[0] > q[0]) return 1;
  if (p[1] < q[1]) return -1;
  if (p[1] > q[1]) return 1;
  return 0;
}

int orc_mesh_1d(double x0, double xmax, int n, int npts, const double *pts, int cap, double *x, double *dx,
                int *nout) {
  if (xmax < x0 || npts < 0) return -1;
  if (npts == 0) {
    if (n <= 0) return -1;
    const double h = (xmax - x0) / n;
    for (int i = 0; i < n && i < cap; i++) {
      if (dx) dx[i] = h;
      if (x) x[i] = x0 + 0.5 * h + i * h;
    }
    *nout = n;
    return 0;
  }
  double *p = (double *)malloc(sizeof(double) * 2 * (size_t)(npts + 2));
  int np = 0;
  for (int q = 0; q < npts; q++) { p[2 * q] = pts[2 * q]; p[2 * q + 1] = pts[2 * q + 1]; }
  qsort(p, (size_t)npts, 2 * sizeof(double), cmp_pair);
  np = npts;
  if (p[0] > x0) {   /* the end points take the nearest control point's width */
    for (int q = np; q > 0; q--) { p[2 * q] = p[2 * q - 2]; p[2 * q + 1] = p[2 * q - 1]; }
    p[0] = x0;   /* p[1] keeps the first point's width */
    np++;
  } else if (p[0] < x0) { free(p); return -1; }
  if (p[2 * (np - 1)] < xmax) {
    p[2 * np] = xmax;
    p[2 * np + 1] = p[2 * (np - 1) + 1];
    np++;
  } else if (p[2 * (np - 1)] > xmax) { free(p); return -1; }
  double Nreal = 0.0;
  for (int s = 0; s < np - 1; s++) {
    const double xa = p[2 * s], ha = p[2 * s + 1], xb = p[2 * s + 2], hb = p[2 * s + 3];
    const double a = (hb - ha) / (xb - xa), b = (ha * xb - hb * xa) / (xb - xa);
    Nreal += (a == 0.0) ? (xb - xa) / b : 1.0 / a * log((a * xb + b) / (a * xa + b));
  }
  int N = (int)(Nreal + 1e-12);
  if (N < 1) N = 1;
  const double kappa = N / Nreal;
  double *xi = (double *)malloc(sizeof(double) * (size_t)(N + 1));
  xi[0] = x0;
  int s = 0;
  double xp1 = p[0], h1 = p[1], xp2 = p[2], h2 = p[3];
  double a = (h2 - h1) / (xp2 - xp1), b = (h1 * xp2 - h2 * xp1) / (xp2 - xp1);
  int err = 0;
  for (int k = 0; k < N; k++) {
    xi[k + 1] = (a == 0.0) ? xi[k] + b / kappa : (xi[k] + b / a) * exp(a / kappa) - b / a;
    if (xi[k + 1] >= xp2) {
      if (k + 1 == N && s + 1 == np - 1) { xi[k + 1] = xp2; break; }
      if (k + 1 == N || s + 1 == np - 1) { err = 1; break; }
      const double C = (a == 0.0) ? (xp2 - xi[k]) / b : 1.0 / a * log((a * xp2 + b) / (a * xi[k] + b));
      s++;
      xp1 = xp2; h1 = h2;
      xp2 = p[2 * (s + 1)]; h2 = p[2 * (s + 1) + 1];
      a = (h2 - h1) / (xp2 - xp1);
      b = (h1 * xp2 - h2 * xp1) / (xp2 - xp1);
      xi[k + 1] = (a == 0.0) ? xp1 + b * (1.0 / kappa - C) : (xp1 + b / a) * exp(a * (1.0 / kappa - C)) - b / a;
      if (xi[k + 1] > xp2) { err = 1; break; }
    }
  }
  if (!err) {
    if (xi[N] < xmax) xi[N] = xmax;
    double xprev = 0.0, wprev = 0.0;
    for (int i = 0; i < N; i++) {
      const double w = xi[i + 1] - xi[i];
      const double xc = (i == 0) ? x0 + 0.5 * w : xprev + 0.5 * (wprev + w);
      if (i < cap) {
        if (dx) dx[i] = w;
        if (x) x[i] = xc;
      }
      xprev = xc;
      wprev = w;
    }
    *nout = N;
  }
  free(xi);
  free(p);
  return err ? -1 : 0;
}
