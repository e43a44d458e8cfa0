/* oracle/embedded_walls.c -- wall records of mesh faces cut by embedded surfaces (CPU oracle: test
 * infrastructure only).
 *
 * Restates SpaceOperator::FindEdgeSurfaceIntersections (SpaceOperator.cpp:3434-3500): over the
 * surfaces, the forward crossing (seen from the face's minus cell) with the smallest distance and
 * the backward crossing with the largest distance from the edge's first node, the first surface
 * winning ties (strict comparisons from +/-DBL_MAX); the wall normal is the crossing triangle's
 * normal (elemNorm[tid]) and the wall velocity xi0 Udot[n0] + xi1 Udot[n1] + xi2 Udot[n2], summed
 * left to right per component as the reference's Vec3D expression is. A face with a crossing in one
 * direction only fails (the reference asserts). */
#include <float.h>
#include <stddef.h>

#include "m2c_oracle.h"

static int crossing(const orc_surface *s, int q, double *nw, double *vw) {
  int tid = s->point_tid[q];
  if (tid < 0 || tid >= s->nelems) return -1;
  const int *n = s->elems + 3 * (size_t)tid;
  for (int c = 0; c < 3; c++)
    if (n[c] < 0 || n[c] >= s->nnodes) return -1;
  const double *xi = s->point_xi + 3 * (size_t)q;
  for (int d = 0; d < 3; d++) {
    nw[d] = s->elem_norm[3 * (size_t)tid + d];
    double a = xi[0] * s->udot[3 * (size_t)n[0] + d];
    double b = xi[1] * s->udot[3 * (size_t)n[1] + d];
    double c = xi[2] * s->udot[3 * (size_t)n[2] + d];
    vw[d] = (a + b) + c;
  }
  return 0;
}

int orc_edge_surface_walls(int nsurf, const orc_surface *surf, int nfaces, const int *xf, const int *xb, int *cut,
                           double *wall) {
  for (int t = 0; t < nfaces; t++) {
    int fsurf = -1, fid = -1, bsurf = -1, bid = -1;
    double fdist = DBL_MAX, bdist = -DBL_MAX;
    for (int s = 0; s < nsurf; s++) {
      int x = xf[(size_t)t * nsurf + s];
      if (x >= surf[s].npoints) return -1;
      if (x >= 0) {
        double dist = surf[s].point_dist[x];
        if (dist < fdist) { fsurf = s; fid = x; fdist = dist; }
      }
      x = xb[(size_t)t * nsurf + s];
      if (x >= surf[s].npoints) return -1;
      if (x >= 0) {
        double dist = surf[s].point_dist[x];
        if (dist > bdist) { bsurf = s; bid = x; bdist = dist; }
      }
    }
    if (fsurf < 0) {
      if (bsurf >= 0) return -1;
      cut[t] = 0;
      continue;
    }
    if (bsurf < 0) return -1;
    cut[t] = 1;
    if (crossing(&surf[fsurf], fid, wall + 12 * (size_t)t, wall + 12 * (size_t)t + 3)) return -1;
    if (crossing(&surf[bsurf], bid, wall + 12 * (size_t)t + 6, wall + 12 * (size_t)t + 9)) return -1;
  }
  return 0;
}
