// embedded_walls.cpp -- the wall records of mesh faces cut by embedded surfaces (host code).
//
// SpaceOperator::FindEdgeSurfaceIntersections (SpaceOperator.cpp:3434-3500): per face the nearest
// forward and backward crossings over all surfaces, then the crossing triangle's normal and its
// nodes' velocities interpolated with the crossing's barycentric coordinates. The records feed
// m2c_set_embedded_faces.
#include <cfloat>
#include <cstddef>

#include "../../include/m2c_hip.h"

namespace {

// normal and velocity of crossing q of surface s (the triangle's normal; xi-weighted node velocities,
// summed in the order the reference's Vec3D expression evaluates them)
bool wall_of(const m2c_surface &s, int q, double *nw, double *vw) {
  const int tid = s.point_tid[q];
  if (tid < 0 || tid >= s.nelems) return false;
  const int *e = s.elems + 3 * (std::size_t)tid;
  for (int c = 0; c < 3; c++)
    if (e[c] < 0 || e[c] >= s.nnodes) return false;
  const double *xi = s.point_xi + 3 * (std::size_t)q;
  for (int d = 0; d < 3; d++) {
    nw[d] = s.elem_norm[3 * (std::size_t)tid + d];
    vw[d] = xi[0] * s.udot[3 * (std::size_t)e[0] + d] + xi[1] * s.udot[3 * (std::size_t)e[1] + d] +
            xi[2] * s.udot[3 * (std::size_t)e[2] + d];
  }
  return true;
}

}  // namespace

extern "C" int m2c_edge_surface_walls(int nsurf, const m2c_surface *surf, int nfaces, const int *xf, const int *xb,
                                      int *cut, double *wall) {
  if (nsurf < 0 || nfaces < 0) return M2C_ERR_ARG;
  if (nfaces == 0) return M2C_OK;
  if (nsurf > 0 && !surf) return M2C_ERR_ARG;
  if (!xf || !xb || !cut || !wall) return M2C_ERR_ARG;
  for (int s = 0; s < nsurf; s++) {
    const m2c_surface &S = surf[s];
    if (S.nnodes < 0 || S.nelems < 0 || S.npoints < 0) return M2C_ERR_ARG;
    if ((S.nnodes && !S.udot) || (S.nelems && (!S.elems || !S.elem_norm)) ||
        (S.npoints && (!S.point_tid || !S.point_xi || !S.point_dist)))
      return M2C_ERR_ARG;
  }
  for (int t = 0; t < nfaces; t++) {
    int fs = -1, fq = -1, bs = -1, bq = -1;
    double fd = DBL_MAX, bd = -DBL_MAX;
    for (int s = 0; s < nsurf; s++) {
      const int qf = xf[(std::size_t)t * nsurf + s], qb = xb[(std::size_t)t * nsurf + s];
      if (qf >= surf[s].npoints || qb >= surf[s].npoints) return M2C_ERR_ARG;
      if (qf >= 0 && surf[s].point_dist[qf] < fd) { fs = s; fq = qf; fd = surf[s].point_dist[qf]; }
      if (qb >= 0 && surf[s].point_dist[qb] > bd) { bs = s; bq = qb; bd = surf[s].point_dist[qb]; }
    }
    if ((fs < 0) != (bs < 0)) return M2C_ERR_ARG;   // a crossing seen from one side only
    cut[t] = fs >= 0;
    if (fs < 0) continue;
    double *w = wall + 12 * (std::size_t)t;
    if (!wall_of(surf[fs], fq, w, w + 3) || !wall_of(surf[bs], bq, w + 6, w + 9)) return M2C_ERR_ARG;
  }
  return M2C_OK;
}
