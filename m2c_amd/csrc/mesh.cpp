// mesh.cpp -- the cells of one mesh axis (host code).
//
// MeshGenerator::ComputeMesh1DUniform (MeshGenerator.cpp:141-153) for an axis given by its cell
// count, and MeshGenerator::ComputeMesh1DNonUniform (MeshGenerator.cpp:163-259) for an axis given
// by control points (coordinate, cell width), between which the cell width varies linearly in x:
// the cell count is the integral of dx/h(x) over the axis rounded down, and the cell boundaries
// follow the exponential map of that integral, restarted at each control point. The steps and
// their floating-point order are the reference's, so the coordinates equal the reference's.
#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include "../../include/m2c_hip.h"
#include "deck.hpp"

extern "C" int m2c_mesh_1d(double x0, double xmax, int n, int npts, const double *pts, int cap, double *x,
                           double *dx, int *nout) {
  if (!nout || cap < 0 || npts < 0 || (npts > 0 && !pts) || xmax < x0) return M2C_ERR_ARG;
  std::vector<double> xs, ws;
  if (npts == 0) {   // ComputeMesh1DUniform
    if (n <= 0) return M2C_ERR_ARG;
    const double h = (xmax - x0) / n;
    ws.assign((size_t)n, h);
    xs.resize((size_t)n);
    for (int i = 0; i < n; i++) xs[i] = x0 + 0.5 * h + i * h;
  } else {
    std::vector<std::pair<double, double>> p;
    for (int q = 0; q < npts; q++) {
      // a width that is not positive and finite has no cell count (the reference divides by it)
      if (!(pts[2 * q + 1] > 0.0) || !std::isfinite(pts[2 * q + 1]) || !std::isfinite(pts[2 * q])) return M2C_ERR_ARG;
      p.emplace_back(pts[2 * q], pts[2 * q + 1]);
    }
    std::sort(p.begin(), p.end());
    // the end points take the nearest control point's width
    if (p.front().first > x0) p.insert(p.begin(), std::make_pair(x0, p.front().second));
    else if (p.front().first < x0) return M2C_ERR_ARG;
    if (p.back().first < xmax) p.push_back(std::make_pair(xmax, p.back().second));
    else if (p.back().first > xmax) return M2C_ERR_ARG;
    const int np = (int)p.size();
    // number of cells
    double Nreal = 0.0, a, b, xp1, xp2, h1, h2;
    for (int i = 0; i < np - 1; i++) {
      xp1 = p[i].first; h1 = p[i].second;
      xp2 = p[i + 1].first; h2 = p[i + 1].second;
      a = (h2 - h1) / (xp2 - xp1);
      b = (h1 * xp2 - h2 * xp1) / (xp2 - xp1);
      Nreal += (a == 0.0) ? (xp2 - xp1) / b : 1.0 / a * log((a * xp2 + b) / (a * xp1 + b));
    }
    if (!(Nreal < 1e9)) return M2C_ERR_ARG;   // beyond any mesh one device holds
    const int N = std::max(1, (int)(Nreal + 1e-12));
    // cell boundaries
    const double kappa = N / Nreal;
    std::vector<double> xi((size_t)N + 1);
    xi[0] = x0;
    int i = 0;
    xp1 = p[i].first; h1 = p[i].second;
    xp2 = p[i + 1].first; h2 = p[i + 1].second;
    a = (h2 - h1) / (xp2 - xp1);
    b = (h1 * xp2 - h2 * xp1) / (xp2 - xp1);
    for (int k = 0; k < N; k++) {
      xi[k + 1] = (a == 0.0) ? xi[k] + b / kappa : (xi[k] + b / a) * exp(a / kappa) - b / a;
      if (xi[k + 1] >= xp2) {
        if (k + 1 == N && i + 1 == np - 1) {   // done
          xi[k + 1] = xp2;
          break;
        } else if (k + 1 == N || i + 1 == np - 1) {
          return M2C_ERR_ARG;
        }
        const double C = (a == 0.0) ? (xp2 - xi[k]) / b : 1.0 / a * log((a * xp2 + b) / (a * xi[k] + b));
        i++;
        xp1 = xp2; h1 = h2;
        xp2 = p[i + 1].first; h2 = p[i + 1].second;
        a = (h2 - h1) / (xp2 - xp1);
        b = (h1 * xp2 - h2 * xp1) / (xp2 - xp1);
        xi[k + 1] = (a == 0.0) ? xp1 + b * (1 / kappa - C) : (xp1 + b / a) * exp(a * (1 / kappa - C)) - b / a;
        if (xi[k + 1] > xp2) return M2C_ERR_ARG;   // control points and widths in conflict
      }
    }
    if (xi[N] < xmax) xi[N] = xmax;
    ws.resize((size_t)N);
    xs.resize((size_t)N);
    for (int q = 0; q < N; q++) ws[q] = xi[q + 1] - xi[q];
    xs[0] = x0 + 0.5 * ws[0];
    for (int q = 0; q < N - 1; q++) xs[q + 1] = xs[q] + 0.5 * (ws[q] + ws[q + 1]);
  }
  const int m = (int)ws.size();
  *nout = m;
  for (int q = 0; q < m && q < cap; q++) {
    if (x) x[q] = xs[q];
    if (dx) dx[q] = ws[q];
  }
  return M2C_OK;
}

int m2c_deck::mesh_axes(const m2c_config &c, std::vector<double> x[3], std::vector<double> w[3]) {
  const double lo[3] = {c.x0, c.y0, c.z0}, hi[3] = {c.x1, c.y1, c.z1};
  const int n[3] = {c.nx, c.ny, c.nz};
  for (int d = 0; d < 3; d++) {
    if (n[d] <= 0) return M2C_ERR_ARG;
    x[d].assign((size_t)n[d], 0.0);
    w[d].assign((size_t)n[d], 0.0);
    int m = 0;
    if (int rc = m2c_mesh_1d(lo[d], hi[d], n[d], c.mesh_npts[d], &c.mesh_pts[d][0][0], n[d], x[d].data(), w[d].data(), &m))
      return rc;
    if (m != n[d]) return M2C_ERR_ARG;
  }
  return M2C_OK;
}
