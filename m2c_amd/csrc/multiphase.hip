// multiphase.hip -- two-material interface faces, level set and end-of-step tasks (gfx950).
//
//   interface faces      SpaceOperator.cpp:2486-2542 (+ y/z analogues): exact two-material Riemann
//                        solve along the mesh normal, clip of the one-sided star states, the
//                        Riemann-solution store of the first stage (RiemannSolutions.h:12-30,
//                        SpaceOperator.cpp:2518-2523) and the two one-sided numerical fluxes
//   level-set stages     LevelSetOperator.cpp:1066-1141 (residual), GradientCalculatorFD3.cpp:54-171,
//                        LevelSetOperator.cpp:934-1033 (ghost rules, evaluated on the fly) and the
//                        stage combinations of TimeIntegrator.cpp:150-690
//   material IDs         MultiPhaseOperator.cpp:127-330
//   reinitialisation     LevelSetReinitializer.cpp:74-172, :297-430, :594-615, :953-1035, :1970-1999
//                        (first layer FIXED)
//   state update         MultiPhaseOperator.cpp:857-1028 (LocalUpdateByRiemannSolutions),
//                        :1156-1320 (FixUnresolvedNodes, failsafe density)
//
// Device fields have no ghost frame: ghost values of V, ID and phi are produced by the boundary
// rules where they are read. Interface faces are gathered once per step into a list (the material
// IDs are fixed during a step); the Riemann pass runs one thread per listed face and writes both
// one-sided fluxes into face-keyed arrays that the fused stage kernel reads.
//
// The reference updates the cells whose ID changed in place, in (k,j,i) order, so a cell sees the
// updated velocity of an earlier changed neighbour. Here every changed cell is first computed from
// the pre-update field; a cell whose computation reads a changed lower neighbour across a face that
// carries a usable Riemann solution is recomputed once that neighbour is final (a few dependency
// rounds, then an ordered single-thread pass for any chain that is still open), which reproduces
// the in-order result exactly.
#include <hip/hip_runtime.h>
#include <type_traits>

#include "fv_kernels.hpp"
#include "multiphase.hpp"
#include "riemann_dev.hpp"
#include "sg_star.hpp"

namespace m2c {
#ifdef M2C_IFACE_MULTI
namespace imulti {   // a second copy of the interface code: its own symbols
#endif

#ifdef M2C_IFACE_MULTI
static __shared__ MpParams s_imp;
#define M2C_MP s_imp
#else
static __constant__ MpParams c_mp;
#define M2C_MP c_mp
#endif

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ long long cidx3(int i, int j, int k) {
  return ((long long)k * M2C_G.ny + j) * M2C_G.nx + i;
}

__device__ __forceinline__ double XC(int d, int i) { return M2C_MP.xc[d][i + 1]; }

// min(dx, dy, dz) of cell (i,j,k): the reinitialisation's pseudo-time step / cfl and the width of
// its sign function (LevelSetReinitializer.cpp:600-603, 988-989)
__device__ __forceinline__ double cell_min_width(int i, int j, int k) {
  if (!M2C_MP.stretched) return M2C_MP.min_dx;
  return smin(M2C_MP.wc[0][i + 1], smin(M2C_MP.wc[1][j + 1], M2C_MP.wc[2][k + 1]));
}

// wave-aggregated append; every lane of the wavefront must call it
__device__ __forceinline__ unsigned wave_append(unsigned *counter, bool pred) {
  unsigned long long m = __ballot(pred);
  const int lane = threadIdx.x & 63;
  unsigned base = 0;
  if (m) {
    const int leader = __ffsll((long long)m) - 1;
    if (lane == leader) base = atomicAdd(counter, (unsigned)__popcll(m));
    base = __shfl((int)base, leader, 64);
  }
  return base + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
}

__device__ __forceinline__ void atomic_max_pos(unsigned long long *dst, double v) {
  atomicMax(dst, (unsigned long long)__double_as_longlong(v));   // v >= 0: bits are ordered
}

__device__ __forceinline__ double bits_to_d(unsigned long long b) { return __longlong_as_double((long long)b); }

// material ID of (possibly ghost) cell (i,j,k): the image cell's (MultiPhaseOperator.cpp:796-819)
__device__ __forceinline__ int cell_id(const int8_t *ID, int i, int j, int k) {
  const Grid &g = M2C_G;
  int ii = clampi(image_idx(i, g.nx, i < 0 ? g.bc[0] == BC_PERIODIC : g.bc[1] == BC_PERIODIC), 0, g.nx - 1);
  int jj = clampi(image_idx(j, g.ny, j < 0 ? g.bc[2] == BC_PERIODIC : g.bc[3] == BC_PERIODIC), 0, g.ny - 1);
  int kk = clampi(image_idx(k, g.nz, k < 0 ? g.bc[4] == BC_PERIODIC : g.bc[5] == BC_PERIODIC), 0, g.nz - 1);
  return ID[cidx3(ii, jj, kk)];
}

// level set at (i,j,k) with at most one index one step outside the box: the face-ghost rules of
// LevelSetOperator.cpp:934-1033 (1 zero-Neumann, 2 linear extrapolation, 3 non-negative,
// 7 periodic)
__device__ inline double phi_at(const double *phi, int i, int j, int k) {
  const Grid &g = M2C_G;
  int dim, n, q;
  long long stride;
  if (i < 0 || i >= g.nx) { dim = 0; n = g.nx; q = i; stride = 1; }
  else if (j < 0 || j >= g.ny) { dim = 1; n = g.ny; q = j; stride = g.nx; }
  else if (k < 0 || k >= g.nz) { dim = 2; n = g.nz; q = k; stride = (long long)g.nx * g.ny; }
  else return phi[cidx3(i, j, k)];
  // (i,j,k) is a ghost along `dim` only; c is its (virtual) linear index
  const long long c = cidx3(i, j, k);
  const int hi = q >= n;
  const int bc = M2C_MP.ls_bc[2 * dim + hi];
  int im;
  if (bc == BC_PERIODIC) im = hi ? 0 : n - 1;
  else im = hi ? n - 1 : 0;
  double v = phi[c + (long long)(im - q) * stride];
  if (bc == 2 || bc == 3) {
    if (n >= 2) {
      const int s = hi ? -1 : 1;
      double r = XC(dim, q), r1 = XC(dim, q + s), r2 = XC(dim, q + 2 * s);
      double f1 = phi[c + s * stride], f2 = phi[c + 2 * s * stride];
      v = f1 + (f2 - f1) / (r2 - r1) * (r - r1);
    }
    if (bc == 3) {
      double d2wall = 0.5 * dabs(XC(dim, im) - XC(dim, q));
      v = smax(-d2wall, v);
    }
  }
  return v;
}

// ---------------------------------------------------------------- interface normal from the level set

// phi at (i,j,k) for the interface normal: periodic axes are wrapped first, then at most one index
// is outside the box and takes the face-ghost rule of phi_at
__device__ inline double phi_nrm(const double *phi, int i, int j, int k) {
  const Grid &g = M2C_G;
  if (i < 0 && M2C_MP.ls_bc[0] == BC_PERIODIC) i += g.nx;
  else if (i >= g.nx && M2C_MP.ls_bc[1] == BC_PERIODIC) i -= g.nx;
  if (j < 0 && M2C_MP.ls_bc[2] == BC_PERIODIC) j += g.ny;
  else if (j >= g.ny && M2C_MP.ls_bc[3] == BC_PERIODIC) j -= g.ny;
  if (k < 0 && M2C_MP.ls_bc[4] == BC_PERIODIC) k += g.nz;
  else if (k >= g.nz && M2C_MP.ls_bc[5] == BC_PERIODIC) k -= g.nz;
  return phi_at(phi, i, j, k);
}

// CentralDifferenceLocal (SpaceOperator.h:218-223)
__device__ __forceinline__ double cdiff_local(double p0, double p1, double p2, double x0, double x1, double x2) {
  const double c0 = -(x2 - x1) / ((x1 - x0) * (x2 - x0));
  const double c1 = 1.0 / (x1 - x0) - 1.0 / (x2 - x1);
  const double c2 = (x1 - x0) / ((x2 - x0) * (x2 - x1));
  return c0 * p0 + c1 * p1 + c2 * p2;
}

// SpaceOperator::CalculateGradientAtCellInterface (SpaceOperator.cpp:3052-3190): the normalised
// gradient of phi at the D-face of cell q = (i,j,k) (between q - e_D and q). The normal component
// is the difference across the face; each transverse one is the central difference (one-sided at
// the ghost layer) in the two cells, interpolated to the face.
template <int D>
__device__ void grad_at_face(const double *phi, const int q[3], double *dir) {
  const Grid &g = M2C_G;
  const int n[3] = {g.nx, g.ny, g.nz};
  int m[3] = {q[0], q[1], q[2]};
  m[D] -= 1;
  auto P = [&](const int *c, int t, int s) {
    int r[3] = {c[0], c[1], c[2]};
    r[t] += s;
    return phi_nrm(phi, r[0], r[1], r[2]);
  };
  dir[D] = (P(q, D, 0) - P(m, D, 0)) / (XC(D, q[D]) - XC(D, m[D]));
  const double x_half = XC(D, q[D]) - 0.5 * M2C_MP.wc[D][q[D] + 1];
  const double dd = XC(D, q[D]) - XC(D, m[D]);
  const double c1 = (XC(D, q[D]) - x_half) / dd;
  const double c2 = (x_half - XC(D, m[D])) / dd;
#pragma unroll
  for (int a = 1; a <= 2; a++) {
    const int tt = a == 1 ? (D == 0 ? 1 : 0) : (D == 2 ? 1 : 2);   // the transverse axes in order
    const int qt = q[tt];
    double d1, d2;
    if (qt < 0) {
      d1 = (P(m, tt, 1) - P(m, tt, 0)) / (XC(tt, qt + 1) - XC(tt, qt));
      d2 = (P(q, tt, 1) - P(q, tt, 0)) / (XC(tt, qt + 1) - XC(tt, qt));
    } else if (qt >= n[tt]) {
      d1 = (P(m, tt, 0) - P(m, tt, -1)) / (XC(tt, qt) - XC(tt, qt - 1));
      d2 = (P(q, tt, 0) - P(q, tt, -1)) / (XC(tt, qt) - XC(tt, qt - 1));
    } else {
      d1 = cdiff_local(P(m, tt, -1), P(m, tt, 0), P(m, tt, 1), XC(tt, qt - 1), XC(tt, qt), XC(tt, qt + 1));
      d2 = cdiff_local(P(q, tt, -1), P(q, tt, 0), P(q, tt, 1), XC(tt, qt - 1), XC(tt, qt), XC(tt, qt + 1));
    }
    dir[tt] = c1 * d1 + c2 * d2;
  }
  const double nrm = dsqrt<double>(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
  if (nrm != 0.0) {
    const double s = 1.0 / nrm;
    dir[0] *= s; dir[1] *= s; dir[2] *= s;
  }
}

// GetNormalForBimaterialRiemann with CalculateGradPhiAtCellInterface (SpaceOperator.cpp:2966-3048)
// for one level set: the level-set gradient at the face, turned when the minus cell is material 0,
// the mesh normal if it vanishes, its axis component made non-negative; AVERAGE adds the mesh normal
// and normalises. myid: the plus cell's material, nbid: the minus cell's.
__device__ void interface_normal(int D, int i, int j, int k, int myid, int nbid, const double *phi, double *dir) {
  dir[0] = dir[1] = dir[2] = 0.0;
  const int mode = M2C_MP.riemann_normal;
  if (mode == 1 || !phi) {
    dir[D] = 1.0;
    return;
  }
  const int q[3] = {i, j, k};
  if (D == 0) grad_at_face<0>(phi, q, dir);
  else if (D == 1) grad_at_face<1>(phi, q, dir);
  else grad_at_face<2>(phi, q, dir);
  if (nbid == 0) { dir[0] *= -1.0; dir[1] *= -1.0; dir[2] *= -1.0; }
  if (dsqrt<double>(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]) == 0.0) dir[D] = 1.0;
  if (dir[D] < 0.0) dir[D] *= -1.0;
  if (mode == 2) {
    dir[D] += 1.0;
    const double s = 1.0 / dsqrt<double>(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    dir[0] *= s; dir[1] *= s; dir[2] *= s;
  }
}

// linear index of (i,j,k) with each coordinate clamped into the box
__device__ __forceinline__ long long clamped_idx(int i, int j, int k) {
  const Grid &g = M2C_G;
  return cidx3(clampi(i, 0, g.nx - 1), clampi(j, 0, g.ny - 1), clampi(k, 0, g.nz - 1));
}

// Cell sweeps over the box: blocks of 64 x 4 (i, j) columns, each thread walking SWEEP_KB planes
// of k; no integer division in the index arithmetic.
constexpr int SWEEP_BX = 64, SWEEP_BY = 4, SWEEP_KB = 4;

static inline dim3 sweep_grid(int nx, int ny, int nz) {
  return dim3((nx + SWEEP_BX - 1) / SWEEP_BX, (ny + SWEEP_BY - 1) / SWEEP_BY, (nz + SWEEP_KB - 1) / SWEEP_KB);
}

__device__ __forceinline__ bool first_thread() {
  return blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0 && threadIdx.y == 0;
}

// block-wide maximum of non-negative v; thread (0,0) of each block folds it into *dst
__device__ __forceinline__ void block_max_to(unsigned long long *dst, double v, bool skip_zero) {
  __shared__ double red[SWEEP_BY];
  for (int off = 32; off > 0; off >>= 1) v = smax(v, __shfl_xor(v, off, 64));
  if (threadIdx.x == 0) red[threadIdx.y] = v;
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    double m = red[0];
    for (int w = 1; w < SWEEP_BY; w++) m = smax(m, red[w]);
    if (!skip_zero || m > 0.0) atomic_max_pos(dst, m);
  }
}

// calls f(i, j, k, idx, inner) for this thread's cells; inner = all six face neighbours are in the box
template <class F>
__device__ __forceinline__ void sweep_cells(F &&f) {
  const Grid &g = M2C_G;
  const int i = blockIdx.x * SWEEP_BX + threadIdx.x, j = blockIdx.y * SWEEP_BY + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const int k0 = blockIdx.z * SWEEP_KB, k1 = min(g.nz, k0 + SWEEP_KB);
  const bool inij = i > 0 && i < g.nx - 1 && j > 0 && j < g.ny - 1;
  const long long plane = (long long)g.nx * g.ny;
  const long long idx0 = ((long long)k0 * g.ny + j) * g.nx + i;
  if (k1 - k0 == SWEEP_KB) {   // full tile: unrolled so that the planes' loads can overlap
#pragma unroll
    for (int q = 0; q < SWEEP_KB; q++) f(i, j, k0 + q, idx0 + q * plane, inij && k0 + q > 0 && k0 + q < g.nz - 1);
  } else {
    long long idx = idx0;
    for (int k = k0; k < k1; k++, idx += plane) f(i, j, k, idx, inij && k > 0 && k < g.nz - 1);
  }
}

// the six face neighbours of cell (i,j,k) = idx, ghosts by the level-set boundary rules
struct Nb6 { double im, ip, jm, jp, km, kp; };
__device__ __forceinline__ Nb6 phi_nb6(const double *phi, int i, int j, int k, long long idx, bool inner) {
  Nb6 n;
  if (M2C_MP.ls_neumann) {   // zero-Neumann ghosts are the clamped image cells
    n.im = phi[clamped_idx(i - 1, j, k)]; n.ip = phi[clamped_idx(i + 1, j, k)];
    n.jm = phi[clamped_idx(i, j - 1, k)]; n.jp = phi[clamped_idx(i, j + 1, k)];
    n.km = phi[clamped_idx(i, j, k - 1)]; n.kp = phi[clamped_idx(i, j, k + 1)];
  } else if (inner) {
    const long long sy = M2C_G.nx, sz = (long long)M2C_G.nx * M2C_G.ny;
    n.im = phi[idx - 1]; n.ip = phi[idx + 1];
    n.jm = phi[idx - sy]; n.jp = phi[idx + sy];
    n.km = phi[idx - sz]; n.kp = phi[idx + sz];
  } else {
    n.im = phi_at(phi, i - 1, j, k); n.ip = phi_at(phi, i + 1, j, k);
    n.jm = phi_at(phi, i, j - 1, k); n.jp = phi_at(phi, i, j + 1, k);
    n.km = phi_at(phi, i, j, k - 1); n.kp = phi_at(phi, i, j, k + 1);
  }
  return n;
}

// velocity of a cell of the stage field S (primitive, or conservative -> ConservativeToPrimitive)
template <int IN_CONS>
__device__ __forceinline__ void cell_velocity(const double *S, const int8_t *ID, long long idx, double *u) {
  const Grid &g = M2C_G;
  if (IN_CONS) {
    const DevEos &m = M2C_EOS[ID[idx]];
    if (m.type == EOS_DUMMY) { u[0] = u[1] = u[2] = 0.0; return; }
    double invRho = 1.0 / S[idx];
    u[0] = S[g.N + idx] * invRho;
    u[1] = S[2 * g.N + idx] * invRho;
    u[2] = S[3 * g.N + idx] * invRho;
  } else {
    u[0] = S[g.N + idx];
    u[1] = S[2 * g.N + idx];
    u[2] = S[3 * g.N + idx];
  }
}

struct W4 { double w0, w1, w2, w3; };
__device__ __forceinline__ W4 fd_w(bool uni, const double *u, const double *tab) {
  W4 w;
  if (uni) { w.w0 = u[0]; w.w1 = u[1]; w.w2 = u[2]; w.w3 = u[3]; }
  else { w.w0 = tab[0]; w.w1 = tab[1]; w.w2 = tab[2]; w.w3 = tab[3]; }
  return w;
}

// upwinded third-order level-set residual -(u . grad phi) of interior cell (i,j,k)
// (LevelSetOperator.cpp:1066-1141 with the FD3 stencils of GradientCalculatorFD3.cpp:54-171).
// NB: 0 = ghosts by the boundary rules, 1 = every index within two cells is in the box (inner2),
// 2 = zero-Neumann boundaries (ghost = clamped image).
template <int NB>
__device__ inline double ls_residual(const double *phi, int i, int j, int k, long long idx, bool inner2,
                                     const double *u) {
  const Grid &g = M2C_G;
  const int NX = g.nx, NY = g.ny, NZ = g.nz;
  const long long sy = NX, sz = (long long)NX * NY;
  auto P = [&](int ii, int jj, int kk) -> double {
    if (NB == 1) return phi[idx + (ii - i) + (jj - j) * sy + (kk - k) * sz];
    if (NB == 2) return phi[clamped_idx(ii, jj, kk)];
    return phi_at(phi, ii, jj, kk);
  };
  const double p0 = phi[idx];
  double a, b, cc, d, e, f;
  if (i - 2 >= -1) {
    const W4 w = fd_w(inner2 && M2C_MP.fd_uni[0], M2C_MP.fdl_u[0], M2C_MP.fdl[0] + 4 * i);
    a = w.w0 * P(i - 2, j, k) + w.w1 * P(i - 1, j, k) + w.w2 * p0 + w.w3 * P(i + 1, j, k);
  } else
    a = (p0 - P(i - 1, j, k)) / (XC(0, i) - XC(0, i - 1));
  if (i + 2 <= NX) {
    const W4 w = fd_w(inner2 && M2C_MP.fd_uni[0], M2C_MP.fdr_u[0], M2C_MP.fdr[0] + 4 * i);
    b = w.w0 * P(i - 1, j, k) + w.w1 * p0 + w.w2 * P(i + 1, j, k) + w.w3 * P(i + 2, j, k);
  } else
    b = (P(i + 1, j, k) - p0) / (XC(0, i + 1) - XC(0, i));
  if (j - 2 >= -1) {
    const W4 w = fd_w(inner2 && M2C_MP.fd_uni[1], M2C_MP.fdl_u[1], M2C_MP.fdl[1] + 4 * j);
    cc = w.w0 * P(i, j - 2, k) + w.w1 * P(i, j - 1, k) + w.w2 * p0 + w.w3 * P(i, j + 1, k);
  } else
    cc = (p0 - P(i, j - 1, k)) / (XC(1, j) - XC(1, j - 1));
  if (j + 2 <= NY) {
    // the right-biased y coefficients are third order only where the column index i < NY-1
    // (GradientCalculatorFD3.cpp:415 tests i, not j)
    const bool third = i < NY - 1 && j < NY - 1;
    const W4 w = third ? fd_w(inner2 && M2C_MP.fd_uni[1], M2C_MP.fdr_u[1], M2C_MP.fdr[1] + 4 * j)
                       : fd_w(inner2 && M2C_MP.fd_uni_r1y, M2C_MP.fdr1y_u, M2C_MP.fdr1y + 4 * j);
    d = w.w0 * P(i, j - 1, k) + w.w1 * p0 + w.w2 * P(i, j + 1, k) + w.w3 * P(i, j + 2, k);
  } else
    d = (P(i, j + 1, k) - p0) / (XC(1, j + 1) - XC(1, j));
  if (k - 2 >= -1) {
    const W4 w = fd_w(inner2 && M2C_MP.fd_uni[2], M2C_MP.fdl_u[2], M2C_MP.fdl[2] + 4 * k);
    e = w.w0 * P(i, j, k - 2) + w.w1 * P(i, j, k - 1) + w.w2 * p0 + w.w3 * P(i, j, k + 1);
  } else
    e = (p0 - P(i, j, k - 1)) / (XC(2, k) - XC(2, k - 1));
  if (k + 2 <= NZ) {
    const W4 w = fd_w(inner2 && M2C_MP.fd_uni[2], M2C_MP.fdr_u[2], M2C_MP.fdr[2] + 4 * k);
    f = w.w0 * P(i, j, k - 1) + w.w1 * p0 + w.w2 * P(i, j, k + 1) + w.w3 * P(i, j, k + 2);
  } else
    f = (P(i, j, k + 1) - p0) / (XC(2, k + 1) - XC(2, k));
  double r = (u[0] >= 0 ? a * u[0] : b * u[0]) + (u[1] >= 0 ? cc * u[1] : d * u[1]) + (u[2] >= 0 ? e * u[2] : f * u[2]);
  return r * -1.0;
}

__device__ __forceinline__ bool step_active(const Clock *clk) { return clk->active != 0; }

// ---------------------------------------------------------------- per-step bookkeeping

__global__ void k_mp_step_begin(MpState *ms, const Clock *clk) {
  if (!step_active(clk)) return;
  ms->epoch += 1;
  ms->nface = 0;
  ms->nrem = 0;
  ms->nchg = 0;
  ms->nunres = 0;
  ms->nseq = 0;
}

__global__ void k_mp_face_reset(MpState *ms) {
  ms->nface = 0;
  ms->nfb = 0;
}

__global__ void k_mp_fb_reset(MpState *ms) { ms->nfb = 0; }

// interface faces of the current IDs: (dir, plus cell) with different active materials on the two
// sides; code = dir*Nf + face_key
// One thread per (i,j) column of the box extended by one cell on the high sides, walking k; cell
// (i,j,k) owns the three faces on its low sides.
__device__ __forceinline__ void iface_list_block(const int8_t *ID, unsigned *faces, MpState *ms, const Clock *clk,
                                                 int ignore_active, int bx, int by, int bz) {
  if (!ignore_active && !step_active(clk)) return;
  const Grid &g = M2C_G;
  if (g.mp_flux == 2) return;
  const long long Nf = (long long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  const int i = bx * SWEEP_BX + threadIdx.x, j = by * SWEEP_BY + threadIdx.y;
  if (i > g.nx || j > g.ny) return;
  auto id_at = [&](int a, int b, int c) -> int {
    if (a >= 0 && a < g.nx && b >= 0 && b < g.ny && c >= 0 && c < g.nz) return ID[cidx3(a, b, c)];
    return cell_id(ID, a, b, c);
  };
  const int k0 = bz * SWEEP_KB, k1 = min(g.nz + 1, k0 + SWEEP_KB);
  for (int k = k0; k < k1; k++) {
    const long long key = ((long long)k * (g.ny + 1) + j) * (g.nx + 1) + i;
    const int idp = id_at(i, j, k);
#pragma unroll
    for (int dir = 0; dir < 3; dir++) {
      const bool valid = (dir == 0) ? (j < g.ny && k < g.nz) : (dir == 1 ? (i < g.nx && k < g.nz) : (i < g.nx && j < g.ny));
      bool pred = false;
      if (valid) {
        const int idm = id_at(i - (dir == 0), j - (dir == 1), k - (dir == 2));
        pred = idm != idp && idm != g.inactive_id && idp != g.inactive_id;
      }
      // the lanes of a wave share j and the k range, so they reach the append together
      unsigned slot = wave_append(&ms->nface, pred);
      if (pred) faces[slot] = (unsigned)(dir * Nf + key);
    }
  }
}

__global__ __launch_bounds__(256) void k_iface_list(const int8_t *ID, unsigned *faces, MpState *ms, const Clock *clk,
                                                     int ignore_active) {
  iface_list_block(ID, faces, ms, clk, ignore_active, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z);
}

// ---------------------------------------------------------------- learned interface solver

// the two sides of a mesh-axis face as stiffened gases (normal velocity along axis d); false unless
// both materials are stiffened gases with admissible states
__device__ __forceinline__ bool sur_sides(const double *vm, int idm, const double *vp, int idp, int d, SgSide &L,
                                          SgSide &R) {
  const DevEos &ml = M2C_EOS[idm], &mr = M2C_EOS[idp];
  if (ml.type != EOS_SG || mr.type != EOS_SG) return false;
  L.rho = vm[0]; L.u = vm[1 + d]; L.p = vm[4]; L.gam = ml.gam; L.pinf = ml.pstiff;
  R.rho = vp[0]; R.u = vp[1 + d]; R.p = vp[4]; R.gam = mr.gam; R.pinf = mr.pstiff;
  return sg_side_init(L) && sg_side_init(R);
}

// Star states of a face from the network's prediction y: the predicted p* is polished by at most
// nmax Newton updates of the closed-form stiffened-gas relations and accepted when the velocity
// continuity residual is <= tol. Outputs follow ComputeRiemannSolution's conventions
// (ExactRiemannSolverBase.cpp:347-514): one-sided star states with the tangential velocity of their
// side, the state at xi = 0 (closed-form sampling, Toro 4.5) and its material. 1 = accepted.
__device__ inline int sur_face(int d, const double *vm, int idm, const double *vp, int idp, const float *y, int nmax,
                               double tol, double *Vs, int *mid, double *Vsm, double *Vsp) {
  SgSide L, R;
  if (!sur_sides(vm, idm, vp, idp, d, L, R)) return 0;
  float x[MLP_IN];
  SgScales sc;
  sg_features(L, R, x, &sc);
  double ps, res;
  int it;
  if (!sg_newton(L, R, sg_decode_p(sc, y[0]), nmax, tol, &ps, &res, &it)) return 0;
  double fl, dfl, fr, dfr;
  sg_wave(L, ps, fl, dfl);
  sg_wave(R, ps, fr, dfr);
  const double us = 0.5 * (L.u + R.u) + 0.5 * (fr - fl);
  const double rl = sg_star_rho(L, ps), rr = sg_star_rho(R, ps);
  if (!(rl > 0.0) || !(rr > 0.0) || !__builtin_isfinite(us)) return 0;
  double w[3];
  sg_sample0(L, R, ps, us, rl, rr, w);
  double utl[3] = {vm[1], vm[2], vm[3]}, utr[3] = {vp[1], vp[2], vp[3]};
  utl[d] = 0.0;
  utr[d] = 0.0;
  *mid = us >= 0 ? idm : idp;
  Vsm[0] = rl; Vsp[0] = rr; Vs[0] = w[0];
  Vsm[4] = ps; Vsp[4] = ps; Vs[4] = w[2];
  for (int q = 0; q < 3; q++) {
    const double un = q == d ? 1.0 : 0.0;
    Vsm[1 + q] = utl[q] + us * un;
    Vsp[1 + q] = utr[q] + us * un;
    const double ut = us > 0 ? utl[q] : (us < 0 ? utr[q] : 0.5 * (utl[q] + utr[q]));
    Vs[1 + q] = w[1] * un + ut;
  }
  return 1;
}

// network inputs of the listed interface faces (the first `cap`); faces without two stiffened
// gases get zeros and take the exact solver
template <int IN_CONS>
__global__ __launch_bounds__(256) void k_iface_feat(const double *S, const int8_t *ID, const unsigned *faces,
                                                     const MpState *ms, float *X, int cap, const Clock *clk,
                                                     int ignore_active) {
  if (!ignore_active && !step_active(clk)) return;
  const Grid &g = M2C_G;
  const long long Nf = (long long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  const unsigned n = min(ms->nface, (unsigned)cap);
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const unsigned code = faces[e];
    const int dir = (int)(code / Nf);
    const long long key = code - (long long)dir * Nf;
    const int i = (int)(key % (g.nx + 1));
    const long long r = key / (g.nx + 1);
    const int j = (int)(r % (g.ny + 1));
    const int k = (int)(r / (g.ny + 1));
    double vm[5], vp[5];
    int idm, idp;
    unsigned fl = 0;
    fetch_cell<IN_CONS, -1>(S, ID, i - (dir == 0), j - (dir == 1), k - (dir == 2), vm, idm, fl);
    fetch_cell<IN_CONS, -1>(S, ID, i, j, k, vp, idp, fl);
    float x[MLP_IN] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    SgSide L, R;
    SgScales sc;
    if (sur_sides(vm, idm, vp, idp, dir, L, R)) sg_features(L, R, x, &sc);
    for (int q = 0; q < MLP_IN; q++) X[(size_t)e * MLP_IN + q] = x[q];
  }
}

// ---------------------------------------------------------------- reconstructed face states

// The two reconstructed states of the DIR-face of cell (i,j,k) (between (i,j,k) - e_DIR, state vm1
// of material idm1, and (i,j,k), state vp0 of material idp0), formed as the stage kernel forms them:
// the slopes of the two cells from the four cells along the axis, the revert rule, and at a domain
// face the boundary rule on the inner state.
template <int DIR, int IN_CONS>
__device__ void face_states(const double *S, const int8_t *ID, int i, int j, int k, const double *vm1, int idm1,
                            const double *vp0, int idp0, double *fm, double *fp, const double *crphi) {
  const Grid &g = M2C_G;
  constexpr int di = DIR == 0, dj = DIR == 1, dk = DIR == 2;
  const int gn = DIR == 0 ? g.nx : (DIR == 1 ? g.ny : g.nz);
  const bool per = g.bc[2 * DIR] == BC_PERIODIC;
  double vm2[5], vp1[5];
  int idm2, idp1;
  unsigned fl = 0;
  fetch_cell<IN_CONS, -1>(S, ID, i - 2 * di, j - 2 * dj, k - 2 * dk, vm2, idm2, fl);
  fetch_cell<IN_CONS, -1>(S, ID, i + di, j + dj, k + dk, vp1, idp1, fl);
  const int gpl = DIR == 0 ? i : (DIR == 1 ? j : k);
  double sgm[5], sgp[5];
  // crphi: the level set at t_n when a constant-reconstruction depth is set
  const bool cm = conrec_cell(crphi, i - di, j - dj, k - dk), cp = conrec_cell(crphi, i, j, k);
  if (gpl == 0 && !per) {
    cell_slopes<DIR, double, true>(vm1, vp0, vp1, idm1, idp0, idp1, sgp, gpl);
    if (cp) zero5(sgp);
    face_state<-1, false, double>(vp0, sgp, idp0, fp);
    ghost_face<DIR, double>(g.bc[2 * DIR], fp, vm1, fm);
  } else if (gpl == gn && !per) {
    cell_slopes<DIR, double, true>(vm2, vm1, vp0, idm2, idm1, idp0, sgm, gpl - 1);
    if (cm) zero5(sgm);
    face_state<-1, true, double>(vm1, sgm, idm1, fm);
    ghost_face<DIR, double>(g.bc[2 * DIR + 1], fm, vp0, fp);
  } else {
    cell_slopes<DIR, double, true>(vm2, vm1, vp0, idm2, idm1, idp0, sgm, gpl - 1);
    cell_slopes<DIR, double, true>(vm1, vp0, vp1, idm1, idp0, idp1, sgp, gpl);
    if (cm) zero5(sgm);
    if (cp) zero5(sgp);
    face_state<-1, true, double>(vm1, sgm, idm1, fm);
    face_state<-1, false, double>(vp0, sgp, idp0, fp);
  }
}

// ---------------------------------------------------------------- interface Riemann pass

template <int FLUX, int DIR>
__device__ __forceinline__ void iface_fluxes(RSolver &S, const double *vm, const double *Vsm, int idm,
                                             const double *Vsp, const double *vp, int idp, const double *Vmid,
                                             int midid, double *lf1, double *lf2) {
  if (M2C_G.mp_flux == 1) {   // EXACT: physical flux of the state at xi = 0 to both sides
    phys_flux<double, -1, DIR>(M2C_EOS[midid], Vmid, lf1);
#pragma unroll
    for (int p = 0; p < 5; p++) lf2[p] = lf1[p];
  } else if (FLUX == FLUX_GODUNOV) {   // the numerical flux is itself an exact solve per side
    godunov_flux<DIR>(S, vm, Vsm, idm, lf1);
    godunov_flux<DIR>(S, Vsp, vp, idp, lf2);
  } else {
    numerical_flux<FLUX, -1, DIR>(vm, Vsm, idm, lf1);
    numerical_flux<FLUX, -1, DIR>(Vsp, vp, idp, lf2);
  }
}

// One thread per listed face (grid-stride). store != 0: keep the star states of this stage as
// the step's Riemann solutions (first stage of a step).
// PASS 0: the exact solver on every listed face. With the learned solver, PASS 1 takes the
// network's checked solution where it is accepted (trivial faces and faces beyond the prediction
// buffer go to the exact solver inline) and lists the rest in fbl; PASS 2 runs the exact solver on
// that list, so waves of the expensive iterative solver hold only the faces that need it.
// The pass over the listed faces as block blk of nblk blocks of 64 threads (k_riemann: the launch's
// grid; the multi-instance pass: the blocks that serve one instance).
template <int FLUX, int IN_CONS, int PASS>
__device__ __forceinline__ void riemann_faces(const double *S, const int8_t *ID, const unsigned *faces, MpState *ms,
                                              double *Fi, double *rsV, unsigned long long *rsMeta, int store,
                                              double *scratch, Status *st, const Clock *clk, int ignore_active,
                                              const float *pred, int sur_cap, int sur_nmax, double sur_tol,
                                              unsigned *fbl, const double *phin, const int nblk, const int blk) {
  if (!ignore_active && !step_active(clk)) return;
  const Grid &g = M2C_G;
  const long long Nf = (long long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  const int T = nblk * blockDim.x;
  const int tid = blk * blockDim.x + threadIdx.x;
  const unsigned n = PASS == 2 ? ms->nfb : ms->nface;
  RSolver S_;
  S_.mat = M2C_EOS;
  S_.cfg = M2C_MP.rcfg;
  S_.overflow = 0;
  const int cap = M2C_MP.path_cap;
  S_.path1.base = scratch + tid;
  S_.path1.cap = cap;
  S_.path1.stride = T;
  S_.path1.n = 0;
  S_.path3.base = scratch + (long long)3 * cap * T + tid;
  S_.path3.cap = cap;
  S_.path3.stride = T;
  S_.path3.n = 0;
  unsigned flags = 0;
  unsigned long long nerr = 0, nacc = 0, nfb = 0;
  const unsigned long long stamp = ms->epoch;
  for (unsigned t = tid; t < n; t += T) {
    const unsigned e = PASS == 2 ? fbl[t] : t;
    const unsigned code = faces[e];
    const int dir = (int)(code / Nf);
    const long long key = code - (long long)dir * Nf;
    const int i = (int)(key % (g.nx + 1));
    const long long r = key / (g.nx + 1);
    const int j = (int)(r % (g.ny + 1));
    const int k = (int)(r / (g.ny + 1));
    double vm[5], vp[5];
    int idm, idp;
    unsigned fl = 0;
    fetch_cell<IN_CONS, -1>(S, ID, i - (dir == 0), j - (dir == 1), k - (dir == 2), vm, idm, fl);
    fetch_cell<IN_CONS, -1>(S, ID, i, j, k, vp, idp, fl);
    flags |= fl;
    if (M2C_G.mp_recon) {   // linear reconstruction at the interface: vr of the minus cell, vl of the plus cell
      double fm[5], fp[5];
      const double *crphi = M2C_G.conrec_depth > 0.0 ? phin : nullptr;
      if (dir == 0) face_states<0, IN_CONS>(S, ID, i, j, k, vm, idm, vp, idp, fm, fp, crphi);
      else if (dir == 1) face_states<1, IN_CONS>(S, ID, i, j, k, vm, idm, vp, idp, fm, fp, crphi);
      else face_states<2, IN_CONS>(S, ID, i, j, k, vm, idm, vp, idp, fm, fp, crphi);
#pragma unroll
      for (int p = 0; p < 5; p++) { vm[p] = fm[p]; vp[p] = fp[p]; }
    }
    double nrm[3];
    interface_normal(dir, i, j, k, idp, idm, phin, nrm);
    double Vmid[5], Vsm[5], Vsp[5];
    int midid = 0;
    int err = 0;
    if (PASS == 1) {
      // a face with equal normal velocity and pressure on both sides is the exact solver's
      // immediate return (ExactRiemannSolverBase.cpp:129), taken as it is
      const bool trivial = vm[1 + dir] == vp[1 + dir] && vm[4] == vp[4];
      const bool in_buf = e < (unsigned)sur_cap;
      bool ok = false;
      if (!trivial && in_buf)
        ok = sur_face(dir, vm, idm, vp, idp, pred + (size_t)e * MLP_OUT, sur_nmax, sur_tol, Vmid, &midid, Vsm, Vsp);
      const bool defer = !trivial && in_buf && !ok;
      const unsigned slot = wave_append(&ms->nfb, defer);   // every active lane of the wave reaches it
      if (defer) {
        fbl[slot] = e;
        continue;
      }
      if (ok) nacc++;
      else err = riemann_solve(S_, nrm, vm, idm, vp, idp, Vmid, &midid, Vsm, Vsp);
    } else {
      if (PASS == 2) nfb++;
      err = riemann_solve(S_, nrm, vm, idm, vp, idp, Vmid, &midid, Vsm, Vsp);
    }
    if (err < 0) flags |= ST_BAD_STATE;
    if (err) nerr++;
    eos_clip<double, -1>(M2C_EOS[idm], Vsm);
    eos_clip<double, -1>(M2C_EOS[idp], Vsp);
    if (store && err == 0) {
#pragma unroll
      for (int p = 0; p < 5; p++) {
        rsV[((long long)dir * 10 + p) * Nf + key] = Vsm[p];
        rsV[((long long)dir * 10 + 5 + p) * Nf + key] = Vsp[p];
      }
      rsMeta[(long long)dir * Nf + key] = (stamp << 16) | ((unsigned long long)(idm & 0xff) << 8) |
                                          (unsigned long long)(idp & 0xff);
    }
    double lf1[5], lf2[5];
    if (dir == 0) iface_fluxes<FLUX, 0>(S_, vm, Vsm, idm, Vsp, vp, idp, Vmid, midid, lf1, lf2);
    else if (dir == 1) iface_fluxes<FLUX, 1>(S_, vm, Vsm, idm, Vsp, vp, idp, Vmid, midid, lf1, lf2);
    else iface_fluxes<FLUX, 2>(S_, vm, Vsm, idm, Vsp, vp, idp, Vmid, midid, lf1, lf2);
    double *f = Fi + (long long)dir * 10 * Nf + key;
#pragma unroll
    for (int p = 0; p < 5; p++) {
      f[p * Nf] = lf1[p];
      f[(5 + p) * Nf] = lf2[p];
    }
  }
  if (S_.overflow) flags |= ST_RIEMANN_PATH;
  if (nerr) atomicAdd(&st->riemann_err, nerr);
  if (nacc) atomicAdd(&ms->sur_accepted, nacc);
  if (nfb) atomicAdd(&ms->sur_fallback, nfb);
  if (flags & (ST_BAD_STATE | ST_RIEMANN_PATH | ST_CLIP_INTERMEDIATE)) atomicOr(&st->bits, flags);
}

template <int FLUX, int IN_CONS, int PASS>
__global__ __launch_bounds__(64) void k_riemann(const double *S, const int8_t *ID, const unsigned *faces,
                                                 MpState *ms, double *Fi, double *rsV, unsigned long long *rsMeta,
                                                 int store, double *scratch, Status *st, const Clock *clk,
                                                 int ignore_active, const float *pred, int sur_cap, int sur_nmax,
                                                 double sur_tol, unsigned *fbl, const double *phin) {
  riemann_faces<FLUX, IN_CONS, PASS>(S, ID, faces, ms, Fi, rsV, rsMeta, store, scratch, st, clk, ignore_active, pred,
                                     sur_cap, sur_nmax, sur_tol, fbl, phin, (int)gridDim.x, (int)blockIdx.x);
}

// ---------------------------------------------------------------- level-set stages

// out = t + (coef*dt) R(phiS; velocity of S), with t = 0*out + 1*X (LS_FIRST), a*X + b*Y (LS_AXBY)
// or X (LS_X). FINAL != 0: also the material-ID update of the step (MultiPhaseOperator.cpp:127-213)
// from sign(phi_n) = sign(X... phin) and the new value; out may alias phin in that case.
// NEUMANN: every level-set boundary is zero-Neumann, so every cell reads its neighbours at their
// clamped indices (for a cell whose stencil stays in the box those are its fixed offsets); one
// path keeps the kernel's registers low. Otherwise cells near the boundary take the ghost rules.
template <int IN_CONS, int MODE, int FINAL, int NEUMANN>
__global__ __launch_bounds__(256) void k_ls_stage(const double *S, const int8_t *IDn, const double *phiS,
                                                   const double *X, const double *Y, double *out, double a, double b,
                                                   double coef, const double *phin, int8_t *IDnew, unsigned *rem,
                                                   MpState *ms, const Clock *clk) {
  if (!step_active(clk)) return;
  const Grid &g = M2C_G;
  const double cdt = coef * clk->dt;
  sweep_cells([&](int i, int j, int k, long long idx, bool) {
    bool swept_out = false;
    double u[3];
    cell_velocity<IN_CONS>(S, IDn, idx, u);
    const bool inner2 = i >= 2 && i < g.nx - 2 && j >= 2 && j < g.ny - 2 && k >= 2 && k < g.nz - 2;
    double R;
    // cells whose stencil stays in the box read their neighbours at fixed offsets (the same values the
    // ghost rules would give them)
    if (NEUMANN) R = ls_residual<2>(phiS, i, j, k, idx, inner2, u);
    else if (inner2) R = ls_residual<1>(phiS, i, j, k, idx, true, u);
    else R = ls_residual<0>(phiS, i, j, k, idx, false, u);
    double t;
    if (MODE == LS_FIRST) {
      double x = X[idx];
      t = (x == 0.0) ? 0.0 * out[idx] + 1.0 * x : x;   // 0*old + 1*x differs from x only in the sign of zero
    } else if (MODE == LS_AXBY) {
      t = a * X[idx] + b * Y[idx];
    } else {
      t = X[idx];
    }
    const double pn = FINAL ? phin[idx] : 0.0;
    const double v = t + cdt * R;
    out[idx] = v;
    if (FINAL) {
      int id = IDn[idx];
      int nid = id;
      if (id != g.inactive_id && !(v * pn > 0.0)) {
        if (v < 0) nid = M2C_MP.ls_matid;
        else swept_out = true;
      }
      IDnew[idx] = (int8_t)nid;
      // every lane of the wave runs the same planes, so the append is reached together
      unsigned slot = wave_append(&ms->nrem, swept_out);
      if (swept_out) rem[slot] = (unsigned)idx;
    }
  });
}

// nodes swept out of the level-set subdomain take a neighbour's untracked ID; when the only
// candidate material is 0 every such node ends with ID 0 whatever the visiting order
// (MultiPhaseOperator.cpp:214-318, orphans -> 0)
__global__ __launch_bounds__(256) void k_id_remaining_single(const unsigned *rem, const MpState *ms, int8_t *IDnew,
                                                              const Clock *clk) {
  if (!step_active(clk)) return;
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < ms->nrem; e += gridDim.x * blockDim.x)
    IDnew[rem[e]] = 0;
}

// bitonic sort of up to MP_SORT_CAP 64-bit keys in LDS by one block of 1024 threads
__device__ void block_sort(unsigned long long *s, int n) {
  int m = 1;
  while (m < n) m <<= 1;
  for (int t = threadIdx.x + n; t < m; t += blockDim.x) s[t] = ~0ull;
  __syncthreads();
  for (int size = 2; size <= m; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < m; t += blockDim.x) {
        int o = t ^ stride;
        if (o > t) {
          bool up = (t & size) == 0;
          unsigned long long x = s[t], y = s[o];
          if ((x > y) == up) { s[t] = y; s[o] = x; }
        }
      }
      __syncthreads();
    }
}

// general case: the reference's passes in Int3 (i,j,k) order, one thread, Gauss-Seidel updates
__global__ __launch_bounds__(1024) void k_id_remaining_seq(const unsigned *rem, MpState *ms, int8_t *IDnew,
                                                            uint8_t *tagR, Status *st, const Clock *clk) {
  if (!step_active(clk)) return;
  const Grid &g = M2C_G;
  __shared__ unsigned long long s[MP_SORT_CAP];
  __shared__ uint8_t gone[MP_SORT_CAP];
  const int n = (int)ms->nrem;
  if (n == 0) return;
  if (n > MP_SORT_CAP) {
    if (threadIdx.x == 0) atomicOr(&st->bits, (unsigned)ST_LIST_OVERFLOW);
    return;
  }
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    long long idx = rem[t];
    int i = (int)(idx % g.nx);
    long long r = idx / g.nx;
    int j = (int)(r % g.ny), k = (int)(r / g.ny);
    s[t] = ((unsigned long long)i * g.ny + j) * g.nz + k;   // Int3(i,j,k) order (Vector3D.h:31-35)
    tagR[idx] = 1;
    gone[t] = 0;
  }
  __syncthreads();
  block_sort(s, n);
  if (threadIdx.x == 0) {
    int left = n;
    for (int iter = 0; iter < 100 && left > 0; iter++) {
      int progress = 0;
      for (int q = 0; q < n; q++) {
        if (gone[q]) continue;
        unsigned long long key = s[q];
        int k = (int)(key % g.nz);
        unsigned long long r = key / g.nz;
        int j = (int)(r % g.ny), i = (int)(r / g.ny);
        int best = -1;
        for (int nk = k - 1; nk <= k + 1; nk++)
          for (int nj = j - 1; nj <= j + 1; nj++)
            for (int ni = i - 1; ni <= i + 1; ni++) {
              if (ni < 0 || ni >= g.nx || nj < 0 || nj >= g.ny || nk < 0 || nk >= g.nz) continue;
              long long nb = cidx3(ni, nj, nk);
              if (tagR[nb]) continue;
              int nid = IDnew[nb];
              if (nid == g.inactive_id || nid == M2C_MP.ls_matid) continue;
              if (best < 0 || nid < best) best = nid;
            }
        if (best < 0) continue;
        long long me = cidx3(i, j, k);
        IDnew[me] = (int8_t)best;
        tagR[me] = 0;
        gone[q] = 1;
        left--;
        progress++;
      }
      if (!progress) {
        for (int q = 0; q < n; q++)
          if (!gone[q]) {
            unsigned long long key = s[q];
            int k = (int)(key % g.nz);
            unsigned long long r = key / g.nz;
            int j = (int)(r % g.ny), i = (int)(r / g.ny);
            IDnew[cidx3(i, j, k)] = 0;
          }
        break;
      }
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < n; t += blockDim.x) tagR[rem[t]] = 0;
}

// ---------------------------------------------------------------- reinitialisation

__global__ void k_reinit_begin(MpState *ms, const Clock *clk, int retry) {
  ReinitState &R = ms->rs;
  if (!retry) {
    int ts = clk->steps + 1;
    R.active = step_active(clk) && M2C_MP.reinit_freq > 0 && ts % M2C_MP.reinit_freq == 0;
    R.detected = 0;
  }
  R.retry = 0;
  R.done = 0;
  R.cur = 0;
  R.accepted = 0;
  R.acc_res_bits = 0ull;
  R.acc_dphi_bits = 0ull;
  R.iter = M2C_MP.reinit_maxits;
  R.maxres_bits = 0ull;
  R.dphi_bits = 0x7FF0000000000000ull;   // +inf
}

// first-layer tags (interior), sign function and the failsafe copy (LevelSetReinitializer.cpp:74-100)
__global__ __launch_bounds__(256) void k_reinit_tag(const double *phi, uint8_t *tag, double *sign, double *phibk,
                                                     MpState *ms) {
  if (!ms->rs.active) return;
  int det = 0;
  sweep_cells([&](int i, int j, int k, long long idx, bool inner) {
    const double p = phi[idx];
    const Nb6 n = phi_nb6(phi, i, j, k, idx, inner);
    int t = 0;
    if (p * n.im <= 0) t = 1;
    else if (p * n.ip <= 0) t = 1;
    else if (p * n.jm <= 0) t = 1;
    else if (p * n.jp <= 0) t = 1;
    else if (p * n.km <= 0) t = 1;
    else if (p * n.kp <= 0) t = 1;
    tag[idx] = (uint8_t)t;
    det |= t;
    const double f = 1.0 * cell_min_width(i, j, k);
    sign[idx] = p / dsqrt<double>(p * p + f * f);
    const double ob = phibk[idx];
    phibk[idx] = 0.0 * ob + 1.0 * p;
  });
  if (__any(det) && threadIdx.x == 0) atomicOr(&ms->rs.detected, 1);
}

__global__ void k_reinit_restore(double *phi, const double *phibk, const MpState *ms) {
  const Grid &g = M2C_G;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < g.N;
       idx += (long long)gridDim.x * blockDim.x)
    phi[idx] = 0.0 * phi[idx] + 1.0 * phibk[idx];
}

// pseudo-time residual of the reinitialisation equation (LevelSetReinitializer.cpp:953-1035) of cell
// (i,j,k) = idx from its value p and its six face neighbours n (ghosts already applied)
// T = double: the drop-in arithmetic; T = float: the mixed form's reinitialisation
template <typename T> struct NbT { T im, ip, jm, jp, km, kp; };

template <typename T>
__device__ __forceinline__ T reinit_rhs_t(T p, const NbT<T> &n, int tg, T sg, int i, int j, int k, T cfl, T *lres) {
  *lres = T(0);
  if (tg != 0) return T(0);
  const T dt = cfl * T(cell_min_width(i, j, k));
  T a, b, c, d, e, f;
  if (M2C_MP.dx_pow2[0]) { a = (p - n.im) * T(M2C_MP.inv_dx[0]); b = (n.ip - p) * T(M2C_MP.inv_dx[0]); }
  else { a = (p - n.im) / T(XC(0, i) - XC(0, i - 1)); b = (n.ip - p) / T(XC(0, i + 1) - XC(0, i)); }
  if (M2C_MP.dx_pow2[1]) { c = (p - n.jm) * T(M2C_MP.inv_dx[1]); d = (n.jp - p) * T(M2C_MP.inv_dx[1]); }
  else { c = (p - n.jm) / T(XC(1, j) - XC(1, j - 1)); d = (n.jp - p) / T(XC(1, j + 1) - XC(1, j)); }
  if (M2C_MP.dx_pow2[2]) { e = (p - n.km) * T(M2C_MP.inv_dx[2]); f = (n.kp - p) * T(M2C_MP.inv_dx[2]); }
  else { e = (p - n.km) / T(XC(2, k) - XC(2, k - 1)); f = (n.kp - p) / T(XC(2, k + 1) - XC(2, k)); }
  T ap = smax(a, T(0)), am = smin(a, T(0)), bp = smax(b, T(0)), bm = smin(b, T(0));
  T cp = smax(c, T(0)), cm = smin(c, T(0)), dp = smax(d, T(0)), dm = smin(d, T(0));
  T ep = smax(e, T(0)), em = smin(e, T(0)), fp = smax(f, T(0)), fm = smin(f, T(0));
  T lr;
  if (p >= T(0))
    lr = dsqrt<T>(smax(ap * ap, bm * bm) + smax(cp * cp, dm * dm) + smax(ep * ep, fm * fm)) - T(1);
  else
    lr = dsqrt<T>(smax(am * am, bp * bp) + smax(cm * cm, dp * dp) + smax(em * em, fp * fp)) - T(1);
  *lres = dabs(lr);
  if (dabs(lr) <= T(1)) return -dt * sg * lr;
  return -dt * sg * lr / dabs(lr);
}

__device__ __forceinline__ double reinit_rhs_nb(double p, const Nb6 &n, int tg, double sg, int i, int j, int k,
                                                double cfl, double *lres) {
  const NbT<double> m = {n.im, n.ip, n.jm, n.jp, n.km, n.kp};
  return reinit_rhs_t<double>(p, m, tg, sg, i, j, k, cfl, lres);
}

__device__ __forceinline__ double reinit_rhs(const double *phi, const uint8_t *tag, const double *sign, int i, int j,
                                             int k, long long idx, bool inner, double cfl, double *lres) {
  if (tag[idx] != 0) { *lres = 0.0; return 0.0; }
  return reinit_rhs_nb(phi[idx], phi_nb6(phi, i, j, k, idx, inner), 0, sign[idx], i, j, k, cfl, lres);
}

// ---- one whole reinitialisation iteration per launch (zero-Neumann level-set boundaries)
//
// The three RK3 stages of an iteration in one pass: a block owns RF_TX x RF_TY columns and marches
// up a chunk of RF_KC planes. Step jz stores one plane of phi (tile + 3-cell halo) and the sign and
// first-layer tag of one plane (tile + 2) into LDS rings, then evaluates stage 1 at plane jz + 2
// (tile + 2 halo), stage 2 at jz + 1 (tile + 1) and the new value at jz (tile), each stage reading its
// input from the ring of the previous one; the next step's global loads are issued before this
// step's arithmetic. Ghosts are the clamped image cells (the zero-Neumann rule of
// LevelSetOperator.cpp:934-1033 applied to Phi, Phi1 and the result), so a neighbour outside the box
// is read at its clamped index. The arithmetic per cell is reinit_rhs_nb with the stage combinations
// of k_reinit_A/B/C; the result goes to the other buffer of a pair (the residual test of stage 1 may
// discard the iteration).
constexpr int RF_TX = 32, RF_TY = 16, RF_KC = 16, RF_NT = RF_TX * RF_TY;
constexpr int RF_W0 = RF_TX + 6, RF_H0 = RF_TY + 6;
constexpr int RF_W1 = RF_TX + 4, RF_H1 = RF_TY + 4;
constexpr int RF_W2 = RF_TX + 2, RF_H2 = RF_TY + 2;
constexpr int RF_R0 = (RF_W0 * RF_H0 + RF_NT - 1) / RF_NT;   // load rounds of a phi plane
constexpr int RF_R1 = (RF_W1 * RF_H1 + RF_NT - 1) / RF_NT;   // ... of a sign / tag plane

template <typename T>
__global__ __launch_bounds__(RF_NT) void k_reinit_fused(T *__restrict__ P0, T *__restrict__ P1,
                                                         const uint8_t *__restrict__ tag,
                                                         const T *__restrict__ sign, MpState *ms) {
  ReinitState &Rs = ms->rs;
  if (!Rs.active || !Rs.detected || Rs.done) return;
  const Grid &g = M2C_G;
  const T *__restrict__ in = Rs.cur ? P1 : P0;
  T *__restrict__ out = Rs.cur ? P0 : P1;
  const T cfl = T(Rs.cfl);
  __shared__ T r0[4][RF_H0][RF_W0];        // phi, planes by p & 3
  __shared__ T r1[3][RF_H1][RF_W1];        // stage 1, planes by p % 3
  __shared__ T r2[3][RF_H2][RF_W2];        // stage 2
  __shared__ T rs[3][RF_H1][RF_W1];        // sign, planes by p % 3
  __shared__ uint8_t rt[3][RF_H1][RF_W1];  // first-layer tag
  const int tid = threadIdx.x;
  const int i0 = blockIdx.x * RF_TX, j0 = blockIdx.y * RF_TY;
  const int kb = blockIdx.z * RF_KC, ke = min(kb + RF_KC, g.nz);
  const long long plane = (long long)g.nx * g.ny;
  T mres = T(0), mdphi = T(0);
  // prefetched values of the next step: phi of plane jz + 3 (region 0), sign and tag of jz + 2 (region 1)
  T pf0[RF_R0], pfs[RF_R1];
  uint8_t pft[RF_R1];
  auto fetch = [&](int jz) {
    const int p0 = jz + 3, p1 = jz + 2;
    const bool l0 = p0 >= max(kb - 3, 0) && p0 < g.nz, l1 = p1 >= max(kb - 2, 0) && p1 < g.nz;
#pragma unroll
    for (int q = 0; q < RF_R0; q++) {
      const int r = tid + q * RF_NT;
      const int x = r % RF_W0, y = r / RF_W0;
      const int i = i0 - 3 + x, j = j0 - 3 + y;
      pf0[q] = (l0 && r < RF_W0 * RF_H0 && i >= 0 && i < g.nx && j >= 0 && j < g.ny)
                   ? in[p0 * plane + (long long)j * g.nx + i] : T(0);
    }
#pragma unroll
    for (int q = 0; q < RF_R1; q++) {
      const int r = tid + q * RF_NT;
      const int x = r % RF_W1, y = r / RF_W1;
      const int i = i0 - 2 + x, j = j0 - 2 + y;
      const bool ok = l1 && r < RF_W1 * RF_H1 && i >= 0 && i < g.nx && j >= 0 && j < g.ny;
      const long long idx = p1 * plane + (long long)j * g.nx + i;
      pfs[q] = ok ? sign[idx] : T(0);
      pft[q] = ok ? tag[idx] : (uint8_t)0;
    }
  };
  // value of stage field `ring` (halo h) at global (i, j) of ring plane slot q, clamped into the box (a
  // block whose 3-cell halo lies inside the box needs no clamping and skips no cell)
#define RF_AT(ring, h, q, ii, jj)                                                                       \
  (EDGE ? ring[q][clampi(jj, 0, g.ny - 1) - (j0 - h)][clampi(ii, 0, g.nx - 1) - (i0 - h)]              \
        : ring[q][(jj) - (j0 - h)][(ii) - (i0 - h)])
  const bool edge = i0 < 3 || j0 < 3 || i0 + RF_TX + 3 > g.nx || j0 + RF_TY + 3 > g.ny;
  auto march = [&](auto edge_c) {
  constexpr bool EDGE = decltype(edge_c)::value;
  // step jz stores phi(jz + 3) and evaluates stage 1 at jz + 2, stage 2 at jz + 1 and the result at jz;
  // the first steps fill the rings: phi from kb - 3, stage 1 from kb - 2, stage 2 from kb - 1
  fetch(kb - 6);
  for (int jz = kb - 6; jz < ke; jz++) {
    {
      const int p0 = jz + 3, p1 = jz + 2;
      const int q0 = p0 & 3, q1 = ((p1 % 3) + 3) % 3;
#pragma unroll
      for (int q = 0; q < RF_R0; q++) {
        const int r = tid + q * RF_NT;
        if (r < RF_W0 * RF_H0) r0[q0][r / RF_W0][r % RF_W0] = pf0[q];
      }
#pragma unroll
      for (int q = 0; q < RF_R1; q++) {
        const int r = tid + q * RF_NT;
        if (r < RF_W1 * RF_H1) { rs[q1][r / RF_W1][r % RF_W1] = pfs[q]; rt[q1][r / RF_W1][r % RF_W1] = pft[q]; }
      }
      (void)p0;
    }
    if (jz + 1 < ke) fetch(jz + 1);
    __syncthreads();
    // stage 1 at plane jz + 2: rphi1a = phi + 1.0 * R(phi)
    const int p1 = jz + 2;
    if (p1 >= max(kb - 2, 0) && p1 < g.nz) {
      const int q = p1 % 3, qc = p1 & 3, qm = max(p1 - 1, 0) & 3, qp = min(p1 + 1, g.nz - 1) & 3;
      for (int r = tid; r < RF_W1 * RF_H1; r += RF_NT) {
        const int x = r % RF_W1, y = r / RF_W1;
        const int i = i0 - 2 + x, j = j0 - 2 + y;
        if (EDGE && (i < 0 || i >= g.nx || j < 0 || j >= g.ny)) continue;
        const T pc = r0[qc][y + 1][x + 1];
        NbT<T> n;
        n.im = RF_AT(r0, 3, qc, i - 1, j); n.ip = RF_AT(r0, 3, qc, i + 1, j);
        n.jm = RF_AT(r0, 3, qc, i, j - 1); n.jp = RF_AT(r0, 3, qc, i, j + 1);
        n.km = r0[qm][y + 1][x + 1]; n.kp = r0[qp][y + 1][x + 1];
        T lres;
        const T R = reinit_rhs_t<T>(pc, n, rt[q][y][x], rs[q][y][x], i, j, p1, cfl, &lres);
        if (lres > mres) mres = lres;
        r1[q][y][x] = pc + T(1) * R;
      }
    }
    __syncthreads();
    // stage 2 at plane jz + 1: rphi1b = (0.25 rphi1a + 0.75 phi) + 0.25 R(rphi1a)
    const int p2 = jz + 1;
    if (p2 >= max(kb - 1, 0) && p2 < g.nz) {
      const int qc = p2 % 3, qm = max(p2 - 1, 0) % 3, qp = min(p2 + 1, g.nz - 1) % 3;
      const int q0 = p2 & 3;
      for (int r = tid; r < RF_W2 * RF_H2; r += RF_NT) {
        const int x = r % RF_W2, y = r / RF_W2;
        const int i = i0 - 1 + x, j = j0 - 1 + y;
        if (EDGE && (i < 0 || i >= g.nx || j < 0 || j >= g.ny)) continue;
        const T pc = r1[qc][y + 1][x + 1];
        NbT<T> n;
        n.im = RF_AT(r1, 2, qc, i - 1, j); n.ip = RF_AT(r1, 2, qc, i + 1, j);
        n.jm = RF_AT(r1, 2, qc, i, j - 1); n.jp = RF_AT(r1, 2, qc, i, j + 1);
        n.km = r1[qm][y + 1][x + 1]; n.kp = r1[qp][y + 1][x + 1];
        T lres;
        const T R = reinit_rhs_t<T>(pc, n, rt[qc][y + 1][x + 1], rs[qc][y + 1][x + 1], i, j, p2, cfl, &lres);
        const T t = T(0.25) * pc + T(0.75) * r0[q0][y + 2][x + 2];
        r2[qc][y][x] = t + T(0.25) * R;
      }
    }
    __syncthreads();
    // the new value at plane jz: (1/3 phi + 2/3 rphi1b) + 2/3 R(rphi1b), and the relative change
    if (jz >= kb) {
      const int qc = jz % 3, qm = max(jz - 1, 0) % 3, qp = min(jz + 1, g.nz - 1) % 3, q0 = jz & 3;
      for (int r = tid; r < RF_TX * RF_TY; r += RF_NT) {
        const int x = r % RF_TX, y = r / RF_TX;
        const int i = i0 + x, j = j0 + y;
        if (EDGE && (i >= g.nx || j >= g.ny)) continue;
        const long long idx = jz * plane + (long long)j * g.nx + i;
        const T pc = r2[qc][y + 1][x + 1];
        NbT<T> n;
        n.im = RF_AT(r2, 1, qc, i - 1, j); n.ip = RF_AT(r2, 1, qc, i + 1, j);
        n.jm = RF_AT(r2, 1, qc, i, j - 1); n.jp = RF_AT(r2, 1, qc, i, j + 1);
        n.km = r2[qm][y + 1][x + 1]; n.kp = r2[qp][y + 1][x + 1];
        T lres;
        const T R = reinit_rhs_t<T>(pc, n, rt[qc][y + 2][x + 2], rs[qc][y + 2][x + 2], i, j, jz, cfl, &lres);
        const T po = r0[q0][y + 3][x + 3];
        const T t = T(1.0 / 3.0) * po + T(2.0 / 3.0) * pc;
        const T pn = t + T(2.0 / 3.0) * R;
        out[idx] = pn;
        T le = dabs(pn - po);
        if (pn != T(0)) le /= dabs(pn);
        mdphi = smax(mdphi, le);
      }
    }
    __syncthreads();
  }
  };
  if (edge) march(std::true_type{});
  else march(std::false_type{});
#undef RF_AT
  // block maxima of the stage-1 residual and the relative change
  __shared__ T red[2][RF_NT / 64];
  for (int off = 32; off > 0; off >>= 1) {
    mres = smax(mres, __shfl_xor(mres, off, 64));
    mdphi = smax(mdphi, __shfl_xor(mdphi, off, 64));
  }
  if ((tid & 63) == 0) { red[0][tid >> 6] = mres; red[1][tid >> 6] = mdphi; }
  __syncthreads();
  if (tid == 0) {
    T a = red[0][0], b = red[1][0];
    for (int w = 1; w < RF_NT / 64; w++) { a = smax(a, red[0][w]); b = smax(b, red[1][w]); }
    if (a > T(0)) atomic_max_pos(&Rs.acc_res_bits, double(a));
    atomic_max_pos(&Rs.acc_dphi_bits, double(b));
  }
}

// after iteration `it` of k_reinit_fused: the residual test of stage 1 (below tolerance: the
// iteration is discarded and the loop ends, LevelSetReinitializer.cpp:109-117), else the new values
// become current and the relative-change test may end the loop (:146-151)
__global__ void k_reinit_decide(MpState *ms, int it) {
  ReinitState &Rs = ms->rs;
  if (!Rs.active || !Rs.detected || Rs.done) return;
  const double res = bits_to_d(Rs.acc_res_bits), dphi = bits_to_d(Rs.acc_dphi_bits);
  Rs.acc_res_bits = 0ull;
  Rs.acc_dphi_bits = 0ull;
  if (res < M2C_MP.reinit_tol) { Rs.done = 1; Rs.iter = it; return; }
  Rs.cur ^= 1;
  Rs.accepted += 1;
  Rs.dphi_bits = __double_as_longlong(dphi);
  if (dphi < M2C_MP.reinit_tol) { Rs.done = 1; Rs.iter = it; }
}

// the failsafe of LevelSetReinitializer.cpp:144-169 after the fused iterations
__global__ void k_reinit_end_fused(MpState *ms, Status *st) {
  ReinitState &Rs = ms->rs;
  if (!Rs.active || !Rs.detected) return;
  if (!Rs.done) Rs.iter = M2C_MP.reinit_maxits;
  ms->reinit_iters += Rs.iter;
  if (Rs.iter < M2C_MP.reinit_maxits) return;
  if (bits_to_d(Rs.dphi_bits) < 0.02) return;
  if (Rs.cfl < 0.02) {
    atomicOr(&st->bits, (unsigned)ST_REINIT_FAIL);
    return;
  }
  Rs.cfl = Rs.cfl / 1.5;
  Rs.retry = 1;
}

// the mixed form: phi and the sign function to FP32 for the iterations
__global__ void k_reinit_to_f32(const double *phi, float *phif, const double *sign, float *signf, const MpState *ms) {
  const ReinitState &Rs = ms->rs;
  if (!Rs.active || !Rs.detected) return;
  const Grid &g = M2C_G;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < g.N;
       idx += (long long)gridDim.x * blockDim.x) {
    phif[idx] = (float)phi[idx];
    signf[idx] = (float)sign[idx];
  }
}

// ... and the FP32 result back into phi when an iteration was accepted
__global__ void k_reinit_settle_f32(double *phi, const float *phif0, const float *phif1, const MpState *ms) {
  const ReinitState &Rs = ms->rs;
  if (!Rs.active || !Rs.detected || !Rs.accepted) return;
  const float *src = Rs.cur ? phif1 : phif0;
  const Grid &g = M2C_G;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < g.N;
       idx += (long long)gridDim.x * blockDim.x)
    phi[idx] = (double)src[idx];
}

// the current buffer of the pair back into phi
__global__ void k_reinit_settle(double *phi, const double *P1, const MpState *ms) {
  const ReinitState &Rs = ms->rs;
  if (!Rs.active || !Rs.detected || !Rs.cur) return;
  const Grid &g = M2C_G;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < g.N;
       idx += (long long)gridDim.x * blockDim.x)
    phi[idx] = P1[idx];
}

// RK3 stage 1 of reinitialisation iteration `it`: rphi1a = (0*rphi1a + 1*phi) + 1*R(phi)
__global__ __launch_bounds__(256) void k_reinit_A(const double *__restrict__ phi, double *__restrict__ rphi1a,
                                                   const uint8_t *__restrict__ tag,
                                                   const double *__restrict__ sign, MpState *ms, int it) {
  ReinitState &Rs = ms->rs;
  if (!Rs.active || !Rs.detected || Rs.done) return;
  if (it > 0 && bits_to_d(Rs.dphi_bits) < M2C_MP.reinit_tol) {   // converged after iteration it-1
    if (first_thread()) { Rs.done = 1; Rs.iter = it - 1; }
    return;
  }
  const double cfl = Rs.cfl;
  double mx = 0.0;
  sweep_cells([&](int i, int j, int k, long long idx, bool inner) {
    double lres;
    double R = reinit_rhs(phi, tag, sign, i, j, k, idx, inner, cfl, &lres);
    if (lres > mx) mx = lres;
    double x = phi[idx];
    double t = (x == 0.0) ? 0.0 * rphi1a[idx] + 1.0 * x : x;
    rphi1a[idx] = t + 1.0 * R;
  });
  block_max_to(&Rs.maxres_bits, mx, true);
}

// stage 2: stop if the residual of stage 1 is below tolerance, else
// rphi1b = (0.25*rphi1a + 0.75*phi) + 0.25*R(rphi1a)
__global__ __launch_bounds__(256) void k_reinit_B(const double *__restrict__ phi, const double *__restrict__ rphi1a,
                                                   double *__restrict__ rphi1b, const uint8_t *__restrict__ tag,
                                                   const double *__restrict__ sign, MpState *ms, int it) {
  ReinitState &Rs = ms->rs;
  if (!Rs.active || !Rs.detected || Rs.done) return;
  if (bits_to_d(Rs.maxres_bits) < M2C_MP.reinit_tol) {
    if (first_thread()) { Rs.done = 1; Rs.iter = it; }
    return;
  }
  if (first_thread()) Rs.dphi_bits = 0ull;
  const double cfl = Rs.cfl;
  sweep_cells([&](int i, int j, int k, long long idx, bool inner) {
    double lres;
    double R = reinit_rhs(rphi1a, tag, sign, i, j, k, idx, inner, cfl, &lres);
    double t = 0.25 * rphi1a[idx] + 0.75 * phi[idx];
    rphi1b[idx] = t + 0.25 * R;
  });
}

// stage 3: phi = (1/3*phi + 2/3*rphi1b) + 2/3*R(rphi1b), and the maximum relative change
__global__ __launch_bounds__(256) void k_reinit_C(double *__restrict__ phi, const double *__restrict__ rphi1b,
                                                   const uint8_t *__restrict__ tag, const double *__restrict__ sign,
                                                   MpState *ms) {
  ReinitState &Rs = ms->rs;
  if (!Rs.active || !Rs.detected || Rs.done) return;
  if (first_thread()) Rs.maxres_bits = 0ull;
  const double cfl = Rs.cfl;
  double mx = 0.0;
  sweep_cells([&](int i, int j, int k, long long idx, bool inner) {
    double lres;
    double R = reinit_rhs(rphi1b, tag, sign, i, j, k, idx, inner, cfl, &lres);
    const double po = phi[idx];
    double t = 1.0 / 3.0 * po + 2.0 / 3.0 * rphi1b[idx];
    const double pn = t + 2.0 / 3.0 * R;
    phi[idx] = pn;
    double le = dabs(pn - po);
    if (pn != 0) le /= dabs(pn);
    mx = smax(mx, le);
  });
  block_max_to(&Rs.dphi_bits, mx, false);
}

// after the last iteration: convergence on the last relative change, else the failsafe of
// LevelSetReinitializer.cpp:144-169 (accept below 0.02, retry with cfl/1.5, fail below cfl 0.02)
__global__ void k_reinit_end(MpState *ms, Status *st) {
  ReinitState &Rs = ms->rs;
  if (!Rs.active || !Rs.detected) return;
  if (!Rs.done) {
    if (bits_to_d(Rs.dphi_bits) < M2C_MP.reinit_tol) Rs.iter = M2C_MP.reinit_maxits - 1;
    else Rs.iter = M2C_MP.reinit_maxits;
  }
  ms->reinit_iters += Rs.iter;
  if (Rs.iter < M2C_MP.reinit_maxits) return;
  if (bits_to_d(Rs.dphi_bits) < 0.02) return;
  if (Rs.cfl < 0.02) {
    atomicOr(&st->bits, (unsigned)ST_REINIT_FAIL);
    return;
  }
  Rs.cfl = Rs.cfl / 1.5;
  Rs.retry = 1;
}

// ---------------------------------------------------------------- state update after interface motion

__global__ __launch_bounds__(256) void k_changed_list(const int8_t *IDn, const int8_t *IDnew, unsigned *chg,
                                                       int *cidx, MpState *ms, const Clock *clk) {
  if (!step_active(clk)) return;
  const Grid &g = M2C_G;
  for (long long base = (long long)blockIdx.x * blockDim.x; base < g.N; base += (long long)gridDim.x * blockDim.x) {
    long long idx = base + threadIdx.x;
    bool pred = idx < g.N && IDnew[idx] != IDn[idx] && IDnew[idx] != g.inactive_id;
    unsigned slot = wave_append(&ms->nchg, pred);
    if (pred) {
      chg[slot] = (unsigned)idx;
      cidx[idx] = (int)slot;
    }
  }
}

// neighbour state for the state update: the value final at this point of the in-order sweep
// (a changed, resolved, final lower neighbour contributes its new value; everything else the
// field after the last stage, ghosts by the boundary rules)
struct NbView {
  const double *V;
  const int8_t *IDn;
  const int *cidx;
  const int *lvl;
  const int8_t *cnt;
  const double *Vnew;
  unsigned ncap;
};

// LocalUpdateByRiemannSolutions (MultiPhaseOperator.cpp:916-1028). `round` >= 0: lower changed
// neighbours must be final before `round` (returns -2 if one is not); round < 0: round 0, the
// pre-update field is used everywhere and *dep reports whether a lower changed neighbour was read.
__device__ int local_update(const NbView &nv, const double *rsV, const unsigned long long *rsMeta,
                            unsigned long long stamp, int i, int j, int k, int id, bool upwind, double *v, int round,
                            bool *dep) {
  const Grid &g = M2C_G;
  const long long Nf = (long long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  int counter = 0;
  double sum_weight = 0.0;
#pragma unroll 1
  for (int s = 0; s < 6; s++) {
    const int dir = s >> 1, plus = s & 1;
    const int fi = i + (dir == 0 && plus), fj = j + (dir == 1 && plus), fk = k + (dir == 2 && plus);
    const long long key = ((long long)fk * (g.ny + 1) + fj) * (g.nx + 1) + fi;
    const unsigned long long meta = rsMeta[(long long)dir * Nf + key];
    if ((meta >> 16) != stamp) continue;
    const int rid = plus ? (int)(meta & 0xff) : (int)((meta >> 8) & 0xff);
    if (rid != id) continue;
    const int ni = i + (dir == 0 ? (plus ? 1 : -1) : 0);
    const int nj = j + (dir == 1 ? (plus ? 1 : -1) : 0);
    const int nk = k + (dir == 2 ? (plus ? 1 : -1) : 0);
    double vn[5];
    bool used_new = false;
    if (!plus && ni >= 0 && nj >= 0 && nk >= 0) {
      const long long nidx = cidx3(ni, nj, nk);
      const int ci = nv.cidx[nidx];
      if (ci >= 0) {   // a changed lower neighbour
        if (round < 0) {
          if (upwind) *dep = true;
        } else {
          const int l = nv.lvl[ci];
          if (l < 0 || l >= round) return -2;
          if (nv.cnt[ci] > 0) {
#pragma unroll
            for (int p = 0; p < 5; p++) vn[p] = nv.Vnew[(long long)5 * ci + p];
            used_new = true;
          }
        }
      }
    }
    if (!used_new) {
      int dummy;
      unsigned fl = 0;
      fetch_cell<0, -1>(nv.V, nv.IDn, ni, nj, nk, vn, dummy, fl);
    }
    const int comp = dir + 1;
    if (upwind && !(plus ? vn[comp] < 0 : vn[comp] > 0)) continue;
    double w;
    if (upwind) {
      double nrm = dsqrt<double>(vn[1] * vn[1] + vn[2] * vn[2] + vn[3] * vn[3]);
      w = (plus ? -vn[comp] : vn[comp]) / nrm;
    } else {
      w = 1.0;
    }
    sum_weight += w;
    const double *rs = rsV + ((long long)dir * 10 + (plus ? 5 : 0)) * Nf + key;
    if (counter == 0) {
#pragma unroll
      for (int p = 0; p < 5; p++) v[p] = w * rs[p * Nf];
    } else {
#pragma unroll
      for (int p = 0; p < 5; p++) v[p] += w * rs[p * Nf];
    }
    counter++;
  }
  if (sum_weight > 0.0) {
    double inv = 1.0 / sum_weight;
#pragma unroll
    for (int p = 0; p < 5; p++) v[p] *= inv;
  }
  return counter;
}

__device__ __forceinline__ void load_v(const double *V, long long idx, double *v) {
#pragma unroll
  for (int p = 0; p < 5; p++) v[p] = V[p * M2C_G.N + idx];
}

// upwinded pass, then the pass without upwinding (MultiPhaseOperator.cpp:881-892)
__device__ int update_cell(const NbView &nv, const double *rsV, const unsigned long long *rsMeta,
                           unsigned long long stamp, long long idx, int id, double *v, int round, bool *dep) {
  const int i = (int)(idx % M2C_G.nx);
  const long long r = idx / M2C_G.nx;
  const int j = (int)(r % M2C_G.ny), k = (int)(r / M2C_G.ny);
  load_v(nv.V, idx, v);
  int cnt = local_update(nv, rsV, rsMeta, stamp, i, j, k, id, true, v, round, dep);
  if (cnt == -2) return -2;
  if (cnt == 0) cnt = local_update(nv, rsV, rsMeta, stamp, i, j, k, id, false, v, round, dep);
  return cnt;
}

// round 0: every changed cell from the pre-update field
__global__ __launch_bounds__(256) void k_state_round0(NbView nv, const unsigned *chg, const int8_t *IDnew,
                                                       const double *rsV, const unsigned long long *rsMeta,
                                                       MpState *ms, int *lvl, int8_t *cnt, double *Vnew,
                                                       const Clock *clk) {
  if (!step_active(clk)) return;
  const unsigned n = ms->nchg;
  const unsigned long long stamp = ms->epoch;
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const long long idx = chg[e];
    double v[5];
    bool dep = false;
    int c = update_cell(nv, rsV, rsMeta, stamp, idx, IDnew[idx], v, -1, &dep);
    cnt[e] = (int8_t)(c > 0 ? 1 : 0);
    lvl[e] = dep ? -1 : 0;
#pragma unroll
    for (int p = 0; p < 5; p++) Vnew[(long long)5 * e + p] = v[p];
  }
}

// dependency round `round`: cells whose changed lower neighbours are all final
__global__ __launch_bounds__(256) void k_state_round(NbView nv, const unsigned *chg, const int8_t *IDnew,
                                                      const double *rsV, const unsigned long long *rsMeta,
                                                      MpState *ms, int *lvl, int8_t *cnt, double *Vnew, int round,
                                                      const Clock *clk) {
  if (!step_active(clk)) return;
  const unsigned n = ms->nchg;
  const unsigned long long stamp = ms->epoch;
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    if (lvl[e] >= 0) continue;
    const long long idx = chg[e];
    double v[5];
    bool dep = false;
    int c = update_cell(nv, rsV, rsMeta, stamp, idx, IDnew[idx], v, round, &dep);
    if (c == -2) continue;
#pragma unroll
    for (int p = 0; p < 5; p++) Vnew[(long long)5 * e + p] = v[p];
    cnt[e] = (int8_t)(c > 0 ? 1 : 0);
    __threadfence();
    lvl[e] = round;
  }
}

// cells still open after the rounds, in (k,j,i) order by one thread
__global__ __launch_bounds__(1024) void k_state_seq(NbView nv, const unsigned *chg, const int8_t *IDnew,
                                                     const double *rsV, const unsigned long long *rsMeta, MpState *ms,
                                                     int *lvl, int8_t *cnt, double *Vnew, Status *st,
                                                     const Clock *clk) {
  if (!step_active(clk)) return;
  __shared__ unsigned long long s[MP_SORT_CAP];
  __shared__ int ns;
  if (threadIdx.x == 0) ns = 0;
  __syncthreads();
  const unsigned n = ms->nchg;
  for (unsigned e = threadIdx.x; e < n; e += blockDim.x)
    if (lvl[e] < 0) {
      int q = atomicAdd(&ns, 1);
      if (q < MP_SORT_CAP) s[q] = ((unsigned long long)chg[e] << 32) | (unsigned long long)e;
    }
  __syncthreads();
  const int m = ns;
  if (m == 0) return;
  if (m > MP_SORT_CAP) {
    if (threadIdx.x == 0) atomicOr(&st->bits, (unsigned)ST_LIST_OVERFLOW);
    return;
  }
  block_sort(s, m);
  if (threadIdx.x == 0) {
    ms->nseq += m;
    const int big = 1 << 30;
    for (int q = 0; q < m; q++) {
      const unsigned e = (unsigned)(s[q] & 0xffffffffull);
      const long long idx = (long long)(s[q] >> 32);
      double v[5];
      bool dep = false;
      int c = update_cell(nv, rsV, rsMeta, ms->epoch, idx, IDnew[idx], v, big, &dep);
      for (int p = 0; p < 5; p++) Vnew[(long long)5 * e + p] = v[p];
      cnt[e] = (int8_t)(c > 0 ? 1 : 0);
      lvl[e] = big - 1;   // final for every later cell of this pass
    }
  }
}

// new values into V; unresolved cells are listed for FixUnresolvedNodes
__global__ __launch_bounds__(256) void k_state_commit(double *V, const unsigned *chg, const int8_t *cnt,
                                                       const double *Vnew, unsigned *unres, MpState *ms,
                                                       const Clock *clk) {
  if (!step_active(clk)) return;
  const unsigned n = ms->nchg;
  for (unsigned base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    unsigned e = base + threadIdx.x;
    bool un = false;
    if (e < n) {
      const long long idx = chg[e];
      if (cnt[e] > 0) {
#pragma unroll
        for (int p = 0; p < 5; p++) V[p * M2C_G.N + idx] = Vnew[(long long)5 * e + p];
      } else {
        un = true;
      }
    }
    unsigned slot = wave_append(&ms->nunres, un);
    if (un) unres[slot] = chg[e];
  }
}

// FixUnresolvedNodes with the failsafe density (MultiPhaseOperator.cpp:1156-1320), in (k,j,i)
// order by one thread; `reset` carries over from node to node as in the reference
__global__ __launch_bounds__(1024) void k_fix_unresolved(double *V, const int8_t *IDnew, const int *cidx,
                                                          const int8_t *cnt, const unsigned *unres, MpState *ms,
                                                          Status *st, const Clock *clk) {
  if (!step_active(clk)) return;
  const Grid &g = M2C_G;
  __shared__ unsigned long long s[MP_SORT_CAP];
  const int n = (int)ms->nunres;
  if (n == 0) return;
  if (n > MP_SORT_CAP) {
    if (threadIdx.x == 0) atomicOr(&st->bits, (unsigned)ST_LIST_OVERFLOW);
    return;
  }
  for (int t = threadIdx.x; t < n; t += blockDim.x) s[t] = unres[t];
  __syncthreads();
  block_sort(s, n);
  if (threadIdx.x != 0) return;
  ms->unresolved_total += n;
  bool reset = false;
  for (int q = 0; q < n; q++) {
    const long long me = (long long)s[q];
    const int i = (int)(me % g.nx);
    const long long r = me / g.nx;
    const int j = (int)(r % g.ny), k = (int)(r / g.ny);
    const int myid = IDnew[me];
    double v[5];
    load_v(V, me, v);
    const double x0[3] = {XC(0, i), XC(1, j), XC(2, k)};
    double sum_weight = 0.0, sum_weight2 = 0.0, vtmp[5] = {0, 0, 0, 0, 0};
    for (int nk = k - 1; nk <= k + 1; nk++)
      for (int nj = j - 1; nj <= j + 1; nj++)
        for (int ni = i - 1; ni <= i + 1; ni++) {
          if (ni < 0 || ni >= g.nx || nj < 0 || nj >= g.ny || nk < 0 || nk >= g.nz) continue;
          const long long nb = cidx3(ni, nj, nk);
          if (IDnew[nb] != myid) continue;
          if (nb == me) continue;
          const int ci = cidx[nb];
          if (ci >= 0 && cnt[ci] == 0) continue;   // also unresolved
          double vn[5];
          load_v(V, nb, vn);
          double v1[3] = {vn[1], vn[2], vn[3]};
          double v1n = dsqrt<double>(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
          if (v1n != 0) { double iv = 1.0 / v1n; v1[0] *= iv; v1[1] *= iv; v1[2] *= iv; }
          double d[3] = {x0[0] - XC(0, ni), x0[1] - XC(1, nj), x0[2] - XC(2, nk)};
          double dn = dsqrt<double>(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
          double idn = 1.0 / dn;
          d[0] *= idn; d[1] *= idn; d[2] *= idn;
          double dot = d[0] * v1[0] + d[1] * v1[1] + d[2] * v1[2];
          double weight = smax(0.0, dot);
          if (weight > 0) {
            sum_weight += weight;
            if (reset) for (int p = 0; p < 5; p++) v[p] += weight * vn[p];
            else { for (int p = 0; p < 5; p++) v[p] = weight * vn[p]; reset = true; }
          }
          for (int p = 0; p < 5; p++) vtmp[p] += 1.0 / dn * vn[p];
          sum_weight2 += 1.0 / dn;
        }
    if (sum_weight > 0) {
      double inv = 1.0 / sum_weight;
      for (int p = 0; p < 5; p++) v[p] *= inv;
    } else if (sum_weight2 > 0) {
      for (int p = 0; p < 5; p++) v[p] = vtmp[p] / sum_weight2;
    } else {
      double density = 0.0;
      for (int layer = 1; layer <= 10; layer++) {
        for (int nk = k - layer; nk <= k + layer; nk++)
          for (int nj = j - layer; nj <= j + layer; nj++)
            for (int ni = i - layer; ni <= i + layer; ni++) {
              if (ni < 0 || ni >= g.nx || nj < 0 || nj >= g.ny || nk < 0 || nk >= g.nz) continue;
              const long long nb = cidx3(ni, nj, nk);
              if (IDnew[nb] != myid) continue;
              if (nb == me) continue;
              const int ci = cidx[nb];
              if (ci >= 0 && cnt[ci] == 0) continue;
              double d[3] = {XC(0, ni) - x0[0], XC(1, nj) - x0[1], XC(2, nk) - x0[2]};
              double dist = dsqrt<double>(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
              sum_weight += 1.0 / dist;
              density += (1.0 / dist) * V[nb];
            }
        if (sum_weight > 0) { v[0] = density / sum_weight; break; }
      }
      if (sum_weight == 0) v[0] = M2C_MP.failsafe_density[myid];
    }
    for (int p = 0; p < 5; p++) V[p * g.N + me] = v[p];
  }
}

// ClipDensityAndPressure + CheckState of the cells whose material changed (unchanged cells were
// clipped with the same material by the last stage); IDs committed; per-step maps cleared
__global__ __launch_bounds__(256) void k_state_finish(double *V, int8_t *IDn, const int8_t *IDnew, const unsigned *chg,
                                                       int *cidx, MpState *ms, Status *st, const Clock *clk) {
  if (!step_active(clk)) return;
  const Grid &g = M2C_G;
  const unsigned n = ms->nchg;
  unsigned flags = 0;
  unsigned long long nclip = 0;
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const long long idx = chg[e];
    const int id = IDnew[idx];
    double v[5];
    load_v(V, idx, v);
    const DevEos &m = M2C_EOS[id];
    if (eos_clip<double, -1>(m, v)) nclip++;
    if (eos_bad<double, -1>(m, v)) flags |= ST_BAD_STATE;
#pragma unroll
    for (int p = 0; p < 5; p++) V[p * g.N + idx] = v[p];
    IDn[idx] = (int8_t)id;
    cidx[idx] = -1;
  }
  if (nclip) atomicAdd(&st->clipped, nclip);
  if (flags) atomicOr(&st->bits, flags);
}

// ComputeRiemannSolution for a batch of independent problems: problem t is solved along the unit
// normal dir[3t..3t+2] (ExactRiemannSolverBase.h:35-39 takes it as double *dir)
__global__ __launch_bounds__(64) void k_riemann_batch(int n, const double *dir, const double *Vm, const int *idm,
                                                       const double *Vp, const int *idp, double *Vs, int *id,
                                                       double *Vsm, double *Vsp, int *status, double *scratch,
                                                       int nthreads) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= nthreads) return;
  RSolver S_;
  S_.mat = M2C_EOS;
  S_.cfg = M2C_MP.rcfg;
  S_.overflow = 0;
  const int cap = M2C_MP.path_cap;
  S_.path1.base = scratch + tid; S_.path1.cap = cap; S_.path1.stride = nthreads; S_.path1.n = 0;
  S_.path3.base = scratch + (long long)3 * cap * nthreads + tid; S_.path3.cap = cap; S_.path3.stride = nthreads;
  S_.path3.n = 0;
  for (int t = tid; t < n; t += nthreads) {
    if (idm[t] < 0 || idm[t] >= M2C_G.nmat || idp[t] < 0 || idp[t] >= M2C_G.nmat) {
      // material index outside the table (callers passing device arrays are not checked on the host)
#pragma unroll
      for (int q = 0; q < 5; q++) Vs[5 * t + q] = Vsm[5 * t + q] = Vsp[5 * t + q] = 0.0;
      id[t] = -1;
      status[t] = -1;
      continue;
    }
    double nrm[3] = {dir[3 * t], dir[3 * t + 1], dir[3 * t + 2]};
    int mid = 0;
    S_.overflow = 0;
    int r = riemann_solve(S_, nrm, Vm + 5 * t, idm[t], Vp + 5 * t, idp[t], Vs + 5 * t, &mid, Vsm + 5 * t, Vsp + 5 * t);
    id[t] = mid;
    status[t] = S_.overflow ? 2 : r;
  }
}

// ComputeOneSidedRiemannSolution for a batch of independent problems (FACE = 0): problem t has the
// unit normal dir[3t..3t+2] and the wall velocity uw[3t..3t+2]. FACE = 1: the fluid side of a face
// normal to axis ax[t] cut by an embedded wall (SpaceOperator.cpp:2438-2484, 2569-2611, 2696-2737):
// the direction of GetNormalForOneSidedRiemann from the wall normal dir[3t..], fb[t] = +1 when the
// fluid cell is on the minus side, the clipped star state in Vsm and the numerical flux along the
// axis, F(V, V*) or F(V*, V), in F. Vs is zero when the wall recedes (id = -1).
template <int FLUX, int FACE>
__global__ __launch_bounds__(64) void k_onesided_batch(int n, const double *dir, const double *Vm, const int *idl,
                                                        const double *uw, const int *ax, const int *fb, int mode,
                                                        double *Vs, int *id, double *Vsm, double *F, int *status,
                                                        double *scratch, int nthreads) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= nthreads) return;
  RSolver S_;
  S_.mat = M2C_EOS;
  S_.cfg = M2C_MP.rcfg;
  S_.overflow = 0;
  const int cap = M2C_MP.path_cap;
  S_.path1.base = scratch + tid; S_.path1.cap = cap; S_.path1.stride = nthreads; S_.path1.n = 0;
  S_.path3.base = scratch + (long long)3 * cap * nthreads + tid; S_.path3.cap = cap; S_.path3.stride = nthreads;
  S_.path3.n = 0;
  for (int t = tid; t < n; t += nthreads) {
    double nrm[3], vs[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, vsm[5];
    const double *v = Vm + 5 * t;
    const int m = idl[t];
    if (FACE) onesided_normal(ax[t], fb[t], mode, dir + 3 * t, nrm);
    else { nrm[0] = dir[3 * t]; nrm[1] = dir[3 * t + 1]; nrm[2] = dir[3 * t + 2]; }
    int mid = 0;
    S_.overflow = 0;
    int r = onesided_solve(S_, nrm, v, m, uw + 3 * t, vs, &mid, vsm);
    if (FACE) {   // an invalid input state (r < 0) leaves the flux at zero
      double f[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      if (r >= 0) {
        eos_clip<double, -1>(M2C_EOS[m], vsm);
        const bool minus = fb[t] > 0;
        const double *a = minus ? v : vsm, *b = minus ? vsm : v;
        if (FLUX == FLUX_GODUNOV) {
          if (ax[t] == 0) godunov_flux<0>(S_, a, b, m, f);
          else if (ax[t] == 1) godunov_flux<1>(S_, a, b, m, f);
          else godunov_flux<2>(S_, a, b, m, f);
        } else if (ax[t] == 0) numerical_flux<FLUX, -1, 0>(a, b, m, f);
        else if (ax[t] == 1) numerical_flux<FLUX, -1, 1>(a, b, m, f);
        else numerical_flux<FLUX, -1, 2>(a, b, m, f);
      }
      for (int q = 0; q < 5; q++) F[5 * t + q] = f[q];
    }
    for (int q = 0; q < 5; q++) { Vs[5 * t + q] = vs[q]; Vsm[5 * t + q] = vsm[q]; }
    id[t] = mid;
    status[t] = S_.overflow ? 2 : r;
  }
}

// ---------------------------------------------------------------- faces cut by embedded walls

// One side of a wall-cut face: the one-sided problem of the side's state v (material m) along
// GetNormalForOneSidedRiemann's direction (fb = +1: the cell is on the minus side), the clipped star
// state, then the numerical flux along DIR between the state and the star state (F(v, v*) on the
// minus side, F(v*, v) on the plus side). Where the reference exits (a negative status) the flux
// is zero. Returns 1 on a solver error (counted as the sweep counts them).
template <int FLUX, int DIR>
__device__ __forceinline__ int wall_side(RSolver &S, int fb, int mode, const double *nwall, const double *vwall,
                                         const double *v, int m, double *F, unsigned &flags) {
  double nrm[3], vs[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, vsm[5];
  onesided_normal(DIR, fb, mode, nwall, nrm);
  int mid = 0;
  S.overflow = 0;
  const int r = onesided_solve(S, nrm, v, m, vwall, vs, &mid, vsm);
  if (S.overflow) flags |= ST_RIEMANN_PATH;
#pragma unroll
  for (int p = 0; p < 5; p++) F[p] = 0.0;
  if (r >= 0) {
    eos_clip<double, -1>(M2C_EOS[m], vsm);
    const double *a = fb > 0 ? v : vsm, *b = fb > 0 ? vsm : v;
    if (FLUX == FLUX_GODUNOV) godunov_flux<DIR>(S, a, b, m, F);
    else numerical_flux<FLUX, -1, DIR>(a, b, m, F);
  }
  return r != 0;
}

// The two fluxes of every mesh face cut by an embedded wall, for the stage that follows
// (SpaceOperator.cpp:2438-2484 i-1/2, :2569-2611 j-1/2, :2696-2737 k-1/2): faces[t] = dir * Nf + key
// (the plus cell's face key), wall[12t..] = nwallf, vwallf, nwallb, vwallb. An active minus cell
// solves against the forward intersection and gives localflux1 (Fi components 0-4), an active plus
// cell against the backward one and gives localflux2 (components 5-9); an inactive side's flux is
// zero. The sides' states are the cell values (ReconstructionAtInterface CONSTANT) or the
// reconstructed face states (LINEAR), which are the cell values too where slopes next to cut faces
// are zeroed. One thread per face (grid-stride), the rarefaction path cache in the thread's slice.
template <int FLUX, int IN_CONS>
__global__ __launch_bounds__(64) void k_wall_faces(const double *S, const int8_t *ID, const unsigned *faces,
                                                   const double *wall, int n, int mode, int recon, double *Fi,
                                                   long long Nf, double *scratch, Status *st, const Clock *clk,
                                                   int ignore_active) {
  if (!ignore_active && !step_active(clk)) return;
  const Grid &g = M2C_G;
  const int T = gridDim.x * blockDim.x;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  RSolver S_;
  S_.mat = M2C_EOS;
  S_.cfg = M2C_MP.rcfg;
  S_.overflow = 0;
  const int cap = M2C_MP.path_cap;
  S_.path1.base = scratch + tid; S_.path1.cap = cap; S_.path1.stride = T; S_.path1.n = 0;
  S_.path3.base = scratch + (long long)3 * cap * T + tid; S_.path3.cap = cap; S_.path3.stride = T; S_.path3.n = 0;
  unsigned flags = 0;
  unsigned long long nerr = 0;
  for (int t = tid; t < n; t += T) {
    const unsigned code = faces[t];
    const int dir = (int)(code / Nf);
    const long long key = code - (long long)dir * Nf;
    const int i = (int)(key % (g.nx + 1));
    const long long r = key / (g.nx + 1);
    const int j = (int)(r % (g.ny + 1));
    const int k = (int)(r / (g.ny + 1));
    double vm[5], vp[5];
    int idm, idp;
    unsigned fl = 0;
    fetch_cell<IN_CONS, -1>(S, ID, i - (dir == 0), j - (dir == 1), k - (dir == 2), vm, idm, fl);
    fetch_cell<IN_CONS, -1>(S, ID, i, j, k, vp, idp, fl);
    flags |= fl;
    if (recon == 1 && !g.slope_zero) {
      double fm[5], fp[5];
      if (dir == 0) face_states<0, IN_CONS>(S, ID, i, j, k, vm, idm, vp, idp, fm, fp, nullptr);
      else if (dir == 1) face_states<1, IN_CONS>(S, ID, i, j, k, vm, idm, vp, idp, fm, fp, nullptr);
      else face_states<2, IN_CONS>(S, ID, i, j, k, vm, idm, vp, idp, fm, fp, nullptr);
#pragma unroll
      for (int p = 0; p < 5; p++) { vm[p] = fm[p]; vp[p] = fp[p]; }
    }
    const double *w = wall + 12 * (long long)t;
    double F1[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, F2[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (dir == 0) {
      if (idm != g.inactive_id) nerr += wall_side<FLUX, 0>(S_, 1, mode, w, w + 3, vm, idm, F1, flags);
      if (idp != g.inactive_id) nerr += wall_side<FLUX, 0>(S_, -1, mode, w + 6, w + 9, vp, idp, F2, flags);
    } else if (dir == 1) {
      if (idm != g.inactive_id) nerr += wall_side<FLUX, 1>(S_, 1, mode, w, w + 3, vm, idm, F1, flags);
      if (idp != g.inactive_id) nerr += wall_side<FLUX, 1>(S_, -1, mode, w + 6, w + 9, vp, idp, F2, flags);
    } else {
      if (idm != g.inactive_id) nerr += wall_side<FLUX, 2>(S_, 1, mode, w, w + 3, vm, idm, F1, flags);
      if (idp != g.inactive_id) nerr += wall_side<FLUX, 2>(S_, -1, mode, w + 6, w + 9, vp, idp, F2, flags);
    }
    double *f = Fi + (long long)dir * 10 * Nf + key;
#pragma unroll
    for (int p = 0; p < 5; p++) {
      f[p * Nf] = F1[p];
      f[(5 + p) * Nf] = F2[p];
    }
  }
  if (nerr) atomicAdd(&st->riemann_err, nerr);
  if (flags) atomicOr(&st->bits, flags);
}

// ---------------------------------------------------------------- Godunov flux on the grid

// FluxFcnGodunov (FluxFcnGodunov.h:37-60) on every face of direction DIR between two cells of one
// active material. The face states are formed as the stage kernel forms them: the slopes of the
// face's two cells from the four cells along the axis, the revert rule, and at a domain face the
// boundary rule on the inner state. Then the exact solver runs along the axis and G[p * nf + f] takes
// the physical flux of the state at xi = 0. Other faces are left as they are (the stage kernel
// does not read them). One thread per face (grid-stride) with the rarefaction path cache in the
// thread's slice of the scratch buffer; faces are indexed (k, j, i) with n_DIR + 1 planes along DIR.
template <int DIR, int IN_CONS>
__global__ __launch_bounds__(64) void k_godunov_faces(const double *S, const int8_t *ID, double *G, long long nf,
                                                       double *scratch, Status *st, const Clock *clk,
                                                       int ignore_active, const double *crphi) {
  if (!ignore_active && !step_active(clk)) return;
  const Grid &g = M2C_G;
  const int T = gridDim.x * blockDim.x;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  RSolver S_;
  S_.mat = M2C_EOS;
  S_.cfg = M2C_MP.rcfg;
  S_.overflow = 0;
  const int cap = M2C_MP.path_cap;
  S_.path1.base = scratch + tid; S_.path1.cap = cap; S_.path1.stride = T; S_.path1.n = 0;
  S_.path3.base = scratch + (long long)3 * cap * T + tid; S_.path3.cap = cap; S_.path3.stride = T; S_.path3.n = 0;
  constexpr int di = DIR == 0, dj = DIR == 1, dk = DIR == 2;
  const int fx = g.nx + di, fy = g.ny + dj;
  unsigned flags = 0;
  for (long long f = tid; f < nf; f += T) {
    const int i = (int)(f % fx);
    const long long r = f / fx;
    const int j = (int)(r % fy);
    const int k = (int)(r / fy);
    double vm1[5], vp0[5];
    int idm1, idp0;
    unsigned fl = 0;
    fetch_cell<IN_CONS, -1>(S, ID, i - di, j - dj, k - dk, vm1, idm1, fl);
    fetch_cell<IN_CONS, -1>(S, ID, i, j, k, vp0, idp0, fl);
    if (idm1 != idp0 || idp0 == g.inactive_id) continue;
    double fm[5], fp[5];
    face_states<DIR, IN_CONS>(S, ID, i, j, k, vm1, idm1, vp0, idp0, fm, fp, crphi);
    double F[5];
    S_.overflow = 0;
    if (godunov_flux<DIR>(S_, fm, fp, idp0, F) < 0) flags |= ST_BAD_STATE;
    if (S_.overflow) flags |= ST_RIEMANN_PATH;
#pragma unroll
    for (int p = 0; p < 5; p++) G[p * nf + f] = F[p];
  }
  if (flags) atomicOr(&st->bits, flags);
}

// FluxFcnGodunov for n independent face-state pairs along axis dir (the pointwise flux batch)
template <int DIR>
__global__ __launch_bounds__(64) void k_godunov_batch(int n, const double *Vm, const double *Vp, int id, double *F,
                                                       double *scratch, int nthreads) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= nthreads) return;
  RSolver S_;
  S_.mat = M2C_EOS;
  S_.cfg = M2C_MP.rcfg;
  S_.overflow = 0;
  const int cap = M2C_MP.path_cap;
  S_.path1.base = scratch + tid; S_.path1.cap = cap; S_.path1.stride = nthreads; S_.path1.n = 0;
  S_.path3.base = scratch + (long long)3 * cap * nthreads + tid; S_.path3.cap = cap; S_.path3.stride = nthreads;
  S_.path3.n = 0;
  for (int t = tid; t < n; t += nthreads) {
    double vm[5], vp[5], f[5];
    for (int p = 0; p < 5; p++) { vm[p] = Vm[5 * (size_t)t + p]; vp[p] = Vp[5 * (size_t)t + p]; }
    godunov_flux<DIR>(S_, vm, vp, id, f);
    for (int p = 0; p < 5; p++) F[5 * (size_t)t + p] = f[p];
  }
}

// pointwise batch of the learned solver: network inputs of every problem
__global__ void k_sur_batch_feat(int n, const int *dir, const double *Vm, const int *idm, const double *Vp,
                                 const int *idp, float *X) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    float x[MLP_IN] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    SgSide L, R;
    SgScales sc;
    if (sur_sides(Vm + 5 * t, idm[t], Vp + 5 * t, idp[t], dir[t], L, R)) sg_features(L, R, x, &sc);
    for (int q = 0; q < MLP_IN; q++) X[(size_t)t * MLP_IN + q] = x[q];
  }
}

// ... then the prediction, checked (PASS 1; failures listed in fbl), and the exact solver on the
// listed problems (PASS 2); path[t] = 1 network, 0 exact
template <int PASS>
__global__ __launch_bounds__(64) void k_sur_batch(int n, const int *dir, const double *Vm, const int *idm,
                                                  const double *Vp, const int *idp, const float *Y, int nmax,
                                                  double tol, double *Vs, int *id, double *Vsm, double *Vsp,
                                                  int *status, int *path, double *scratch, int nthreads,
                                                  unsigned *fbl, unsigned *nfb) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= nthreads) return;
  if (PASS == 1) {
    for (int t = tid; t < n; t += nthreads) {
      int mid = 0;
      const double *vm = Vm + 5 * t, *vp = Vp + 5 * t;
      const bool trivial = vm[1 + dir[t]] == vp[1 + dir[t]] && vm[4] == vp[4];   // the exact solver's immediate return
      const bool ok = !trivial && sur_face(dir[t], vm, idm[t], vp, idp[t], Y + (size_t)t * MLP_OUT, nmax, tol,
                                           Vs + 5 * t, &mid, Vsm + 5 * t, Vsp + 5 * t);
      const unsigned slot = wave_append(nfb, !ok);
      if (ok) {
        id[t] = mid;
        status[t] = 0;
        path[t] = 1;
      } else {
        fbl[slot] = (unsigned)t;
      }
    }
    return;
  }
  const int nl = (int)*nfb;
  RSolver S_;
  S_.mat = M2C_EOS;
  S_.cfg = M2C_MP.rcfg;
  S_.overflow = 0;
  const int cap = M2C_MP.path_cap;
  S_.path1.base = scratch + tid; S_.path1.cap = cap; S_.path1.stride = nthreads; S_.path1.n = 0;
  S_.path3.base = scratch + (long long)3 * cap * nthreads + tid; S_.path3.cap = cap; S_.path3.stride = nthreads;
  S_.path3.n = 0;
  for (int q = tid; q < nl; q += nthreads) {
    const int t = (int)fbl[q];
    int mid = 0;
    double nrm[3] = {0.0, 0.0, 0.0};
    nrm[dir[t]] = 1.0;
    S_.overflow = 0;
    int r = riemann_solve(S_, nrm, Vm + 5 * t, idm[t], Vp + 5 * t, idp[t], Vs + 5 * t, &mid, Vsm + 5 * t, Vsp + 5 * t);
    id[t] = mid;
    status[t] = S_.overflow ? 2 : r;
    path[t] = 0;
  }
}

static int hchk(hipError_t e) { return e == hipSuccess ? 0 : 2; }

#ifdef M2C_IFACE_MULTI
// ---------------------------------------------------------------- multi-instance interface pass
//
// Each block finds its instance, copies the instance's grid, multiphase parameters and materials
// into the LDS tables the interface code reads in this unit, and runs the instance's pass as one
// of its nblk blocks (so a thread solves the faces, and uses the scratch slice, it would in the
// instance's own launch of nblk blocks).
template <typename T> __device__ __forceinline__ void lds_copy(T *dst, const T *src) {
  static_assert(sizeof(T) % 4 == 0, "word copy");
  unsigned *d = reinterpret_cast<unsigned *>(dst);
  const unsigned *q = reinterpret_cast<const unsigned *>(src);
  for (int w = threadIdx.x; w < (int)(sizeof(T) / 4); w += blockDim.x) d[w] = q[w];
}

// the face lists of every instance: counters reset (one thread per instance), then the list sweep,
// block b of the launch serving the instance whose list blocks hold it
__global__ void k_iface_reset_multi(const IfaceInst *tab, int ninst, int relist) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ninst) return;
  MpState *ms = tab[t].ms;
  if (relist) ms->nface = 0;
  ms->nfb = 0;
}

__global__ __launch_bounds__(256) void k_iface_list_multi(const IfaceInst *tab, int ninst) {
  int inst = 0;
  while (inst + 1 < ninst && (int)blockIdx.x >= tab[inst + 1].lblk0) inst++;
  const IfaceInst &I = tab[inst];
  lds_copy(&s_igrid, &I.grid);
  __syncthreads();
  const int b = (int)blockIdx.x - I.lblk0;
  if (b < 0 || b >= I.lnblk) return;
  const int nbx = (s_igrid.nx + SWEEP_BX) / SWEEP_BX, nby = (s_igrid.ny + SWEEP_BY) / SWEEP_BY;
  iface_list_block(I.ID, I.faces, I.ms, I.clk, 0, b % nbx, (b / nbx) % nby, b / (nbx * nby));
}

template <int FLUX, int IN_CONS>
__global__ __launch_bounds__(64) void k_riemann_multi(const IfaceInst *tab, int ninst, int stage, int store) {
  int inst = 0;
  while (inst + 1 < ninst && (int)blockIdx.x >= tab[inst + 1].blk0) inst++;
  const IfaceInst &I = tab[inst];
  lds_copy(&s_igrid, &I.grid);
  lds_copy(&s_imp, &I.mp);
  for (int m = 0; m < 8; m++) lds_copy(&s_ieos[m], &I.eos[m]);
  __syncthreads();
  const int blk = (int)blockIdx.x - I.blk0;
  if (blk < 0 || blk >= I.nblk) return;
  riemann_faces<FLUX, IN_CONS, 0>(I.S[stage], I.ID, I.faces, I.ms, I.Fi, I.rsV, I.rsMeta, store && I.store, I.scratch,
                                  I.st, I.clk, 0, nullptr, 0, 0, 0.0, nullptr, I.phin, I.nblk, blk);
}
#endif

#ifdef M2C_IFACE_MULTI
}  // namespace imulti
#endif
}  // namespace m2c

#ifdef M2C_IFACE_MULTI
using namespace m2c;
using namespace m2c::imulti;

extern "C" int m2c_mp_launch_iface_list_multi(hipStream_t s, const IfaceInst *dtab, int ninst, int nlist_blocks,
                                              int relist) {
  if (ninst <= 0 || nlist_blocks <= 0) return 1;
  hipLaunchKernelGGL(k_iface_reset_multi, dim3((ninst + 63) / 64), dim3(64), 0, s, dtab, ninst, relist);
  if (relist)
    hipLaunchKernelGGL(k_iface_list_multi, dim3(nlist_blocks), dim3(SWEEP_BX, SWEEP_BY), 0, s, dtab, ninst);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

extern "C" int m2c_mp_launch_iface_multi(hipStream_t s, const IfaceInst *dtab, int ninst, int nblk_total, int flux,
                                         int in_cons, int stage, int store) {
  if (ninst <= 0 || nblk_total <= 0 || stage < 0 || stage > 2) return 1;
  const dim3 gr(nblk_total), bl(64);
#define M2C_RM(F)                                                                                               \
  if (in_cons) hipLaunchKernelGGL((k_riemann_multi<F, 1>), gr, bl, 0, s, dtab, ninst, stage, store);           \
  else hipLaunchKernelGGL((k_riemann_multi<F, 0>), gr, bl, 0, s, dtab, ninst, stage, store)
  if (flux == FLUX_HLLC) { M2C_RM(FLUX_HLLC); }
  else if (flux == FLUX_ROE) { M2C_RM(FLUX_ROE); }
  else if (flux == FLUX_LLF) { M2C_RM(FLUX_LLF); }
  else return 1;
#undef M2C_RM
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
#else

using namespace m2c;

extern "C" int m2c_mp_upload(const Grid &grid, const DevEos *eos, const MpParams &mp) {
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_grid), &grid, sizeof(Grid)) != hipSuccess) return 2;
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_eos), eos, sizeof(DevEos) * 8) != hipSuccess) return 2;
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_mp), &mp, sizeof(MpParams)) != hipSuccess) return 2;
  return 0;
}

// the same three tables copied on a stream (sources in pinned host memory that outlive the copy):
// the switch between the two-material instances of a batch
extern "C" int m2c_mp_upload_async(hipStream_t s, const Grid *grid, const DevEos *eos, const MpParams *mp) {
  if (hipMemcpyToSymbolAsync(HIP_SYMBOL(c_grid), grid, sizeof(Grid), 0, hipMemcpyHostToDevice, s) != hipSuccess)
    return 2;
  if (hipMemcpyToSymbolAsync(HIP_SYMBOL(c_eos), eos, sizeof(DevEos) * 8, 0, hipMemcpyHostToDevice, s) != hipSuccess)
    return 2;
  if (hipMemcpyToSymbolAsync(HIP_SYMBOL(c_mp), mp, sizeof(MpParams), 0, hipMemcpyHostToDevice, s) != hipSuccess)
    return 2;
  return 0;
}

static int nblocks(long long n, int bs, int cap) {
  long long b = (n + bs - 1) / bs;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (int)b;
}

extern "C" int m2c_mp_iface_list_blocks(int nx, int ny, int nz) {
  const dim3 g = sweep_grid(nx + 1, ny + 1, nz + 1);
  return (int)(g.x * g.y * g.z);
}

extern "C" int m2c_mp_launch_iface_list(const MpLaunch &L, int ignore_active, int relist) {
  if (relist) {
    hipLaunchKernelGGL(k_mp_face_reset, dim3(1), dim3(1), 0, L.stream, L.ms);
    hipLaunchKernelGGL(k_iface_list, sweep_grid(L.nx + 1, L.ny + 1, L.nz + 1), dim3(SWEEP_BX, SWEEP_BY), 0, L.stream,
                       L.ID, L.faces, L.ms, L.clk, ignore_active);
  } else {
    hipLaunchKernelGGL(k_mp_fb_reset, dim3(1), dim3(1), 0, L.stream, L.ms);
  }
  return hchk(hipGetLastError());
}

extern "C" int m2c_mp_launch_iface(const MpLaunch &L, const double *S, int in_cons, int store, int ignore_active,
                                   int relist) {
  if (int rc = m2c_mp_launch_iface_list(L, ignore_active, relist)) return rc;
  dim3 gr(L.riemann_threads / 64), bl(64);
  const float *pred = nullptr;
  if (L.sur_on) {   // network inputs and predictions of the listed faces (count read on the device)
    const dim3 gf(nblocks(L.sur_cap, 256, 4096));
    if (in_cons)
      hipLaunchKernelGGL((k_iface_feat<1>), gf, dim3(256), 0, L.stream, S, L.ID, L.faces, L.ms, L.surX, L.sur_cap,
                         L.clk, ignore_active);
    else
      hipLaunchKernelGGL((k_iface_feat<0>), gf, dim3(256), 0, L.stream, S, L.ID, L.faces, L.ms, L.surX, L.sur_cap,
                         L.clk, ignore_active);
    if (m2c_mlp_launch_infer(L.sur_w, L.stream, L.surX, L.surY, L.sur_cap, &L.ms->nface)) return 2;
    pred = L.surY;
  }
#define M2C_RIEMANN_P(F, IC, P)                                                                                 \
  hipLaunchKernelGGL((k_riemann<F, IC, P>), gr, bl, 0, L.stream, S, L.ID, L.faces, L.ms, L.Fi, L.rsV, L.rsMeta,    \
                     store, L.scratch, L.st, L.clk, ignore_active, pred, L.sur_cap, L.sur_nmax, L.sur_tol, L.surFb, \
                     L.phin)
#define M2C_RIEMANN_IC(F, IC)                                                                                   \
  if (pred) {                                                                                                   \
    M2C_RIEMANN_P(F, IC, 1);                                                                                    \
    M2C_RIEMANN_P(F, IC, 2);                                                                                    \
  } else {                                                                                                      \
    M2C_RIEMANN_P(F, IC, 0);                                                                                    \
  }
#define M2C_RIEMANN(F)                                                                                          \
  if (in_cons) { M2C_RIEMANN_IC(F, 1) } else { M2C_RIEMANN_IC(F, 0) }
  if (L.flux == FLUX_HLLC) { M2C_RIEMANN(FLUX_HLLC) }
  else if (L.flux == FLUX_ROE) { M2C_RIEMANN(FLUX_ROE) }
  else if (L.flux == FLUX_GODUNOV) {   // no learned solver with this flux
    if (in_cons) { M2C_RIEMANN_P(FLUX_GODUNOV, 1, 0); } else { M2C_RIEMANN_P(FLUX_GODUNOV, 0, 0); }
  }
  else { M2C_RIEMANN(FLUX_LLF) }
#undef M2C_RIEMANN
#undef M2C_RIEMANN_IC
#undef M2C_RIEMANN_P
  return hchk(hipGetLastError());
}

extern "C" int m2c_mp_launch_riemann_batch(hipStream_t s, int n, const double *dir, const double *Vm, const int *idm,
                                           const double *Vp, const int *idp, double *Vs, int *id, double *Vsm,
                                           double *Vsp, int *status, double *scratch, int nthreads) {
  hipLaunchKernelGGL(k_riemann_batch, dim3((nthreads + 63) / 64), dim3(64), 0, s, n, dir, Vm, idm, Vp, idp, Vs, id,
                     Vsm, Vsp, status, scratch, nthreads);
  return hchk(hipGetLastError());
}

extern "C" int m2c_mp_launch_onesided_batch(hipStream_t s, int flux, int face, int n, const double *dir,
                                            const double *Vm, const int *idl, const double *uw, const int *ax,
                                            const int *fb, int mode, double *Vs, int *id, double *Vsm, double *F,
                                            int *status, double *scratch, int nthreads) {
  const dim3 gr((nthreads + 63) / 64), bl(64);
#define M2C_OS(FL, FA)                                                                                           \
  hipLaunchKernelGGL((k_onesided_batch<FL, FA>), gr, bl, 0, s, n, dir, Vm, idl, uw, ax, fb, mode, Vs, id, Vsm, F, \
                     status, scratch, nthreads)
  if (!face) M2C_OS(FLUX_HLLC, 0);
  else if (flux == FLUX_HLLC) M2C_OS(FLUX_HLLC, 1);
  else if (flux == FLUX_ROE) M2C_OS(FLUX_ROE, 1);
  else if (flux == FLUX_GODUNOV) M2C_OS(FLUX_GODUNOV, 1);
  else M2C_OS(FLUX_LLF, 1);
#undef M2C_OS
  return hchk(hipGetLastError());
}

extern "C" int m2c_mp_launch_walls(hipStream_t s, int flux, int in_cons, const double *S, const int8_t *ID,
                                   const unsigned *faces, const double *wall, int n, int mode, int recon, double *Fi,
                                   long long Nf, double *scratch, int nthreads, Status *st, const Clock *clk,
                                   int ignore_active) {
  if (n <= 0) return 0;
  const dim3 gr((nthreads + 63) / 64), bl(64);
#define M2C_WF(FL, IC)                                                                                        \
  hipLaunchKernelGGL((k_wall_faces<FL, IC>), gr, bl, 0, s, S, ID, faces, wall, n, mode, recon, Fi, Nf, scratch, \
                     st, clk, ignore_active)
  if (in_cons) {
    if (flux == FLUX_HLLC) M2C_WF(FLUX_HLLC, 1);
    else if (flux == FLUX_ROE) M2C_WF(FLUX_ROE, 1);
    else if (flux == FLUX_GODUNOV) M2C_WF(FLUX_GODUNOV, 1);
    else M2C_WF(FLUX_LLF, 1);
  } else {
    if (flux == FLUX_HLLC) M2C_WF(FLUX_HLLC, 0);
    else if (flux == FLUX_ROE) M2C_WF(FLUX_ROE, 0);
    else if (flux == FLUX_GODUNOV) M2C_WF(FLUX_GODUNOV, 0);
    else M2C_WF(FLUX_LLF, 0);
  }
#undef M2C_WF
  return hchk(hipGetLastError());
}

extern "C" int m2c_mp_launch_godunov_faces(hipStream_t s, int in_cons, const double *S, const int8_t *ID,
                                           double *const *G, const long long *nf, double *scratch, int nthreads,
                                           Status *st, const Clock *clk, int ignore_active, const double *crphi) {
  const dim3 gr((nthreads + 63) / 64), bl(64);
#define M2C_GF(D, IC)                                                                                           \
  hipLaunchKernelGGL((k_godunov_faces<D, IC>), gr, bl, 0, s, S, ID, G[D], nf[D], scratch, st, clk, ignore_active, \
                     crphi)
  if (in_cons) { M2C_GF(0, 1); M2C_GF(1, 1); M2C_GF(2, 1); }
  else { M2C_GF(0, 0); M2C_GF(1, 0); M2C_GF(2, 0); }
#undef M2C_GF
  return hchk(hipGetLastError());
}

extern "C" int m2c_mp_launch_godunov_batch(hipStream_t s, int dir, int n, const double *Vm, const double *Vp, int id,
                                           double *F, double *scratch, int nthreads) {
  const dim3 gr((nthreads + 63) / 64), bl(64);
  if (dir == 0) hipLaunchKernelGGL(k_godunov_batch<0>, gr, bl, 0, s, n, Vm, Vp, id, F, scratch, nthreads);
  else if (dir == 1) hipLaunchKernelGGL(k_godunov_batch<1>, gr, bl, 0, s, n, Vm, Vp, id, F, scratch, nthreads);
  else hipLaunchKernelGGL(k_godunov_batch<2>, gr, bl, 0, s, n, Vm, Vp, id, F, scratch, nthreads);
  return hchk(hipGetLastError());
}

extern "C" int m2c_mp_launch_sur_batch(hipStream_t s, int n, const int *dir, const double *Vm, const int *idm,
                                       const double *Vp, const int *idp, float *X, float *Y, const MlpWeights &w,
                                       int nmax, double tol, double *Vs, int *id, double *Vsm, double *Vsp,
                                       int *status, int *path, double *scratch, int nthreads, unsigned *fbl,
                                       unsigned *nfb) {
  hipLaunchKernelGGL(k_sur_batch_feat, dim3(nblocks(n, 256, 4096)), dim3(256), 0, s, n, dir, Vm, idm, Vp, idp, X);
  if (hchk(hipGetLastError())) return 2;
  if (m2c_mlp_launch_infer(w, s, X, Y, n, nullptr)) return 2;
  hipLaunchKernelGGL(k_sur_batch<1>, dim3((nthreads + 63) / 64), dim3(64), 0, s, n, dir, Vm, idm, Vp, idp, Y, nmax,
                     tol, Vs, id, Vsm, Vsp, status, path, scratch, nthreads, fbl, nfb);
  hipLaunchKernelGGL(k_sur_batch<2>, dim3((nthreads + 63) / 64), dim3(64), 0, s, n, dir, Vm, idm, Vp, idp, Y, nmax,
                     tol, Vs, id, Vsm, Vsp, status, path, scratch, nthreads, fbl, nfb);
  return hchk(hipGetLastError());
}

extern "C" int m2c_mp_step_begin(const MpLaunch &L) {
  hipLaunchKernelGGL(k_mp_step_begin, dim3(1), dim3(1), 0, L.stream, L.ms, L.clk);
  return hchk(hipGetLastError());
}

extern "C" int m2c_mp_launch_ls(const MpLaunch &L, int in_cons, int mode, int final_, const double *S,
                                const double *phiS, const double *X, const double *Y, double *out, double a, double b,
                                double coef, const double *phin) {
  const dim3 gr = sweep_grid(L.nx, L.ny, L.nz), bl(SWEEP_BX, SWEEP_BY);
#define M2C_LS(IC, MO, FI)                                                                                   \
  do {                                                                                                       \
    if (L.ls_neumann)                                                                                        \
      hipLaunchKernelGGL((k_ls_stage<IC, MO, FI, 1>), gr, bl, 0, L.stream, S, L.ID, phiS, X, Y, out, a, b,    \
                         coef, phin, L.IDnew, L.rem, L.ms, L.clk);                                           \
    else                                                                                                     \
      hipLaunchKernelGGL((k_ls_stage<IC, MO, FI, 0>), gr, bl, 0, L.stream, S, L.ID, phiS, X, Y, out, a, b,    \
                         coef, phin, L.IDnew, L.rem, L.ms, L.clk);                                           \
  } while (0)
  if (final_) {
    if (in_cons) {
      if (mode == LS_AXBY) M2C_LS(1, LS_AXBY, 1); else if (mode == LS_X) M2C_LS(1, LS_X, 1); else M2C_LS(1, LS_FIRST, 1);
    } else {
      if (mode == LS_AXBY) M2C_LS(0, LS_AXBY, 1); else if (mode == LS_X) M2C_LS(0, LS_X, 1); else M2C_LS(0, LS_FIRST, 1);
    }
  } else {
    if (in_cons) {
      if (mode == LS_AXBY) M2C_LS(1, LS_AXBY, 0); else if (mode == LS_X) M2C_LS(1, LS_X, 0); else M2C_LS(1, LS_FIRST, 0);
    } else {
      if (mode == LS_AXBY) M2C_LS(0, LS_AXBY, 0); else if (mode == LS_X) M2C_LS(0, LS_X, 0); else M2C_LS(0, LS_FIRST, 0);
    }
  }
#undef M2C_LS
  return hchk(hipGetLastError());
}

extern "C" int m2c_mp_launch_id_remaining(const MpLaunch &L, int single) {
  if (single)
    hipLaunchKernelGGL(k_id_remaining_single, dim3(64), dim3(256), 0, L.stream, L.rem, L.ms, L.IDnew, L.clk);
  else
    hipLaunchKernelGGL(k_id_remaining_seq, dim3(1), dim3(1024), 0, L.stream, L.rem, L.ms, L.IDnew, L.tagR, L.st,
                       L.clk);
  return hchk(hipGetLastError());
}

extern "C" int m2c_mp_launch_reinit(const MpLaunch &L, double *phi, int maxits, int retry) {
  dim3 gr(nblocks(L.N, 256, 8192)), bl(256);
  const dim3 gs = sweep_grid(L.nx, L.ny, L.nz), bs(SWEEP_BX, SWEEP_BY);
  hipLaunchKernelGGL(k_reinit_begin, dim3(1), dim3(1), 0, L.stream, L.ms, L.clk, retry);
  if (retry) hipLaunchKernelGGL(k_reinit_restore, gr, bl, 0, L.stream, phi, L.phibk, L.ms);
  else hipLaunchKernelGGL(k_reinit_tag, gs, bs, 0, L.stream, phi, L.tag, L.sign, L.phibk, L.ms);
  if (L.ls_neumann) {   // every boundary zero-Neumann: one fused launch per iteration
    const dim3 gf((L.nx + RF_TX - 1) / RF_TX, (L.ny + RF_TY - 1) / RF_TY, (L.nz + RF_KC - 1) / RF_KC);
    if (L.ls_f32) hipLaunchKernelGGL(k_reinit_to_f32, gr, bl, 0, L.stream, phi, L.phif0, L.sign, L.signf, L.ms);
    for (int it = 0; it < maxits; it++) {
      if (L.ls_f32)
        hipLaunchKernelGGL(k_reinit_fused<float>, gf, dim3(RF_NT), 0, L.stream, L.phif0, L.phif1, L.tag, L.signf,
                           L.ms);
      else
        hipLaunchKernelGGL(k_reinit_fused<double>, gf, dim3(RF_NT), 0, L.stream, phi, L.rphi1a, L.tag, L.sign, L.ms);
      hipLaunchKernelGGL(k_reinit_decide, dim3(1), dim3(1), 0, L.stream, L.ms, it);
    }
    hipLaunchKernelGGL(k_reinit_end_fused, dim3(1), dim3(1), 0, L.stream, L.ms, L.st);
    if (L.ls_f32) hipLaunchKernelGGL(k_reinit_settle_f32, gr, bl, 0, L.stream, phi, L.phif0, L.phif1, L.ms);
    else hipLaunchKernelGGL(k_reinit_settle, gr, bl, 0, L.stream, phi, L.rphi1a, L.ms);
    return hchk(hipGetLastError());
  }
  for (int it = 0; it < maxits; it++) {
    hipLaunchKernelGGL(k_reinit_A, gs, bs, 0, L.stream, phi, L.rphi1a, L.tag, L.sign, L.ms, it);
    hipLaunchKernelGGL(k_reinit_B, gs, bs, 0, L.stream, phi, L.rphi1a, L.rphi1b, L.tag, L.sign, L.ms, it);
    hipLaunchKernelGGL(k_reinit_C, gs, bs, 0, L.stream, phi, L.rphi1b, L.tag, L.sign, L.ms);
  }
  hipLaunchKernelGGL(k_reinit_end, dim3(1), dim3(1), 0, L.stream, L.ms, L.st);
  return hchk(hipGetLastError());
}

extern "C" int m2c_mp_launch_state_update(const MpLaunch &L, double *V, int rounds) {
  dim3 grN(nblocks(L.N, 256, 16384)), bl(256);
  hipLaunchKernelGGL(k_changed_list, grN, bl, 0, L.stream, L.ID, L.IDnew, L.chg, L.cidx, L.ms, L.clk);
  NbView nv;
  nv.V = V; nv.IDn = L.ID; nv.cidx = L.cidx; nv.lvl = L.lvl; nv.cnt = L.cnt; nv.Vnew = L.Vnew; nv.ncap = 0;
  dim3 grc(1024);
  hipLaunchKernelGGL(k_state_round0, grc, bl, 0, L.stream, nv, L.chg, L.IDnew, L.rsV, L.rsMeta, L.ms, L.lvl, L.cnt,
                     L.Vnew, L.clk);
  for (int r = 1; r <= rounds; r++)
    hipLaunchKernelGGL(k_state_round, grc, bl, 0, L.stream, nv, L.chg, L.IDnew, L.rsV, L.rsMeta, L.ms, L.lvl, L.cnt,
                       L.Vnew, r, L.clk);
  hipLaunchKernelGGL(k_state_seq, dim3(1), dim3(1024), 0, L.stream, nv, L.chg, L.IDnew, L.rsV, L.rsMeta, L.ms, L.lvl,
                     L.cnt, L.Vnew, L.st, L.clk);
  hipLaunchKernelGGL(k_state_commit, grc, bl, 0, L.stream, V, L.chg, L.cnt, L.Vnew, L.unres, L.ms, L.clk);
  hipLaunchKernelGGL(k_fix_unresolved, dim3(1), dim3(1024), 0, L.stream, V, L.IDnew, L.cidx, L.cnt, L.unres, L.ms,
                     L.st, L.clk);
  hipLaunchKernelGGL(k_state_finish, grc, bl, 0, L.stream, V, L.ID, L.IDnew, L.chg, L.cidx, L.ms, L.st, L.clk);
  return hchk(hipGetLastError());
}
#endif  // M2C_IFACE_MULTI
