// m2c_hip.hip -- C-ABI context of the MI355X-native finite-volume update.
//
// Owns the device fields (structure of arrays, x fastest, no ghost frame), the HIP stream and
// the device-side time-loop clock; launches the fused stage kernels of fv_kernels.hpp.
// Entry points are declared in include/m2c_hip.h.
#include <hip/hip_runtime.h>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/m2c_hip.h"
#include "multiphase.hpp"
#include "batch_ctx.hpp"
#include "stage_launch.hpp"

using namespace m2c;

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      fprintf(stderr, "m2c_hip: %s failed: %s\n", #x, hipGetErrorString(e_));       \
      return M2C_ERR_HIP;                                                           \
    }                                                                               \
  } while (0)

struct DiffTab {
  const double *x[3];          // coordinates, index -1..n stored at [q+1]
  const double *wl[3], *wr[3]; // interpolation weights of face q-1/2, q = 0..n
  const double *cc[3];         // central node-derivative coefficients, 3 per node q = 0..n-1
  double clo[3], chi[3];       // one-sided coefficients at q = 0 and q = n-1
  int visc, heat;
};

struct DiffMat {
  int visc[M2C_MAX_MAT];
  double mu[M2C_MAX_MAT], bulk[M2C_MAX_MAT], suT0[M2C_MAX_MAT], suS[M2C_MAX_MAT], kappa[M2C_MAX_MAT];
  TLaw tl;
};

struct m2c_ctx {
  m2c_config cfg;
  Grid grid;
  DevEos eos[8];
  hipStream_t stream = nullptr;
  long long N = 0;
  double *V = nullptr, *U1 = nullptr, *U2 = nullptr, *R = nullptr;
  double *Fi = nullptr;
  long long Nf = 0;
  int8_t *ID = nullptr;
  Clock *clk = nullptr;
  Status *st = nullptr;
  int kchunk = 1;
  int variant = 1;             // 0 HLLC+SG, 1 HLLC generic EOS, 2 Roe, 3 LLF; +4: mixed precision; 8 Godunov;
                               // 9 HLLC+SG with material interfaces, 10 its mixed form
  dim3 grid3;
  // two materials / level set
  bool iface = false;          // exact Riemann pass at material interfaces
  bool ls = false;             // level set advected and material IDs updated
  bool id_single = false;      // the only material a swept-out node can take is 0
  MpParams mp;
  MpLaunch L;
  double *phiA = nullptr, *phiB = nullptr, *phiD = nullptr;
  double *xc = nullptr, *fd = nullptr;   // coordinates and FD3 tables (device)
  std::vector<void *> mp_bufs;
  // viscous and heat-diffusion terms
  bool diff = false;
  double *Dv[3] = {nullptr, nullptr, nullptr}, *Dh[3] = {nullptr, nullptr, nullptr};
  double *Vprim = nullptr;     // primitive copy of a conservative stage field for the face pass
  long long nf[3] = {0, 0, 0};
  DiffTab dtab;
  DiffMat dmat;
  // profiling: events around each diffusion pass (m2c_time_steps), and its last average
  std::vector<hipEvent_t> *diff_ev = nullptr;
  size_t diff_ev_i = 0;
  double diff_ms = 0.0;
  // Godunov flux: per-direction face fluxes of the face pass and the exact solver's path cache
  bool godunov = false;
  double *G[3] = {nullptr, nullptr, nullptr};
  long long gnf[3] = {0, 0, 0};
  double *gscratch = nullptr;
  int gthreads = 0;
  // mesh faces cut by embedded walls (m2c_set_embedded_faces): face codes dir * Nf + key, wall
  // records (12 doubles each), the per-face-key cut mask, the exact solver's path cache
  int eb_n = 0, eb_mode = 1, eb_recon = 0, eb_threads = 0;
  int variant0 = 1;            // the kernel variant chosen at creation (walls force a generic one)
  unsigned *eb_faces = nullptr;
  double *eb_wall = nullptr, *eb_scratch = nullptr;
  uint8_t *eb_mask = nullptr;
  // attached to a batch (m2c_ctx_batch_attach): the context's own stream, clock and status wait here
  // while the batch's are in use; pinned copies of the interface and level-set constants for the
  // asynchronous switch between instances; the autotuner's copy of the state
  hipStream_t own_stream = nullptr;
  Clock *own_clk = nullptr;
  Status *own_st = nullptr;
  struct Bound {
    Grid g;
    DevEos eos[8];
    MpParams mp;
  } *bound = nullptr;
  double *bk_V = nullptr, *bk_phi = nullptr;
  int8_t *bk_ID = nullptr;
  MpState *bk_ms = nullptr;
  // mesh: cell centres and widths of each axis (host), and on a stretched mesh the device tables
  // the kernels read (widths with two ghosts per side, limiter coefficients per cell)
  std::vector<double> mx[3], mw[3];
  double *d_mesh = nullptr;
  int *d_meshK = nullptr;
};

static const unsigned long long kDblMaxBits = 0x7FEFFFFFFFFFFFFFull;

static int ensure_riemann_constants(m2c_ctx *c);
static void riemann_only_params(const m2c_config &cfg, MpParams &mp);
static bool riemann_batch_args_ok(const m2c_ctx *c, int n, const int *idm, const int *idp);
static int setup_godunov(m2c_ctx *c);
static TLaw build_tlaw(const m2c_config &cfg);

// ---------------------------------------------------------------- small kernels

__global__ void k_clock_init(Clock *clk, double t, double tmax, double cfl, int maxsteps) {
  clk->t = t;
  clk->dt = 0.0;
  clk->tmax = tmax;
  clk->cfl = cfl;
  clk->minbits = kDblMaxBits;
  clk->steps = 0;
  clk->maxsteps = maxsteps;
  clk->active = 1;
}

__global__ void k_clock_set(Clock *clk, double tmax, int maxsteps_more) {
  clk->tmax = tmax;
  clk->maxsteps = clk->steps + maxsteps_more;
  clk->minbits = kDblMaxBits;
  clk->active = 1;
}

// dt for the next step from the CFL minimum (SpaceOperator.cpp:2219-2225, Main.cpp:535-539).
// advance != 0: first account for the step just taken (t += dt, Main.cpp:566-567).
__global__ void k_clock_finalize(Clock *clk, int advance) {
  if (advance) {
    if (!clk->active) return;
    clk->t += clk->dt;
    clk->steps += 1;
  }
  double mn = __longlong_as_double((long long)clk->minbits);
  double dt = clk->cfl * mn;
  if (clk->t + dt > clk->tmax) dt = clk->tmax - clk->t;
  clk->dt = dt;
  clk->active = (clk->t < clk->tmax && clk->steps < clk->maxsteps) ? 1 : 0;
  clk->minbits = kDblMaxBits;
}

__global__ void k_clock_fixed_dt(Clock *clk, double dt) {
  clk->dt = dt;
  clk->active = 1;
  clk->minbits = kDblMaxBits;
}

// dx/(|u|+c) minimum over the active cells (SpaceOperator.cpp:2128-2198)
// EK: the equation of state when every material has it (stiffened gas), -1 = dispatch per cell
template <int EK>
__global__ __launch_bounds__(256) void k_cfl_min(const double *__restrict__ V, const int8_t *__restrict__ ID,
                                                 Clock *clk, Status *st) {
  const Grid &g = c_grid;
  double mn = DBL_MAX;
  unsigned flags = 0;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < g.N;
       idx += (long long)gridDim.x * blockDim.x) {
    int id = ID[idx];
    if (id == g.inactive_id) continue;
    const DevEos &m = c_eos[id];
    double v[5];
#pragma unroll
    for (int p = 0; p < 5; p++) v[p] = V[p * g.N + idx];
    double c2 = eos_c2<double, EK>(m, v[0], eos_e<double, EK>(m, v[0], v[4]));
    if (c2 < 0) { flags |= ST_NEG_C2; continue; }
    double cs = dsqrt(c2);
    double lf = dabs(v[1]) + cs, lg = dabs(v[2]) + cs, lh = dabs(v[3]) + cs;
    double hx = g.dx, hy = g.dy, hz = g.dz;
    if (EK < 0 && g.stretched) {   // the cell's own widths
      const long long r = idx / g.nx;
      hx = g.w[0][idx - r * g.nx + 2]; hy = g.w[1][r % g.ny + 2]; hz = g.w[2][r / g.ny + 2];
    }
    double loc = smin(hx / lf, smin(hy / lg, hz / lh));
    if (loc < mn) mn = loc;
  }
  // block minimum first: one atomic per block on the single address
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) mn = smin(mn, __shfl_xor(mn, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mn;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double m = smin(smin(red[0], red[1]), smin(red[2], red[3]));
    atomicMin(&clk->minbits, (unsigned long long)__double_as_longlong(m));
  }
  if (flags) atomicOr(&st->bits, flags);
}

template <int FLUX, int DIR>
__global__ void k_flux_batch(int n, const double *Vm, const double *Vp, int id, double *F) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double a[5], b[5], f[5];
  for (int p = 0; p < 5; p++) { a[p] = Vm[5 * t + p]; b[p] = Vp[5 * t + p]; }
  numerical_flux<FLUX, -1, DIR>(a, b, id, f);
  for (int p = 0; p < 5; p++) F[5 * t + p] = f[p];
}

template <int FLUX>
static void launch_flux_batch(int dir, dim3 g, hipStream_t s, int n, const double *a, const double *b, int id,
                              double *f) {
  if (dir == 0) hipLaunchKernelGGL((k_flux_batch<FLUX, 0>), g, dim3(256), 0, s, n, a, b, id, f);
  else if (dir == 1) hipLaunchKernelGGL((k_flux_batch<FLUX, 1>), g, dim3(256), 0, s, n, a, b, id, f);
  else hipLaunchKernelGGL((k_flux_batch<FLUX, 2>), g, dim3(256), 0, s, n, a, b, id, f);
}

// Plain streaming copy (16-byte vectors) used to measure and calibrate HBM traffic.
__global__ __launch_bounds__(256) void k_stream_copy(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                     long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// AoS host layout <-> SoA device layout
__global__ void k_aos_to_soa(const double *in, double *out, long long N) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x)
    for (int p = 0; p < 5; p++) out[p * N + i] = in[5 * i + p];
}
__global__ void k_soa_to_aos(const double *in, double *out, long long N) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x)
    for (int p = 0; p < 5; p++) out[5 * i + p] = in[p * N + i];
}

__global__ void k_copy_gated(const double *src, double *dst, long long n, const Clock *clk) {
  if (!clk->active) return;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// ---------------------------------------------------------------- host helpers

// derivative-at-0 weights of the 4-point interpolant through offsets d[0..3] with d[c] = 0
// (closed form of the Taylor system solved by GradientCalculatorFD3.cpp:451-466)
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
  double sum = 0.0;
  for (int m = 0; m < 4; m++)
    if (m != cidx) sum = sum + 1.0 / d[m];
  w[cidx] = -sum;
}

static void build_eos(const m2c_material &in, DevEos &m) {
  memset(&m, 0, sizeof(m));
  m.type = in.eos;
  m.rhomin = in.rhomin; m.pmin = in.pmin; m.rhomax = in.rhomax; m.pmax = in.pmax;
  const double *p = in.p;
  switch (in.eos) {
  case M2C_EOS_SG:
    m.gam = p[0]; m.pstiff = p[1];
    m.gam1 = m.gam - 1.0;
    m.gam_pstiff = m.gam * m.pstiff;            // VarFcnSG.h:64 evaluates gam*Pstiff
    break;
  case M2C_EOS_NASG:
    m.gam = p[0]; m.pc = p[1]; m.b = p[2]; m.q = p[3];
    m.gam1 = m.gam - 1.0; m.invgam1 = 1.0 / m.gam1; m.gam_pc = m.gam * m.pc;   // VarFcnNASG.h:104-109
    break;
  case M2C_EOS_MG:
    m.rho0 = p[0]; m.c0 = p[1]; m.Gamma0 = p[2]; m.s = p[3]; m.e0 = p[4];
    m.r0c0c0 = m.rho0 * m.c0 * m.c0; m.G0half = 0.5 * m.Gamma0; m.G0r0 = m.Gamma0 * m.rho0;  // VarFcnMG.h:117-119
    break;
  case M2C_EOS_JWL:
    m.omega = p[0]; m.A1 = p[1]; m.A2 = p[2]; m.rho0 = p[5];
    m.R1r0 = p[3] * m.rho0; m.R2r0 = p[4] * m.rho0;
    m.oR1r0 = m.omega / m.R1r0; m.oR2r0 = m.omega / m.R2r0;                    // VarFcnJWL.h:95-98
    break;
  case M2C_EOS_DUMMY:
    m.d_p0 = p[0]; m.d_rho0 = p[1]; m.d_e0 = p[2];
    break;
  default: break;
  }
}

static int slope_k(double A, double B) {   // Reconstructor.cpp:103-121
  int k;
  for (k = 2; k < 1000; k++) {
    double rhs = 2.0 / (1.0 + std::pow((k - 1.0) / k, k - 1) / k) * std::min(A, 1.0);
    if (B <= rhs) break;
  }
  return k;
}

// The grid and material tables live in __constant__ memory, shared by all contexts of the
// process; the context whose tables are resident is remembered and switched on demand (every
// entry point leaves its stream idle before returning, so a switch never races a kernel).
static m2c_ctx *g_resident = nullptr;

static int upload_constants(m2c_ctx *c) {
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_grid), &c->grid, sizeof(Grid)));
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_eos), c->eos, sizeof(c->eos)));
  if ((c->iface || c->ls || c->godunov || c->eb_n) && m2c_mp_upload(c->grid, c->eos, c->mp)) return M2C_ERR_HIP;
  if (m2c_stage_upload_hllc_sg(c->grid, c->eos) || m2c_stage_upload_hllc_gen(c->grid, c->eos) ||
      m2c_stage_upload_god_gen(c->grid, c->eos) ||
      m2c_stage_upload_roe_gen(c->grid, c->eos) || m2c_stage_upload_llf_gen(c->grid, c->eos) ||
      m2c_stage_upload_hllc_sg_f32(c->grid, c->eos) || m2c_stage_upload_hllc_gen_f32(c->grid, c->eos) ||
      m2c_stage_upload_roe_gen_f32(c->grid, c->eos) || m2c_stage_upload_llf_gen_f32(c->grid, c->eos) ||
      m2c_stage_upload_hllc_sg_mp(c->grid, c->eos) || m2c_stage_upload_hllc_sg_mp_f32(c->grid, c->eos))
    return M2C_ERR_HIP;
  g_resident = c;
  return M2C_OK;
}

static int ensure_constants(m2c_ctx *c) {
  HIPCHK(hipSetDevice(c->cfg.device));
  if (g_resident == c) return M2C_OK;
  HIPCHK(hipDeviceSynchronize());
  return upload_constants(c);
}

// device buffers owned by the context (freed with it)
template <typename T> static int mp_alloc(m2c_ctx *c, T **p, size_t count, int fill) {
  HIPCHK(hipMalloc((void **)p, sizeof(T) * (count ? count : 1)));
  HIPCHK(hipMemset(*p, fill, sizeof(T) * (count ? count : 1)));
  c->mp_bufs.push_back((void *)*p);
  return M2C_OK;
}

// ---------------------------------------------------------------- viscous and heat-diffusion terms
//
// ViscosityOperator::AddDiffusionFluxes (ViscosityOperator.cpp:200-404) and HeatDiffusionOperator::
// AddDiffusionFluxes (HeatDiffusionOperator.cpp:64-139): one thread per face computes the face
// velocity (InterpolatorLinear, Interpolator.cpp:32-181), the velocity gradient (central differences
// across the face; along the face the node derivatives of GradientCalculatorCentral.cpp:22-207, with
// its one-sided rows next to the ghost layer, interpolated to the face) and the stresses of
// ViscoFcn.h, and the harmonic-mean heat flux. The stage kernel adds them after the advective sum.
// Reproduced as the reference writes them: the k-1/2 face of two materials takes the upper cell's
// viscosity function for both sides (ViscosityOperator.cpp:386); Sutherland's law evaluates the
// temperature at (rho, p) (ViscoFcn.h:255); the j-1/2 and k-1/2 heat fluxes divide by the i-1/2
// face's diffusivity sum, carried over from the last cell that had one on the north ghost row and
// the top ghost plane (HeatDiffusionOperator.cpp:106-127).

// material id of (possibly ghost) cell (i, j, k): the face ghosts take their image's
__device__ __forceinline__ int id_at(const int8_t *__restrict__ ID, int i, int j, int k) {
  const Grid &g = c_grid;
  int ii = clampi(image_idx(i, g.nx, (i < 0 ? g.bc[0] : g.bc[1]) == BC_PERIODIC), 0, g.nx - 1);
  int jj = clampi(image_idx(j, g.ny, (j < 0 ? g.bc[2] : g.bc[3]) == BC_PERIODIC), 0, g.ny - 1);
  int kk = clampi(image_idx(k, g.nz, (k < 0 ? g.bc[4] : g.bc[5]) == BC_PERIODIC), 0, g.nz - 1);
  return ID[((long long)kk * g.ny + jj) * g.nx + ii];
}

template <int IN_CONS>
__device__ __forceinline__ void dcell(const double *__restrict__ S, const int8_t *__restrict__ ID, int i, int j,
                                      int k, double *V, int &id) {
  unsigned fl = 0;
  fetch_cell<IN_CONS, -1>(S, ID, i, j, k, V, id, fl);
}

// d u_m / d x_d at node (i, j, k), m = 0..2 (velocity), given the node's own velocity uc
template <int IN_CONS>
__device__ void node_deriv(const DiffTab &t, const double *__restrict__ S, const int8_t *__restrict__ ID, int d,
                           int i, int j, int k, const double *uc, double *out) {
  const Grid &g = c_grid;
  const int n = d == 0 ? g.nx : (d == 1 ? g.ny : g.nz);
  const int q = d == 0 ? i : (d == 1 ? j : k);
  bool ghostrow;
  if (d == 0) ghostrow = (j == -1 || j == g.ny || k == -1 || k == g.nz);
  else if (d == 1) ghostrow = (i == -1 || i == g.nx || k == -1 || k == g.nz);
  else ghostrow = (i == -1 || i == g.nx || j == -1 || j == g.ny);
  const int di = d == 0, dj = d == 1, dk = d == 2;
  double W[5];
  int idd;
  if (q == 0 && ghostrow) {
    dcell<IN_CONS>(S, ID, i + di, j + dj, k + dk, W, idd);
    const double coeff = t.clo[d], c1 = -coeff, c2 = coeff;
    for (int m = 0; m < 3; m++) out[m] = c1 * uc[m] + c2 * W[1 + m];
    return;
  }
  if (q == n - 1 && ghostrow) {
    dcell<IN_CONS>(S, ID, i - di, j - dj, k - dk, W, idd);
    const double coeff = t.chi[d], c0 = -coeff, c1 = coeff;
    for (int m = 0; m < 3; m++) out[m] = c0 * W[1 + m] + c1 * uc[m];
    return;
  }
  double P[5];
  dcell<IN_CONS>(S, ID, i - di, j - dj, k - dk, W, idd);
  dcell<IN_CONS>(S, ID, i + di, j + dj, k + dk, P, idd);
  const double c0 = t.cc[d][3 * q], c1 = t.cc[d][3 * q + 1], c2 = t.cc[d][3 * q + 2];
  for (int m = 0; m < 3; m++) out[m] = c0 * W[1 + m] + c1 * uc[m] + c2 * P[1 + m];
}

// EvaluateViscousFluxFunction_F/G/H (ViscoFcn.h:61-110), Constant and Sutherland stresses
template <int DIR>
__device__ void visc_flux(const DiffMat &dm, int mid, const double D[3][3], const double *V3, double rho, double p,
                          double *F) {
  if (dm.visc[mid] == 0) {
    F[0] = F[1] = F[2] = F[3] = F[4] = 0.0;
    return;
  }
  double dyna = dm.mu[mid];
  if (dm.visc[mid] == 2) {
    const double T = tlaw_temperature(dm.tl, c_eos[mid], mid, rho, p);
    dyna = dm.mu[mid] * pow(T / dm.suT0[mid], 1.5) * (dm.suT0[mid] + dm.suS[mid]) / (T + dm.suS[mid]);
  }
  const double bulk = dm.bulk[mid];
  const double *dVdx = D[0], *dVdy = D[1], *dVdz = D[2];
  const double div = dVdx[0] + dVdy[1] + dVdz[2];
  double t0, t1, t2;
  if (DIR == 0) {
    t0 = bulk * div + dyna * (2.0 * dVdx[0] - 2.0 / 3.0 * div);
    t1 = dyna * (dVdx[1] + dVdy[0]);
    t2 = dyna * (dVdx[2] + dVdz[0]);
  } else if (DIR == 1) {
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

template <int DIR, int IN_CONS>
__global__ __launch_bounds__(256) void k_diff_faces(const double *__restrict__ S, const int8_t *__restrict__ ID,
                                                    DiffTab t, DiffMat dm, double *__restrict__ Dv,
                                                    double *__restrict__ Dh, long long nf) {
  const Grid &g = c_grid;
  const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int nxf = g.nx + (DIR == 0), nyf = g.ny + (DIR == 1);
  const int i = (int)(f % nxf), j = (int)((f / nxf) % nyf), k = (int)(f / ((long long)nxf * nyf));
  const int di = DIR == 0, dj = DIR == 1, dk = DIR == 2;
  double Vh[5], Vl[5];
  int idh, idl;
  dcell<IN_CONS>(S, ID, i, j, k, Vh, idh);
  dcell<IN_CONS>(S, ID, i - di, j - dj, k - dk, Vl, idl);
  const int q = DIR == 0 ? i : (DIR == 1 ? j : k);
  const double *x = t.x[DIR] + 1;
  const double area = DIR == 0 ? g.ax : (DIR == 1 ? g.ay : g.az);
  const int inactive = g.inactive_id;
  double out[10];
#pragma unroll
  for (int c = 0; c < 10; c++) out[c] = 0.0;
  if (t.visc && idh != inactive) {
    const double wl = t.wl[DIR][q], wr = t.wr[DIR][q], dd = x[q] - x[q - 1];
    double V3[3], D[3][3];
    for (int m = 0; m < 3; m++) {
      V3[m] = wl * Vl[1 + m] + wr * Vh[1 + m];
      D[DIR][m] = (Vh[1 + m] - Vl[1 + m]) / dd;
    }
    for (int ax = 0; ax < 3; ax++) {
      if (ax == DIR) continue;
      double Dl[3], Dhh[3];
      node_deriv<IN_CONS>(t, S, ID, ax, i - di, j - dj, k - dk, Vl + 1, Dl);
      node_deriv<IN_CONS>(t, S, ID, ax, i, j, k, Vh + 1, Dhh);
      for (int m = 0; m < 3; m++) D[ax][m] = wl * Dl[m] + wr * Dhh[m];
    }
    double F[5];
    if (idl == idh) {
      visc_flux<DIR>(dm, idh, D, V3, 0.5 * (Vh[0] + Vl[0]), 0.5 * (Vh[4] + Vl[4]), F);
      for (int c = 0; c < 5; c++) out[c] = out[5 + c] = F[c] * area;
    } else {
      visc_flux<DIR>(dm, idh, D, V3, Vh[0], Vh[4], F);
      for (int c = 0; c < 5; c++) out[c] = area * F[c];
      if (idl != inactive) {
        visc_flux<DIR>(dm, DIR == 2 ? idh : idl, D, V3, Vl[0], Vl[4], F);
        for (int c = 0; c < 5; c++) out[5 + c] = area * F[c];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 10; c++) Dv[c * nf + f] = out[c];
  double hf = 0.0;
  if (t.heat && idh != inactive) {
    const double myk = dm.kappa[idh], neighk = dm.kappa[idl];
    double denom;
    if (DIR == 0) denom = myk + neighk;
    else if ((DIR == 1 && j < g.ny) || (DIR == 2 && k < g.nz)) denom = myk + dm.kappa[id_at(ID, i - 1, j, k)];
    else {
      // the last i-1/2 face computed before this ghost row / plane: cell (nx, ny-1, k') with
      // k' = k on the north ghost row, nz-1 on the top ghost plane
      const int kk = DIR == 1 ? k : g.nz - 1;
      denom = dm.kappa[id_at(ID, g.nx, g.ny - 1, kk)] + dm.kappa[id_at(ID, g.nx - 1, g.ny - 1, kk)];
    }
    const double eh = eos_e<double, -1>(c_eos[idh], Vh[0], Vh[4]);
    const double el = eos_e<double, -1>(c_eos[idl], Vl[0], Vl[4]);
    const double Th = tlaw_temperature(dm.tl, c_eos[idh], idh, Vh[0], eh);
    const double Tl = tlaw_temperature(dm.tl, c_eos[idl], idl, Vl[0], el);
    const double dT = (Th - Tl) / (x[q] - x[q - 1]);
    double flux = (denom == 0) ? 0.0 : 2.0 * myk * neighk / denom * dT;
    flux *= area;
    hf = flux;
  }
  Dh[f] = hf;
}

static int setup_diffusion(m2c_ctx *c) {
  const m2c_config &cfg = c->cfg;
  bool visc = false, heat = false;
  for (int m = 0; m < cfg.nmat; m++) {
    visc |= cfg.mat[m].viscosity != 0;
    heat |= cfg.mat[m].kappa != 0.0;
  }
  c->diff = visc || heat;
  if (!c->diff) return M2C_OK;
  const Grid &g = c->grid;
  const int n[3] = {g.nx, g.ny, g.nz};
  c->nf[0] = (long long)(g.nx + 1) * g.ny * g.nz;
  c->nf[1] = (long long)g.nx * (g.ny + 1) * g.nz;
  c->nf[2] = (long long)g.nx * g.ny * (g.nz + 1);
  int rc;
  for (int d = 0; d < 3; d++) {
    if ((rc = mp_alloc(c, &c->Dv[d], (size_t)(10 * c->nf[d]), 0))) return rc;
    if ((rc = mp_alloc(c, &c->Dh[d], (size_t)c->nf[d], 0))) return rc;
  }
  if ((rc = mp_alloc(c, &c->Vprim, (size_t)(5 * c->N), 0))) return rc;
  // coordinates x[-1..n] (MeshGenerator.cpp:146-158, ghosts SpaceOperator.cpp:218-265), interpolation
  // weights (Interpolator.cpp:150-181) and node-derivative coefficients (GradientCalculatorCentral.cpp:
  // 249-424), evaluated as the reference evaluates them
  const double lo[3] = {cfg.x0, cfg.y0, cfg.z0}, h[3] = {g.dx, g.dy, g.dz};
  std::vector<double> tab;
  size_t ox[3], owl[3], owr[3], occ[3];
  DiffTab t;
  for (int d = 0; d < 3; d++) {
    std::vector<double> x(n[d] + 2);
    for (int q = 0; q < n[d]; q++) x[q + 1] = lo[d] + 0.5 * h[d] + q * h[d];
    x[0] = x[1] - h[d];
    x[n[d] + 1] = x[n[d]] + h[d];
    const double *X = x.data() + 1;
    ox[d] = tab.size();
    tab.insert(tab.end(), x.begin(), x.end());
    owl[d] = tab.size();
    for (int q = 0; q <= n[d]; q++) {
      const double xm = X[q] - 0.5 * h[d], dd = X[q] - X[q - 1];
      tab.push_back((X[q] - xm) / dd);
    }
    owr[d] = tab.size();
    for (int q = 0; q <= n[d]; q++) {
      const double xm = X[q] - 0.5 * h[d], dd = X[q] - X[q - 1];
      tab.push_back((xm - X[q - 1]) / dd);
    }
    occ[d] = tab.size();
    for (int q = 0; q < n[d]; q++) {
      const double d0 = X[q] - X[q - 1], d1 = X[q + 1] - X[q], d2 = d0 + d1;
      tab.push_back(-d1 / (d0 * d2));
      tab.push_back(1.0 / d0 - 1.0 / d1);
      tab.push_back(d0 / (d1 * d2));
    }
    t.clo[d] = 1.0 / (X[1] - X[0]);
    t.chi[d] = 1.0 / (X[n[d] - 1] - X[n[d] - 2]);
  }
  double *dt;
  if ((rc = mp_alloc(c, &dt, tab.size(), 0))) return rc;
  HIPCHK(hipMemcpy(dt, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
  for (int d = 0; d < 3; d++) {
    t.x[d] = dt + ox[d];
    t.wl[d] = dt + owl[d];
    t.wr[d] = dt + owr[d];
    t.cc[d] = dt + occ[d];
  }
  t.visc = visc;
  t.heat = heat;
  c->dtab = t;
  DiffMat &dm = c->dmat;
  memset(&dm, 0, sizeof(dm));
  for (int m = 0; m < cfg.nmat; m++) {
    dm.visc[m] = cfg.mat[m].viscosity;
    dm.mu[m] = cfg.mat[m].mu;
    dm.bulk[m] = cfg.mat[m].bulk;
    dm.suT0[m] = cfg.mat[m].su_T0;
    dm.suS[m] = cfg.mat[m].su_S;
    dm.kappa[m] = cfg.mat[m].kappa;
  }
  dm.tl = build_tlaw(cfg);
  return M2C_OK;
}

// ConservativeToPrimitive and the clip of every cell of a conservative stage field, once, as the
// stencil fetch does them (the ghost rules then act on the primitive values, as they do there)
__global__ __launch_bounds__(256) void k_prim_field(const double *__restrict__ U, const int8_t *__restrict__ ID,
                                                    double *__restrict__ V) {
  const Grid &g = c_grid;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.N) return;
  double u[5], v[5];
#pragma unroll
  for (int p = 0; p < 5; p++) u[p] = U[p * g.N + c];
  const DevEos &m = c_eos[ID[c]];
  eos_c2p<-1>(m, u, v);
  eos_clip<double, -1>(m, v);
#pragma unroll
  for (int p = 0; p < 5; p++) V[p * g.N + c] = v[p];
}

static int diffusion_pass(m2c_ctx *c, const double *S, int in_cons) {
  if (in_cons) {   // the face kernels fetch ~10 cells per face: convert each cell once instead
    hipLaunchKernelGGL(k_prim_field, dim3((unsigned)((c->N + 255) / 256)), dim3(256), 0, c->stream, S, c->ID,
                       c->Vprim);
    S = c->Vprim;
  }
  for (int d = 0; d < 3; d++) {
    dim3 gb((unsigned)((c->nf[d] + 255) / 256)), bs(256);
    if (d == 0) hipLaunchKernelGGL((k_diff_faces<0, 0>), gb, bs, 0, c->stream, S, c->ID, c->dtab, c->dmat, c->Dv[0], c->Dh[0], c->nf[0]);
    else if (d == 1) hipLaunchKernelGGL((k_diff_faces<1, 0>), gb, bs, 0, c->stream, S, c->ID, c->dtab, c->dmat, c->Dv[1], c->Dh[1], c->nf[1]);
    else hipLaunchKernelGGL((k_diff_faces<2, 0>), gb, bs, 0, c->stream, S, c->ID, c->dtab, c->dmat, c->Dv[2], c->Dh[2], c->nf[2]);
  }
  HIPCHK(hipGetLastError());
  return M2C_OK;
}

// Phi(n) for the constant-reconstruction depth (MultiPhase.ConstantReconstructionDepth), else null
static const double *conrec_phi(const m2c_ctx *c) {
  return (c->ls && c->cfg.mp_conrec_depth > 0.0) ? c->phiA : nullptr;
}

static int launch_stage(m2c_ctx *c, int in_cons, int mode, const double *S, const double *Vn, const double *Us,
                        double *Out, double a, double b, double cc, int first, int ignore_active) {
  StageArgs s;
  if (c->diff) {
    const bool prof = c->diff_ev && c->diff_ev_i + 2 <= c->diff_ev->size();
    if (prof) hipEventRecord((*c->diff_ev)[c->diff_ev_i++], c->stream);
    if (int rc = diffusion_pass(c, S, in_cons)) return rc;
    if (prof) hipEventRecord((*c->diff_ev)[c->diff_ev_i++], c->stream);
    for (int d = 0; d < 3; d++) {
      s.Dv[d] = c->Dv[d];
      s.Dh[d] = c->Dh[d];
    }
  }
  if (c->godunov) {   // the exact solve of every same-material face, before the stage
    const bool prof = c->diff_ev && c->diff_ev_i + 2 <= c->diff_ev->size();
    if (prof) hipEventRecord((*c->diff_ev)[c->diff_ev_i++], c->stream);
    if (m2c_mp_launch_godunov_faces(c->stream, in_cons, S, c->ID, c->G, c->gnf, c->gscratch, c->gthreads, c->st,
                                    c->clk, ignore_active, conrec_phi(c)))
      return M2C_ERR_HIP;
    if (prof) hipEventRecord((*c->diff_ev)[c->diff_ev_i++], c->stream);
    for (int d = 0; d < 3; d++) s.G[d] = c->G[d];
  }
  if (c->eb_n) {   // the two fluxes of every wall-cut face, before the stage reads them
    if (m2c_mp_launch_walls(c->stream, c->cfg.flux, in_cons, S, c->ID, c->eb_faces, c->eb_wall, c->eb_n, c->eb_mode,
                            c->eb_recon, c->Fi, c->Nf, c->eb_scratch, c->eb_threads, c->st, c->clk, ignore_active))
      return M2C_ERR_HIP;
    s.Wm = c->eb_mask;
  }
  s.phin = conrec_phi(c);
  s.no_cfl = c->ls ? 1 : 0;
  s.S = S; s.Vn = Vn; s.Us = Us; s.Out = Out; s.ID = c->ID;
  s.Fi = c->Fi; s.Nf = c->Nf;
  s.clk = c->clk; s.st = c->st;
  s.a = a; s.b = b; s.c = cc; s.first = first;
  s.kchunk = c->kchunk;
  s.ignore_active = ignore_active;
  int rc;
  switch (c->variant) {
  case 0: rc = m2c_stage_launch_hllc_sg(in_cons, mode, c->grid3, c->stream, s); break;
  case 1: rc = m2c_stage_launch_hllc_gen(in_cons, mode, c->grid3, c->stream, s); break;
  case 2: rc = m2c_stage_launch_roe_gen(in_cons, mode, c->grid3, c->stream, s); break;
  case 3: rc = m2c_stage_launch_llf_gen(in_cons, mode, c->grid3, c->stream, s); break;
  case 4: rc = m2c_stage_launch_hllc_sg_f32(in_cons, mode, c->grid3, c->stream, s); break;
  case 5: rc = m2c_stage_launch_hllc_gen_f32(in_cons, mode, c->grid3, c->stream, s); break;
  case 6: rc = m2c_stage_launch_roe_gen_f32(in_cons, mode, c->grid3, c->stream, s); break;
  case 7: rc = m2c_stage_launch_llf_gen_f32(in_cons, mode, c->grid3, c->stream, s); break;
  case 9: rc = m2c_stage_launch_hllc_sg_mp(in_cons, mode, c->grid3, c->stream, s); break;
  case 10: rc = m2c_stage_launch_hllc_sg_mp_f32(in_cons, mode, c->grid3, c->stream, s); break;
  default: rc = m2c_stage_launch_god_gen(in_cons, mode, c->grid3, c->stream, s); break;
  }
  if (rc == 1) return M2C_ERR_ARG;
  if (rc) return M2C_ERR_HIP;
  HIPCHK(hipGetLastError());
  return M2C_OK;
}

static void launch_cfl(m2c_ctx *c);

static int mp_chk(int rc) { return rc ? M2C_ERR_HIP : M2C_OK; }

// Interface Riemann pass on the stage field S (first stage: keep the Riemann solutions).
// relist = 0: the interface faces listed by an earlier stage of the same step (the IDs change only
// at the end of a step)
static int iface_pass(m2c_ctx *c, const double *S, int in_cons, int store, int ignore_active, int relist = 1) {
  if (!c->iface) return M2C_OK;
  return mp_chk(m2c_mp_launch_iface(c->L, S, in_cons, store && c->ls, ignore_active, relist));
}

// End-of-step tasks of TimeIntegratorBase::UpdateSolutionAfterTimeStepping (TimeIntegrator.cpp:
// 695-797): IDs of swept nodes, reinitialisation, states of cells whose material changed, clip,
// then the CFL minimum of the updated field. `reinit_now`: the host's view of whether this step
// is a reinitialisation step (the device decides again from its own step counter). The head (IDs
// and the first reinitialisation attempt), the host-side retries of the reinitialisation failsafe
// and the tail (state update) are separate so that a batch can run the head of every instance
// before it waits on any of them.
static int end_of_step_head(m2c_ctx *c, bool reinit_now) {
  if (mp_chk(m2c_mp_launch_id_remaining(c->L, c->id_single ? 1 : 0))) return M2C_ERR_HIP;
  if (reinit_now && mp_chk(m2c_mp_launch_reinit(c->L, c->phiA, c->cfg.reinit_maxits, 0))) return M2C_ERR_HIP;
  return M2C_OK;
}

static int end_of_step_retries(m2c_ctx *c, bool reinit_now) {
  if (!reinit_now) return M2C_OK;
  for (int attempt = 0; attempt < 64; attempt++) {
    MpState h;
    HIPCHK(hipMemcpyAsync(&h, c->L.ms, sizeof(MpState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!h.rs.retry) break;
    if (mp_chk(m2c_mp_launch_reinit(c->L, c->phiA, c->cfg.reinit_maxits, 1))) return M2C_ERR_HIP;
  }
  return M2C_OK;
}

static int end_of_step_tail(m2c_ctx *c) {
  return mp_chk(m2c_mp_launch_state_update(c->L, c->V, 4));
}

static int end_of_step(m2c_ctx *c, bool reinit_now) {
  int rc;
  if ((rc = end_of_step_head(c, reinit_now)) || (rc = end_of_step_retries(c, reinit_now)) ||
      (rc = end_of_step_tail(c)))
    return rc;
  launch_cfl(c);
  return M2C_OK;
}

// Stage s of an RK2 (TimeIntegrator.cpp:291-428) or RK3 (:496-690) step: the fluid stage's fields
// and coefficients, U_out = (a U_s + b U_n) + c dt R(S), and the level-set stage interleaved with it
// (phi_out = (a' X + b' Y) + c' dt R(phi_S; V(S)); the last stage also updates the material IDs).
struct StageSpec {
  int in_cons, mode;           // mode 1: intermediate stage (U out), 2: final stage (V out)
  const double *S, *Vn, *Us;
  double *Out;
  double a, b, c;
  int first;
  // level set
  int ls_mode, ls_final;
  const double *phiS, *X, *Y;
  double *phiOut;
  double la, lb, lc;
  const double *phin;
};

static int nstages(const m2c_ctx *c);

static void stage_spec(const m2c_ctx *c, int s, StageSpec &p) {
  memset(&p, 0, sizeof(p));
  const bool rk2 = c->cfg.ts == M2C_TS_RK2;
  if (s == 0) {
    // U1 = U_n + dt R(V_n); Phi1 = (0*Phi1 + 1*Phi) + dt R(Phi; V_n)
    p = {0, 1, c->V, c->V, nullptr, c->U1, 0.0, 0.0, 1.0, 1,
         LS_FIRST, 0, c->phiA, c->phiA, nullptr, c->phiB, 0.0, 0.0, 1.0, nullptr};
  } else if (rk2) {
    // U(n+1) = 0.5 U1 + 0.5 U_n + 0.5 dt R(V1); Phi(n+1) = (0.5 Phi + 0.5 Phi1) + 0.5 dt R(Phi1; V1)
    p = {1, 2, c->U1, c->V, c->U1, c->V, 0.5, 0.5, 0.5, 0,
         LS_AXBY, 1, c->phiB, c->phiA, c->phiB, c->phiA, 0.5, 0.5, 0.5, c->phiA};
  } else if (s == 1) {
    // U2 = 0.25 U1 + 0.75 U_n + 0.25 dt R(V1)
    p = {1, 1, c->U1, c->V, c->U1, c->U2, 0.25, 0.75, 0.25, 0,
         LS_AXBY, 0, c->phiB, c->phiB, c->phiA, c->phiD, 0.25, 0.75, 0.25, nullptr};
  } else {
    // U(n+1) = 2/3 U2 + 1/3 U_n + 2/3 dt R(V2)
    p = {1, 2, c->U2, c->V, c->U2, c->V, 2.0 / 3.0, 1.0 / 3.0, 2.0 / 3.0, 0,
         LS_AXBY, 1, c->phiD, c->phiA, c->phiD, c->phiA, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, c->phiA};
  }
}

// what precedes the fluid stage s: the step's start and the interface Riemann pass (the first stage
// keeps the Riemann solutions and lists the interface faces; later stages reuse the list)
static int stage_before(m2c_ctx *c, int s, const StageSpec &p, int ignore_active) {
  if (s == 0 && (c->iface || c->ls) && mp_chk(m2c_mp_step_begin(c->L))) return M2C_ERR_HIP;
  return iface_pass(c, p.S, p.in_cons, s == 0 ? 1 : 0, ignore_active, s == 0 ? 1 : 0);
}

// what follows the fluid stage s: the level-set stage
static int stage_after(m2c_ctx *c, const StageSpec &p) {
  if (!c->ls) return M2C_OK;
  return mp_chk(m2c_mp_launch_ls(c->L, p.in_cons, p.ls_mode, p.ls_final, p.S, p.phiS, p.X, p.Y, p.phiOut, p.la,
                                 p.lb, p.lc, p.phin));
}

// One time step: the Runge-Kutta stages (TimeIntegrator.cpp FE :150-226, RK2 :291-428,
// RK3 :496-690), with the level-set stages interleaved as there. `ev` (optional) receives a
// start/stop event pair per stage launch. `time_step`: the step's number as the host counts it.
static int enqueue_step(m2c_ctx *c, int ignore_active, std::vector<hipEvent_t> *ev, long long time_step = 0) {
  int ts = c->cfg.ts;
  int rc;
  auto rec = [&](int i) { if (ev) hipEventRecord((*ev)[i], c->stream); };
  const bool reinit_now = c->ls && c->cfg.reinit_freq > 0 && time_step > 0 && time_step % c->cfg.reinit_freq == 0;
  const MpLaunch &L = c->L;
  if (ts == M2C_TS_FE) {
    if (c->iface || c->ls) if (mp_chk(m2c_mp_step_begin(c->L))) return M2C_ERR_HIP;
    if ((rc = iface_pass(c, c->V, 0, 1, ignore_active))) return rc;
    rec(0);
    rc = launch_stage(c, 0, 2, c->V, c->V, nullptr, c->U1, 0, 0, 1.0, 1, ignore_active);
    rec(1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_copy_gated, dim3(2048), dim3(256), 0, c->stream, c->U1, c->V, 5 * c->N, c->clk);
    if (c->ls) {
      // Phi(n+1) = 1*Phi + dt R(Phi; V(n+1)) into B, IDs from sign(Phi(n)) and sign(Phi(n+1))
      if (mp_chk(m2c_mp_launch_ls(L, 0, LS_X, 1, c->V, c->phiA, c->phiA, nullptr, c->phiB, 0, 0, 1.0, c->phiA)))
        return M2C_ERR_HIP;
      hipLaunchKernelGGL(k_copy_gated, dim3(2048), dim3(256), 0, c->stream, c->phiB, c->phiA, c->N, c->clk);
      return end_of_step(c, reinit_now);
    }
    return M2C_OK;
  }
  const int ns = nstages(c);
  for (int s = 0; s < ns; s++) {
    StageSpec p;
    stage_spec(c, s, p);
    if ((rc = stage_before(c, s, p, ignore_active))) return rc;
    rec(2 * s);
    rc = launch_stage(c, p.in_cons, p.mode, p.S, p.Vn, p.Us, p.Out, p.a, p.b, p.c, p.first, ignore_active);
    rec(2 * s + 1);
    if (rc) return rc;
    if ((rc = stage_after(c, p))) return rc;
  }
  return c->ls ? end_of_step(c, reinit_now) : M2C_OK;
}

static int nstages(const m2c_ctx *c) { return c->cfg.ts == M2C_TS_FE ? 1 : (c->cfg.ts == M2C_TS_RK2 ? 2 : 3); }

static void launch_cfl(m2c_ctx *c) {
  int nb = (int)std::min<long long>((c->N + 255) / 256, 4096);
  // the stiffened-gas kernel variants (0, 4, 9, 10) have only stiffened gases and no inactive material
  const bool sg = c->variant == 0 || c->variant == 4 || c->variant == 9 || c->variant == 10;
  if (sg) hipLaunchKernelGGL(k_cfl_min<EOS_SG>, dim3(nb), dim3(256), 0, c->stream, c->V, c->ID, c->clk, c->st);
  else hipLaunchKernelGGL(k_cfl_min<-1>, dim3(nb), dim3(256), 0, c->stream, c->V, c->ID, c->clk, c->st);
}


// Cell centres and widths of every axis (MeshGenerator::ComputeMeshCoordinatesAndDeltas,
// MeshGenerator.cpp:12-131); on a stretched mesh also the device tables of the stage kernels: per
// axis the widths of cells -2..n+1 (a ghost takes its boundary cell's width,
// SpaceOperator::PopulateGhostBoundaryCoordinates, SpaceOperator.cpp:219-300) and the limiter
// coefficients of Reconstructor::Setup (Reconstructor.cpp:53-121) per cell.
static int setup_mesh(m2c_ctx *c) {
  const m2c_config &cfg = c->cfg;
  const double lo[3] = {cfg.x0, cfg.y0, cfg.z0}, hi[3] = {cfg.x1, cfg.y1, cfg.z1};
  const int n[3] = {cfg.nx, cfg.ny, cfg.nz};
  for (int d = 0; d < 3; d++) {
    c->mx[d].assign((size_t)n[d], 0.0);
    c->mw[d].assign((size_t)n[d], 0.0);
    int m = 0;
    if (int rc = m2c_mesh_1d(lo[d], hi[d], n[d], cfg.mesh_npts[d], &cfg.mesh_pts[d][0][0], n[d], c->mx[d].data(),
                             c->mw[d].data(), &m))
      return rc;
    if (m != n[d]) return M2C_ERR_ARG;
  }
  if (!c->grid.stretched) return M2C_OK;
  size_t nd = 0, nk = 0;
  for (int d = 0; d < 3; d++) { nd += (size_t)(n[d] + 4) + 2 * (size_t)n[d]; nk += (size_t)n[d]; }
  std::vector<double> h(nd);
  std::vector<int> hk(nk);
  size_t ow[3], oa[3], ob[3], ok[3], q = 0, qk = 0;
  for (int d = 0; d < 3; d++) {
    const std::vector<double> &w = c->mw[d];
    auto W = [&](int i) { return w[(size_t)std::min(std::max(i, 0), n[d] - 1)]; };   // ghosts: boundary widths
    ow[d] = q;
    for (int i = -2; i <= n[d] + 1; i++) h[q++] = W(i);
    oa[d] = q; q += (size_t)n[d];
    ob[d] = q; q += (size_t)n[d];
    ok[d] = qk; qk += (size_t)n[d];
    for (int i = 0; i < n[d]; i++) {
      const double A = (W(i - 1) + W(i)) / (W(i) + W(i + 1));
      const double B = 2.0 * W(i) / (W(i) + W(i + 1));
      h[oa[d] + i] = A;
      h[ob[d] + i] = B;
      hk[ok[d] + i] = cfg.limiter == M2C_LIM_VANALBADA ? slope_k(A, B) : 0;
    }
  }
  HIPCHK(hipMalloc(&c->d_mesh, sizeof(double) * nd));
  HIPCHK(hipMalloc(&c->d_meshK, sizeof(int) * nk));
  HIPCHK(hipMemcpy(c->d_mesh, h.data(), sizeof(double) * nd, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_meshK, hk.data(), sizeof(int) * nk, hipMemcpyHostToDevice));
  for (int d = 0; d < 3; d++) {
    c->grid.w[d] = c->d_mesh + ow[d];
    c->grid.cA[d] = c->d_mesh + oa[d];
    c->grid.cB[d] = c->d_mesh + ob[d];
    c->grid.cK[d] = c->d_meshK + ok[d];
  }
  return M2C_OK;
}

// Buffers and tables of the interface pass and the level set (per-cell arrays without a ghost
// frame; face arrays keyed by the face's plus cell).
static int setup_multiphase(m2c_ctx *c) {
  if (!c->iface && !c->ls) return M2C_OK;
  const m2c_config *cfg = &c->cfg;
  const Grid &g = c->grid;
  MpParams &mp = c->mp;
  memset(&mp, 0, sizeof(mp));
  MpLaunch &L = c->L;
  memset(&L, 0, sizeof(L));
  const long long N = c->N;
  const long long Nf = (long long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  if (3 * Nf >= (1ll << 32)) return M2C_ERR_UNSUPPORTED;
  c->Nf = Nf;
  L.stream = c->stream;
  L.N = N;
  L.Nf = Nf;
  L.nx = g.nx;
  L.ny = g.ny;
  L.nz = g.nz;
  L.flux = cfg->flux;
  L.ID = c->ID;
  L.st = c->st;
  L.clk = c->clk;
  mp.ls_matid = cfg->ls_matid;
  for (int sd = 0; sd < 6; sd++) mp.ls_bc[sd] = cfg->ls_bc[sd];
  mp.ls_neumann = 1;
  for (int sd = 0; sd < 6; sd++) mp.ls_neumann &= (cfg->ls_bc[sd] == 1);
  L.ls_neumann = mp.ls_neumann;
  mp.reinit_freq = cfg->reinit_freq;
  mp.reinit_maxits = cfg->reinit_maxits;
  mp.reinit_tol = cfg->reinit_tol;
  mp.min_dx = std::min(g.dx, std::min(g.dy, g.dz));
  mp.sign_factor = 1.0 * std::min(g.dx, std::min(g.dy, g.dz));
  const m2c_riemann_params &rp = cfg->riemann;
  mp.rcfg.maxIts_main = rp.maxIts_main; mp.rcfg.maxIts_bracket = rp.maxIts_bracket;
  mp.rcfg.maxIts_shock = rp.maxIts_shock; mp.rcfg.numSteps_rarefaction = rp.numSteps_rarefaction;
  mp.rcfg.tol_main = rp.tol_main; mp.rcfg.tol_shock = rp.tol_shock; mp.rcfg.tol_rarefaction = rp.tol_rarefaction;
  mp.rcfg.min_pressure = rp.min_pressure; mp.rcfg.failure_threshold = rp.failure_threshold;
  mp.rcfg.pressure_at_failure = rp.pressure_at_failure;
  mp.path_cap = 1024;
  mp.riemann_normal = cfg->mp_riemann_normal;
  for (int m = 0; m < cfg->nmat; m++) mp.failsafe_density[m] = cfg->mat[m].failsafe_density;
  int rc;
  if ((rc = mp_alloc(c, &L.ms, 1, 0))) return rc;
  {
    MpState h;
    memset(&h, 0, sizeof(h));
    h.rs.cfl = cfg->reinit_cfl;
    HIPCHK(hipMemcpy(L.ms, &h, sizeof(h), hipMemcpyHostToDevice));
  }
  if (c->iface) {
    long long nthr = std::min<long long>(16384, ((3 * Nf + 63) / 64) * 64);
    L.riemann_threads = (int)std::max<long long>(64, nthr);
    if ((rc = mp_alloc(c, &L.faces, (size_t)(3 * Nf), 0))) return rc;
    if ((rc = mp_alloc(c, &c->Fi, (size_t)(30 * Nf), 0))) return rc;
    L.Fi = c->Fi;
    if ((rc = mp_alloc(c, &L.rsV, (size_t)(30 * Nf), 0))) return rc;
    if ((rc = mp_alloc(c, &L.rsMeta, (size_t)(3 * Nf), 0))) return rc;
    if ((rc = mp_alloc(c, &L.scratch, (size_t)6 * mp.path_cap * L.riemann_threads, 0))) return rc;
  }
  if (c->ls) {
    if ((rc = mp_alloc(c, &c->phiA, (size_t)N, 0))) return rc;
    if ((rc = mp_alloc(c, &c->phiB, (size_t)N, 0))) return rc;
    L.phin = c->phiA;   // Phi(n) throughout a step's fluid stages
    if (cfg->ts == M2C_TS_RK3 && (rc = mp_alloc(c, &c->phiD, (size_t)N, 0))) return rc;
    if ((rc = mp_alloc(c, &L.IDnew, (size_t)N, 0))) return rc;
    if ((rc = mp_alloc(c, &L.rem, (size_t)N, 0))) return rc;
    if ((rc = mp_alloc(c, &L.tagR, (size_t)N, 0))) return rc;
    if ((rc = mp_alloc(c, &L.tag, (size_t)N, 0))) return rc;
    if ((rc = mp_alloc(c, &L.sign, (size_t)N, 0))) return rc;
    if ((rc = mp_alloc(c, &L.phibk, (size_t)N, 0))) return rc;
    if ((rc = mp_alloc(c, &L.rphi1a, (size_t)N, 0))) return rc;
    if ((rc = mp_alloc(c, &L.rphi1b, (size_t)N, 0))) return rc;
    if (cfg->precision == M2C_PREC_MIXED && L.ls_neumann) {   // the mixed form reinitialises in FP32
      L.ls_f32 = 1;
      if ((rc = mp_alloc(c, &L.phif0, (size_t)N, 0))) return rc;
      if ((rc = mp_alloc(c, &L.phif1, (size_t)N, 0))) return rc;
      if ((rc = mp_alloc(c, &L.signf, (size_t)N, 0))) return rc;
    }
    if ((rc = mp_alloc(c, &L.chg, (size_t)N, 0))) return rc;
    if ((rc = mp_alloc(c, &L.cidx, (size_t)N, 0xff))) return rc;
    if ((rc = mp_alloc(c, &L.lvl, (size_t)N, 0))) return rc;
    if ((rc = mp_alloc(c, &L.cnt, (size_t)N, 0))) return rc;
    if ((rc = mp_alloc(c, &L.Vnew, (size_t)(5 * N), 0))) return rc;
    if ((rc = mp_alloc(c, &L.unres, (size_t)N, 0))) return rc;
    // coordinates x[-1..n] (MeshGenerator.cpp:146-158, ghosts SpaceOperator.cpp:188-230) and the
    // FD3 coefficients of every index along each axis (GradientCalculatorFD3.cpp:300-469)
    const int n[3] = {g.nx, g.ny, g.nz};
    std::vector<double> xc, fd;
    size_t xo[3], wo[3], lo_off[3], ro_off[3], r1y_off;
    for (int d = 0; d < 3; d++) {
      xo[d] = xc.size();
      std::vector<double> x(n[d] + 2);
      for (int i = 0; i < n[d]; i++) x[i + 1] = c->mx[d][i];
      x[0] = x[1] - c->mw[d][0];
      x[n[d] + 1] = x[n[d]] + c->mw[d][n[d] - 1];
      xc.insert(xc.end(), x.begin(), x.end());
      const double *xi = x.data() + 1;
      lo_off[d] = fd.size();
      for (int i = 0; i < n[d]; i++) {
        double w[4];
        if (i > 0) {
          double dd[4] = {xi[i - 2] - xi[i], xi[i - 1] - xi[i], 0.0, xi[i + 1] - xi[i]};
          fd_weights(dd, 2, w);
        } else {
          double dx = xi[i] - xi[i - 1];
          w[0] = 0.0; w[1] = -1.0 / dx; w[2] = 1.0 / dx; w[3] = 0.0;
        }
        fd.insert(fd.end(), w, w + 4);
      }
      ro_off[d] = fd.size();
      for (int i = 0; i < n[d]; i++) {
        double w[4];
        if (i < n[d] - 1) {
          double dd[4] = {xi[i - 1] - xi[i], 0.0, xi[i + 1] - xi[i], xi[i + 2] - xi[i]};
          fd_weights(dd, 1, w);
        } else {
          double dx = xi[i + 1] - xi[i];
          w[0] = 0.0; w[1] = -1.0 / dx; w[2] = 1.0 / dx; w[3] = 0.0;
        }
        fd.insert(fd.end(), w, w + 4);
      }
    }
    // cell widths with one ghost per side (the boundary cell's), index i + 1
    for (int d = 0; d < 3; d++) {
      wo[d] = xc.size();
      xc.push_back(c->mw[d][0]);
      xc.insert(xc.end(), c->mw[d].begin(), c->mw[d].end());
      xc.push_back(c->mw[d][n[d] - 1]);
    }
    r1y_off = fd.size();
    {
      const double *yj = xc.data() + xo[1] + 1;
      for (int j = 0; j < n[1]; j++) {
        double dy = yj[j + 1] - yj[j];
        double w[4] = {0.0, -1.0 / dy, 1.0 / dy, 0.0};
        fd.insert(fd.end(), w, w + 4);
      }
    }
    if ((rc = mp_alloc(c, &c->xc, xc.size(), 0))) return rc;
    if ((rc = mp_alloc(c, &c->fd, fd.size(), 0))) return rc;
    HIPCHK(hipMemcpy(c->xc, xc.data(), sizeof(double) * xc.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->fd, fd.data(), sizeof(double) * fd.size(), hipMemcpyHostToDevice));
    mp.stretched = g.stretched;
    for (int d = 0; d < 3; d++) {
      mp.xc[d] = c->xc + xo[d];
      mp.wc[d] = c->xc + wo[d];
      mp.fdl[d] = c->fd + lo_off[d];
      mp.fdr[d] = c->fd + ro_off[d];
    }
    mp.fdr1y = c->fd + r1y_off;
    // uniform interior coefficients (bitwise equal over indices 2..n-3)
    auto uniform4 = [&](size_t off, int nn, double *out) -> int {
      if (nn < 5) return 0;
      const double *w0 = fd.data() + off + 4 * 2;
      for (int i = 3; i <= nn - 3; i++)
        if (memcmp(fd.data() + off + 4 * i, w0, 4 * sizeof(double))) return 0;
      memcpy(out, w0, 4 * sizeof(double));
      return 1;
    };
    for (int d = 0; d < 3; d++) {
      const double *x = xc.data() + xo[d];
      const double hd = x[1] - x[0];
      int e2, ok = std::frexp(hd, &e2) == 0.5;
      for (int i = 1; ok && i <= n[d]; i++) ok = (x[i + 1] - x[i]) == hd;
      mp.dx_pow2[d] = ok;
      mp.inv_dx[d] = ok ? 1.0 / hd : 0.0;
    }
    for (int d = 0; d < 3; d++)
      mp.fd_uni[d] = uniform4(lo_off[d], n[d], mp.fdl_u[d]) & uniform4(ro_off[d], n[d], mp.fdr_u[d]);
    mp.fd_uni_r1y = uniform4(r1y_off, n[1], mp.fdr1y_u);
  }
  return M2C_OK;
}

// ---------------------------------------------------------------- C ABI

extern "C" {

const char *m2c_version(void) { return "m2c_hip 0.1 (gfx950)"; }

// Grid constants, material tables and kernel variant of a configuration (shared with the batched
// translation unit). Returns an m2c_status code.
int m2c_fill_grid(const m2c_config *cfg, Grid *gp, DevEos *eos, int *variant) {
  if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1 || cfg->nmat < 1 || cfg->nmat > M2C_MAX_MAT) return M2C_ERR_ARG;
  if (cfg->flux == M2C_FLUX_GODUNOV && cfg->precision != M2C_PREC_FP64) return M2C_ERR_UNSUPPORTED;
  if (cfg->precision != M2C_PREC_FP64 && cfg->precision != M2C_PREC_MIXED) return M2C_ERR_ARG;
  for (int s = 0; s < 6; s++) {
    int b = cfg->bc[s];
    if (b < M2C_BC_INLET || b > M2C_BC_PERIODIC) return M2C_ERR_ARG;
    if (b == M2C_BC_OVERSET) return M2C_ERR_UNSUPPORTED;
    if ((b == M2C_BC_PERIODIC) != (cfg->bc[s ^ 1] == M2C_BC_PERIODIC)) return M2C_ERR_ARG;
  }
  Grid &g = *gp;
  memset(&g, 0, sizeof(g));
  g.nx = cfg->nx; g.ny = cfg->ny; g.nz = cfg->nz;
  g.N = (long long)cfg->nx * cfg->ny * cfg->nz;
  for (int s = 0; s < 6; s++) g.bc[s] = cfg->bc[s];
  for (int p = 0; p < 5; p++) { g.inlet[p] = cfg->inlet[p]; g.outlet[p] = cfg->outlet[p]; }
  // stretched axes (Mesh.ControlPointX/Y/Z): the configured cell count must be the generator's
  {
    const double lo[3] = {cfg->x0, cfg->y0, cfg->z0}, hi[3] = {cfg->x1, cfg->y1, cfg->z1};
    const int n[3] = {cfg->nx, cfg->ny, cfg->nz};
    for (int d = 0; d < 3; d++) {
      if (cfg->mesh_npts[d] < 0 || cfg->mesh_npts[d] > M2C_MAX_MESH_PTS) return M2C_ERR_ARG;
      if (cfg->mesh_npts[d] == 0) continue;
      int m = 0;
      if (int rc = m2c_mesh_1d(lo[d], hi[d], n[d], cfg->mesh_npts[d], &cfg->mesh_pts[d][0][0], 0, nullptr, nullptr, &m))
        return rc;
      if (m != n[d]) return M2C_ERR_ARG;
      g.stretched = 1;
    }
    // the reference has no periodic axes; their ghost widths would be the far side's
    for (int sd = 0; sd < 6 && g.stretched; sd++)
      if (cfg->bc[sd] == M2C_BC_PERIODIC) return M2C_ERR_UNSUPPORTED;
  }
  // uniform mesh, MeshGenerator.cpp:146-158; volume and areas as SpaceOperator.cpp:166-177
  g.dx = (cfg->x1 - cfg->x0) / cfg->nx;
  g.dy = (cfg->y1 - cfg->y0) / cfg->ny;
  g.dz = (cfg->z1 - cfg->z0) / cfg->nz;
  double vol = g.dx * g.dy * g.dz;
  g.inv_mvol = 1.0 / (-vol);
  g.ax = g.dy * g.dz; g.ay = g.dx * g.dz; g.az = g.dx * g.dy;
  double d[3] = {g.dx, g.dy, g.dz};
  for (int a = 0; a < 3; a++) {   // Reconstructor.cpp:69-81
    g.A[a] = (d[a] + d[a]) / (d[a] + d[a]);
    g.B[a] = 2.0 * d[a] / (d[a] + d[a]);
    g.K[a] = cfg->limiter == M2C_LIM_VANALBADA ? slope_k(g.A[a], g.B[a]) : 0;
  }
  g.flux = cfg->flux; g.limiter = cfg->limiter; g.rec_linear = cfg->rec_linear;
  g.slope_zero = cfg->slope_zero_near_interface;
  g.alpha = cfg->gminmod_alpha; g.roe_delta = cfg->roe_delta;
  g.conrec_depth = cfg->mp_conrec_depth;
  g.inactive_id = cfg->inactive_id; g.nmat = cfg->nmat;
  g.mp_flux = cfg->mp_flux; g.mp_recon = cfg->mp_recon;
  int nactive = 0;
  for (int m = 0; m < cfg->nmat; m++) {
    build_eos(cfg->mat[m], eos[m]);
    if (m != cfg->inactive_id) nactive++;
  }
  g.has_interfaces = nactive > 1;
  bool all_sg = true;
  for (int m = 0; m < cfg->nmat; m++) all_sg = all_sg && cfg->mat[m].eos == M2C_EOS_SG;
  int v;
  // HLLC over one or two stiffened gases: the compile-time stiffened-gas variants, with the
  // material-interface code (9, 10) when the two materials meet
  // (a stretched mesh runs the run-time-EOS variants, which read the per-cell mesh tables)
  if (cfg->flux == M2C_FLUX_HLLC)
    v = (all_sg && cfg->inactive_id < 0 && cfg->nmat <= 2 && !g.stretched) ? (g.has_interfaces ? 9 : 0) : 1;
  else v = cfg->flux == M2C_FLUX_ROE ? 2 : 3;
  if (cfg->precision == M2C_PREC_MIXED) v = (v == 9) ? 10 : v + 4;
  if (cfg->flux == M2C_FLUX_GODUNOV) v = 8;
  *variant = v;
  return M2C_OK;
}

int m2c_ctx_create(const m2c_config *cfg, m2c_ctx **out) {
  if (!cfg || !out) return M2C_ERR_ARG;
  Grid g0;
  DevEos e0[8];
  int v0 = 0;
  if (int rc = m2c_fill_grid(cfg, &g0, e0, &v0)) return rc;
  HIPCHK(hipSetDevice(cfg->device));
  m2c_ctx *c = new m2c_ctx();
  c->cfg = *cfg;
  c->grid = g0;
  memcpy(c->eos, e0, sizeof(e0));
  c->variant = v0;
  c->variant0 = v0;
  c->N = g0.N;
  Grid &g = c->grid;
  c->iface = g.has_interfaces && cfg->mp_flux != 2;
  c->ls = cfg->ls_matid >= 0;
  c->godunov = cfg->flux == M2C_FLUX_GODUNOV;
  int bad = 0;
  // RiemannNormal LevelSet / Average: the reference takes the gradient of the level set of the
  // face's non-zero material (CalculateGradPhiAtCellInterface asserts that it has one)
  if (cfg->mp_riemann_normal < 0 || cfg->mp_riemann_normal > 2) bad = M2C_ERR_ARG;
  if (c->iface && cfg->mp_riemann_normal != 1) {
    if (!c->ls) bad = bad ? bad : M2C_ERR_ARG;
    for (int m = 0; m < cfg->nmat; m++)
      if (m != 0 && m != cfg->ls_matid && m != cfg->inactive_id) bad = bad ? bad : M2C_ERR_UNSUPPORTED;
  }
  if (c->ls) {
    if (cfg->ls_matid >= cfg->nmat || cfg->ls_matid == cfg->inactive_id) bad = M2C_ERR_ARG;
    for (int sd = 0; sd < 6; sd++) {
      int b = cfg->ls_bc[sd];
      if (b != 1 && b != 2 && b != 3 && b != M2C_BC_PERIODIC) bad = bad ? bad : M2C_ERR_ARG;
      if ((b == M2C_BC_PERIODIC) != (cfg->bc[sd] == M2C_BC_PERIODIC)) bad = bad ? bad : M2C_ERR_ARG;
    }
    if (cfg->reinit_freq > 0 && (cfg->reinit_maxits < 1 || !(cfg->reinit_cfl > 0))) bad = bad ? bad : M2C_ERR_ARG;
    int nuni = 0, only = -1;
    for (int m = 0; m < cfg->nmat; m++)
      if (m != cfg->ls_matid && m != cfg->inactive_id) { nuni++; only = m; }
    c->id_single = (nuni == 1 && only == 0);
  }
  // viscous and heat-diffusion terms: two or more cells along every axis (with one the reference
  // reads unpopulated edge ghosts); Sutherland viscosity and heat diffusion need a temperature law
  for (int m = 0; m < cfg->nmat; m++) {
    const m2c_material &mm = cfg->mat[m];
    if (mm.viscosity < 0 || mm.viscosity > 2) bad = bad ? bad : M2C_ERR_ARG;
    const bool d = mm.viscosity != 0 || mm.kappa != 0.0;
    if (d && (g.nx < 2 || g.ny < 2 || g.nz < 2)) bad = bad ? bad : M2C_ERR_UNSUPPORTED;
    if (d && cfg->inactive_id >= 0) bad = bad ? bad : M2C_ERR_UNSUPPORTED;
  }
  {
    bool needT = false;
    for (int m = 0; m < cfg->nmat; m++) needT |= cfg->mat[m].viscosity == 2 || cfg->mat[m].kappa != 0.0;
    for (int m = 0; needT && m < cfg->nmat; m++)
      if (cfg->mat[m].eos == M2C_EOS_JWL || cfg->mat[m].eos == M2C_EOS_DUMMY) bad = bad ? bad : M2C_ERR_ARG;
  }
  if (g.stretched) {   // stretched meshes: no diffusion terms (their face tables assume one spacing)
    for (int m = 0; m < cfg->nmat; m++)
      if (cfg->mat[m].viscosity != 0 || cfg->mat[m].kappa != 0.0) bad = bad ? bad : M2C_ERR_UNSUPPORTED;
  }
  if (bad) {
    delete c;
    return bad;
  }

  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  size_t bytes = sizeof(double) * 5 * (size_t)c->N;
  HIPCHK(hipMalloc(&c->V, bytes));
  HIPCHK(hipMalloc(&c->U1, bytes));
  HIPCHK(hipMalloc(&c->R, bytes));
  if (cfg->ts == M2C_TS_RK3) HIPCHK(hipMalloc(&c->U2, bytes));
  HIPCHK(hipMalloc(&c->ID, (size_t)c->N));
  HIPCHK(hipMemset(c->ID, 0, (size_t)c->N));
  HIPCHK(hipMemset(c->V, 0, bytes));
  HIPCHK(hipMalloc(&c->clk, sizeof(Clock)));
  HIPCHK(hipMalloc(&c->st, sizeof(Status)));
  HIPCHK(hipMemset(c->st, 0, sizeof(Status)));
  if (int rc = setup_mesh(c)) return rc;
  if (int rc = setup_multiphase(c)) return rc;
  if (int rc = setup_diffusion(c)) return rc;
  if (int rc = setup_godunov(c)) return rc;
  if (upload_constants(c)) return M2C_ERR_HIP;
  hipLaunchKernelGGL(k_clock_init, dim3(1), dim3(1), 0, c->stream, c->clk, 0.0, 0.0, cfg->cfl, 0);

  // launch shape: 32x8 column tiles marching up z chunks
  m2c_set_kchunk(c, 0);
  HIPCHK(hipStreamSynchronize(c->stream));
  *out = c;
  return M2C_OK;
}

int m2c_set_kchunk(m2c_ctx *c, int kchunk) {
  if (!c || kchunk < 0) return M2C_ERR_ARG;
  const Grid &g = c->grid;
  const int gx = (g.nx + TX - 1) / TX, gy = (g.ny + TY - 1) / TY;
  if (kchunk == 0) {   // the default: z chunks such that the grid has >= ~2048 blocks
    const long long tiles = (long long)gx * gy;
    const int nzc = (int)std::max<long long>(1, std::min<long long>(g.nz, (2048 + tiles - 1) / tiles));
    kchunk = (g.nz + nzc - 1) / nzc;
  }
  c->kchunk = std::min(kchunk, g.nz);
  c->grid3 = dim3(gx, gy, (g.nz + c->kchunk - 1) / c->kchunk);
  return M2C_OK;
}

int m2c_ctx_destroy(m2c_ctx *c) {
  if (!c) return M2C_OK;
  if (g_resident == c) g_resident = nullptr;
  hipStreamSynchronize(c->stream);
  if (c->own_stream) {   // attached to a batch: take back the context's own stream, clock and status
    c->stream = c->L.stream = c->own_stream;
    c->clk = c->L.clk = c->own_clk;
    c->st = c->L.st = c->own_st;
  }
  if (c->bound) hipHostFree(c->bound);
  hipFree(c->d_mesh); hipFree(c->d_meshK);
  hipFree(c->bk_V); hipFree(c->bk_phi); hipFree(c->bk_ID); hipFree(c->bk_ms);
  hipFree(c->V); hipFree(c->U1); hipFree(c->U2); hipFree(c->R);
  hipFree(c->ID); hipFree(c->clk); hipFree(c->st);
  for (void *p : c->mp_bufs) hipFree(p);
  hipFree(c->eb_faces); hipFree(c->eb_wall); hipFree(c->eb_mask); hipFree(c->eb_scratch);
  hipStreamDestroy(c->stream);
  delete c;
  return M2C_OK;
}

static void free_walls(m2c_ctx *c) {
  hipFree(c->eb_faces); hipFree(c->eb_wall); hipFree(c->eb_mask); hipFree(c->eb_scratch);
  c->eb_faces = nullptr; c->eb_wall = nullptr; c->eb_mask = nullptr; c->eb_scratch = nullptr;
  c->eb_n = 0;
  c->eb_threads = 0;
  c->variant = c->variant0;
}

int m2c_set_embedded_faces(m2c_ctx *c, int n, const int *face, const double *wall, int normal_mode, int recon) {
  if (!c || n < 0 || (n > 0 && (!face || !wall))) return M2C_ERR_ARG;
  if (normal_mode < 0 || normal_mode > 2 || recon < 0 || recon > 1) return M2C_ERR_ARG;
  const Grid &g = c->grid;
  if (n > 0 && g.stretched) return M2C_ERR_UNSUPPORTED;
  if (n > 0) {
    // one active material beside walls: no interfaces or level set, no diffusion terms, no
    // periodic axis (a wrapped ghost cell's cut faces are not tracked)
    if (c->iface || c->ls || c->diff) return M2C_ERR_UNSUPPORTED;
    for (int sd = 0; sd < 6; sd++)
      if (c->cfg.bc[sd] == M2C_BC_PERIODIC) return M2C_ERR_UNSUPPORTED;
  }
  const long long Nf = (long long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  if (3 * Nf >= (1ll << 32)) return M2C_ERR_UNSUPPORTED;
  std::vector<uint8_t> mask;
  std::vector<unsigned> codes((size_t)n);
  if (n > 0) {
    mask.assign((size_t)Nf, 0);
    const int N[3] = {g.nx, g.ny, g.nz};
    for (int t = 0; t < n; t++) {
      const int d = face[4 * t], q[3] = {face[4 * t + 1], face[4 * t + 2], face[4 * t + 3]};
      if (d < 0 || d > 2) return M2C_ERR_ARG;
      for (int a = 0; a < 3; a++) {
        const bool ok = a == d ? (q[a] >= 1 && q[a] <= N[a] - 1) : (q[a] >= 0 && q[a] < N[a]);
        if (!ok) return M2C_ERR_ARG;   // faces on the domain boundary are not cut faces
      }
      for (int v = 0; v < 12; v++)
        if (!std::isfinite(wall[12 * t + v])) return M2C_ERR_ARG;
      const long long key = ((long long)q[2] * (g.ny + 1) + q[1]) * (g.nx + 1) + q[0];
      if (mask[(size_t)key] & (1u << d)) return M2C_ERR_ARG;   // listed twice
      mask[(size_t)key] |= (uint8_t)(1u << d);
      codes[(size_t)t] = (unsigned)((long long)d * Nf + key);
    }
  }
  HIPCHK(hipSetDevice(c->cfg.device));
  HIPCHK(hipStreamSynchronize(c->stream));
  free_walls(c);
  if (n == 0) {
    g_resident = nullptr;
    return ensure_constants(c);
  }
  const int threads = (int)std::min<long long>(16384, ((long long)n + 63) / 64 * 64);
  if (!c->godunov) riemann_only_params(c->cfg, c->mp);
  int rc = M2C_OK;
  if (!c->Fi) {
    c->Nf = Nf;
    if ((rc = mp_alloc(c, &c->Fi, (size_t)(30 * Nf), 0))) return rc;
  }
  if (hipMalloc(&c->eb_faces, sizeof(unsigned) * (size_t)n) != hipSuccess ||
      hipMalloc(&c->eb_wall, sizeof(double) * 12 * (size_t)n) != hipSuccess ||
      hipMalloc(&c->eb_mask, (size_t)Nf) != hipSuccess ||
      hipMalloc(&c->eb_scratch, sizeof(double) * 6 * (size_t)c->mp.path_cap * threads) != hipSuccess ||
      hipMemcpy(c->eb_faces, codes.data(), sizeof(unsigned) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->eb_wall, wall, sizeof(double) * 12 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->eb_mask, mask.data(), (size_t)Nf, hipMemcpyHostToDevice) != hipSuccess) {
    free_walls(c);
    return M2C_ERR_HIP;
  }
  c->eb_n = n;
  c->eb_mode = normal_mode;
  c->eb_recon = recon;
  c->eb_threads = threads;
  // the compile-time stiffened-gas variant has no cut-face code: the generic one of the same flux
  if (c->variant == 0) c->variant = 1;
  else if (c->variant == 4) c->variant = 5;
  g_resident = nullptr;   // the exact solver's constants go up with the context's
  return ensure_constants(c);
}

int m2c_upload_state(m2c_ctx *c, const double *V, const int32_t *id) {
  if (!c || !V) return M2C_ERR_ARG;
  if (int rc0 = ensure_constants(c)) return rc0;
  size_t bytes = sizeof(double) * 5 * (size_t)c->N;
  HIPCHK(hipMemcpyAsync(c->R, V, bytes, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_aos_to_soa, dim3(1024), dim3(256), 0, c->stream, c->R, c->V, c->N);
  if (id) {
    std::vector<int8_t> id8((size_t)c->N);
    for (long long i = 0; i < c->N; i++) {
      if (id[i] < 0 || id[i] >= c->cfg.nmat) return M2C_ERR_ARG;
      id8[i] = (int8_t)id[i];
    }
    HIPCHK(hipMemcpyAsync(c->ID, id8.data(), (size_t)c->N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  } else {
    HIPCHK(hipMemsetAsync(c->ID, 0, (size_t)c->N, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return M2C_OK;
}

int m2c_download_state(m2c_ctx *c, double *V, int32_t *id) {
  if (!c) return M2C_ERR_ARG;
  if (int rc0 = ensure_constants(c)) return rc0;
  size_t bytes = sizeof(double) * 5 * (size_t)c->N;
  if (V) {
    hipLaunchKernelGGL(k_soa_to_aos, dim3(1024), dim3(256), 0, c->stream, c->V, c->R, c->N);
    HIPCHK(hipMemcpyAsync(V, c->R, bytes, hipMemcpyDeviceToHost, c->stream));
  }
  if (id) {
    std::vector<int8_t> id8((size_t)c->N);
    HIPCHK(hipMemcpyAsync(id8.data(), c->ID, (size_t)c->N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (long long i = 0; i < c->N; i++) id[i] = id8[i];
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return M2C_OK;
}

int m2c_compute_residual(m2c_ctx *c, double *R) {
  if (!c) return M2C_ERR_ARG;
  if (int rc0 = ensure_constants(c)) return rc0;
  int rc = iface_pass(c, c->V, 0, 0, 1);
  if (rc) return rc;
  rc = launch_stage(c, 0, 0, c->V, c->V, nullptr, c->U1, 0, 0, 0, 0, 1);
  if (rc) return rc;
  if (R) {
    hipLaunchKernelGGL(k_soa_to_aos, dim3(1024), dim3(256), 0, c->stream, c->U1, c->R, c->N);
    HIPCHK(hipMemcpyAsync(R, c->R, sizeof(double) * 5 * (size_t)c->N, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return M2C_OK;
}

int m2c_time_step_size(m2c_ctx *c, double *dt, double *cfl) {
  if (!c) return M2C_ERR_ARG;
  if (int rc0 = ensure_constants(c)) return rc0;
  Clock h;
  HIPCHK(hipMemcpyAsync(&h, c->clk, sizeof(Clock), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  hipLaunchKernelGGL(k_clock_set, dim3(1), dim3(1), 0, c->stream, c->clk, INFINITY, 1 << 30);
  launch_cfl(c);
  hipLaunchKernelGGL(k_clock_finalize, dim3(1), dim3(1), 0, c->stream, c->clk, 0);
  Clock h2;
  HIPCHK(hipMemcpyAsync(&h2, c->clk, sizeof(Clock), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (dt) *dt = h2.dt;
  if (cfl) *cfl = c->cfg.cfl;
  return M2C_OK;
}

int m2c_advance_dt(m2c_ctx *c, double dt) {
  if (!c) return M2C_ERR_ARG;
  if (int rc0 = ensure_constants(c)) return rc0;
  hipLaunchKernelGGL(k_clock_set, dim3(1), dim3(1), 0, c->stream, c->clk, INFINITY, 1 << 30);
  hipLaunchKernelGGL(k_clock_fixed_dt, dim3(1), dim3(1), 0, c->stream, c->clk, dt);
  Clock h;
  HIPCHK(hipMemcpyAsync(&h, c->clk, sizeof(Clock), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  int rc = enqueue_step(c, 1, nullptr, (long long)h.steps + 1);
  if (rc) return rc;
  hipLaunchKernelGGL(k_clock_finalize, dim3(1), dim3(1), 0, c->stream, c->clk, 1);
  HIPCHK(hipStreamSynchronize(c->stream));
  return M2C_OK;
}

int m2c_run(m2c_ctx *c, double tmax, int maxsteps, int sync_every) {
  if (!c || maxsteps < 0) return M2C_ERR_ARG;
  if (int rc0 = ensure_constants(c)) return rc0;
  hipLaunchKernelGGL(k_clock_set, dim3(1), dim3(1), 0, c->stream, c->clk, tmax, maxsteps);
  launch_cfl(c);
  hipLaunchKernelGGL(k_clock_finalize, dim3(1), dim3(1), 0, c->stream, c->clk, 0);
  long long step0 = 0;
  if (c->ls) {
    Clock h;
    HIPCHK(hipMemcpyAsync(&h, c->clk, sizeof(Clock), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    step0 = h.steps;
  }
  for (int s = 0; s < maxsteps; s++) {
    int rc = enqueue_step(c, 0, nullptr, step0 + s + 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_clock_finalize, dim3(1), dim3(1), 0, c->stream, c->clk, 1);
    if (sync_every > 0 && (s + 1) % sync_every == 0) {
      Clock h;
      HIPCHK(hipMemcpyAsync(&h, c->clk, sizeof(Clock), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      if (!h.active) break;
    }
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return M2C_OK;
}

int m2c_time_steps(m2c_ctx *c, int nsteps, double *ms_stage, double *ms_step) {
  if (!c || nsteps < 1) return M2C_ERR_ARG;
  if (int rc0 = ensure_constants(c)) return rc0;
  int ns = nstages(c);
  std::vector<hipEvent_t> ev(2 * ns * nsteps);
  for (auto &e : ev) HIPCHK(hipEventCreate(&e));
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  hipLaunchKernelGGL(k_clock_set, dim3(1), dim3(1), 0, c->stream, c->clk, INFINITY, nsteps + 1);
  launch_cfl(c);
  hipLaunchKernelGGL(k_clock_finalize, dim3(1), dim3(1), 0, c->stream, c->clk, 0);
  HIPCHK(hipEventRecord(a, c->stream));
  long long step0 = 0;
  if (c->ls) {
    Clock h;
    HIPCHK(hipMemcpyAsync(&h, c->clk, sizeof(Clock), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    step0 = h.steps;
  }
  // the stage events bracket the face pass before the stage (viscous / heat terms, Godunov fluxes)
  // too; its own events let it be taken out
  const bool pre = c->diff || c->godunov;
  std::vector<hipEvent_t> dev(pre ? 2 * ns * nsteps * (1 + (c->diff && c->godunov)) : 0);
  for (auto &e : dev) HIPCHK(hipEventCreate(&e));
  c->diff_ev = pre ? &dev : nullptr;
  c->diff_ev_i = 0;
  for (int s = 0; s < nsteps; s++) {
    std::vector<hipEvent_t> sub(ev.begin() + 2 * ns * s, ev.begin() + 2 * ns * (s + 1));
    int rc = enqueue_step(c, 0, &sub, step0 + s + 1);
    if (rc) {
      c->diff_ev = nullptr;
      return rc;
    }
    hipLaunchKernelGGL(k_clock_finalize, dim3(1), dim3(1), 0, c->stream, c->clk, 1);
  }
  c->diff_ev = nullptr;
  HIPCHK(hipEventRecord(b, c->stream));
  HIPCHK(hipEventSynchronize(b));
  float tot = 0.f;
  HIPCHK(hipEventElapsedTime(&tot, a, b));
  double ks = 0.0, kd = 0.0;
  for (int i = 0; i < ns * nsteps; i++) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]));
    ks += t;
  }
  for (size_t i = 0; i + 1 < c->diff_ev_i; i += 2) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, dev[i], dev[i + 1]));
    kd += t;
  }
  ks -= kd;
  c->diff_ms = c->diff_ev_i ? kd / (ns * nsteps) : 0.0;
  for (auto &e : dev) hipEventDestroy(e);
  for (auto &e : ev) hipEventDestroy(e);
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (ms_stage) *ms_stage = ks / (ns * nsteps);
  if (ms_step) *ms_step = tot / nsteps;
  return M2C_OK;
}

int m2c_time_diffusion_pass(m2c_ctx *c, double *ms) {
  if (!c || !ms) return M2C_ERR_ARG;
  *ms = c->diff_ms;
  return M2C_OK;
}

int m2c_get_status(m2c_ctx *c, m2c_status *out) {
  if (!c || !out) return M2C_ERR_ARG;
  if (int rc0 = ensure_constants(c)) return rc0;
  Status s;
  Clock k;
  HIPCHK(hipMemcpyAsync(&s, c->st, sizeof(Status), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(&k, c->clk, sizeof(Clock), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  memset(out, 0, sizeof(*out));
  out->error = (s.bits & (ST_BAD_STATE | ST_NAN_FLUX | ST_NEG_C2 | ST_RIEMANN_PATH | ST_LIST_OVERFLOW |
                          ST_REINIT_FAIL)) ? M2C_ERR_STATE
             : ((s.bits & ST_CLIP_INTERMEDIATE) ? M2C_ERR_UNSUPPORTED : M2C_OK);
  out->reverted_faces = (long long)s.reverted;
  out->clipped_cells = (long long)s.clipped;
  out->riemann_errors = (long long)s.riemann_err;
  out->steps = k.steps;
  out->t = k.t;
  out->dt = k.dt;
  out->cfl = c->cfg.cfl;
  if (c->ls || c->iface) {
    MpState h;
    HIPCHK(hipMemcpy(&h, c->L.ms, sizeof(MpState), hipMemcpyDeviceToHost));
    out->reinit_iterations = h.reinit_iters;
    out->unresolved_cells = h.unresolved_total;
    out->surrogate_accepted = (long long)h.sur_accepted;
    out->surrogate_fallbacks = (long long)h.sur_fallback;
  }
  return M2C_OK;
}

int m2c_set_surrogate(m2c_ctx *c, const m2c_mlp *net, int newton_max) {
  if (!c || newton_max < 0) return M2C_ERR_ARG;
  if (!net) {
    c->L.sur_on = 0;
    return M2C_OK;
  }
  if (!c->iface) return M2C_ERR_ARG;                      // no material interfaces in this context
  if (c->cfg.mp_flux != 0) return M2C_ERR_UNSUPPORTED;    // the network serves the numerical-flux form
  if (c->godunov) return M2C_ERR_UNSUPPORTED;             // ... of the approximate fluxes
  if (c->cfg.mp_riemann_normal != 1) return M2C_ERR_UNSUPPORTED;   // it solves along the mesh axes
  if (c->cfg.mp_recon != 0) return M2C_ERR_UNSUPPORTED;            // ... between the cell states
  MlpWeights w;
  int dev = -1;
  if (m2c_mlp_get_weights(net, &w, &dev)) return M2C_ERR_ARG;
  if (dev != c->cfg.device) return M2C_ERR_ARG;
  if (int rc0 = ensure_constants(c)) return rc0;
  if (!c->L.surX) {
    const long long cap = std::min<long long>(3 * c->Nf, 1ll << 22);
    int rc;
    if ((rc = mp_alloc(c, &c->L.surX, (size_t)cap * MLP_IN, 0))) return rc;
    if ((rc = mp_alloc(c, &c->L.surY, (size_t)cap * MLP_OUT, 0))) return rc;
    if ((rc = mp_alloc(c, &c->L.surFb, (size_t)cap, 0))) return rc;
    c->L.sur_cap = (int)cap;
  }
  const double tol = c->cfg.riemann.surrogate_residual_tol;
  c->L.sur_w = w;
  c->L.sur_nmax = newton_max;
  c->L.sur_tol = tol > 0.0 ? tol : 1e-6;
  c->L.sur_on = 1;
  return M2C_OK;
}

int m2c_surrogate_riemann_batch(m2c_ctx *c, const m2c_mlp *net, int newton_max, int n, const int *dir,
                                const double *Vm, const int *idm, const double *Vp, const int *idp, double *Vs,
                                int *id, double *Vsm, double *Vsp, int *status, int *path) {
  if (!c || !net || newton_max < 0 || n < 0 ||
      (n > 0 && (!dir || !Vm || !idm || !Vp || !idp || !Vs || !id || !Vsm || !Vsp || !status || !path)))
    return M2C_ERR_ARG;
  if (n == 0) return M2C_OK;
  for (int t = 0; t < n; t++) {
    if (dir[t] < 0 || dir[t] > 2) return M2C_ERR_ARG;
    if (idm[t] < 0 || idm[t] >= c->cfg.nmat || idp[t] < 0 || idp[t] >= c->cfg.nmat) return M2C_ERR_ARG;
  }
  MlpWeights w;
  int dev = -1;
  if (m2c_mlp_get_weights(net, &w, &dev) || dev != c->cfg.device) return M2C_ERR_ARG;
  if (int rc0 = ensure_constants(c)) return rc0;
  if (int rc0 = ensure_riemann_constants(c)) return rc0;
  const double tol = c->cfg.riemann.surrogate_residual_tol > 0.0 ? c->cfg.riemann.surrogate_residual_tol : 1e-6;
  const int nthreads = (int)std::min<long long>(4096, ((long long)n + 63) / 64 * 64);
  const size_t b5 = sizeof(double) * 5 * (size_t)n, bi = sizeof(int) * (size_t)n;
  const size_t bxy = sizeof(float) * (size_t)n * (MLP_IN + MLP_OUT);
  const size_t sbytes = sizeof(double) * 6 * (size_t)c->mp.path_cap * nthreads;
  char *d = nullptr;
  HIPCHK(hipMalloc(&d, 5 * b5 + 7 * bi + sbytes + bxy + 64));
  double *dVm = (double *)d, *dVp = dVm + 5 * (size_t)n, *dVs = dVp + 5 * (size_t)n, *dVsm = dVs + 5 * (size_t)n,
         *dVsp = dVsm + 5 * (size_t)n;
  double *scratch = dVsp + 5 * (size_t)n;
  int *ddir = (int *)(scratch + 6 * (size_t)c->mp.path_cap * nthreads), *didm = ddir + n, *didp = didm + n,
      *did = didp + n, *dst = did + n, *dpath = dst + n;
  unsigned *dfb = (unsigned *)(dpath + n);
  float *dX = (float *)(dfb + n), *dY = dX + (size_t)n * MLP_IN;
  unsigned *dnfb = (unsigned *)(dY + (size_t)n * MLP_OUT);
  int rc = M2C_OK;
  if (hipMemsetAsync(dnfb, 0, sizeof(unsigned), c->stream) != hipSuccess) rc = M2C_ERR_HIP;
  if (rc || hipMemcpyAsync(dVm, Vm, b5, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(dVp, Vp, b5, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(ddir, dir, bi, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(didm, idm, bi, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(didp, idp, bi, hipMemcpyHostToDevice, c->stream) != hipSuccess)
    rc = M2C_ERR_HIP;
  if (!rc && m2c_mp_launch_sur_batch(c->stream, n, ddir, dVm, didm, dVp, didp, dX, dY, w, newton_max, tol, dVs, did,
                                     dVsm, dVsp, dst, dpath, scratch, nthreads, dfb, dnfb))
    rc = M2C_ERR_HIP;
  if (!rc && (hipMemcpyAsync(Vs, dVs, b5, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
              hipMemcpyAsync(Vsm, dVsm, b5, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
              hipMemcpyAsync(Vsp, dVsp, b5, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
              hipMemcpyAsync(id, did, bi, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
              hipMemcpyAsync(status, dst, bi, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
              hipMemcpyAsync(path, dpath, bi, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
              hipStreamSynchronize(c->stream) != hipSuccess))
    rc = M2C_ERR_HIP;
  hipFree(d);
  return rc;
}

int m2c_upload_levelset(m2c_ctx *c, const double *phi) {
  if (!c || !phi) return M2C_ERR_ARG;
  if (!c->ls) return M2C_ERR_ARG;
  if (int rc0 = ensure_constants(c)) return rc0;
  HIPCHK(hipMemcpyAsync(c->phiA, phi, sizeof(double) * (size_t)c->N, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return M2C_OK;
}

int m2c_download_levelset(m2c_ctx *c, double *phi) {
  if (!c || !phi) return M2C_ERR_ARG;
  if (!c->ls) return M2C_ERR_ARG;
  if (int rc0 = ensure_constants(c)) return rc0;
  HIPCHK(hipMemcpyAsync(phi, c->phiA, sizeof(double) * (size_t)c->N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return M2C_OK;
}

int m2c_reset_time(m2c_ctx *c, double t) {
  if (!c) return M2C_ERR_ARG;
  if (int rc0 = ensure_constants(c)) return rc0;
  hipLaunchKernelGGL(k_clock_init, dim3(1), dim3(1), 0, c->stream, c->clk, t, 0.0, c->cfg.cfl, 0);
  HIPCHK(hipMemsetAsync(c->st, 0, sizeof(Status), c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return M2C_OK;
}

int m2c_synchronize(m2c_ctx *c) {
  if (!c) return M2C_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->stream));
  return M2C_OK;
}

int m2c_flux_batch(m2c_ctx *c, int dir, int n, const double *Vm, const double *Vp, int id, double *F) {
  if (!c || n < 0 || dir < 0 || dir > 2 || id < 0 || id >= c->cfg.nmat) return M2C_ERR_ARG;
  if (n == 0) return M2C_OK;
  if (int rc0 = ensure_constants(c)) return rc0;
  double *d;
  size_t b = sizeof(double) * 5 * (size_t)n;
  HIPCHK(hipMalloc(&d, 3 * b));
  HIPCHK(hipMemcpyAsync(d, Vm, b, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d + 5 * (size_t)n, Vp, b, hipMemcpyHostToDevice, c->stream));
  dim3 gb((n + 255) / 256);
  if (c->cfg.flux == M2C_FLUX_GODUNOV) {   // an exact solve per pair: the solver's path cache in scratch
    if (int rc0 = ensure_riemann_constants(c)) { hipFree(d); return rc0; }
    const int nthreads = (int)std::min<long long>(4096, ((long long)n + 63) / 64 * 64);
    double *scr = nullptr;
    HIPCHK(hipMalloc(&scr, sizeof(double) * 6 * (size_t)c->mp.path_cap * nthreads));
    int rc = m2c_mp_launch_godunov_batch(c->stream, dir, n, d, d + 5 * (size_t)n, id, d + 10 * (size_t)n, scr,
                                         nthreads);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipFree(scr));
    if (rc) { hipFree(d); return M2C_ERR_HIP; }
  } else if (c->cfg.flux == M2C_FLUX_HLLC) launch_flux_batch<FLUX_HLLC>(dir, gb, c->stream, n, d, d + 5 * (size_t)n, id, d + 10 * (size_t)n);
  else if (c->cfg.flux == M2C_FLUX_ROE) launch_flux_batch<FLUX_ROE>(dir, gb, c->stream, n, d, d + 5 * (size_t)n, id, d + 10 * (size_t)n);
  else launch_flux_batch<FLUX_LLF>(dir, gb, c->stream, n, d, d + 5 * (size_t)n, id, d + 10 * (size_t)n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(F, d + 10 * (size_t)n, b, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipFree(d));
  return M2C_OK;
}

// ---------------------------------------------------------------- pointwise VarFcnBase / interface flux

// VarFcnBase's pointwise set (VarFcnBase.h:62-294) of material id[t] at point t; the operation and
// the in/out strides are described with m2c_varfcn_op. CheckState and clipping also report a flag.
__global__ void k_varfcn(int op, int n, const int *__restrict__ id, const double *__restrict__ in,
                         double *__restrict__ out, int *__restrict__ flag) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const DevEos &m = c_eos[id[t]];
  switch (op) {
  case M2C_VF_PRESSURE: out[t] = eos_p<double, -1>(m, in[2 * t], in[2 * t + 1]); break;
  case M2C_VF_INTERNAL_ENERGY: out[t] = eos_e<double, -1>(m, in[2 * t], in[2 * t + 1]); break;
  case M2C_VF_DPDRHO: out[t] = eos_dpdrho<double, -1>(m, in[2 * t], in[2 * t + 1]); break;
  case M2C_VF_BIG_GAMMA: out[t] = eos_G<double, -1>(m, in[2 * t]); break;
  case M2C_VF_SOUND_SPEED_SQUARE:
    // the dummy material's GetDpdrho/GetBigGamma return 0 and its c^2 is DBL_MIN (VarFcnDummy.h:43)
    out[t] = eos_c2<double, -1>(m, in[2 * t], in[2 * t + 1]);
    break;
  case M2C_VF_SOUND_SPEED: {   // the reference exits for c^2 <= 0 (VarFcnBase.h:203-209): NaN and flag 1
    const double c2 = eos_c2<double, -1>(m, in[2 * t], in[2 * t + 1]);
    out[t] = c2 <= 0 ? __builtin_nan("") : __builtin_sqrt(c2);
    if (flag) flag[t] = c2 <= 0;
    break;
  }
  case M2C_VF_ENTHALPY: {      // h = e + p/rho (VarFcnBase.h:240-244)
    const double r = in[2 * t], pp = in[2 * t + 1];
    out[t] = eos_e<double, -1>(m, r, pp) + pp / r;
    break;
  }
  case M2C_VF_TOTAL_ENTHALPY: out[t] = eos_H<double, -1>(m, in + 5 * t); break;
  case M2C_VF_MACH: {          // VarFcnBase.h:224-236; c^2 < 0 exits in the reference: NaN and flag 1
    const double *V = in + 5 * t;
    double c = eos_c2<double, -1>(m, V[0], eos_e<double, -1>(m, V[0], V[4]));
    const bool bad = c < 0;
    c = bad ? __builtin_nan("") : __builtin_sqrt(c);
    out[t] = __builtin_sqrt(V[1] * V[1] + V[2] * V[2] + V[3] * V[3]) / c;
    if (flag) flag[t] = bad;
    break;
  }
  case M2C_VF_C2P: {
    double V[5];
    eos_c2p<-1>(m, in + 5 * t, V);
#pragma unroll
    for (int q = 0; q < 5; q++) out[5 * t + q] = V[q];
    break;
  }
  case M2C_VF_P2C: {
    double U[5];
    eos_p2c<double, -1>(m, in + 5 * t, U);
#pragma unroll
    for (int q = 0; q < 5; q++) out[5 * t + q] = U[q];
    break;
  }
  case M2C_VF_CHECK_STATE: {
    const int bad = eos_bad<double, -1>(m, in + 5 * t) ? 1 : 0;
    out[t] = bad;
    if (flag) flag[t] = bad;
    break;
  }
  case M2C_VF_CLIP: {          // ClipDensityAndPressure(V) (VarFcnBase.h:259-294); the dummy resets V
    double V[5];
#pragma unroll
    for (int q = 0; q < 5; q++) V[q] = in[5 * t + q];
    const int cl = eos_clip<double, -1>(m, V) ? 1 : 0;
#pragma unroll
    for (int q = 0; q < 5; q++) out[5 * t + q] = V[q];
    if (flag) flag[t] = cl;
    break;
  }
  default: break;
  }
}

static int varfcn_strides(int op, int *sin, int *sout) {
  switch (op) {
  case M2C_VF_PRESSURE: case M2C_VF_INTERNAL_ENERGY: case M2C_VF_DPDRHO: case M2C_VF_BIG_GAMMA:
  case M2C_VF_SOUND_SPEED_SQUARE: case M2C_VF_SOUND_SPEED: case M2C_VF_ENTHALPY:
    *sin = 2; *sout = 1; return 0;
  case M2C_VF_TOTAL_ENTHALPY: case M2C_VF_MACH: case M2C_VF_CHECK_STATE:
    *sin = 5; *sout = 1; return 0;
  case M2C_VF_C2P: case M2C_VF_P2C: case M2C_VF_CLIP:
    *sin = 5; *sout = 5; return 0;
  default: return -1;
  }
}

int m2c_varfcn_batch(m2c_ctx *c, int op, int n, const int32_t *id, const double *in, double *out, int32_t *flag) {
  int sin, sout;
  if (!c || n < 0 || varfcn_strides(op, &sin, &sout)) return M2C_ERR_ARG;
  if (n > 0 && (!id || !in || !out)) return M2C_ERR_ARG;
  for (int t = 0; t < n; t++)
    if (id[t] < 0 || id[t] >= c->cfg.nmat) return M2C_ERR_ARG;
  if (n == 0) return M2C_OK;
  if (int rc0 = ensure_constants(c)) return rc0;
  const size_t bin = sizeof(double) * sin * (size_t)n, bout = sizeof(double) * sout * (size_t)n,
               bi = sizeof(int) * (size_t)n;
  char *d = nullptr;
  HIPCHK(hipMalloc(&d, bin + bout + 2 * bi));
  double *din = (double *)d, *dout = din + sin * (size_t)n;
  int *did = (int *)(dout + sout * (size_t)n), *dfl = did + n;
  HIPCHK(hipMemcpyAsync(din, in, bin, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(did, id, bi, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(dfl, 0, bi, c->stream));
  hipLaunchKernelGGL(k_varfcn, dim3((n + 255) / 256), dim3(256), 0, c->stream, op, n, did, din, dout, dfl);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, dout, bout, hipMemcpyDeviceToHost, c->stream));
  if (flag) HIPCHK(hipMemcpyAsync(flag, dfl, bi, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipFree(d));
  return M2C_OK;
}

// FluxFcnLLF::ComputeNumericalFluxAtMaterialInterface (FluxFcnLLF.h:112-185), the flux the
// reference's SpaceOperator takes between two materials with MultiPhase.Flux = LLF
__global__ void k_iface_flux_batch(int dir, int n, const double *__restrict__ Vm, const int *__restrict__ idm,
                                   const double *__restrict__ Vp, const int *__restrict__ idp, double *__restrict__ F) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double a[5], b[5], f[5];
#pragma unroll
  for (int q = 0; q < 5; q++) { a[q] = Vm[5 * t + q]; b[q] = Vp[5 * t + q]; }
  const DevEos &mm = c_eos[idm[t]], &mp = c_eos[idp[t]];
  if (dir == 0) flux_llf2<double, -1, -1, 0>(mm, mp, a, b, f);
  else if (dir == 1) flux_llf2<double, -1, -1, 1>(mm, mp, a, b, f);
  else flux_llf2<double, -1, -1, 2>(mm, mp, a, b, f);
#pragma unroll
  for (int q = 0; q < 5; q++) F[5 * t + q] = f[q];
}

int m2c_interface_flux_batch(m2c_ctx *c, int dir, int n, const double *Vm, const int32_t *idm, const double *Vp,
                             const int32_t *idp, double *F) {
  if (!c || n < 0 || dir < 0 || dir > 2) return M2C_ERR_ARG;
  if (n > 0 && (!Vm || !idm || !Vp || !idp || !F)) return M2C_ERR_ARG;
  if (!riemann_batch_args_ok(c, n, idm, idp)) return M2C_ERR_ARG;
  if (n == 0) return M2C_OK;
  if (int rc0 = ensure_constants(c)) return rc0;
  const size_t b = sizeof(double) * 5 * (size_t)n, bi = sizeof(int) * (size_t)n;
  char *d = nullptr;
  HIPCHK(hipMalloc(&d, 3 * b + 2 * bi));
  double *dm = (double *)d, *dp = dm + 5 * (size_t)n, *dF = dp + 5 * (size_t)n;
  int *dim_ = (int *)(dF + 5 * (size_t)n), *dip = dim_ + n;
  HIPCHK(hipMemcpyAsync(dm, Vm, b, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(dp, Vp, b, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(dim_, idm, bi, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(dip, idp, bi, hipMemcpyHostToDevice, c->stream));
  const dim3 g((n + 255) / 256), bl(256);
  hipLaunchKernelGGL(k_iface_flux_batch, g, bl, 0, c->stream, dir, n, dm, dim_, dp, dip, dF);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(F, dF, b, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipFree(d));
  return M2C_OK;
}

// the temperature-law constants of the materials (TLaw, physics.hpp)
static TLaw build_tlaw(const m2c_config &cfg) {
  TLaw t;
  memset(&t, 0, sizeof(t));
  for (int m = 0; m < cfg.nmat; m++) {
    const m2c_material &mm = cfg.mat[m];
    const double cv = mm.tlaw[0], cp = mm.tlaw[4];
    t.invcv[m] = cv == 0.0 ? 0.0 : 1.0 / cv;
    t.invcp[m] = cp == 0.0 ? 0.0 : 1.0 / cp;
    t.T0[m] = mm.tlaw[1]; t.e0[m] = mm.tlaw[2]; t.rho0[m] = mm.tlaw[3]; t.h0[m] = mm.tlaw[5];
    if (mm.eos == M2C_EOS_SG) t.method[m] = (cv > 0 && mm.tlaw[3] > 0) ? 3 : ((cp > 0 && cv <= 0.0) ? 2 : 1);
    else if (mm.eos == M2C_EOS_MG) { t.method[m] = (cp > 0 && cv <= 0.0) ? 2 : 1; t.e0[m] = mm.p[4]; }
    else if (mm.eos == M2C_EOS_NASG) t.method[m] = 4;
  }
  return t;
}

// e(rho, p) of each point's material and, with a temperature law, T(rho, e)
__global__ void k_eos_derived(int n, const double *__restrict__ rho, const double *__restrict__ p,
                              const int *__restrict__ id, TLaw t, double *__restrict__ e, double *__restrict__ T) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = id[i];
  const DevEos &m = c_eos[k];
  const double r = rho[i];
  const double ee = eos_e<double, -1>(m, r, p[i]);
  e[i] = ee;
  T[i] = tlaw_temperature(t, m, k, r, ee);
}

int m2c_eos_derived(m2c_ctx *c, int n, const double *rho, const double *p, const int32_t *id, double *e, double *T) {
  if (!c || n < 0 || (n > 0 && (!rho || !p || !id || !e))) return M2C_ERR_ARG;
  for (int i = 0; i < n; i++)
    if (id[i] < 0 || id[i] >= c->cfg.nmat) return M2C_ERR_ARG;
  if (n == 0) return M2C_OK;
  if (int rc0 = ensure_constants(c)) return rc0;
  const size_t b = sizeof(double) * (size_t)n;
  char *d;
  HIPCHK(hipMalloc(&d, 4 * b + sizeof(int) * (size_t)n));
  double *dr = (double *)d, *dp = dr + n, *de = dp + n, *dT = de + n;
  int *di = (int *)(dT + n);
  HIPCHK(hipMemcpyAsync(dr, rho, b, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(dp, p, b, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(di, id, sizeof(int) * (size_t)n, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_eos_derived, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, dr, dp, di, build_tlaw(c->cfg),
                     de, dT);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(e, de, b, hipMemcpyDeviceToHost, c->stream));
  if (T) HIPCHK(hipMemcpyAsync(T, dT, b, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipFree(d));
  return M2C_OK;
}

// the exact solver's constants live with the interface kernels: a context without interfaces
// uploads them before a pointwise batch
static void riemann_only_params(const m2c_config &cfg, MpParams &mp) {
  memset(&mp, 0, sizeof(mp));
  const m2c_riemann_params &rp = cfg.riemann;
  mp.rcfg.maxIts_main = rp.maxIts_main; mp.rcfg.maxIts_bracket = rp.maxIts_bracket;
  mp.rcfg.maxIts_shock = rp.maxIts_shock; mp.rcfg.numSteps_rarefaction = rp.numSteps_rarefaction;
  mp.rcfg.tol_main = rp.tol_main; mp.rcfg.tol_shock = rp.tol_shock; mp.rcfg.tol_rarefaction = rp.tol_rarefaction;
  mp.rcfg.min_pressure = rp.min_pressure; mp.rcfg.failure_threshold = rp.failure_threshold;
  mp.rcfg.pressure_at_failure = rp.pressure_at_failure;
  mp.path_cap = 1024;
  mp.riemann_normal = 1;
}

static int ensure_riemann_constants(m2c_ctx *c) {
  if (c->iface || c->ls || c->godunov || c->eb_n) return M2C_OK;   // uploaded with the context's constants
  riemann_only_params(c->cfg, c->mp);
  if (m2c_mp_upload(c->grid, c->eos, c->mp)) return M2C_ERR_HIP;
  return M2C_OK;
}

// Buffers of the Godunov face pass: five flux components per face of each direction (faces indexed
// (k, j, i) with n_d + 1 planes along d) and the exact solver's rarefaction path cache, one slice
// per thread of the pass.
static int setup_godunov(m2c_ctx *c) {
  if (!c->godunov) return M2C_OK;
  const Grid &g = c->grid;
  if (!c->iface && !c->ls) riemann_only_params(c->cfg, c->mp);
  c->gnf[0] = (long long)(g.nx + 1) * g.ny * g.nz;
  c->gnf[1] = (long long)g.nx * (g.ny + 1) * g.nz;
  c->gnf[2] = (long long)g.nx * g.ny * (g.nz + 1);
  int rc;
  for (int d = 0; d < 3; d++)
    if ((rc = mp_alloc(c, &c->G[d], (size_t)(5 * c->gnf[d]), 0))) return rc;
  const long long mx = std::max(c->gnf[0], std::max(c->gnf[1], c->gnf[2]));
  // every face is an iterative solve: one wave per SIMD of the 256 CUs (65 536 threads, 48 KB of
  // path cache each); three waves per SIMD (the pass's 168 VGPRs allow it) measured no faster
  c->gthreads = (int)std::max<long long>(64, std::min<long long>(65536, (mx + 63) / 64 * 64));
  return mp_alloc(c, &c->gscratch, (size_t)6 * c->mp.path_cap * c->gthreads, 0);
}

static bool riemann_batch_args_ok(const m2c_ctx *c, int n, const int *idm, const int *idp) {
  for (int t = 0; t < n; t++)
    if (idm[t] < 0 || idm[t] >= c->cfg.nmat || idp[t] < 0 || idp[t] >= c->cfg.nmat) return false;
  return true;
}

int m2c_riemann_batch(m2c_ctx *c, int n, const double *dir, const double *Vm, const int *idm, const double *Vp,
                      const int *idp, double *Vs, int *id, double *Vsm, double *Vsp, int *status) {
  if (!c || n < 0 || (n > 0 && (!dir || !Vm || !idm || !Vp || !idp || !Vs || !id || !Vsm || !Vsp || !status)))
    return M2C_ERR_ARG;
  if (n == 0) return M2C_OK;
  if (!riemann_batch_args_ok(c, n, idm, idp)) return M2C_ERR_ARG;
  if (int rc0 = ensure_constants(c)) return rc0;
  if (int rc0 = ensure_riemann_constants(c)) return rc0;
  const int nthreads = (int)std::min<long long>(4096, ((long long)n + 63) / 64 * 64);
  size_t b5 = sizeof(double) * 5 * (size_t)n, b3 = sizeof(double) * 3 * (size_t)n, bi = sizeof(int) * (size_t)n;
  size_t sbytes = sizeof(double) * 6 * (size_t)c->mp.path_cap * nthreads;
  char *d = nullptr;
  HIPCHK(hipMalloc(&d, 5 * b5 + b3 + 4 * bi + sbytes));
  double *dVm = (double *)d, *dVp = dVm + 5 * (size_t)n, *dVs = dVp + 5 * (size_t)n, *dVsm = dVs + 5 * (size_t)n,
         *dVsp = dVsm + 5 * (size_t)n, *ddir = dVsp + 5 * (size_t)n, *scratch = ddir + 3 * (size_t)n;
  int *didm = (int *)(scratch + 6 * (size_t)c->mp.path_cap * nthreads), *didp = didm + n, *did = didp + n,
      *dst = did + n;
  HIPCHK(hipMemcpyAsync(dVm, Vm, b5, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(dVp, Vp, b5, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(ddir, dir, b3, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(didm, idm, bi, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(didp, idp, bi, hipMemcpyHostToDevice, c->stream));
  int rc = m2c_mp_launch_riemann_batch(c->stream, n, ddir, dVm, didm, dVp, didp, dVs, did, dVsm, dVsp, dst, scratch,
                                       nthreads);
  if (rc) { hipFree(d); return M2C_ERR_HIP; }
  HIPCHK(hipMemcpyAsync(Vs, dVs, b5, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(Vsm, dVsm, b5, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(Vsp, dVsp, b5, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(id, did, bi, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(status, dst, bi, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipFree(d));
  return M2C_OK;
}

// ---------------------------------------------------------------- m2c_riemann_batch_noalloc
//
// The exact solver's rarefaction-path cache for callers that cannot allocate: a fixed device
// array (each GPU the library runs on holds its own copy until the process ends) serving
// kNoAllocThreads concurrent solves of kNoAllocPathCap entries per side. Its device address is
// looked up once per device; host code passes that address, never the array's name.
constexpr int kNoAllocThreads = 512;
constexpr int kNoAllocPathCap = 1024;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[(size_t)6 * kNoAllocPathCap * kNoAllocThreads];
static double *g_scratch_ptr[kMaxDevices];

static int noalloc_scratch(int device, double **out) {
  if (device < 0 || device >= kMaxDevices) return M2C_ERR_ARG;
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = (double *)ptr;
  }
  *out = g_scratch_ptr[device];
  return M2C_OK;
}

int m2c_riemann_batch_noalloc(m2c_ctx *c, int n, const double *dir, const double *Vm, const int *idm, const double *Vp,
                              const int *idp, double *Vs, int *id, double *Vsm, double *Vsp, int *status) {
  if (!c || n < 0 || (n > 0 && (!dir || !Vm || !idm || !Vp || !idp || !Vs || !id || !Vsm || !Vsp || !status)))
    return M2C_ERR_ARG;
  if (c->mp.path_cap > kNoAllocPathCap) {
    fprintf(stderr, "m2c_riemann_batch_noalloc: rarefaction path cache of %d entries exceeds the fixed %d\n",
            c->mp.path_cap, kNoAllocPathCap);
    return M2C_ERR_ARG;
  }
  if (n == 0) return M2C_OK;
  if (int rc0 = ensure_constants(c)) return rc0;   // sets the context's device current
  if (int rc0 = ensure_riemann_constants(c)) return rc0;
  double *scr = nullptr;
  if (int rc0 = noalloc_scratch(c->cfg.device, &scr)) return rc0;
  const int nthreads = (int)std::min<long long>(kNoAllocThreads, ((long long)n + 63) / 64 * 64);
  if (m2c_mp_launch_riemann_batch(c->stream, n, dir, Vm, idm, Vp, idp, Vs, id, Vsm, Vsp, status, scr, nthreads))
    return M2C_ERR_HIP;
  HIPCHK(hipStreamSynchronize(c->stream));
  return M2C_OK;
}

// shared by the one-sided entry points: face = 0 the plain solver, 1 a wall-cut face
static int onesided_batch(m2c_ctx *c, int face, int n, const double *dir, const double *Vm, const int *idl,
                          const double *uw, const int *ax, const int *fb, int mode, double *Vs, int *id, double *Vsm,
                          double *F, int *status) {
  if (n == 0) return M2C_OK;
  for (int t = 0; t < n; t++) {
    if (idl[t] < 0 || idl[t] >= c->cfg.nmat) return M2C_ERR_ARG;
    if (face && (ax[t] < 0 || ax[t] > 2 || (fb[t] != 1 && fb[t] != -1))) return M2C_ERR_ARG;
  }
  if (int rc0 = ensure_constants(c)) return rc0;
  if (int rc0 = ensure_riemann_constants(c)) return rc0;
  const int nthreads = (int)std::min<long long>(4096, ((long long)n + 63) / 64 * 64);
  const size_t b5 = sizeof(double) * 5 * (size_t)n, b3 = sizeof(double) * 3 * (size_t)n, bi = sizeof(int) * (size_t)n;
  const size_t sbytes = sizeof(double) * 6 * (size_t)c->mp.path_cap * nthreads;
  char *d = nullptr;
  HIPCHK(hipMalloc(&d, 4 * b5 + 2 * b3 + 5 * bi + 16 + sbytes));
  double *dVm = (double *)d, *dVs = dVm + 5 * (size_t)n, *dVsm = dVs + 5 * (size_t)n, *dF = dVsm + 5 * (size_t)n;
  double *ddir = dF + 5 * (size_t)n, *duw = ddir + 3 * (size_t)n;
  int *didl = (int *)(duw + 3 * (size_t)n), *dax = didl + n, *dfb = dax + n, *did = dfb + n, *dst = did + n;
  double *scratch = (double *)(((uintptr_t)(dst + n) + 15) & ~(uintptr_t)15);
  HIPCHK(hipMemcpyAsync(dVm, Vm, b5, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(ddir, dir, b3, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(duw, uw, b3, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(didl, idl, bi, hipMemcpyHostToDevice, c->stream));
  if (face) {
    HIPCHK(hipMemcpyAsync(dax, ax, bi, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(dfb, fb, bi, hipMemcpyHostToDevice, c->stream));
  }
  int rc = m2c_mp_launch_onesided_batch(c->stream, c->cfg.flux, face, n, ddir, dVm, didl, duw, dax, dfb, mode, dVs,
                                        did, dVsm, dF, dst, scratch, nthreads);
  if (rc) { hipFree(d); return M2C_ERR_HIP; }
  if (Vs) HIPCHK(hipMemcpyAsync(Vs, dVs, b5, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(Vsm, dVsm, b5, hipMemcpyDeviceToHost, c->stream));
  if (face) HIPCHK(hipMemcpyAsync(F, dF, b5, hipMemcpyDeviceToHost, c->stream));
  if (id) HIPCHK(hipMemcpyAsync(id, did, bi, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(status, dst, bi, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipFree(d));
  return M2C_OK;
}

int m2c_onesided_riemann_batch(m2c_ctx *c, int n, const double *dir, const double *Vm, const int *idl,
                               const double *Ustar, double *Vs, int *id, double *Vsm, int *status) {
  if (!c || n < 0 || (n > 0 && (!dir || !Vm || !idl || !Ustar || !Vs || !id || !Vsm || !status))) return M2C_ERR_ARG;
  return onesided_batch(c, 0, n, dir, Vm, idl, Ustar, nullptr, nullptr, 1, Vs, id, Vsm, nullptr, status);
}

int m2c_wall_face_flux_batch(m2c_ctx *c, int n, const int *axis, const int *side, int normal_mode,
                             const double *nwall, const double *vwall, const double *V, const int *id, double *F,
                             double *Vstar, int *status) {
  if (!c || n < 0 || normal_mode < 0 || normal_mode > 2) return M2C_ERR_ARG;
  if (n > 0 && (!axis || !side || !nwall || !vwall || !V || !id || !F || !Vstar || !status)) return M2C_ERR_ARG;
  return onesided_batch(c, 1, n, nwall, V, id, vwall, axis, side, normal_mode, nullptr, nullptr, Vstar, F, status);
}

int m2c_stream_copy(m2c_ctx *c, long long nbytes, int reps, double *ms_avg) {
  if (!c || nbytes < 16 || reps < 1) return M2C_ERR_ARG;
  if (int rc0 = ensure_constants(c)) return rc0;
  long long n = nbytes / 16;
  double2 *a = nullptr, *b = nullptr;
  HIPCHK(hipMalloc(&a, n * 16));
  HIPCHK(hipMalloc(&b, n * 16));
  HIPCHK(hipMemsetAsync(a, 0, n * 16, c->stream));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  int nb = 256 * 8 * 4;
  hipLaunchKernelGGL(k_stream_copy, dim3(nb), dim3(256), 0, c->stream, a, b, n);   // warm
  HIPCHK(hipEventRecord(e0, c->stream));
  for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_stream_copy, dim3(nb), dim3(256), 0, c->stream, a, b, n);
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  HIPCHK(hipFree(a));
  HIPCHK(hipFree(b));
  if (ms_avg) *ms_avg = ms / reps;
  return M2C_OK;
}

int m2c_device_state(m2c_ctx *c, double **V) {
  if (!c || !V) return M2C_ERR_ARG;
  *V = c->V;
  return M2C_OK;
}


// ---------------------------------------------------------------- two-material instances of a batch

int m2c_ctx_batch_attach(m2c_ctx *c, hipStream_t s, Clock *clk, Status *st) {
  if (!c || !s || !clk || !st || c->own_stream) return M2C_ERR_ARG;
  if (c->diff || c->godunov || c->eb_n || c->L.sur_on || c->cfg.ts == M2C_TS_FE) return M2C_ERR_UNSUPPORTED;
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipHostMalloc((void **)&c->bound, sizeof(m2c_ctx::Bound), hipHostMallocDefault));
  c->bound->g = c->grid;
  memcpy(c->bound->eos, c->eos, sizeof(c->eos));
  c->bound->mp = c->mp;
  c->own_stream = c->stream;
  c->own_clk = c->clk;
  c->own_st = c->st;
  c->stream = c->L.stream = s;
  c->clk = c->L.clk = clk;
  c->st = c->L.st = st;
  if (g_resident == c) g_resident = nullptr;
  return M2C_OK;
}

int m2c_ctx_batch_stage_args(m2c_ctx *c, int s, StageArgs *a, const double **V, const int8_t **ID, int *variant) {
  if (!c || !a || s < 0 || s >= nstages(c)) return M2C_ERR_ARG;
  StageSpec p;
  stage_spec(c, s, p);
  StageArgs q;
  q.phin = conrec_phi(c);
  q.no_cfl = c->ls ? 1 : 0;
  q.S = p.S; q.Vn = p.Vn; q.Us = p.Us; q.Out = p.Out; q.ID = c->ID;
  q.Fi = c->Fi; q.Nf = c->Nf;
  q.clk = c->clk; q.st = c->st;
  q.a = p.a; q.b = p.b; q.c = p.c; q.first = p.first;
  q.kchunk = c->kchunk;
  q.ignore_active = 0;
  *a = q;
  if (V) *V = c->V;
  if (ID) *ID = c->ID;
  if (variant) *variant = c->variant;
  return M2C_OK;
}

int m2c_ctx_batch_bind(m2c_ctx *c) {
  if (!c || !c->bound) return M2C_ERR_ARG;
  // the single-context entry points re-upload their tables on their next call
  g_resident = nullptr;
  return mp_chk(m2c_mp_upload_async(c->stream, &c->bound->g, c->bound->eos, &c->bound->mp));
}

int m2c_ctx_batch_before(m2c_ctx *c, int s) {
  StageSpec p;
  stage_spec(c, s, p);
  return stage_before(c, s, p, 0);
}

int m2c_ctx_batch_iface_inst(m2c_ctx *c, IfaceInst *out, int *ok, int *flux) {
  if (!c || !out || !ok || !flux || !c->bound) return M2C_ERR_ARG;
  memset(out, 0, sizeof(*out));
  const MpLaunch &L = c->L;
  *flux = L.flux;
  // an instance without interfaces takes no part (nblk = 0)
  *ok = !c->iface || (!L.sur_on && (L.flux == FLUX_HLLC || L.flux == FLUX_ROE || L.flux == FLUX_LLF) &&
                      nstages(c) <= 3 && L.riemann_threads >= 64 && L.riemann_threads % 64 == 0);
  out->grid = c->bound->g;
  out->mp = c->bound->mp;
  for (int m = 0; m < 8; m++) out->eos[m] = c->bound->eos[m];
  for (int s = 0; s < nstages(c) && s < 3; s++) {
    StageSpec p;
    stage_spec(c, s, p);
    out->S[s] = p.S;
    if (p.in_cons != (s == 0 ? 0 : 1)) *ok = 0;
  }
  out->ID = L.ID;
  out->faces = L.faces;
  out->ms = L.ms;
  out->Fi = L.Fi;
  out->rsV = L.rsV;
  out->rsMeta = L.rsMeta;
  out->scratch = L.scratch;
  out->st = L.st;
  out->clk = L.clk;
  out->phin = L.phin;
  out->store = c->ls ? 1 : 0;
  out->nblk = c->iface ? L.riemann_threads / 64 : 0;
  out->lnblk = m2c_mp_iface_list_blocks(c->grid.nx, c->grid.ny, c->grid.nz);
  return M2C_OK;
}

int m2c_ctx_batch_step_begin(m2c_ctx *c) {
  if ((c->iface || c->ls) && mp_chk(m2c_mp_step_begin(c->L))) return M2C_ERR_HIP;
  return M2C_OK;
}

int m2c_ctx_batch_after(m2c_ctx *c, int s) {
  StageSpec p;
  stage_spec(c, s, p);
  return stage_after(c, p);
}

static bool reinit_step(const m2c_ctx *c, long long time_step) {
  return c->ls && c->cfg.reinit_freq > 0 && time_step > 0 && time_step % c->cfg.reinit_freq == 0;
}

int m2c_ctx_batch_end_head(m2c_ctx *c, long long time_step) {
  return c->ls ? end_of_step_head(c, reinit_step(c, time_step)) : M2C_OK;
}

int m2c_ctx_batch_end_retries(m2c_ctx *c, long long time_step) {
  return c->ls ? end_of_step_retries(c, reinit_step(c, time_step)) : M2C_OK;
}

int m2c_ctx_batch_end_tail(m2c_ctx *c) { return c->ls ? end_of_step_tail(c) : M2C_OK; }

int m2c_ctx_batch_backup(m2c_ctx *c, int restore) {
  if (!c) return M2C_ERR_ARG;
  const size_t nv = sizeof(double) * 5 * (size_t)c->N;
  if (!c->bk_V) {
    if (restore) return M2C_ERR_ARG;
    HIPCHK(hipMalloc(&c->bk_V, nv));
    HIPCHK(hipMalloc(&c->bk_ID, (size_t)c->N));
    if (c->ls) HIPCHK(hipMalloc(&c->bk_phi, sizeof(double) * (size_t)c->N));
    if (c->iface || c->ls) HIPCHK(hipMalloc(&c->bk_ms, sizeof(MpState)));
  }
  auto cp = [&](void *dst, const void *src, size_t n) {
    return hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToDevice, c->stream);
  };
  if (!restore) {
    HIPCHK(cp(c->bk_V, c->V, nv));
    HIPCHK(cp(c->bk_ID, c->ID, (size_t)c->N));
    if (c->ls) HIPCHK(cp(c->bk_phi, c->phiA, sizeof(double) * (size_t)c->N));
    if (c->bk_ms) HIPCHK(cp(c->bk_ms, c->L.ms, sizeof(MpState)));
  } else {
    HIPCHK(cp(c->V, c->bk_V, nv));
    HIPCHK(cp(c->ID, c->bk_ID, (size_t)c->N));
    if (c->ls) HIPCHK(cp(c->phiA, c->bk_phi, sizeof(double) * (size_t)c->N));
    if (c->bk_ms) HIPCHK(cp(c->L.ms, c->bk_ms, sizeof(MpState)));
  }
  return M2C_OK;
}

}  // extern "C"
