// m2c_hip.hip -- C-ABI context of the MI355X-native finite-volume update.
//
// Owns the device fields (structure of arrays, x fastest, no ghost frame), the HIP stream and
// the device-side time-loop clock; launches the fused stage kernels of fv_kernels.hpp.
// Entry points are declared in include/m2c_hip.h.
#include <hip/hip_runtime.h>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/m2c_hip.h"
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
  int variant = 1;             // 0 HLLC+SG, 1 HLLC generic EOS, 2 Roe, 3 LLF
  dim3 grid3;
};

static const unsigned long long kDblMaxBits = 0x7FEFFFFFFFFFFFFFull;

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
    double c2 = eos_c2<double, -1>(m, v[0], eos_e<double, -1>(m, v[0], v[4]));
    if (c2 < 0) { flags |= ST_NEG_C2; continue; }
    double cs = dsqrt(c2);
    double lf = dabs(v[1]) + cs, lg = dabs(v[2]) + cs, lh = dabs(v[3]) + cs;
    double loc = smin(g.dx / lf, smin(g.dy / lg, g.dz / lh));
    if (loc < mn) mn = loc;
  }
  warp_min_atomic(&clk->minbits, mn);
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

// ---------------------------------------------------------------- host helpers

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
  if (m2c_stage_upload_hllc_sg(c->grid, c->eos) || m2c_stage_upload_hllc_gen(c->grid, c->eos) ||
      m2c_stage_upload_roe_gen(c->grid, c->eos) || m2c_stage_upload_llf_gen(c->grid, c->eos))
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

static int launch_stage(m2c_ctx *c, int in_cons, int mode, const double *S, const double *Vn, const double *Us,
                        double *Out, double a, double b, double cc, int first, int ignore_active) {
  StageArgs s;
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
  default: rc = m2c_stage_launch_llf_gen(in_cons, mode, c->grid3, c->stream, s); break;
  }
  if (rc == 1) return M2C_ERR_ARG;
  if (rc) return M2C_ERR_HIP;
  HIPCHK(hipGetLastError());
  return M2C_OK;
}

// One time step: the Runge-Kutta stages (TimeIntegrator.cpp FE :150-226, RK2 :291-428,
// RK3 :496-690). `ev` (optional) receives a start/stop event pair per stage launch.
static int enqueue_step(m2c_ctx *c, int ignore_active, std::vector<hipEvent_t> *ev) {
  int ts = c->cfg.ts;
  int rc;
  auto rec = [&](int i) { if (ev) hipEventRecord((*ev)[i], c->stream); };
  if (ts == M2C_TS_FE) {
    rec(0);
    rc = launch_stage(c, 0, 2, c->V, c->V, nullptr, c->V, 0, 0, 1.0, 1, ignore_active);
    rec(1);
    return rc;
  }
  // stage 1: U1 = U_n + dt R(V_n)
  rec(0);
  rc = launch_stage(c, 0, 1, c->V, c->V, nullptr, c->U1, 0, 0, 1.0, 1, ignore_active);
  rec(1);
  if (rc) return rc;
  if (ts == M2C_TS_RK2) {
    // U(n+1) = 0.5 U1 + 0.5 U_n + 0.5 dt R(V1)
    rec(2);
    rc = launch_stage(c, 1, 2, c->U1, c->V, c->U1, c->V, 0.5, 0.5, 0.5, 0, ignore_active);
    rec(3);
    return rc;
  }
  // RK3: U2 = 0.25 U1 + 0.75 U_n + 0.25 dt R(V1); U(n+1) = 2/3 U2 + 1/3 U_n + 2/3 dt R(V2)
  rec(2);
  rc = launch_stage(c, 1, 1, c->U1, c->V, c->U1, c->U2, 0.25, 0.75, 0.25, 0, ignore_active);
  rec(3);
  if (rc) return rc;
  rec(4);
  rc = launch_stage(c, 1, 2, c->U2, c->V, c->U2, c->V, 2.0 / 3.0, 1.0 / 3.0, 2.0 / 3.0, 0, ignore_active);
  rec(5);
  return rc;
}

static int nstages(const m2c_ctx *c) { return c->cfg.ts == M2C_TS_FE ? 1 : (c->cfg.ts == M2C_TS_RK2 ? 2 : 3); }

static void launch_cfl(m2c_ctx *c) {
  int nb = (int)std::min<long long>((c->N + 255) / 256, 4096);
  hipLaunchKernelGGL(k_cfl_min, dim3(nb), dim3(256), 0, c->stream, c->V, c->ID, c->clk, c->st);
}

// ---------------------------------------------------------------- C ABI

extern "C" {

const char *m2c_version(void) { return "m2c_hip 0.1 (gfx950)"; }

int m2c_ctx_create(const m2c_config *cfg, m2c_ctx **out) {
  if (!cfg || !out) return M2C_ERR_ARG;
  if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1 || cfg->nmat < 1 || cfg->nmat > M2C_MAX_MAT) return M2C_ERR_ARG;
  if (cfg->flux == M2C_FLUX_GODUNOV) return M2C_ERR_UNSUPPORTED;
  if (cfg->precision != M2C_PREC_FP64) return M2C_ERR_UNSUPPORTED;
  for (int s = 0; s < 6; s++) {
    int b = cfg->bc[s];
    if (b < M2C_BC_INLET || b > M2C_BC_PERIODIC) return M2C_ERR_ARG;
    if (b == M2C_BC_OVERSET) return M2C_ERR_UNSUPPORTED;
    if ((b == M2C_BC_PERIODIC) != (cfg->bc[s ^ 1] == M2C_BC_PERIODIC)) return M2C_ERR_ARG;
  }
  HIPCHK(hipSetDevice(cfg->device));
  m2c_ctx *c = new m2c_ctx();
  c->cfg = *cfg;
  Grid &g = c->grid;
  memset(&g, 0, sizeof(g));
  g.nx = cfg->nx; g.ny = cfg->ny; g.nz = cfg->nz;
  g.N = (long long)cfg->nx * cfg->ny * cfg->nz;
  c->N = g.N;
  for (int s = 0; s < 6; s++) g.bc[s] = cfg->bc[s];
  for (int p = 0; p < 5; p++) { g.inlet[p] = cfg->inlet[p]; g.outlet[p] = cfg->outlet[p]; }
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
  g.inactive_id = cfg->inactive_id; g.nmat = cfg->nmat;
  g.mp_flux = cfg->mp_flux; g.mp_recon = cfg->mp_recon;
  int nactive = 0;
  for (int m = 0; m < cfg->nmat; m++) {
    build_eos(cfg->mat[m], c->eos[m]);
    if (m != cfg->inactive_id) nactive++;
  }
  g.has_interfaces = nactive > 1;
  bool all_sg = true;
  for (int m = 0; m < cfg->nmat; m++) all_sg = all_sg && cfg->mat[m].eos == M2C_EOS_SG;
  if (cfg->flux == M2C_FLUX_HLLC) c->variant = (all_sg && cfg->inactive_id < 0 && !g.has_interfaces) ? 0 : 1;
  else c->variant = cfg->flux == M2C_FLUX_ROE ? 2 : 3;
  if (g.has_interfaces && cfg->mp_flux != 2) {
    delete c;
    return M2C_ERR_UNSUPPORTED;   // exact two-material Riemann pass: see m2c_interface.hip
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
  if (upload_constants(c)) return M2C_ERR_HIP;
  hipLaunchKernelGGL(k_clock_init, dim3(1), dim3(1), 0, c->stream, c->clk, 0.0, 0.0, cfg->cfl, 0);

  // launch shape: 32x8 column tiles; z split into chunks so that the grid has >= ~2048 blocks
  int gx = (g.nx + TX - 1) / TX, gy = (g.ny + TY - 1) / TY;
  long long tiles = (long long)gx * gy;
  int nzc = (int)std::max<long long>(1, std::min<long long>(g.nz, (2048 + tiles - 1) / tiles));
  c->kchunk = (g.nz + nzc - 1) / nzc;
  nzc = (g.nz + c->kchunk - 1) / c->kchunk;
  c->grid3 = dim3(gx, gy, nzc);
  HIPCHK(hipStreamSynchronize(c->stream));
  *out = c;
  return M2C_OK;
}

int m2c_ctx_destroy(m2c_ctx *c) {
  if (!c) return M2C_OK;
  if (g_resident == c) g_resident = nullptr;
  hipStreamSynchronize(c->stream);
  hipFree(c->V); hipFree(c->U1); hipFree(c->U2); hipFree(c->R); hipFree(c->Fi);
  hipFree(c->ID); hipFree(c->clk); hipFree(c->st);
  hipStreamDestroy(c->stream);
  delete c;
  return M2C_OK;
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
  int rc = launch_stage(c, 0, 0, c->V, c->V, nullptr, c->U1, 0, 0, 0, 0, 1);
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
  int rc = enqueue_step(c, 1, nullptr);
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
  for (int s = 0; s < maxsteps; s++) {
    int rc = enqueue_step(c, 0, nullptr);
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
  for (int s = 0; s < nsteps; s++) {
    std::vector<hipEvent_t> sub(ev.begin() + 2 * ns * s, ev.begin() + 2 * ns * (s + 1));
    int rc = enqueue_step(c, 0, &sub);
    if (rc) return rc;
    hipLaunchKernelGGL(k_clock_finalize, dim3(1), dim3(1), 0, c->stream, c->clk, 1);
  }
  HIPCHK(hipEventRecord(b, c->stream));
  HIPCHK(hipEventSynchronize(b));
  float tot = 0.f;
  HIPCHK(hipEventElapsedTime(&tot, a, b));
  double ks = 0.0;
  for (int i = 0; i < ns * nsteps; i++) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]));
    ks += t;
  }
  for (auto &e : ev) hipEventDestroy(e);
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (ms_stage) *ms_stage = ks / (ns * nsteps);
  if (ms_step) *ms_step = tot / nsteps;
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
  out->error = (s.bits & (ST_BAD_STATE | ST_NAN_FLUX | ST_NEG_C2)) ? M2C_ERR_STATE
             : ((s.bits & ST_CLIP_INTERMEDIATE) ? M2C_ERR_UNSUPPORTED : M2C_OK);
  out->reverted_faces = (long long)s.reverted;
  out->clipped_cells = (long long)s.clipped;
  out->riemann_errors = (long long)s.riemann_err;
  out->steps = k.steps;
  out->t = k.t;
  out->dt = k.dt;
  out->cfl = c->cfg.cfl;
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
  if (c->cfg.flux == M2C_FLUX_HLLC) launch_flux_batch<FLUX_HLLC>(dir, gb, c->stream, n, d, d + 5 * (size_t)n, id, d + 10 * (size_t)n);
  else if (c->cfg.flux == M2C_FLUX_ROE) launch_flux_batch<FLUX_ROE>(dir, gb, c->stream, n, d, d + 5 * (size_t)n, id, d + 10 * (size_t)n);
  else launch_flux_batch<FLUX_LLF>(dir, gb, c->stream, n, d, d + 5 * (size_t)n, id, d + 10 * (size_t)n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(F, d + 10 * (size_t)n, b, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipFree(d));
  return M2C_OK;
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

}  // extern "C"
