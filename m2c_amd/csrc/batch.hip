// batch.hip -- many independent instances advanced by one persistent launch per RK stage.
//
// Each instance is a single-material finite-volume problem with its own grid, boundary
// conditions, state and CFL-adaptive time step (its own device clock). A stage launch drains a
// queue of (instance, tile, z-chunk) work items: one queue per XCD (a workgroup's XCD is its
// block index modulo 8, the hardware's round-robin), holding a contiguous range of work items so
// that neighbouring tiles of an instance share an XCD's L2; a block whose queue is empty steals
// from the next XCD's. A block loads the grid of the instance it works on into LDS and runs the
// same fused stage body as the single-instance kernel (fv_kernels.hpp), so a batched instance
// evolves bit for bit as it would alone.
//
// A whole time step (queue resets, the stage launches, the per-instance dt update) is captured
// once into a hipGraph and replayed; the autotuner times candidate (z-chunk depth, resident blocks)
// pairs on a copy of the state and keeps the fastest.
#define M2C_BATCHED
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/m2c_hip.h"
#include "fv_kernels.hpp"

namespace m2c {

constexpr int NXCD = 8;

struct BatchInst {
  Grid grid;
  StageArgs args[3];           // per stage of the time scheme
  int gx, gy, nzc;             // tiles in x, y; z chunks
  unsigned tile0;              // first work item of this instance
  long long cell0;             // offset of the instance's cells in the concatenated fields
};

struct TileQueue {
  unsigned head[NXCD];
  unsigned start[NXCD];
  unsigned end[NXCD];
};

__device__ __forceinline__ void load_grid(const Grid &src) {
  const int n = (int)(sizeof(Grid) / sizeof(int));
  const int *s = reinterpret_cast<const int *>(&src);
  int *d = reinterpret_cast<int *>(&s_bgrid);
  for (int t = threadIdx.x; t < n; t += blockDim.x) d[t] = s[t];
}

__device__ __forceinline__ int find_instance(const BatchInst *in, int n, unsigned tile) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (in[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int IN_CONS, int MODE, int FLUX, int EK, typename TC>
__global__ __launch_bounds__(NTHR, 2) void k_batch_stage(const BatchInst *insts, int ninst, TileQueue *q,
                                                          int stage) {
  __shared__ int s_tile, s_inst;
  const int myq = blockIdx.x % NXCD;
  for (;;) {
    if (threadIdx.x == 0) {
      int t = -1;
      for (int r = 0; r < NXCD && t < 0; r++) {
        const int qq = (myq + r) % NXCD;
        if (q->head[qq] >= q->end[qq]) continue;
        unsigned v = atomicAdd(&q->head[qq], 1u);
        if (v < q->end[qq]) t = (int)v;
      }
      s_tile = t;
      s_inst = t >= 0 ? find_instance(insts, ninst, (unsigned)t) : 0;
    }
    __syncthreads();
    const int t = s_tile;
    if (t < 0) break;
    const BatchInst &B = insts[s_inst];
    load_grid(B.grid);
    __syncthreads();
    const unsigned local = (unsigned)t - B.tile0;
    const int per = B.gx * B.gy;
    const int bz = (int)(local / per);
    const int rem = (int)(local % per);
    const StageArgs a = B.args[stage];
    stage_body<IN_CONS, MODE, FLUX, EK, TC>(a, rem % B.gx, rem / B.gx, bz);
    __syncthreads();
  }
}

__global__ void k_queue_reset(TileQueue *q) {
  if (threadIdx.x < NXCD) q->head[threadIdx.x] = q->start[threadIdx.x];
}

__global__ void k_batch_clock_set(Clock *clk, int n, int maxsteps_more) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  clk[i].tmax = INFINITY;
  clk[i].maxsteps = clk[i].steps + maxsteps_more;
  clk[i].minbits = 0x7FEFFFFFFFFFFFFFull;
  clk[i].active = 1;
}

__global__ void k_batch_clock_init(Clock *clk, int n, double cfl) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  clk[i].t = 0.0; clk[i].dt = 0.0; clk[i].tmax = 0.0; clk[i].cfl = cfl;
  clk[i].minbits = 0x7FEFFFFFFFFFFFFFull; clk[i].steps = 0; clk[i].maxsteps = 0; clk[i].active = 1;
}

// the dt update of every instance (the single-instance k_clock_finalize, Main.cpp:535-567)
__global__ void k_batch_clock_finalize(Clock *clk, int n, int advance) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Clock &c = clk[i];
  if (advance) {
    if (!c.active) return;
    c.t += c.dt;
    c.steps += 1;
  }
  double mn = __longlong_as_double((long long)c.minbits);
  double dt = c.cfl * mn;
  if (c.t + dt > c.tmax) dt = c.tmax - c.t;
  c.dt = dt;
  c.active = (c.t < c.tmax && c.steps < c.maxsteps) ? 1 : 0;
  c.minbits = 0x7FEFFFFFFFFFFFFFull;
}

// CFL minimum of every instance; blockIdx.y = instance (SpaceOperator.cpp:2128-2198)
__global__ __launch_bounds__(256) void k_batch_cfl(const BatchInst *insts, const double *V, const int8_t *ID,
                                                    Clock *clk, Status *st) {
  const BatchInst &B = insts[blockIdx.y];
  load_grid(B.grid);
  __syncthreads();
  const Grid &g = s_bgrid;
  const double *Vi = V + 5 * B.cell0;
  const int8_t *IDi = ID + B.cell0;
  double mn = DBL_MAX;
  unsigned flags = 0;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < g.N;
       idx += (long long)gridDim.x * blockDim.x) {
    int id = IDi[idx];
    if (id == g.inactive_id) continue;
    const DevEos &m = c_eos[id];
    double v[5];
#pragma unroll
    for (int p = 0; p < 5; p++) v[p] = Vi[p * g.N + idx];
    double c2 = eos_c2<double, -1>(m, v[0], eos_e<double, -1>(m, v[0], v[4]));
    if (c2 < 0) { flags |= ST_NEG_C2; continue; }
    double cs = dsqrt(c2);
    double lf = dabs(v[1]) + cs, lg = dabs(v[2]) + cs, lh = dabs(v[3]) + cs;
    double loc = smin(g.dx / lf, smin(g.dy / lg, g.dz / lh));
    if (loc < mn) mn = loc;
  }
  warp_min_atomic(&clk[blockIdx.y].minbits, mn);
  if (flags) atomicOr(&st[blockIdx.y].bits, flags);
}

__global__ void k_batch_copy(const double *src, double *dst, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void k_batch_aos_to_soa(const double *in, double *out, long long N) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x)
    for (int p = 0; p < 5; p++) out[p * N + i] = in[5 * i + p];
}
__global__ void k_batch_soa_to_aos(const double *in, double *out, long long N) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x)
    for (int p = 0; p < 5; p++) out[5 * i + p] = in[p * N + i];
}

}  // namespace m2c

using namespace m2c;

// ---------------------------------------------------------------- host side

// grid and material tables of a configuration: defined with the single-instance context
extern "C" int m2c_fill_grid(const m2c_config *cfg, Grid *g, DevEos *eos, int *variant);

struct m2c_batch {
  std::vector<m2c_config> cfg;
  std::vector<BatchInst> hinst;
  int n = 0, ts = M2C_TS_RK2, variant = 0, nstage = 2, flux = M2C_FLUX_HLLC;
  DevEos eos[8];
  hipStream_t stream = nullptr;
  long long cells = 0;
  double *V = nullptr, *U1 = nullptr, *U2 = nullptr, *stage_buf = nullptr, *backup = nullptr;
  int8_t *ID = nullptr;
  Clock *clk = nullptr;
  Status *st = nullptr;
  BatchInst *dinst = nullptr;
  TileQueue *q = nullptr;
  unsigned ntiles = 0;
  int kchunk = 8, nblocks = 512;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

#define BHIP(x)                                                                   \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) {                                                       \
      fprintf(stderr, "m2c_batch: %s failed: %s\n", #x, hipGetErrorString(e_));   \
      return M2C_ERR_HIP;                                                         \
    }                                                                             \
  } while (0)

static int batch_upload_tables(m2c_batch *b) {
  BHIP(hipMemcpyToSymbol(HIP_SYMBOL(c_eos), b->eos, sizeof(b->eos)));
  return M2C_OK;
}

// work items and queues for z chunks of depth kchunk
static int batch_layout(m2c_batch *b, int kchunk) {
  unsigned t = 0;
  for (int i = 0; i < b->n; i++) {
    BatchInst &B = b->hinst[i];
    const Grid &g = B.grid;
    B.gx = (g.nx + TX - 1) / TX;
    B.gy = (g.ny + TY - 1) / TY;
    B.nzc = (g.nz + kchunk - 1) / kchunk;
    B.tile0 = t;
    for (int s = 0; s < 3; s++) B.args[s].kchunk = kchunk;
    t += (unsigned)(B.gx * B.gy * B.nzc);
  }
  b->ntiles = t;
  b->kchunk = kchunk;
  TileQueue hq;
  for (int x = 0; x < NXCD; x++) {
    hq.start[x] = (unsigned)((unsigned long long)t * x / NXCD);
    hq.end[x] = (unsigned)((unsigned long long)t * (x + 1) / NXCD);
    hq.head[x] = hq.start[x];
  }
  BHIP(hipMemcpy(b->dinst, b->hinst.data(), sizeof(BatchInst) * b->n, hipMemcpyHostToDevice));
  BHIP(hipMemcpy(b->q, &hq, sizeof(hq), hipMemcpyHostToDevice));
  if (b->exec) { hipGraphExecDestroy(b->exec); b->exec = nullptr; }
  if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
  return M2C_OK;
}

template <int FLUX, int EK, typename TC>
static void launch_stage_t(m2c_batch *b, int in_cons, int mode, int stage) {
  dim3 gr(b->nblocks), bl(NTHR);
  if (in_cons == 0 && mode == 1)
    hipLaunchKernelGGL((k_batch_stage<0, 1, FLUX, EK, TC>), gr, bl, 0, b->stream, b->dinst, b->n, b->q, stage);
  else if (in_cons == 1 && mode == 1)
    hipLaunchKernelGGL((k_batch_stage<1, 1, FLUX, EK, TC>), gr, bl, 0, b->stream, b->dinst, b->n, b->q, stage);
  else
    hipLaunchKernelGGL((k_batch_stage<1, 2, FLUX, EK, TC>), gr, bl, 0, b->stream, b->dinst, b->n, b->q, stage);
}

static void launch_stage(m2c_batch *b, int in_cons, int mode, int stage) {
  // the single-instance context's variants: Roe and LLF run the generic-EOS kernels
  switch (b->variant) {
  case 0: launch_stage_t<FLUX_HLLC, 0, double>(b, in_cons, mode, stage); break;
  case 1: launch_stage_t<FLUX_HLLC, -1, double>(b, in_cons, mode, stage); break;
  case 2: launch_stage_t<FLUX_ROE, -1, double>(b, in_cons, mode, stage); break;
  case 3: launch_stage_t<FLUX_LLF, -1, double>(b, in_cons, mode, stage); break;
  case 4: launch_stage_t<FLUX_HLLC, 0, float>(b, in_cons, mode, stage); break;
  default: launch_stage_t<FLUX_HLLC, -1, float>(b, in_cons, mode, stage); break;
  }
}

// one time step of every instance: RK2 (TimeIntegrator.cpp:291-428) or RK3 (:496-690)
static void enqueue_batch_step(m2c_batch *b) {
  for (int s = 0; s < b->nstage; s++) {
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, b->stream, b->q);
    const int in_cons = s == 0 ? 0 : 1;
    const int mode = s == b->nstage - 1 ? 2 : 1;
    launch_stage(b, in_cons, mode, s);
  }
  hipLaunchKernelGGL(k_batch_clock_finalize, dim3((b->n + 63) / 64), dim3(64), 0, b->stream, b->clk, b->n, 1);
}

static void enqueue_batch_cfl(m2c_batch *b, int maxsteps_more) {
  hipLaunchKernelGGL(k_batch_clock_set, dim3((b->n + 63) / 64), dim3(64), 0, b->stream, b->clk, b->n, maxsteps_more);
  hipLaunchKernelGGL(k_batch_cfl, dim3(64, b->n), dim3(256), 0, b->stream, b->dinst, b->V, b->ID, b->clk, b->st);
  hipLaunchKernelGGL(k_batch_clock_finalize, dim3((b->n + 63) / 64), dim3(64), 0, b->stream, b->clk, b->n, 0);
}

extern "C" {

int m2c_batch_create(const m2c_config *cfgs, int n, m2c_batch **out) {
  if (!cfgs || n < 1 || !out) return M2C_ERR_ARG;
  m2c_batch *b = new m2c_batch();
  b->n = n;
  b->cfg.assign(cfgs, cfgs + n);
  b->ts = cfgs[0].ts;
  b->flux = cfgs[0].flux;
  if (b->ts != M2C_TS_RK2 && b->ts != M2C_TS_RK3) { delete b; return M2C_ERR_UNSUPPORTED; }
  b->nstage = b->ts == M2C_TS_RK2 ? 2 : 3;
  b->hinst.resize(n);
  for (int i = 0; i < n; i++) {
    const m2c_config &c = cfgs[i];
    // one kernel variant and one material table for the whole batch
    if (c.ts != b->ts || c.flux != b->flux || c.precision != cfgs[0].precision || c.nmat != cfgs[0].nmat ||
        c.device != cfgs[0].device || memcmp(c.mat, cfgs[0].mat, sizeof(c.mat)) != 0 || c.ls_matid >= 0) {
      delete b;
      return M2C_ERR_UNSUPPORTED;
    }
    for (int m = 0; m < c.nmat; m++)      // inviscid, adiabatic instances only
      if (c.mat[m].viscosity != 0 || c.mat[m].kappa != 0.0) {
        delete b;
        return M2C_ERR_UNSUPPORTED;
      }
    int variant = 0;
    DevEos eos[8];
    int rc = m2c_fill_grid(&c, &b->hinst[i].grid, eos, &variant);
    if (rc) { delete b; return rc; }
    if (b->hinst[i].grid.has_interfaces) { delete b; return M2C_ERR_UNSUPPORTED; }
    if (i == 0) { b->variant = variant; memcpy(b->eos, eos, sizeof(eos)); }
    else if (variant != b->variant) { delete b; return M2C_ERR_UNSUPPORTED; }
    b->hinst[i].cell0 = b->cells;
    b->cells += b->hinst[i].grid.N;
  }
  // HLLC (0, 1), Roe (2) and LLF (3) in FP64, HLLC in the mixed form (4, 5)
  if (b->variant < 0 || b->variant > 5) { delete b; return M2C_ERR_UNSUPPORTED; }
  // from here on a failure releases everything allocated so far
#define BHIP_D(x)                                                                                  \
  do {                                                                                             \
    if ((x) != hipSuccess) { m2c_batch_destroy(b); return M2C_ERR_HIP; }                            \
  } while (0)
  BHIP_D(hipSetDevice(cfgs[0].device));
  BHIP_D(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  size_t bytes = sizeof(double) * 5 * (size_t)b->cells;
  BHIP_D(hipMalloc(&b->V, bytes));
  BHIP_D(hipMalloc(&b->U1, bytes));
  if (b->nstage == 3) BHIP_D(hipMalloc(&b->U2, bytes));
  BHIP_D(hipMalloc(&b->stage_buf, bytes));
  BHIP_D(hipMalloc(&b->ID, (size_t)b->cells));
  BHIP_D(hipMemset(b->ID, 0, (size_t)b->cells));
  BHIP_D(hipMemset(b->V, 0, bytes));
  BHIP_D(hipMalloc(&b->clk, sizeof(Clock) * n));
  BHIP_D(hipMalloc(&b->st, sizeof(Status) * n));
  BHIP_D(hipMemset(b->st, 0, sizeof(Status) * n));
  BHIP_D(hipMalloc(&b->dinst, sizeof(BatchInst) * n));
  BHIP_D(hipMalloc(&b->q, sizeof(TileQueue)));
  // stage arguments of every instance (the same buffers as the single-instance context)
  for (int i = 0; i < n; i++) {
    BatchInst &B = b->hinst[i];
    const long long o = 5 * B.cell0;
    double *V = b->V + o, *U1 = b->U1 + o, *U2 = b->U2 ? b->U2 + o : nullptr;
    for (int s = 0; s < 3; s++) {
      StageArgs &a = B.args[s];
      memset(&a, 0, sizeof(a));
      a.ID = b->ID + B.cell0;
      a.clk = b->clk + i;
      a.st = b->st + i;
      a.ignore_active = 0;
      a.no_cfl = 0;
    }
    // stage 1: U1 = U_n + dt R(V_n)
    B.args[0].S = V; B.args[0].Vn = V; B.args[0].Out = U1; B.args[0].c = 1.0; B.args[0].first = 1;
    if (b->nstage == 2) {
      StageArgs &a = B.args[1];   // U(n+1) = 0.5 U1 + 0.5 U_n + 0.5 dt R(V1)
      a.S = U1; a.Vn = V; a.Us = U1; a.Out = V; a.a = 0.5; a.b = 0.5; a.c = 0.5;
    } else {
      StageArgs &a1 = B.args[1];  // U2 = 0.25 U1 + 0.75 U_n + 0.25 dt R(V1)
      a1.S = U1; a1.Vn = V; a1.Us = U1; a1.Out = U2; a1.a = 0.25; a1.b = 0.75; a1.c = 0.25;
      StageArgs &a2 = B.args[2];  // U(n+1) = 2/3 U2 + 1/3 U_n + 2/3 dt R(V2)
      a2.S = U2; a2.Vn = V; a2.Us = U2; a2.Out = V; a2.a = 2.0 / 3.0; a2.b = 1.0 / 3.0; a2.c = 2.0 / 3.0;
    }
  }
  if (int rc = batch_upload_tables(b)) { m2c_batch_destroy(b); return rc; }
  if (int rc = batch_layout(b, 8)) { m2c_batch_destroy(b); return rc; }
  b->nblocks = 512;
  hipLaunchKernelGGL(k_batch_clock_init, dim3((n + 63) / 64), dim3(64), 0, b->stream, b->clk, n, cfgs[0].cfl);
  for (int i = 0; i < n; i++)
    if (cfgs[i].cfl != cfgs[0].cfl) {   // per-instance CFL numbers
      Clock h;
      memset(&h, 0, sizeof(h));
      h.cfl = cfgs[i].cfl;
      h.minbits = 0x7FEFFFFFFFFFFFFFull;
      h.active = 1;
      BHIP_D(hipMemcpyAsync(b->clk + i, &h, sizeof(h), hipMemcpyHostToDevice, b->stream));
    }
  BHIP_D(hipStreamSynchronize(b->stream));
#undef BHIP_D
  *out = b;
  return M2C_OK;
}

int m2c_batch_destroy(m2c_batch *b) {
  if (!b) return M2C_OK;
  if (b->stream) hipStreamSynchronize(b->stream);
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  hipFree(b->V); hipFree(b->U1); hipFree(b->U2); hipFree(b->stage_buf); hipFree(b->backup);
  hipFree(b->ID); hipFree(b->clk); hipFree(b->st); hipFree(b->dinst); hipFree(b->q);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return M2C_OK;
}

int m2c_batch_upload(m2c_batch *b, int i, const double *V, const int32_t *id) {
  if (!b || i < 0 || i >= b->n || !V) return M2C_ERR_ARG;
  BHIP(hipSetDevice(b->cfg[0].device));
  const BatchInst &B = b->hinst[i];
  const long long N = B.grid.N;
  // the IDs are checked before anything is copied: M2C_ERR_ARG leaves the instance as it was
  std::vector<int8_t> id8(id ? (size_t)N : 0);
  for (long long c = 0; id && c < N; c++) {
    if (id[c] < 0 || id[c] >= b->cfg[i].nmat) return M2C_ERR_ARG;
    id8[c] = (int8_t)id[c];
  }
  BHIP(hipMemcpyAsync(b->stage_buf, V, sizeof(double) * 5 * N, hipMemcpyHostToDevice, b->stream));
  hipLaunchKernelGGL(k_batch_aos_to_soa, dim3(1024), dim3(256), 0, b->stream, b->stage_buf, b->V + 5 * B.cell0, N);
  BHIP(hipGetLastError());
  if (id) BHIP(hipMemcpyAsync(b->ID + B.cell0, id8.data(), (size_t)N, hipMemcpyHostToDevice, b->stream));
  BHIP(hipStreamSynchronize(b->stream));
  return M2C_OK;
}

int m2c_batch_download(m2c_batch *b, int i, double *V, int32_t *id) {
  if (!b || i < 0 || i >= b->n) return M2C_ERR_ARG;
  BHIP(hipSetDevice(b->cfg[0].device));
  const BatchInst &B = b->hinst[i];
  const long long N = B.grid.N;
  if (V) {
    hipLaunchKernelGGL(k_batch_soa_to_aos, dim3(1024), dim3(256), 0, b->stream, b->V + 5 * B.cell0, b->stage_buf, N);
    BHIP(hipMemcpyAsync(V, b->stage_buf, sizeof(double) * 5 * N, hipMemcpyDeviceToHost, b->stream));
  }
  if (id) {
    std::vector<int8_t> id8((size_t)N);
    BHIP(hipMemcpyAsync(id8.data(), b->ID + B.cell0, (size_t)N, hipMemcpyDeviceToHost, b->stream));
    BHIP(hipStreamSynchronize(b->stream));
    for (long long c = 0; c < N; c++) id[c] = id8[c];
  }
  BHIP(hipStreamSynchronize(b->stream));
  return M2C_OK;
}

int m2c_batch_run(m2c_batch *b, int nsteps, int use_graph, double *ms) {
  if (!b || nsteps < 0) return M2C_ERR_ARG;
  BHIP(hipSetDevice(b->cfg[0].device));
  if (int rc = batch_upload_tables(b)) return rc;
  enqueue_batch_cfl(b, nsteps);
  hipEvent_t e0, e1;
  BHIP(hipEventCreate(&e0));
  BHIP(hipEventCreate(&e1));
  BHIP(hipEventRecord(e0, b->stream));
  if (use_graph) {
    if (!b->exec) {
      BHIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      enqueue_batch_step(b);
      BHIP(hipStreamEndCapture(b->stream, &b->graph));
      BHIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    for (int s = 0; s < nsteps; s++) BHIP(hipGraphLaunch(b->exec, b->stream));
  } else {
    for (int s = 0; s < nsteps; s++) enqueue_batch_step(b);
  }
  BHIP(hipGetLastError());
  BHIP(hipEventRecord(e1, b->stream));
  BHIP(hipEventSynchronize(e1));
  float t = 0.f;
  BHIP(hipEventElapsedTime(&t, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (ms) *ms = t;
  return M2C_OK;
}

int m2c_batch_status(m2c_batch *b, int i, m2c_status *out) {
  if (!b || i < 0 || i >= b->n || !out) return M2C_ERR_ARG;
  Status s;
  Clock k;
  BHIP(hipMemcpyAsync(&s, b->st + i, sizeof(Status), hipMemcpyDeviceToHost, b->stream));
  BHIP(hipMemcpyAsync(&k, b->clk + i, sizeof(Clock), hipMemcpyDeviceToHost, b->stream));
  BHIP(hipStreamSynchronize(b->stream));
  memset(out, 0, sizeof(*out));
  out->error = (s.bits & (ST_BAD_STATE | ST_NAN_FLUX | ST_NEG_C2)) ? M2C_ERR_STATE
             : ((s.bits & ST_CLIP_INTERMEDIATE) ? M2C_ERR_UNSUPPORTED : M2C_OK);
  out->reverted_faces = (long long)s.reverted;
  out->clipped_cells = (long long)s.clipped;
  out->steps = k.steps;
  out->t = k.t;
  out->dt = k.dt;
  out->cfl = k.cfl;
  return M2C_OK;
}

int m2c_batch_autotune(m2c_batch *b, int trial_steps, m2c_batch_tuning *out) {
  if (!b || trial_steps < 1) return M2C_ERR_ARG;
  BHIP(hipSetDevice(b->cfg[0].device));
  const size_t bytes = sizeof(double) * 5 * (size_t)b->cells;
  if (!b->backup) BHIP(hipMalloc(&b->backup, bytes));
  // tune on a copy of the state; the clocks and counters are restored afterwards
  std::vector<Clock> hclk(b->n);
  std::vector<Status> hst(b->n);
  BHIP(hipMemcpy(hclk.data(), b->clk, sizeof(Clock) * b->n, hipMemcpyDeviceToHost));
  BHIP(hipMemcpy(hst.data(), b->st, sizeof(Status) * b->n, hipMemcpyDeviceToHost));
  hipLaunchKernelGGL(k_batch_copy, dim3(4096), dim3(256), 0, b->stream, b->V, b->backup, 5 * b->cells);
  const int kchunks[] = {4, 8, 16, 32};
  const int blocks[] = {256, 512, 1024};
  double best = 1e300;
  int bk = b->kchunk, bb = b->nblocks, ncand = 0;
  for (int kc : kchunks)
    for (int nb : blocks) {
      if (int rc = batch_layout(b, kc)) return rc;
      b->nblocks = nb;
      hipLaunchKernelGGL(k_batch_copy, dim3(4096), dim3(256), 0, b->stream, b->backup, b->V, 5 * b->cells);
      BHIP(hipMemcpy(b->clk, hclk.data(), sizeof(Clock) * b->n, hipMemcpyHostToDevice));
      double ms = 0.0;
      if (int rc = m2c_batch_run(b, trial_steps, 1, &ms)) return rc;
      ncand++;
      if (ms < best) { best = ms; bk = kc; bb = nb; }
    }
  if (int rc = batch_layout(b, bk)) return rc;
  b->nblocks = bb;
  hipLaunchKernelGGL(k_batch_copy, dim3(4096), dim3(256), 0, b->stream, b->backup, b->V, 5 * b->cells);
  BHIP(hipMemcpy(b->clk, hclk.data(), sizeof(Clock) * b->n, hipMemcpyHostToDevice));
  BHIP(hipMemcpy(b->st, hst.data(), sizeof(Status) * b->n, hipMemcpyHostToDevice));
  BHIP(hipStreamSynchronize(b->stream));
  if (out) {
    out->kchunk = bk;
    out->blocks = bb;
    out->ms_per_step = best / trial_steps;
    out->candidates = ncand;
  }
  return M2C_OK;
}

int m2c_batch_set_tuning(m2c_batch *b, int kchunk, int blocks) {
  if (!b || kchunk < 1 || blocks < 1) return M2C_ERR_ARG;
  if (int rc = batch_layout(b, kchunk)) return rc;
  b->nblocks = blocks;
  return M2C_OK;
}

}  // extern "C"
