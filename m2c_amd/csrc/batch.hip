// batch.hip -- many independent instances advanced by one persistent launch per RK stage.
//
// Each instance is a single- or two-material finite-volume problem with its own grid, boundary
// conditions, state and CFL-adaptive time step (its own device clock). A two-material instance
// (level set, exact Riemann solves at its material interface) keeps a single-instance context for
// its interface-face lists, level set and end-of-step buffers (batch_ctx.hpp): its fluid stages run
// in the same persistent launches as every other instance, and the batch runs its interface pass,
// level-set stages and end-of-step tasks around them. A stage launch drains a
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
#include "batch_ctx.hpp"
#include "batch_kern.hpp"

namespace m2c {

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

// CFL minimum of every instance (list null) or of the listed ones; blockIdx.y = instance or list
// entry (SpaceOperator.cpp:2128-2198)
__global__ __launch_bounds__(256) void k_batch_cfl(const BatchInst *insts, const int *list, Clock *clk, Status *st) {
  const int inst = list ? list[blockIdx.y] : (int)blockIdx.y;
  const BatchInst &B = insts[inst];
  load_grid(B.grid);
  __syncthreads();
  const Grid &g = s_bgrid;
  const double *Vi = B.Vc;
  const int8_t *IDi = B.IDc;
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
  warp_min_atomic(&clk[inst].minbits, mn);
  if (flags) atomicOr(&st[inst].bits, flags);
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
  long long cells = 0;          // cells of the single-material instances (concatenated fields)
  double *V = nullptr, *U1 = nullptr, *U2 = nullptr, *stage_buf = nullptr, *backup = nullptr;
  int8_t *ID = nullptr;
  Clock *clk = nullptr;
  Status *st = nullptr;
  BatchInst *dinst = nullptr;
  TileQueue *q = nullptr;
  // with two-material instances: the single-material instances' stage launch apart, with the
  // variant without the material-interface code (variant_s >= 0), each group its own instance
  // array (tiles numbered within the group) and queue
  int variant_s = -1, n_s = 0, n_m = 0;
  BatchInst *dinst_s = nullptr, *dinst_m = nullptr;
  TileQueue *q_s = nullptr, *q_m = nullptr;
  unsigned ntiles = 0;
  int kchunk = 8, nblocks = 512, shape = 0;   // shape: 0 = 32 x 8 column tiles, 1 = 16 x 16
  // two-material instances: their contexts (null for a single-material instance) and their indices
  std::vector<m2c_ctx *> mpc;
  std::vector<int> mp_list;
  int *d_mp_list = nullptr;
  // their exact interface solves in one launch per stage (null: one launch per instance)
  IfaceInst *d_iface = nullptr;
  int n_iface = 0, iface_blocks = 0, list_blocks = 0, iface_flux = -1;
  long long steps = 0;          // time steps taken (the reinitialisation schedule)
  // captured steps: every step (single-material batches) or, with two-material instances, a step
  // without reinitialisation, and the parts before and after the host's reinitialisation retries
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraphExec_t exec_rhead = nullptr, exec_rtail = nullptr;
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
  if (m2c_bstage_upload_hllc_sg(b->eos) || m2c_bstage_upload_hllc_gen(b->eos) || m2c_bstage_upload_roe_gen(b->eos) ||
      m2c_bstage_upload_llf_gen(b->eos) || m2c_bstage_upload_hllc_sg_f32(b->eos) ||
      m2c_bstage_upload_hllc_gen_f32(b->eos) || m2c_bstage_upload_hllc_sg_mp(b->eos) ||
      m2c_bstage_upload_hllc_sg_mp_f32(b->eos))
    return M2C_ERR_HIP;
  return M2C_OK;
}

static void drop_graphs(m2c_batch *b) {
  if (b->exec) { hipGraphExecDestroy(b->exec); b->exec = nullptr; }
  if (b->exec_rhead) { hipGraphExecDestroy(b->exec_rhead); b->exec_rhead = nullptr; }
  if (b->exec_rtail) { hipGraphExecDestroy(b->exec_rtail); b->exec_rtail = nullptr; }
  if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
}

// work items and queues for tiles of the given shape and z chunks of depth kchunk
static int batch_layout(m2c_batch *b, int kchunk, int shape) {
  int tx, ty;
  batch_tile_dims(shape, &tx, &ty);
  b->shape = shape;
  unsigned t = 0;
  for (int i = 0; i < b->n; i++) {
    BatchInst &B = b->hinst[i];
    const Grid &g = B.grid;
    B.gx = (g.nx + tx - 1) / tx;
    B.gy = (g.ny + ty - 1) / ty;
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
  if (b->variant_s >= 0) {   // the two groups' arrays and queues
    for (int grp = 0; grp < 2; grp++) {
      std::vector<BatchInst> h;
      unsigned tg = 0;
      for (int i = 0; i < b->n; i++) {
        if ((b->mpc[i] != nullptr) != (grp == 1)) continue;
        BatchInst B = b->hinst[i];
        B.tile0 = tg;
        tg += (unsigned)(B.gx * B.gy * B.nzc);
        h.push_back(B);
      }
      TileQueue gq;
      for (int x = 0; x < NXCD; x++) {
        gq.start[x] = (unsigned)((unsigned long long)tg * x / NXCD);
        gq.end[x] = (unsigned)((unsigned long long)tg * (x + 1) / NXCD);
        gq.head[x] = gq.start[x];
      }
      BHIP(hipMemcpy(grp ? b->dinst_m : b->dinst_s, h.data(), sizeof(BatchInst) * h.size(), hipMemcpyHostToDevice));
      BHIP(hipMemcpy(grp ? b->q_m : b->q_s, &gq, sizeof(gq), hipMemcpyHostToDevice));
    }
  }
  drop_graphs(b);
  return M2C_OK;
}

static int launch_variant(m2c_batch *b, int variant, const BatchInst *dinst, int n, TileQueue *q, int in_cons,
                          int mode, int stage) {
  // the single-instance context's variants: Roe and LLF run the generic-EOS kernels; 9 and 10 are
  // the stiffened-gas variants with the material-interface code (two-material instances)
  const int sh = b->shape;
  int rc;
  hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, b->stream, q);
  switch (variant) {
#define M2C_BL(tag) m2c_bstage_launch_##tag(sh, in_cons, mode, b->nblocks, b->stream, dinst, n, q, stage)
  case 0: rc = M2C_BL(hllc_sg); break;
  case 1: rc = M2C_BL(hllc_gen); break;
  case 2: rc = M2C_BL(roe_gen); break;
  case 3: rc = M2C_BL(llf_gen); break;
  case 4: rc = M2C_BL(hllc_sg_f32); break;
  case 9: rc = M2C_BL(hllc_sg_mp); break;
  case 10: rc = M2C_BL(hllc_sg_mp_f32); break;
  default: rc = M2C_BL(hllc_gen_f32); break;
#undef M2C_BL
  }
  return rc ? M2C_ERR_HIP : M2C_OK;
}

static int launch_stage(m2c_batch *b, int in_cons, int mode, int stage) {
  if (b->variant_s < 0) return launch_variant(b, b->variant, b->dinst, b->n, b->q, in_cons, mode, stage);
  if (int rc = launch_variant(b, b->variant_s, b->dinst_s, b->n_s, b->q_s, in_cons, mode, stage)) return rc;
  return launch_variant(b, b->variant, b->dinst_m, b->n_m, b->q_m, in_cons, mode, stage);
}

static bool reinit_step(const m2c_batch *b, long long time_step) {
  for (int i : b->mp_list) {
    const m2c_config &c = b->cfg[i];
    if (c.ls_matid >= 0 && c.reinit_freq > 0 && time_step > 0 && time_step % c.reinit_freq == 0) return true;
  }
  return false;
}

// the fluid stages of every instance (RK2, TimeIntegrator.cpp:291-428; RK3, :496-690), each
// two-material instance's interface pass before and level-set stage after each of them, and the
// head of its end-of-step tasks (IDs of swept nodes, the first reinitialisation attempt)
static int enqueue_batch_head(m2c_batch *b, long long time_step) {
  for (int s = 0; s < b->nstage; s++) {
    if (b->d_iface) {   // every instance's face list, then every instance's interface solves, one launch each
      for (int i : b->mp_list)
        if (s == 0 && m2c_ctx_batch_step_begin(b->mpc[i])) return M2C_ERR_HIP;
      if (m2c_mp_launch_iface_list_multi(b->stream, b->d_iface, b->n_iface, b->list_blocks, s == 0 ? 1 : 0))
        return M2C_ERR_HIP;
      if (m2c_mp_launch_iface_multi(b->stream, b->d_iface, b->n_iface, b->iface_blocks, b->iface_flux, s == 0 ? 0 : 1, s,
                                    s == 0 ? 1 : 0))
        return M2C_ERR_HIP;
    } else {
      for (int i : b->mp_list)
        if (m2c_ctx_batch_bind(b->mpc[i]) || m2c_ctx_batch_before(b->mpc[i], s)) return M2C_ERR_HIP;
    }
    const int in_cons = s == 0 ? 0 : 1;
    const int mode = s == b->nstage - 1 ? 2 : 1;
    if (int rc = launch_stage(b, in_cons, mode, s)) return rc;
    for (int i : b->mp_list)
      if (m2c_ctx_batch_bind(b->mpc[i]) || m2c_ctx_batch_after(b->mpc[i], s)) return M2C_ERR_HIP;
  }
  for (int i : b->mp_list)
    if (m2c_ctx_batch_bind(b->mpc[i]) || m2c_ctx_batch_end_head(b->mpc[i], time_step)) return M2C_ERR_HIP;
  return M2C_OK;
}

// the tail: two-material instances' state update after interface motion and CFL minimum (the
// single-material instances' minimum comes from their final stage), then every instance's dt
static int enqueue_batch_tail(m2c_batch *b) {
  for (int i : b->mp_list)
    if (m2c_ctx_batch_bind(b->mpc[i]) || m2c_ctx_batch_end_tail(b->mpc[i])) return M2C_ERR_HIP;
  if (!b->mp_list.empty())
    hipLaunchKernelGGL(k_batch_cfl, dim3(256, (unsigned)b->mp_list.size()), dim3(256), 0, b->stream, b->dinst,
                       b->d_mp_list, b->clk, b->st);
  hipLaunchKernelGGL(k_batch_clock_finalize, dim3((b->n + 63) / 64), dim3(64), 0, b->stream, b->clk, b->n, 1);
  return M2C_OK;
}

static int batch_retries(m2c_batch *b, long long time_step) {
  for (int i : b->mp_list)
    if (m2c_ctx_batch_bind(b->mpc[i]) || m2c_ctx_batch_end_retries(b->mpc[i], time_step)) return M2C_ERR_HIP;
  return M2C_OK;
}

static void enqueue_batch_cfl(m2c_batch *b, int maxsteps_more) {
  hipLaunchKernelGGL(k_batch_clock_set, dim3((b->n + 63) / 64), dim3(64), 0, b->stream, b->clk, b->n, maxsteps_more);
  hipLaunchKernelGGL(k_batch_cfl, dim3(256, b->n), dim3(256), 0, b->stream, b->dinst, (const int *)nullptr, b->clk,
                     b->st);
  hipLaunchKernelGGL(k_batch_clock_finalize, dim3((b->n + 63) / 64), dim3(64), 0, b->stream, b->clk, b->n, 0);
}

// capture what `fn` enqueues on the batch stream into an executable graph
template <class F> static int capture(m2c_batch *b, hipGraphExec_t *exec, F fn) {
  hipGraph_t g = nullptr;
  BHIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = fn();
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) { if (g) hipGraphDestroy(g); return rc; }
  BHIP(e);
  e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  BHIP(e);
  return M2C_OK;
}

// one time step of every instance
static int batch_step(m2c_batch *b, int use_graph) {
  const long long ts = b->steps + 1;
  int rc;
  if (b->mp_list.empty()) {
    if (use_graph) {
      if (!b->exec && (rc = capture(b, &b->exec, [&] {
            int r = enqueue_batch_head(b, ts);
            return r ? r : enqueue_batch_tail(b);
          })))
        return rc;
      BHIP(hipGraphLaunch(b->exec, b->stream));
    } else if ((rc = enqueue_batch_head(b, ts)) || (rc = enqueue_batch_tail(b))) {
      return rc;
    }
  } else if (!reinit_step(b, ts)) {
    if (use_graph) {
      if (!b->exec && (rc = capture(b, &b->exec, [&] {
            int r = enqueue_batch_head(b, ts);
            return r ? r : enqueue_batch_tail(b);
          })))
        return rc;
      BHIP(hipGraphLaunch(b->exec, b->stream));
    } else if ((rc = enqueue_batch_head(b, ts)) || (rc = enqueue_batch_tail(b))) {
      return rc;
    }
  } else {
    // a reinitialisation step: the host reads each instance's failsafe flag between head and tail
    if (use_graph) {
      if (!b->exec_rhead && (rc = capture(b, &b->exec_rhead, [&] { return enqueue_batch_head(b, ts); }))) return rc;
      if (!b->exec_rtail && (rc = capture(b, &b->exec_rtail, [&] { return enqueue_batch_tail(b); }))) return rc;
      BHIP(hipGraphLaunch(b->exec_rhead, b->stream));
      if ((rc = batch_retries(b, ts))) return rc;
      BHIP(hipGraphLaunch(b->exec_rtail, b->stream));
    } else if ((rc = enqueue_batch_head(b, ts)) || (rc = batch_retries(b, ts)) || (rc = enqueue_batch_tail(b))) {
      return rc;
    }
  }
  b->steps += 1;
  return M2C_OK;
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
  b->mpc.assign(n, nullptr);
  std::vector<int> variants(n);
  int nmat_max = 0;
  for (int i = 0; i < n; i++) {
    const m2c_config &c = cfgs[i];
    // one kernel variant and one material table for the whole batch: the materials an instance
    // defines are the leading entries of the batch's table
    const int nm = std::min(c.nmat, cfgs[0].nmat);
    if (c.ts != b->ts || c.flux != b->flux || c.precision != cfgs[0].precision || c.device != cfgs[0].device ||
        c.nmat < 1 || c.nmat > M2C_MAX_MAT || memcmp(c.mat, cfgs[0].mat, sizeof(m2c_material) * nm) != 0) {
      delete b;
      return M2C_ERR_UNSUPPORTED;
    }
    for (int m = 0; m < c.nmat; m++)      // inviscid, adiabatic instances only
      if (c.mat[m].viscosity != 0 || c.mat[m].kappa != 0.0) {
        delete b;
        return M2C_ERR_UNSUPPORTED;
      }
    DevEos eos[8];
    int rc = m2c_fill_grid(&c, &b->hinst[i].grid, eos, &variants[i]);
    if (rc) { delete b; return rc; }
    if (b->hinst[i].grid.stretched) { delete b; return M2C_ERR_UNSUPPORTED; }   // uniform meshes only
    // two-material instances: materials meeting at interface faces, or a level set
    if (b->hinst[i].grid.has_interfaces || c.ls_matid >= 0) b->mp_list.push_back(i);
    if (c.nmat > nmat_max) { nmat_max = c.nmat; memcpy(b->eos, eos, sizeof(eos)); }
  }
  // with two-material instances the batch's variant is the one with the material-interface code
  // (the stiffened-gas ones: 9 for FP64, 10 for the mixed form; the run-time-EOS variants have it)
  const bool anymp = !b->mp_list.empty();
  for (int i = 0; i < n; i++) {
    int v = variants[i];
    if (anymp && v == 0) v = 9;
    if (anymp && v == 4) v = 10;
    if (i == 0) b->variant = v;
    else if (v != b->variant) { delete b; return M2C_ERR_UNSUPPORTED; }
  }
  // HLLC (0, 1, 9), Roe (2) and LLF (3) in FP64, HLLC in the mixed form (4, 5, 10)
  if (b->variant < 0 || (b->variant > 5 && b->variant != 9 && b->variant != 10)) { delete b; return M2C_ERR_UNSUPPORTED; }
  // the single-material instances of a batch with two-material ones keep the stiffened-gas variant
  // without the interface code in a launch of their own (the interface code costs them ~40 %)
  b->n_m = (int)b->mp_list.size();
  b->n_s = n - b->n_m;
  if (anymp && b->n_s > 0 && (b->variant == 9 || b->variant == 10)) b->variant_s = b->variant == 9 ? 0 : 4;
  std::vector<char> two(n, 0);
  for (int i : b->mp_list) two[i] = 1;
  for (int i = 0; i < n; i++) {
    if (two[i]) continue;
    b->hinst[i].cell0 = b->cells;
    b->cells += b->hinst[i].grid.N;
  }
  // from here on a failure releases everything allocated so far
#define BHIP_D(x)                                                                                  \
  do {                                                                                             \
    if ((x) != hipSuccess) { m2c_batch_destroy(b); return M2C_ERR_HIP; }                            \
  } while (0)
  BHIP_D(hipSetDevice(cfgs[0].device));
  BHIP_D(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  const long long cells = std::max<long long>(b->cells, 1);
  size_t bytes = sizeof(double) * 5 * (size_t)cells;
  BHIP_D(hipMalloc(&b->V, bytes));
  BHIP_D(hipMalloc(&b->U1, bytes));
  if (b->nstage == 3) BHIP_D(hipMalloc(&b->U2, bytes));
  BHIP_D(hipMalloc(&b->ID, (size_t)cells));
  BHIP_D(hipMemset(b->ID, 0, (size_t)cells));
  BHIP_D(hipMemset(b->V, 0, bytes));
  long long stage_cells = cells;   // the host-transfer buffer fits the largest instance
  for (int i = 0; i < n; i++) stage_cells = std::max<long long>(stage_cells, b->hinst[i].grid.N);
  BHIP_D(hipMalloc(&b->stage_buf, sizeof(double) * 5 * (size_t)stage_cells));
  BHIP_D(hipMalloc(&b->clk, sizeof(Clock) * n));
  BHIP_D(hipMalloc(&b->st, sizeof(Status) * n));
  BHIP_D(hipMemset(b->st, 0, sizeof(Status) * n));
  BHIP_D(hipMalloc(&b->dinst, sizeof(BatchInst) * n));
  BHIP_D(hipMalloc(&b->q, sizeof(TileQueue)));
  if (b->variant_s >= 0) {
    BHIP_D(hipMalloc(&b->dinst_s, sizeof(BatchInst) * b->n_s));
    BHIP_D(hipMalloc(&b->dinst_m, sizeof(BatchInst) * b->n_m));
    BHIP_D(hipMalloc(&b->q_s, sizeof(TileQueue)));
    BHIP_D(hipMalloc(&b->q_m, sizeof(TileQueue)));
  }
  BHIP_D(hipMalloc(&b->d_mp_list, sizeof(int) * std::max<size_t>(1, b->mp_list.size())));
  if (anymp)
    BHIP_D(hipMemcpy(b->d_mp_list, b->mp_list.data(), sizeof(int) * b->mp_list.size(), hipMemcpyHostToDevice));
  // stage arguments of every instance (the same buffers as the single-instance context)
  for (int i = 0; i < n; i++) {
    BatchInst &B = b->hinst[i];
    if (two[i]) {
      m2c_ctx *c = nullptr;
      int rc = m2c_ctx_create(&cfgs[i], &c);
      if (rc) { m2c_batch_destroy(b); return rc; }
      b->mpc[i] = c;
      if ((rc = m2c_ctx_batch_attach(c, b->stream, b->clk + i, b->st + i))) { m2c_batch_destroy(b); return rc; }
      int v = 0;
      for (int s = 0; s < b->nstage; s++)
        if ((rc = m2c_ctx_batch_stage_args(c, s, &B.args[s], &B.Vc, &B.IDc, &v))) { m2c_batch_destroy(b); return rc; }
      if (v != b->variant) { m2c_batch_destroy(b); return M2C_ERR_UNSUPPORTED; }
      continue;
    }
    const long long o = 5 * B.cell0;
    double *V = b->V + o, *U1 = b->U1 + o, *U2 = b->U2 ? b->U2 + o : nullptr;
    B.Vc = V;
    B.IDc = b->ID + B.cell0;
    for (int s = 0; s < 3; s++) {
      StageArgs &a = B.args[s];
      a = StageArgs();
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
  if (anymp) {   // one multi-instance interface pass per stage when every instance can join it
    std::vector<IfaceInst> tab;
    int blocks = 0, lblocks = 0, flux = -1;
    bool all = true;
    for (int i : b->mp_list) {
      IfaceInst t;
      int ok = 0, fl = 0;
      if (int rc = m2c_ctx_batch_iface_inst(b->mpc[i], &t, &ok, &fl)) { m2c_batch_destroy(b); return rc; }
      if (!ok || (t.nblk > 0 && flux >= 0 && fl != flux)) { all = false; break; }
      if (t.nblk == 0) continue;
      flux = fl;
      t.blk0 = blocks;
      blocks += t.nblk;
      t.lblk0 = lblocks;
      lblocks += t.lnblk;
      tab.push_back(t);
    }
    if (all && !tab.empty()) {
      BHIP_D(hipMalloc(&b->d_iface, sizeof(IfaceInst) * tab.size()));
      BHIP_D(hipMemcpy(b->d_iface, tab.data(), sizeof(IfaceInst) * tab.size(), hipMemcpyHostToDevice));
      b->n_iface = (int)tab.size();
      b->iface_blocks = blocks;
      b->list_blocks = lblocks;
      b->iface_flux = flux;
    }
  }
  if (int rc = batch_upload_tables(b)) { m2c_batch_destroy(b); return rc; }
  if (int rc = batch_layout(b, 8, 0)) { m2c_batch_destroy(b); return rc; }
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
      BHIP_D(hipStreamSynchronize(b->stream));
    }
  BHIP_D(hipStreamSynchronize(b->stream));
#undef BHIP_D
  *out = b;
  return M2C_OK;
}

int m2c_batch_destroy(m2c_batch *b) {
  if (!b) return M2C_OK;
  if (b->stream) hipStreamSynchronize(b->stream);
  for (m2c_ctx *c : b->mpc) m2c_ctx_destroy(c);
  drop_graphs(b);
  hipFree(b->V); hipFree(b->U1); hipFree(b->U2); hipFree(b->stage_buf); hipFree(b->backup);
  hipFree(b->ID); hipFree(b->clk); hipFree(b->st); hipFree(b->dinst); hipFree(b->q); hipFree(b->d_mp_list);
  hipFree(b->d_iface);
  hipFree(b->dinst_s); hipFree(b->dinst_m); hipFree(b->q_s); hipFree(b->q_m);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return M2C_OK;
}

int m2c_batch_upload(m2c_batch *b, int i, const double *V, const int32_t *id) {
  if (!b || i < 0 || i >= b->n || !V) return M2C_ERR_ARG;
  if (b->mpc[i]) return m2c_upload_state(b->mpc[i], V, id);
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
  if (b->mpc[i]) return m2c_download_state(b->mpc[i], V, id);
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

int m2c_batch_upload_levelset(m2c_batch *b, int i, const double *phi) {
  if (!b || i < 0 || i >= b->n || !b->mpc[i]) return M2C_ERR_ARG;
  return m2c_upload_levelset(b->mpc[i], phi);
}

int m2c_batch_download_levelset(m2c_batch *b, int i, double *phi) {
  if (!b || i < 0 || i >= b->n || !b->mpc[i]) return M2C_ERR_ARG;
  return m2c_download_levelset(b->mpc[i], phi);
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
  int rc = M2C_OK;
  for (int s = 0; s < nsteps && !rc; s++) rc = batch_step(b, use_graph);
  if (!rc && hipGetLastError() != hipSuccess) rc = M2C_ERR_HIP;
  BHIP(hipEventRecord(e1, b->stream));
  BHIP(hipEventSynchronize(e1));
  float t = 0.f;
  BHIP(hipEventElapsedTime(&t, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (ms) *ms = t;
  return rc;
}

int m2c_batch_status(m2c_batch *b, int i, m2c_status *out) {
  if (!b || i < 0 || i >= b->n || !out) return M2C_ERR_ARG;
  if (b->mpc[i]) {   // the context reads the clock and status the batch owns
    int rc = m2c_get_status(b->mpc[i], out);
    out->cfl = b->cfg[i].cfl;
    return rc;
  }
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
  const size_t bytes = sizeof(double) * 5 * (size_t)std::max<long long>(b->cells, 1);
  if (!b->backup) BHIP(hipMalloc(&b->backup, bytes));
  // tune on a copy of the state; the clocks, counters and step count are restored afterwards
  std::vector<Clock> hclk(b->n);
  std::vector<Status> hst(b->n);
  BHIP(hipMemcpy(hclk.data(), b->clk, sizeof(Clock) * b->n, hipMemcpyDeviceToHost));
  BHIP(hipMemcpy(hst.data(), b->st, sizeof(Status) * b->n, hipMemcpyDeviceToHost));
  const long long steps0 = b->steps;
  hipLaunchKernelGGL(k_batch_copy, dim3(4096), dim3(256), 0, b->stream, b->V, b->backup, 5 * b->cells);
  for (int i : b->mp_list)
    if (int rc = m2c_ctx_batch_backup(b->mpc[i], 0)) return rc;
  auto restore = [&]() -> int {
    hipLaunchKernelGGL(k_batch_copy, dim3(4096), dim3(256), 0, b->stream, b->backup, b->V, 5 * b->cells);
    for (int i : b->mp_list)
      if (int rc = m2c_ctx_batch_backup(b->mpc[i], 1)) return rc;
    BHIP(hipStreamSynchronize(b->stream));
    BHIP(hipMemcpy(b->clk, hclk.data(), sizeof(Clock) * b->n, hipMemcpyHostToDevice));
    BHIP(hipMemcpy(b->st, hst.data(), sizeof(Status) * b->n, hipMemcpyHostToDevice));
    b->steps = steps0;
    return M2C_OK;
  };
  const int kchunks[] = {4, 8, 16, 32};
  const int blocks[] = {256, 512, 1024};
  double best = 1e300;
  int bk = b->kchunk, bb = b->nblocks, bs = b->shape, ncand = 0;
  for (int sh = 0; sh < BATCH_TILES; sh++)
    for (int kc : kchunks)
      for (int nb : blocks) {
        if (int rc = batch_layout(b, kc, sh)) return rc;
        b->nblocks = nb;
        if (int rc = restore()) return rc;
        double ms = 0.0;
        if (int rc = m2c_batch_run(b, trial_steps, 1, &ms)) return rc;
        ncand++;
        if (ms < best) { best = ms; bk = kc; bb = nb; bs = sh; }
      }
  if (int rc = batch_layout(b, bk, bs)) return rc;
  b->nblocks = bb;
  if (int rc = restore()) return rc;
  BHIP(hipStreamSynchronize(b->stream));
  if (out) {
    out->kchunk = bk;
    out->blocks = bb;
    batch_tile_dims(bs, &out->tile_x, &out->tile_y);
    out->ms_per_step = best / trial_steps;
    out->candidates = ncand;
  }
  return M2C_OK;
}

int m2c_batch_set_tuning(m2c_batch *b, int kchunk, int blocks) {
  if (!b || kchunk < 1 || blocks < 1) return M2C_ERR_ARG;
  if (int rc = batch_layout(b, kchunk, b->shape)) return rc;
  b->nblocks = blocks;
  return M2C_OK;
}

int m2c_batch_set_tile(m2c_batch *b, int tile_x, int tile_y) {
  if (!b) return M2C_ERR_ARG;
  for (int sh = 0; sh < BATCH_TILES; sh++) {
    int tx, ty;
    batch_tile_dims(sh, &tx, &ty);
    if (tx == tile_x && ty == tile_y) return batch_layout(b, b->kchunk, sh);
  }
  return M2C_ERR_UNSUPPORTED;
}

}  // extern "C"
