// batch_kern.hpp -- device side of the batched runner shared by its host unit (batch.hip) and the
// per-variant stage units (batch_stage.hip): the instance table, the per-XCD tile queues and the
// persistent stage kernel.
#pragma once
#include <hip/hip_runtime.h>

// every unit that includes this header defines M2C_BATCHED before its first include of
// fv_kernels.hpp (a block reads the grid of its instance from LDS)
#include "fv_kernels.hpp"
#ifndef M2C_BATCHED_LAYOUT
#error "define M2C_BATCHED before including fv_kernels.hpp in a batched unit"
#endif

namespace m2c {

constexpr int NXCD = 8;

struct BatchInst {
  Grid grid;
  StageArgs args[3];           // per stage of the time scheme
  int gx, gy, nzc;             // tiles in x, y; z chunks
  unsigned tile0;              // first work item of this instance
  long long cell0;             // offset of the instance's cells in the concatenated fields
  const double *Vc;            // the instance's primitive state and IDs (the CFL pass)
  const int8_t *IDc;
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

// tile shapes of the batched kernel (the autotuner picks one): 32 x 8 and 16 x 16 columns
using BatchTile0 = TileShape<32, 8>;
using BatchTile1 = TileShape<16, 16>;
constexpr int BATCH_TILES = 2;
inline void batch_tile_dims(int shape, int *tx, int *ty) {
  *tx = shape ? BatchTile1::TX : BatchTile0::TX;
  *ty = shape ? BatchTile1::TY : BatchTile0::TY;
}

template <int IN_CONS, int MODE, int FLUX, int EK, typename TC, bool MP, class TS>
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
    stage_body<IN_CONS, MODE, FLUX, EK, TC, MP, TS>(a, rem % B.gx, rem / B.gx, bz);
    __syncthreads();
  }
}

}  // namespace m2c

// per-variant stage units: the launch of one stage of every instance (shape 0: 32 x 8 tiles, 1: 16 x 16)
// and the upload of that unit's material table
#define M2C_BSTAGE_DECL(tag)                                                                              \
  extern "C" int m2c_bstage_launch_##tag(int shape, int in_cons, int mode, int nblocks, hipStream_t s,     \
                                          const m2c::BatchInst *insts, int n, m2c::TileQueue *q, int stage); \
  extern "C" int m2c_bstage_upload_##tag(const m2c::DevEos *eos);
M2C_BSTAGE_DECL(hllc_sg)
M2C_BSTAGE_DECL(hllc_gen)
M2C_BSTAGE_DECL(roe_gen)
M2C_BSTAGE_DECL(llf_gen)
M2C_BSTAGE_DECL(hllc_sg_f32)
M2C_BSTAGE_DECL(hllc_gen_f32)
M2C_BSTAGE_DECL(hllc_sg_mp)
M2C_BSTAGE_DECL(hllc_sg_mp_f32)
#undef M2C_BSTAGE_DECL
