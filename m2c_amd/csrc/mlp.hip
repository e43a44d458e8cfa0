// mlp.hip -- the learned interface solver on gfx950: bf16 MFMA network kernels (inference,
// training forward and backward, weight gradients, Adam), training samples from the exact
// two-material Riemann solver, and the network's C ABI.
//
// MFMA: v_mfma_f32_16x16x32_bf16. A (16 x 32): lane l holds row l%16, k = 8*(l/16) .. +7;
// B (32 x 16): lane l holds column l%16, the same k; C (16 x 16, FP32): lane l holds column l%16,
// rows 4*(l/16) .. +3.
//
// Forward: one workgroup (4 waves) per tile of 64 samples runs all five layers; activations stay in
// LDS (two 64 x 264 bf16 buffers), each wave computes 64 output columns of a hidden layer, weights
// stream from L2. Training stores the post-activation H and SiLU' of every hidden layer transposed
// ([feature][sample], bf16), the backward kernel chains dZ_l = (dZ_{l+1} W_{l+1}) * SiLU'(Z_l) per
// tile in LDS, and the weight gradients dW_l = dZ_{l+1}^T H_l are split-K MFMA GEMMs over the batch.
// A row of ones appended to every H (column 10 of the padded input) makes the same GEMMs produce
// the bias gradients.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/m2c_hip.h"
#include "mlp.hpp"
#include "riemann_dev.hpp"

namespace m2c {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int LDS_S = MLP_H + 8;   // bf16 per LDS activation row: 528 B, rows start 4 banks apart

__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float silu_f(float z, float *dsilu) {
  const float s = 1.0f / (1.0f + __expf(-z));
  *dsilu = s * (1.0f + z * (1.0f - s));
  return z * s;
}

// acc[mt][nt] = in[64 x K] (LDS rows) * W[c0 + 16 nt + (0..15)][0..K)^T for the wave's 64 columns
template <int K>
__device__ __forceinline__ void tile_mma(const __bf16 *in, const __bf16 *W, int c0, f32x4 (&acc)[4][4]) {
  const int lane = threadIdx.x & 63, r = lane & 15, kq = (lane >> 4) * 8;
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
  for (int k0 = 0; k0 < K; k0 += 32) {
    bf16x8 a[4], b[4];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) b[nt] = *(const bf16x8 *)(W + (size_t)(c0 + nt * 16 + r) * K + k0 + kq);
#pragma unroll
    for (int mt = 0; mt < 4; mt++) a[mt] = *(const bf16x8 *)(in + (mt * 16 + r) * LDS_S + k0 + kq);
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
      for (int nt = 0; nt < 4; nt++) acc[mt][nt] = mfma16(a[mt], b[nt], acc[mt][nt]);
  }
}

struct FwdArgs {
  const float *X;            // [rows][10]
  float *Y;                  // [rows][3] or NULL (training)
  int n;                     // rows (the capacity when n_dev is set)
  const unsigned *n_dev;     // device-side row count or NULL
  // training only
  const float *T;            // targets [N][3]; NaN marks a sample without a solution
  unsigned long long pa, pc, pmask;   // batch row q is sample (pa * (base + q) + pc) & pmask
  long long base;
  int B;                     // batch size = row length of the transposed buffers
  __bf16 *XT;                // [32][B]
  __bf16 *HT[5], *DT[5];     // l = 1..4: HT [MLP_HROWS][B], DT [256][B]
  __bf16 *dZT5;              // [16][B]
  float gscale;              // d loss / d y = gscale * (y - t)
  float *loss;               // += sum of squared errors
};

__device__ __forceinline__ long long train_sample(const FwdArgs &a, int q) {
  return (long long)((a.pa * (unsigned long long)(a.base + q) + a.pc) & a.pmask);
}

// one hidden layer of a tile, in place: every wave reads all K input columns of the tile's rows, so
// the outputs are written only after the whole workgroup has finished reading
template <int K, int TRAIN>
__device__ __forceinline__ void fwd_hidden(__bf16 *act, const __bf16 *W, const float *bias, __bf16 *HT, __bf16 *DT,
                                           int B, int r0) {
  const int lane = threadIdx.x & 63, c0 = (threadIdx.x >> 6) * 64;
  __bf16 *out = act;
  f32x4 acc[4][4];
  tile_mma<K>(act, W, c0, acc);
  __syncthreads();
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    const int col = c0 + nt * 16 + (lane & 15);
    const float bc = bias[col];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      const int row = mt * 16 + 4 * (lane >> 4);
      bf16x4 hv, dv;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        float d;
        const float h = silu_f(acc[mt][nt][i] + bc, &d);
        hv[i] = (__bf16)h;
        dv[i] = (__bf16)d;
        out[(row + i) * LDS_S + col] = hv[i];
      }
      if (TRAIN) {
        *(bf16x4 *)(HT + (size_t)col * B + r0 + row) = hv;
        *(bf16x4 *)(DT + (size_t)col * B + r0 + row) = dv;
      }
    }
  }
}

// the 64-sample forward with one activation buffer (33 KB): TRAIN = 1 is the training forward (two
// workgroups per CU, more live registers for the stored activations); TRAIN = 0 evaluates the same
// network at four workgroups per CU (inference runs k_mlp_infer below)
template <int TRAIN>
__global__ __launch_bounds__(256, TRAIN ? 2 : 4) void k_mlp_fwd(MlpWeights w, FwdArgs a) {
  __shared__ __attribute__((aligned(16))) __bf16 buf[1][MLP_TM * LDS_S];
  const int count = a.n_dev ? min((int)*a.n_dev, a.n) : a.n;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, r = lane & 15, kq = (lane >> 4) * 8;
  for (int tile = blockIdx.x; tile * MLP_TM < count; tile += gridDim.x) {
    const int r0 = tile * MLP_TM;
    for (int t = threadIdx.x; t < MLP_TM * MLP_K0; t += 256) {
      const int row = t / MLP_K0, col = t % MLP_K0, q = r0 + row;
      float x = 0.f;
      if (q < count && col < MLP_IN) {
        const long long s = TRAIN ? train_sample(a, q) : q;
        x = a.X[s * MLP_IN + col];
      }
      const __bf16 xb = (__bf16)x;
      buf[0][row * LDS_S + col] = xb;
      if (TRAIN && col < MLP_IN) a.XT[(size_t)col * a.B + q] = xb;
    }
    __syncthreads();
    fwd_hidden<MLP_K0, TRAIN>(buf[0], w.W[0], w.bias[0], a.HT[1], a.DT[1], a.B, r0);
    __syncthreads();
    fwd_hidden<MLP_H, TRAIN>(buf[0], w.W[1], w.bias[1], a.HT[2], a.DT[2], a.B, r0);
    __syncthreads();
    fwd_hidden<MLP_H, TRAIN>(buf[0], w.W[2], w.bias[2], a.HT[3], a.DT[3], a.B, r0);
    __syncthreads();
    fwd_hidden<MLP_H, TRAIN>(buf[0], w.W[3], w.bias[3], a.HT[4], a.DT[4], a.B, r0);
    __syncthreads();
    // output layer: wave wv takes rows 16 wv .. +15, 16 (padded) output columns
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
    for (int k0 = 0; k0 < MLP_H; k0 += 32) {
      const bf16x8 av = *(const bf16x8 *)(buf[0] + (wv * 16 + r) * LDS_S + k0 + kq);
      const bf16x8 bv = *(const bf16x8 *)(w.W[4] + (size_t)r * MLP_H + k0 + kq);
      acc = mfma16(av, bv, acc);
    }
    const int col = lane & 15;
    const float bc = w.bias[4][col];
    float sq = 0.f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int q = r0 + wv * 16 + 4 * (lane >> 4) + i;
      const float y = acc[i] + bc;
      if (!TRAIN) {
        if (col < MLP_OUT && q < count) a.Y[(size_t)q * MLP_OUT + col] = y;
      } else {
        float g = 0.f;
        if (col < MLP_OUT) {
          const float t = a.T[train_sample(a, q) * MLP_OUT + col];
          if (t == t) {
            const float d = y - t;
            sq += d * d;
            g = a.gscale * d;
          }
          if (a.Y) a.Y[(size_t)q * MLP_OUT + col] = y;
        }
        a.dZT5[(size_t)col * a.B + q] = (__bf16)g;
      }
    }
    if (TRAIN) {
      for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off, 64);
      if (lane == 0) atomicAdd(a.loss, sq);
    }
    __syncthreads();   // buf[0] is restaged by the next tile
  }
}

// ---- inference: 128-sample tiles in MFMA fragment order -----------------------------------------
//
// The product is formed transposed, Z^T = W X^T: the A operand is 16 output features x 32 inputs of
// W, the B operand 32 inputs x 16 samples of the activation tile, so a lane's accumulator holds
// four consecutive features of one sample. Weights (MlpWeights::WF) and the LDS activation tile are
// both stored in fragment order: every A or B fragment a wave reads is one contiguous kilobyte (64
// lanes x 16 bytes), and the epilogue's four features of a sample are one 8-byte LDS store. Each
// wave computes 64 features for all 128 samples (4 x 8 MFMA tiles, 128 accumulator registers):
// every weight fragment read from L2 feeds eight MFMAs, and the fragments of consecutive K steps
// alternate between two register sets so that a step waits only for loads issued a step earlier.
constexpr int MLP_TI = 128;   // samples per inference tile

// bf16 offset of (sample s, feature f) in a fragment-ordered tile with ks_n K steps of 32 features:
// B fragment (sample tile s/16, K step f/32) at 512 (s/16 * ks_n + f/32), lane s%16 + 16 (f%32/8)
__device__ __forceinline__ int frag_off(int s, int f, int ks_n) {
  return (((s >> 4) * ks_n + (f >> 5)) * 64 + (s & 15) + 16 * ((f & 31) >> 3)) * 8 + (f & 7);
}

template <int KS>
__device__ __forceinline__ void mma_step_inf(const __bf16 *act, int ks, const bf16x8 (&a)[4], f32x4 (&acc)[4][8]) {
  const int lane = threadIdx.x & 63;
  bf16x8 b[8];
#pragma unroll
  for (int nt = 0; nt < 8; nt++) b[nt] = *(const bf16x8 *)(act + ((nt * KS + ks) * 64 + lane) * 8);
#pragma unroll
  for (int nt = 0; nt < 8; nt++)
#pragma unroll
    for (int mt = 0; mt < 4; mt++) acc[mt][nt] = mfma16(a[mt], b[nt], acc[mt][nt]);
}

// acc = W[f0 .. f0+63][0..K) x act[0..127][0..K)^T
template <int K>
__device__ __forceinline__ void tile_mma_inf(const __bf16 *act, const __bf16 *WF, int f0, f32x4 (&acc)[4][8]) {
  constexpr int KS = K / 32;
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < 8; nt++) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  // the fragment of feature tile f0/16 + mt and K step ks: 1 KB at wp + (mt KS + ks) 512
  const __bf16 *wp = WF + ((size_t)(f0 / 16) * KS * 64 + lane) * 8;
  bf16x8 a0[4], a1[4];
#pragma unroll
  for (int mt = 0; mt < 4; mt++) a0[mt] = *(const bf16x8 *)(wp + (size_t)mt * KS * 512);
#pragma unroll 1
  for (int ks = 0; ks < KS; ks += 2) {
    if (ks + 1 < KS) {
#pragma unroll
      for (int mt = 0; mt < 4; mt++) a1[mt] = *(const bf16x8 *)(wp + (size_t)(mt * KS + ks + 1) * 512);
    }
    mma_step_inf<KS>(act, ks, a0, acc);
    if (ks + 1 < KS) {
      if (ks + 2 < KS) {
#pragma unroll
        for (int mt = 0; mt < 4; mt++) a0[mt] = *(const bf16x8 *)(wp + (size_t)(mt * KS + ks + 2) * 512);
      }
      mma_step_inf<KS>(act, ks + 1, a1, acc);
    }
  }
}

// SiLU for inference: the hardware exponential and reciprocal (about one FP32 ulp), then bf16
__device__ __forceinline__ float silu_inf(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}

// one hidden layer of an inference tile, in place (outputs, 256 features in fragment order, written
// after every wave has read the layer's inputs)
template <int K>
__device__ __forceinline__ void inf_hidden(__bf16 *act, const __bf16 *WF, const float *bias) {
  const int lane = threadIdx.x & 63, f0 = (threadIdx.x >> 6) * 64;
  f32x4 acc[4][8];
  tile_mma_inf<K>(act, WF, f0, acc);
  __syncthreads();
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    const int f = f0 + mt * 16 + 4 * (lane >> 4);
    const float b0 = bias[f], b1 = bias[f + 1], b2 = bias[f + 2], b3 = bias[f + 3];
#pragma unroll
    for (int nt = 0; nt < 8; nt++) {
      bf16x4 hv;
      hv[0] = (__bf16)silu_inf(acc[mt][nt][0] + b0);
      hv[1] = (__bf16)silu_inf(acc[mt][nt][1] + b1);
      hv[2] = (__bf16)silu_inf(acc[mt][nt][2] + b2);
      hv[3] = (__bf16)silu_inf(acc[mt][nt][3] + b3);
      *(bf16x4 *)(act + frag_off(nt * 16 + (lane & 15), f, MLP_H / 32)) = hv;
    }
  }
}

// persistent inference: two 128-sample tiles per CU at a time (64 KB of LDS each)
__global__ __launch_bounds__(256, 2) void k_mlp_infer(MlpWeights w, FwdArgs a) {
  __shared__ __attribute__((aligned(16))) __bf16 buf[MLP_TI * MLP_H];
  const int count = a.n_dev ? min((int)*a.n_dev, a.n) : a.n;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int tile = blockIdx.x; tile * MLP_TI < count; tile += gridDim.x) {
    const int r0 = tile * MLP_TI;
    for (int t = threadIdx.x; t < MLP_TI * MLP_K0; t += 256) {   // inputs: one K step of 32
      const int row = t / MLP_K0, col = t % MLP_K0, q = r0 + row;
      const float x = (q < count && col < MLP_IN) ? a.X[(size_t)q * MLP_IN + col] : 0.f;
      buf[frag_off(row, col, 1)] = (__bf16)x;
    }
    __syncthreads();
    inf_hidden<MLP_K0>(buf, w.WF[0], w.bias[0]);
    __syncthreads();
    inf_hidden<MLP_H>(buf, w.WF[1], w.bias[1]);
    __syncthreads();
    inf_hidden<MLP_H>(buf, w.WF[2], w.bias[2]);
    __syncthreads();
    inf_hidden<MLP_H>(buf, w.WF[3], w.bias[3]);
    __syncthreads();
    // output layer: the 16 (padded) output rows of W4 as A, wave wv's sample tiles 2 wv, 2 wv + 1
    const int r = lane & 15, kq = (lane >> 4) * 8;
    f32x4 acc[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll 4
    for (int ks = 0; ks < MLP_H / 32; ks++) {
      const bf16x8 av = *(const bf16x8 *)(w.W[4] + (size_t)r * MLP_H + 32 * ks + kq);
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const bf16x8 bv = *(const bf16x8 *)(buf + (((2 * wv + h) * (MLP_H / 32) + ks) * 64 + lane) * 8);
        acc[h] = mfma16(av, bv, acc[h]);
      }
    }
    // lane: sample 32 wv + 16 h + lane%16, outputs 4 (lane/16) .. +3; outputs 0..2 are the network's
    if ((lane >> 4) == 0) {
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int q = r0 + wv * 32 + h * 16 + r;
        if (q < count) {
#pragma unroll
          for (int o = 0; o < MLP_OUT; o++) a.Y[(size_t)q * MLP_OUT + o] = acc[h][o] + w.bias[4][o];
        }
      }
    }
    __syncthreads();   // buf is restaged by the next tile
  }
}

struct BwdArgs {
  int B;
  const __bf16 *dZT5;
  const __bf16 *DT[5];
  __bf16 *dZT[5];     // l = 1..4: [256][B]
};

// dZ_l of this tile: (dZ_{l+1} [64 x K] from LDS) * WT [256][K]^T times SiLU'(Z_l)
template <int K>
__device__ __forceinline__ void bwd_hidden(const __bf16 *in, __bf16 *out, const __bf16 *WT, const __bf16 *DT,
                                           __bf16 *dZT, int B, int r0) {
  const int lane = threadIdx.x & 63, c0 = (threadIdx.x >> 6) * 64;
  f32x4 acc[4][4];
  tile_mma<K>(in, WT, c0, acc);
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    const int col = c0 + nt * 16 + (lane & 15);
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      const int row = mt * 16 + 4 * (lane >> 4);
      const bf16x4 dv = *(const bf16x4 *)(DT + (size_t)col * B + r0 + row);
      bf16x4 zv;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        zv[i] = (__bf16)(acc[mt][nt][i] * (float)dv[i]);
        out[(row + i) * LDS_S + col] = zv[i];
      }
      *(bf16x4 *)(dZT + (size_t)col * B + r0 + row) = zv;
    }
  }
}

__global__ __launch_bounds__(256, 2) void k_mlp_bwd(MlpWeights w, BwdArgs a) {
  __shared__ __attribute__((aligned(16))) __bf16 buf[2][MLP_TM * LDS_S];
  const int r0 = blockIdx.x * MLP_TM;
  for (int t = threadIdx.x; t < MLP_TM * 32; t += 256) {
    const int row = t / 32, col = t % 32;
    buf[0][row * LDS_S + col] = col < MLP_O4 ? a.dZT5[(size_t)col * a.B + r0 + row] : (__bf16)0.f;
  }
  __syncthreads();
  bwd_hidden<32>(buf[0], buf[1], w.WT[4], a.DT[4], a.dZT[4], a.B, r0);
  __syncthreads();
  bwd_hidden<MLP_H>(buf[1], buf[0], w.WT[3], a.DT[3], a.dZT[3], a.B, r0);
  __syncthreads();
  bwd_hidden<MLP_H>(buf[0], buf[1], w.WT[2], a.DT[2], a.dZT[2], a.B, r0);
  __syncthreads();
  bwd_hidden<MLP_H>(buf[1], buf[0], w.WT[1], a.DT[1], a.dZT[1], a.B, r0);
}

// split-K weight gradient: part[split][o][i] = sum over the split's samples of dZT[o][s] * HT[i][s].
// One wave per 16 x 16 tile; a workgroup covers 16 rows x 64 columns.
__global__ __launch_bounds__(256) void k_mlp_dw(const __bf16 *dZT, const __bf16 *HT, int O, int Ip, int B, int kc,
                                                 float *part) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, r = lane & 15, kq = (lane >> 4) * 8;
  const int o0 = blockIdx.y * 16, i0 = blockIdx.x * 64 + wv * 16;
  if (i0 >= Ip) return;
  const size_t s0 = (size_t)blockIdx.z * kc;
  const __bf16 *pa = dZT + (size_t)(o0 + r) * B + s0 + kq;
  const __bf16 *pb = HT + (size_t)(i0 + r) * B + s0 + kq;
  f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
  int k = 0;
  for (; k + 64 <= kc; k += 64) {
    acc0 = mfma16(*(const bf16x8 *)(pa + k), *(const bf16x8 *)(pb + k), acc0);
    acc1 = mfma16(*(const bf16x8 *)(pa + k + 32), *(const bf16x8 *)(pb + k + 32), acc1);
  }
  for (; k < kc; k += 32) acc0 = mfma16(*(const bf16x8 *)(pa + k), *(const bf16x8 *)(pb + k), acc0);
  float *dst = part + ((size_t)blockIdx.z * O + o0 + 4 * (lane >> 4)) * Ip + i0 + (lane & 15);
#pragma unroll
  for (int i = 0; i < 4; i++) dst[(size_t)i * Ip] = acc0[i] + acc1[i];
}

struct GradParts {
  const float *part[5];
  int O[5], Ip[5], ones[5];
  int splits;
};

__device__ __forceinline__ int layer_of(int p) {
  int l = 0;
  while (l < 4 && p >= mlp_woff(l + 1)) l++;
  return l;
}

// grad[p] = sum over splits of the partial sums (weights, and the ones column for biases)
__global__ void k_mlp_reduce(GradParts g, float *grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_NPARAM) return;
  const int l = layer_of(p);
  int o, i;
  if (p < mlp_boff(l)) { o = (p - mlp_woff(l)) / mlp_in(l); i = (p - mlp_woff(l)) % mlp_in(l); }
  else { o = p - mlp_boff(l); i = g.ones[l]; }
  float s = 0.f;
  for (int sp = 0; sp < g.splits; sp++) s += g.part[l][((size_t)sp * g.O[l] + o) * g.Ip[l] + i];
  grad[p] = s;
}

struct NetPtrs {
  __bf16 *W[5], *WT[5], *WF[4];
  float *bias[5];
};

__device__ __forceinline__ void write_copies(const NetPtrs &n, int p, float v) {
  const int l = layer_of(p);
  if (p < mlp_boff(l)) {
    const int o = (p - mlp_woff(l)) / mlp_in(l), i = (p - mlp_woff(l)) % mlp_in(l);
    const int kpad = l == 0 ? MLP_K0 : MLP_H;
    n.W[l][(size_t)o * kpad + i] = (__bf16)v;
    if (l >= 1) n.WT[l][(size_t)i * (l == 4 ? 32 : MLP_H) + o] = (__bf16)v;
    if (l <= 3) {   // A-fragment order (MlpWeights::WF)
      const int lane = o % 16 + 16 * ((i % 32) / 8);
      n.WF[l][((size_t)((o / 16) * (kpad / 32) + i / 32) * 64 + lane) * 8 + i % 8] = (__bf16)v;
    }
  } else {
    n.bias[l][p - mlp_boff(l)] = v;
  }
}

__global__ void k_mlp_refresh(NetPtrs n, const float *master) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < MLP_NPARAM) write_copies(n, p, master[p]);
}

// Adam (Kingma & Ba 2015) with bias correction, then the bf16 / transposed copies
__global__ void k_mlp_adam(NetPtrs n, float *master, float *m, float *v, const float *grad, float lr, float b1,
                           float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_NPARAM) return;
  const float g = grad[p];
  const float mm = b1 * m[p] + (1.f - b1) * g;
  const float vv = b2 * v[p] + (1.f - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  const float w = master[p] - lr * (mm * c1) / (sqrtf(vv * c2) + eps);
  master[p] = w;
  write_copies(n, p, w);
}

// ---------------------------------------------------------------- training samples

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {   // splitmix64 finaliser
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double u01(unsigned long long seed, long long s, int j) {
  return (double)(mix64(seed ^ mix64((unsigned long long)s * 64ull + (unsigned long long)j)) >> 11) * 0x1.0p-53;
}

__device__ __forceinline__ void sg_eos(DevEos &m, double gam, double pinf) {
  memset(&m, 0, sizeof(m));
  m.type = EOS_SG;
  m.rhomin = 0.0;
  m.pmin = -1.7976931348623157e308;
  m.rhomax = m.pmax = 1.7976931348623157e308;
  m.gam = gam;
  m.pstiff = pinf;
  m.gam1 = gam - 1.0;
  m.gam_pstiff = gam * pinf;
}

// the 12 raw doubles of a sample: rhoL uL pL gamL pinfL rhoR uR pR gamR pinfR | p* u*  (then rho*L rho*R)
constexpr int RAW = 14;

struct SampleArgs {
  long long n;
  unsigned long long seed;
  float *X, *T;
  double *raw;               // [n][RAW] or NULL
  RiemannCfg cfg;
  double *scratch;
  int cap, nthreads;
  unsigned long long *next;  // work counter (zero at launch): waves claim 64 samples at a time
  long long s0;              // index of the first sample (the random stream is per global index)
  const int *order;          // the chunk's samples in solve order (NULL: index order)
};

// Sample s of the stream: sides are a gas (gamma 1.1-1.7, pinf 0) or a liquid (gamma 4-7.5, pinf
// 1e8-1e9) with even odds; p_L log-uniform in 1e4..1e9, p_R/p_L log-uniform in 1e-4..1e4, rho_L
// log-uniform in 0.1..1e3, rho_R/rho_L log-uniform in 1e-3..1e3, u_R - u_L uniform in +-(c_L + c_R).
__device__ __forceinline__ void sg_draw(unsigned long long sd, long long sg, SgSide &L, SgSide &R) {
  const bool liqL = u01(sd, sg, 0) < 0.5, liqR = u01(sd, sg, 1) < 0.5;
  L.gam = liqL ? 4.0 + 3.5 * u01(sd, sg, 2) : 1.1 + 0.6 * u01(sd, sg, 2);
  R.gam = liqR ? 4.0 + 3.5 * u01(sd, sg, 3) : 1.1 + 0.6 * u01(sd, sg, 3);
  L.pinf = liqL ? pow(10.0, 8.0 + u01(sd, sg, 4)) : 0.0;
  R.pinf = liqR ? pow(10.0, 8.0 + u01(sd, sg, 5)) : 0.0;
  L.p = pow(10.0, 4.0 + 5.0 * u01(sd, sg, 6));
  R.p = L.p * pow(10.0, 4.0 * (2.0 * u01(sd, sg, 7) - 1.0));
  L.rho = pow(10.0, -1.0 + 4.0 * u01(sd, sg, 8));
  R.rho = L.rho * pow(10.0, 3.0 * (2.0 * u01(sd, sg, 9) - 1.0));
  L.u = 0.0; R.u = 0.0;
  sg_side_init(L);
  sg_side_init(R);
  const double cs = L.c + R.c;
  L.u = cs * (u01(sd, sg, 11) - 0.5);
  R.u = L.u + cs * (2.0 * u01(sd, sg, 10) - 1.0);
}

// Solve-order key of a sample (ascending = solved first). Exact solves differ in cost by three
// orders of magnitude, and a wave of 64 lanes takes as long as its slowest lane, so the samples are
// solved in the order of a predicted cost: lanes of one wave then finish close together. The slow
// solves are near-vacuum expansions, where the two rarefaction curves (Toro 4.7 in effective
// pressure) have not met at a pressure of 1e-6 sqrt(p_L p_R): their key is in [-2, -1), lower for
// a gas side holding a smaller share of c_L + c_R, which measured slower. Every other sample keys on
// how far the curves are from meeting there (>= -1). Only the order depends on the key; each
// sample's result does not.
__global__ void k_sg_keys(long long n, unsigned long long seed, long long s0, float *key, int *idx) {
  const long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (s >= n) return;
  SgSide L, R;
  sg_draw(seed, s0 + s, L, R);
  const double cs = L.c + R.c;
  const double pf = 1e-6 * sqrt(L.p * R.p);
  auto rare = [&](const SgSide &q) {
    return 2.0 * q.c / (q.gam - 1.0) * (pow((pf + q.pinf) / q.P, (q.gam - 1.0) / (2.0 * q.gam)) - 1.0);
  };
  const double gap = -(rare(L) + rare(R) + (R.u - L.u)) / cs;
  double k;
  if (gap < 0.0) {
    const double gshare = L.pinf == 0.0 ? L.c / cs : (R.pinf == 0.0 ? R.c / cs : 1.0);
    k = -2.0 + 0.999 * gshare;
  } else {
    k = fmin(gap, 1e30) - 1.0;
  }
  key[s] = (float)k;
  idx[s] = (int)s;
}

// Waves claim 64 samples at a time, in the chunk's solve order. Targets are NaN where the exact
// solver fails or the star state is inadmissible.
__global__ __launch_bounds__(64) void k_sg_samples(SampleArgs a) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= a.nthreads) return;
  DevEos mats[2];
  RSolver S_;
  S_.mat = mats;
  S_.cfg = a.cfg;
  S_.overflow = 0;
  S_.path1.base = a.scratch + tid; S_.path1.cap = a.cap; S_.path1.stride = a.nthreads; S_.path1.n = 0;
  S_.path3.base = a.scratch + (long long)3 * a.cap * a.nthreads + tid; S_.path3.cap = a.cap;
  S_.path3.stride = a.nthreads; S_.path3.n = 0;
  const int lane = threadIdx.x & 63;
  for (;;) {
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(a.next, 64ull);
    base = __shfl(base, 0, 64);
    if (base >= (unsigned long long)a.n) break;
    const long long q = (long long)base + lane;
    if (q < a.n) {   // the claim above is reached by the whole wave
      const long long s = a.order ? (long long)a.order[q] : q;
      const unsigned long long sd = a.seed;
      const long long sg = a.s0 + s;   // global sample index: the random draws
      SgSide L, R;
      sg_draw(sd, sg, L, R);
      sg_eos(mats[0], L.gam, L.pinf);
      sg_eos(mats[1], R.gam, R.pinf);
      const double Vm[5] = {L.rho, L.u, 0.0, 0.0, L.p}, Vp[5] = {R.rho, R.u, 0.0, 0.0, R.p};
      const double nrm[3] = {1.0, 0.0, 0.0};
      double Vs[5], Vsm[5], Vsp[5];
      int mid = 0;
      S_.overflow = 0;
      const int rc = riemann_solve(S_, nrm, Vm, 0, Vp, 1, Vs, &mid, Vsm, Vsp);
      float x[MLP_IN];
      SgScales sc;
      sg_features(L, R, x, &sc);
      for (int j = 0; j < MLP_IN; j++) a.X[s * MLP_IN + j] = x[j];
      const double ps = Vsm[4], us = Vsm[1];
      // near-vacuum star states (p* + min pinf < 1e-6 sqrt(P_L P_R)) are left to the exact solver
      const bool ok = rc == 0 && !S_.overflow && ps + sc.pimin > 1e-6 * sc.Ps && Vsm[0] > 0.0 && Vsp[0] > 0.0 &&
                      __builtin_isfinite(ps) && __builtin_isfinite(us);
      const float nanf_ = __builtin_nanf("");
      a.T[s * MLP_OUT + 0] = ok ? (float)(0.25 * log((ps + sc.pimin) / sc.Ps)) : nanf_;
      a.T[s * MLP_OUT + 1] = ok ? (float)((us - sc.ubar) / sc.cs) : nanf_;
      a.T[s * MLP_OUT + 2] = ok ? (float)(0.25 * log(Vsm[0] / L.rho)) : nanf_;
      if (a.raw) {
        double *o = a.raw + s * RAW;
        o[0] = L.rho; o[1] = L.u; o[2] = L.p; o[3] = L.gam; o[4] = L.pinf;
        o[5] = R.rho; o[6] = R.u; o[7] = R.p; o[8] = R.gam; o[9] = R.pinf;
        const double qnan = __builtin_nan("");
        o[10] = ok ? ps : qnan; o[11] = ok ? us : qnan; o[12] = ok ? Vsm[0] : qnan; o[13] = ok ? Vsp[0] : qnan;
      }
    }
  }
}

// Acceptance of the network's p* (decoded from Y) and of the acoustic estimate after at most nmax
// Newton updates; also the mean number of updates each needs (capped at 50).
// cnt: [0] valid samples, [1] accepted (network), [2] accepted (acoustic), [3] updates (network),
// [4] updates (acoustic)
__global__ void k_sg_accept(long long n, const double *raw, const float *Y, int nmax, double tol,
                            unsigned long long *cnt) {
  unsigned long long c[5] = {0, 0, 0, 0, 0};
  for (long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x; s < n; s += (long long)gridDim.x * blockDim.x) {
    const double *o = raw + s * RAW;
    if (!(o[10] == o[10])) continue;
    SgSide L, R;
    L.rho = o[0]; L.u = o[1]; L.p = o[2]; L.gam = o[3]; L.pinf = o[4];
    R.rho = o[5]; R.u = o[6]; R.p = o[7]; R.gam = o[8]; R.pinf = o[9];
    if (!sg_side_init(L) || !sg_side_init(R)) continue;
    float x[MLP_IN];
    SgScales sc;
    sg_features(L, R, x, &sc);
    c[0]++;
    const double p_net = sg_decode_p(sc, Y[s * MLP_OUT]);
    const double p_ac = exp(4.0 * (double)x[9]) * sc.Ps - sc.pimin;
    double ps, res;
    int it;
    c[1] += sg_newton(L, R, p_net, nmax, tol, &ps, &res, &it);
    c[2] += sg_newton(L, R, p_ac, nmax, tol, &ps, &res, &it);
    sg_newton(L, R, p_net, 50, tol, &ps, &res, &it);
    c[3] += it;
    sg_newton(L, R, p_ac, 50, tol, &ps, &res, &it);
    c[4] += it;
  }
  for (int j = 0; j < 5; j++) {
    unsigned long long v = c[j];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(&cnt[j], v);
  }
}

__global__ void k_fill_rows(__bf16 *dst, int row, int B, float v) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < B) dst[(size_t)row * B + t] = (__bf16)v;
}

}  // namespace m2c

using namespace m2c;

// ---------------------------------------------------------------- host side

#define MCHK(x)                         \
  do {                                  \
    if ((x) != hipSuccess) return M2C_ERR_HIP; \
  } while (0)

struct m2c_mlp {
  int device;
  hipStream_t stream;
  MlpNet net;
  void *params_block;   // one allocation for all parameter buffers
  // training workspace (batch B)
  int B, splits, kc;
  void *ws_block;
  __bf16 *XT, *HT[5], *DT[5], *dZT[5], *dZT5;
  float *part[5], *loss;
};

static NetPtrs net_ptrs(const MlpNet &n) {
  NetPtrs p;
  for (int l = 0; l < 5; l++) { p.W[l] = n.W[l]; p.WT[l] = n.WT[l]; p.bias[l] = n.bias[l]; }
  for (int l = 0; l < 4; l++) p.WF[l] = n.WF[l];
  return p;
}

static unsigned long long host_mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static int mlp_alloc_params(m2c_mlp *m) {
  // sizes in bytes
  const size_t fp = align256(sizeof(float) * MLP_NPARAM);
  size_t wbytes[5], wtbytes[5], bbytes[5];
  for (int l = 0; l < 5; l++) {
    const int rows = l == 4 ? MLP_O4 : MLP_H, cols = l == 0 ? MLP_K0 : MLP_H;
    wbytes[l] = align256(sizeof(__bf16) * rows * cols);
    wtbytes[l] = l == 0 ? 0 : align256(sizeof(__bf16) * MLP_H * (l == 4 ? 32 : MLP_H));
    bbytes[l] = align256(sizeof(float) * (l == 4 ? MLP_O4 : MLP_H));
  }
  size_t total = 4 * fp;
  for (int l = 0; l < 5; l++) total += wbytes[l] + wtbytes[l] + bbytes[l];
  for (int l = 0; l < 4; l++) total += wbytes[l];   // the fragment-ordered copies
  char *d = nullptr;
  MCHK(hipMalloc(&d, total));
  MCHK(hipMemset(d, 0, total));
  m->params_block = d;
  m->net.master = (float *)d; d += fp;
  m->net.adam_m = (float *)d; d += fp;
  m->net.adam_v = (float *)d; d += fp;
  m->net.grad = (float *)d; d += fp;
  for (int l = 0; l < 5; l++) {
    m->net.W[l] = (__bf16 *)d; d += wbytes[l];
    m->net.WT[l] = l == 0 ? nullptr : (__bf16 *)d; d += wtbytes[l];
    m->net.bias[l] = (float *)d; d += bbytes[l];
  }
  for (int l = 0; l < 4; l++) { m->net.WF[l] = (__bf16 *)d; d += wbytes[l]; }
  m->net.adam_t = 0;
  return M2C_OK;
}

static int mlp_refresh(m2c_mlp *m) {
  hipLaunchKernelGGL(k_mlp_refresh, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, m->stream, net_ptrs(m->net),
                     m->net.master);
  MCHK(hipGetLastError());
  return M2C_OK;
}

static void mlp_free_ws(m2c_mlp *m) {
  if (m->ws_block) hipFree(m->ws_block);
  m->ws_block = nullptr;
  m->B = 0;
}

// training workspace for batch B (multiple of 64)
static int mlp_ensure_ws(m2c_mlp *m, int B) {
  if (m->B == B && m->ws_block) return M2C_OK;
  mlp_free_ws(m);
  int kc = B;   // split-K chunk: divides B, a multiple of 64, at most 2048 where B allows
  while (kc > 2048 && (kc / 2) % MLP_TM == 0) kc /= 2;
  const int splits = B / kc;
  const size_t hb = sizeof(__bf16) * (size_t)B;
  size_t total = align256(32 * hb) + align256(MLP_O4 * hb) + 256;
  for (int l = 1; l <= 4; l++) total += align256(MLP_HROWS * hb) + 2 * align256(MLP_H * hb);
  const int O[5] = {MLP_H, MLP_H, MLP_H, MLP_H, MLP_O4}, Ip[5] = {MLP_K0, MLP_HROWS, MLP_HROWS, MLP_HROWS, MLP_HROWS};
  for (int l = 0; l < 5; l++) total += align256(sizeof(float) * (size_t)splits * O[l] * Ip[l]);
  char *d = nullptr;
  MCHK(hipMalloc(&d, total));
  MCHK(hipMemset(d, 0, total));
  m->ws_block = d;
  m->XT = (__bf16 *)d; d += align256(32 * hb);
  m->dZT5 = (__bf16 *)d; d += align256(MLP_O4 * hb);
  m->loss = (float *)d; d += 256;
  m->HT[0] = m->DT[0] = m->dZT[0] = nullptr;
  for (int l = 1; l <= 4; l++) {
    m->HT[l] = (__bf16 *)d; d += align256(MLP_HROWS * hb);
    m->DT[l] = (__bf16 *)d; d += align256(MLP_H * hb);
    m->dZT[l] = (__bf16 *)d; d += align256(MLP_H * hb);
  }
  for (int l = 0; l < 5; l++) {
    m->part[l] = (float *)d;
    d += align256(sizeof(float) * (size_t)splits * O[l] * Ip[l]);
  }
  // rows of ones: input column MLP_ONES0 and row MLP_H of every hidden activation
  hipLaunchKernelGGL(k_fill_rows, dim3((B + 255) / 256), dim3(256), 0, m->stream, m->XT, MLP_ONES0, B, 1.0f);
  for (int l = 1; l <= 4; l++)
    hipLaunchKernelGGL(k_fill_rows, dim3((B + 255) / 256), dim3(256), 0, m->stream, m->HT[l], MLP_H, B, 1.0f);
  MCHK(hipGetLastError());
  m->B = B;
  m->kc = kc;
  m->splits = splits;
  return M2C_OK;
}

static FwdArgs train_args(m2c_mlp *m, const float *X, const float *T, unsigned long long pa, unsigned long long pc,
                          unsigned long long pmask, long long base, float gscale, float *Y) {
  FwdArgs a;
  memset(&a, 0, sizeof(a));
  a.X = X; a.Y = Y; a.n = m->B; a.n_dev = nullptr; a.T = T;
  a.pa = pa; a.pc = pc; a.pmask = pmask; a.base = base; a.B = m->B;
  a.XT = m->XT;
  for (int l = 0; l < 5; l++) { a.HT[l] = m->HT[l]; a.DT[l] = m->DT[l]; }
  a.dZT5 = m->dZT5;
  a.gscale = gscale;
  a.loss = m->loss;
  return a;
}

// forward + backward + weight gradients of one batch into net.grad
static int mlp_grad_step(m2c_mlp *m, const FwdArgs &fa) {
  const MlpWeights w = m->net.view();
  const int tiles = m->B / MLP_TM;
  hipLaunchKernelGGL((k_mlp_fwd<1>), dim3(tiles), dim3(256), 0, m->stream, w, fa);
  BwdArgs ba;
  ba.B = m->B;
  ba.dZT5 = m->dZT5;
  for (int l = 0; l < 5; l++) { ba.DT[l] = m->DT[l]; ba.dZT[l] = m->dZT[l]; }
  hipLaunchKernelGGL(k_mlp_bwd, dim3(tiles), dim3(256), 0, m->stream, w, ba);
  const __bf16 *A[5] = {m->dZT[1], m->dZT[2], m->dZT[3], m->dZT[4], m->dZT5};
  const __bf16 *H[5] = {m->XT, m->HT[1], m->HT[2], m->HT[3], m->HT[4]};
  const int O[5] = {MLP_H, MLP_H, MLP_H, MLP_H, MLP_O4}, Ip[5] = {MLP_K0, MLP_HROWS, MLP_HROWS, MLP_HROWS, MLP_HROWS};
  GradParts g;
  for (int l = 0; l < 5; l++) {
    hipLaunchKernelGGL(k_mlp_dw, dim3((Ip[l] + 63) / 64, O[l] / 16, m->splits), dim3(256), 0, m->stream, A[l], H[l],
                       O[l], Ip[l], m->B, m->kc, m->part[l]);
    g.part[l] = m->part[l];
    g.O[l] = O[l];
    g.Ip[l] = Ip[l];
    g.ones[l] = l == 0 ? MLP_ONES0 : MLP_H;
  }
  g.splits = m->splits;
  hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, m->stream, g, m->net.grad);
  MCHK(hipGetLastError());
  return M2C_OK;
}

extern "C" {

int m2c_mlp_create(int device, unsigned long long seed, m2c_mlp **out) {
  if (!out) return M2C_ERR_ARG;
  *out = nullptr;
  MCHK(hipSetDevice(device));
  m2c_mlp *m = new m2c_mlp();
  memset(m, 0, sizeof(*m));
  m->device = device;
  if (hipStreamCreate(&m->stream) != hipSuccess) { delete m; return M2C_ERR_HIP; }
  if (int rc = mlp_alloc_params(m)) { hipStreamDestroy(m->stream); delete m; return rc; }
  // U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases, from splitmix64(seed, index)
  std::vector<float> p(MLP_NPARAM);
  for (int l = 0; l < 5; l++) {
    const double a = 1.0 / std::sqrt((double)mlp_in(l));
    const int end = mlp_boff(l) + mlp_out(l);
    for (int q = mlp_woff(l); q < end; q++) {
      const double u = (double)(host_mix64(seed ^ host_mix64((unsigned long long)q)) >> 11) * 0x1.0p-53;
      p[q] = (float)((2.0 * u - 1.0) * a);
    }
  }
  if (hipMemcpy(m->net.master, p.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice) != hipSuccess ||
      mlp_refresh(m) || hipStreamSynchronize(m->stream) != hipSuccess) {
    hipFree(m->params_block);
    hipStreamDestroy(m->stream);
    delete m;
    return M2C_ERR_HIP;
  }
  *out = m;
  return M2C_OK;
}

int m2c_mlp_destroy(m2c_mlp *m) {
  if (!m) return M2C_OK;
  hipSetDevice(m->device);
  mlp_free_ws(m);
  hipFree(m->params_block);
  hipStreamDestroy(m->stream);
  delete m;
  return M2C_OK;
}

int m2c_mlp_nparams(void) { return MLP_NPARAM; }

int m2c_mlp_get_params(m2c_mlp *m, float *p) {
  if (!m || !p) return M2C_ERR_ARG;
  MCHK(hipSetDevice(m->device));
  MCHK(hipStreamSynchronize(m->stream));
  MCHK(hipMemcpy(p, m->net.master, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost));
  return M2C_OK;
}

int m2c_mlp_set_params(m2c_mlp *m, const float *p) {
  if (!m || !p) return M2C_ERR_ARG;
  MCHK(hipSetDevice(m->device));
  MCHK(hipStreamSynchronize(m->stream));
  MCHK(hipMemcpy(m->net.master, p, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
  MCHK(hipMemset(m->net.adam_m, 0, sizeof(float) * MLP_NPARAM));
  MCHK(hipMemset(m->net.adam_v, 0, sizeof(float) * MLP_NPARAM));
  m->net.adam_t = 0;
  if (int rc = mlp_refresh(m)) return rc;
  MCHK(hipStreamSynchronize(m->stream));
  return M2C_OK;
}

int m2c_mlp_forward(m2c_mlp *m, int n, const float *X, float *Y) {
  if (!m || n < 0 || (n > 0 && (!X || !Y))) return M2C_ERR_ARG;
  if (n == 0) return M2C_OK;
  MCHK(hipSetDevice(m->device));
  float *d = nullptr;
  MCHK(hipMalloc(&d, sizeof(float) * (size_t)n * (MLP_IN + MLP_OUT)));
  float *dY = d + (size_t)n * MLP_IN;
  int rc = M2C_OK;
  if (hipMemcpyAsync(d, X, sizeof(float) * (size_t)n * MLP_IN, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
      m2c_mlp_launch_infer(m->net.view(), m->stream, d, dY, n, nullptr) ||
      hipMemcpyAsync(Y, dY, sizeof(float) * (size_t)n * MLP_OUT, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
      hipStreamSynchronize(m->stream) != hipSuccess)
    rc = M2C_ERR_HIP;
  hipFree(d);
  return rc;
}

// loss = sum over rows and outputs of (y - t)^2 / n (rows with NaN targets contribute nothing) and its
// gradient with respect to the parameters (flat layout); n a multiple of 64
int m2c_mlp_gradient(m2c_mlp *m, int n, const float *X, const float *T, double *loss, float *grad) {
  if (!m || n <= 0 || n % MLP_TM || !X || !T || !loss || !grad) return M2C_ERR_ARG;
  MCHK(hipSetDevice(m->device));
  if (int rc = mlp_ensure_ws(m, n)) return rc;
  float *d = nullptr;
  MCHK(hipMalloc(&d, sizeof(float) * (size_t)n * (MLP_IN + MLP_OUT)));
  float *dT = d + (size_t)n * MLP_IN;
  int rc = M2C_OK;
  float hl = 0.f;
  if (hipMemcpyAsync(d, X, sizeof(float) * (size_t)n * MLP_IN, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
      hipMemcpyAsync(dT, T, sizeof(float) * (size_t)n * MLP_OUT, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
      hipMemsetAsync(m->loss, 0, sizeof(float), m->stream) != hipSuccess)
    rc = M2C_ERR_HIP;
  if (!rc) {
    // identity gather: pa = 1, pc = 0, mask covers n
    unsigned long long mask = 1;
    while (mask < (unsigned long long)n) mask <<= 1;
    const FwdArgs fa = train_args(m, d, dT, 1ull, 0ull, mask - 1, 0, 2.0f / (float)n, nullptr);
    rc = mlp_grad_step(m, fa);
  }
  if (!rc && (hipMemcpyAsync(grad, m->net.grad, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, m->stream) !=
                  hipSuccess ||
              hipMemcpyAsync(&hl, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
              hipStreamSynchronize(m->stream) != hipSuccess))
    rc = M2C_ERR_HIP;
  hipFree(d);
  *loss = (double)hl / n;
  return rc;
}

static RiemannCfg default_rcfg() {
  RiemannCfg c;   // ExactRiemannSolverData defaults, IoData.cpp:1345-1357
  c.maxIts_main = 200; c.maxIts_bracket = 100; c.maxIts_shock = 200; c.numSteps_rarefaction = 200;
  c.tol_main = 1e-4; c.tol_shock = 1e-12; c.tol_rarefaction = 1e-8;
  c.min_pressure = -1e8; c.failure_threshold = 0.2; c.pressure_at_failure = 1e-8;
  return c;
}

// samples in launches of at most 2^19 (a progress line on stderr after each when verbose); each
// launch solves its samples in the order of k_sg_keys
static int gen_samples_dev(hipStream_t s, long long n, unsigned long long seed, float *X, float *T, double *raw,
                           int verbose = 0) {
  // one wave per SIMD at this kernel's register use: 4 SIMDs x 256 CUs x 64 lanes
  const int cap = 1024, nthreads = 65536;
  const long long chunk = 1ll << 19;
  const size_t sb = sizeof(double) * 6 * (size_t)cap * nthreads;
  size_t tb = 0;
  if (hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (const float *)nullptr, (float *)nullptr, (const int *)nullptr,
                                         (int *)nullptr, (int)chunk, 0, 32, s) != hipSuccess)
    return M2C_ERR_HIP;
  const size_t kb = align256(sizeof(float) * chunk), ib = align256(sizeof(int) * chunk);
  char *buf = nullptr;
  MCHK(hipMalloc(&buf, sb + 256 + 2 * kb + 2 * ib + align256(tb)));
  double *scratch = (double *)buf;
  unsigned long long *next = (unsigned long long *)(buf + sb);
  float *key = (float *)(buf + sb + 256), *key_sorted = (float *)((char *)key + kb);
  int *idx = (int *)((char *)key_sorted + kb), *order = (int *)((char *)idx + ib);
  void *tmp = (char *)order + ib;
  int rc = M2C_OK;
  auto t0 = std::chrono::steady_clock::now();
  for (long long s0 = 0; s0 < n && !rc; s0 += chunk) {
    const long long m = n - s0 < chunk ? n - s0 : chunk;
    if (hipMemsetAsync(next, 0, sizeof(unsigned long long), s) != hipSuccess) { rc = M2C_ERR_HIP; break; }
    hipLaunchKernelGGL(k_sg_keys, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, m, seed, s0, key, idx);
    if (hipGetLastError() != hipSuccess ||
        hipcub::DeviceRadixSort::SortPairs(tmp, tb, key, key_sorted, idx, order, (int)m, 0, 32, s) != hipSuccess) {
      rc = M2C_ERR_HIP;
      break;
    }
    SampleArgs a;
    a.n = m; a.seed = seed; a.X = X + (size_t)s0 * MLP_IN; a.T = T + (size_t)s0 * MLP_OUT;
    a.raw = raw ? raw + (size_t)s0 * RAW : nullptr; a.cfg = default_rcfg(); a.scratch = scratch;
    a.cap = cap; a.nthreads = nthreads; a.next = next; a.s0 = s0; a.order = order;
    hipLaunchKernelGGL(k_sg_samples, dim3(nthreads / 64), dim3(64), 0, s, a);
    rc = hipGetLastError() == hipSuccess && hipStreamSynchronize(s) == hipSuccess ? M2C_OK : M2C_ERR_HIP;
    if (verbose) {
      const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      fprintf(stderr, "m2c_mlp: %lld / %lld exact-solver samples, %.1f s\n", s0 + m, n, el);
      fflush(stderr);
    }
  }
  hipFree(buf);
  return rc;
}

// n training samples from the exact solver (X [n][10], T [n][3], raw [n][14] or NULL; host arrays)
int m2c_mlp_samples(m2c_mlp *m, long long n, unsigned long long seed, float *X, float *T, double *raw) {
  if (!m || n <= 0 || !X || !T) return M2C_ERR_ARG;
  MCHK(hipSetDevice(m->device));
  char *d = nullptr;
  const size_t bx = sizeof(float) * (size_t)n * MLP_IN, bt = sizeof(float) * (size_t)n * MLP_OUT,
               br = raw ? sizeof(double) * (size_t)n * RAW : 0;
  MCHK(hipMalloc(&d, align256(bx) + align256(bt) + br));
  float *dX = (float *)d, *dT = (float *)(d + align256(bx));
  double *dR = raw ? (double *)(d + align256(bx) + align256(bt)) : nullptr;
  int rc = gen_samples_dev(m->stream, n, seed, dX, dT, dR);
  if (!rc && (hipMemcpy(X, dX, bx, hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(T, dT, bt, hipMemcpyDeviceToHost) != hipSuccess ||
              (raw && hipMemcpy(raw, dR, br, hipMemcpyDeviceToHost) != hipSuccess)))
    rc = M2C_ERR_HIP;
  hipFree(d);
  return rc;
}

int m2c_mlp_train(m2c_mlp *m, const m2c_train_config *cfg, m2c_train_report *rep) {
  if (!m || !cfg || !rep) return M2C_ERR_ARG;
  const long long N = cfg->n_samples, NH = cfg->n_holdout;
  if (N < MLP_TM || (N & (N - 1)) || cfg->batch <= 0 || cfg->batch % MLP_TM || cfg->batch > N || NH < 0 ||
      NH % MLP_TM || cfg->epochs <= 0)
    return M2C_ERR_ARG;
  memset(rep, 0, sizeof(*rep));
  MCHK(hipSetDevice(m->device));
  if (int rc = mlp_ensure_ws(m, cfg->batch)) return rc;
  const size_t bx = align256(sizeof(float) * (size_t)(N + NH) * MLP_IN),
               bt = align256(sizeof(float) * (size_t)(N + NH) * MLP_OUT),
               br = align256(sizeof(double) * (size_t)(NH > 0 ? NH : 1) * RAW),
               by = align256(sizeof(float) * (size_t)(NH > 0 ? NH : 1) * MLP_OUT);
  char *d = nullptr;
  MCHK(hipMalloc(&d, bx + bt + br + by + 256));
  float *dX = (float *)d, *dT = (float *)(d + bx);
  double *dR = (double *)(d + bx + bt);
  float *dY = (float *)(d + bx + bt + br);
  unsigned long long *dcnt = (unsigned long long *)(d + bx + bt + br + by);
  int rc = M2C_OK;
  auto t0 = std::chrono::steady_clock::now();
  rc = gen_samples_dev(m->stream, N, cfg->seed, dX, dT, nullptr, cfg->verbose);
  if (!rc && NH > 0)
    rc = gen_samples_dev(m->stream, NH, cfg->seed ^ 0x5eed5eed5eedull, dX + (size_t)N * MLP_IN,
                         dT + (size_t)N * MLP_OUT, dR, cfg->verbose);
  auto t1 = std::chrono::steady_clock::now();
  rep->data_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  // valid training samples
  if (!rc) {
    std::vector<float> Th((size_t)N * MLP_OUT);
    if (hipMemcpy(Th.data(), dT, sizeof(float) * Th.size(), hipMemcpyDeviceToHost) != hipSuccess) rc = M2C_ERR_HIP;
    long long v = 0;
    for (long long s = 0; s < N; s++) v += Th[(size_t)s * MLP_OUT] == Th[(size_t)s * MLP_OUT];
    rep->valid_samples = v;
  }
  const int B = cfg->batch;
  const long long steps_per_epoch = N / B;
  const long long total_steps = steps_per_epoch * cfg->epochs;
  const float gscale = 2.0f / (float)B;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = M2C_ERR_HIP;
  const NetPtrs np = net_ptrs(m->net);
  long long step = 0;
  float step_ms_total = 0.f;
  for (int ep = 0; !rc && ep < cfg->epochs; ep++) {
    const unsigned long long pa = (host_mix64(cfg->seed + 977ull * (ep + 1)) | 1ull), pc = host_mix64(cfg->seed + ep);
    if (hipMemsetAsync(m->loss, 0, sizeof(float), m->stream) != hipSuccess) { rc = M2C_ERR_HIP; break; }
    if (hipEventRecord(e0, m->stream) != hipSuccess) { rc = M2C_ERR_HIP; break; }
    for (long long k = 0; k < steps_per_epoch && !rc; k++, step++) {
      const FwdArgs fa = train_args(m, dX, dT, pa, pc, (unsigned long long)(N - 1), k * B, gscale, nullptr);
      rc = mlp_grad_step(m, fa);
      const double frac = (double)step / (double)(total_steps > 1 ? total_steps - 1 : 1);
      const float lr = (float)(cfg->lr_min + 0.5 * (cfg->lr - cfg->lr_min) * (1.0 + std::cos(M_PI * frac)));
      m->net.adam_t++;
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)m->net.adam_t)));
      const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)m->net.adam_t)));
      hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, m->stream, np, m->net.master,
                         m->net.adam_m, m->net.adam_v, m->net.grad, lr, b1, b2, 1e-8f, c1, c2);
      if (hipGetLastError() != hipSuccess) rc = M2C_ERR_HIP;
    }
    float hl = 0.f, ms = 0.f;
    if (!rc && (hipEventRecord(e1, m->stream) != hipSuccess ||
                hipMemcpyAsync(&hl, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
                hipStreamSynchronize(m->stream) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      rc = M2C_ERR_HIP;
    step_ms_total += ms;
    const double el = (double)hl / (double)(steps_per_epoch * B);
    if (ep == 0) rep->first_epoch_loss = el;
    rep->last_epoch_loss = el;
    if (cfg->verbose) {
      fprintf(stderr, "m2c_mlp: epoch %d loss %.6f (%.1f ms)\n", ep + 1, el, ms);
      fflush(stderr);
    }
  }
  rep->steps = step;
  rep->train_ms = step_ms_total;
  rep->samples_per_s = step_ms_total > 0 ? (double)step * B / (step_ms_total * 1e-3) : 0.0;
  // held-out loss and acceptance
  if (!rc && NH > 0) {
    const float *hX = dX + (size_t)N * MLP_IN;
    rc = m2c_mlp_launch_infer(m->net.view(), m->stream, hX, dY, (int)NH, nullptr);
    if (!rc && (hipMemsetAsync(dcnt, 0, 5 * sizeof(unsigned long long), m->stream) != hipSuccess)) rc = M2C_ERR_HIP;
    if (!rc) {
      hipLaunchKernelGGL(k_sg_accept, dim3(256), dim3(256), 0, m->stream, NH, (const double *)dR, (const float *)dY,
                         cfg->newton_max, cfg->tol, dcnt);
      if (hipGetLastError() != hipSuccess) rc = M2C_ERR_HIP;
    }
    std::vector<float> Yh((size_t)NH * MLP_OUT), Th((size_t)NH * MLP_OUT);
    unsigned long long cnt[5] = {0, 0, 0, 0, 0};
    if (!rc && (hipMemcpy(Yh.data(), dY, sizeof(float) * Yh.size(), hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(Th.data(), dT + (size_t)N * MLP_OUT, sizeof(float) * Th.size(), hipMemcpyDeviceToHost) !=
                    hipSuccess ||
                hipMemcpy(cnt, dcnt, sizeof(cnt), hipMemcpyDeviceToHost) != hipSuccess))
      rc = M2C_ERR_HIP;
    double se = 0.0;
    long long nv = 0;
    for (long long s = 0; s < NH; s++) {
      if (!(Th[(size_t)s * MLP_OUT] == Th[(size_t)s * MLP_OUT])) continue;
      nv++;
      for (int j = 0; j < MLP_OUT; j++) {
        const double dd = (double)Yh[(size_t)s * MLP_OUT + j] - (double)Th[(size_t)s * MLP_OUT + j];
        se += dd * dd;
      }
    }
    rep->holdout_loss = nv ? se / nv : 0.0;
    if (cnt[0]) {
      rep->holdout_accept_mlp = (double)cnt[1] / cnt[0];
      rep->holdout_accept_acoustic = (double)cnt[2] / cnt[0];
      rep->holdout_iters_mlp = (double)cnt[3] / cnt[0];
      rep->holdout_iters_acoustic = (double)cnt[4] / cnt[0];
    }
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipFree(d);
  return rc;
}

// average duration of one inference launch over n device-resident rows (reps launches)
int m2c_mlp_bench_infer(m2c_mlp *m, int n, int reps, double *ms_per_launch) {
  if (!m || n <= 0 || reps <= 0 || !ms_per_launch) return M2C_ERR_ARG;
  MCHK(hipSetDevice(m->device));
  float *d = nullptr;
  MCHK(hipMalloc(&d, sizeof(float) * (size_t)n * (MLP_IN + MLP_OUT)));
  MCHK(hipMemsetAsync(d, 0, sizeof(float) * (size_t)n * MLP_IN, m->stream));
  hipEvent_t e0, e1;
  MCHK(hipEventCreate(&e0));
  MCHK(hipEventCreate(&e1));
  int rc = m2c_mlp_launch_infer(m->net.view(), m->stream, d, d + (size_t)n * MLP_IN, n, nullptr);
  float ms = 0.f;
  if (!rc && (hipEventRecord(e0, m->stream) != hipSuccess)) rc = M2C_ERR_HIP;
  for (int r = 0; r < reps && !rc; r++) rc = m2c_mlp_launch_infer(m->net.view(), m->stream, d, d + (size_t)n * MLP_IN, n, nullptr);
  if (!rc && (hipEventRecord(e1, m->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
              hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
    rc = M2C_ERR_HIP;
  *ms_per_launch = ms / reps;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(d);
  return rc;
}

__global__ void k_fill_synthetic(float *X, float *T, long long n, unsigned long long seed) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    for (int j = 0; j < MLP_IN; j++) X[t * MLP_IN + j] = (float)(2.0 * u01(seed, t, j) - 1.0);
    for (int j = 0; j < MLP_OUT; j++) T[t * MLP_OUT + j] = (float)(2.0 * u01(seed, t, 16 + j) - 1.0);
  }
}

// average duration (ms) of one training step (forward, backward, weight gradients, reduction, Adam)
// over `steps` steps with batch B on device-resident synthetic rows (the parameters change)
int m2c_mlp_bench_train(m2c_mlp *m, int B, int steps, double *ms_per_step) {
  if (!m || B <= 0 || B % MLP_TM || steps <= 0 || !ms_per_step) return M2C_ERR_ARG;
  MCHK(hipSetDevice(m->device));
  if (int rc = mlp_ensure_ws(m, B)) return rc;
  const long long N = (long long)B * 8;
  float *d = nullptr;
  MCHK(hipMalloc(&d, sizeof(float) * (size_t)N * (MLP_IN + MLP_OUT)));
  float *dT = d + (size_t)N * MLP_IN;
  hipLaunchKernelGGL(k_fill_synthetic, dim3(1024), dim3(256), 0, m->stream, d, dT, N, 7ull);
  const NetPtrs np = net_ptrs(m->net);
  hipEvent_t e0, e1;
  MCHK(hipEventCreate(&e0));
  MCHK(hipEventCreate(&e1));
  int rc = M2C_OK;
  float ms = 0.f;
  for (int k = -2; k < steps && !rc; k++) {   // two untimed steps first
    if (k == 0 && hipEventRecord(e0, m->stream) != hipSuccess) rc = M2C_ERR_HIP;
    const long long kk = k < 0 ? k + 2 : k;
    const FwdArgs fa = train_args(m, d, dT, 1ull, 0ull, (unsigned long long)(N - 1), (kk % 8) * B, 2.0f / B, nullptr);
    if (!rc) rc = mlp_grad_step(m, fa);
    m->net.adam_t++;
    const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)m->net.adam_t)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)m->net.adam_t)));
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, m->stream, np, m->net.master,
                       m->net.adam_m, m->net.adam_v, m->net.grad, 1e-4f, 0.9f, 0.999f, 1e-8f, c1, c2);
  }
  if (!rc && (hipEventRecord(e1, m->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
              hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
    rc = M2C_ERR_HIP;
  *ms_per_step = ms / steps;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(d);
  return rc;
}

int m2c_mlp_launch_infer(const MlpWeights &w, hipStream_t stream, const float *X, float *Y, int n_cap,
                         const unsigned *n_dev) {
  if (n_cap <= 0) return M2C_OK;
  FwdArgs a;
  memset(&a, 0, sizeof(a));
  a.X = X; a.Y = Y; a.n = n_cap; a.n_dev = n_dev;
  const int tiles = (n_cap + MLP_TI - 1) / MLP_TI;
  const int grid = tiles < 512 ? tiles : 512;   // two resident tiles on each of the 256 CUs
  hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(256), 0, stream, w, a);
  return hipGetLastError() == hipSuccess ? M2C_OK : M2C_ERR_HIP;
}

int m2c_mlp_get_weights(const m2c_mlp *m, MlpWeights *out, int *device) {
  if (!m || !out || !device) return M2C_ERR_ARG;
  *out = m->net.view();
  *device = m->device;
  return M2C_OK;
}

}  // extern "C"
