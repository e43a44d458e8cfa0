// mlp.hpp -- the learned interface solver: a 10-256-256-256-256-3 SiLU network evaluated and
// trained with bf16 MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulation) on gfx950.
//
// Parameters (FP32 master copy, flat): W0 [256][10], b0 [256], W1..W3 [256][256] with b1..b3,
// W4 [3][256], b4 [3] -- 200 963 values, each W row-major [out][in].
#pragma once
#include <hip/hip_runtime.h>

#include "sg_star.hpp"

namespace m2c {

constexpr int MLP_K0 = 32;      // layer-0 inputs padded to one MFMA K step
constexpr int MLP_O4 = 16;      // output rows padded to one MFMA tile
constexpr int MLP_TM = 64;      // samples per workgroup tile
constexpr int MLP_ONES0 = 10;   // layer-0 input column that holds 1.0 in the training copy (bias gradient)
constexpr int MLP_HROWS = MLP_H + 16;   // hidden activations stored with a row of ones (bias gradient)

// offsets of W_l and b_l in the flat parameter vector
__host__ __device__ constexpr int mlp_in(int l) { return l == 0 ? MLP_IN : MLP_H; }
__host__ __device__ constexpr int mlp_out(int l) { return l == 4 ? MLP_OUT : MLP_H; }
__host__ __device__ constexpr int mlp_woff(int l) {
  return l == 0 ? 0 : mlp_woff(l - 1) + mlp_out(l - 1) * mlp_in(l - 1) + mlp_out(l - 1);
}
__host__ __device__ constexpr int mlp_boff(int l) { return mlp_woff(l) + mlp_out(l) * mlp_in(l); }
static_assert(mlp_boff(4) + MLP_OUT == MLP_NPARAM, "parameter layout");

// device copies of the weights used by the kernels
struct MlpWeights {
  const __bf16 *W[5];    // W[0] [256][32], W[1..3] [256][256], W[4] [16][256] (zero padding)
  const __bf16 *WT[5];   // WT[1..3] [256][256], WT[4] [256][32]: transposes for the backward pass
  // WF[0..3]: W[l] in MFMA A-fragment order for the inference kernel: the fragment of feature tile
  // ft (rows 16 ft .. +15) and K step ks (inputs 32 ks .. +31) is 64 lanes x 8 bf16 at
  // ((ft * K/32 + ks) * 64 + lane) * 8, lane = row % 16 + 16 * (input % 32 / 8)
  const __bf16 *WF[4];
  const float *bias[5];          // bias[4] padded to 16
};

struct MlpNet {
  float *master, *adam_m, *adam_v, *grad;   // flat FP32, MLP_NPARAM each
  __bf16 *W[5], *WT[5], *WF[4];
  float *bias[5];
  long long adam_t;
  MlpWeights view() const {
    MlpWeights w;
    for (int l = 0; l < 5; l++) { w.W[l] = W[l]; w.WT[l] = WT[l]; w.bias[l] = bias[l]; }
    for (int l = 0; l < 4; l++) w.WF[l] = WF[l];
    return w;
  }
};

}  // namespace m2c

extern "C" {
// Network inference Y[n][3] = net(X[n][10]) on `stream`; with n_dev != NULL the row count is read on
// the device (at most n_cap rows).
int m2c_mlp_launch_infer(const m2c::MlpWeights &w, hipStream_t stream, const float *X, float *Y, int n_cap,
                         const unsigned *n_dev);
struct m2c_mlp;
// the device weights and device ordinal of a network (for the interface kernels of a context)
int m2c_mlp_get_weights(const struct m2c_mlp *m, m2c::MlpWeights *out, int *device);
}
