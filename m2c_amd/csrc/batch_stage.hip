// batch_stage.hip -- one variant of the batched persistent stage kernel, in both tile shapes.
// Compiled with -DM2C_FLUX=<flux code> -DM2C_EK=<eos kind or -1> -DM2C_TC=<double|float>
// -DM2C_MP=<0|1> (material-interface code) -DM2C_TAG=<name>.
#define M2C_BATCHED
#include "batch_kern.hpp"

#define M2C_CAT2(a, b) a##b
#define M2C_CAT(a, b) M2C_CAT2(a, b)

using namespace m2c;

extern "C" int M2C_CAT(m2c_bstage_upload_, M2C_TAG)(const DevEos *eos) {
  return hipMemcpyToSymbol(HIP_SYMBOL(c_eos), eos, sizeof(DevEos) * 8) == hipSuccess ? 0 : 2;
}

template <class TS>
static void launch(int in_cons, int mode, dim3 gr, hipStream_t s, const BatchInst *insts, int n, TileQueue *q,
                   int stage) {
  constexpr bool MP = M2C_MP != 0;
  const dim3 bl(NTHR);
  if (in_cons == 0 && mode == 1)
    hipLaunchKernelGGL((k_batch_stage<0, 1, M2C_FLUX, M2C_EK, M2C_TC, MP, TS>), gr, bl, 0, s, insts, n, q, stage);
  else if (in_cons == 1 && mode == 1)
    hipLaunchKernelGGL((k_batch_stage<1, 1, M2C_FLUX, M2C_EK, M2C_TC, MP, TS>), gr, bl, 0, s, insts, n, q, stage);
  else
    hipLaunchKernelGGL((k_batch_stage<1, 2, M2C_FLUX, M2C_EK, M2C_TC, MP, TS>), gr, bl, 0, s, insts, n, q, stage);
}

extern "C" int M2C_CAT(m2c_bstage_launch_, M2C_TAG)(int shape, int in_cons, int mode, int nblocks, hipStream_t s,
                                                    const BatchInst *insts, int n, TileQueue *q, int stage) {
  if (nblocks < 1 || !(mode == 1 || (mode == 2 && in_cons == 1))) return 1;
  const dim3 gr(nblocks);
  if (shape == 0) launch<BatchTile0>(in_cons, mode, gr, s, insts, n, q, stage);
  else launch<BatchTile1>(in_cons, mode, gr, s, insts, n, q, stage);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
