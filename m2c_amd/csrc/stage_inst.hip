// stage_inst.hip -- one (flux, equation-of-state) variant of the fused stage kernel.
// Compiled with -DM2C_FLUX=<flux code> -DM2C_EK=<eos kind or -1> -DM2C_TAG=<name> and
// -DM2C_TC=<double|float> (arithmetic type of reconstruction and fluxes).
#include "stage_launch.hpp"

#define M2C_CAT2(a, b) a##b
#define M2C_CAT(a, b) M2C_CAT2(a, b)

using namespace m2c;

#ifndef M2C_TC
#define M2C_TC double
#endif
// M2C_MP=1: the material-interface code (interface fluxes, constant-reconstruction depth) also in a
// variant whose equation of state is fixed at compile time (two stiffened gases, say)
#ifndef M2C_MP
#define M2C_MP (M2C_EK < 0)
#endif

int M2C_CAT(m2c_stage_upload_, M2C_TAG)(const Grid &g, const DevEos *eos) {
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_grid), &g, sizeof(Grid)) != hipSuccess) return 2;
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_eos), eos, sizeof(DevEos) * 8) != hipSuccess) return 2;
  return 0;
}

int M2C_CAT(m2c_stage_launch_, M2C_TAG)(int in_cons, int mode, dim3 grid, hipStream_t s, const StageArgs &a) {
  dim3 blk(NTHR);
  if (in_cons == 0 && mode == 0) hipLaunchKernelGGL((stage_kernel<0, 0, M2C_FLUX, M2C_EK, M2C_TC, (M2C_MP != 0)>), grid, blk, 0, s, a);
  else if (in_cons == 0 && mode == 1) hipLaunchKernelGGL((stage_kernel<0, 1, M2C_FLUX, M2C_EK, M2C_TC, (M2C_MP != 0)>), grid, blk, 0, s, a);
  else if (in_cons == 0 && mode == 2) hipLaunchKernelGGL((stage_kernel<0, 2, M2C_FLUX, M2C_EK, M2C_TC, (M2C_MP != 0)>), grid, blk, 0, s, a);
  else if (in_cons == 1 && mode == 1) hipLaunchKernelGGL((stage_kernel<1, 1, M2C_FLUX, M2C_EK, M2C_TC, (M2C_MP != 0)>), grid, blk, 0, s, a);
  else if (in_cons == 1 && mode == 2) hipLaunchKernelGGL((stage_kernel<1, 2, M2C_FLUX, M2C_EK, M2C_TC, (M2C_MP != 0)>), grid, blk, 0, s, a);
  else return 1;
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
