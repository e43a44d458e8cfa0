// stage_launch.hpp -- host-side entry points of the separately compiled stage-kernel variants.
//
// stage_inst.hip is compiled once per (flux, equation-of-state) variant with M2C_FLUX/M2C_EK/
// M2C_TAG defined; each object exports a launcher and an upload function for its own copy of
// the constant tables.
#pragma once
#include <hip/hip_runtime.h>
#include "fv_kernels.hpp"

#define M2C_DECLARE_VARIANT(TAG)                                                                   \
  int m2c_stage_launch_##TAG(int in_cons, int mode, dim3 grid, hipStream_t s, const m2c::StageArgs &a); \
  int m2c_stage_upload_##TAG(const m2c::Grid &g, const m2c::DevEos *eos);

M2C_DECLARE_VARIANT(hllc_sg)
M2C_DECLARE_VARIANT(hllc_gen)
M2C_DECLARE_VARIANT(roe_gen)
M2C_DECLARE_VARIANT(llf_gen)
M2C_DECLARE_VARIANT(hllc_sg_f32)
M2C_DECLARE_VARIANT(hllc_gen_f32)
M2C_DECLARE_VARIANT(roe_gen_f32)
M2C_DECLARE_VARIANT(llf_gen_f32)
M2C_DECLARE_VARIANT(god_gen)
M2C_DECLARE_VARIANT(hllc_sg_mp)
M2C_DECLARE_VARIANT(hllc_sg_mp_f32)
