// batch_ctx.hpp -- internal interface between the batched runner (batch.hip) and the
// single-instance context (m2c_hip.hip) for two-material instances of a batch.
//
// A two-material instance keeps a full context (its state, level set, interface-face lists,
// Riemann solutions and end-of-step buffers). Attached to a batch, the context's kernels run on
// the batch's stream and read the clock and status the batch owns; its fluid stages run inside the
// batch's persistent stage launches (the StageArgs below), and the batch calls the interface
// Riemann pass, the level-set stage and the end-of-step tasks of every such instance around them,
// after binding the instance's interface and level-set constants with asynchronous copies.
#pragma once
#include <hip/hip_runtime.h>

#include "fv_kernels.hpp"
#include "multiphase.hpp"

struct m2c_ctx;

extern "C" {
int m2c_ctx_batch_attach(m2c_ctx *c, hipStream_t s, m2c::Clock *clk, m2c::Status *st);
// the fluid stage s as the context would launch it, and the fields the batch's CFL pass reads
int m2c_ctx_batch_stage_args(m2c_ctx *c, int s, m2c::StageArgs *a, const double **V, const int8_t **ID,
                             int *variant);
int m2c_ctx_batch_bind(m2c_ctx *c);
int m2c_ctx_batch_before(m2c_ctx *c, int s);
// the instance's part of a multi-instance interface pass: *ok = 0 when its pass cannot join one
// (the learned solver, the Godunov flux); *flux its flux
int m2c_ctx_batch_iface_inst(m2c_ctx *c, m2c::IfaceInst *out, int *ok, int *flux);
// the step's start of the instance's interface and level-set counters (its first stage)
int m2c_ctx_batch_step_begin(m2c_ctx *c);
int m2c_ctx_batch_after(m2c_ctx *c, int s);
int m2c_ctx_batch_end_head(m2c_ctx *c, long long time_step);
int m2c_ctx_batch_end_retries(m2c_ctx *c, long long time_step);
int m2c_ctx_batch_end_tail(m2c_ctx *c);
// restore = 0: copy the state, IDs, level set and interface counters aside; 1: copy them back
int m2c_ctx_batch_backup(m2c_ctx *c, int restore);
}
