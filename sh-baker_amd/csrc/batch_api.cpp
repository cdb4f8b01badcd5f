// Batched bake, autotuner and path-replay adjoint entry points.
#include "shb_ctx.h"

extern "C" {

int shb_bake_batched(shb_ctx* const* ctxs, int32_t num_jobs, const shb_bake_cfg* cfgs,
                     float* const* sh_outs, float* const* vis_outs, shb_bake_stats* stats) {
  if (!ctxs || num_jobs <= 0 || !ctxs[0]) return SHB_ERR_INVALID_ARG;
  ctxs[0]->c.err = "shb_bake_batched: not available in this build yet";
  return SHB_ERR_UNSUPPORTED;
}

int shb_autotune(shb_ctx* ctx, const shb_bake_cfg* cfg, shb_tuning* best) {
  if (!ctx) return SHB_ERR_INVALID_ARG;
  ctx->c.err = "shb_autotune: not available in this build yet";
  return SHB_ERR_UNSUPPORTED;
}

int shb_bake_adjoint(shb_ctx* ctx, const shb_bake_cfg* cfg, const float* dloss_dsh,
                     float* dloss_dlight, float* dloss_dalbedo, shb_bake_stats* stats) {
  if (!ctx) return SHB_ERR_INVALID_ARG;
  ctx->c.err = "shb_bake_adjoint: not available in this build yet";
  return SHB_ERR_UNSUPPORTED;
}

}  // extern "C"
