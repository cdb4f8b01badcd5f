// The wavefront bake driver shared by shb_bake, shb_bake_lightmap,
// shb_bake_batched and the autotuner (not part of the ABI).
#ifndef SHB_BAKE_RUN_H_
#define SHB_BAKE_RUN_H_

#include <vector>

#include "api_util.h"

namespace shb {

// One lightmap of a bake: its context holds the scene, the surface and the
// outputs (d_sh, d_vis); prepare_job fills its valid-texel list.
struct JobIn {
  Ctx* c = nullptr;
  shb_bake_cfg cfg{};
  uint32_t stream = 0;  // random-stream id (0 for the plain entry points)
  uint32_t valid = 0;   // valid texels (set by prepare_job)
  // the job's texels: entries [list_offset, list_offset + valid) of its
  // context's texel list (a tile of a tiled bake; 0 for a whole surface)
  uint32_t list_offset = 0;
};

// Validates job j's context and config, zeroes its outputs and builds its
// valid-texel list on stream `st` (the job context's own stream when null);
// the valid-texel count lands in *count_dev. Nothing waits: the caller
// synchronises `st` and copies the counts. Errors go to the job's context.
int prepare_job_async(JobIn& j, hipStream_t st, uint32_t* count_dev);

// prepare_job_async on the job's own stream, then waits and reads j.valid.
int prepare_job(JobIn& j);

// Runs the bake of every job on x's stream: one set of wavefront batches
// whose path slots may mix jobs (jobs of several scenes traverse their
// concatenated BVHs; jobs that share a context share its BVH). With
// `graph`, the launch sequence is captured once in a hipGraph and replayed
// while the plan and buffers are unchanged. Errors are reported on x.
// Path-replay adjoint inputs: dL/dSH on the device (27 floats per
// supersampled texel) and the zeroed FP64 gradient sums (3 per light, then 3
// per material).
struct AdjIn {
  const float* dsh = nullptr;
  double* grad = nullptr;
  uint32_t nlights = 0;
  uint32_t nparam = 0;
};

// With `adj` (one job only), the replay adjoint runs instead of the bake:
// the same paths from the same counters, gradients instead of SH sums.
int run_bake(Ctx& x, std::vector<JobIn>& jobs, bool graph, BakeRun* run,
             const AdjIn* adj = nullptr);

}  // namespace shb

#endif  // SHB_BAKE_RUN_H_
