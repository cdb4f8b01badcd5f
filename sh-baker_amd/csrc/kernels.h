// Kernel argument blocks and launchers shared by the host API and the HIP
// kernels of the baker.
#ifndef SHB_KERNELS_H_
#define SHB_KERNELS_H_

#include "../../include/shb.h"
#include "shb_common.h"

namespace shb {

struct TraceArgs {
  SceneView scene;
  const float4* rays;         // 2 float4 per ray slot: (o, tnear), (d, tfar)
  const uint32_t* queue;      // ray slots to trace, or null for 0..count-1
  const uint32_t* count_ptr;  // device count of queue entries, or null
  uint32_t count_const;       // count when count_ptr is null
  int has_nodes;              // 0: empty scene
  uint32_t max_steps;         // per-ray traversal step cap (node count + 1)
  uint32_t stack_entries;     // LDS stack entries per lane (tree depth - 1)
  uint32_t* head;             // persistent form: queue position counter (zeroed)
  uint32_t refill;            // persistent form: idle lanes that trigger a refill
  int num_cus;                // persistent form: compute units of the device
  uint32_t* error;            // [0] bits: 1 stack overflow, 2 step cap; [1..3] first ray
  uint32_t* ray_counts;       // optional: node visits, triangle tests per slot
  unsigned long long* totals; // optional: 64 pairs of summed (nodes, tris)
};

struct BakeArgs {
  SceneView scene;
  const float4* surface;      // 3 float4 per supersampled texel
  const uint32_t* texel_list; // valid texels, ascending
  uint32_t texel_begin;       // first rank of this batch in texel_list
  uint32_t texels_in_batch;   // T
  uint32_t spp;
  uint32_t bounces;
  uint32_t P;                 // path slots (>= T * spp)
  uint32_t job;
  uint64_t seed;
  float4* ray;                // 2 per slot: extension ray
  float4* hit;                // 1 per slot: (t, id, u, v)
  float4* thr;                // throughput, branch id in w
  float4* li;                 // radiance along the primary ray; w = sky hit
  float4* nee0;               // texel NEE radiance; w = visible
  float4* nee0_dir;           // texel NEE direction (unnormalised)
  float4* dir0;               // primary direction
  float4* sray;               // 2 per shadow slot; 2P slots (vertex, texel)
  float4* scontrib;           // vertex NEE contribution per slot
  uint32_t* shadow_queue;
  uint32_t* shadow_count;
  float* sh_out;              // 27 per supersampled texel
  float* vis_out;             // 1 per supersampled texel
};

#ifndef SHB_HOST_ONLY
// Traversal launches: `capacity` is the largest possible queue length (the
// grid covers it; threads past the device-side count exit).
hipError_t launch_trace_hits(const TraceArgs& a, float4* hits, bool count, uint32_t capacity,
                             int block, hipStream_t st);
hipError_t launch_trace_occluded(const TraceArgs& a, uint32_t* occ, bool count,
                                 uint32_t capacity, int block, hipStream_t st);
hipError_t launch_trace_bake_shadow(const TraceArgs& a, const BakeArgs& b, bool count,
                                    uint32_t capacity, int block, hipStream_t st);
hipError_t launch_hit_details(const SceneView& sc, const float4* rays, const float4* hits,
                              uint32_t n, shb_hit* out, hipStream_t st);
hipError_t launch_generate(const BakeArgs& b, int grid, hipStream_t st);
hipError_t launch_shade(const BakeArgs& b, int depth, const uint32_t* in_q,
                        const uint32_t* in_n, uint32_t* out_q, uint32_t* out_n, int grid,
                        hipStream_t st);
hipError_t launch_project(const BakeArgs& b, hipStream_t st);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t st);
hipError_t launch_fill_f32(float* p, float v, size_t n, hipStream_t st);

#endif  // SHB_HOST_ONLY

// image / raster kernels (image_kernels.hip)
struct RasterArgs {
  const float* luv;           // 2 per vertex (flattened scene arrays)
  const float* wpos;          // 3 per vertex, world space
  const float* wnrm;          // 3 per vertex, world space
  const float* wtan;          // 4 per vertex, world space
  const uint32_t* tri_vidx;   // 3 global vertex indices per global triangle
  const int32_t* tri_mat;     // material per global triangle
  uint32_t num_tris;
  int32_t W, H;               // supersampled size
  const uint32_t* tri_scan;   // exclusive scan of bbox texel counts (n+1)
  uint32_t* win;              // per texel: last regular writer + 1 (0: none)
  uint32_t* fb;               // per texel: first fallback writer + 1 (0: none)
  uint32_t* tri_hit;          // per triangle: covered any texel centre
  float4* surface;            // 3 float4 per texel (out)
};

#ifndef SHB_HOST_ONLY
// counts: num_tris + 1 uint32 scratch; scan: num_tris + 1 uint64 scratch.
hipError_t launch_raster(const RasterArgs& a, uint32_t* counts, unsigned long long* scan,
                         hipStream_t st);
hipError_t launch_validity_mask(const float4* surface, size_t n, uint8_t* mask,
                                hipStream_t st);
hipError_t launch_downsample(const float* in, int w, int h, int scale, int ch, float* out,
                             hipStream_t st);
hipError_t launch_downsample_mask(const uint8_t* in, int w, int h, int scale, uint8_t* out,
                                  hipStream_t st);
hipError_t launch_dilate(float* a, float* b, uint8_t* ma, uint8_t* mb, int w, int h, int ch,
                         int passes, uint32_t* changed, bool* result_in_a, hipStream_t st);
// Ascending list of the indices whose mask byte is set; block_tmp holds
// compact_blocks(n) + 1 uint32 (count written to list[n]).
uint32_t compact_blocks(uint32_t n);
hipError_t launch_compact_valid(const uint8_t* mask, uint32_t n, uint32_t* list,
                                uint32_t* block_tmp, hipStream_t st);

#endif  // SHB_HOST_ONLY

}  // namespace shb

#endif  // SHB_KERNELS_H_
