// Kernel argument blocks and launchers shared by the host API and the HIP
// kernels of the baker.
#ifndef SHB_KERNELS_H_
#define SHB_KERNELS_H_

#include "../../include/shb.h"
#include "shb_common.h"

namespace shb {

constexpr int kMaxJobs = 64;  // lightmaps in one batched bake

// One lightmap of a bake: an entry of the device job table.
struct JobDev {
  SceneView scene;             // the job's scene tables (shading, lights, materials)
  const float4* surface;       // 3 float4 per supersampled texel
  const uint32_t* texel_list;  // the job's valid texels, ascending
  float* sh_out;               // 27 per supersampled texel
  float* vis_out;              // 1 per supersampled texel
  uint64_t seed;
  uint32_t spp;
  uint32_t bounces;
  uint32_t stream;             // random-stream id (the counter's job word)
  uint32_t nls;                // light samples per NEE (BakeConfig::num_light_samples)
};

// Path slots of one wavefront batch: segment s holds texels
// [texel_begin[s], texel_begin[s] + T_s) of job[s] (ranks in its texel
// list), T_s = tex_base[s+1] - tex_base[s], as slots
// [slot_base[s], slot_base[s+1]) laid out sample-major (slot = base +
// sample * T_s + texel), so one texel's samples are T_s slots apart.
struct BatchMap {
  uint32_t nseg;
  uint32_t pad;
  uint32_t slot_base[kMaxJobs + 1];
  uint32_t tex_base[kMaxJobs + 1];
  uint32_t job[kMaxJobs];
  uint32_t texel_begin[kMaxJobs];
  uint32_t root[kMaxJobs];     // the job's BVH root in the traversal arrays
};

// Largest s with base[s] <= v (base[0] = 0, n entries, ascending).
__device__ __forceinline__ uint32_t find_segment(const uint32_t* base, uint32_t n, uint32_t v) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (base[mid] <= v)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

constexpr uint32_t kNoRoot = 0xffffffffu;  // a job without triangles

struct TraceArgs {
  SceneView scene;            // nodes / tris used (the batch's traversal arrays)
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
  // root node per ray: root0 when nseg <= 1, else map->root[] of the
  // segment of the ray's root path slot (slot < P: itself; P <= slot < Pc:
  // a branch slot, root link[slot - P].x; slot >= Pc: a texel shadow ray of
  // slot - Pc)
  const BatchMap* map;
  uint32_t nseg;
  uint32_t root0;
  uint32_t P;
  uint32_t Pc;
  const uint4* link;
  // persistent form: test the triangles of a step's nodes spread over the
  // wave's lanes (needs fewer than 2^26 triangles)
  uint32_t wave_tris;
  // light samples per NEE (the batch's largest): shadow slots j * Pc + p
  // are vertex rays of path slot p, nls * Pc + j * P + p texel rays of p
  uint32_t nls;
};

struct BakeArgs {
  const JobDev* jobs;         // device job table
  const BatchMap* map;        // this batch's slot map (device)
  uint32_t nseg;              // segments (1: the fields below describe it)
  uint32_t job0;              // nseg == 1: its job, first texel rank, texels
  uint32_t texel_begin0;
  uint32_t T0;
  uint32_t T;                 // texels in the batch
  uint32_t P;                 // path slots of the batch's samples (roots)
  // Branch slots: the transmitted parts of alpha < 1 splits take spare slots
  // P + k, k < S (Pc = P + S slots in all per-slot arrays); link[k] =
  // (root slot, parent slot, spawn depth, 0); *spawn_count counts the
  // branches asked for, so a count above S means some were dropped.
  uint32_t S;
  uint32_t Pc;
  uint4* link;
  uint32_t* spawn_count;
  uint32_t unit_scales;       // every job's light and albedo scales are all 1
  // light samples per NEE, the batch's largest (a job's own: JobDev::nls).
  // Per-pick arrays hold nls entries per slot: nee0 / nee0_dir at j * P + p,
  // scontrib at j * Pc + p; shadow slots j * Pc + p (vertex) and
  // nls * Pc + j * P + p (texel). With nls > 1, scontrib.w is the pick's
  // visible flag (set by the shadow trace, summed by k_nee_gather).
  uint32_t nls;
  float4* ray;                // 2 per slot: extension ray
  float4* hit;                // 1 per slot: (t, id, u, v)
  float4* thr;                // throughput, branch id in w
  float4* li;                 // radiance along the primary ray; w = sky hit
  float4* nee0;               // texel NEE radiance; w = visible
  float4* nee0_dir;           // texel NEE direction (unnormalised)
  float4* dir0;               // primary direction
  float4* sray;               // 2 per shadow slot; Pc + P slots (vertex, texel)
  float4* scontrib;           // vertex NEE contribution per slot
  uint32_t* shadow_queue;
  uint32_t* shadow_count;
  uint32_t* resolved_count;   // shadow queries of this depth answered untraced
  // path-replay adjoint (single job): dL/dSH per supersampled texel
  // (27 floats), gradient sums (3 per light, then 3 per material), the
  // per-slot projection weights, the materials of each path's vertices
  // (hist[depth * Pc + slot]) and the visible flags of vertex NEE rays
  const float* adj_dsh;
  double* adj_grad;
  uint32_t adj_nlights;
  uint32_t adj_nparam;
  float4* adj_gvec;
  uint16_t* adj_hist;
  uint32_t* adj_vflag;
};

// Gradient accumulation of a block: float sums in LDS (when the parameters
// fit), flushed once per block into FP64 sums in global memory.
struct GradAcc {
  float* lds;     // null: add straight to global memory
  double* g;
  uint32_t n;
#ifndef SHB_HOST_ONLY
  __device__ void init(float* mem, double* global, uint32_t count) {
    lds = mem;
    g = global;
    n = count;
    if (lds)
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) lds[i] = 0.f;
    __syncthreads();
  }
  __device__ void add(uint32_t i, float v) {
    if (v == 0.f) return;
    if (lds)
      atomicAdd(&lds[i], v);
    else
      atomicAdd(&g[i], (double)v);
  }
  __device__ void flush() {
    __syncthreads();
    if (lds)
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
        if (lds[i] != 0.f) atomicAdd(&g[i], (double)lds[i]);
  }
#endif
};
constexpr uint32_t kAdjLdsParams = 8192;  // LDS accumulators up to 32 KB
// adj_hist entry of a vertex whose path was transmitted (alpha < 1)
constexpr uint16_t kHistTransmit = 0xFFFF;
constexpr uint32_t kNoSlot = 0xffffffffu;

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
// With b.adj_grad set, the adjoint forms of generate / shade / project run.
hipError_t launch_generate(const BakeArgs& b, int grid, hipStream_t st);
hipError_t launch_shade(const BakeArgs& b, int depth, const uint32_t* in_q,
                        const uint32_t* in_n, uint32_t* out_q, uint32_t* out_n, int grid,
                        hipStream_t st);
hipError_t launch_project(const BakeArgs& b, hipStream_t st);
// Adds the radiance of every branch slot spawned at `depth` to its parent's
// (one launch per depth, deepest first).
hipError_t launch_merge_branches(const BakeArgs& b, int depth, int grid, hipStream_t st);
// Adjoint of the visible vertex NEE terms of depth `depth` (shadow queue).
hipError_t launch_adjoint_vertex(const BakeArgs& b, int depth, const uint32_t* sh_count, int grid,
                                 hipStream_t st);
hipError_t launch_trace_bake_shadow_adj(const TraceArgs& a, const BakeArgs& b, bool count,
                                        uint32_t capacity, int block, hipStream_t st);
// b.nls > 1: the shadow trace flags visible picks; k_nee_gather then adds
// each shaded slot's visible vertex picks (in pick order, / its job's
// nls) to its radiance. The slots are those k_shade took at `depth`.
hipError_t launch_trace_bake_shadow_multi(const TraceArgs& a, const BakeArgs& b, bool count,
                                          uint32_t capacity, int block, hipStream_t st);
hipError_t launch_nee_gather(const BakeArgs& b, const uint32_t* in_q, const uint32_t* in_n,
                             int grid, hipStream_t st);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t st);
// Copies n BVH nodes, shifting child and triangle bases by node_off / tri_off
// (concatenated trees of a batched bake).
hipError_t launch_relocate_nodes(const CWNode* src, CWNode* dst, uint32_t n, uint32_t node_off,
                                 uint32_t tri_off, hipStream_t st);
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

// Launch shape of the LatLong environment projection: gx x gy blocks of 256
// columns x rows_per_block rows; gx * gy * 27 FP64 partials.
struct EnvGrid {
  int gx = 0, gy = 0, rows_per_block = 0;
};
EnvGrid env_grid(int w, int h);
// Bytes of the fixed partial-sum scratch (2 MiB): (gx * gy + 1) * 27 FP64
// must fit, i.e. up to 9708 blocks (a 16384 x 8192 map takes 8192).
constexpr size_t kEnvScratchBytes = (size_t)2 << 20;

#ifndef SHB_HOST_ONLY
// SH9 of an RGB(A) equirectangular map (channels >= 3) into out (27 FP64)
// and / or out_f (27 floats); partials: g.gx * g.gy * 27 FP64.
hipError_t launch_env_latlong(const float* px, int w, int h, int ch, const EnvGrid& g,
                              double* partials, double* out, float* out_f, hipStream_t st);
// Fixed device scratch of shb_env_sh_latlong_noalloc: kEnvScratchBytes per
// GPU, its address looked up once per device (with that device current) by
// env_scratch_init; env_scratch_ptr returns it (null before init).
hipError_t env_scratch_init(int device);
void* env_scratch_ptr(int device);
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
