// Context state of the baker (host side of the C ABI).
#ifndef SHB_CTX_H_
#define SHB_CTX_H_

#include <string>
#include <vector>

#include "../../include/shb.h"
#include "bvh_build.h"
#include "scene_host.h"
#include "kernels.h"

namespace shb {

// Grow-only device buffer.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipError_t reserve(size_t n) {
    if (n <= bytes) return hipSuccess;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&ptr, n);
    if (e == hipSuccess) bytes = n;
    return e;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  std::string err;
  shb_tuning tuning{256, 1 << 25, 12, 8};
  // device-resident result of the last shb_bake_lightmap (base resolution)
  const float* lm_sh = nullptr;
  const float* lm_vis = nullptr;
  const uint8_t* lm_mask = nullptr;
  size_t lm_texels = 0;

  // scene
  bool has_scene = false;
  HostScene hs;
  BvhBuildOut bvh;
  DevBuf d_nodes, d_tris, d_shade, d_tri_info, d_mats, d_texels, d_lut, d_lights, d_area,
      d_lscale, d_ascale;
  // raster inputs
  DevBuf d_luv, d_wpos, d_wnrm, d_wtan, d_tri_vidx, d_tri_mat;
  // surface
  bool has_surface = false;
  int width = 0, height = 0, ss = 1;
  DevBuf d_surface;  // 3 float4 per supersampled texel
  // bake scratch
  DevBuf d_paths;    // path state arrays
  DevBuf d_queues;
  DevBuf d_counters;
  DevBuf d_work;     // misc small device scratch (error flags, totals)
  DevBuf d_list, d_mask, d_tmp;
  DevBuf d_sh, d_vis;         // supersampled outputs
  DevBuf d_img_a, d_img_b, d_mask_a, d_mask_b;
  DevBuf d_rays, d_hits, d_hit_out, d_raycounts;
  DevBuf d_adj;               // adjoint accumulators
  // branch links of alpha < 1 splits; spare slots per batch that the last
  // bake needed (the next bake starts with at least as many)
  DevBuf d_link;
  uint32_t spare_slots = 0;
  // batched / graph bakes run on this context: job table, batch maps,
  // concatenated BVHs of several jobs, the captured launch sequence
  DevBuf d_jobs, d_maps, d_cnodes, d_ctris, d_vcount;
  // parameters of shb_set_params (host copies) and the adjoint's inputs
  std::vector<float> lscale_h, ascale_h;
  DevBuf d_dsh, d_grad;
  hipGraphExec_t graph_exec = nullptr;
  std::string graph_key;
  // environment map of the last shb_env_sh_latlong call, its partial sums
  DevBuf d_env, d_env_part;
  int env_w = 0, env_h = 0, env_ch = 0;

  SceneView view() const {
    SceneView v;
    v.nodes = d_nodes.as<CWNode>();
    v.tris = d_tris.as<float4>();
    v.shade = d_shade.as<float4>();
    v.tri_info = d_tri_info.as<int4>();
    v.mats = d_mats.as<MaterialDev>();
    v.texels = d_texels.as<uint8_t>();
    v.srgb_lut = d_lut.as<float>();
    v.lights = d_lights.as<LightDev>();
    v.area_tris = d_area.as<AreaTri>();
    v.light_scale = d_lscale.as<float>();
    v.albedo_scale = d_ascale.as<float>();
    v.num_lights = (int32_t)hs.lights.size();
    v.num_tris = (int32_t)hs.num_tris;
    v.has_nodes = bvh.nodes.empty() ? 0 : 1;
    v.emitter_test = 1;
    return v;
  }
};

}  // namespace shb

struct shb_ctx {
  shb::Ctx c;
};

#endif  // SHB_CTX_H_
