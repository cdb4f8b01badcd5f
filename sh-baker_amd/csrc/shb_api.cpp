// C ABI of the baker (include/shb.h): context, scene upload, raster, bake
// orchestration, post-bake image steps, fixed-ray queries.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "api_util.h"
#include "bake_run.h"
#include "scene_host.h"
#include "shb_ctx.h"

using namespace shb;

namespace {

int build_scene(Ctx& c, const shb_scene_desc* d) {
  std::string err;
  int rc = build_host_scene(d, &c.hs, &err);
  if (rc) c.err = err;
  return rc;
}

int upload_bvh(Ctx& c) {
  build_cwbvh(c.hs.world_xyz.data(), c.hs.num_tris, c.tuning.bvh_width, &c.bvh);
  // the group stack holds at most one entry per level below the root
  if (c.bvh.max_depth > kMaxBvhDepth)
    return fail(c, SHB_ERR_LIMIT,
                "BVH depth " + std::to_string(c.bvh.max_depth) + " exceeds the traversal stack (" +
                    std::to_string(kMaxBvhDepth) + " levels) even with median splits: use a "
                    "wider BVH (tuning.bvh_width)");
  if (c.bvh.nodes.size() >= kMaxTreeNodes)
    return fail(c, SHB_ERR_LIMIT,
                "BVH of " + std::to_string(c.bvh.nodes.size()) + " nodes: at most " +
                    std::to_string(kMaxTreeNodes - 1) + " (about 160 M triangles)");
  // the device form of the tree: device meta bytes, nodes kNodeBytes apart
  std::vector<uint8_t> nb(c.bvh.nodes.size() * (size_t)kNodeBytes, 0);
  for (size_t i = 0; i < c.bvh.nodes.size(); ++i) {
    const CWNode d = device_node(c.bvh.nodes[i]);
    std::memcpy(nb.data() + i * kNodeBytes, &d, sizeof(CWNode));
  }
  SHB_TRY(c, upload(c.d_nodes, nb, c.stream));
  SHB_TRY(c, hipStreamSynchronize(c.stream));  // nb is a temporary
  SHB_TRY(c, upload(c.d_tris, c.bvh.tris, c.stream));
  // each area-light triangle's record in the leaf order of this tree (the
  // shading kernels answer shadow queries that end on it with one test)
  std::vector<uint32_t> leaf_of(c.hs.num_tris, kNoLeaf);
  for (size_t i = 0; 3 * i < c.bvh.tris.size(); ++i) {
    const uint32_t id = as_uint(c.bvh.tris[3 * i].w);
    if (id < c.hs.num_tris) leaf_of[id] = (uint32_t)i;
  }
  for (size_t k = 0; k < c.hs.area_tris.size(); ++k) {
    AreaTri& at = c.hs.area_tris[k];
    const uint32_t g = c.hs.area_tri_global[k];
    at.leaf = (!c.bvh.nodes.empty() && g < c.hs.num_tris) ? leaf_of[g] : kNoLeaf;
    for (int j = 0; j < 3; ++j)
      at.rec[j] = at.leaf == kNoLeaf ? make_float4(0.f, 0.f, 0.f, 0.f)
                                     : c.bvh.tris[3 * (size_t)at.leaf + j];
  }
  SHB_TRY(c, upload(c.d_area, c.hs.area_tris, c.stream));
  return SHB_OK;
}

}  // namespace

namespace {
// The plain bake of one context: its own job, random stream 0.
int bake_one(Ctx& c, const shb_bake_cfg* cfg, BakeRun* run) {
  if (!cfg) return fail(c, SHB_ERR_INVALID_ARG, "cfg is NULL");
  std::vector<JobIn> jobs(1);
  jobs[0].c = &c;
  jobs[0].cfg = *cfg;
  int rc = prepare_job(jobs[0]);
  if (rc) return rc;
  return run_bake(c, jobs, (cfg->flags & SHB_BAKE_USE_GRAPH) != 0, run);
}
}  // namespace

// =====================================================================
extern "C" {

const char* shb_version(void) { return "shb-hip 0.1 (gfx950)"; }

int shb_create(shb_ctx** out, int device) {
  if (!out) return SHB_ERR_INVALID_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return SHB_ERR_DEVICE;
  if (device < 0 || device >= n) return SHB_ERR_INVALID_ARG;
  if (hipSetDevice(device) != hipSuccess) return SHB_ERR_DEVICE;
  shb_ctx* h = new (std::nothrow) shb_ctx();
  if (!h) return SHB_ERR_OOM;
  Ctx& c = h->c;
  c.device = device;
  if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return SHB_ERR_DEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    (void)hipStreamDestroy(c.stream);
    delete h;
    return SHB_ERR_DEVICE;
  }
  c.num_cus = prop.multiProcessorCount;
  // address of this GPU's copy of the fixed scratch (once per device)
  if (env_scratch_init(device) != hipSuccess) {
    (void)hipStreamDestroy(c.stream);
    delete h;
    return SHB_ERR_DEVICE;
  }
  // default parameters: all scales 1 (reproduces the reference bake)
  *out = h;
  return SHB_OK;
}

void shb_destroy(shb_ctx* h) {
  if (!h) return;
  Ctx& c = h->c;
  (void)hipSetDevice(c.device);
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  DevBuf* bufs[] = {&c.d_nodes,  &c.d_tris,     &c.d_shade,   &c.d_tri_info, &c.d_mats,
                    &c.d_texels, &c.d_lut,      &c.d_lights,  &c.d_area,     &c.d_lscale,
                    &c.d_ascale, &c.d_luv,      &c.d_wpos,    &c.d_wnrm,     &c.d_wtan,
                    &c.d_tri_vidx, &c.d_tri_mat, &c.d_surface, &c.d_paths,   &c.d_queues,
                    &c.d_counters, &c.d_work,    &c.d_list,     &c.d_mask,
                    &c.d_tmp,    &c.d_sh,       &c.d_vis,     &c.d_img_a,    &c.d_img_b,
                    &c.d_mask_a, &c.d_mask_b,   &c.d_rays,    &c.d_hits,     &c.d_hit_out,
                    &c.d_raycounts, &c.d_adj,   &c.d_jobs,    &c.d_maps,     &c.d_cnodes,
                    &c.d_ctris,  &c.d_vcount,  &c.d_dsh,     &c.d_grad,
                    &c.d_env,    &c.d_env_part, &c.d_link};
  if (c.graph_exec) (void)hipGraphExecDestroy(c.graph_exec);
  c.graph_exec = nullptr;
  for (DevBuf* b : bufs) b->release();
  if (c.stream) (void)hipStreamDestroy(c.stream);
  delete h;
}

const char* shb_last_error(const shb_ctx* h) { return h ? h->c.err.c_str() : "ctx is NULL"; }

int shb_upload_scene(shb_ctx* h, const shb_scene_desc* scene) {
  if (!h) return SHB_ERR_INVALID_ARG;
  Ctx& c = h->c;
  c.err.clear();
  SHB_TRY(c, hipSetDevice(c.device));
  int rc = build_scene(c, scene);
  if (rc) return rc;
  c.has_surface = false;
  c.lm_sh = nullptr;
  rc = upload_bvh(c);
  if (rc) return rc;
  std::vector<float> lut(256);
  for (int i = 0; i < 256; ++i) lut[i] = srgb_to_linear((uint8_t)i);
  HostScene& hs = c.hs;
  SHB_TRY(c, upload(c.d_shade, hs.shade, c.stream));
  SHB_TRY(c, upload(c.d_tri_info, hs.tri_info, c.stream));
  SHB_TRY(c, upload(c.d_mats, hs.mats, c.stream));
  SHB_TRY(c, upload(c.d_texels, hs.texels, c.stream));
  SHB_TRY(c, upload(c.d_lut, lut, c.stream));
  SHB_TRY(c, upload(c.d_lights, hs.lights, c.stream));
  std::vector<float> ones_l(3 * std::max<size_t>(hs.lights.size(), 1), 1.f);
  std::vector<float> ones_m(3 * std::max<size_t>(hs.mats.size(), 1), 1.f);
  SHB_TRY(c, upload(c.d_lscale, ones_l, c.stream));
  SHB_TRY(c, upload(c.d_ascale, ones_m, c.stream));
  c.lscale_h = ones_l;
  c.ascale_h = ones_m;
  SHB_TRY(c, upload(c.d_luv, hs.luv, c.stream));
  SHB_TRY(c, upload(c.d_wpos, hs.wpos, c.stream));
  SHB_TRY(c, upload(c.d_wnrm, hs.wnrm, c.stream));
  SHB_TRY(c, upload(c.d_wtan, hs.wtan, c.stream));
  SHB_TRY(c, upload(c.d_tri_vidx, hs.tri_vidx, c.stream));
  SHB_TRY(c, upload(c.d_tri_mat, hs.tri_mat, c.stream));
  SHB_TRY(c, hipStreamSynchronize(c.stream));
  c.has_scene = true;
  return SHB_OK;
}

int shb_set_params(shb_ctx* h, const float* light_scale, const float* albedo_scale) {
  if (!h) return SHB_ERR_INVALID_ARG;
  Ctx& c = h->c;
  if (!c.has_scene) return fail(c, SHB_ERR_NO_SCENE, "no scene uploaded");
  SHB_TRY(c, hipSetDevice(c.device));
  const size_t nl = 3 * c.hs.lights.size(), nm = 3 * c.hs.mats.size();
  auto ok = [](const float* v, size_t n) {
    for (size_t i = 0; i < n; ++i)
      if (!(v[i] >= 0.f) || !std::isfinite(v[i])) return false;
    return true;
  };
  if ((light_scale && !ok(light_scale, nl)) || (albedo_scale && !ok(albedo_scale, nm)))
    return fail(c, SHB_ERR_INVALID_ARG, "scales must be finite and >= 0");
  if (light_scale && nl) {
    c.lscale_h.assign(light_scale, light_scale + nl);
    SHB_TRY(c, hipMemcpyAsync(c.d_lscale.ptr, c.lscale_h.data(), 4 * nl, hipMemcpyHostToDevice,
                              c.stream));
  }
  if (albedo_scale && nm) {
    c.ascale_h.assign(albedo_scale, albedo_scale + nm);
    SHB_TRY(c, hipMemcpyAsync(c.d_ascale.ptr, c.ascale_h.data(), 4 * nm, hipMemcpyHostToDevice,
                              c.stream));
  }
  SHB_TRY(c, hipStreamSynchronize(c.stream));
  return SHB_OK;
}

static int check_raster_cfg(Ctx& c, const shb_raster_cfg* cfg) {
  if (!cfg) return fail(c, SHB_ERR_INVALID_ARG, "raster cfg is NULL");
  if (cfg->width <= 0 || cfg->height <= 0 || cfg->supersample_scale <= 0)
    return fail(c, SHB_ERR_INVALID_ARG, "width, height and supersample_scale must be > 0");
  uint64_t n = (uint64_t)cfg->width * cfg->supersample_scale * cfg->height * cfg->supersample_scale;
  if (n > (1ull << 31)) return fail(c, SHB_ERR_LIMIT, "lightmap larger than 2^31 texels");
  return SHB_OK;
}

int shb_rasterize(shb_ctx* h, const shb_raster_cfg* cfg, shb_surface_point* out) {
  if (!h) return SHB_ERR_INVALID_ARG;
  Ctx& c = h->c;
  c.err.clear();
  if (!c.has_scene) return fail(c, SHB_ERR_NO_SCENE, "no scene uploaded");
  int rc = check_raster_cfg(c, cfg);
  if (rc) return rc;
  if (!c.hs.raster_ready)
    return fail(c, SHB_ERR_INVALID_ARG,
                "every geometry needs lightmap uvs, normals and tangents to rasterize "
                "(src/rasterizer.cpp:86-88)");
  SHB_TRY(c, hipSetDevice(c.device));
  const int W = cfg->width * cfg->supersample_scale, H = cfg->height * cfg->supersample_scale;
  const size_t npx = (size_t)W * H;
  const uint32_t n = c.hs.num_tris;
  SHB_TRY(c, c.d_surface.reserve(npx * 3 * sizeof(float4)));
  SHB_TRY(c, c.d_img_a.reserve(npx * 4 * 2 + 16));  // win + fb
  SHB_TRY(c, c.d_list.reserve(std::max<size_t>((size_t)(n + 1) * 8, 64)));  // tri_hit + counts
  SHB_TRY(c, c.d_tmp.reserve((size_t)(n + 1) * 8 + 64));  // scan
  RasterArgs a{};
  a.luv = c.d_luv.as<float>();
  a.wpos = c.d_wpos.as<float>();
  a.wnrm = c.d_wnrm.as<float>();
  a.wtan = c.d_wtan.as<float>();
  a.tri_vidx = c.d_tri_vidx.as<uint32_t>();
  a.tri_mat = c.d_tri_mat.as<int32_t>();
  a.num_tris = n;
  a.W = W;
  a.H = H;
  a.win = c.d_img_a.as<uint32_t>();
  a.fb = a.win + npx;
  a.tri_hit = c.d_list.as<uint32_t>();
  a.surface = c.d_surface.as<float4>();
  uint32_t* counts = a.tri_hit + (n + 1);
  SHB_TRY(c, launch_raster(a, counts, c.d_tmp.as<unsigned long long>(), c.stream));
  c.width = cfg->width;
  c.height = cfg->height;
  c.ss = cfg->supersample_scale;
  c.has_surface = true;
  c.lm_sh = nullptr;
  if (out) {
    SHB_TRY(c, hipMemcpyAsync(out, c.d_surface.ptr, npx * sizeof(shb_surface_point),
                              hipMemcpyDeviceToHost, c.stream));
  }
  SHB_TRY(c, hipStreamSynchronize(c.stream));
  return SHB_OK;
}

int shb_set_surface(shb_ctx* h, const shb_raster_cfg* cfg, const shb_surface_point* pts) {
  if (!h) return SHB_ERR_INVALID_ARG;
  Ctx& c = h->c;
  c.err.clear();
  int rc = check_raster_cfg(c, cfg);
  if (rc) return rc;
  if (!pts) return fail(c, SHB_ERR_INVALID_ARG, "points is NULL");
  SHB_TRY(c, hipSetDevice(c.device));
  size_t npx = (size_t)cfg->width * cfg->supersample_scale * cfg->height * cfg->supersample_scale;
  for (size_t i = 0; i < npx; ++i) {
    int m = pts[i].material_id;
    if (m >= 0 && (!c.has_scene || (uint32_t)m >= c.hs.mats.size()))
      return fail(c, SHB_ERR_INVALID_ARG,
                  "surface point " + std::to_string(i) + ": material_id out of range");
  }
  SHB_TRY(c, c.d_surface.reserve(npx * sizeof(shb_surface_point)));
  SHB_TRY(c, hipMemcpyAsync(c.d_surface.ptr, pts, npx * sizeof(shb_surface_point),
                            hipMemcpyHostToDevice, c.stream));
  SHB_TRY(c, hipStreamSynchronize(c.stream));
  c.width = cfg->width;
  c.height = cfg->height;
  c.ss = cfg->supersample_scale;
  c.has_surface = true;
  c.lm_sh = nullptr;
  return SHB_OK;
}

int shb_validity_mask(shb_ctx* h, uint8_t* mask_out) {
  if (!h) return SHB_ERR_INVALID_ARG;
  Ctx& c = h->c;
  if (!c.has_surface) return fail(c, SHB_ERR_NO_SURFACE, "no surface");
  if (!mask_out) return fail(c, SHB_ERR_INVALID_ARG, "mask_out is NULL");
  SHB_TRY(c, hipSetDevice(c.device));
  size_t npx = (size_t)c.width * c.ss * c.height * c.ss;
  SHB_TRY(c, c.d_mask.reserve(std::max<size_t>(npx, 16)));
  SHB_TRY(c, launch_validity_mask(c.d_surface.as<float4>(), npx, c.d_mask.as<uint8_t>(), c.stream));
  SHB_TRY(c, hipMemcpyAsync(mask_out, c.d_mask.ptr, npx, hipMemcpyDeviceToHost, c.stream));
  SHB_TRY(c, hipStreamSynchronize(c.stream));
  return SHB_OK;
}

int shb_bake(shb_ctx* h, const shb_bake_cfg* cfg, float* sh_out, float* vis_out,
             shb_bake_stats* stats) {
  if (!h) return SHB_ERR_INVALID_ARG;
  Ctx& c = h->c;
  c.err.clear();
  double t0 = now_s();
  SHB_TRY(c, hipSetDevice(c.device));
  BakeRun run;
  int rc = bake_one(c, cfg, &run);
  if (rc) return rc;
  size_t npx = (size_t)c.width * c.ss * c.height * c.ss;
  if (sh_out)
    SHB_TRY(c, hipMemcpyAsync(sh_out, c.d_sh.ptr, npx * 27 * 4, hipMemcpyDeviceToHost, c.stream));
  if (vis_out)
    SHB_TRY(c, hipMemcpyAsync(vis_out, c.d_vis.ptr, npx * 4, hipMemcpyDeviceToHost, c.stream));
  SHB_TRY(c, hipStreamSynchronize(c.stream));
  fill_stats(stats, run, (uint32_t)cfg->samples, t0);
  return SHB_OK;
}

int shb_bake_lightmap(shb_ctx* h, const shb_bake_cfg* cfg, int32_t dilation, float* sh_out,
                      float* vis_out, uint8_t* mask_out, shb_bake_stats* stats) {
  if (!h) return SHB_ERR_INVALID_ARG;
  Ctx& c = h->c;
  c.err.clear();
  if (dilation < 0) return fail(c, SHB_ERR_INVALID_ARG, "dilation must be >= 0");
  double t0 = now_s();
  SHB_TRY(c, hipSetDevice(c.device));
  BakeRun run;
  int rc = bake_one(c, cfg, &run);
  if (rc) return rc;
  hipStream_t st = c.stream;
  const int w = c.width, hgt = c.height, ss = c.ss;
  const size_t n = (size_t)w * hgt, npx = n * ss * ss;
  SHB_TRY(c, c.d_img_a.reserve(n * 28 * 4 + 16));
  SHB_TRY(c, c.d_img_b.reserve(n * 28 * 4 + 16));
  SHB_TRY(c, c.d_mask_a.reserve(n + 16));
  SHB_TRY(c, c.d_mask_b.reserve(n + 16));
  float* sa = c.d_img_a.as<float>();
  float* sb = c.d_img_b.as<float>();
  float* va = sa + n * 27;
  float* vb = sb + n * 27;
  uint8_t* ma = c.d_mask_a.as<uint8_t>();
  uint8_t* mb = c.d_mask_b.as<uint8_t>();
  // main.cpp:138-152: mask, then downsample SH / visibility / mask
  SHB_TRY(c, launch_validity_mask(c.d_surface.as<float4>(), npx, c.d_mask.as<uint8_t>(), st));
  if (ss > 1) {
    SHB_TRY(c, launch_downsample(c.d_sh.as<float>(), w, hgt, ss, 27, sa, st));
    SHB_TRY(c, launch_downsample(c.d_vis.as<float>(), w, hgt, ss, 1, va, st));
    SHB_TRY(c, launch_downsample_mask(c.d_mask.as<uint8_t>(), w, hgt, ss, ma, st));
  } else {
    SHB_TRY(c, hipMemcpyAsync(sa, c.d_sh.ptr, n * 27 * 4, hipMemcpyDeviceToDevice, st));
    SHB_TRY(c, hipMemcpyAsync(va, c.d_vis.ptr, n * 4, hipMemcpyDeviceToDevice, st));
    SHB_TRY(c, hipMemcpyAsync(ma, c.d_mask.ptr, n, hipMemcpyDeviceToDevice, st));
  }
  // main.cpp:155-163: SH dilation grows the mask; the visibility dilation
  // starts from that grown mask.
  float* sh_res = sa;
  float* vis_res = va;
  uint8_t* m_res = ma;
  if (dilation > 0) {
    bool in_a = true;
    SHB_TRY(c, launch_dilate(sa, sb, ma, mb, w, hgt, 27, dilation, nullptr, &in_a, st));
    sh_res = in_a ? sa : sb;
    uint8_t* m_sh = in_a ? ma : mb;
    uint8_t* m_other = in_a ? mb : ma;
    float* v_src = va;
    float* v_other = vb;
    bool vin = true;
    SHB_TRY(c, launch_dilate(v_src, v_other, m_sh, m_other, w, hgt, 1, dilation, nullptr, &vin, st));
    vis_res = vin ? v_src : v_other;
    m_res = vin ? m_sh : m_other;
  }
  c.lm_sh = sh_res;
  c.lm_vis = vis_res;
  c.lm_mask = m_res;
  c.lm_texels = n;
  if (sh_out) SHB_TRY(c, hipMemcpyAsync(sh_out, sh_res, n * 27 * 4, hipMemcpyDeviceToHost, st));
  if (vis_out) SHB_TRY(c, hipMemcpyAsync(vis_out, vis_res, n * 4, hipMemcpyDeviceToHost, st));
  if (mask_out) SHB_TRY(c, hipMemcpyAsync(mask_out, m_res, n, hipMemcpyDeviceToHost, st));
  SHB_TRY(c, hipStreamSynchronize(st));
  fill_stats(stats, run, (uint32_t)cfg->samples, t0);
  return SHB_OK;
}

int shb_read_lightmap(shb_ctx* h, float* sh_out, float* vis_out, uint8_t* mask_out) {
  if (!h) return SHB_ERR_INVALID_ARG;
  Ctx& c = h->c;
  c.err.clear();
  if (!c.lm_sh) return fail(c, SHB_ERR_NO_SURFACE, "no lightmap baked in this context yet");
  SHB_TRY(c, hipSetDevice(c.device));
  const size_t n = c.lm_texels;
  if (sh_out) SHB_TRY(c, hipMemcpyAsync(sh_out, c.lm_sh, n * 27 * 4, hipMemcpyDeviceToHost, c.stream));
  if (vis_out) SHB_TRY(c, hipMemcpyAsync(vis_out, c.lm_vis, n * 4, hipMemcpyDeviceToHost, c.stream));
  if (mask_out) SHB_TRY(c, hipMemcpyAsync(mask_out, c.lm_mask, n, hipMemcpyDeviceToHost, c.stream));
  SHB_TRY(c, hipStreamSynchronize(c.stream));
  return SHB_OK;
}

int shb_downsample(shb_ctx* h, int32_t width, int32_t height, int32_t scale, int32_t channels,
                   const float* in, float* out) {
  if (!h) return SHB_ERR_INVALID_ARG;
  Ctx& c = h->c;
  c.lm_sh = nullptr;
  if (width <= 0 || height <= 0 || scale <= 0 || channels <= 0 || !in || !out)
    return fail(c, SHB_ERR_INVALID_ARG, "bad downsample arguments");
  SHB_TRY(c, hipSetDevice(c.device));
  size_t n_out = (size_t)width * height * channels, n_in = n_out * scale * scale;
  SHB_TRY(c, c.d_img_a.reserve(n_in * 4));
  SHB_TRY(c, c.d_img_b.reserve(n_out * 4));
  SHB_TRY(c, hipMemcpyAsync(c.d_img_a.ptr, in, n_in * 4, hipMemcpyHostToDevice, c.stream));
  SHB_TRY(c, launch_downsample(c.d_img_a.as<float>(), width, height, scale, channels,
                               c.d_img_b.as<float>(), c.stream));
  SHB_TRY(c, hipMemcpyAsync(out, c.d_img_b.ptr, n_out * 4, hipMemcpyDeviceToHost, c.stream));
  SHB_TRY(c, hipStreamSynchronize(c.stream));
  return SHB_OK;
}

int shb_downsample_mask(shb_ctx* h, int32_t width, int32_t height, int32_t scale,
                        const uint8_t* in, uint8_t* out) {
  if (!h) return SHB_ERR_INVALID_ARG;
  Ctx& c = h->c;
  c.lm_sh = nullptr;
  if (width <= 0 || height <= 0 || scale <= 0 || !in || !out)
    return fail(c, SHB_ERR_INVALID_ARG, "bad downsample_mask arguments");
  SHB_TRY(c, hipSetDevice(c.device));
  size_t n_out = (size_t)width * height, n_in = n_out * scale * scale;
  SHB_TRY(c, c.d_mask_a.reserve(n_in));
  SHB_TRY(c, c.d_mask_b.reserve(n_out));
  SHB_TRY(c, hipMemcpyAsync(c.d_mask_a.ptr, in, n_in, hipMemcpyHostToDevice, c.stream));
  SHB_TRY(c, launch_downsample_mask(c.d_mask_a.as<uint8_t>(), width, height, scale,
                                    c.d_mask_b.as<uint8_t>(), c.stream));
  SHB_TRY(c, hipMemcpyAsync(out, c.d_mask_b.ptr, n_out, hipMemcpyDeviceToHost, c.stream));
  SHB_TRY(c, hipStreamSynchronize(c.stream));
  return SHB_OK;
}

int shb_dilate(shb_ctx* h, int32_t width, int32_t height, int32_t channels, float* data,
               uint8_t* mask, int32_t passes) {
  if (!h) return SHB_ERR_INVALID_ARG;
  Ctx& c = h->c;
  c.lm_sh = nullptr;
  if (width < 0 || height < 0 || channels <= 0 || passes < 0)
    return fail(c, SHB_ERR_INVALID_ARG, "bad dilate arguments");
  size_t n = (size_t)width * height;
  if (n == 0 || passes == 0) return SHB_OK;  // dilation.cpp:11 / empty loop
  if (!data || !mask) return fail(c, SHB_ERR_INVALID_ARG, "data or mask is NULL");
  SHB_TRY(c, hipSetDevice(c.device));
  SHB_TRY(c, c.d_img_a.reserve(n * channels * 4));
  SHB_TRY(c, c.d_img_b.reserve(n * channels * 4));
  SHB_TRY(c, c.d_mask_a.reserve(n));
  SHB_TRY(c, c.d_mask_b.reserve(n));
  SHB_TRY(c, hipMemcpyAsync(c.d_img_a.ptr, data, n * channels * 4, hipMemcpyHostToDevice, c.stream));
  SHB_TRY(c, hipMemcpyAsync(c.d_mask_a.ptr, mask, n, hipMemcpyHostToDevice, c.stream));
  bool in_a = true;
  SHB_TRY(c, launch_dilate(c.d_img_a.as<float>(), c.d_img_b.as<float>(), c.d_mask_a.as<uint8_t>(),
                           c.d_mask_b.as<uint8_t>(), width, height, channels, passes, nullptr,
                           &in_a, c.stream));
  SHB_TRY(c, hipMemcpyAsync(data, in_a ? c.d_img_a.ptr : c.d_img_b.ptr, n * channels * 4,
                            hipMemcpyDeviceToHost, c.stream));
  SHB_TRY(c, hipMemcpyAsync(mask, in_a ? c.d_mask_a.ptr : c.d_mask_b.ptr, n,
                            hipMemcpyDeviceToHost, c.stream));
  SHB_TRY(c, hipStreamSynchronize(c.stream));
  return SHB_OK;
}

int shb_bvh_export(shb_ctx* h, void* nodes, uint32_t* num_nodes, void* tris, uint32_t* num_tris) {
  if (!h || !num_nodes || !num_tris) return SHB_ERR_INVALID_ARG;
  Ctx& c = h->c;
  c.err.clear();
  if (!c.has_scene) return fail(c, SHB_ERR_NO_SCENE, "no scene uploaded");
  const uint32_t nn = (uint32_t)c.bvh.nodes.size();
  const uint32_t nt = (uint32_t)(c.bvh.tris.size() / 3);
  if ((nodes && *num_nodes < nn) || (tris && *num_tris < nt))
    return fail(c, SHB_ERR_INVALID_ARG,
                "buffers hold fewer than " + std::to_string(nn) + " nodes / " +
                    std::to_string(nt) + " triangles");
  if (nodes && nn) std::memcpy(nodes, c.bvh.nodes.data(), (size_t)nn * sizeof(CWNode));
  if (tris && nt) std::memcpy(tris, c.bvh.tris.data(), (size_t)nt * 3 * sizeof(float4));
  *num_nodes = nn;
  *num_tris = nt;
  return SHB_OK;
}

int shb_trace_rays(shb_ctx* h, const shb_ray* rays, uint32_t n, int32_t any_hit, shb_hit* hits,
                   uint32_t* node_visits, uint32_t* tri_tests) {
  if (!h) return SHB_ERR_INVALID_ARG;
  Ctx& c = h->c;
  c.err.clear();
  if (!c.has_scene) return fail(c, SHB_ERR_NO_SCENE, "no scene uploaded");
  if (n == 0) return SHB_OK;
  if (!rays || !hits) return fail(c, SHB_ERR_INVALID_ARG, "rays or hits is NULL");
  SHB_TRY(c, hipSetDevice(c.device));
  hipStream_t st = c.stream;
  SHB_TRY(c, c.d_rays.reserve((size_t)n * sizeof(shb_ray)));
  SHB_TRY(c, c.d_hits.reserve((size_t)n * sizeof(float4)));
  SHB_TRY(c, c.d_hit_out.reserve((size_t)n * sizeof(shb_hit)));
  SHB_TRY(c, c.d_work.reserve(kWorkBytes));
  const bool count = node_visits || tri_tests;
  if (count) SHB_TRY(c, c.d_raycounts.reserve((size_t)n * 8));
  SHB_TRY(c, hipMemcpyAsync(c.d_rays.ptr, rays, (size_t)n * sizeof(shb_ray), hipMemcpyHostToDevice, st));
  SHB_TRY(c, c.d_counters.reserve(64));
  SHB_TRY(c, launch_fill_u32(c.d_work.as<uint32_t>(), 0u, kWorkBytes / 4, st));
  SHB_TRY(c, launch_fill_u32(c.d_counters.as<uint32_t>(), 0u, 16, st));
  TraceArgs ta = trace_args(c);
  if (ta.refill) ta.head = c.d_counters.as<uint32_t>();
  ta.rays = c.d_rays.as<float4>();
  ta.queue = nullptr;
  ta.count_ptr = nullptr;
  ta.count_const = n;
  ta.ray_counts = count ? c.d_raycounts.as<uint32_t>() : nullptr;
  if (any_hit) {
    SHB_TRY(c, launch_trace_occluded(ta, c.d_hits.as<uint32_t>(), count, n, c.tuning.block_size, st));
    std::vector<uint32_t> occ(n);
    SHB_TRY(c, hipMemcpyAsync(occ.data(), c.d_hits.ptr, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    SHB_TRY(c, hipStreamSynchronize(st));
    for (uint32_t i = 0; i < n; ++i) {
      std::memset(&hits[i], 0, sizeof(shb_hit));
      hits[i].prim_id = occ[i] ? 0 : -1;
      hits[i].geom_id = -1;
      hits[i].material_id = -1;
    }
  } else {
    SHB_TRY(c, launch_trace_hits(ta, c.d_hits.as<float4>(), count, n, c.tuning.block_size, st));
    SHB_TRY(c, launch_hit_details(c.view(), c.d_rays.as<float4>(), c.d_hits.as<float4>(), n,
                                  c.d_hit_out.as<shb_hit>(), st));
    SHB_TRY(c, hipMemcpyAsync(hits, c.d_hit_out.ptr, (size_t)n * sizeof(shb_hit), hipMemcpyDeviceToHost, st));
    SHB_TRY(c, hipStreamSynchronize(st));
  }
  if (count) {
    std::vector<uint32_t> rc2((size_t)n * 2);
    SHB_TRY(c, hipMemcpy(rc2.data(), c.d_raycounts.ptr, rc2.size() * 4, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; ++i) {
      if (node_visits) node_visits[i] = rc2[2 * i];
      if (tri_tests) tri_tests[i] = rc2[2 * i + 1];
    }
  }
  return check_overflow(c, c.bvh.nodes.size() + 1);
}

int shb_set_tuning(shb_ctx* h, const shb_tuning* t) {
  if (!h || !t) return SHB_ERR_INVALID_ARG;
  Ctx& c = h->c;
  if (t->block_size != 64 && t->block_size != 128 && t->block_size != 256)
    return fail(c, SHB_ERR_INVALID_ARG, "block_size must be 64, 128 or 256");
  if (t->batch_paths < 1024 || t->batch_paths > (1 << 26))
    return fail(c, SHB_ERR_INVALID_ARG, "batch_paths must be in [1024, 2^26]");
  if (t->queue_chunk < 0 || t->queue_chunk > 64)
    return fail(c, SHB_ERR_INVALID_ARG, "queue_chunk must be in [0, 64]");
  if (t->bvh_width < 2 || t->bvh_width > 8)
    return fail(c, SHB_ERR_INVALID_ARG, "bvh_width must be in [2, 8]");
  bool rebuild = c.has_scene && t->bvh_width != c.tuning.bvh_width;
  c.tuning = *t;
  if (rebuild) {
    SHB_TRY(c, hipSetDevice(c.device));
    int rc = upload_bvh(c);
    if (rc) return rc;
    SHB_TRY(c, hipStreamSynchronize(c.stream));
  }
  return SHB_OK;
}

int shb_get_tuning(const shb_ctx* h, shb_tuning* t) {
  if (!h || !t) return SHB_ERR_INVALID_ARG;
  *t = h->c.tuning;
  return SHB_OK;
}

}  // extern "C"
