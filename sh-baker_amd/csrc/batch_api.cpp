// Batched bake and autotuner entry points (no reference counterpart: the
// reference bakes one lightmap per process, src/main.cpp:40-193).
#include <algorithm>
#include <set>
#include <string>
#include <vector>

#include "bake_run.h"

using namespace shb;

extern "C" {

int shb_bake_batched(shb_ctx* const* ctxs, int32_t num_jobs, const shb_bake_cfg* cfgs,
                     float* const* sh_outs, float* const* vis_outs, shb_bake_stats* stats) {
  if (!ctxs || num_jobs <= 0 || !ctxs[0]) return SHB_ERR_INVALID_ARG;
  Ctx& x = ctxs[0]->c;
  x.err.clear();
  if (!cfgs) return fail(x, SHB_ERR_INVALID_ARG, "cfgs is NULL");
  if (num_jobs > kMaxJobs)
    return fail(x, SHB_ERR_LIMIT, "at most " + std::to_string(kMaxJobs) + " jobs per batched bake");
  const double t0 = now_s();
  std::set<const shb_ctx*> seen;
  std::vector<JobIn> jobs((size_t)num_jobs);
  for (int32_t i = 0; i < num_jobs; ++i) {
    if (!ctxs[i]) return fail(x, SHB_ERR_INVALID_ARG, "job " + std::to_string(i) + ": ctx is NULL");
    if (!seen.insert(ctxs[i]).second)
      return fail(x, SHB_ERR_INVALID_ARG,
                  "job " + std::to_string(i) + ": a context may appear once (it holds the outputs)");
    if (ctxs[i]->c.device != x.device)
      return fail(x, SHB_ERR_INVALID_ARG,
                  "job " + std::to_string(i) + ": every context must be on the same device");
    if ((cfgs[i].flags & SHB_BAKE_COUNT_VISITS) != (cfgs[0].flags & SHB_BAKE_COUNT_VISITS))
      return fail(x, SHB_ERR_INVALID_ARG, "SHB_BAKE_COUNT_VISITS must be set on all jobs or none");
    jobs[i].c = &ctxs[i]->c;
    jobs[i].cfg = cfgs[i];
    jobs[i].stream = 0;  // same stream as a plain bake of the job: identical results
  }
  // every job's texel list on the batch stream, one wait for all counts
  // (work a job context still has queued on its own stream comes first)
  SHB_TRY(x, hipSetDevice(x.device));
  for (int32_t i = 0; i < num_jobs; ++i) SHB_TRY(x, hipStreamSynchronize(jobs[i].c->stream));
  SHB_TRY(x, x.d_vcount.reserve((size_t)num_jobs * 4 + 16));
  for (int32_t i = 0; i < num_jobs; ++i) {
    int rc = prepare_job_async(jobs[i], x.stream, x.d_vcount.as<uint32_t>() + i);
    if (rc) return fail(x, rc, "job " + std::to_string(i) + ": " + jobs[i].c->err);
  }
  std::vector<uint32_t> counts((size_t)num_jobs);
  SHB_TRY(x, hipMemcpyAsync(counts.data(), x.d_vcount.ptr, (size_t)num_jobs * 4,
                            hipMemcpyDeviceToHost, x.stream));
  SHB_TRY(x, hipStreamSynchronize(x.stream));
  for (int32_t i = 0; i < num_jobs; ++i) jobs[i].valid = counts[i];
  BakeRun run;
  int rc = run_bake(x, jobs, true, &run);
  if (rc) return rc;
  for (int32_t i = 0; i < num_jobs; ++i) {
    Ctx& c = *jobs[i].c;
    const size_t npx = (size_t)c.width * c.ss * c.height * c.ss;
    if (sh_outs && sh_outs[i])
      SHB_TRY(x, hipMemcpyAsync(sh_outs[i], c.d_sh.ptr, npx * 27 * 4, hipMemcpyDeviceToHost,
                                x.stream));
    if (vis_outs && vis_outs[i])
      SHB_TRY(x, hipMemcpyAsync(vis_outs[i], c.d_vis.ptr, npx * 4, hipMemcpyDeviceToHost,
                                x.stream));
  }
  SHB_TRY(x, hipStreamSynchronize(x.stream));
  if (stats) {
    fill_stats(stats, run, 0, t0);
    uint64_t paths = 0;
    for (const JobIn& j : jobs) paths += (uint64_t)j.valid * (uint64_t)j.cfg.samples;
    stats->paths = paths;
  }
  return SHB_OK;
}

int shb_bake_tiled(shb_ctx* h, const shb_bake_cfg* cfg, const shb_tile* tiles,
                   int32_t num_tiles, float* sh_out, float* vis_out, shb_bake_stats* stats) {
  if (!h) return SHB_ERR_INVALID_ARG;
  Ctx& c = h->c;
  c.err.clear();
  if (!cfg) return fail(c, SHB_ERR_INVALID_ARG, "cfg is NULL");
  if (!tiles || num_tiles <= 0) return fail(c, SHB_ERR_INVALID_ARG, "no tiles");
  if (num_tiles > kMaxJobs)
    return fail(c, SHB_ERR_LIMIT, "at most " + std::to_string(kMaxJobs) + " tiles per bake");
  if (!c.has_surface)
    return fail(c, SHB_ERR_NO_SURFACE, "no surface: call shb_rasterize or shb_set_surface");
  const int64_t W = (int64_t)c.width * c.ss, H = (int64_t)c.height * c.ss;
  for (int32_t i = 0; i < num_tiles; ++i) {
    const shb_tile& t = tiles[i];
    if (t.w <= 0 || t.h <= 0 || t.x0 < 0 || t.y0 < 0 || (int64_t)t.x0 + t.w > W ||
        (int64_t)t.y0 + t.h > H)
      return fail(c, SHB_ERR_INVALID_ARG,
                  "tile " + std::to_string(i) + " is empty or outside the " + std::to_string(W) +
                      " x " + std::to_string(H) + " supersampled surface");
    for (int32_t k = 0; k < i; ++k) {
      const shb_tile& u = tiles[k];
      if (t.x0 < u.x0 + u.w && u.x0 < t.x0 + t.w && t.y0 < u.y0 + u.h && u.y0 < t.y0 + t.h)
        return fail(c, SHB_ERR_INVALID_ARG,
                    "tiles " + std::to_string(k) + " and " + std::to_string(i) + " overlap");
    }
  }
  const double t0 = now_s();
  // the whole surface's valid texels (ascending), then each tile's share of
  // them, row by row: a tile row's texels are one range of the ascending list
  std::vector<JobIn> jobs((size_t)num_tiles);
  jobs[0].c = &c;
  jobs[0].cfg = *cfg;
  int rc = prepare_job(jobs[0]);
  if (rc) return rc;
  const uint32_t V = jobs[0].valid;
  std::vector<uint32_t> list(V), tiled;
  tiled.reserve(V);
  if (V)
    SHB_TRY(c, hipMemcpy(list.data(), c.d_list.ptr, (size_t)V * 4, hipMemcpyDeviceToHost));
  for (int32_t i = 0; i < num_tiles; ++i) {
    const shb_tile& t = tiles[i];
    JobIn& j = jobs[i];
    j.c = &c;
    j.cfg = *cfg;
    j.stream = 0;  // the plain bake's stream: identical results
    j.list_offset = (uint32_t)tiled.size();
    for (int64_t y = t.y0; y < (int64_t)t.y0 + t.h; ++y) {
      const uint32_t lo = (uint32_t)(y * W + t.x0), hi = (uint32_t)(y * W + t.x0 + t.w);
      auto a = std::lower_bound(list.begin(), list.end(), lo);
      auto b = std::lower_bound(a, list.end(), hi);
      tiled.insert(tiled.end(), a, b);
    }
    j.valid = (uint32_t)tiled.size() - j.list_offset;
  }
  if (!tiled.empty())
    SHB_TRY(c, hipMemcpy(c.d_list.ptr, tiled.data(), tiled.size() * 4, hipMemcpyHostToDevice));
  BakeRun run;
  rc = run_bake(c, jobs, true, &run);
  if (rc) return rc;
  const size_t npx = (size_t)W * H;
  if (sh_out)
    SHB_TRY(c, hipMemcpyAsync(sh_out, c.d_sh.ptr, npx * 27 * 4, hipMemcpyDeviceToHost, c.stream));
  if (vis_out)
    SHB_TRY(c, hipMemcpyAsync(vis_out, c.d_vis.ptr, npx * 4, hipMemcpyDeviceToHost, c.stream));
  SHB_TRY(c, hipStreamSynchronize(c.stream));
  if (stats) {
    fill_stats(stats, run, 0, t0);
    stats->valid_texels = tiled.size();
    stats->paths = (uint64_t)tiled.size() * (uint64_t)cfg->samples;
  }
  return SHB_OK;
}

int shb_autotune(shb_ctx* h, const shb_bake_cfg* cfg, shb_tuning* best) {
  if (!h) return SHB_ERR_INVALID_ARG;
  Ctx& c = h->c;
  c.err.clear();
  if (!cfg) return fail(c, SHB_ERR_INVALID_ARG, "cfg is NULL");
  // short bakes of the current scene and surface: same bounces, at most 16
  // samples per texel (the per-path work does not depend on spp)
  shb_bake_cfg probe = *cfg;
  probe.samples = std::min(cfg->samples, 16);
  probe.flags = 0;
  std::vector<JobIn> jobs(1);
  jobs[0].c = &c;
  jobs[0].cfg = probe;
  int rc = prepare_job(jobs[0]);
  if (rc) return rc;
  auto time_it = [&](double* sec) -> int {
    double b = 1e30;
    for (int k = 0; k < 2; ++k) {
      BakeRun run;
      int r = run_bake(c, jobs, false, &run);
      if (r) return r;
      b = std::min(b, run.dev_s);
    }
    *sec = b;
    return SHB_OK;
  };
  const shb_tuning start = c.tuning;
  shb_tuning cur = start;
  double best_s = 1e30;
  shb_tuning best_t = start;
  const int widths[] = {8, 4};
  const int blocks[] = {64, 256};
  const int chunks[] = {0, 8, 12, 16, 32};
  for (int w : widths) {
    cur.bvh_width = w;
    rc = shb_set_tuning(h, &cur);
    if (rc) return rc;
    for (int b : blocks)
      for (int q : chunks) {
        cur.block_size = b;
        cur.queue_chunk = q;
        c.tuning = cur;
        double s = 0;
        rc = time_it(&s);
        if (rc) return rc;
        if (s < best_s) {
          best_s = s;
          best_t = cur;
        }
      }
  }
  // batch size for the chosen launch shape, judged at the caller's spp
  jobs[0].cfg = *cfg;
  jobs[0].cfg.flags = 0;
  cur = best_t;
  rc = shb_set_tuning(h, &cur);
  if (rc) return rc;
  best_s = 1e30;
  const int batch[] = {1 << 22, 1 << 24, 1 << 25};
  for (int bp : batch) {
    cur.batch_paths = bp;
    c.tuning = cur;
    double s = 0;
    rc = time_it(&s);
    if (rc) return rc;
    if (s < best_s) {
      best_s = s;
      best_t = cur;
    }
  }
  rc = shb_set_tuning(h, &best_t);
  if (rc) return rc;
  if (best) *best = best_t;
  return SHB_OK;
}

int shb_bake_adjoint(shb_ctx* h, const shb_bake_cfg* cfg, const float* dloss_dsh,
                     float* dloss_dlight, float* dloss_dalbedo, shb_bake_stats* stats) {
  if (!h) return SHB_ERR_INVALID_ARG;
  Ctx& c = h->c;
  c.err.clear();
  if (!cfg) return fail(c, SHB_ERR_INVALID_ARG, "cfg is NULL");
  if (!dloss_dsh) return fail(c, SHB_ERR_INVALID_ARG, "dloss_dsh is NULL");
  if (!c.has_scene) return fail(c, SHB_ERR_NO_SCENE, "no scene uploaded");
  for (float v : c.lscale_h)
    if (!(v > 0.f)) return fail(c, SHB_ERR_INVALID_ARG, "the adjoint needs light scales > 0");
  for (float v : c.ascale_h)
    if (!(v > 0.f)) return fail(c, SHB_ERR_INVALID_ARG, "the adjoint needs albedo scales > 0");
  const double t0 = now_s();
  std::vector<JobIn> jobs(1);
  jobs[0].c = &c;
  jobs[0].cfg = *cfg;
  jobs[0].cfg.flags &= ~(uint32_t)SHB_BAKE_COUNT_VISITS;
  int rc = prepare_job(jobs[0]);
  if (rc) return rc;
  const size_t npx = (size_t)c.width * c.ss * c.height * c.ss;
  const uint32_t nl = (uint32_t)c.hs.lights.size(), nm = (uint32_t)c.hs.mats.size();
  if (nm > 65535) return fail(c, SHB_ERR_LIMIT, "the adjoint supports up to 65535 materials");
  const uint32_t nparam = 3 * (nl + nm);
  SHB_TRY(c, c.d_dsh.reserve(std::max<size_t>(npx * 27 * sizeof(float), 16)));
  SHB_TRY(c, c.d_grad.reserve(std::max<size_t>((size_t)nparam * sizeof(double), 16)));
  SHB_TRY(c, hipMemcpyAsync(c.d_dsh.ptr, dloss_dsh, npx * 27 * sizeof(float),
                            hipMemcpyHostToDevice, c.stream));
  SHB_TRY(c, launch_fill_u32(c.d_grad.as<uint32_t>(), 0u, (size_t)nparam * 2, c.stream));
  AdjIn adj;
  adj.dsh = c.d_dsh.as<float>();
  adj.grad = c.d_grad.as<double>();
  adj.nlights = nl;
  adj.nparam = nparam;
  BakeRun run;
  rc = run_bake(c, jobs, (cfg->flags & SHB_BAKE_USE_GRAPH) != 0, &run, &adj);
  if (rc) return rc;
  std::vector<double> g(nparam);
  if (nparam)
    SHB_TRY(c, hipMemcpy(g.data(), c.d_grad.ptr, (size_t)nparam * sizeof(double),
                         hipMemcpyDeviceToHost));
  if (dloss_dlight)
    for (uint32_t i = 0; i < 3 * nl; ++i) dloss_dlight[i] = (float)g[i];
  if (dloss_dalbedo)
    for (uint32_t i = 0; i < 3 * nm; ++i) dloss_dalbedo[i] = (float)g[3 * nl + i];
  fill_stats(stats, run, (uint32_t)cfg->samples, t0);
  return SHB_OK;
}

}  // extern "C"

