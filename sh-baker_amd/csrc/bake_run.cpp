// Wavefront bake driver (see bake_run.h). Reference: BakeSHLightMap,
// src/baker.cpp:186-304 (one lightmap); the batched form bakes several
// lightmaps through the same batches.
#include "bake_run.h"

#include <algorithm>
#include <cstring>
#include <string>

namespace shb {

int prepare_job_async(JobIn& j, hipStream_t stream, uint32_t* count_dev) {
  Ctx& c = *j.c;
  const shb_bake_cfg* cfg = &j.cfg;
  c.lm_sh = nullptr;  // the device buffers below are about to be reused
  if (!c.has_scene) return fail(c, SHB_ERR_NO_SCENE, "no scene uploaded");
  if (!c.has_surface)
    return fail(c, SHB_ERR_NO_SURFACE, "no surface: call shb_rasterize or shb_set_surface");
  if (cfg->samples < 1 || cfg->samples > kMaxSamples)
    return fail(c, SHB_ERR_INVALID_ARG, "samples must be in [1, 65535]");
  if (cfg->bounces < 0 || cfg->bounces > kMaxBounces)
    return fail(c, SHB_ERR_INVALID_ARG, "bounces must be in [0, 15]");
  if (cfg->num_light_samples < 1 || cfg->num_light_samples > kMaxLightSamples)
    return fail(c, SHB_ERR_INVALID_ARG,
                "num_light_samples must be in [1, " + std::to_string(kMaxLightSamples) + "]");
  if (c.hs.bad_material)
    return fail(c, SHB_ERR_INVALID_ARG,
                "a geometry's material_id does not name a material (the bake shades every hit)");
  SHB_TRY(c, hipSetDevice(c.device));
  const size_t npx = (size_t)c.width * c.ss * c.height * c.ss;
  hipStream_t st = stream ? stream : c.stream;
  SHB_TRY(c, c.d_sh.reserve(std::max<size_t>(npx * 27 * sizeof(float), 16)));
  SHB_TRY(c, c.d_vis.reserve(std::max<size_t>(npx * sizeof(float), 16)));
  SHB_TRY(c, launch_fill_f32(c.d_sh.as<float>(), 0.f, npx * 27, st));
  SHB_TRY(c, launch_fill_f32(c.d_vis.as<float>(), 0.f, npx, st));
  SHB_TRY(c, c.d_mask.reserve(std::max<size_t>(npx, 16)));
  SHB_TRY(c, c.d_list.reserve(std::max<size_t>(npx * 4 + 16, 64)));
  SHB_TRY(c, c.d_tmp.reserve(((size_t)compact_blocks((uint32_t)npx) + 1) * 4 + 64));
  SHB_TRY(c, launch_validity_mask(c.d_surface.as<float4>(), npx, c.d_mask.as<uint8_t>(), st));
  SHB_TRY(c, launch_compact_valid(c.d_mask.as<uint8_t>(), (uint32_t)npx, c.d_list.as<uint32_t>(),
                                  c.d_tmp.as<uint32_t>(), st));
  SHB_TRY(c, hipMemcpyAsync(count_dev, c.d_list.as<uint32_t>() + npx, 4,
                            hipMemcpyDeviceToDevice, st));
  return SHB_OK;
}

int prepare_job(JobIn& j) {
  Ctx& c = *j.c;
  SHB_TRY(c, hipSetDevice(c.device));
  SHB_TRY(c, c.d_work.reserve(kWorkBytes));
  // the count goes through the error word's slot 1, read before any bake
  uint32_t* cnt = c.d_work.as<uint32_t>() + 1;
  int rc = prepare_job_async(j, c.stream, cnt);
  if (rc) return rc;
  uint32_t V = 0;
  SHB_TRY(c, hipMemcpyAsync(&V, cnt, 4, hipMemcpyDeviceToHost, c.stream));
  SHB_TRY(c, hipStreamSynchronize(c.stream));
  j.valid = V;
  return SHB_OK;
}

namespace {

struct Seg {
  uint32_t job, begin, T, spp;
};

// Job j's texels continue segment g (of job i) in memory: same context and
// settings, and j's texel-list entries start where g's end (the tiles of one
// surface). g may then run on into j's entries, which read as ranks of job
// i past its own valid count; per texel nothing changes (random streams are
// keyed by the texel, outputs are the context's), and the batch keeps one
// segment, so it runs the one-job kernel forms.
bool continues(const JobIn& i, const Seg& g, const JobIn& j) {
  return i.c == j.c && i.stream == j.stream && i.cfg.samples == j.cfg.samples &&
         i.cfg.bounces == j.cfg.bounces && i.cfg.num_light_samples == j.cfg.num_light_samples &&
         i.cfg.flags == j.cfg.flags && i.cfg.seed == j.cfg.seed &&
         (uint64_t)i.list_offset + g.begin + g.T == (uint64_t)j.list_offset;
}

// Greedy packing of every job's valid texels into batches of at most
// batch_paths path slots (a texel's samples never split across batches).
// A job whose texels continue the batch's last segment (continues())
// extends that segment instead of opening one.
std::vector<std::vector<Seg>> plan_batches(const std::vector<JobIn>& jobs, uint64_t cap) {
  std::vector<std::vector<Seg>> out;
  std::vector<Seg> cur;
  uint64_t used = 0;
  auto flush = [&]() {
    if (!cur.empty()) out.push_back(cur);
    cur.clear();
    used = 0;
  };
  for (uint32_t j = 0; j < jobs.size(); ++j) {
    const uint32_t spp = (uint32_t)jobs[j].cfg.samples;
    uint32_t left = jobs[j].valid, begin = 0;
    while (left > 0) {
      // job j's texels continue the last segment: extend it (its T stays
      // below 2^32: at most cap / spp)
      const bool extend =
          begin == 0 && !cur.empty() && continues(jobs[cur.back().job], cur.back(), jobs[j]);
      if (cur.size() == (size_t)kMaxJobs && !extend) flush();
      uint64_t room = cap > used ? cap - used : 0;
      uint32_t t = (uint32_t)std::min<uint64_t>(left, room / spp);
      if (t == 0) {
        if (!cur.empty()) {
          flush();
          continue;
        }
        t = 1;  // one texel's samples exceed the batch size
      }
      if (extend) {
        cur.back().T += t;
        used += (uint64_t)t * spp;
        begin += t;
        left -= t;
        continue;
      }
      cur.push_back({j, begin, t, spp});
      used += (uint64_t)t * spp;
      begin += t;
      left -= t;
    }
  }
  flush();
  return out;
}

// Launch recorder: launches, or (dry run) appends each launch's arguments
// to a key that identifies the recorded sequence.
struct Rec {
  bool dry = false;
  std::string* key = nullptr;
  template <class T>
  void add(int kind, const T& v) {
    key->append(reinterpret_cast<const char*>(&kind), sizeof(kind));
    key->append(reinterpret_cast<const char*>(&v), sizeof(T));
  }
};

// Spare path slots of one batch at most.
constexpr uint64_t kMaxSpareSlots = 1ull << 30;

// One pass of the bake. With transparent materials in some job every batch
// gets S spare path slots for the transmitted rays of alpha < 1 splits:
// at least `spare_min`, the context's hint and an eighth of the batch;
// *S_out and *demand (the most branch slots any batch asked for) tell the
// caller whether some were dropped.
int run_pass(Ctx& x, std::vector<JobIn>& jobs, bool graph, BakeRun* run, const AdjIn* adj,
             uint64_t spare_min, uint32_t* S_out, uint32_t* demand) {
  SHB_TRY(x, hipSetDevice(x.device));
  hipStream_t st = x.stream;
  const uint32_t nj = (uint32_t)jobs.size();
  const bool count = (jobs[0].cfg.flags & SHB_BAKE_COUNT_VISITS) != 0;
  run->valid = 0;
  for (const JobIn& j : jobs) run->valid += j.valid;

  // traversal arrays: the distinct scenes of the jobs (a context may carry
  // several jobs, e.g. the tiles of one surface). One scene uses its own
  // BVH; several get their trees concatenated (child and triangle bases
  // shifted), each ray starting at its job's root.
  std::vector<Ctx*> scenes_u;
  std::vector<uint32_t> scene_of(nj);
  for (uint32_t k = 0; k < nj; ++k) {
    auto it = std::find(scenes_u.begin(), scenes_u.end(), jobs[k].c);
    scene_of[k] = (uint32_t)(it - scenes_u.begin());
    if (it == scenes_u.end()) scenes_u.push_back(jobs[k].c);
  }
  const uint32_t nu = (uint32_t)scenes_u.size();
  std::vector<uint32_t> roots(nj, kNoRoot), uroot(nu, kNoRoot);
  const CWNode* tnodes = nullptr;
  const float4* ttris = nullptr;
  size_t total_nodes = 0, total_tris = 0;
  int depth = 1;
  for (const Ctx* c : scenes_u) {
    total_nodes += c->bvh.nodes.size();
    total_tris += c->bvh.tris.size() / 3;
    depth = std::max(depth, c->bvh.max_depth);
  }
  if (total_nodes >= kMaxTreeNodes)
    return fail(x, SHB_ERR_LIMIT,
                "the batch's scenes hold " + std::to_string(total_nodes) +
                    " BVH nodes together: at most " + std::to_string(kMaxTreeNodes - 1) +
                    "; bake fewer scenes per batch");
  if (nu == 1) {
    tnodes = scenes_u[0]->d_nodes.as<CWNode>();
    ttris = scenes_u[0]->d_tris.as<float4>();
    if (!scenes_u[0]->bvh.nodes.empty()) uroot[0] = 0;
  } else {
    SHB_TRY(x, x.d_cnodes.reserve(std::max<size_t>(total_nodes * kNodeBytes, 16)));
    SHB_TRY(x, x.d_ctris.reserve(std::max<size_t>(total_tris * 3 * sizeof(float4), 16)));
    size_t no = 0, to = 0;
    for (uint32_t u = 0; u < nu; ++u) {
      const Ctx& c = *scenes_u[u];
      const size_t nn = c.bvh.nodes.size(), nt = c.bvh.tris.size() / 3;
      if (nn) {
        uroot[u] = (uint32_t)no;
        SHB_TRY(x, launch_relocate_nodes(
                       c.d_nodes.as<CWNode>(),
                       reinterpret_cast<CWNode*>(x.d_cnodes.as<uint8_t>() + no * kNodeBytes),
                       (uint32_t)nn, (uint32_t)no, (uint32_t)to, st));
        SHB_TRY(x, hipMemcpyAsync(x.d_ctris.as<float4>() + 3 * to, c.d_tris.ptr,
                                  nt * 3 * sizeof(float4), hipMemcpyDeviceToDevice, st));
      }
      no += nn;
      to += nt;
    }
    tnodes = x.d_cnodes.as<CWNode>();
    ttris = x.d_ctris.as<float4>();
  }
  for (uint32_t k = 0; k < nj; ++k) roots[k] = uroot[scene_of[k]];

  // job table
  std::vector<JobDev> jt(nj);
  for (uint32_t k = 0; k < nj; ++k) {
    Ctx& c = *jobs[k].c;
    JobDev& J = jt[k];
    std::memset(&J, 0, sizeof(J));
    J.scene = c.view();
    J.scene.emitter_test = (jobs[k].cfg.flags & SHB_BAKE_TRACE_ALL_SHADOW) ? 0 : 1;
    J.surface = c.d_surface.as<float4>();
    J.texel_list = c.d_list.as<uint32_t>() + jobs[k].list_offset;
    J.sh_out = c.d_sh.as<float>();
    J.vis_out = c.d_vis.as<float>();
    J.seed = jobs[k].cfg.seed;
    J.spp = (uint32_t)jobs[k].cfg.samples;
    J.bounces = (uint32_t)jobs[k].cfg.bounces;
    J.stream = jobs[k].stream;
    J.nls = (uint32_t)jobs[k].cfg.num_light_samples;
  }
  SHB_TRY(x, x.d_jobs.reserve(std::max<size_t>(nj * sizeof(JobDev), 16)));
  SHB_TRY(x, hipMemcpyAsync(x.d_jobs.ptr, jt.data(), nj * sizeof(JobDev), hipMemcpyHostToDevice,
                            st));

  // batches and their slot maps
  const auto batches = plan_batches(jobs, (uint64_t)x.tuning.batch_paths);
  const uint32_t nb = (uint32_t)batches.size();
  std::vector<BatchMap> maps(std::max<uint32_t>(nb, 1));
  std::vector<uint32_t> maxB(nb, 0);
  uint32_t Pmax = 0;
  for (uint32_t bi = 0; bi < nb; ++bi) {
    BatchMap& m = maps[bi];
    std::memset(&m, 0, sizeof(m));
    m.nseg = (uint32_t)batches[bi].size();
    for (uint32_t s = 0; s < m.nseg; ++s) {
      const Seg& g = batches[bi][s];
      m.slot_base[s + 1] = m.slot_base[s] + g.T * g.spp;
      m.tex_base[s + 1] = m.tex_base[s] + g.T;
      m.job[s] = g.job;
      m.texel_begin[s] = g.begin;
      m.root[s] = roots[g.job];
      maxB[bi] = std::max(maxB[bi], (uint32_t)jobs[g.job].cfg.bounces);
    }
    Pmax = std::max(Pmax, m.slot_base[m.nseg]);
  }
  bool split = false;
  for (const JobIn& j : jobs) split = split || j.c->hs.has_transparency;
  uint32_t S = 0;
  if (split) {
    const uint64_t want = std::max<uint64_t>(
        {spare_min, (uint64_t)x.spare_slots, ((uint64_t)Pmax + 7) / 8, (uint64_t)256});
    if (want > kMaxSpareSlots || (uint64_t)Pmax * 2 + want > 0xffffffffull)
      return fail(x, SHB_ERR_LIMIT,
                  "alpha < 1 splits need more than " + std::to_string(kMaxSpareSlots) +
                      " spare path slots per batch: lower tuning.batch_paths");
    S = (uint32_t)want;
  }
  *S_out = S;
  const uint32_t Pcmax = Pmax + S;
  // every scale 1 in every job: the kernels skip the scale products
  bool unit = true;
  for (const JobIn& j : jobs) {
    for (float v : j.c->lscale_h) unit = unit && v == 1.0f;
    for (float v : j.c->ascale_h) unit = unit && v == 1.0f;
  }
  SHB_TRY(x, x.d_maps.reserve(maps.size() * sizeof(BatchMap)));
  SHB_TRY(x, hipMemcpyAsync(x.d_maps.ptr, maps.data(), maps.size() * sizeof(BatchMap),
                            hipMemcpyHostToDevice, st));

  // light samples per NEE: the largest of the jobs' (per-pick arrays)
  uint32_t nls = 1;
  for (const JobIn& j : jobs) nls = std::max(nls, (uint32_t)j.cfg.num_light_samples);
  if (adj && nls > 1)
    return fail(x, SHB_ERR_UNSUPPORTED, "the adjoint takes one light sample per NEE");
  // path state per slot, Pc = P + S slots: extension ray 2, hit, throughput,
  // radiance, texel NEE value and direction nls each, primary direction,
  // shadow rays 4 nls (vertex and texel, 2 float4 each), vertex NEE
  // contribution nls: 6 + 7 nls float4 (13 with one light sample; +1 for the
  // adjoint's projection weights). Queues: two extension queues of Pc, the
  // shadow queue of nls (Pc + P) (vertex and texel rays); branch links.
  const size_t f4_per_slot = adj ? 14 : 6 + 7 * (size_t)nls;
  SHB_TRY(x,
          x.d_paths.reserve(std::max<size_t>((size_t)Pcmax * f4_per_slot * sizeof(float4), 16)));
  uint32_t maxB_all = 0;
  for (uint32_t v : maxB) maxB_all = std::max(maxB_all, v);
  // adjoint: vertex materials (u16 per depth and slot), visible flags
  const size_t hist_bytes = ((size_t)(maxB_all + 1) * Pcmax * 2 + 15) / 16 * 16;
  if (adj) SHB_TRY(x, x.d_adj.reserve(hist_bytes + (size_t)Pcmax * 4 + 16));
  SHB_TRY(x, x.d_queues.reserve(std::max<size_t>(
                 ((size_t)Pcmax * 2 + (size_t)nls * (Pcmax + Pmax)) * sizeof(uint32_t), 16)));
  if (S) SHB_TRY(x, x.d_link.reserve((size_t)S * sizeof(uint4)));
  SHB_TRY(x, x.d_counters.reserve(std::max<size_t>((size_t)nb * kCntStride * 4, 64)));
  SHB_TRY(x, x.d_work.reserve(kWorkBytes));
  float4* base = x.d_paths.as<float4>();
  uint32_t* q = x.d_queues.as<uint32_t>();
  unsigned long long* totals = reinterpret_cast<unsigned long long*>(x.d_work.as<uint32_t>() + 4);
  unsigned long long* totals_sh = totals + 128;
  // one thread per path slot for the shading stages (their loops stride by
  // the grid): a grid of 16 blocks per CU looping over the slots ran the
  // bake 3 % slower
  const int sgrid = std::max(1, (int)(((uint64_t)Pmax + 255) / 256));
  const int mgrid = std::max(1, std::min<int>((int)((S + 255) / 256), x.num_cus * 16));

  TraceArgs tbase{};
  tbase.scene.nodes = tnodes;
  tbase.scene.tris = ttris;
  tbase.scene.has_nodes = total_nodes ? 1 : 0;
  tbase.has_nodes = total_nodes ? 1 : 0;
  tbase.wave_tris = total_tris < kWaveTrisMaxRecords ? 1u : 0u;
  tbase.max_steps = (uint32_t)total_nodes + 1;
  tbase.stack_entries = (uint32_t)std::max(1, std::min(depth - 1, kStackLds));
  tbase.refill = (uint32_t)x.tuning.queue_chunk;
  tbase.num_cus = x.num_cus;
  tbase.error = x.d_work.as<uint32_t>();
  tbase.nls = nls;

  // event pairs around each traversal launch; kind 0 extension, 1 shadow
  std::vector<hipEvent_t> tev;
  std::vector<int> tkind;
  const bool time_traces = !graph;
  auto trace_ev = [&](int kind) {
    if (!time_traces) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) e = nullptr;
    tev.push_back(e);
    tkind.push_back(kind);
    if (e) (void)hipEventRecord(e, st);
  };

  // the launch sequence of the whole bake
  auto record = [&](Rec& r) -> int {
    auto fill = [&](uint32_t* p, size_t n) -> hipError_t {
      if (r.dry) {
        r.add(1, p);
        r.add(1, n);
        return hipSuccess;
      }
      return launch_fill_u32(p, 0u, n, st);
    };
    SHB_TRY(x, fill(x.d_counters.as<uint32_t>(), (size_t)nb * kCntStride));
    SHB_TRY(x, fill(x.d_work.as<uint32_t>(), kWorkBytes / 4));
    if (adj)
      SHB_TRY(x, fill(reinterpret_cast<uint32_t*>(x.d_adj.as<uint8_t>() + hist_bytes), Pcmax));
    for (uint32_t bi = 0; bi < nb; ++bi) {
      const BatchMap& m = maps[bi];
      uint32_t* cnt = x.d_counters.as<uint32_t>() + (size_t)bi * kCntStride;
      const uint32_t P = m.slot_base[m.nseg];
      const uint32_t Pc = P + S;
      BakeArgs b{};
      b.jobs = x.d_jobs.as<JobDev>();
      b.map = x.d_maps.as<BatchMap>() + bi;
      b.nseg = m.nseg;
      b.job0 = m.job[0];
      b.texel_begin0 = m.texel_begin[0];
      b.T0 = m.tex_base[1] - m.tex_base[0];
      b.T = m.tex_base[m.nseg];
      b.P = P;
      b.S = S;
      b.Pc = Pc;
      b.link = S ? x.d_link.as<uint4>() : nullptr;
      b.spawn_count = cnt + kCntSpawn;
      b.unit_scales = unit ? 1u : 0u;
      b.nls = nls;
      b.ray = base;
      b.hit = base + 2 * (size_t)Pc;
      b.thr = base + 3 * (size_t)Pc;
      b.li = base + 4 * (size_t)Pc;
      b.nee0 = base + 5 * (size_t)Pc;
      b.nee0_dir = base + (5 + (size_t)nls) * Pc;
      b.dir0 = base + (5 + 2 * (size_t)nls) * Pc;
      b.sray = base + (6 + 2 * (size_t)nls) * Pc;
      b.scontrib = base + (6 + 6 * (size_t)nls) * Pc;
      uint32_t* qext[2] = {q, q + Pc};
      b.shadow_queue = q + 2 * (size_t)Pc;
      b.shadow_count = cnt + kCntSh;
      b.resolved_count = cnt + kCntRes;
      if (adj) {
        b.adj_dsh = adj->dsh;
        b.adj_grad = adj->grad;
        b.adj_nlights = adj->nlights;
        b.adj_nparam = adj->nparam;
        b.adj_gvec = base + 13 * (size_t)Pc;
        b.adj_hist = x.d_adj.as<uint16_t>();
        b.adj_vflag = reinterpret_cast<uint32_t*>(x.d_adj.as<uint8_t>() + hist_bytes);
      }
      TraceArgs tb = tbase;
      tb.map = b.map;
      tb.nseg = m.nseg;
      tb.root0 = m.root[0];
      tb.P = P;
      tb.Pc = Pc;
      tb.link = b.link;
      if (r.dry) r.add(2, b);
      else SHB_TRY(x, launch_generate(b, sgrid, st));
      if (!r.dry) SHB_TRY(x, stage_done(st, "generate", bi, -1));
      for (uint32_t d = 0; d <= maxB[bi]; ++d) {
        // this bounce's extension queue (depth 0 traces the primary rays in
        // slot order)
        const uint32_t* ext_q = d == 0 ? nullptr : qext[d & 1];
        TraceArgs ta = tb;
        ta.rays = b.ray;
        ta.queue = ext_q;
        ta.count_ptr = d == 0 ? nullptr : cnt + kCntExt + d;
        ta.count_const = P;
        ta.totals = count ? totals : nullptr;
        ta.head = ta.refill ? cnt + kCntHeadExt + d : nullptr;
        if (r.dry) {
          r.add(3, ta);
        } else {
          trace_ev(0);
          SHB_TRY(x, launch_trace_hits(ta, b.hit, count, Pc, x.tuning.block_size, st));
          trace_ev(0);
          ++run->launches_ext;
          SHB_TRY(x, stage_done(st, "extend", bi, (int)d));
        }
        b.shadow_count = cnt + kCntSh + d;
        b.resolved_count = cnt + kCntRes + d;
        if (r.dry) {
          r.add(4, b);
          r.add(4, d);
        } else {
          SHB_TRY(x, launch_shade(b, (int)d, ext_q, cnt + kCntExt + d, qext[(d + 1) & 1],
                                  cnt + kCntExt + d + 1, sgrid, st));
          SHB_TRY(x, stage_done(st, "shade", bi, (int)d));
        }
        TraceArgs sa = tb;
        sa.rays = b.sray;
        sa.queue = b.shadow_queue;
        sa.count_ptr = cnt + kCntSh + d;
        sa.totals = count ? totals_sh : nullptr;
        sa.head = sa.refill ? cnt + kCntHeadSh + d : nullptr;
        if (r.dry) {
          r.add(5, sa);
        } else {
          trace_ev(1);
          const uint32_t sh_cap = nls * (Pc + P);
          if (adj)
            SHB_TRY(x,
                    launch_trace_bake_shadow_adj(sa, b, count, sh_cap, x.tuning.block_size, st));
          else if (nls > 1)
            SHB_TRY(x, launch_trace_bake_shadow_multi(sa, b, count, sh_cap, x.tuning.block_size,
                                                      st));
          else
            SHB_TRY(x, launch_trace_bake_shadow(sa, b, count, sh_cap, x.tuning.block_size, st));
          trace_ev(1);
          ++run->launches_sh;
          SHB_TRY(x, stage_done(st, "shadow", bi, (int)d));
          if (nls > 1) SHB_TRY(x, launch_nee_gather(b, ext_q, cnt + kCntExt + d, sgrid, st));
          if (adj) SHB_TRY(x, launch_adjoint_vertex(b, (int)d, cnt + kCntSh + d, sgrid, st));
        }
      }
      // branch radiance into the parents, deepest spawn depth first (splits
      // happen above the last bounce only); the adjoint needs no radiance
      if (S && !adj)
        for (int d = (int)maxB[bi] - 1; d >= 0; --d) {
          if (r.dry) {
            r.add(7, b);
            r.add(7, d);
          } else {
            SHB_TRY(x, launch_merge_branches(b, d, mgrid, st));
          }
        }
      if (r.dry) r.add(6, b);
      else {
        SHB_TRY(x, launch_project(b, st));
        SHB_TRY(x, stage_done(st, "project", bi, -1));
      }
    }
    return SHB_OK;
  };

  hipEvent_t e0, e1;
  SHB_TRY(x, hipEventCreate(&e0));
  SHB_TRY(x, hipEventCreate(&e1));
  SHB_TRY(x, hipEventRecord(e0, st));
  int rc = SHB_OK;
  if (graph && !debug_stages()) {
    std::string key;
    Rec dry;
    dry.dry = true;
    dry.key = &key;
    rc = record(dry);
    if (rc) return rc;
    if (!x.graph_exec || key != x.graph_key) {
      if (x.graph_exec) {
        (void)hipGraphExecDestroy(x.graph_exec);
        x.graph_exec = nullptr;
        x.graph_key.clear();
      }
      SHB_TRY(x, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      Rec live;
      rc = record(live);
      hipGraph_t g = nullptr;
      hipError_t ce = hipStreamEndCapture(st, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      SHB_TRY(x, ce);
      hipError_t ie = hipGraphInstantiate(&x.graph_exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      SHB_TRY(x, ie);
      x.graph_key = key;
      run->graph_captured = true;
    }
    SHB_TRY(x, hipGraphLaunch(x.graph_exec, st));
    // launch counts as recorded
    for (uint32_t bi = 0; bi < nb; ++bi) {
      run->launches_ext += maxB[bi] + 1;
      run->launches_sh += maxB[bi] + 1;
    }
  } else {
    Rec live;
    rc = record(live);
    if (rc) return rc;
  }
  SHB_TRY(x, hipEventRecord(e1, st));
  SHB_TRY(x, hipEventSynchronize(e1));
  float ms = 0.f;
  SHB_TRY(x, hipEventElapsedTime(&ms, e0, e1));
  run->dev_s = ms * 1e-3;
  double tr = 0, tr_sh = 0;
  for (size_t i = 0; i + 1 < tev.size(); i += 2) {
    float m = 0.f;
    if (tev[i] && tev[i + 1] && hipEventElapsedTime(&m, tev[i], tev[i + 1]) == hipSuccess) {
      tr += m;
      if (tkind[i] == 1) tr_sh += m;
    }
  }
  for (hipEvent_t e : tev)
    if (e) (void)hipEventDestroy(e);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  run->trace_s = tr * 1e-3;
  run->trace_sh_s = tr_sh * 1e-3;
  // ray counts
  std::vector<uint32_t> hc((size_t)nb * kCntStride);
  if (nb)
    SHB_TRY(x, hipMemcpy(hc.data(), x.d_counters.ptr, hc.size() * 4, hipMemcpyDeviceToHost));
  *demand = 0;
  for (uint32_t bi = 0; bi < nb; ++bi) {
    const uint32_t* k = hc.data() + (size_t)bi * kCntStride;
    *demand = std::max(*demand, k[kCntSpawn]);
    run->branches += std::min(k[kCntSpawn], S);
    run->rays_ext += maps[bi].slot_base[maps[bi].nseg];
    for (uint32_t d = 1; d <= maxB[bi]; ++d) run->rays_ext += k[kCntExt + d];
    for (uint32_t d = 0; d <= maxB[bi]; ++d) run->rays_sh += k[kCntSh + d];
    for (uint32_t d = 0; d <= maxB[bi]; ++d) run->rays_sh_res += k[kCntRes + d];
  }
  std::vector<uint32_t> hw(kWorkBytes / 4);
  SHB_TRY(x, hipMemcpy(hw.data(), x.d_work.ptr, kWorkBytes, hipMemcpyDeviceToHost));
  if (hw[0]) return check_overflow(x, total_nodes + 1);
  if (count) {
    const unsigned long long* t2 = reinterpret_cast<const unsigned long long*>(hw.data() + 4);
    for (int k = 0; k < 64; ++k) {
      run->node_visits += t2[2 * k] + t2[128 + 2 * k];
      run->tri_tests += t2[2 * k + 1] + t2[128 + 2 * k + 1];
      run->node_visits_sh += t2[128 + 2 * k];
      run->tri_tests_sh += t2[128 + 2 * k + 1];
    }
  }
  return SHB_OK;
}

}  // namespace

int run_bake(Ctx& x, std::vector<JobIn>& jobs, bool graph, BakeRun* run, const AdjIn* adj) {
  if (adj && jobs.size() != 1) return fail(x, SHB_ERR_INVALID_ARG, "the adjoint runs one job");
  uint64_t want = 0;
  for (uint32_t pass = 1;; ++pass) {
    BakeRun r;
    uint32_t S = 0, demand = 0;
    const int rc = run_pass(x, jobs, graph, &r, adj, want, &S, &demand);
    if (rc) return rc;
    r.passes = pass;
    *run = r;
    if (demand <= S) return SHB_OK;
    if (S == 0)  // splits need a material with alpha < 255, which sets S
      return fail(x, SHB_ERR_DEVICE, "alpha split in a scene without transparent texels");
    // Some transmitted rays found no spare slot (and their own splits were
    // never asked for): bake again with room for more than the demand seen.
    // Every path depends only on its counters, so the pass that fits gives
    // the result any larger pool would.
    want = std::max<uint64_t>(2ull * S, (uint64_t)demand + demand / 4);
    x.spare_slots = (uint32_t)std::min<uint64_t>(want, kMaxSpareSlots);
    if (adj) SHB_TRY(x, launch_fill_u32(reinterpret_cast<uint32_t*>(adj->grad), 0u,
                                        2 * (size_t)adj->nparam, x.stream));
  }
}

}  // namespace shb
