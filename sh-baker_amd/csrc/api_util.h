// Internal helpers shared by the C ABI sources (not part of the ABI).
#ifndef SHB_API_UTIL_H_
#define SHB_API_UTIL_H_

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "shb_ctx.h"

namespace shb {

#define SHB_TRY(ctx, expr)                                                   \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess) {                                                  \
      (ctx).err = std::string(#expr) + ": " + hipGetErrorString(_e);         \
      return _e == hipErrorOutOfMemory ? SHB_ERR_OOM : SHB_ERR_DEVICE;       \
    }                                                                        \
  } while (0)

inline int fail(Ctx& c, int code, const std::string& msg) {
  c.err = msg;
  return code;
}

template <class T>
hipError_t upload(DevBuf& b, const std::vector<T>& v, hipStream_t st) {
  size_t n = std::max<size_t>(v.size() * sizeof(T), 16);
  hipError_t e = b.reserve(n);
  if (e != hipSuccess) return e;
  if (!v.empty())
    return hipMemcpyAsync(b.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st);
  return hipSuccess;
}

inline double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

inline TraceArgs trace_args(Ctx& c) {
  TraceArgs a{};
  a.scene = c.view();
  a.has_nodes = c.bvh.nodes.empty() ? 0 : 1;
  a.max_steps = (uint32_t)c.bvh.nodes.size() + 1;
  a.stack_entries = (uint32_t)std::max(1, std::min(c.bvh.max_depth - 1, kStackLds));
  a.refill = (uint32_t)c.tuning.queue_chunk;
  a.num_cus = c.num_cus;
  a.head = nullptr;  // set per launch when the persistent form is used
  a.error = c.d_work.as<uint32_t>();
  a.map = nullptr;   // one tree, every ray starts at its root
  a.nseg = 1;
  a.root0 = c.bvh.nodes.empty() ? kNoRoot : 0u;
  a.P = 0xffffffffu;
  a.Pc = 0xffffffffu;
  a.link = nullptr;
  a.wave_tris = c.bvh.tris.size() / 3 < kWaveTrisMaxRecords ? 1u : 0u;
  return a;
}

// Reads the traversal error word; `cap` is the step cap of the launches.
inline int check_overflow(Ctx& c, size_t cap) {
  uint32_t e[4] = {0, 0, 0, 0};
  SHB_TRY(c, hipMemcpyAsync(e, c.d_work.ptr, sizeof(e), hipMemcpyDeviceToHost, c.stream));
  SHB_TRY(c, hipStreamSynchronize(c.stream));
  if (e[0] == 0) return SHB_OK;
  std::string what = (e[0] & 1u) ? "traversal stack overflow" : "";
  if (e[0] & 2u) what += std::string(what.empty() ? "" : " and ") + "traversal step cap reached";
  return fail(c, SHB_ERR_LIMIT,
              what + " (first ray slot " + std::to_string((long long)e[1] - 1) + ", stack depth " +
                  std::to_string(e[2]) + ", " + std::to_string(e[3]) + " steps, cap " +
                  std::to_string(cap) + ")");
}

// Counter layout per batch (uint32): extension-queue counts by depth, shadow
// counts by depth, queue heads of the persistent extension and shadow
// traversals by depth, branch slots asked for by alpha < 1 splits, shadow
// queries answered without traversal by depth.
constexpr int kCntExt = 0, kCntSh = 20, kCntHeadExt = 40, kCntHeadSh = 56, kCntSpawn = 72,
              kCntRes = 76, kCntStride = 96;

// SHB_DEBUG_STAGES=1: synchronise after every bake stage and log its time to
// stderr (diagnostics; off by default).
inline bool debug_stages() {
  static int v = -1;
  if (v < 0) {
    const char* e = getenv("SHB_DEBUG_STAGES");
    v = (e && e[0] == '1') ? 1 : 0;
  }
  return v == 1;
}

inline hipError_t stage_done(hipStream_t st, const char* name, uint32_t batch, int depth) {
  if (!debug_stages()) return hipSuccess;
  static double last = 0;
  hipError_t e = hipStreamSynchronize(st);
  double t = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  fprintf(stderr, "[shb] batch %u depth %d %-8s %.3f ms (%s)\n", batch, depth, name,
          last > 0 ? (t - last) * 1e3 : 0.0, hipGetErrorString(e));
  fflush(stderr);
  last = t;
  return e;
}

// d_work layout: [0] error bits, [1..3] first failing ray (slot + 1, stack
// depth, steps); then 64 pairs of u64 visit totals of the extension
// traversals, then 64 pairs of the shadow traversals.
constexpr size_t kWorkBytes = 16 + 2 * 64 * 16;

struct BakeRun {
  uint64_t rays_ext = 0, rays_sh = 0, rays_sh_res = 0, valid = 0, node_visits = 0,
           tri_tests = 0;
  uint64_t node_visits_sh = 0, tri_tests_sh = 0;
  double dev_s = 0, trace_s = 0, trace_sh_s = 0;
  uint32_t launches_ext = 0, launches_sh = 0;
  uint64_t branches = 0;
  uint32_t passes = 0;
  bool graph_captured = false;
};

inline void fill_stats(shb_bake_stats* s, const BakeRun& r, uint32_t spp, double t0) {
  if (!s) return;
  s->valid_texels = r.valid;
  s->paths = r.valid * spp;
  s->rays_extension = r.rays_ext;
  s->rays_shadow = r.rays_sh + r.rays_sh_res;
  s->rays_shadow_resolved = r.rays_sh_res;
  s->node_visits = r.node_visits;
  s->tri_tests = r.tri_tests;
  s->seconds_total = now_s() - t0;
  s->seconds_device = r.dev_s;
  s->seconds_trace = r.trace_s;
  s->node_visits_shadow = r.node_visits_sh;
  s->tri_tests_shadow = r.tri_tests_sh;
  s->seconds_trace_shadow = r.trace_sh_s;
  s->launches_extension = r.launches_ext;
  s->branches = r.branches;
  s->passes = r.passes;
  s->reserved = 0;
  s->launches_shadow = r.launches_sh;
}


}  // namespace shb

#endif  // SHB_API_UTIL_H_
