// Host BVH builder (see bvh_build.h).
//
// 1. Binary BVH by binned SAH (32 bins per axis on centroids, leaves of at
//    most 3 triangles), built in parallel: large ranges binned and
//    partitioned (stably) by all threads, the rest as subtrees by worker
//    threads; the tree is the same for any thread count.
// 2. Collapse into nodes of up to `width` children by the surface-area cost
//    optimisation of Ylitie, Karras, Laine (HPG 2017, section 3.1): for every
//    binary node n and slot budget i <= width, the cheapest way to represent
//    n's subtree in at most i child slots of a wide node -- as one leaf (at
//    most 3 triangles), as one wide internal node, or split between its two
//    children -- with cost area x (c_node per internal node, c_tri per
//    triangle test). A greedy collapse (open the largest child) left most
//    nodes half empty (262k-triangle synthetic room: 40 % of the nodes with
//    2 children).
// 3. Each node's children are placed in octant slots: slot s stands for the
//    direction with sign bit a set when the child lies on the negative side of
//    axis a from the parent centre, assigned greedily by the best dot product
//    of (child centre - parent centre) with the slot's sign vector.
// 4. Child boxes are quantised outward to 8 bits per plane on a power-of-two
//    grid anchored at the node's lower corner; every decoded box is checked in
//    double precision to contain the exact child box.
#include "bvh_build.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <deque>
#include <thread>

namespace shb {
namespace {

struct Box {
  float lo[3] = {INFINITY, INFINITY, INFINITY};
  float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  void grow(const Box& b) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], b.lo[a]);
      hi[a] = std::max(hi[a], b.hi[a]);
    }
  }
  float area() const {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (!(dx >= 0.f)) return 0.f;
    return dx * dy + dy * dz + dz * dx;
  }
};

struct B2Node {
  Box box;
  int left = -1, right = -1;
  uint32_t first = 0, count = 0;  // leaf range in `order`
  bool leaf() const { return left < 0; }
};

// A triangle reference: its box and id, 32 bytes, partitioned in place so
// the binning passes read memory in order.
struct Ref {
  float lo[3];
  float hi[3];
  uint32_t id;
  uint32_t pad;
  float c(int a) const { return 0.5f * (lo[a] + hi[a]); }
};

struct Builder {
  std::vector<Ref> refs;
  std::vector<Ref> scratch;  // the parallel partition's buffer (refs.size())
  std::vector<uint32_t> order;  // triangle ids in leaf order (refs[k].id)
  std::vector<B2Node> nodes;

  struct Item {
    int node;  // index in the node array it is built into
    uint32_t first, count;
    int depth;  // binary depth (root 0)
  };
  // SAH splits above this binary depth; deeper ranges split at their object
  // median, which bounds the tree's depth (kept deterministic: stable sort)
  int sah_depth = INT_MAX;

  // Bounds of item `it`, stored in nodes[it.node]; unless it becomes a leaf,
  // partitions its range of `order` by the binned-SAH split and appends its
  // two children to `nodes` and `work`.
  static constexpr int NB = 32;
  struct Bins {
    uint32_t cnt[3][NB];
    Box bb[3][NB];
  };

  // Bounds and centroid bounds of order[first, first + count) (min / max:
  // the result does not depend on how the range is cut into chunks).
  static void grow_ref(Box& b, const Ref& r) {
    for (int a = 0; a < 3; ++a) {
      b.lo[a] = std::min(b.lo[a], r.lo[a]);
      b.hi[a] = std::max(b.hi[a], r.hi[a]);
    }
  }
  void bounds(uint32_t first, uint32_t count, Box& box, Box& cbox) const {
    for (uint32_t k = first; k < first + count; ++k) {
      const Ref& r = refs[k];
      grow_ref(box, r);
      for (int a = 0; a < 3; ++a) {
        float c = r.c(a);
        cbox.lo[a] = std::min(cbox.lo[a], c);
        cbox.hi[a] = std::max(cbox.hi[a], c);
      }
    }
  }
  void bin(uint32_t first, uint32_t count, const Box& cbox, Bins& B) const {
    float kscale[3];
    bool live[3];
    for (int a = 0; a < 3; ++a) {
      for (int b = 0; b < NB; ++b) B.cnt[a][b] = 0;
      float span = cbox.hi[a] - cbox.lo[a];
      live[a] = span > 0.f;
      kscale[a] = live[a] ? NB / span : 0.f;
    }
    for (uint32_t k = first; k < first + count; ++k) {
      const Ref& r = refs[k];
      for (int a = 0; a < 3; ++a) {
        if (!live[a]) continue;
        int b = std::min(NB - 1, (int)((r.c(a) - cbox.lo[a]) * kscale[a]));
        B.cnt[a][b]++;
        grow_ref(B.bb[a][b], r);
      }
    }
  }

  // Bounds and bins of a range, cut into `threads` chunks when large.
  void bounds_and_bins(const Item& it, int threads, Box& box, Box& cbox, Bins& B) const {
    if (threads <= 1) {
      bounds(it.first, it.count, box, cbox);
      bin(it.first, it.count, cbox, B);
      return;
    }
    const uint32_t chunk = (it.count + threads - 1) / threads;
    std::vector<Box> bx(threads), cb(threads);
    std::vector<Bins> bs(threads);
    auto run = [&](auto&& f) {
      std::vector<std::thread> pool;
      for (int t = 0; t < threads; ++t) {
        const uint32_t f0 = it.first + std::min(it.count, t * chunk);
        const uint32_t f1 = it.first + std::min(it.count, (t + 1) * chunk);
        pool.emplace_back([&, t, f0, f1]() { f(t, f0, f1 - f0); });
      }
      for (auto& th : pool) th.join();
    };
    run([&](int t, uint32_t f0, uint32_t c) { bounds(f0, c, bx[t], cb[t]); });
    for (int t = 0; t < threads; ++t) {
      box.grow(bx[t]);
      cbox.grow(cb[t]);
    }
    run([&](int t, uint32_t f0, uint32_t c) { bin(f0, c, cbox, bs[t]); });
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < NB; ++b) {
        B.cnt[a][b] = 0;
        for (int t = 0; t < threads; ++t) {
          B.cnt[a][b] += bs[t].cnt[a][b];
          B.bb[a][b].grow(bs[t].bb[a][b]);
        }
      }
  }

  // Stable partition of refs[it.first, +count) by `left`, cut into
  // `threads` chunks when threads > 1 (counts, offsets, scatter through a
  // buffer): the same order for any thread count. Returns the left count.
  template <class Pred>
  uint32_t stable_split(const Item& it, int threads, Pred left) {
    Ref* base = refs.data() + it.first;
    if (threads <= 1) {
      thread_local std::vector<Ref> tmp;
      tmp.clear();
      uint32_t l = 0;
      for (uint32_t k = 0; k < it.count; ++k) {
        if (left(base[k]))
          base[l++] = base[k];
        else
          tmp.push_back(base[k]);
      }
      std::copy(tmp.begin(), tmp.end(), base + l);
      return l;
    }
    const uint32_t chunk = (it.count + threads - 1) / threads;
    std::vector<uint32_t> nl(threads, 0), offl(threads), offr(threads);
    Ref* buf = scratch.data() + it.first;  // this range's part of the buffer
    auto run = [&](auto&& f) {
      std::vector<std::thread> pool;
      for (int t = 0; t < threads; ++t) pool.emplace_back([&, t]() { f(t); });
      for (auto& th : pool) th.join();
    };
    auto range = [&](int t, uint32_t& a, uint32_t& b) {
      a = std::min(it.count, t * chunk);
      b = std::min(it.count, (t + 1) * chunk);
    };
    run([&](int t) {
      uint32_t a, b;
      range(t, a, b);
      uint32_t c = 0;
      for (uint32_t k = a; k < b; ++k) c += left(base[k]) ? 1u : 0u;
      nl[t] = c;
    });
    uint32_t L = 0;
    for (int t = 0; t < threads; ++t) {
      offl[t] = L;
      L += nl[t];
    }
    uint32_t R = L;
    for (int t = 0; t < threads; ++t) {
      uint32_t a, b;
      range(t, a, b);
      offr[t] = R;
      R += (b - a) - nl[t];
    }
    run([&](int t) {
      uint32_t a, b;
      range(t, a, b);
      uint32_t pl = offl[t], pr = offr[t];
      for (uint32_t k = a; k < b; ++k) buf[left(base[k]) ? pl++ : pr++] = base[k];
    });
    run([&](int t) {
      uint32_t a, b;
      range(t, a, b);
      std::copy(buf + a, buf + b, base + a);
    });
    return L;
  }

  void split_item(const Item& it, std::vector<B2Node>& nodes_out, std::vector<Item>& work,
                  int threads = 1) {
    Box box, cbox;
    Bins B;
    bounds_and_bins(it, threads, box, cbox, B);
    nodes_out[it.node].box = box;
    nodes_out[it.node].first = it.first;
    nodes_out[it.node].count = it.count;
    if (it.count <= 1) return;
    if (it.depth >= sah_depth) {
      // depth budget spent: halve the range along its widest centroid axis
      if (it.count <= 3) return;  // a leaf
      int ax = 0;
      for (int a = 1; a < 3; ++a)
        if (cbox.hi[a] - cbox.lo[a] > cbox.hi[ax] - cbox.lo[ax]) ax = a;
      std::stable_sort(refs.begin() + it.first, refs.begin() + it.first + it.count,
                       [ax](const Ref& p, const Ref& q) { return p.c(ax) < q.c(ax); });
      const uint32_t mid = it.first + it.count / 2;
      int l = (int)nodes_out.size();
      nodes_out.push_back(B2Node());
      int r = (int)nodes_out.size();
      nodes_out.push_back(B2Node());
      nodes_out[it.node].left = l;
      nodes_out[it.node].right = r;
      work.push_back({r, mid, it.first + it.count - mid, it.depth + 1});
      work.push_back({l, it.first, mid - it.first, it.depth + 1});
      return;
    }
    int best_axis = -1, best_bin = -1;
    float best_cost = INFINITY;
    for (int a = 0; a < 3; ++a) {
      float span = cbox.hi[a] - cbox.lo[a];
      if (!(span > 0.f)) continue;
      const uint32_t* cnt = B.cnt[a];
      const Box* bb = B.bb[a];
      float la[NB];
      uint32_t lc[NB];
      Box acc;
      uint32_t c = 0;
      for (int b = 0; b < NB; ++b) {
        acc.grow(bb[b]);
        c += cnt[b];
        la[b] = acc.area();
        lc[b] = c;
      }
      acc = Box();
      c = 0;
      for (int b = NB - 1; b >= 1; --b) {
        acc.grow(bb[b]);
        c += cnt[b];
        if (lc[b - 1] == 0 || c == 0) continue;
        float cost = la[b - 1] * lc[b - 1] + acc.area() * c;
        if (cost < best_cost) {
          best_cost = cost;
          best_axis = a;
          best_bin = b - 1;
        }
      }
    }
    float leaf_cost = box.area() * it.count;
    // SAH with traversal cost ~ 1 triangle test per unit of parent area.
    if (it.count <= 3 && !(best_cost + box.area() < leaf_cost)) return;
    uint32_t mid;
    if (best_axis < 0) {
      mid = it.first + it.count / 2;
    } else {
      float span = cbox.hi[best_axis] - cbox.lo[best_axis];
      float kscale = NB / span;
      const float lo = cbox.lo[best_axis];
      mid = it.first + stable_split(it, threads, [&](const Ref& r) {
              int b = std::min(NB - 1, (int)((r.c(best_axis) - lo) * kscale));
              return b <= best_bin;
            });
      if (mid == it.first || mid == it.first + it.count) mid = it.first + it.count / 2;
    }
    int l = (int)nodes_out.size();
    nodes_out.push_back(B2Node());
    int r = (int)nodes_out.size();
    nodes_out.push_back(B2Node());
    nodes_out[it.node].left = l;
    nodes_out[it.node].right = r;
    work.push_back({r, mid, it.first + it.count - mid, it.depth + 1});
    work.push_back({l, it.first, mid - it.first, it.depth + 1});
  }

  // Top-down build. The top levels are split here until the open ranges
  // are small or numerous enough; the remaining subtrees are built by
  // worker threads into their own node arrays and appended (children
  // always after their parent). Every range sees the same permutation of
  // `order` as in a serial build, so the tree does not depend on the
  // thread count.
  void build(const float* xyz, uint32_t n) {
    refs.resize(n);
    scratch.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
      const float* t = xyz + 9 * (size_t)i;
      Ref& r = refs[i];
      for (int a = 0; a < 3; ++a) {
        r.lo[a] = INFINITY;
        r.hi[a] = -INFINITY;
      }
      for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 3; ++a) {
          r.lo[a] = std::min(r.lo[a], t[3 * k + a]);
          r.hi[a] = std::max(r.hi[a], t[3 * k + a]);
        }
      r.id = i;
      r.pad = 0;
    }
    nodes.reserve(2 * (size_t)n + 1);
    nodes.push_back(B2Node());
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned T = std::min(16u, hw);
    // ranges below this many triangles go to the subtree workers (at least
    // 4 ranges per worker); larger ranges are split here, their bins made
    // by all threads
    const uint32_t hand_off = std::max<uint32_t>(1024u, n / (4u * T));
    const uint32_t kParallelBins = 1u << 16;
    std::vector<Item> work{{0, 0, n, 0}}, open;
    while (!work.empty()) {
      Item it = work.back();
      work.pop_back();
      if (T > 1 && it.count < hand_off) {
        open.push_back(it);
        continue;
      }
      split_item(it, nodes, work, T > 1 && it.count >= kParallelBins ? (int)T : 1);
    }
    if (open.empty()) {
      finish_order();
      return;
    }
    std::sort(open.begin(), open.end(),
              [](const Item& a, const Item& b) { return a.count > b.count; });
    std::vector<std::vector<B2Node>> local(open.size());
    std::atomic<size_t> next{0};
    auto worker = [&]() {
      for (;;) {
        const size_t j = next.fetch_add(1);
        if (j >= open.size()) return;
        std::vector<B2Node>& ln = local[j];
        ln.reserve(2 * (size_t)open[j].count);
        ln.push_back(B2Node());
        std::vector<Item> w{{0, open[j].first, open[j].count, open[j].depth}};
        while (!w.empty()) {
          Item it = w.back();
          w.pop_back();
          split_item(it, ln, w);
        }
      }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < T; ++t) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
    for (size_t j = 0; j < open.size(); ++j) {
      const std::vector<B2Node>& ln = local[j];
      const int base = (int)nodes.size() - 1;  // local index k >= 1 -> base + k
      auto remap = [&](int k) { return k < 0 ? -1 : (k == 0 ? open[j].node : base + k); };
      B2Node root = ln[0];
      root.left = remap(root.left);
      root.right = remap(root.right);
      nodes[open[j].node] = root;
      for (size_t k = 1; k < ln.size(); ++k) {
        B2Node nd = ln[k];
        nd.left = remap(nd.left);
        nd.right = remap(nd.right);
        nodes.push_back(nd);
      }
    }
    finish_order();
  }

  void finish_order() {
    order.resize(refs.size());
    for (size_t k = 0; k < refs.size(); ++k) order[k] = refs[k].id;
  }
};

// Outward 8-bit quantisation of [clo, chi] on the grid p + q * 2^e.
void quantize_axis(float p, int e, float clo, float chi, uint8_t* qlo, uint8_t* qhi) {
  double s = std::ldexp(1.0, e);
  double lo = std::floor(((double)clo - (double)p) / s);
  double hi = std::ceil(((double)chi - (double)p) / s);
  lo = std::max(0.0, std::min(255.0, lo));
  hi = std::max(0.0, std::min(255.0, hi));
  while (lo > 0 && (double)p + lo * s > (double)clo) lo -= 1;
  while (hi < 255 && (double)p + hi * s < (double)chi) hi += 1;
  *qlo = (uint8_t)lo;
  *qhi = (uint8_t)hi;
}

int axis_exponent(float lo, float hi) {
  double ext = (double)hi - (double)lo;
  if (!(ext > 0.0)) return -126;
  int e = (int)std::ceil(std::log2(ext / 255.0));
  while (std::ldexp(255.0, e) < ext) ++e;
  return std::max(-126, std::min(127, e));
}

// The wide tree of binary tree b (see the file comment, steps 2-4).
void collapse(const Builder& b, const float* tri_xyz, uint32_t n, int width, BvhBuildOut* out) {

  // ---- cost-optimal collapse (see the file comment)
  // cost[n * W + i - 1]: subtree n in at most i slots; how[...]: 0 leaf,
  // 1 wide internal node, 2 split (children get k and i - k slots, k in
  // split[...]), 3 as with i - 1 slots.
  const int W = width;
  const size_t N2 = b.nodes.size();
  constexpr float kNodeCost = 1.0f, kTriCost = 0.35f;
  std::vector<float> cost(N2 * W);
  std::vector<uint8_t> how(N2 * W), split(N2 * W);
  std::vector<uint8_t> wide_split(N2, 1);  // as a wide node: left child's share of W slots
  {
    // children after parents in `nodes`: a reverse sweep visits children first
    for (size_t ni = N2; ni-- > 0;) {
      const B2Node& nd = b.nodes[ni];
      const float A = nd.box.area();
      float* c = &cost[ni * W];
      uint8_t* h = &how[ni * W];
      uint8_t* sp = &split[ni * W];
      if (nd.leaf()) {
        for (int i = 0; i < W; ++i) {
          c[i] = A * kTriCost * (float)nd.count;
          h[i] = 0;
          sp[i] = 0;
        }
        continue;
      }
      const float* cl = &cost[(size_t)nd.left * W];
      const float* cr = &cost[(size_t)nd.right * W];
      // distribute over j slots (j >= 2): best k for the left child
      float dist[9];
      uint8_t dk[9];
      for (int j = 2; j <= W; ++j) {
        float best = INFINITY;
        int bk = 1;
        for (int k = 1; k < j; ++k) {
          float v = cl[k - 1] + cr[j - k - 1];
          if (v < best) {
            best = v;
            bk = k;
          }
        }
        dist[j] = best;
        dk[j] = (uint8_t)bk;
      }
      const float c_int = A * kNodeCost + dist[W];
      wide_split[ni] = dk[W];
      float c1 = c_int;
      uint8_t h1 = 1;
      if (nd.count <= 3) {
        const float c_leaf = A * kTriCost * (float)nd.count;
        if (c_leaf <= c_int) {
          c1 = c_leaf;
          h1 = 0;
        }
      }
      c[0] = c1;
      h[0] = h1;
      sp[0] = 0;
      for (int i = 2; i <= W; ++i) {
        if (dist[i] < c[i - 2]) {
          c[i - 1] = dist[i];
          h[i - 1] = 2;
          sp[i - 1] = dk[i];
        } else {
          c[i - 1] = c[i - 2];
          h[i - 1] = 3;
          sp[i - 1] = 0;
        }
      }
    }
  }
  // the slots of a wide internal node rooted at binary node n: its subtree
  // split over W slots; each part is a leaf (kid, true) or a wide node
  // (kid, false)
  // gather(n, i, ...) appends the parts of subtree n over at most i slots,
  // left subtree first (an explicit stack; at most 8 parts)
  auto gather = [&](int n0, int i0, int* kid, bool* leaf, int& m) {
    int st_n[16], st_i[16], sp = 0;
    st_n[sp] = n0;
    st_i[sp++] = i0;
    while (sp > 0) {
      const int n = st_n[--sp];
      int i = st_i[sp];
      const uint8_t* h = &how[(size_t)n * W];
      while (i > 1 && h[i - 1] == 3) --i;
      if (i == 1 || h[i - 1] != 2) {
        kid[m] = n;
        leaf[m] = h[0] == 0;
        ++m;
        continue;
      }
      const int k = split[(size_t)n * W + i - 1];
      st_n[sp] = b.nodes[n].right;  // popped after the left part
      st_i[sp++] = i - k;
      st_n[sp] = b.nodes[n].left;
      st_i[sp++] = k;
    }
  };

  struct Pending {
    int b2;       // binary node collapsed into this wide node
    uint32_t at;  // index in out->nodes
    int depth;
  };
  out->tris.reserve(3 * (size_t)n);
  out->nodes.reserve(N2 / 2 + 1);
  out->nodes.push_back(CWNode());
  std::deque<Pending> q;
  q.push_back({0, 0, 1});
  while (!q.empty()) {
    Pending pd = q.front();
    q.pop_front();
    out->max_depth = std::max(out->max_depth, pd.depth);
    int kids[8];
    bool kid_leaf[8];
    int m = 0;
    const B2Node& root = b.nodes[pd.b2];
    if (root.leaf()) {
      kids[0] = pd.b2;  // the whole tree is one leaf
      kid_leaf[0] = true;
      m = 1;
    } else {
      // a wide node: its subtree over W slots, the left child taking the
      // share that minimised the cost
      const int k = wide_split[pd.b2];
      gather(root.left, k, kids, kid_leaf, m);
      gather(root.right, W - k, kids, kid_leaf, m);
    }
    Box nb;
    for (int k = 0; k < m; ++k) nb.grow(b.nodes[kids[k]].box);
    // octant slot assignment
    float pc[3], cc[8][3];
    for (int a = 0; a < 3; ++a) pc[a] = 0.5f * (nb.lo[a] + nb.hi[a]);
    for (int k = 0; k < m; ++k)
      for (int a = 0; a < 3; ++a)
        cc[k][a] = 0.5f * (b.nodes[kids[k]].box.lo[a] + b.nodes[kids[k]].box.hi[a]) - pc[a];
    int slot_of[8];
    bool kid_done[8] = {false}, slot_used[8] = {false};
    float score[8][8];  // child k in slot s: (child centre - node centre) . slot sign vector
    for (int k = 0; k < m; ++k)
      for (int s = 0; s < 8; ++s) {
        float sc = 0.f;
        for (int a = 0; a < 3; ++a) sc += ((s >> a) & 1) ? -cc[k][a] : cc[k][a];
        score[k][s] = sc;
      }
    for (int round = 0; round < m; ++round) {
      float best = -INFINITY;
      int bk = -1, bs = -1;
      for (int k = 0; k < m; ++k) {
        if (kid_done[k]) continue;
        for (int s = 0; s < 8; ++s) {
          if (slot_used[s]) continue;
          const float sc = score[k][s];
          if (sc > best) {
            best = sc;
            bk = k;
            bs = s;
          }
        }
      }
      kid_done[bk] = true;
      slot_used[bs] = true;
      slot_of[bk] = bs;
    }
    int slot_kid[8];  // index into kids
    for (int s = 0; s < 8; ++s) slot_kid[s] = -1;
    for (int k = 0; k < m; ++k) slot_kid[slot_of[k]] = k;

    CWNode node;
    std::memset(&node, 0, sizeof(node));
    node.px = nb.lo[0];
    node.py = nb.lo[1];
    node.pz = nb.lo[2];
    int ex = axis_exponent(nb.lo[0], nb.hi[0]);
    int ey = axis_exponent(nb.lo[1], nb.hi[1]);
    int ez = axis_exponent(nb.lo[2], nb.hi[2]);
    node.ex = (int8_t)ex;
    node.ey = (int8_t)ey;
    node.ez = (int8_t)ez;
    node.child_base = (uint32_t)out->nodes.size();
    node.tri_base = (uint32_t)(out->tris.size() / 3);
    uint32_t rank = 0, tri_off = 0;
    for (int s = 0; s < 8; ++s) {
      int k = slot_kid[s];
      if (k < 0) continue;
      const B2Node& c = b.nodes[kids[k]];
      quantize_axis(node.px, ex, c.box.lo[0], c.box.hi[0], &node.qlox[s], &node.qhix[s]);
      quantize_axis(node.py, ey, c.box.lo[1], c.box.hi[1], &node.qloy[s], &node.qhiy[s]);
      quantize_axis(node.pz, ez, c.box.lo[2], c.box.hi[2], &node.qloz[s], &node.qhiz[s]);
      // leaves hold at most 3 triangles (the collapse only makes those)
      bool as_leaf = kid_leaf[k];
      if (as_leaf) {
        node.meta[s] = (uint8_t)((c.count << 5) | tri_off);
        for (uint32_t t = 0; t < c.count; ++t) {
          uint32_t id = b.order[c.first + t];
          const float* v = tri_xyz + 9 * (size_t)id;
          out->tris.push_back(make_float4(v[0], v[1], v[2], as_float(id)));
          out->tris.push_back(make_float4(v[3], v[4], v[5], 0.f));
          out->tris.push_back(make_float4(v[6], v[7], v[8], 0.f));
        }
        tri_off += c.count;
      } else {
        node.meta[s] = (uint8_t)(0x80 | rank);
        node.imask |= (uint8_t)(1u << s);
        ++rank;
      }
    }
    // allocate internal children contiguously, in slot order
    for (int s = 0; s < 8; ++s) {
      int k = slot_kid[s];
      if (k < 0 || !(node.imask & (1u << s))) continue;
      uint32_t at = (uint32_t)out->nodes.size();
      out->nodes.push_back(CWNode());
      q.push_back({kids[k], at, pd.depth + 1});
    }
    out->nodes[pd.at] = node;
  }
}

}  // namespace

void build_cwbvh(const float* tri_xyz, uint32_t n, int width, BvhBuildOut* out) {
  auto t0 = std::chrono::steady_clock::now();
  width = std::max(2, std::min(8, width));
  // SAH everywhere first; when the wide tree comes out deeper than the
  // traversal stack allows, again with the SAH limited to the top levels
  // (object-median splits below), down to a median tree
  static const int kBudgets[] = {INT_MAX, 32, 24, 16, 12, 8, 4, 0};
  for (int k = 0;; ++k) {
    out->nodes.clear();
    out->tris.clear();
    out->max_depth = 0;
    out->sah_depth = kBudgets[k];
    if (n == 0) break;
    Builder b;
    b.sah_depth = kBudgets[k];
    b.build(tri_xyz, n);
    collapse(b, tri_xyz, n, width, out);
    if (out->max_depth <= kMaxBvhDepth || k + 1 == (int)(sizeof(kBudgets) / sizeof(int))) break;
  }
  auto t1 = std::chrono::steady_clock::now();
  out->seconds = std::chrono::duration<double>(t1 - t0).count();
}

}  // namespace shb
