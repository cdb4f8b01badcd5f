// Host BVH builder (see bvh_build.h).
//
// 1. Binary BVH by binned SAH (32 bins per axis on centroids, leaves of at
//    most 3 triangles).
// 2. Greedy collapse into nodes of up to `width` children: repeatedly open the
//    largest-area internal child (Ylitie et al. 2017 use an optimal DP; the
//    greedy form is simpler and close in practice).
// 3. Each node's children are placed in octant slots: slot s stands for the
//    direction with sign bit a set when the child lies on the negative side of
//    axis a from the parent centre, assigned greedily by the best dot product
//    of (child centre - parent centre) with the slot's sign vector.
// 4. Child boxes are quantised outward to 8 bits per plane on a power-of-two
//    grid anchored at the node's lower corner; every decoded box is checked in
//    double precision to contain the exact child box.
#include "bvh_build.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>

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

struct Builder {
  std::vector<Box> tbox;
  std::vector<float> cen;  // 3 per triangle
  std::vector<uint32_t> order;
  std::vector<B2Node> nodes;

  void build(const float* xyz, uint32_t n) {
    tbox.resize(n);
    cen.resize(3 * (size_t)n);
    order.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
      const float* t = xyz + 9 * (size_t)i;
      Box b;
      for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 3; ++a) {
          b.lo[a] = std::min(b.lo[a], t[3 * k + a]);
          b.hi[a] = std::max(b.hi[a], t[3 * k + a]);
        }
      tbox[i] = b;
      for (int a = 0; a < 3; ++a) cen[3 * (size_t)i + a] = 0.5f * (b.lo[a] + b.hi[a]);
      order[i] = i;
    }
    nodes.reserve(2 * (size_t)n + 1);
    nodes.push_back(B2Node());
    struct Item {
      int node;
      uint32_t first, count;
    };
    std::vector<Item> work{{0, 0, n}};
    constexpr int NB = 32;
    while (!work.empty()) {
      Item it = work.back();
      work.pop_back();
      Box box, cbox;
      for (uint32_t k = it.first; k < it.first + it.count; ++k) {
        uint32_t id = order[k];
        box.grow(tbox[id]);
        for (int a = 0; a < 3; ++a) {
          float c = cen[3 * (size_t)id + a];
          cbox.lo[a] = std::min(cbox.lo[a], c);
          cbox.hi[a] = std::max(cbox.hi[a], c);
        }
      }
      nodes[it.node].box = box;
      nodes[it.node].first = it.first;
      nodes[it.node].count = it.count;
      if (it.count <= 1) continue;
      int best_axis = -1, best_bin = -1;
      float best_cost = INFINITY;
      for (int a = 0; a < 3; ++a) {
        float span = cbox.hi[a] - cbox.lo[a];
        if (!(span > 0.f)) continue;
        float kscale = NB / span;
        uint32_t cnt[NB] = {0};
        Box bb[NB];
        for (uint32_t k = it.first; k < it.first + it.count; ++k) {
          uint32_t id = order[k];
          int b = std::min(NB - 1, (int)((cen[3 * (size_t)id + a] - cbox.lo[a]) * kscale));
          cnt[b]++;
          bb[b].grow(tbox[id]);
        }
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
      if (it.count <= 3 && !(best_cost + box.area() < leaf_cost)) continue;
      uint32_t mid;
      if (best_axis < 0) {
        mid = it.first + it.count / 2;
      } else {
        float span = cbox.hi[best_axis] - cbox.lo[best_axis];
        float kscale = NB / span;
        auto pv = std::partition(order.begin() + it.first,
                                 order.begin() + it.first + it.count, [&](uint32_t id) {
                                   int b = std::min(NB - 1, (int)((cen[3 * (size_t)id + best_axis] -
                                                                   cbox.lo[best_axis]) * kscale));
                                   return b <= best_bin;
                                 });
        mid = (uint32_t)(pv - order.begin());
        if (mid == it.first || mid == it.first + it.count) mid = it.first + it.count / 2;
      }
      int l = (int)nodes.size();
      nodes.push_back(B2Node());
      int r = (int)nodes.size();
      nodes.push_back(B2Node());
      nodes[it.node].left = l;
      nodes[it.node].right = r;
      work.push_back({r, mid, it.first + it.count - mid});
      work.push_back({l, it.first, mid - it.first});
    }
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

}  // namespace

void build_cwbvh(const float* tri_xyz, uint32_t n, int width, BvhBuildOut* out) {
  auto t0 = std::chrono::steady_clock::now();
  width = std::max(2, std::min(8, width));
  out->nodes.clear();
  out->tris.clear();
  out->max_depth = 0;
  if (n == 0) {
    out->seconds = 0;
    return;
  }
  Builder b;
  b.build(tri_xyz, n);

  struct Pending {
    int b2;       // BVH2 node collapsed into this wide node
    uint32_t at;  // index in out->nodes
    int depth;
  };
  out->nodes.push_back(CWNode());
  std::deque<Pending> q;
  q.push_back({0, 0, 1});
  while (!q.empty()) {
    Pending pd = q.front();
    q.pop_front();
    out->max_depth = std::max(out->max_depth, pd.depth);
    // gather up to `width` children by opening the largest internal child
    std::vector<int> kids;
    const B2Node& root = b.nodes[pd.b2];
    if (root.leaf()) {
      kids.push_back(pd.b2);  // single-leaf tree
    } else {
      kids.push_back(root.left);
      kids.push_back(root.right);
      while ((int)kids.size() < width) {
        int best = -1;
        float ba = -1.f;
        for (int k = 0; k < (int)kids.size(); ++k) {
          const B2Node& c = b.nodes[kids[k]];
          if (c.leaf()) continue;
          float a = c.box.area();
          if (a > ba) {
            ba = a;
            best = k;
          }
        }
        if (best < 0) break;
        int opened = kids[best];
        kids[best] = b.nodes[opened].left;
        kids.push_back(b.nodes[opened].right);
      }
    }
    Box nb;
    for (int k : kids) nb.grow(b.nodes[k].box);
    // octant slot assignment
    float pc[3], cc[8][3];
    for (int a = 0; a < 3; ++a) pc[a] = 0.5f * (nb.lo[a] + nb.hi[a]);
    int m = (int)kids.size();
    for (int k = 0; k < m; ++k)
      for (int a = 0; a < 3; ++a)
        cc[k][a] = 0.5f * (b.nodes[kids[k]].box.lo[a] + b.nodes[kids[k]].box.hi[a]) - pc[a];
    int slot_of[8];
    bool kid_done[8] = {false}, slot_used[8] = {false};
    for (int round = 0; round < m; ++round) {
      float best = -INFINITY;
      int bk = -1, bs = -1;
      for (int k = 0; k < m; ++k) {
        if (kid_done[k]) continue;
        for (int s = 0; s < 8; ++s) {
          if (slot_used[s]) continue;
          float sc = 0.f;
          for (int a = 0; a < 3; ++a) sc += ((s >> a) & 1) ? -cc[k][a] : cc[k][a];
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
    int slot_kid[8];
    for (int s = 0; s < 8; ++s) slot_kid[s] = -1;
    for (int k = 0; k < m; ++k) slot_kid[slot_of[k]] = kids[k];

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
      const B2Node& c = b.nodes[k];
      quantize_axis(node.px, ex, c.box.lo[0], c.box.hi[0], &node.qlox[s], &node.qhix[s]);
      quantize_axis(node.py, ey, c.box.lo[1], c.box.hi[1], &node.qloy[s], &node.qhiy[s]);
      quantize_axis(node.pz, ez, c.box.lo[2], c.box.hi[2], &node.qloz[s], &node.qhiz[s]);
      // The binary build makes leaves of at most 3 triangles.
      bool as_leaf = c.leaf();
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
      q.push_back({k, at, pd.depth + 1});
    }
    out->nodes[pd.at] = node;
  }
  auto t1 = std::chrono::steady_clock::now();
  out->seconds = std::chrono::duration<double>(t1 - t0).count();
}

}  // namespace shb
