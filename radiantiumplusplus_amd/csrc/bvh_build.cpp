// Host binned-SAH BVH builder: a binary tree built with 32 centroid bins per
// axis, parallel over independent subtrees, then collapsed to 4-wide nodes by
// repeatedly opening the largest-area child.  Node boxes are widened by a few
// ulps so the device slab test never rejects a primitive the exact
// primitive test would accept.
#include "bvh_build.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <thread>

namespace rad {
namespace {

struct Box {
  float lo[3] = {kInf, kInf, kInf};
  float hi[3] = {-kInf, -kInf, -kInf};
  void grow(const Box& b) {
    for (int a = 0; a < 3; a++) {
      lo[a] = std::min(lo[a], b.lo[a]);
      hi[a] = std::max(hi[a], b.hi[a]);
    }
  }
  void grow_pt(const float* p) {
    for (int a = 0; a < 3; a++) {
      lo[a] = std::min(lo[a], p[a]);
      hi[a] = std::max(hi[a], p[a]);
    }
  }
  float area() const {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (dx < 0) return 0.f;
    return 2.f * (dx * dy + dy * dz + dz * dx);
  }
};

struct Node2 {
  Box box;
  int32_t left = -1, right = -1;
  uint32_t start = 0, count = 0;
};

struct Ctx {
  const float* bounds;
  std::vector<float> cent;  // 3 per prim
  std::vector<uint32_t> ref;
  int max_leaf;
};

constexpr int kBins = 32;

// Build a subtree over ref[begin, end) into `out` (root at out[first]).
int BuildRange(Ctx& c, std::vector<Node2>& out, uint32_t begin, uint32_t end, int depth);

Box RangeBox(const Ctx& c, uint32_t begin, uint32_t end, Box* cbox) {
  Box b, cb;
  for (uint32_t i = begin; i < end; i++) {
    uint32_t p = c.ref[i];
    const float* bb = c.bounds + 6 * size_t(p);
    for (int a = 0; a < 3; a++) {
      b.lo[a] = std::min(b.lo[a], bb[a]);
      b.hi[a] = std::max(b.hi[a], bb[3 + a]);
    }
    cb.grow_pt(&c.cent[3 * size_t(p)]);
  }
  if (cbox) *cbox = cb;
  return b;
}

int BuildRange(Ctx& c, std::vector<Node2>& out, uint32_t begin, uint32_t end, int depth) {
  Box cb;
  Node2 node;
  node.box = RangeBox(c, begin, end, &cb);
  uint32_t n = end - begin;
  int idx = (int)out.size();
  out.push_back(node);
  if (n <= 1 || depth >= 60) {
    out[idx].start = begin;
    out[idx].count = n;
    return idx;
  }
  // binned SAH
  float bestCost = kInf;
  int bestAxis = -1, bestBin = -1;
  for (int a = 0; a < 3; a++) {
    float lo = cb.lo[a], hi = cb.hi[a];
    if (!(hi > lo)) continue;
    float scale = kBins / (hi - lo);
    Box bb[kBins];
    uint32_t cnt[kBins] = {0};
    for (uint32_t i = begin; i < end; i++) {
      uint32_t p = c.ref[i];
      int k = std::min(kBins - 1, int((c.cent[3 * size_t(p) + a] - lo) * scale));
      const float* pb = c.bounds + 6 * size_t(p);
      Box t;
      for (int q = 0; q < 3; q++) {
        t.lo[q] = pb[q];
        t.hi[q] = pb[3 + q];
      }
      bb[k].grow(t);
      cnt[k]++;
    }
    float leftA[kBins];
    uint32_t leftN[kBins];
    Box acc;
    uint32_t an = 0;
    for (int k = 0; k < kBins; k++) {
      acc.grow(bb[k]);
      an += cnt[k];
      leftA[k] = acc.area();
      leftN[k] = an;
    }
    acc = Box();
    an = 0;
    for (int k = kBins - 1; k > 0; k--) {
      acc.grow(bb[k]);
      an += cnt[k];
      if (leftN[k - 1] == 0 || an == 0) continue;
      float cost = leftA[k - 1] * leftN[k - 1] + acc.area() * an;
      if (cost < bestCost) {
        bestCost = cost;
        bestAxis = a;
        bestBin = k;
      }
    }
  }
  float parentA = node.box.area();
  float leafCost = float(n);
  float splitCost = 1.0f + (parentA > 0 ? bestCost / parentA : kInf);
  if ((int)n <= c.max_leaf && (bestAxis < 0 || leafCost <= splitCost)) {
    out[idx].start = begin;
    out[idx].count = n;
    return idx;
  }
  uint32_t mid;
  if (bestAxis < 0) {
    mid = begin + n / 2;  // all centroids coincide: split the list
  } else {
    float lo = cb.lo[bestAxis];
    float scale = kBins / (cb.hi[bestAxis] - lo);
    auto it = std::partition(c.ref.begin() + begin, c.ref.begin() + end, [&](uint32_t p) {
      int k = std::min(kBins - 1, int((c.cent[3 * size_t(p) + bestAxis] - lo) * scale));
      return k < bestBin;
    });
    mid = uint32_t(it - c.ref.begin());
    if (mid == begin || mid == end) mid = begin + n / 2;
  }
  // Large independent halves are built in parallel into private vectors.
  if (n > 200000 && depth < 8) {
    std::vector<Node2> lv, rv;
    auto fut = std::async(std::launch::async, [&] { BuildRange(c, lv, begin, mid, depth + 1); });
    BuildRange(c, rv, mid, end, depth + 1);
    fut.get();
    int lbase = (int)out.size();
    for (auto& q : lv) {
      if (q.count == 0) {
        q.left += lbase;
        q.right += lbase;
      }
      out.push_back(q);
    }
    int rbase = (int)out.size();
    for (auto& q : rv) {
      if (q.count == 0) {
        q.left += rbase;
        q.right += rbase;
      }
      out.push_back(q);
    }
    out[idx].left = lbase;
    out[idx].right = rbase;
  } else {
    int l = BuildRange(c, out, begin, mid, depth + 1);
    int r = BuildRange(c, out, mid, end, depth + 1);
    out[idx].left = l;
    out[idx].right = r;
  }
  return idx;
}

float Widen(float v, bool up) {
  // a few ulps outward, so slab tests stay conservative
  float m = std::fabs(v) * 4.0f * 1.1920929e-7f + 1e-30f;
  return up ? v + m : v - m;
}

struct Collapser {
  const std::vector<Node2>& n2;
  std::vector<BvhNode4>& out;
  int Emit(int root) {
    int idx = (int)out.size();
    out.emplace_back();
    int kids[4];
    int nk = 0;
    if (n2[root].count > 0) {
      kids[nk++] = root;  // a leaf root becomes a single-child node
    } else {
      kids[nk++] = n2[root].left;
      kids[nk++] = n2[root].right;
      while (nk < 4) {
        int best = -1;
        float bestA = -1;
        for (int i = 0; i < nk; i++) {
          const Node2& q = n2[kids[i]];
          if (q.count == 0 && q.box.area() > bestA) {
            bestA = q.box.area();
            best = i;
          }
        }
        if (best < 0) break;
        int k = kids[best];
        kids[best] = n2[k].left;
        kids[nk++] = n2[k].right;
      }
    }
    int child[4] = {-1, -1, -1, -1};
    uint32_t count[4] = {0, 0, 0, 0};
    for (int i = 0; i < nk; i++) {
      const Node2& q = n2[kids[i]];
      if (q.count > 0) {
        child[i] = (int)q.start;
        count[i] = q.count;
      } else {
        child[i] = Emit(kids[i]);
      }
    }
    BvhNode4& o = out[idx];
    for (int i = 0; i < 4; i++) {
      if (i < nk) {
        const Box& b = n2[kids[i]].box;
        o.lox[i] = Widen(b.lo[0], false);
        o.loy[i] = Widen(b.lo[1], false);
        o.loz[i] = Widen(b.lo[2], false);
        o.hix[i] = Widen(b.hi[0], true);
        o.hiy[i] = Widen(b.hi[1], true);
        o.hiz[i] = Widen(b.hi[2], true);
      } else {
        o.lox[i] = o.loy[i] = o.loz[i] = kInf;
        o.hix[i] = o.hiy[i] = o.hiz[i] = -kInf;
      }
      o.child[i] = child[i];
      o.count[i] = count[i];
    }
    return idx;
  }
};

}  // namespace

void BuildBvh4Host(const float* bounds, uint32_t count, int max_leaf, BvhBuildResult& out) {
  auto t0 = std::chrono::steady_clock::now();
  Ctx c;
  c.bounds = bounds;
  c.max_leaf = max_leaf > 0 ? std::min(max_leaf, 15) : 4;
  c.cent.resize(3 * size_t(count));
  c.ref.resize(count);
  for (uint32_t i = 0; i < count; i++) {
    const float* b = bounds + 6 * size_t(i);
    for (int a = 0; a < 3; a++) c.cent[3 * size_t(i) + a] = 0.5f * (b[a] + b[3 + a]);
    c.ref[i] = i;
  }
  std::vector<Node2> n2;
  n2.reserve(count ? 2 * size_t(count) / std::max(1, c.max_leaf / 2) : 1);
  out.nodes.clear();
  if (count == 0) {
    BvhNode4 e{};
    for (int i = 0; i < 4; i++) {
      e.lox[i] = e.loy[i] = e.loz[i] = kInf;
      e.hix[i] = e.hiy[i] = e.hiz[i] = -kInf;
      e.child[i] = -1;
      e.count[i] = 0;
    }
    out.nodes.push_back(e);
  } else {
    BuildRange(c, n2, 0, count, 0);
    out.nodes.reserve(n2.size() / 2 + 1);
    Collapser col{n2, out.nodes};
    col.Emit(0);
  }
  out.order = std::move(c.ref);
  out.depth = BvhDepth(out.nodes);
  out.stack_need = BvhStackNeed(out.nodes);
  out.sah_cost = BvhSahCost(out.nodes);
  out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

static float NodeArea(const BvhNode4& n, int i) {
  float dx = n.hix[i] - n.lox[i], dy = n.hiy[i] - n.loy[i], dz = n.hiz[i] - n.loz[i];
  if (dx < 0) return 0;
  return 2.f * (dx * dy + dy * dz + dz * dx);
}

float BvhSahCost(const std::vector<BvhNode4>& nodes) {
  if (nodes.empty()) return 0;
  float rootA = 0;
  {
    float lo[3] = {kInf, kInf, kInf}, hi[3] = {-kInf, -kInf, -kInf};
    const BvhNode4& r = nodes[0];
    for (int i = 0; i < 4; i++) {
      if (r.hix[i] < r.lox[i]) continue;
      lo[0] = std::min(lo[0], r.lox[i]);
      lo[1] = std::min(lo[1], r.loy[i]);
      lo[2] = std::min(lo[2], r.loz[i]);
      hi[0] = std::max(hi[0], r.hix[i]);
      hi[1] = std::max(hi[1], r.hiy[i]);
      hi[2] = std::max(hi[2], r.hiz[i]);
    }
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    rootA = 2.f * (dx * dy + dy * dz + dz * dx);
  }
  if (!(rootA > 0)) return 0;
  double cost = 1.0;  // root traversal
  for (const auto& n : nodes) {
    for (int i = 0; i < 4; i++) {
      if (n.child[i] < 0) continue;
      double a = NodeArea(n, i) / rootA;
      cost += n.count[i] > 0 ? a * n.count[i] : a * 1.0;
    }
  }
  return float(cost);
}

uint32_t BvhDepth(const std::vector<BvhNode4>& nodes) {
  if (nodes.empty()) return 0;
  std::vector<std::pair<int, uint32_t>> st{{0, 1}};
  uint32_t d = 0;
  while (!st.empty()) {
    auto [n, k] = st.back();
    st.pop_back();
    d = std::max(d, k);
    for (int i = 0; i < 4; i++)
      if (nodes[n].child[i] >= 0 && nodes[n].count[i] == 0) st.push_back({nodes[n].child[i], k + 1});
  }
  return d;
}

uint32_t BvhStackNeed(const std::vector<BvhNode4>& nodes) {
  if (nodes.empty()) return 0;
  // post-order over the inner nodes with an explicit stack
  std::vector<uint32_t> need(nodes.size(), 0);
  std::vector<std::pair<uint32_t, bool>> st{{0u, false}};
  while (!st.empty()) {
    auto [n, expanded] = st.back();
    st.pop_back();
    const BvhNode4& nd = nodes[n];
    if (!expanded) {
      st.push_back({n, true});
      for (int i = 0; i < 4; i++)
        if (nd.child[i] >= 0 && nd.count[i] == 0) st.push_back({uint32_t(nd.child[i]), false});
      continue;
    }
    uint32_t inner = 0, deepest = 0;
    for (int i = 0; i < 4; i++)
      if (nd.child[i] >= 0 && nd.count[i] == 0) {
        inner++;
        deepest = std::max(deepest, need[nd.child[i]]);
      }
    need[n] = inner ? inner - 1 + deepest : 0;
  }
  return need[0];
}

}  // namespace rad
