// Host BVH builder: binned SAH (16 bins per axis over centroid bounds, all three axes),
// leaves of up to 16 primitives, built as a binary tree and then collapsed into 4-wide nodes (the
// child with the largest surface area is opened first), which the device reads compressed to 64
// bytes (8-bit child bounds against a per-node origin and power-of-two steps, rounded outward).
// The tree topology is this library's own; the reference builds a different tree
// (accelStruct/Bvh.cpp:51-79,611-771). First hits agree up to exact/near ties because
// both return the closest primitive under the same FP32 intersection arithmetic.
#include "rtg_bvh.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace rtg {

namespace {

struct Bounds {
  float mn[3], mx[3];
  void reset() {
    for (int i = 0; i < 3; i++) {
      mn[i] = INFINITY;
      mx[i] = -INFINITY;
    }
  }
  void grow(const float* a, const float* b) {
    for (int i = 0; i < 3; i++) {
      mn[i] = std::min(mn[i], a[i]);
      mx[i] = std::max(mx[i], b[i]);
    }
  }
  void growp(const float* p) { grow(p, p); }
  float area() const {
    float d0 = mx[0] - mn[0], d1 = mx[1] - mn[1], d2 = mx[2] - mn[2];
    if (d0 < 0.f || d1 < 0.f || d2 < 0.f) return 0.f;
    return 2.f * (d0 * d1 + d1 * d2 + d0 * d2);
  }
};

struct Builder {
  const BuildPrim* prims;
  std::vector<uint32_t> idx;
  std::vector<float> nodes;  // 16 floats per interior node
  uint32_t base;             // leaf-order offset of this primitive set
  int maxDepth = 0;

  uint32_t make_leaf(int begin, int end) { return make_leaf_ref(base + (uint32_t)begin, (uint32_t)(end - begin)); }

  void bounds_of(int begin, int end, Bounds& b, Bounds& cb) const {
    b.reset();
    cb.reset();
    for (int i = begin; i < end; i++) {
      const BuildPrim& p = prims[idx[i]];
      b.grow(p.bmin, p.bmax);
      cb.growp(p.c);
    }
  }

  // returns the ref of the subtree over idx[begin, end)
  uint32_t build(int begin, int end, int depth) {
    int n = end - begin;
    Bounds b, cb;
    bounds_of(begin, end, b, cb);
    if (n <= 1) return make_leaf(begin, end);
    const int NB = 16;
    float bestCost = INFINITY;
    int bestAxis = -1, bestBin = -1;
    float parentArea = b.area();
    for (int axis = 0; axis < 3; axis++) {
      float lo = cb.mn[axis], hi = cb.mx[axis];
      if (!(hi > lo)) continue;
      float scale = NB / (hi - lo);
      Bounds bb[NB];
      int cnt[NB];
      for (int k = 0; k < NB; k++) {
        bb[k].reset();
        cnt[k] = 0;
      }
      for (int i = begin; i < end; i++) {
        const BuildPrim& p = prims[idx[i]];
        int k = std::min(NB - 1, std::max(0, (int)((p.c[axis] - lo) * scale)));
        bb[k].grow(p.bmin, p.bmax);
        cnt[k]++;
      }
      float rightArea[NB];
      int rightCnt[NB];
      Bounds acc;
      acc.reset();
      int c = 0;
      for (int k = NB - 1; k > 0; k--) {
        if (cnt[k]) acc.grow(bb[k].mn, bb[k].mx);
        c += cnt[k];
        rightArea[k] = acc.area();
        rightCnt[k] = c;
      }
      acc.reset();
      c = 0;
      for (int k = 0; k < NB - 1; k++) {
        if (cnt[k]) acc.grow(bb[k].mn, bb[k].mx);
        c += cnt[k];
        if (c == 0 || rightCnt[k + 1] == 0) continue;
        float cost = 1.0f + (c * acc.area() + rightCnt[k + 1] * rightArea[k + 1]) / std::max(parentArea, 1e-30f);
        if (cost < bestCost) {
          bestCost = cost;
          bestAxis = axis;
          bestBin = k;
        }
      }
    }
    float leafCost = (float)n;
    if (n <= kMaxLeafPrims && (bestAxis < 0 || leafCost <= bestCost)) return make_leaf(begin, end);
    int mid;
    if (bestAxis < 0) {
      // all centroids coincide: split by index
      mid = begin + n / 2;
    } else {
      float lo = cb.mn[bestAxis], hi = cb.mx[bestAxis];
      float scale = NB / (hi - lo);
      auto it = std::partition(idx.begin() + begin, idx.begin() + end, [&](uint32_t id) {
        const BuildPrim& p = prims[id];
        int k = std::min(NB - 1, std::max(0, (int)((p.c[bestAxis] - lo) * scale)));
        return k <= bestBin;
      });
      mid = (int)(it - idx.begin());
      if (mid == begin || mid == end) mid = begin + n / 2;
    }
    uint32_t me = (uint32_t)(nodes.size() / 16);
    nodes.resize(nodes.size() + 16, 0.f);
    maxDepth = std::max(maxDepth, depth + 1);
    uint32_t l = build(begin, mid, depth + 1);
    uint32_t r = build(mid, end, depth + 1);
    Bounds lb, rb, tmp;
    bounds_of(begin, mid, lb, tmp);
    bounds_of(mid, end, rb, tmp);
    float* nd = &nodes[(size_t)me * 16];
    write_slot(nd, lb, l);
    write_slot(nd + 8, rb, r);
    return me;
  }
  static void write_slot(float* s, const Bounds& b, uint32_t ref) {
    s[0] = b.mn[0]; s[1] = b.mn[1]; s[2] = b.mn[2];
    s[3] = b.mx[0]; s[4] = b.mx[1]; s[5] = b.mx[2];
    std::memcpy(&s[6], &ref, 4);
    s[7] = 0.f;
  }
};

// ---------------------------------------------------------------- 4-wide collapse
struct Child {
  float mn[3], mx[3];
  uint32_t ref;
  float area() const {
    float d0 = mx[0] - mn[0], d1 = mx[1] - mn[1], d2 = mx[2] - mn[2];
    return 2.f * (d0 * d1 + d1 * d2 + d0 * d2);
  }
};

struct Collapser {
  const std::vector<float>& bin;  // 16 floats per binary node
  std::vector<float> out;         // 32 floats per 4-wide node
  int max_depth = 0;              // 4-wide levels
  int max_stack = 0;              // most traversal stack entries any path can hold

  explicit Collapser(const std::vector<float>& b) : bin(b) {}

  void children(uint32_t ref, Child c[2]) const {
    const float* nd = &bin[(size_t)ref * 16];
    for (int k = 0; k < 2; k++) {
      const float* sl = nd + 8 * k;
      for (int a = 0; a < 3; a++) {
        c[k].mn[a] = sl[a];
        c[k].mx[a] = sl[3 + a];
      }
      std::memcpy(&c[k].ref, &sl[6], 4);
    }
  }

  // 4-wide node over the binary interior node `ref`; `stk` = stack entries its ancestors can leave
  uint32_t collapse(uint32_t ref, int depth, int stk) {
    Child ch[4];
    children(ref, ch);
    int n = 2;
    while (n < 4) {
      int best = -1;
      float ba = -1.f;
      for (int i = 0; i < n; i++)
        if (!(ch[i].ref & kLeafBit) && ch[i].area() > ba) {
          ba = ch[i].area();
          best = i;
        }
      if (best < 0) break;
      Child sub[2];
      children(ch[best].ref, sub);
      ch[best] = sub[0];
      ch[n++] = sub[1];
    }
    const uint32_t me = (uint32_t)(out.size() / 32);
    out.resize(out.size() + 32, 0.f);
    max_depth = std::max(max_depth, depth + 1);
    // a visit pushes every hit child but the nearest
    const int here = stk + (n - 1);
    max_stack = std::max(max_stack, here);
    uint32_t refs[4] = {kEmptyChild, kEmptyChild, kEmptyChild, kEmptyChild};
    for (int i = 0; i < n; i++) refs[i] = (ch[i].ref & kLeafBit) ? ch[i].ref : collapse(ch[i].ref, depth + 1, here);
    float* nd = &out[(size_t)me * 32];  // after the recursion: `out` may have grown
    for (int j = 0; j < 4; j++) {
      const bool v = j < n;
      for (int a = 0; a < 3; a++) {
        nd[4 * a + j] = v ? ch[j].mn[a] : 0.f;
        nd[12 + 4 * a + j] = v ? ch[j].mx[a] : 0.f;
      }
      std::memcpy(&nd[24 + j], &refs[j], 4);
    }
    return me;
  }
};

}  // namespace

BvhResult build_bvh(const std::vector<BuildPrim>& tris, const std::vector<BuildPrim>& spheres) {
  BvhResult res;
  size_t total = tris.size() + spheres.size();
  if (total >= (1u << 27)) throw std::runtime_error("too many primitives for the BVH reference encoding");
  res.empty = total == 0;
  Bounds all;
  all.reset();
  for (auto& p : tris) all.grow(p.bmin, p.bmax);
  for (auto& p : spheres) all.grow(p.bmin, p.bmax);
  for (int i = 0; i < 3; i++) {
    res.root_min[i] = all.mn[i];
    res.root_max[i] = all.mx[i];
  }
  if (res.empty) return res;
  // one builder over a shared node array: triangles, then spheres
  Builder bt;
  bt.prims = tris.data();
  bt.idx.resize(tris.size());
  for (size_t i = 0; i < tris.size(); i++) bt.idx[i] = (uint32_t)i;
  bt.base = 0;
  bool both = !tris.empty() && !spheres.empty();
  if (both) bt.nodes.resize(16, 0.f);  // node 0 joins the two subtrees
  uint32_t triRef = 0, sphRef = 0;
  if (!tris.empty()) triRef = bt.build(0, (int)tris.size(), both ? 1 : 0);
  int depth = bt.maxDepth;
  res.order_tri = bt.idx;
  if (!spheres.empty()) {
    Builder bs;
    bs.prims = spheres.data();
    bs.idx.resize(spheres.size());
    for (size_t i = 0; i < spheres.size(); i++) bs.idx[i] = (uint32_t)i;
    bs.base = (uint32_t)tris.size();
    bs.nodes.swap(bt.nodes);
    sphRef = bs.build(0, (int)spheres.size(), both ? 1 : 0);
    bt.nodes.swap(bs.nodes);
    depth = std::max(depth, bs.maxDepth);
    res.order_sph = bs.idx;
  }
  if (both) {
    Bounds tb, sb;
    tb.reset();
    sb.reset();
    for (auto& p : tris) tb.grow(p.bmin, p.bmax);
    for (auto& p : spheres) sb.grow(p.bmin, p.bmax);
    Builder::write_slot(&bt.nodes[0], tb, triRef);
    Builder::write_slot(&bt.nodes[8], sb, sphRef);
    res.root_ref = 0;
    depth = std::max(depth, 1);
  } else {
    res.root_ref = tris.empty() ? sphRef : triRef;
  }
  (void)depth;
  if (res.root_ref & kLeafBit) {
    res.max_depth = 0;
    res.max_stack = 0;
  } else {
    Collapser cl(bt.nodes);
    res.root_ref = cl.collapse(res.root_ref, 0, 0);
    res.nodes.swap(cl.out);
    res.max_depth = cl.max_depth;
    res.max_stack = cl.max_stack;
  }
  quantize_nodes(res.nodes, res.qnodes);
  return res;
}

// ---------------------------------------------------------------- node compression
namespace {

inline float pow2f(int e) { return std::ldexp(1.f, e); }

// One axis of a node: origin = the smallest child lower bound; the step 2^e is the smallest power of
// two for which every child's bounds round outward into 8 bits (starting from 2^ceil(log2(extent /
// 255)), never below the float spacing at the node's coordinates, so that a step of q changes the
// decoded value). qlo/qhi: the largest / smallest codes whose decoded bound is at or below / above
// the exact bound.
void quantize_axis(const float* lo, const float* hi, const bool* used, float& origin, int& e, uint8_t* qlo,
                   uint8_t* qhi) {
  origin = INFINITY;
  float top = -INFINITY;
  for (int j = 0; j < 4; j++)
    if (used[j]) {
      origin = std::min(origin, lo[j]);
      top = std::max(top, hi[j]);
    }
  if (!(origin <= top)) {  // no used child (not reached for interior nodes)
    origin = 0.f;
    e = 0;
    for (int j = 0; j < 4; j++) qlo[j] = qhi[j] = 0;
    return;
  }
  const float mag = std::max(std::fabs(origin), std::fabs(top));
  int e_ulp = -126;
  if (mag > 0.f) {
    int ex;
    std::frexp(mag, &ex);  // mag = m 2^ex, m in [0.5, 1): float spacing there 2^(ex - 24)
    e_ulp = std::max(-126, ex - 24);
  }
  const double ext = (double)top - (double)origin;
  e = e_ulp;
  if (ext > 0.0) e = std::max(e_ulp, (int)std::ceil(std::log2(ext / 255.0)));
  for (;; e++) {
    if (e > 127) throw std::runtime_error("BVH node too large to compress");
    const float step = pow2f(e);
    bool ok = true;
    for (int j = 0; j < 4 && ok; j++) {
      if (!used[j]) {
        qlo[j] = qhi[j] = 0;
        continue;
      }
      double a = std::floor(((double)lo[j] - origin) / step);
      long ql = (long)std::max(0.0, std::min(255.0, a));
      while (ql > 0 && dequantize_bound(origin, (uint32_t)ql, step) > lo[j]) ql--;
      double b = std::ceil(((double)hi[j] - origin) / step);
      long qh = (long)std::max(0.0, std::min(256.0, b));
      while (qh <= 255 && dequantize_bound(origin, (uint32_t)qh, step) < hi[j]) qh++;
      if (qh > 255 || dequantize_bound(origin, (uint32_t)ql, step) > lo[j]) {
        ok = false;
        break;
      }
      qlo[j] = (uint8_t)ql;
      qhi[j] = (uint8_t)qh;
    }
    if (ok) return;
  }
}

}  // namespace

void quantize_nodes(const std::vector<float>& nodes, std::vector<uint32_t>& qnodes) {
  const size_t n = nodes.size() / 32;
  qnodes.assign(n * 16, 0u);
  for (size_t i = 0; i < n; i++) {
    const float* nd = &nodes[i * 32];
    uint32_t* q = &qnodes[i * 16];
    uint32_t refs[4];
    std::memcpy(refs, nd + 24, 16);
    bool used[4];
    for (int j = 0; j < 4; j++) used[j] = refs[j] != kEmptyChild;
    uint32_t ebits = 0;
    for (int a = 0; a < 3; a++) {
      float origin;
      int e;
      uint8_t ql[4], qh[4];
      quantize_axis(nd + 4 * a, nd + 12 + 4 * a, used, origin, e, ql, qh);
      std::memcpy(&q[a], &origin, 4);
      ebits |= (uint32_t)(e + 127) << (8 * a);
      q[4 + a] = (uint32_t)ql[0] | ((uint32_t)ql[1] << 8) | ((uint32_t)ql[2] << 16) | ((uint32_t)ql[3] << 24);
      q[7 + a] = (uint32_t)qh[0] | ((uint32_t)qh[1] << 8) | ((uint32_t)qh[2] << 16) | ((uint32_t)qh[3] << 24);
    }
    q[3] = ebits;
    std::memcpy(&q[12], refs, 16);
  }
}

}  // namespace rtg
