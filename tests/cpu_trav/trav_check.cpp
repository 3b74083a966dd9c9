// CPU check of the BVH builder (4-wide collapse) and of the device traversal code itself
// (trav_begin / trav_step / stack helpers from rtg_device.h, compiled for the host through
// device_cpu.h): random scenes and rays, closest and any hit against brute force over all
// primitives with the same intersection arithmetic, stack depth against the builder's bound, and
// every primitive reachable exactly once. Meant to run under AddressSanitizer.
// Usage: trav_check <seed> <ntri> <nsph> <nrays>; prints a summary line, exit 1 on failure.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "rtg_bvh.h"
#include "rtg_device.h"

using namespace rtg;
using namespace rtg::dev;

static int g_max_sp = 0;

int main(int argc, char** argv) {
  const uint32_t seed = argc > 1 ? (uint32_t)atoi(argv[1]) : 1;
  const int ntri = argc > 2 ? atoi(argv[2]) : 5000;
  const int nsph = argc > 3 ? atoi(argv[3]) : 0;
  const int nrays = argc > 4 ? atoi(argv[4]) : 20000;
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> U(0.f, 1.f);
  // triangles: mostly small ones in the unit cube, some large, some axis-aligned quads
  std::vector<float> tv((size_t)ntri * 9);
  for (int i = 0; i < ntri; i++) {
    float* v = &tv[(size_t)i * 9];
    if (i % 50 == 0 && i + 1 < ntri) {  // axis-aligned quad split in two (shared edge, coplanar)
      int ax = (int)(U(rng) * 3) % 3;
      float c[3] = {U(rng), U(rng), U(rng)}, s = 0.05f + 0.3f * U(rng);
      int a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
      float q[4][3];
      for (int k = 0; k < 4; k++) {
        q[k][ax] = c[ax];
        q[k][a1] = c[a1] + ((k == 1 || k == 2) ? s : 0.f);
        q[k][a2] = c[a2] + ((k >= 2) ? s : 0.f);
      }
      const int t0[3] = {0, 1, 2}, t1[3] = {0, 2, 3};
      for (int k = 0; k < 3; k++)
        for (int a = 0; a < 3; a++) {
          v[3 * k + a] = q[t0[k]][a];
          v[9 + 3 * k + a] = q[t1[k]][a];
        }
      i++;
      continue;
    }
    float c[3] = {U(rng), U(rng), U(rng)};
    float s = (i % 97 == 0) ? 0.5f : 0.02f + 0.05f * U(rng);
    for (int k = 0; k < 3; k++)
      for (int a = 0; a < 3; a++) v[3 * k + a] = c[a] + s * (U(rng) - 0.5f);
  }
  std::vector<float> sc((size_t)nsph * 4);  // center, radius
  for (int i = 0; i < nsph; i++) {
    for (int a = 0; a < 3; a++) sc[4 * i + a] = U(rng);
    sc[4 * i + 3] = 0.01f + 0.08f * U(rng);
  }
  std::vector<BuildPrim> tp(ntri), sp(nsph);
  for (int i = 0; i < ntri; i++) {
    const float* v = &tv[(size_t)i * 9];
    for (int a = 0; a < 3; a++) {
      tp[i].bmin[a] = std::min(v[a], std::min(v[3 + a], v[6 + a]));
      tp[i].bmax[a] = std::max(v[a], std::max(v[3 + a], v[6 + a]));
      tp[i].c[a] = 0.5f * (tp[i].bmin[a] + tp[i].bmax[a]);
    }
  }
  for (int i = 0; i < nsph; i++)
    for (int a = 0; a < 3; a++) {
      sp[i].bmin[a] = sc[4 * i + a] - sc[4 * i + 3];
      sp[i].bmax[a] = sc[4 * i + a] + sc[4 * i + 3];
      sp[i].c[a] = sc[4 * i + a];
    }
  BvhResult bvh = build_bvh(tp, sp);
  // leaf-order primitive data, as the upload lays it out
  std::vector<float> ltri((size_t)ntri * 9);
  for (int i = 0; i < ntri; i++) std::memcpy(&ltri[(size_t)i * 9], &tv[(size_t)bvh.order_tri[i] * 9], 36);
  std::vector<float4> lsph((size_t)nsph * 5);
  for (int i = 0; i < nsph; i++) {
    const float* c = &sc[4 * (size_t)bvh.order_sph[i]];
    lsph[5 * i + 0] = make_float4(1, 0, 0, -c[0]);  // world -> object: translation by -center
    lsph[5 * i + 1] = make_float4(0, 1, 0, -c[1]);
    lsph[5 * i + 2] = make_float4(0, 0, 1, -c[2]);
    lsph[5 * i + 3] = make_float4(0, 0, 0, 1);
    lsph[5 * i + 4] = make_float4(c[3], 0, 0, 0);
  }
  DevScene S;
  std::memset(&S, 0, sizeof(S));
  S.nodes = (const float4*)bvh.qnodes.data();  // the compressed nodes the device traverses
  S.root_ref = bvh.root_ref;
  S.empty = bvh.empty ? 1 : 0;
  for (int a = 0; a < 3; a++) {
    S.root_min[a] = bvh.root_min[a];
    S.root_max[a] = bvh.root_max[a];
  }
  S.tri_v = ltri.data();
  S.spheres = lsph.data();
  S.ntri = ntri;
  S.nsph = nsph;
  const size_t nnodes = bvh.nodes.size() / 32;
  // structure: every primitive in exactly one reachable leaf; refs in range; boxes contain leaves
  std::vector<int> seen(ntri + nsph, 0);
  std::vector<uint32_t> work = {bvh.root_ref};
  bool bad = false;
  while (!work.empty()) {
    uint32_t r = work.back();
    work.pop_back();
    if (r & kLeafBit) {
      for (uint32_t k = 0; k < leaf_count(r); k++) {
        uint32_t p = leaf_first(r) + k;
        if (p >= (uint32_t)(ntri + nsph)) { bad = true; continue; }
        seen[p]++;
      }
      continue;
    }
    if (r >= nnodes) { bad = true; continue; }
    const uint32_t* rf = (const uint32_t*)&bvh.nodes[(size_t)r * 32 + 24];
    for (int j = 0; j < 4; j++)
      if (rf[j] != kEmptyChild) work.push_back(rf[j]);
  }
  for (int i = 0; i < ntri + nsph; i++) bad = bad || seen[i] != 1;
  if (bad) { printf("FAIL structure\n"); return 1; }
  const int cap = std::max(16, bvh.max_stack + 1);
  std::vector<uint2> ls(kLdsStack);
  std::vector<uint2> gs(cap > kLdsStack ? cap - kLdsStack : 1);
  TravStack stk;
  stk.l = ls.data();
  stk.lstride = 1;
  stk.g = gs.data();
  stk.gstride = 1;
  stk.cap = argc > 5 ? atoi(argv[5]) : cap;  // a sixth argument forces a smaller stack (overflow check)
  uint32_t spills = 0, ovf = 0;
  stk.spill = &spills;
  stk.ovf = &ovf;
  long mism = 0, ties = 0, anymism = 0, hits = 0;
  uint64_t nodes = 0, prims = 0;
  for (int q = 0; q < nrays; q++) {
    Ray r;
    float o[3], d[3];
    for (int a = 0; a < 3; a++) o[a] = -0.2f + 1.4f * U(rng);
    for (int a = 0; a < 3; a++) d[a] = 2.f * U(rng) - 1.f;
    if (q % 7 == 0) d[q % 3] = 0.f;           // zero direction components
    if (q % 11 == 0) { d[0] = 0.f; d[1] = 0.f; d[2] = (q & 1) ? 1.f : -1.f; }
    float l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (!(l > 0.f)) { d[0] = 1.f; l = 1.f; }
    r.o = mk(o[0], o[1], o[2]);
    r.d = mk(d[0] / l, d[1] / l, d[2] / l);
    r.inv = mk(1.f / r.d.x, 1.f / r.d.y, 1.f / r.d.z);
    r.tmin = (q % 3 == 0) ? EPS : 1e-3f * U(rng);
    r.tmax = (q % 5 == 0) ? 0.3f * U(rng) : 1e8f;
    for (int any = 0; any < 2; any++) {
      TravState s;
      s.r = r;
      uint32_t nv = 0, np = 0;
      if (trav_begin(S, s)) {
        const SceneGeo G = scene_geo(S);
        while (true) {
          bool done = any ? trav_step<true>(G, s, stk, nv, np) : trav_step<false>(G, s, stk, nv, np);
          g_max_sp = std::max(g_max_sp, s.sp);
          if (done) break;
        }
      }
      nodes += nv;
      prims += np;
      // brute force, leaf order, same arithmetic and acceptance rules
      Ray b = r;
      adapt_eps(b);
      int bp = -1;
      float bt = 0.f;
      bool banyhit = false;
      if (!(b.tmax < b.tmin)) {
        for (int i = 0; i < ntri; i++) {
          float t, b1, b2;
          if (tri_hit(b, &ltri[(size_t)i * 9], t, b1, b2)) { bp = i; bt = t; banyhit = true; b.tmax = any ? b.tmax : t; }
        }
        for (int i = 0; i < nsph; i++) {
          float t;
          if (sphere_hit(b, &lsph[5 * (size_t)i], t)) { bp = ntri + i; bt = t; banyhit = true; b.tmax = any ? b.tmax : t; }
        }
      }
      if (any) {
        if (s.hit != banyhit) anymism++;
      } else {
        if (s.hit) hits++;
        if (s.hit != (bp >= 0)) {
          mism++;
        } else if (s.hit && s.h.prim != bp) {
          if (std::fabs(s.h.t - bt) <= 4e-7f * std::max(1.f, std::fabs(bt))) ties++;
          else mism++;
        }
      }
    }
  }
  // The fused child test on the compressed nodes accepts every child box that the reference's slab
  // test accepts on the builder's exact float box (it may accept a little more), and its entry
  // distance is never larger: checked for every node and child against a subset of the rays.
  long fused_viol = 0, fused_extra = 0, fused_tests = 0;
  {
    std::mt19937 rng2(seed + 77);
    const int nprop = std::max(50, (int)std::min<size_t>((size_t)nrays, 3000000 / std::max<size_t>(1, nnodes)));
    for (int q = 0; q < nprop; q++) {
      Ray r;
      float o[3], d[3];
      for (int a = 0; a < 3; a++) o[a] = -0.2f + 1.4f * U(rng2);
      for (int a = 0; a < 3; a++) d[a] = 2.f * U(rng2) - 1.f;
      if (q % 5 == 0) d[q % 3] *= 1e-6f;  // nearly axis-parallel directions: large reciprocals
      float l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      r.o = mk(o[0], o[1], o[2]);
      r.d = mk(d[0] / l, d[1] / l, d[2] / l);
      r.inv = mk(1.f / r.d.x, 1.f / r.d.y, 1.f / r.d.z);
      r.tmin = (q % 3 == 0) ? EPS : 1e-3f * U(rng2);
      r.tmax = (q % 4 == 0) ? 0.5f * U(rng2) : 1e8f;
      adapt_eps(r);
      if (has_zero_dir(r)) continue;
      for (size_t ni = 0; ni < nnodes; ni++) {
        const float* nd = &bvh.nodes[ni * 32];
        const uint4* qn = reinterpret_cast<const uint4*>(&bvh.qnodes[ni * 16]);
        bool h[4];
        float t[4];
        node_children(qn[0], qn[1], qn[2], r, false, h[0], h[1], h[2], h[3], t[0], t[1], t[2], t[3]);
        const uint32_t* rf = (const uint32_t*)(nd + 24);
        for (int j = 0; j < 4; j++) {
          if (rf[j] == kEmptyChild) continue;
          float tr;
          const bool hr = slab(r, nd[j], nd[4 + j], nd[8 + j], nd[12 + j], nd[16 + j], nd[20 + j], tr);
          fused_tests++;
          if (hr && (!h[j] || t[j] > tr)) fused_viol++;
          if (!hr && h[j]) fused_extra++;
        }
      }
    }
  }
  printf("fused child tests %ld: violations %ld, extra acceptances %ld\n", fused_tests, fused_viol, fused_extra);
  printf("seed %u tris %d spheres %d rays %d: nodes4 %zu depth4 %d stack bound %d max sp %d hits %ld "
         "mismatches %ld ties %ld any-hit mismatches %ld nodes/ray %.2f prims/ray %.2f\n",
         seed, ntri, nsph, nrays, nnodes, bvh.max_depth, bvh.max_stack, g_max_sp, hits, mism, ties, anymism,
         (double)nodes / (2.0 * nrays), (double)prims / (2.0 * nrays));
  printf("spilled pushes %u overflow %u\n", spills, ovf);
  if (argc > 5) {  // forced small stack: the overflow must be flagged
    printf(ovf ? "OVERFLOW FLAGGED\n" : "FAIL overflow not flagged\n");
    return ovf ? 0 : 1;
  }
  const bool ok = fused_viol == 0 && g_max_sp <= bvh.max_stack && mism <= nrays / 10000 && anymism <= nrays / 10000 &&
                  ties <= nrays / 1000 && hits > 0 && ovf == 0 && (g_max_sp <= kLdsStack || spills > 0);
  printf(ok ? "OK\n" : "FAIL\n");
  return ok ? 0 : 1;
}
