// Host side of librtg: device context, scene upload (BVH, primitive tables, lights),
// the wavefront render loop and the C ABI declared in include/rtg.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rtg.h"
#include "rtg_bvh.h"
#include "rtg_host.h"
#include "rtg_internal.h"

using namespace rtg;

namespace rtg {

[[noreturn]] void fail(int code, const std::string& msg) { throw RtgError(code, msg); }

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    int code = (e == hipErrorOutOfMemory) ? RTG_ERR_OOM : RTG_ERR_HIP;
    fail(code, std::string(what) + ": " + hipGetErrorString(e));
  }
}

void* dev_alloc(rtg_ctx* ctx, size_t bytes, std::vector<void*>& owned) {
  void* p = nullptr;
  if (bytes == 0) bytes = 16;
  hip_check(hipMalloc(&p, bytes), "hipMalloc");
  owned.push_back(p);
  (void)ctx;
  return p;
}

template <class T>
T* upload(rtg_ctx* ctx, const std::vector<T>& v, std::vector<void*>& owned, uint64_t& bytes) {
  size_t n = v.size() * sizeof(T);
  T* p = (T*)dev_alloc(ctx, n, owned);
  if (n) hip_check(hipMemcpyAsync(p, v.data(), n, hipMemcpyHostToDevice, ctx->stream), "upload");
  bytes += n ? n : 16;
  return p;
}

static void cdf_build(const std::vector<float>& w, std::vector<float>& out, float& sum) {
  out.clear();
  out.push_back(0.f);
  sum = 0.f;
  for (float x : w) {
    sum += x;
    out.push_back(sum);
  }
  if (sum != 0.f) {
    float f = 1.f / sum;
    for (size_t i = 1; i < out.size(); i++) out[i] *= f;
    out.back() = 1.f;
  }
}

static inline float lum(const float* c) { return 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2]; }
static inline float avg3(float a) {  // Color(a).average()
  float s = 0.f;
  s += a; s += a; s += a;
  return s / 3.f;
}

static_assert(INTEGRATOR_PATH_MIS == RTG_INTEGRATOR_PATH_MIS && INTEGRATOR_VOLPATH == RTG_INTEGRATOR_VOLPATH,
              "integrator ids");
static_assert(MAT_CONDUCTOR == RTG_MAT_CONDUCTOR && MAT_DIELECTRIC == RTG_MAT_DIELECTRIC &&
                  MAT_ROUGH_DIELECTRIC == RTG_MAT_ROUGH_DIELECTRIC,
              "device material types follow rtg.h");
// rtg_material -> the device record (texture/ConstantTexture.h; bsdf constructors). Smooth
// conductors and dielectrics are shaded by the rough kernels of their kind (shade class 1, 2).
DevMaterial dev_material(const rtg_material& m) {
  if (m.type < RTG_MAT_DIFFUSE || m.type > RTG_MAT_DIELECTRIC) fail(RTG_ERR_INVALID, "unknown material type");
  if ((m.type == RTG_MAT_ROUGH_CONDUCTOR || m.type == RTG_MAT_ROUGH_DIELECTRIC) &&
      m.distribution != RTG_DIST_GGX && m.distribution != RTG_DIST_BECKMANN)
    fail(RTG_ERR_INVALID, "distribution must be ggx or beckmann");
  DevMaterial d;
  std::memset(&d, 0, sizeof(d));
  d.type = m.type;
  d.cls = m.type == RTG_MAT_CONDUCTOR ? 1 : (m.type == RTG_MAT_DIELECTRIC ? 2 : m.type);
  d.dist = m.distribution;
  d.alpha = avg3(m.alpha);
  d.eta_int = m.eta_int;
  d.eta_ext = m.eta_ext;
  for (int c = 0; c < 3; c++) {
    d.albedo[c] = m.albedo[c];
    d.eta[c] = m.eta[c];
    d.k[c] = m.k[c];
    d.trans[c] = m.transmitted[c];
  }
  // rtg_material's 1-based texture references -> 0-based, -1 for none (range-checked by the scene)
  d.tex[0] = m.tex_albedo - 1;
  d.tex[1] = m.tex_alpha - 1;
  d.tex[2] = m.tex_transmitted - 1;
  d.tex[3] = -1;
  return d;
}

// 4x4 inverse in double, rounded to float (row-major)
static void inverse44(const float* a, float* out) {
  double m[4][8];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      m[i][j] = a[4 * i + j];
      m[i][j + 4] = (i == j) ? 1.0 : 0.0;
    }
  for (int c = 0; c < 4; c++) {
    int p = c;
    for (int r = c + 1; r < 4; r++)
      if (std::fabs(m[r][c]) > std::fabs(m[p][c])) p = r;
    if (m[p][c] == 0.0) fail(RTG_ERR_INVALID, "singular sphere transform");
    if (p != c)
      for (int j = 0; j < 8; j++) std::swap(m[p][j], m[c][j]);
    double inv = 1.0 / m[c][c];
    for (int j = 0; j < 8; j++) m[c][j] *= inv;
    for (int r = 0; r < 4; r++)
      if (r != c) {
        double f = m[r][c];
        for (int j = 0; j < 8; j++) m[r][j] -= f * m[c][j];
      }
  }
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) out[4 * i + j] = (float)m[i][j + 4];
}

static inline float xpt(const float* M, int row, float x, float y, float z) {
  return M[4 * row] * x + M[4 * row + 1] * y + M[4 * row + 2] * z + M[4 * row + 3];
}

// Camera matrices: core/Transform.cpp:199-227 (lookAt) and camera/NoriPerspective.cpp:10-37,
// evaluated in double and rounded to float.
void camera_matrices(const rtg_render_desc& d, float* s2c, float* c2w) {
  auto nrm = [](double v[3]) {
    double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (!(l > 0.0)) fail(RTG_ERR_INVALID, "degenerate camera frame");
    v[0] /= l; v[1] /= l; v[2] /= l;
  };
  auto crs = [](const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  if (d.use_camera_matrix) {
    for (int i = 0; i < 16; i++) c2w[i] = d.camera_to_world[i];
  } else {
    double dir[3] = {(double)d.target[0] - d.origin[0], (double)d.target[1] - d.origin[1],
                     (double)d.target[2] - d.origin[2]};
    nrm(dir);
    double up[3] = {d.up[0], d.up[1], d.up[2]};
    nrm(up);
    double left[3], nu[3];
    crs(up, dir, left);
    nrm(left);
    crs(dir, left, nu);
    nrm(nu);
    std::memset(c2w, 0, 16 * sizeof(float));
    for (int i = 0; i < 3; i++) {
      c2w[4 * i + 0] = (float)left[i];
      c2w[4 * i + 1] = (float)nu[i];
      c2w[4 * i + 2] = (float)dir[i];
      c2w[4 * i + 3] = d.origin[i];
    }
    c2w[15] = 1.f;
  }
  double aspect = d.width / (double)d.height;
  double recip = 1.0 / ((double)d.far_clip - d.near_clip);
  double cot = 1.0 / std::tan((d.fov * 3.14159265358979323846 / 180.0) / 2.0);
  double P[4][4] = {{cot, 0, 0, 0},
                    {0, cot, 0, 0},
                    {0, 0, d.far_clip * recip, -(double)d.near_clip * d.far_clip * recip},
                    {0, 0, 1, 0}};
  double S[4][4] = {{-0.5, 0, 0, 0}, {0, -0.5 * aspect, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
  double T[4][4] = {{1, 0, 0, -1.0}, {0, 1, 0, -1.0 / aspect}, {0, 0, 1, 0}, {0, 0, 0, 1}};
  double ST[4][4] = {}, M[4][4] = {};
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      for (int k = 0; k < 4; k++) ST[i][j] += S[i][k] * T[k][j];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      for (int k = 0; k < 4; k++) M[i][j] += ST[i][k] * P[k][j];
  double m[4][8];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      m[i][j] = M[i][j];
      m[i][j + 4] = (i == j);
    }
  for (int c = 0; c < 4; c++) {
    int p = c;
    for (int r = c + 1; r < 4; r++)
      if (std::fabs(m[r][c]) > std::fabs(m[p][c])) p = r;
    for (int j = 0; j < 8; j++) std::swap(m[p][j], m[c][j]);
    double iv = 1.0 / m[c][c];
    for (int j = 0; j < 8; j++) m[c][j] *= iv;
    for (int r = 0; r < 4; r++)
      if (r != c) {
        double f = m[r][c];
        for (int j = 0; j < 8; j++) m[r][j] -= f * m[c][j];
      }
  }
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) s2c[4 * i + j] = (float)m[i][j + 4];
}

// ---------------------------------------------------------------- scene upload
rtg_scene* build_scene(rtg_ctx* ctx, const rtg_scene_desc* desc) {
  if (!desc) fail(RTG_ERR_INVALID, "null scene description");
  if (desc->num_meshes < 0 || desc->num_spheres < 0 || desc->num_materials <= 0)
    fail(RTG_ERR_INVALID, "scene needs at least one material and non-negative counts");
  if (desc->light_strategy != RTG_LIGHTS_DEFAULT && desc->light_strategy != RTG_LIGHTS_UNIFORM &&
      desc->light_strategy != RTG_LIGHTS_WEIGHTED && desc->light_strategy != RTG_LIGHTS_ALL)
    fail(RTG_ERR_INVALID, "unknown light_strategy");
  // the reference's default strategy (integrator/Integrator.cpp:14) for a zero-initialised field
  const int strategy = desc->light_strategy == RTG_LIGHTS_DEFAULT ? RTG_LIGHTS_UNIFORM : desc->light_strategy;
  int slot = -1;
  for (int i = 0; i < kMaxScenes; i++)
    if (!ctx->scene_slots[i]) {
      slot = i;
      break;
    }
  if (slot < 0) fail(RTG_ERR_OOM, "too many resident scenes");
  const int nmat = desc->num_materials;
  // materials (texture/ConstantTexture.h; bsdf constructors)
  std::vector<DevMaterial> mats(nmat);
  for (int i = 0; i < nmat; i++) {
    mats[i] = dev_material(desc->materials[i]);
    for (int k = 0; k < 3; k++)
      if (mats[i].tex[k] < -1 || mats[i].tex[k] >= desc->num_textures)
        fail(RTG_ERR_INVALID, "material " + std::to_string(i) + ": texture reference out of range");
  }
  // image textures (texture/ImageTexture.cpp:7-64)
  std::vector<DevTexture> texs;
  std::vector<float> texels;
  if (desc->num_textures < 0 || (desc->num_textures > 0 && !desc->textures))
    fail(RTG_ERR_INVALID, "bad texture table");
  for (int i = 0; i < desc->num_textures; i++) {
    const rtg_texture_desc& T = desc->textures[i];
    if (!T.rgb || T.width < 2 || T.height < 2) fail(RTG_ERR_INVALID, "texture " + std::to_string(i) + ": needs at least 2x2 texels");
    if ((size_t)T.width * T.height * 3 > ((size_t)1 << 31) - texels.size())
      fail(RTG_ERR_INVALID, "textures above 2^31 values in all");
    DevTexture d;
    std::memset(&d, 0, sizeof(d));
    d.w = T.width;
    d.h = T.height;
    d.invert_x = T.invert_x ? 1 : 0;
    d.invert_y = T.invert_y ? 1 : 0;
    d.uv_scale = std::fabs(T.uv_scale);
    d.value_scale = T.value_scale;
    d.offset = (uint32_t)texels.size();
    texels.insert(texels.end(), T.rgb, T.rgb + (size_t)T.width * T.height * 3);
    texs.push_back(d);
  }
  // triangles in global id order
  std::vector<BuildPrim> tprims;
  std::vector<float> tv;           // 9 per triangle, global order
  std::vector<uint32_t> tmat, tlight, tmesh;
  std::vector<int32_t> tvidx;      // global vertex ids
  std::vector<uint8_t> thasn;
  std::vector<float> vnormals, vuvs;
  std::vector<uint8_t> thasuv;
  std::vector<DevLight> lights;
  std::vector<float> emit_tri, emit_nrm, emit_cdf;
  int vbase = 0;
  bool anyNormals = false, anyUvs = false;
  for (int mi = 0; mi < desc->num_meshes; mi++)
    if (desc->meshes[mi].uvs && desc->meshes[mi].num_triangles > 0) anyUvs = true;
  for (int mi = 0; mi < desc->num_meshes; mi++) {
    const rtg_mesh_desc& M = desc->meshes[mi];
    if (M.num_triangles < 0 || M.num_vertices < 0 || (M.num_triangles > 0 && (!M.indices || !M.vertices)))
      fail(RTG_ERR_INVALID, "mesh " + std::to_string(mi) + ": missing arrays");
    uint32_t lightId = 0xffffffffu;
    if (M.emission) {
      DevLight L;
      std::memset(&L, 0, sizeof(L));
      L.type = LIGHT_MESH;
      L.emit_offset = (int)(emit_tri.size() / 9);
      L.emit_count = M.num_triangles;
      L.cdf_offset = (int)emit_cdf.size();
      for (int c = 0; c < 3; c++) L.radiance[c] = M.emission[c];
      L.has_normals = M.normals ? 1 : 0;
      std::vector<float> areas(M.num_triangles);
      float area = 0.f;
      for (int t = 0; t < M.num_triangles; t++) {
        const float* p[3];
        for (int k = 0; k < 3; k++) {
          int vi = M.indices[3 * t + k];
          p[k] = M.vertices + 3 * vi;
          for (int c = 0; c < 3; c++) emit_tri.push_back(p[k][c]);
          for (int c = 0; c < 3; c++) emit_nrm.push_back(M.normals ? M.normals[3 * vi + c] : 0.f);
        }
        // Triangle.h:128-135 (Eigen reduction order a0 + (a1 + a2) for the norm)
        float e1[3], e2[3];
        for (int c = 0; c < 3; c++) {
          e1[c] = p[1][c] - p[0][c];
          e2[c] = p[2][c] - p[0][c];
        }
        float cx = e1[1] * e2[2] - e1[2] * e2[1];
        float cy = e1[2] * e2[0] - e1[0] * e2[2];
        float cz = e1[0] * e2[1] - e1[1] * e2[0];
        float s = 0.5f * std::sqrt(cx * cx + (cy * cy + cz * cz));
        areas[t] = s;
        area += s;
      }
      std::vector<float> c;
      float sum;
      cdf_build(areas, c, sum);
      emit_cdf.insert(emit_cdf.end(), c.begin(), c.end());
      L.area = area;
      lights.push_back(L);
      lightId = (uint32_t)(lights.size() - 1);
    }
    if (M.normals) anyNormals = true;
    for (int t = 0; t < M.num_triangles; t++) {
      BuildPrim bp;
      float v[9];
      for (int k = 0; k < 3; k++) {
        int vi = M.indices[3 * t + k];
        if (vi < 0 || vi >= M.num_vertices) fail(RTG_ERR_INVALID, "mesh index out of range");
        for (int c = 0; c < 3; c++) v[3 * k + c] = M.vertices[3 * vi + c];
        tvidx.push_back(vbase + vi);
      }
      for (int c = 0; c < 3; c++) {
        bp.bmin[c] = std::min(std::min(v[c], v[3 + c]), v[6 + c]);
        bp.bmax[c] = std::max(std::max(v[c], v[3 + c]), v[6 + c]);
        bp.c[c] = (v[c] + v[3 + c] + v[6 + c]) * (1.f / 3.f);
      }
      tprims.push_back(bp);
      tv.insert(tv.end(), v, v + 9);
      int mat = M.tri_material ? M.tri_material[t] : M.material;
      if (mat < -1 || mat >= nmat) fail(RTG_ERR_INVALID, "material id out of range");
      tmat.push_back(mat < 0 ? 0xffffffffu : (uint32_t)mat);
      tlight.push_back(lightId);
      tmesh.push_back((uint32_t)mi);
      thasn.push_back(M.normals ? 1 : 0);
      thasuv.push_back(M.uvs ? 1 : 0);
    }
    for (int i = 0; i < M.num_vertices; i++)
      for (int c = 0; c < 3; c++) vnormals.push_back(M.normals ? M.normals[3 * i + c] : 0.f);
    if (anyUvs)
      for (int i = 0; i < M.num_vertices; i++)
        for (int c = 0; c < 2; c++) vuvs.push_back(M.uvs ? M.uvs[2 * i + c] : 0.f);
    vbase += M.num_vertices;
  }
  // spheres
  std::vector<BuildPrim> sprims;
  std::vector<float> sdata;  // 20 floats per sphere, global order
  for (int si = 0; si < desc->num_spheres; si++) {
    const rtg_sphere_desc& s = desc->spheres[si];
    if (s.material < -1 || s.material >= nmat) fail(RTG_ERR_INVALID, "sphere material out of range");
    float inv[16];
    inverse44(s.to_world, inv);
    float r = s.radius;
    // GeometricShape.h:87 + Transform.h:141-153: 8 transformed corners of the local box
    float corners[8][3] = {{-r, -r, -r}, {-r, -r, r}, {-r, r, r}, {r, r, r},
                           {r, -r, -r}, {r, r, -r}, {r, -r, r}, {-r, r, -r}};
    BuildPrim bp;
    for (int c = 0; c < 3; c++) {
      bp.bmin[c] = 1e8f;
      bp.bmax[c] = -1e8f;
    }
    const float* W = s.to_world;
    for (auto& cn : corners) {
      float p[3];
      float w = W[12] * cn[0] + W[13] * cn[1] + W[14] * cn[2] + W[15];
      for (int c = 0; c < 3; c++) p[c] = xpt(W, c, cn[0], cn[1], cn[2]);
      if (w != 1.f && w != 0.f)
        for (int c = 0; c < 3; c++) p[c] = p[c] / w;
      for (int c = 0; c < 3; c++) {
        bp.bmin[c] = std::min(bp.bmin[c], p[c]);
        bp.bmax[c] = std::max(bp.bmax[c], p[c]);
      }
    }
    for (int c = 0; c < 3; c++) bp.c[c] = xpt(W, c, 0.f, 0.f, 0.f);
    sprims.push_back(bp);
    for (int k = 0; k < 16; k++) sdata.push_back(inv[k]);
    sdata.push_back(r);
    sdata.push_back(0.f);
    sdata.push_back(0.f);
    sdata.push_back(0.f);
  }
  const int ntri = (int)tprims.size(), nsph = (int)sprims.size();
  BvhResult bvh = build_bvh(tprims, sprims);
  if (bvh.max_stack + 1 > kMaxStack) fail(RTG_ERR_INVALID, "BVH deeper than the traversal stack");
  // leaf-order tables
  std::vector<float> lv((size_t)ntri * 9);
  std::vector<uint4> info((size_t)ntri + nsph);
  std::vector<int32_t> lvidx;
  std::vector<uint8_t> lhasn, lhasuv;
  if (anyNormals || anyUvs) lvidx.resize((size_t)ntri * 3);
  if (anyNormals) lhasn.resize(ntri);
  if (anyUvs) lhasuv.resize(ntri);
  uint32_t sphBase = (uint32_t)ntri;
  for (int i = 0; i < ntri; i++) {
    uint32_t g = bvh.order_tri[i];
    std::memcpy(&lv[(size_t)i * 9], &tv[(size_t)g * 9], 9 * sizeof(float));
    info[i] = make_uint4(g, tmat[g], tlight[g], tmesh[g]);
    if (anyNormals || anyUvs)
      for (int k = 0; k < 3; k++) lvidx[(size_t)i * 3 + k] = tvidx[(size_t)g * 3 + k];
    if (anyNormals) lhasn[i] = thasn[g];
    if (anyUvs) lhasuv[i] = thasuv[g];
  }
  std::vector<float4> lsph((size_t)nsph * 5);
  for (int i = 0; i < nsph; i++) {
    uint32_t g = bvh.order_sph[i];
    const float* sd = &sdata[(size_t)g * 20];
    for (int k = 0; k < 5; k++) lsph[(size_t)i * 5 + k] = make_float4(sd[4 * k], sd[4 * k + 1], sd[4 * k + 2], sd[4 * k + 3]);
    const rtg_sphere_desc& s = desc->spheres[g];
    info[sphBase + i] = make_uint4(sphBase + g, s.material < 0 ? 0xffffffffu : (uint32_t)s.material, 0xffffffffu,
                                   0xffffffffu);
  }
  // environment light (light/EnvironmentLight.cpp:8-44,169-174)
  std::unique_ptr<rtg_scene> sc(new rtg_scene());
  sc->ctx = ctx;
  sc->slot = slot;
  DevScene& D = sc->dev;
  std::memset(&D, 0, sizeof(D));
  D.env_light = -1;
  std::vector<float> env_map, row_cdf, marg_cdf, env_w;
  std::vector<int32_t> marg_guide, row_guide;
  if (desc->env_rgb) {
    int W = desc->env_width, H = desc->env_height;
    if (W < 2 || H < 2) fail(RTG_ERR_INVALID, "environment map must be at least 2x2");
    env_map.assign(desc->env_rgb, desc->env_rgb + (size_t)W * H * 3);
    std::vector<float> margW;
    float scale = desc->env_scale;
    const float PI = 3.14159265358979323846f;
    for (int i = 0; i < H; i++) {
      float weight = std::sin(PI * (i + 0.5f) / H);
      std::vector<float> rw(W);
      for (int j = 0; j < W; j++) {
        const float* px = &env_map[((size_t)i * W + j) * 3];
        float c[3] = {px[0] * scale, px[1] * scale, px[2] * scale};
        rw[j] = lum(c);
      }
      std::vector<float> rc;
      float rsum;
      cdf_build(rw, rc, rsum);
      row_cdf.insert(row_cdf.end(), rc.begin(), rc.end());
      env_w.push_back(weight);
      margW.push_back(rsum * weight);
    }
    float msum;
    cdf_build(margW, marg_cdf, msum);
    cdf_guide(marg_cdf.data(), H + 1, marg_guide);
    for (int i = 0; i < H; i++) cdf_guide(row_cdf.data() + (size_t)i * (W + 1), W + 1, row_guide);
    D.env_w = W;
    D.env_h = H;
    D.env_factor = scale;
    D.env_marg_sum = msum;
    // bounding sphere of the scene box (EnvironmentLight.cpp:169-174)
    float c[3], r1 = 0.f, r2 = 0.f;
    for (int k = 0; k < 3; k++) c[k] = 0.5f * (bvh.root_max[k] + bvh.root_min[k]);
    {
      float a0 = bvh.root_min[0] - c[0], a1 = bvh.root_min[1] - c[1], a2 = bvh.root_min[2] - c[2];
      float b0 = bvh.root_max[0] - c[0], b1 = bvh.root_max[1] - c[1], b2 = bvh.root_max[2] - c[2];
      r1 = std::sqrt(a0 * a0 + (a1 * a1 + a2 * a2));
      r2 = std::sqrt(b0 * b0 + (b1 * b1 + b2 * b2));
    }
    D.env_radius = 2 * std::max(r1, r2);
    DevLight L;
    std::memset(&L, 0, sizeof(L));
    L.type = LIGHT_ENV;
    lights.push_back(L);
    D.env_light = (int)lights.size() - 1;
  }
  // light selection CDF (integrator/Integrator.cpp:24-38)
  std::vector<float> lw, lcdf;
  float lsum;
  for (auto& L : lights) {
    if (strategy == RTG_LIGHTS_WEIGHTED) {
      const float PI = 3.14159265358979323846f;
      float pw[3];
      if (L.type == LIGHT_MESH) {
        for (int c = 0; c < 3; c++) pw[c] = PI * L.radiance[c] * L.area;
      } else {
        for (int c = 0; c < 3; c++) pw[c] = 4 * PI * D.env_radius * D.env_radius * 1.f;
      }
      lw.push_back(lum(pw));
    } else {
      lw.push_back(1.f / lights.size());
    }
  }
  cdf_build(lw, lcdf, lsum);
  if (strategy == RTG_LIGHTS_ALL && (int)lights.size() > kAllLights)
    fail(RTG_ERR_INVALID, "the all_lights strategy supports at most " + std::to_string(kAllLights) + " lights");
  // upload
  uint64_t bytes = 0;
  auto& own = sc->owned;
  D.nodes = (const float4*)upload(ctx, bvh.qnodes, own, bytes);
  D.root_ref = bvh.root_ref;
  D.empty = bvh.empty ? 1 : 0;
  for (int k = 0; k < 3; k++) {
    D.root_min[k] = bvh.root_min[k];
    D.root_max[k] = bvh.root_max[k];
  }
  D.tri_v = upload(ctx, lv, own, bytes);
  D.prim_info = upload(ctx, info, own, bytes);
  if (anyNormals || anyUvs) D.tri_vidx = upload(ctx, lvidx, own, bytes);
  if (anyUvs) {
    D.vuv = upload(ctx, vuvs, own, bytes);
    D.tri_has_uv = upload(ctx, lhasuv, own, bytes);
  }
  if (!texs.empty()) {
    D.textures = upload(ctx, texs, own, bytes);
    D.texels = upload(ctx, texels, own, bytes);
  }
  if (anyNormals) {
    D.tri_has_n = upload(ctx, lhasn, own, bytes);
    D.vnormals = upload(ctx, vnormals, own, bytes);
  }
  D.spheres = upload(ctx, lsph, own, bytes);
  D.ntri = ntri;
  D.nsph = nsph;
  D.mats = upload(ctx, mats, own, bytes);
  D.nmat = nmat;
  D.nlights = (int)lights.size();
  D.lights = upload(ctx, lights, own, bytes);
  D.light_cdf = upload(ctx, lcdf, own, bytes);
  D.strategy = strategy;
  D.emit_tri = upload(ctx, emit_tri, own, bytes);
  D.emit_nrm = upload(ctx, emit_nrm, own, bytes);
  D.emit_cdf = upload(ctx, emit_cdf, own, bytes);
  D.env_map = upload(ctx, env_map, own, bytes);
  D.env_row_cdf = upload(ctx, row_cdf, own, bytes);
  D.env_marg_cdf = upload(ctx, marg_cdf, own, bytes);
  D.env_weight = upload(ctx, env_w, own, bytes);
  D.env_marg_guide = upload(ctx, marg_guide, own, bytes);
  D.env_row_guide = upload(ctx, row_guide, own, bytes);
  // camera medium: GridVolume data as given (GridVolume.cpp:7-64), majorant of
  // Heterogeneous.cpp:40 (GridVolume::getMaxSigmaT x scale)
  if (desc->medium_homogeneous) {
    // medium/Homogeneous.cpp:7-22 (the scale multiplies both coefficients)
    if (desc->medium_density) fail(RTG_ERR_INVALID, "a homogeneous medium takes no density grid");
    const float scale = desc->medium_scale > 0.f ? desc->medium_scale : 1.f;
    for (int k = 0; k < 3; k++) {
      const float a = desc->medium_sigma_a[k], sc = desc->medium_sigma_s[k];
      if (!(std::isfinite(a) && std::isfinite(sc) && a >= 0.f && sc >= 0.f))
        fail(RTG_ERR_INVALID, "medium coefficients must be finite and non-negative");
      D.sig_a[k] = scale != 1.f ? a * scale : a;
      D.sig_s[k] = scale != 1.f ? sc * scale : sc;
      D.sig_t[k] = D.sig_a[k] + D.sig_s[k];
    }
    if (!std::isfinite(desc->medium_g) || desc->medium_g >= 1.f || desc->medium_g <= -1.f)
      fail(RTG_ERR_INVALID, "phase function g must lie in (-1, 1)");
    D.hg_g = desc->medium_g;
    D.gscale = scale;
    D.medium_kind = 2;
  }
  if (desc->medium_density) {
    const int* r = desc->medium_res;
    if (r[0] < 2 || r[1] < 2 || r[2] < 2) fail(RTG_ERR_INVALID, "medium grid must be at least 2x2x2");
    const size_t nv = (size_t)r[0] * r[1] * r[2];
    if (nv > ((size_t)1 << 31)) fail(RTG_ERR_INVALID, "medium grid above 2^31 values");
    float mx = desc->medium_density[0];
    for (size_t i = 0; i < nv; i++) {
      const float v = desc->medium_density[i];
      if (!std::isfinite(v)) fail(RTG_ERR_INVALID, "medium density must be finite");
      if (v > mx) mx = v;
    }
    const float scale = desc->medium_scale > 0.f ? desc->medium_scale : 1.f;
    if (!std::isfinite(desc->medium_g) || desc->medium_g >= 1.f || desc->medium_g <= -1.f)
      fail(RTG_ERR_INVALID, "phase function g must lie in (-1, 1)");
    float* g = (float*)dev_alloc(ctx, nv * sizeof(float), own);
    hip_check(hipMemcpyAsync(g, desc->medium_density, nv * sizeof(float), hipMemcpyHostToDevice, ctx->stream),
              "medium upload");
    bytes += nv * sizeof(float);
    D.grid = g;
    for (int k = 0; k < 3; k++) {
      D.gres[k] = r[k];
      D.albedo[k] = desc->medium_albedo[k];
    }
    D.gscale = scale;
    D.sigma_max = mx * scale;
    D.hg_g = desc->medium_g;
    D.medium_kind = 1;
  }
  hip_check(hipMemcpyAsync(ctx->scenes_dev + slot, &D, sizeof(DevScene), hipMemcpyHostToDevice, ctx->stream),
            "scene table");
  hip_check(hipStreamSynchronize(ctx->stream), "scene upload");
  sc->bytes = bytes;
  sc->stack = std::min(kMaxStack, std::max(kLdsStack, bvh.max_stack + 1));
  sc->ntri = ntri;
  sc->nsph = nsph;
  sc->ntex = (int)texs.size();
  sc->all_lights = strategy == RTG_LIGHTS_ALL;
  sc->mats.assign(desc->materials, desc->materials + desc->num_materials);
  ctx->scene_slots[slot] = sc.get();
  return sc.release();
}

void release_scene(rtg_ctx* ctx, rtg_scene* s) {
  if (s->users > 0) fail(RTG_ERR_INVALID, "scene is used by queued or running requests");
  (void)hipStreamSynchronize(ctx->stream);
  for (void* p : s->owned) (void)hipFree(p);
  s->owned.clear();
  if (s->slot >= 0 && s->slot < kMaxScenes && ctx->scene_slots[s->slot] == s) ctx->scene_slots[s->slot] = nullptr;
  delete s;
}

// ---------------------------------------------------------------- pool
void ensure_pool(rtg_ctx* ctx, int capacity) {
  if (ctx->pool_owned.size() && ctx->pool_alloc >= capacity) {
    // a smaller pool runs on the first `capacity` slots of the arrays already allocated (every
    // per-slot array, the all-lights ones included, holds pool_alloc slots)
    if (ctx->pool.capacity != capacity) {
      (void)hipStreamSynchronize(ctx->stream);
      ctx->pool.capacity = capacity;
    }
    return;
  }
  (void)hipStreamSynchronize(ctx->stream);
  graphs_clear(ctx);
  ctx->pool_epoch++;
  for (void* p : ctx->pool_owned) (void)hipFree(p);
  ctx->pool_owned.clear();
  DevPool P;
  std::memset(&P, 0, sizeof(P));
  P.capacity = capacity;
  auto A = [&](size_t bytes) { return dev_alloc(ctx, bytes, ctx->pool_owned); };
  size_t n = (size_t)capacity;
  P.state = (uint32_t*)A(n * 4);
  P.flags = (uint32_t*)A(n * 4);
  P.pixel = (uint32_t*)A(n * 4);
  P.sample = (uint32_t*)A(n * 4);
  P.ray_o = (float4*)A(n * 16);
  P.ray_d = (float4*)A(n * 16);
  P.hit = (float4*)A(n * 16);
  P.T = (float4*)A(n * 16);
  P.F = (float4*)A(n * 16);
  P.L = (float4*)A(n * 16);
  P.sh_d = (float4*)A(n * 16);
  P.sh_c = (float4*)A(n * 16);
  P.vol_a = (float4*)A(n * 16);
  P.vol_b = (float4*)A(n * 16);
  P.q_ext = (uint32_t*)A(n * 4);
  P.q_shade = (uint32_t*)A(n * 4 * 5);
  P.q_shadow = (uint32_t*)A(n * 4);
  P.cls_q = (uint8_t*)A(n);
  P.shflag = (uint8_t*)A(n);
  P.counters = (uint32_t*)A(C_COUNT * 4);
  P.chunk_of = (uint32_t*)A((n / kClaimChunk + 1) * 4);
  ctx->pool = P;
  ctx->pool_alloc = capacity;
}

// Shadow-ray storage of the all-lights strategy (kAllLights per slot), allocated on first use.
static void ensure_all_lights(rtg_ctx* ctx) {
  if (ctx->pool.sh_d2) return;
  (void)hipStreamSynchronize(ctx->stream);
  graphs_clear(ctx);
  ctx->pool_epoch++;
  auto A = [&](size_t bytes) { return dev_alloc(ctx, bytes, ctx->pool_owned); };
  const size_t n = (size_t)ctx->pool_alloc * kAllLights;
  ctx->pool.sh_d2 = (float4*)A(n * 16);
  ctx->pool.sh_c2 = (float4*)A(n * 16);
  ctx->pool.q_shadow2 = (uint32_t*)A(n * 4);
}

int stack_for_resident(rtg_ctx* ctx) {
  int s = 16;
  for (auto* sc : ctx->scene_slots)
    if (sc) s = std::max(s, sc->stack);
  return s;
}

// Adds the accumulating pool counters of an earlier phase of a render (C0) to C: work and ray
// counts, NaN samples and stack spills add, error flags combine.
static void merge_counters(uint32_t* C, const uint32_t* C0) {
  for (int i : {C_EXTRAYS, C_SHRAYS, C_NODES, C_PRIMS, C_SHNODES, C_SHPRIMS}) {
    uint64_t a, b;
    std::memcpy(&a, &C[i], 8);
    std::memcpy(&b, &C0[i], 8);
    a += b;
    std::memcpy(&C[i], &a, 8);
  }
  C[C_NAN] += C0[C_NAN];
  C[C_SPILL] += C0[C_SPILL];
  C[C_OVF] |= C0[C_OVF];
  C[C_GATHER] |= C0[C_GATHER];
}

LaunchCfg launch_cfg(rtg_ctx* ctx) {
  LaunchCfg cfg;
  cfg.grid = ctx->grid;
  cfg.trace_blocks = ctx->trace_blocks;
  cfg.block = 256;
  cfg.stack = stack_for_resident(ctx);
  cfg.gstack = ctx->gstack;
  cfg.refill = ctx->refill;
  cfg.regen_blocks = std::max(64, 4 * ctx->cu_count);
  kernel_grids(cfg, ctx->cu_count);
  return cfg;
}

void graphs_clear(rtg_ctx* ctx) {
  for (auto& g : ctx->graphs) (void)hipGraphExecDestroy(g.exec);
  ctx->graphs.clear();
}

void run_iterations(rtg_ctx* ctx, int mode, const LaunchCfg& cfg, int count, hipEvent_t* ev) {
  if (mode != IT_PLAIN) ev = nullptr;
  auto launch = [&]() {
    for (int b = 0; b < count; b++) {
      if (mode == IT_VOL)
        launch_iteration_vol(ctx->pool, ctx->scenes_dev, ctx->jobs_dev, ctx->order_dev, cfg, ctx->stream);
      else if (mode == IT_MIXED)
        launch_iteration_mixed(ctx->pool, ctx->scenes_dev, ctx->jobs_dev, ctx->order_dev, cfg, ctx->stream);
      else
        launch_iteration(ctx->pool, ctx->scenes_dev, ctx->jobs_dev, ctx->order_dev, cfg, ctx->stream,
                         ev ? ev + (size_t)b * kIterEvents : nullptr);
    }
  };
  // per-stage events are read with hipEventElapsedTime, which does not accept events recorded by a
  // graph's event nodes: profiled batches launch directly
  if (!ctx->use_graphs || ev) {
    launch();
    return;
  }
  GraphKey k;
  std::memset(&k, 0, sizeof(k));
  k.mode = mode;
  k.count = count;
  k.capacity = ctx->pool.capacity;
  k.all_lights = ctx->pool.all_lights;
  k.count_work = ctx->pool.count_work;
  k.grid = cfg.grid;
  k.trace_blocks = cfg.trace_blocks;
  k.stack = cfg.stack;
  k.refill = cfg.refill;
  k.epoch = ctx->pool_epoch;
  for (auto& g : ctx->graphs)
    if (g.key == k) {
      hip_check(hipGraphLaunch(g.exec, ctx->stream), "hipGraphLaunch");
      ctx->graph_launches++;
      return;
    }
  // Capture: the batch's kernels only (no memset nodes; the order and job table are uploaded
  // before the launch, outside the graph).
  if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed) != hipSuccess) {
    (void)hipGetLastError();
    ctx->use_graphs = false;  // capture unavailable: direct launches from now on
    launch();
    return;
  }
  launch();
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
  hipGraphExec_t exec = nullptr;
  if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (graph) (void)hipGraphDestroy(graph);
  hip_check(e, "graph capture");
  if (ctx->graphs.size() >= 16) {  // a bounded cache: drop the oldest configuration
    (void)hipGraphExecDestroy(ctx->graphs.front().exec);
    ctx->graphs.erase(ctx->graphs.begin());
  }
  ctx->graphs.push_back(GraphEntry{k, exec});
  ctx->graph_captures++;
  hip_check(hipGraphLaunch(exec, ctx->stream), "hipGraphLaunch");
  ctx->graph_launches++;
}

DevJob make_job(rtg_ctx* ctx, rtg_scene* scene, const rtg_render_desc& d, float4* accum) {
  if (d.width <= 0 || d.height <= 0 || d.spp <= 0) fail(RTG_ERR_INVALID, "width, height and spp must be positive");
  if (d.sampler != RTG_SAMPLER_RANDOM && d.sampler != RTG_SAMPLER_STRATIFIED) fail(RTG_ERR_INVALID, "unknown sampler");
  if (d.max_depth < 0 || d.max_depth > 4000 || d.min_depth < 0) fail(RTG_ERR_INVALID, "depth out of range");
  if (!(d.fov > 0.f && d.fov < 180.f)) fail(RTG_ERR_INVALID, "fov out of range");
  DevJob J;
  std::memset(&J, 0, sizeof(J));
  camera_matrices(d, J.s2c, J.c2w);
  J.inv_w = 1.f / d.width;
  J.inv_h = 1.f / d.height;
  J.near_clip = d.near_clip;
  J.far_clip = d.far_clip;
  J.width = d.width;
  J.height = d.height;
  int spp = d.spp;
  int sq = 1;
  if (d.sampler == RTG_SAMPLER_STRATIFIED) {
    // StratifiedSampler.cpp:23-35: a non-square count rounds up to a square
    int i = 0;
    for (i = 0; i < spp; i++)
      if (i * i >= spp) break;
    spp = i * i;
    sq = i;
  }
  J.spp = spp;
  J.sqrt_spp = sq;
  J.sampler = d.sampler;
  J.dimension = d.dimension > 0 ? d.dimension : 4;
  J.min_depth = d.min_depth;
  J.max_depth = d.max_depth;
  J.seed = d.seed;
  J.total = (uint64_t)d.width * d.height * (uint64_t)spp;
  J.next = 0;
  J.done = 0;
  J.accum = accum;
  J.scene = scene->slot;
  J.active = 1;
  J.tiles_x = (d.width + 7) / 8;
  J.integrator = d.integrator;
  if (d.guiding < RTG_GUIDING_OFF || d.guiding > RTG_GUIDING_EXIT) fail(RTG_ERR_INVALID, "unknown guiding mode");
  J.exit_mis = (d.guiding != RTG_GUIDING_OFF && d.integrator == RTG_INTEGRATOR_PATH_MIS) ? 1 : 0;
  float lut[kFilterLut];
  filter_setup(d, J, lut);
  (void)ctx;
  return J;
}

// Reconstruction filter of a render (filter/ReconstructionFilter.cpp:26-39 and the filters'
// getValue): validates the description, fills the job's filter fields and the 32x32 weight table
// (evaluated in float at (x rx / 32, y ry / 32), as the reference tabulates it).
void filter_setup(const rtg_render_desc& d, DevJob& J, float* lut) {
  if (d.filter < RTG_FILTER_BOX || d.filter > RTG_FILTER_MITCHELL) fail(RTG_ERR_INVALID, "unknown filter");
  const float rx = d.filter_radius[0] > 0.f ? d.filter_radius[0] : 0.5f;
  const float ry = d.filter_radius[1] > 0.f ? d.filter_radius[1] : 0.5f;
  if (!(rx <= 8.f && ry <= 8.f)) fail(RTG_ERR_INVALID, "filter radius above 8");
  float p0 = d.filter_param[0], p1 = d.filter_param[1];
  if (d.filter == RTG_FILTER_GAUSSIAN) {  // GaussianFilter.cpp:21-27 defaults
    if (!(p0 > 0.f)) p0 = 1.f / (2.f * 0.5f * 0.5f);
    if (!(p1 > 0.f)) p1 = rx;
  }
  auto gauss = [&](float v) { return std::max(0.f, std::exp(-p0 * v * v) - std::exp(-p0 * p1 * p1)); };
  auto mitchell = [&](float v) {
    const float B = p0, C = p1;
    v = std::fabs(2.f * v);
    if (v > 1.f)
      return ((-B - 6 * C) * v * v * v + (6 * B + 30 * C) * v * v + (-12 * B - 48 * C) * v + 8 * B + 24 * C) *
             (1.f / 6.f);
    return ((12 - 9 * B - 6 * C) * v * v * v + (-18 + 12 * B + 6 * C) * v * v + (6 - 2 * B)) * (1.f / 6.f);
  };
  for (int x = 0; x < 32; x++)
    for (int y = 0; y < 32; y++) {
      const float fx = x * rx / 32, fy = y * ry / 32;
      float v = 1.f;
      if (d.filter == RTG_FILTER_GAUSSIAN) v = gauss(fx) * gauss(fy);
      if (d.filter == RTG_FILTER_MITCHELL) v = mitchell(fx / rx) * mitchell(fy / ry);
      lut[x + y * 32] = v;
    }
  const bool plain_box = d.filter == RTG_FILTER_BOX && rx == 0.5f && ry == 0.5f;
  J.filter = plain_box ? 0 : 1;
  J.rx = rx;
  J.ry = ry;
  J.ox = (int)std::ceil(rx - 0.5f);  // Screen.cpp:101-111
  J.oy = (int)std::ceil(ry - 0.5f);
  J.lut = nullptr;
}

// Uploads a tabulated filter into the context's table of job slot `slot` and points the job at it.
void upload_job_filter(rtg_ctx* ctx, const rtg_render_desc& d, DevJob& J, int slot) {
  if (!J.filter) return;
  float lut[kFilterLut];
  DevJob tmp;
  filter_setup(d, tmp, lut);
  float* dst = ctx->lut_dev + (size_t)slot * kFilterLut;
  hip_check(hipMemcpy(dst, lut, sizeof(lut), hipMemcpyHostToDevice), "filter table upload");
  J.lut = dst;
}

}  // namespace rtg

// ==================================================================== C ABI
namespace rtg {
// Error messages are per calling thread: a failing call on one thread never races another
// thread's message (the serving entry points run concurrently with a step).
static thread_local std::string t_err;
void set_error(const char* msg) { t_err = msg; }
}  // namespace rtg

#define RTG_GUARD(ctx, body)                                   \
  try {                                                        \
    body;                                                      \
    return RTG_OK;                                             \
  } catch (const RtgError& e) {                                \
    set_error(e.what());                                       \
    return e.code;                                             \
  } catch (const std::exception& e) {                          \
    set_error(e.what());                                       \
    return RTG_ERR_INVALID;                                    \
  }

extern "C" {

int rtg_ctx_create(int device, rtg_ctx** out) {
  if (!out) return RTG_ERR_INVALID;
  *out = nullptr;
  rtg_ctx* ctx = new rtg_ctx();
  try {
    int n = 0;
    hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n) fail(RTG_ERR_INVALID, "no such device");
    ctx->device = device;
    hip_check(hipSetDevice(device), "hipSetDevice");
    hip_check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    ctx->cu_count = prop.multiProcessorCount;
    ctx->arch = prop.gcnArchName;
    if (ctx->arch.rfind("gfx950", 0) != 0) fail(RTG_ERR_INVALID, "librtg is built for gfx950, device is " + ctx->arch);
    hip_check(hipMalloc(&ctx->scenes_dev, sizeof(DevScene) * kMaxScenes), "hipMalloc scenes");
    hip_check(hipMalloc(&ctx->jobs_dev, sizeof(DevJob) * kMaxJobs), "hipMalloc jobs");
    // every job-table entry starts inactive with nothing claimed (a render uploads only its own)
    launch_fill_zero(ctx->jobs_dev, sizeof(DevJob) * kMaxJobs, ctx->stream);
    hip_check(hipMalloc(&ctx->order_dev, sizeof(int32_t) * (kMaxJobs + 1)), "hipMalloc job order");
    hip_check(hipMalloc(&ctx->lut_dev, sizeof(float) * kFilterLut * kMaxJobs), "hipMalloc filter tables");
    hip_check(hipHostMalloc(&ctx->pinned, 4096, hipHostMallocDefault), "hipHostMalloc");
    ctx->scene_slots.assign(kMaxScenes, nullptr);
    ctx->pool_slots = 1 << 21;
    ctx->grid = std::max(64, ctx->cu_count * 5);
    // tuning knobs of the persistent traversal: lanes idle before a wave refills (1..64), and
    // waves per SIMD of a traversal launch (blocks of 128 = 2 waves; 4 SIMDs per CU)
    if (const char* e = std::getenv("RTG_TRACE_REFILL")) ctx->refill = std::min(64, std::max(1, std::atoi(e)));
    int waves = kTraceWaves;
    if (const char* e = std::getenv("RTG_TRACE_WAVES")) waves = std::min(kMaxTraceWaves, std::max(1, std::atoi(e)));
    ctx->trace_waves = waves;
    ctx->trace_blocks = std::max(32, ctx->cu_count * 2 * waves);
    // traversal stack spill area for every lane of a traversal launch of up to kMaxTraceWaves waves
    size_t lanes = (size_t)std::max(32, ctx->cu_count * 2 * kMaxTraceWaves) * 128;
    hip_check(hipMalloc(&ctx->gstack, lanes * (kMaxStack - kLdsStack) * sizeof(uint2)), "hipMalloc stack spill");
    hip_check(hipMalloc(&ctx->tap_ctr, 2 * sizeof(uint32_t)), "hipMalloc trace counters");
    // device address of this GPU's static scratch array (the allocation-free trace's scratch); a
    // failed lookup leaves it null, and only rtg_trace_device_noalloc then fails
    ctx->scratch = scratch_address(device);
  } catch (const RtgError& e) {
    int code = e.code;
    delete ctx;
    return code;
  }
  *out = ctx;
  return RTG_OK;
}

int rtg_ctx_destroy(rtg_ctx* ctx) {
  if (!ctx) return RTG_ERR_INVALID;
  if (ctx->server) (void)rtg_server_destroy(ctx->server);  // drops its requests' scene references
  guide_destroy(ctx);
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto* s : ctx->scene_slots)
    if (s) {
      try {
        release_scene(ctx, s);
      } catch (const std::exception&) {
      }
    }
  graphs_clear(ctx);
  for (void* p : ctx->pool_owned) (void)hipFree(p);
  if (ctx->scenes_dev) (void)hipFree(ctx->scenes_dev);
  if (ctx->jobs_dev) (void)hipFree(ctx->jobs_dev);
  if (ctx->gstack) (void)hipFree(ctx->gstack);
  if (ctx->tap_ctr) (void)hipFree(ctx->tap_ctr);
  if (ctx->order_dev) (void)hipFree(ctx->order_dev);
  if (ctx->lut_dev) (void)hipFree(ctx->lut_dev);
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  for (void* p : ctx->scratch_owned) (void)hipFree(p);
  for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return RTG_OK;
}

const char* rtg_last_error(rtg_ctx* ctx) {
  (void)ctx;
  return t_err.c_str();
}

int rtg_ctx_set_pool(rtg_ctx* ctx, int32_t slots) {
  if (!ctx) return RTG_ERR_INVALID;
  RTG_GUARD(ctx, {
    if (slots < 0 || slots > (1 << 26)) fail(RTG_ERR_INVALID, "pool size out of range");
    ctx->pool_slots = slots == 0 ? (1 << 21) : ((slots + 255) / 256) * 256;
  })
}

int rtg_ctx_get_tuning(rtg_ctx* ctx, rtg_tuning* out) {
  if (!ctx || !out) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  out->pool_slots = ctx->pool_slots;
  out->trace_waves = ctx->trace_waves;
  out->trace_refill = ctx->refill;
  out->graphs = ctx->use_graphs ? 1 : 0;
  return RTG_OK;
}

static void apply_tuning(rtg_ctx* ctx, const rtg_tuning& t) {
  if (t.pool_slots < 256 || t.pool_slots > (1 << 26)) fail(RTG_ERR_INVALID, "pool_slots out of range");
  if (t.trace_waves < 1 || t.trace_waves > kMaxTraceWaves) fail(RTG_ERR_INVALID, "trace_waves out of range");
  if (t.trace_refill < 1 || t.trace_refill > 64) fail(RTG_ERR_INVALID, "trace_refill out of range");
  ctx->pool_slots = ((t.pool_slots + 255) / 256) * 256;
  ctx->trace_waves = t.trace_waves;
  ctx->trace_blocks = std::max(32, ctx->cu_count * 2 * t.trace_waves);
  ctx->refill = t.trace_refill;
  ctx->use_graphs = t.graphs != 0;
}

int rtg_ctx_set_tuning(rtg_ctx* ctx, const rtg_tuning* t) {
  if (!ctx || !t) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  RTG_GUARD(ctx, {
    if (ctx->server) fail(RTG_ERR_INVALID, "the context is serving requests; tune it before creating the server");
    apply_tuning(ctx, *t);
  })
}

int rtg_ctx_set_profiling(rtg_ctx* ctx, int32_t flags) {
  if (!ctx) return RTG_ERR_INVALID;
  ctx->profiling = flags;
  return RTG_OK;
}

int rtg_scene_upload(rtg_ctx* ctx, const rtg_scene_desc* desc, rtg_scene** out) {
  if (!ctx || !out) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  RTG_GUARD(ctx, {
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    *out = build_scene(ctx, desc);
  })
}

int rtg_scene_release(rtg_ctx* ctx, rtg_scene* scene) {
  if (!ctx || !scene) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  RTG_GUARD(ctx, { release_scene(ctx, scene); })
}

int rtg_scene_query(rtg_scene* scene, rtg_scene_info* info) {
  if (!scene || !info) return RTG_ERR_INVALID;
  std::memset(info, 0, sizeof(*info));
  info->triangles = scene->ntri;
  info->spheres = scene->nsph;
  info->materials = scene->dev.nmat;
  info->lights = scene->dev.nlights;
  info->env_width = scene->dev.env_w;
  info->env_height = scene->dev.env_h;
  for (int k = 0; k < 3; k++) {
    info->bbox_min[k] = scene->dev.root_min[k];
    info->bbox_max[k] = scene->dev.root_max[k];
  }
  info->textures = scene->ntex;
  info->traversal_stack = scene->stack;
  info->medium_kind = scene->dev.medium_kind;
  for (int k = 0; k < 3; k++) {
    info->medium_sigma_a[k] = scene->dev.sig_a[k];
    info->medium_sigma_s[k] = scene->dev.sig_s[k];
  }
  if (scene->dev.medium_kind == 2) {
    info->medium_g = scene->dev.hg_g;
    info->medium_scale = scene->dev.gscale;
  }
  if (scene->dev.grid) {
    for (int k = 0; k < 3; k++) {
      info->medium_res[k] = scene->dev.gres[k];
      info->medium_albedo[k] = scene->dev.albedo[k];
    }
    info->medium_g = scene->dev.hg_g;
    info->medium_scale = scene->dev.gscale;
    info->medium_sigma_max = scene->dev.sigma_max;
  }
  return RTG_OK;
}

int rtg_scene_material(rtg_scene* scene, int32_t index, rtg_material* out) {
  if (!scene || !out) return RTG_ERR_INVALID;
  if (index < 0 || index >= (int32_t)scene->mats.size()) {
    set_error("material index out of range");
    return RTG_ERR_INVALID;
  }
  *out = scene->mats[(size_t)index];
  return RTG_OK;
}

int rtg_scene_bytes(rtg_scene* scene, uint64_t* bytes) {
  if (!scene || !bytes) return RTG_ERR_INVALID;
  *bytes = scene->bytes;
  return RTG_OK;
}

static void render_impl(rtg_ctx* ctx, rtg_scene* scene, const rtg_render_desc* desc, float* out_dev, rtg_stats* st) {
  if (!scene || !desc) fail(RTG_ERR_INVALID, "null scene or render description");
  if (ctx->server)
    fail(RTG_ERR_INVALID, "the context is serving requests, which do not take guided or all-lights renders");
  hip_check(hipSetDevice(ctx->device), "hipSetDevice");
  const bool guided = desc->guiding == RTG_GUIDING_LEARNED;  // the network (exit proposals: per job)
  const bool vol = desc->integrator == RTG_INTEGRATOR_VOLPATH;
  if (desc->integrator != RTG_INTEGRATOR_PATH_MIS && !vol) fail(RTG_ERR_INVALID, "unknown integrator");
  if (vol && desc->guiding) fail(RTG_ERR_INVALID, "guiding is not available with the volumetric integrator");
  // all_lights: PathTracingMIS samples every light; VolPathTracing always samples one
  // (Integrator::sampleAttenuatedLightDirect, with uniform selection for this strategy)
  const bool all_lights = scene->all_lights && !vol;
  if (all_lights && guided) fail(RTG_ERR_INVALID, "guiding is not available with the all_lights strategy");
  DevJob J = make_job(ctx, scene, *desc, (float4*)out_dev);
  upload_job_filter(ctx, *desc, J, 0);
  // Learned guiding trains between iteration batches and samples plainly until its first steps are
  // done: a pool that holds most of the render at once would draw nearly every sample before the
  // network exists, so guided renders use at most kGuidedPoolMax slots.
  const int pool_cap = guided ? std::min(ctx->pool_slots, kGuidedPoolMax) : ctx->pool_slots;
  // A pool with a slot for every sample (in whole claim chunks) holds the whole render at once: with
  // the box filter, pixels are then summed from the slots after the last iteration (k_gather). The
  // decision takes the pool this render runs on, not the context's setting.
  J.gather = (J.filter == 0 && (uint64_t)pool_cap >= J.total && pool_cap % kClaimChunk == 0) ? 1 : 0;
  // pinned page: [0] claim counter and done count, [16] guide ring and inference counts, [40] guide
  // loss, [48] trace tap counters (rtg_trace), [56] claim limit upload, [64] claim order, [256] job
  const uint64_t* next_h = (const uint64_t*)ctx->pinned;
  uint64_t* done = (uint64_t*)ctx->pinned + 1;
  uint64_t* up_total = (uint64_t*)((char*)ctx->pinned + 56);
  static_assert(offsetof(DevJob, done) == offsetof(DevJob, next) + 8, "claim counter and done count adjacent");
  unsigned long long* ring_cnt = (unsigned long long*)((char*)ctx->pinned + 16);
  float* loss_h = (float*)((char*)ctx->pinned + 40);
  int32_t* up_order = (int32_t*)((char*)ctx->pinned + 64);
  DevJob* up_job = (DevJob*)((char*)ctx->pinned + 256);
  static_assert(256 + sizeof(DevJob) <= 4096, "pinned staging page too small");
  up_order[0] = 1;  // one job, in slot 0
  up_order[1] = 0;
  *up_job = J;
  hip_check(hipMemcpyAsync(ctx->order_dev, up_order, 2 * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream),
            "order");
  ensure_pool(ctx, pool_cap);
  if (J.gather) {
    if ((uint64_t)ctx->pool.capacity < J.total) fail(RTG_ERR_HIP, "gather render on a pool smaller than its samples");
    // every sample chunk unclaimed until this render's first regeneration claims it
    launch_fill_zero(ctx->pool.chunk_of, (size_t)(J.total + kClaimChunk - 1) / kClaimChunk * sizeof(uint32_t),
                     ctx->stream);
  }
  if (all_lights) ensure_all_lights(ctx);
  ctx->pool.all_lights = all_lights ? 1 : 0;
  ctx->pool.count_work = (ctx->profiling & RTG_PROFILE_COUNT) ? 1 : 0;
  LaunchCfg cfg = launch_cfg(ctx);
  launch_reset_pool(ctx->pool, ctx->stream);
  hip_check(hipMemcpyAsync(ctx->jobs_dev, up_job, sizeof(DevJob), hipMemcpyHostToDevice, ctx->stream), "job upload");
  GuideState* gs = nullptr;
  DevGuide G;
  std::memset(&G, 0, sizeof(G));
  // Guided renders sample plainly (recording training vertices) for the first kWarmSteps training
  // steps, which run kWarmPerBatch at a time on the newest samples (a sample may serve several
  // steps); then one-sample MIS with the guide, with one step per kGuideBatch new training samples
  // of the iteration batch (at least one, at most kMaxStepsPerBatch) until kMaxSteps, after which
  // training and the recording that feeds it stop.
  const uint64_t kWarmSteps = 64, kMaxSteps = kGuideMaxSteps;
  const int kWarmPerBatch = kGuideStepsPerBatch;
  const int kMaxStepsPerBatch = kWarmPerBatch;
  // On a context whose pool is larger than the guided cap, the network trains in a first phase on
  // the capped pool; once training ends (or 1/kGuidePhaseDiv of the samples is claimed) that
  // pool stops claiming and drains, and the rest of the render runs on the full pool with the
  // trained network, inference only. Phase: 0 single pool, 1 training, 2 draining, 3 inference.
  int phase = (guided && !J.gather && ctx->pool_slots > pool_cap) ? 1 : 0;
  uint64_t cut = J.total;  // samples of the training phase
  uint32_t C0[C_COUNT] = {};  // pool counters at the end of the training phase
  if (guided) {
    gs = guide_state(ctx);
    guide_prepare_render(ctx, gs, scene, desc->seed, G);
    G.sample = 0;
  }
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "event");
  hip_check(hipEventCreate(&e1), "event");
  hip_check(hipEventRecord(e0, ctx->stream), "event");
  uint64_t iters = 0;
  const int batch = 4;
  const int kEv = kIterEvents;   // events per iteration (8 stages)
  const int kTrainEv = 2 * kMaxStepsPerBatch;
  const bool prof = (ctx->profiling & RTG_PROFILE_EVENTS) != 0 && !vol;
  double stage_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double train_ms = 0.0;
  uint64_t train_steps = 0, trained_upto = 0, ring_total = 0;
  double last_loss = 0.0;
  auto wall0 = std::chrono::steady_clock::now();
  const size_t nev = (size_t)batch * kEv + kTrainEv;
  if (prof && ctx->events.size() < nev) {
    size_t have = ctx->events.size();
    ctx->events.resize(nev);
    for (size_t i = have; i < ctx->events.size(); i++) hip_check(hipEventCreate(&ctx->events[i]), "event");
  }
  int pending_train = 0;  // training steps enqueued before the batch now in flight
  while (true) {
    if (guided) {
      // the guide's state changes between batches: direct launches
      for (int b = 0; b < batch; b++)
        launch_iteration(ctx->pool, ctx->scenes_dev, ctx->jobs_dev, ctx->order_dev, cfg, ctx->stream,
                         prof ? &ctx->events[(size_t)b * kEv] : nullptr, &G);
    } else {
      run_iterations(ctx, vol ? IT_VOL : IT_PLAIN, cfg, batch, prof ? ctx->events.data() : nullptr);
    }
    iters += batch;
    hip_check(hipGetLastError(), "kernel launch");
    hip_check(hipMemcpyAsync(ctx->pinned, &ctx->jobs_dev[0].next, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                             ctx->stream),
              "done readback");
    if (guided) {
      hip_check(hipMemcpyAsync(ring_cnt, gs->ring_count, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                               ctx->stream),
                "ring readback");
      hip_check(hipMemcpyAsync(loss_h, gs->S.scal + 1, sizeof(float), hipMemcpyDeviceToHost, ctx->stream),
                "loss readback");
    }
    hip_check(hipStreamSynchronize(ctx->stream), "render iteration");
    if (prof) {
      for (int b = 0; b < batch; b++)
        for (int k = 0; k < 8; k++) {
          float ms = 0.f;
          hip_check(hipEventElapsedTime(&ms, ctx->events[(size_t)b * kEv + k], ctx->events[(size_t)b * kEv + k + 1]),
                    "event time");
          stage_ms[k] += ms;
        }
      for (int t = 0; t < pending_train; t++) {
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, ctx->events[(size_t)batch * kEv + 2 * t],
                                      ctx->events[(size_t)batch * kEv + 2 * t + 1]),
                  "event time");
        train_ms += ms;
      }
    }
    if (pending_train) last_loss = *loss_h;
    pending_train = 0;
    if (*done >= J.total) break;
    if (phase == 2 && *done >= cut) {
      // the capped pool has drained: the rest of the render on the full pool, inference only (the
      // pool reset clears the counters, so the training phase's are kept on the host)
      hip_check(hipMemcpy(C0, ctx->pool.counters, sizeof(C0), hipMemcpyDeviceToHost), "counters");
      ensure_pool(ctx, ctx->pool_slots);
      ctx->pool.count_work = (ctx->profiling & RTG_PROFILE_COUNT) ? 1 : 0;
      cfg = launch_cfg(ctx);
      launch_reset_pool(ctx->pool, ctx->stream);
      G = guide_infer_phase(ctx, gs, G, ctx->pool.capacity);
      *up_total = J.total;
      hip_check(hipMemcpyAsync(&ctx->jobs_dev[0].total, up_total, sizeof(uint64_t), hipMemcpyHostToDevice,
                               ctx->stream),
                "claim limit");
      phase = 3;
      continue;
    }
    if (guided) {
      // train on the newest samples: one step per full batch of new samples (at least one), at
      // most kMaxStepsPerBatch per iteration batch
      ring_total = *ring_cnt;
      const uint64_t B = kGuideBatch;
      int steps = 0;
      if (G.record && ring_total >= B && ring_total > trained_upto) {
        const uint64_t fresh = std::max<uint64_t>(1, (ring_total - trained_upto) / B);
        steps = train_steps < kWarmSteps ? kWarmPerBatch
                                         : (int)std::min<uint64_t>(fresh, (uint64_t)kMaxStepsPerBatch);
        steps = (int)std::min<uint64_t>((uint64_t)steps, kMaxSteps - train_steps);
      }
      // step t trains on the newest samples, shifted back by t half-batches (within the ring)
      const uint64_t span = std::min<uint64_t>(ring_total, (uint64_t)G.ring_mask + 1);
      for (int t = 0; t < steps; t++) {
        const uint64_t back = ((uint64_t)t * (B / 2)) % (span - B + 1);
        const uint64_t base = ring_total - B - back;
        if (prof) (void)hipEventRecord(ctx->events[(size_t)batch * kEv + 2 * t], ctx->stream);
        guide_train_step(ctx, gs, G, gs->ring, base, G.ring_mask, (int)B, 1e-3f);
        if (prof) (void)hipEventRecord(ctx->events[(size_t)batch * kEv + 2 * t + 1], ctx->stream);
      }
      if (steps) {
        trained_upto = ring_total;
        train_steps += steps;
        pending_train = steps;
        if (train_steps >= kWarmSteps) G.sample = 1;
        if (train_steps >= kMaxSteps) G.record = 0;
      }
      if (phase == 1 && (!G.record || *next_h >= J.total / kGuidePhaseDiv) && *next_h < J.total) {
        // end of the training phase: no further claims on the capped pool (no kernel runs now, so
        // the claim counter read back is current)
        cut = *next_h;
        *up_total = cut;
        hip_check(hipMemcpyAsync(&ctx->jobs_dev[0].total, up_total, sizeof(uint64_t), hipMemcpyHostToDevice,
                                 ctx->stream),
                  "claim limit");
        G.record = 0;
        phase = 2;
      }
    }
    double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
    if (iters > 200000 || wall > 600.0) fail(RTG_ERR_HIP, "render did not finish (iteration or time limit)");
  }
  if (J.gather) {
    launch_gather(ctx->jobs_dev, 0, ctx->pool, J.width * J.height, ctx->stream);
    hip_check(hipGetLastError(), "gather launch");
  }
  hip_check(hipEventRecord(e1, ctx->stream), "event");
  hip_check(hipEventSynchronize(e1), "event");
  ctx->pool.all_lights = 0;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  uint32_t C[C_COUNT];
  hip_check(hipMemcpy(C, ctx->pool.counters, sizeof(C), hipMemcpyDeviceToHost), "counters");
  if (phase == 3) merge_counters(C, C0);
  if (C[C_OVF]) fail(RTG_ERR_HIP, "BVH traversal stack overflow (the builder's depth bound was exceeded)");
  if (C[C_GATHER]) fail(RTG_ERR_HIP, "gather met a sample chunk outside the path pool");
  if (st) {
    auto u64 = [&](int i) {
      uint64_t v;
      std::memcpy(&v, &C[i], 8);
      return v;
    };
    std::memset(st, 0, sizeof(*st));
    st->samples = *done;
    st->extension_rays = u64(C_EXTRAYS);
    st->shadow_rays = u64(C_SHRAYS);
    st->nan_samples = C[C_NAN];
    st->iterations = iters;
    st->milliseconds = ms;
    st->nodes_visited = u64(C_NODES);
    st->prims_tested = u64(C_PRIMS);
    st->shadow_nodes_visited = u64(C_SHNODES);
    st->shadow_prims_tested = u64(C_SHPRIMS);
    st->stack_spills = C[C_SPILL];
    for (int k = 0; k < 8; k++) {
      st->stage_ms[k] = stage_ms[k];
      st->stage_launches[k] = prof ? iters : 0;
    }
    if (guided) {
      st->guide_train_steps = train_steps;
      st->guide_samples = *ring_cnt;
      st->guide_loss = last_loss;
      st->guide_train_ms = train_ms;
      st->guide_queries = ring_cnt[1];
    }
  }
}

// With a server on the context, a render that the server can take (no learned guiding, not
// all-lights PathTracingMIS) joins its mix as a request and the call waits for it; other renders fail.
static bool servable(const rtg_scene* scene, const rtg_render_desc* desc) {
  return scene && desc && desc->guiding != RTG_GUIDING_LEARNED &&
         !(scene->all_lights && desc->integrator == RTG_INTEGRATOR_PATH_MIS);
}

int rtg_autotune(rtg_ctx* ctx, rtg_scene* scene, const rtg_render_desc* desc, uint64_t trial_samples, rtg_tuning* best,
                 double* best_samples_per_s) {
  if (!ctx) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  RTG_GUARD(ctx, {
    if (!scene || !desc) fail(RTG_ERR_INVALID, "null scene or render description");
    if (ctx->server) fail(RTG_ERR_INVALID, "the context is serving requests; tune it before creating the server");
    if (desc->width <= 0 || desc->height <= 0) fail(RTG_ERR_INVALID, "bad size");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    rtg_render_desc d = *desc;
    d.guiding = 0;
    const uint64_t npix = (uint64_t)d.width * d.height;
    d.spp = (int32_t)std::max<uint64_t>(1, std::min<uint64_t>(1u << 20, (trial_samples ? trial_samples : (1u << 24)) / npix));
    if (d.sampler == RTG_SAMPLER_STRATIFIED) {  // the count the sampler will use (rounded up to a square)
      int q = 1;
      while (q * q < d.spp) q++;
      d.spp = q * q;
    }
    const size_t bytes = npix * 16;
    void* film = nullptr;
    hip_check(hipMalloc(&film, bytes), "hipMalloc tuning film");
    const int saved_prof = ctx->profiling;
    ctx->profiling = 0;
    rtg_tuning cur;
    cur.pool_slots = ctx->pool_slots;
    cur.trace_waves = ctx->trace_waves;
    cur.trace_refill = ctx->refill;
    cur.graphs = ctx->use_graphs ? 1 : 0;
    // samples per second of a configuration: the faster of two renders after a warm-up render
    auto rate = [&](const rtg_tuning& t) {
      apply_tuning(ctx, t);
      double bestr = 0.0;
      for (int rep = 0; rep < 3; rep++) {
        launch_fill_zero(film, bytes, ctx->stream);
        rtg_stats st;
        const auto t0 = std::chrono::steady_clock::now();
        render_impl(ctx, scene, &d, (float*)film, &st);
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (rep > 0) bestr = std::max(bestr, (double)st.samples / s);
      }
      return bestr;
    };
    try {
      double r_best = rate(cur);
      rtg_tuning t_best = cur;
      // coordinate search: the pool size first (it changes the launch shapes the most: a render
      // whose samples are all in flight takes its fewest, largest iterations), then the traversal
      // waves and the refill threshold at that pool (the best waves depend on the pool).
      // Pools up to 2^26 slots (~10 GB at ~150 B per slot); a pool larger than twice the trial's
      // sample count could not fill and is not tried.
      const uint64_t trial_total = npix * (uint64_t)d.spp;
      for (int p : {1 << 20, 3 << 19, 1 << 21, 3 << 20, 1 << 22, 3 << 21, 1 << 23, 3 << 22, 1 << 24, 3 << 23,
                    1 << 25, 3 << 24, 1 << 26}) {
        rtg_tuning t = t_best;
        t.pool_slots = p;
        if (p == t_best.pool_slots || (uint64_t)p > 2 * trial_total) continue;
        const double r = rate(t);
        if (r > r_best) { r_best = r; t_best = t; }
      }
      for (int w : {3, 4, 5, 6, 8}) {
        rtg_tuning t = t_best;
        t.trace_waves = w;
        if (w == t_best.trace_waves) continue;
        const double r = rate(t);
        if (r > r_best) { r_best = r; t_best = t; }
      }
      for (int f : {4, 8, 16, 32}) {
        rtg_tuning t = t_best;
        t.trace_refill = f;
        if (f == t_best.trace_refill) continue;
        const double r = rate(t);
        if (r > r_best) { r_best = r; t_best = t; }
      }
      apply_tuning(ctx, t_best);
      if (best) *best = t_best;
      if (best_samples_per_s) *best_samples_per_s = r_best;
    } catch (...) {
      ctx->profiling = saved_prof;
      (void)hipFree(film);
      throw;
    }
    ctx->profiling = saved_prof;
    (void)hipFree(film);
  })
}

int rtg_render_device(rtg_ctx* ctx, rtg_scene* scene, const rtg_render_desc* desc, float* out_dev, rtg_stats* st) {
  if (!ctx) return RTG_ERR_INVALID;
  rtg_server* srv = nullptr;
  {
    std::lock_guard<std::mutex> lk(ctx->mtx);
    srv = ctx->server;
    if (!srv || !servable(scene, desc)) {
      RTG_GUARD(ctx, {
        if (!out_dev) fail(RTG_ERR_INVALID, "null output");
        render_impl(ctx, scene, desc, out_dev, st);
      })
    }
  }
  if (!out_dev) {
    set_error("null output");
    return RTG_ERR_INVALID;
  }
  return serve_render(srv, scene, desc, nullptr, out_dev, st);
}

static int render_blocking_impl(rtg_ctx* ctx, rtg_scene* scene, const rtg_render_desc* desc, float* out_rgbw,
                                rtg_stats* st);

int rtg_render(rtg_ctx* ctx, rtg_scene* scene, const rtg_render_desc* desc, float* out_rgbw, rtg_stats* st) {
  if (!ctx) return RTG_ERR_INVALID;
  rtg_server* srv = nullptr;
  {
    std::lock_guard<std::mutex> lk(ctx->mtx);
    srv = ctx->server;
    if (!srv || !servable(scene, desc)) return render_blocking_impl(ctx, scene, desc, out_rgbw, st);
  }
  if (!out_rgbw) {
    set_error("null output");
    return RTG_ERR_INVALID;
  }
  return serve_render(srv, scene, desc, out_rgbw, nullptr, st);
}

static int render_blocking_impl(rtg_ctx* ctx, rtg_scene* scene, const rtg_render_desc* desc, float* out_rgbw,
                                rtg_stats* st) {
  RTG_GUARD(ctx, {
    if (!out_rgbw || !desc) fail(RTG_ERR_INVALID, "null output");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    size_t bytes = (size_t)desc->width * desc->height * 16;
    if (desc->width <= 0 || desc->height <= 0) fail(RTG_ERR_INVALID, "bad size");
    void* acc = nullptr;
    hip_check(hipMalloc(&acc, bytes), "hipMalloc film");
    try {
      launch_fill_zero(acc, bytes, ctx->stream);
      render_impl(ctx, scene, desc, (float*)acc, st);
      hip_check(hipMemcpy(out_rgbw, acc, bytes, hipMemcpyDeviceToHost), "film readback");
    } catch (...) {
      (void)hipFree(acc);
      throw;
    }
    (void)hipFree(acc);
  })
}

int rtg_trace(rtg_ctx* ctx, rtg_scene* scene, const float* rays, int32_t n, int32_t any_hit, int32_t* prim, float* t,
              float* uv, float* p, float* nrm) {
  if (!ctx) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  RTG_GUARD(ctx, {
    if (!scene || !rays || !prim || n < 0) fail(RTG_ERR_INVALID, "bad trace arguments");
    if (n == 0) return RTG_OK;
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    std::vector<void*> own;
    float* dr = (float*)dev_alloc(ctx, (size_t)n * 32, own);
    int32_t* dp = (int32_t*)dev_alloc(ctx, (size_t)n * 4, own);
    float* dt = (float*)dev_alloc(ctx, (size_t)n * 4, own);
    float* duv = (float*)dev_alloc(ctx, (size_t)n * 8, own);
    float* dpp = (float*)dev_alloc(ctx, (size_t)n * 12, own);
    float* dn = (float*)dev_alloc(ctx, (size_t)n * 12, own);
    try {
      hip_check(hipMemcpyAsync(dr, rays, (size_t)n * 32, hipMemcpyHostToDevice, ctx->stream), "rays");
      LaunchCfg cfg;
      cfg.grid = ctx->grid;
      cfg.trace_blocks = ctx->trace_blocks;
      cfg.block = 128;
      cfg.stack = scene->stack;
      cfg.gstack = ctx->gstack;
      cfg.refill = ctx->refill;
      launch_fill_zero(ctx->tap_ctr, 2 * sizeof(uint32_t), ctx->stream);
      launch_trace_tap(ctx->scenes_dev + scene->slot, dr, n, any_hit, dp, dt, duv, dpp, dn, ctx->tap_ctr, cfg,
                       ctx->stream);
      hip_check(hipGetLastError(), "trace launch");
      uint32_t* ctr_h = (uint32_t*)((char*)ctx->pinned + 48);
      hip_check(hipMemcpyAsync(ctr_h, ctx->tap_ctr, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream),
                "trace counters");
      hip_check(hipMemcpyAsync(prim, dp, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream), "prim");
      if (t) hip_check(hipMemcpyAsync(t, dt, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream), "t");
      if (uv) hip_check(hipMemcpyAsync(uv, duv, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream), "uv");
      if (p) hip_check(hipMemcpyAsync(p, dpp, (size_t)n * 12, hipMemcpyDeviceToHost, ctx->stream), "p");
      if (nrm) hip_check(hipMemcpyAsync(nrm, dn, (size_t)n * 12, hipMemcpyDeviceToHost, ctx->stream), "n");
      hip_check(hipStreamSynchronize(ctx->stream), "trace");
      ctx->last_trace_spills = ctr_h[0];
      if (ctr_h[1]) fail(RTG_ERR_HIP, "BVH traversal stack overflow (the builder's depth bound was exceeded)");
    } catch (...) {
      for (void* q : own) (void)hipFree(q);
      throw;
    }
    for (void* q : own) (void)hipFree(q);
  })
}

int rtg_sampler_batch(rtg_ctx* ctx, const rtg_render_desc* desc, const uint32_t* pixel_sample, int32_t n,
                      int32_t n2, int32_t n1, float* out2, float* out1) {
  if (!ctx) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  RTG_GUARD(ctx, {
    if (!desc || !pixel_sample || n < 0 || n2 < 0 || n1 < 0 || (n2 && !out2) || (n1 && !out1))
      fail(RTG_ERR_INVALID, "bad sampler tap arguments");
    if (n == 0) return RTG_OK;
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    rtg_scene dummy;
    dummy.slot = 0;
    DevJob J = make_job(ctx, &dummy, *desc, nullptr);  // validates and rounds the sample count
    std::vector<void*> own;
    uint32_t* dps = (uint32_t*)dev_alloc(ctx, (size_t)n * 8, own);
    float* d2 = n2 ? (float*)dev_alloc(ctx, (size_t)n * n2 * 8, own) : nullptr;
    float* d1 = n1 ? (float*)dev_alloc(ctx, (size_t)n * n1 * 4, own) : nullptr;
    try {
      hip_check(hipMemcpyAsync(dps, pixel_sample, (size_t)n * 8, hipMemcpyHostToDevice, ctx->stream), "pairs");
      launch_sampler_tap(J, dps, n, n2, n1, d2, d1, ctx->stream);
      hip_check(hipGetLastError(), "sampler tap launch");
      if (n2) hip_check(hipMemcpyAsync(out2, d2, (size_t)n * n2 * 8, hipMemcpyDeviceToHost, ctx->stream), "2d");
      if (n1) hip_check(hipMemcpyAsync(out1, d1, (size_t)n * n1 * 4, hipMemcpyDeviceToHost, ctx->stream), "1d");
      hip_check(hipStreamSynchronize(ctx->stream), "sampler tap");
    } catch (...) {
      for (void* q : own) (void)hipFree(q);
      throw;
    }
    for (void* q : own) (void)hipFree(q);
  })
}

int rtg_trace_device_noalloc(rtg_ctx* ctx, rtg_scene* scene, const float* rays_device, int32_t n, int32_t any_hit,
                             int32_t* prim_device, float* t_device, float* uv_device, float* p_device,
                             float* nrm_device) {
  if (!ctx) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  RTG_GUARD(ctx, {
    if (!scene || !rays_device || !prim_device || n < 0) fail(RTG_ERR_INVALID, "bad trace arguments");
    if (scene->stack > kNoAllocStack)
      fail(RTG_ERR_INVALID, "the scene's traversal stack bound (" + std::to_string(scene->stack) +
                                " entries) exceeds the " + std::to_string(kNoAllocStack) +
                                " the allocation-free trace's scratch holds; use rtg_trace");
    if (!ctx->scratch) fail(RTG_ERR_HIP, "hipGetSymbolAddress of the scratch array failed at context creation");
    if (n == 0) return RTG_OK;
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    launch_trace_tap_noalloc(ctx->scenes_dev + scene->slot, rays_device, n, any_hit, prim_device, t_device, uv_device,
                             p_device, nrm_device, ctx->scratch, scene->stack, ctx->refill, ctx->stream);
    hip_check(hipGetLastError(), "trace launch");
    uint32_t* ctr_h = (uint32_t*)((char*)ctx->pinned + 48);
    hip_check(hipMemcpyAsync(ctr_h, ctx->scratch, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream),
              "trace counters");
    hip_check(hipStreamSynchronize(ctx->stream), "trace");
    ctx->last_trace_spills = ctr_h[0];
    if (ctr_h[1]) fail(RTG_ERR_HIP, "BVH traversal stack overflow (the builder's depth bound was exceeded)");
  })
}

int rtg_trace_spills(rtg_ctx* ctx, uint64_t* spilled_pushes) {
  if (!ctx || !spilled_pushes) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  *spilled_pushes = ctx->last_trace_spills;
  return RTG_OK;
}

int rtg_tonemap(rtg_ctx* ctx, const float* rgbw, int32_t width, int32_t height, int32_t mode, float exposure,
                uint8_t* out_rgb) {
  if (!ctx) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  RTG_GUARD(ctx, {
    if (!rgbw || !out_rgb || width <= 0 || height <= 0 || mode < RTG_TONEMAP_CLAMP || mode > RTG_TONEMAP_REINHARD)
      fail(RTG_ERR_INVALID, "bad tone-mapping arguments");
    if ((int64_t)width * height > (1ll << 28)) fail(RTG_ERR_INVALID, "image too large");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const int n = width * height;
    // gammaToneMapper (core/main.cpp:193): exposure = pow(2., (exposure - 0.5) * 20) in double
    const float expo = (float)std::pow(2., ((double)exposure - 0.5) * 20);
    std::vector<void*> own;
    float4* dfilm = (float4*)dev_alloc(ctx, (size_t)n * 16, own);
    uint8_t* dout = (uint8_t*)dev_alloc(ctx, (size_t)n * 3, own);
    unsigned int* dmax = (unsigned int*)dev_alloc(ctx, 4, own);
    try {
      const unsigned int zero = 0;
      hip_check(hipMemcpyAsync(dfilm, rgbw, (size_t)n * 16, hipMemcpyHostToDevice, ctx->stream), "film");
      hip_check(hipMemcpyAsync(dmax, &zero, 4, hipMemcpyHostToDevice, ctx->stream), "max");
      launch_tonemap(dfilm, n, mode, expo, dmax, dout, ctx->stream);
      hip_check(hipGetLastError(), "tonemap launch");
      hip_check(hipMemcpyAsync(out_rgb, dout, (size_t)n * 3, hipMemcpyDeviceToHost, ctx->stream), "out");
      hip_check(hipStreamSynchronize(ctx->stream), "tonemap");
    } catch (...) {
      for (void* q : own) (void)hipFree(q);
      throw;
    }
    for (void* q : own) (void)hipFree(q);
  })
}

// values per item of rtg_medium_batch's input and output, by op
static const int kMediumIn[5] = {3, 8, 8, 6, 3}, kMediumOut[5] = {1, 3, 3, 1, 4};

int rtg_medium_batch(rtg_ctx* ctx, rtg_scene* scene, int32_t op, const float* in, int32_t n, const float* u,
                     int32_t k, float* out) {
  if (!ctx) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  RTG_GUARD(ctx, {
    if (!scene || !in || !out || n < 0 || op < 0 || op > 4) fail(RTG_ERR_INVALID, "bad medium arguments");
    if (op <= 2 && scene->dev.medium_kind != 1)
      fail(RTG_ERR_INVALID, "density and tracking queries need a heterogeneous (grid) medium");
    if (scene->dev.medium_kind == 0) fail(RTG_ERR_INVALID, "the scene has no medium");
    if ((op == 1 || op == 2) && (k < 0 || (k > 0 && !u))) fail(RTG_ERR_INVALID, "tracking needs k draws per ray");
    if (op == 4 && (!u || k != 2)) fail(RTG_ERR_INVALID, "phase sampling needs 2 draws per item");
    if (n == 0) return RTG_OK;
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const size_t nin = (size_t)n * kMediumIn[op];
    const size_t nout = (size_t)n * kMediumOut[op];
    const size_t nu = (op == 1 || op == 2 || op == 4) ? (size_t)n * k : 0;
    std::vector<void*> own;
    float* din = (float*)dev_alloc(ctx, nin * 4, own);
    float* du = (float*)dev_alloc(ctx, std::max<size_t>(nu, 1) * 4, own);
    float* dout = (float*)dev_alloc(ctx, nout * 4, own);
    try {
      hip_check(hipMemcpyAsync(din, in, nin * 4, hipMemcpyHostToDevice, ctx->stream), "in");
      if (nu) hip_check(hipMemcpyAsync(du, u, nu * 4, hipMemcpyHostToDevice, ctx->stream), "u");
      launch_medium_tap(ctx->scenes_dev + scene->slot, op, din, n, du, nu ? k : 0, dout, ctx->stream);
      hip_check(hipGetLastError(), "medium launch");
      hip_check(hipMemcpyAsync(out, dout, nout * 4, hipMemcpyDeviceToHost, ctx->stream), "out");
      hip_check(hipStreamSynchronize(ctx->stream), "medium");
    } catch (...) {
      for (void* q : own) (void)hipFree(q);
      throw;
    }
    for (void* q : own) (void)hipFree(q);
  })
}

// floats per item of the light tap's input and output, by op
static const int kLightIn[4] = {3, 12, 3, 1};
static const int kLightOut[4] = {14, 1, 3, 3};
int rtg_light_batch(rtg_ctx* ctx, rtg_scene* scene, int32_t op, int32_t light, const float* in, int32_t n,
                    const float* u, float* out) {
  if (!ctx) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  RTG_GUARD(ctx, {
    if (!scene || !in || !out || n < 0 || op < 0 || op > 3) fail(RTG_ERR_INVALID, "bad light arguments");
    if ((op == 0 || op == 1) && (light < 0 || light >= scene->dev.nlights)) fail(RTG_ERR_INVALID, "no such light");
    if (op == 0 && !u) fail(RTG_ERR_INVALID, "light sampling needs 2 draws per point");
    if (op == 2 && scene->dev.env_w <= 0) fail(RTG_ERR_INVALID, "the scene has no environment map");
    if (op == 3 && scene->dev.nlights <= 0) fail(RTG_ERR_INVALID, "the scene has no lights");
    if (n == 0) return RTG_OK;
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const size_t nin = (size_t)n * kLightIn[op];
    const size_t nout = (size_t)n * kLightOut[op];
    const size_t nu = op == 0 ? (size_t)n * 2 : 0;
    std::vector<void*> own;
    float* din = (float*)dev_alloc(ctx, nin * 4, own);
    float* du = (float*)dev_alloc(ctx, std::max<size_t>(nu, 1) * 4, own);
    float* dout = (float*)dev_alloc(ctx, nout * 4, own);
    try {
      hip_check(hipMemcpyAsync(din, in, nin * 4, hipMemcpyHostToDevice, ctx->stream), "in");
      if (nu) hip_check(hipMemcpyAsync(du, u, nu * 4, hipMemcpyHostToDevice, ctx->stream), "u");
      launch_light_tap(ctx->scenes_dev + scene->slot, op, light, din, du, n, dout, ctx->stream);
      hip_check(hipGetLastError(), "light launch");
      hip_check(hipMemcpyAsync(out, dout, nout * 4, hipMemcpyDeviceToHost, ctx->stream), "out");
      hip_check(hipStreamSynchronize(ctx->stream), "light");
    } catch (...) {
      for (void* q : own) (void)hipFree(q);
      throw;
    }
    for (void* q : own) (void)hipFree(q);
  })
}

int rtg_camera_rays(rtg_ctx* ctx, const rtg_render_desc* desc, const float* pxy, int32_t n, float* rays) {
  if (!ctx) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  RTG_GUARD(ctx, {
    if (!desc || !pxy || !rays || n < 0) fail(RTG_ERR_INVALID, "bad camera arguments");
    const rtg_render_desc& d = *desc;
    if (d.width <= 0 || d.height <= 0) fail(RTG_ERR_INVALID, "width and height must be positive");
    if (!(d.fov > 0.f && d.fov < 180.f)) fail(RTG_ERR_INVALID, "fov out of range");
    if (n == 0) return RTG_OK;
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    // the camera fields of a render's job, as make_job fills them
    DevJob J;
    std::memset(&J, 0, sizeof(J));
    camera_matrices(d, J.s2c, J.c2w);
    J.inv_w = 1.f / d.width;
    J.inv_h = 1.f / d.height;
    J.near_clip = d.near_clip;
    J.far_clip = d.far_clip;
    J.width = d.width;
    J.height = d.height;
    std::vector<void*> own;
    float* dp = (float*)dev_alloc(ctx, (size_t)n * 2 * 4, own);
    float* dr = (float*)dev_alloc(ctx, (size_t)n * 8 * 4, own);
    try {
      hip_check(hipMemcpyAsync(dp, pxy, (size_t)n * 2 * 4, hipMemcpyHostToDevice, ctx->stream), "pxy");
      launch_camera_tap(J, dp, n, dr, ctx->stream);
      hip_check(hipGetLastError(), "camera launch");
      hip_check(hipMemcpyAsync(rays, dr, (size_t)n * 8 * 4, hipMemcpyDeviceToHost, ctx->stream), "rays");
      hip_check(hipStreamSynchronize(ctx->stream), "camera");
    } catch (...) {
      for (void* q : own) (void)hipFree(q);
      throw;
    }
    for (void* q : own) (void)hipFree(q);
  })
}

int rtg_bsdf_batch(rtg_ctx* ctx, const rtg_material* material, int32_t op, const float* wi, const float* wo_or_u,
                   int32_t n, float* out_dir, float* f, float* pdf, float* rel_eta) {
  if (!ctx) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  RTG_GUARD(ctx, {
    if (!material || !wi || !wo_or_u || !f || !pdf || n < 0 || op < 0 || op > 3)
      fail(RTG_ERR_INVALID, "bad bsdf arguments");
    if ((op == 1 || op == 3) && (!out_dir || !rel_eta)) fail(RTG_ERR_INVALID, "sample needs out_dir and rel_eta");
    if (op == 3 && material->type != RTG_MAT_ROUGH_DIELECTRIC)
      fail(RTG_ERR_INVALID, "exit proposals need a rough dielectric");
    if (op == 2 && (!rel_eta || (material->type != RTG_MAT_DIFFUSE && material->type != RTG_MAT_ROUGH_CONDUCTOR)))
      fail(RTG_ERR_INVALID, "sampling density needs rel_eta and a diffuse or conductor material");
    if (n == 0) return RTG_OK;
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const DevMaterial d = dev_material(*material);
    std::vector<void*> own;
    DevMaterial* dm = (DevMaterial*)dev_alloc(ctx, sizeof(DevMaterial), own);
    float* dwi = (float*)dev_alloc(ctx, (size_t)n * 12, own);
    size_t w2 = (size_t)n * (op == 1 ? 8 : 12);  // wo or (ux, uy); op 3: (technique, ux, uy)
    float* dwo = (float*)dev_alloc(ctx, w2, own);
    float* ddir = (float*)dev_alloc(ctx, (size_t)n * 12, own);
    float* df = (float*)dev_alloc(ctx, (size_t)n * 12, own);
    float* dp = (float*)dev_alloc(ctx, (size_t)n * 4, own);
    float* de = (float*)dev_alloc(ctx, (size_t)n * 4, own);
    try {
      hip_check(hipMemcpyAsync(dm, &d, sizeof(d), hipMemcpyHostToDevice, ctx->stream), "mat");
      hip_check(hipMemcpyAsync(dwi, wi, (size_t)n * 12, hipMemcpyHostToDevice, ctx->stream), "wi");
      hip_check(hipMemcpyAsync(dwo, wo_or_u, w2, hipMemcpyHostToDevice, ctx->stream), "wo");
      launch_bsdf_tap(dm, op, dwi, dwo, n, ddir, df, dp, de, ctx->stream);
      hip_check(hipGetLastError(), "bsdf launch");
      hip_check(hipMemcpyAsync(f, df, (size_t)n * 12, hipMemcpyDeviceToHost, ctx->stream), "f");
      hip_check(hipMemcpyAsync(pdf, dp, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream), "pdf");
      if (op == 1 || op == 3)
        hip_check(hipMemcpyAsync(out_dir, ddir, (size_t)n * 12, hipMemcpyDeviceToHost, ctx->stream), "dir");
      if (op >= 1)
        hip_check(hipMemcpyAsync(rel_eta, de, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream), "eta");
      hip_check(hipStreamSynchronize(ctx->stream), "bsdf");
    } catch (...) {
      for (void* q : own) (void)hipFree(q);
      throw;
    }
    for (void* q : own) (void)hipFree(q);
  })
}

}  // extern "C"
