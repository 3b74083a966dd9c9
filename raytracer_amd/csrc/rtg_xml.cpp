// Scene files in the reference's XML schema (core/XMLLoader.cpp:33-904), for the types the GPU
// path renders: PathTracingMIS, RandomSampler / StratifiedSampler, NoriPerspective with the box
// filter, Mesh (OBJ, io/ObjLoader.cpp:97-289) / Sphere / Cube / Plane shapes, Diffuse /
// RoughConductor / RoughDielectric / Conductor / Dielectric BSDFs with constant parameters, area
// lights on meshes, and an environment light from a Radiance RGBE (.hdr) image. Anything else fails with RTG_ERR_PARSE and names the element.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "rtg_cie.h"
#include "rtg_host.h"

using namespace rtg;

namespace {

[[noreturn]] void perr(const std::string& m) { fail(RTG_ERR_PARSE, m); }

// Color3::setFromSampledSpectrum (core/Color.cpp:423-443): the spectrum, piecewise linear between
// its samples and constant beyond its ends (Color.h:239-248), is read at each nanometre of the CIE
// table and summed against x, y, z; the sums are scaled by (830 - 360) / (kYIntegral * 471) and
// turned into linear sRGB by the matrix of Color.h:196-205, negative channels clamped to zero.
// All in float, in the reference's order of operations.
void spectrum_to_rgb(const float* lam, const float* val, int n, float rgb[3]) {
  float xyz[3] = {0.f, 0.f, 0.f};
  for (int i = 0; i < rtg_cie::kSamples; i++) {
    const float l = (float)(rtg_cie::kFirstNm + i);
    float v;
    if (l <= lam[0]) {
      v = val[0];
    } else if (l >= lam[n - 1]) {
      v = val[n - 1];
    } else {
      // bisection for the first index whose wavelength exceeds l; the segment starts one before it,
      // clamped to [0, n - 2] (Color.h:219-234: also defined on unsorted input)
      int first = 0, len = n;
      while (len > 0) {
        const int half = len >> 1, mid = first + half;
        if (lam[mid] <= l) {
          first = mid + 1;
          len -= half + 1;
        } else {
          len = half;
        }
      }
      const int o = std::min(std::max(first - 1, 0), n - 2);
      const float t = (l - lam[o]) / (lam[o + 1] - lam[o]);
      v = (1.f - t) * val[o] + t * val[o + 1];
    }
    xyz[0] += v * rtg_cie::kCmf[i][0];
    xyz[1] += v * rtg_cie::kCmf[i][1];
    xyz[2] += v * rtg_cie::kCmf[i][2];
  }
  const float scale = (float)(rtg_cie::kFirstNm + rtg_cie::kSamples - 1 - rtg_cie::kFirstNm) /
                      (rtg_cie::kYIntegral * rtg_cie::kSamples);
  for (float& c : xyz) c *= scale;
  rgb[0] = 3.240479f * xyz[0] - 1.537150f * xyz[1] - 0.498535f * xyz[2];
  rgb[1] = -0.969256f * xyz[0] + 1.875991f * xyz[1] + 0.041556f * xyz[2];
  rgb[2] = 0.055648f * xyz[0] - 0.204043f * xyz[1] + 1.057311f * xyz[2];
  for (int c = 0; c < 3; c++) rgb[c] = std::max(rgb[c], 0.f);
}

// IORHelper::extractAndConvertSpectrumFromFile (tools/Common.cpp:83-110): "lambda value" lines,
// '#' lines skipped. A line of two words that are not numbers (Fe.k.spd's header) enters as the
// pair (0, 0), as failed stream extraction leaves it in the reference; a line of fewer than two
// words (blank) is skipped, where the reference reads an undefined pair.
void read_spd(const std::string& path, float rgb[3]) {
  std::ifstream f(path);
  if (!f) perr("spectrum file " + path + " does not exist");
  std::vector<float> lam, val;
  std::string line;
  while (std::getline(f, line)) {
    if (!line.empty() && line[0] == '#') continue;
    std::istringstream ls(line);
    std::string a, b;
    if (!(ls >> a >> b)) continue;
    auto num = [](const std::string& w) {
      std::istringstream is(w);
      float x = 0.f;
      if (!(is >> x)) x = 0.f;
      return x;
    };
    lam.push_back(num(a));
    val.push_back(num(b));
  }
  if (lam.size() < 2) perr("spectrum file " + path + ": fewer than two samples");
  spectrum_to_rgb(lam.data(), val.data(), (int)lam.size(), rgb);
}

// object types are looked up case-insensitively (factory/ObjectFactoryManager.h:82-89)
std::string lower(const std::string& s) {
  std::string r = s;
  for (char& c : r) c = (char)std::tolower((unsigned char)c);
  return r;
}

// ---------------------------------------------------------------- minimal XML
struct XNode {
  std::string name;
  std::map<std::string, std::string> attr;
  std::string text;
  std::vector<XNode> kids;
  const std::string* get(const char* k) const {
    auto it = attr.find(k);
    return it == attr.end() ? nullptr : &it->second;
  }
};

class XmlParser {
 public:
  explicit XmlParser(const std::string& s) : s_(s) {}
  XNode parse() {
    if (s_.compare(0, 3, "\xEF\xBB\xBF") == 0) p_ = 3;  // UTF-8 byte-order mark
    skip_misc();
    if (p_ >= s_.size() || s_[p_] != '<') perr("xml: no root element");
    XNode root = element();
    return root;
  }

 private:
  const std::string& s_;
  size_t p_ = 0;
  void ws() {
    while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) p_++;
  }
  bool starts(const char* t) const { return s_.compare(p_, std::strlen(t), t) == 0; }
  void skip_until(const char* t) {
    size_t q = s_.find(t, p_);
    if (q == std::string::npos) perr(std::string("xml: unterminated construct, missing ") + t);
    p_ = q + std::strlen(t);
  }
  // declarations, comments, processing instructions and whitespace between elements
  void skip_misc() {
    while (true) {
      ws();
      if (starts("<?")) skip_until("?>");
      else if (starts("<!--")) skip_until("-->");
      else if (starts("<!")) skip_until(">");
      else return;
    }
  }
  static std::string unescape(const std::string& v) {
    std::string o;
    for (size_t i = 0; i < v.size(); i++) {
      if (v[i] == '&') {
        static const char* ents[][2] = {{"&lt;", "<"}, {"&gt;", ">"}, {"&amp;", "&"}, {"&quot;", "\""}, {"&apos;", "'"}};
        bool done = false;
        for (auto& e : ents)
          if (v.compare(i, std::strlen(e[0]), e[0]) == 0) {
            o += e[1];
            i += std::strlen(e[0]) - 1;
            done = true;
            break;
          }
        if (!done) o += v[i];
      } else {
        o += v[i];
      }
    }
    return o;
  }
  std::string name() {
    size_t b = p_;
    while (p_ < s_.size() && (std::isalnum((unsigned char)s_[p_]) || s_[p_] == '_' || s_[p_] == '-' ||
                              s_[p_] == ':' || s_[p_] == '.'))
      p_++;
    if (b == p_) perr("xml: expected a name at offset " + std::to_string(p_));
    return s_.substr(b, p_ - b);
  }
  XNode element() {
    XNode n;
    p_++;  // '<'
    n.name = name();
    while (true) {
      ws();
      if (p_ >= s_.size()) perr("xml: unexpected end in <" + n.name + ">");
      if (starts("/>")) {
        p_ += 2;
        return n;
      }
      if (s_[p_] == '>') {
        p_++;
        break;
      }
      std::string k = name();
      ws();
      if (p_ >= s_.size() || s_[p_] != '=') perr("xml: expected '=' after attribute " + k);
      p_++;
      ws();
      char q = p_ < s_.size() ? s_[p_] : 0;
      if (q != '"' && q != '\'') perr("xml: unquoted attribute " + k);
      size_t e = s_.find(q, p_ + 1);
      if (e == std::string::npos) perr("xml: unterminated attribute " + k);
      n.attr[k] = unescape(s_.substr(p_ + 1, e - p_ - 1));
      p_ = e + 1;
    }
    // content
    while (true) {
      if (p_ >= s_.size()) perr("xml: missing </" + n.name + ">");
      if (starts("<!--")) {
        skip_until("-->");
      } else if (starts("<![CDATA[")) {
        size_t e = s_.find("]]>", p_);
        if (e == std::string::npos) perr("xml: unterminated CDATA");
        n.text += s_.substr(p_ + 9, e - p_ - 9);
        p_ = e + 3;
      } else if (starts("</")) {
        p_ += 2;
        std::string c = name();
        if (c != n.name) perr("xml: </" + c + "> closes <" + n.name + ">");
        ws();
        if (p_ >= s_.size() || s_[p_] != '>') perr("xml: bad closing tag " + c);
        p_++;
        return n;
      } else if (starts("<?")) {
        skip_until("?>");
      } else if (s_[p_] == '<') {
        n.kids.push_back(element());
      } else {
        size_t e = s_.find('<', p_);
        if (e == std::string::npos) e = s_.size();
        n.text += unescape(s_.substr(p_, e - p_));
        p_ = e;
      }
    }
  }
};

std::string type_of(const XNode& n) {
  const std::string* t = n.get("type");
  return t ? lower(*t) : std::string();
}

// ---------------------------------------------------------------- values
std::vector<double> numbers(const std::string& s) {
  std::vector<double> v;
  std::string t;
  for (char c : s) t += (c == ',') ? ' ' : c;
  std::istringstream is(t);
  std::string tok;
  while (is >> tok) {
    char* end = nullptr;
    double x = std::strtod(tok.c_str(), &end);
    if (!end || *end) perr("bad number '" + tok + "'");
    v.push_back(x);
  }
  return v;
}
double number(const std::string& s) {
  auto v = numbers(s);
  if (v.size() != 1) perr("expected one number, got '" + s + "'");
  return v[0];
}

// row-major 4x4 in double
struct M4 {
  double m[16];
  static M4 id() {
    M4 r;
    for (int i = 0; i < 16; i++) r.m[i] = (i % 5 == 0) ? 1.0 : 0.0;
    return r;
  }
};
M4 mul(const M4& a, const M4& b) {
  M4 r;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      double s = 0;
      for (int k = 0; k < 4; k++) s += a.m[4 * i + k] * b.m[4 * k + j];
      r.m[4 * i + j] = s;
    }
  return r;
}
bool invert(const M4& a, M4& out) {
  double m[4][8];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 8; j++) m[i][j] = j < 4 ? a.m[4 * i + j] : (j - 4 == i ? 1.0 : 0.0);
  for (int c = 0; c < 4; c++) {
    int p = c;
    for (int r = c + 1; r < 4; r++)
      if (std::fabs(m[r][c]) > std::fabs(m[p][c])) p = r;
    if (std::fabs(m[p][c]) < 1e-300) return false;
    for (int j = 0; j < 8; j++) std::swap(m[c][j], m[p][j]);
    double d = m[c][c];
    for (int j = 0; j < 8; j++) m[c][j] /= d;
    for (int r = 0; r < 4; r++)
      if (r != c) {
        double f = m[r][c];
        for (int j = 0; j < 8; j++) m[r][j] -= f * m[c][j];
      }
  }
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) out.m[4 * i + j] = m[i][j + 4];
  return true;
}

// xyz children (<x>, <y>, <z>) with a default value
void xyz(const XNode& n, double def, double v[3]) {
  v[0] = v[1] = v[2] = def;
  for (const XNode& c : n.kids) {
    if (c.name == "x") v[0] = number(c.text);
    else if (c.name == "y") v[1] = number(c.text);
    else if (c.name == "z") v[2] = number(c.text);
    else perr("unknown transform component <" + c.name + ">");
  }
}
void vec3(const std::string& s, double v[3]) {
  auto a = numbers(s);
  if (a.size() != 3) perr("expected three numbers, got '" + s + "'");
  v[0] = a[0]; v[1] = a[1]; v[2] = a[2];
}

// core/Transform.cpp:116-227 (angles in degrees; lookAt columns left, up, dir, origin)
M4 rot_axis(int axis, double deg) {
  const double a = deg * 3.14159265358979323846 / 180.0, c = std::cos(a), s = std::sin(a);
  M4 r = M4::id();
  if (axis == 0) { r.m[5] = c; r.m[6] = -s; r.m[9] = s; r.m[10] = c; }
  if (axis == 1) { r.m[0] = c; r.m[2] = s; r.m[8] = -s; r.m[10] = c; }
  if (axis == 2) { r.m[0] = c; r.m[1] = -s; r.m[4] = s; r.m[5] = c; }
  return r;
}
void normalize3(double v[3]) {
  double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (l > 0) { v[0] /= l; v[1] /= l; v[2] /= l; }
}
void cross3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
M4 look_at(const double o[3], const double t[3], const double up_in[3]) {
  double dir[3] = {t[0] - o[0], t[1] - o[1], t[2] - o[2]}, up[3] = {up_in[0], up_in[1], up_in[2]}, left[3], nu[3];
  normalize3(dir);
  normalize3(up);
  cross3(up, dir, left);
  normalize3(left);
  cross3(dir, left, nu);
  normalize3(nu);
  M4 r = M4::id();
  for (int i = 0; i < 3; i++) {
    r.m[4 * i + 0] = left[i];
    r.m[4 * i + 1] = nu[i];
    r.m[4 * i + 2] = dir[i];
    r.m[4 * i + 3] = o[i];
  }
  return r;
}

// XMLLoader::handleTransform (core/XMLLoader.cpp:843-904): each child multiplies from the left
M4 parse_transform(const XNode& n) {
  M4 T = M4::id();
  for (const XNode& c : n.kids) {
    M4 A = M4::id();
    if (c.name == "rotation") {
      double r[3];
      xyz(c, 0.0, r);
      A = mul(mul(rot_axis(0, r[0]), rot_axis(1, r[1])), rot_axis(2, r[2]));
    } else if (c.name == "translation") {
      double t[3];
      xyz(c, 0.0, t);
      A.m[3] = t[0]; A.m[7] = t[1]; A.m[11] = t[2];
    } else if (c.name == "scale") {
      double s[3];
      xyz(c, 1.0, s);
      A.m[0] = s[0]; A.m[5] = s[1]; A.m[10] = s[2];
    } else if (c.name == "matrix") {
      const std::string* v = c.get("value");
      if (!v) perr("<matrix> without value");
      auto a = numbers(*v);
      if (a.size() != 16) perr("<matrix> needs 16 values");
      for (int i = 0; i < 16; i++) A.m[i] = a[i];
    } else if (c.name == "lookAt") {
      double o[3] = {0, 0, 0}, t[3] = {0, 0, 1}, u[3] = {0, 1, 0};
      for (const XNode& k : c.kids) {
        if (k.name == "origin") vec3(k.text, o);
        else if (k.name == "lookAt") vec3(k.text, t);
        else if (k.name == "up") vec3(k.text, u);
        else perr("unknown lookAt component <" + k.name + ">");
      }
      A = look_at(o, t, u);
    } else {
      perr("unknown transform element <" + c.name + ">");
    }
    T = mul(A, T);
  }
  return T;
}

// Property list of an object (XMLLoader::handleProperty, core/XMLLoader.cpp:687-813)
struct Props {
  std::map<std::string, std::string> val;   // name -> text value
  std::map<std::string, std::string> kind;  // name -> element name (integer, color, ...)
  std::map<std::string, M4> xf;
  std::map<std::string, int> tex;           // name -> image texture (0-based) in place of a value
  std::string owner;
  bool boolean(const std::string& k, bool d) const {
    auto it = val.find(k);
    if (it == val.end()) return d;
    const std::string v = lower(it->second);
    if (v == "true" || v == "1") return true;
    if (v == "false" || v == "0") return false;
    perr(owner + ": '" + k + "' must be true or false");
  }
  bool has(const std::string& k) const { return val.count(k) || xf.count(k); }
  std::string str(const std::string& k, const std::string& d) const {
    auto it = val.find(k);
    return it == val.end() ? d : it->second;
  }
  double real(const std::string& k, double d) const {
    auto it = val.find(k);
    return it == val.end() ? d : number(it->second);
  }
  int integer(const std::string& k, int d) const { return (int)std::lround(real(k, d)); }
  void color(const std::string& k, const float d[3], float out[3]) const {
    auto it = val.find(k);
    if (it == val.end()) {
      out[0] = d[0]; out[1] = d[1]; out[2] = d[2];
      return;
    }
    if (kind.at(k) == "spectrum") perr(owner + ": spectral '" + k + "' values are not supported (give an RGB color)");
    auto v = numbers(it->second);
    if (v.size() == 1) v = {v[0], v[0], v[0]};
    if (v.size() != 3) perr(owner + ": color '" + k + "' needs 3 values");
    out[0] = (float)v[0]; out[1] = (float)v[1]; out[2] = (float)v[2];
  }
  M4 transform(const std::string& k) const {
    auto it = xf.find(k);
    return it == xf.end() ? M4::id() : it->second;
  }
};

struct Loader {
  std::string dir;  // directory of the scene file (relative paths)
  std::map<std::string, XNode> ref_bsdf;
  std::map<std::string, M4> ref_xf;
  // output
  std::vector<rtg_material> mats;
  struct Mesh {
    std::vector<float> v, n, uv;
    std::vector<int32_t> idx;
    int material = 0;
    bool emit = false;
    float emission[3] = {0, 0, 0};
  };
  std::vector<Mesh> meshes;
  std::vector<rtg_sphere_desc> spheres;
  std::vector<float> env;
  int env_w = 0, env_h = 0;
  float env_scale = 1.f;
  rtg_render_desc rd{};
  int strategy = RTG_LIGHTS_UNIFORM;
  bool have_camera = false, have_integrator = false;
  // image textures (their texels in texdata; rgb pointers are set when the scene is built)
  std::vector<rtg_texture_desc> textures;
  std::vector<std::vector<float>> texdata;
  // camera medium
  std::vector<float> grid;
  int gres[3] = {0, 0, 0};
  float medium_albedo[3] = {0.f, 0.f, 0.f};
  float medium_g = 0.f, medium_scale = 1.f;
  bool medium_homogeneous = false;
  float medium_sigma_a[3] = {0.f, 0.f, 0.f}, medium_sigma_s[3] = {0.f, 0.f, 0.f};

  // IORHelper::evalConductorIOR (tools/Common.cpp:44-81): a known metal name reads
  // data/IOR/<symbol>.eta.spd and .k.spd (searched beside the scene file, then in the working
  // directory) through the spectral conversion; no name, or an unknown one, takes the 'eta' and
  // 'absorption' colors.
  void conductor_ior(Props& P, rtg_material& m) {
    static const float zero[3] = {0.f, 0.f, 0.f};
    static const std::map<std::string, std::string> symbol = {
        {"gold", "Au"},      {"silver", "Ag"}, {"aluminium", "Al"},   {"copper", "Cu"},
        {"chromium", "Cr"},  {"iron", "Fe"},   {"nickel", "Ni_palik"}, {"tin", "SnTe_palik"}};
    const std::string name = P.str("material", "");
    const auto it = symbol.find(name);
    if (it == symbol.end()) {
      if (!name.empty())
        std::fprintf(stderr, "rtg: material %s not found, using the eta and absorption colors\n", name.c_str());
      P.color("eta", zero, m.eta);
      P.color("absorption", zero, m.k);
      return;
    }
    read_spd(resolve("./data/IOR/" + it->second + ".eta.spd"), m.eta);
    read_spd(resolve("./data/IOR/" + it->second + ".k.spd"), m.k);
  }

  std::string resolve(const std::string& p) const {
    if (p.empty() || p[0] == '/') return p;
    std::string q = p;
    if (q.compare(0, 2, "./") == 0) q = q.substr(2);
    std::string a = dir + "/" + q;
    std::ifstream f(a);
    if (f.good()) return a;
    // the reference also searches the program's working directory
    return p;
  }

  Props props(const XNode& n, const std::string& owner) {
    Props P;
    P.owner = owner;
    for (const XNode& c : n.kids) {
      if (c.name == "bsdf" || c.name == "light" || c.name == "filter" || c.name == "medium" || c.name == "bssrdf" ||
          c.name == "volume" || c.name == "phase")
        continue;  // handled by the caller
      const std::string* nm = c.get("name");
      if (c.name == "texture") {
        // XMLLoader::handleTexture (core/XMLLoader.cpp:556-571): a ConstantTexture is kept as its
        // color, an ImageTexture (or a ScaleTexture around one) becomes an image texture
        if (!nm) perr(owner + ": <texture> without a name");
        float col[3];
        int ti = texture(c, owner + " texture " + *nm, 1.f, col);
        if (ti >= 0) {
          P.tex[*nm] = ti;
        } else {
          char buf[128];
          std::snprintf(buf, sizeof(buf), "%.9g %.9g %.9g", col[0], col[1], col[2]);
          P.val[*nm] = buf;
          P.kind[*nm] = "color";
        }
        continue;
      }
      if (c.name == "transform") {
        if (const std::string* r = c.get("ref")) {
          auto it = ref_xf.find(*r);
          if (it == ref_xf.end()) perr(owner + ": unknown transform ref " + *r);
          P.xf[nm ? *nm : "toWorld"] = it->second;
        } else {
          P.xf[nm ? *nm : "toWorld"] = parse_transform(c);
        }
        continue;
      }
      if (!nm) perr(owner + ": property <" + c.name + "> without a name");
      const std::string* v = c.get("value");
      if (!v) perr(owner + ": property " + *nm + " without a value");
      static const char* kinds[] = {"bool", "integer", "double", "real", "string", "point", "color", "spectrum",
                                    "vector"};
      bool ok = false;
      for (const char* k : kinds) ok = ok || c.name == k;
      if (!ok) perr(owner + ": unsupported property <" + c.name + "> (" + *nm + ")");
      P.val[*nm] = *v;
      P.kind[*nm] = c.name;
    }
    return P;
  }

  // A texture element, times `scale`: returns its image-texture index, or -1 for a constant
  // (then `col`). ConstantTexture "albedo" (texture/ConstantTexture.cpp:6-10); ScaleTexture
  // "scale" x "texture" (texture/ScaleTexture.cpp:4-22); ImageTexture "path", "invertX",
  // "invertY", "scale", "gamma", "channel" (texture/ImageTexture.cpp:7-20) over a Radiance .hdr
  // image, which the reference decodes with stb_image at gamma 1 (io/ImageLoader.cpp:40-100,
  // 183-215: "channel" r, g, b or a replicates one channel; an RGBE file has no alpha, so a is 0).
  int texture(const XNode& c, const std::string& owner, float scale, float col[3]) {
    const std::string* ty = c.get("type");
    const std::string t = ty ? lower(*ty) : std::string("?");
    if (t == "scaletexture") {
      // its own properties are read here: props() would also load the inner texture
      const XNode* inner = nullptr;
      double s = 1.0;
      for (const XNode& k : c.kids) {
        const std::string* nm = k.get("name");
        if (k.name == "texture" && nm && *nm == "texture") inner = &k;
        if (nm && *nm == "scale" && k.get("value")) s = number(*k.get("value"));
      }
      if (!inner) perr(owner + ": ScaleTexture without a texture named 'texture'");
      return texture(*inner, owner + " texture", scale * (float)s, col);
    }
    Props T = props(c, owner);
    if (t == "constanttexture") {
      static const float half[3] = {0.5f, 0.5f, 0.5f};
      T.color("albedo", half, col);
      for (int k = 0; k < 3; k++) col[k] *= scale;
      return -1;
    }
    if (t != "imagetexture")
      perr(owner + ": texture type " + (ty ? *ty : std::string("?")) +
           " is not supported (ConstantTexture, ImageTexture, ScaleTexture)");
    const std::string path = resolve(T.str("path", ""));
    const double gamma = T.real("gamma", 0.0);
    if (gamma != 0.0 && gamma != 1.0) perr(owner + ": only gamma 0 or 1 is supported for .hdr images");
    const std::string ch = lower(T.str("channel", "all"));
    if (ch != "all" && ch != "r" && ch != "g" && ch != "b" && ch != "a")
      perr(owner + ": channel must be all, r, g, b or a");
    std::vector<float> px;
    int W = 0, H = 0;
    read_hdr(path, px, W, H);
    if (W < 2 || H < 2) perr(owner + ": image must be at least 2x2");
    if (ch != "all") {
      const int k = ch == "r" ? 0 : ch == "g" ? 1 : ch == "b" ? 2 : -1;
      for (size_t i = 0; i < px.size(); i += 3) {
        const float v = k < 0 ? 0.f : px[i + k];
        px[i] = px[i + 1] = px[i + 2] = v;
      }
    }
    rtg_texture_desc d;
    std::memset(&d, 0, sizeof(d));
    d.width = W;
    d.height = H;
    d.uv_scale = (float)std::fabs(T.real("scale", 1.0));
    d.invert_x = T.boolean("invertX", false) ? 1 : 0;
    d.invert_y = T.boolean("invertY", false) ? 1 : 0;
    d.value_scale = scale;
    textures.push_back(d);
    texdata.push_back(std::move(px));
    return (int)textures.size() - 1;
  }

  int bsdf(const XNode& n) {
    const XNode* b = &n;
    if (const std::string* r = n.get("ref")) {
      auto it = ref_bsdf.find(*r);
      if (it == ref_bsdf.end()) perr("unknown bsdf ref " + *r);
      b = &it->second;
    }
    const std::string* type = b->get("type");
    if (!type) perr("<bsdf> without a type");
    const std::string lt = lower(*type);
    Props P = props(*b, "bsdf " + *type);
    rtg_material m;
    std::memset(&m, 0, sizeof(m));
    static const float one[3] = {1, 1, 1}, half[3] = {0.5f, 0.5f, 0.5f}, zero[3] = {0, 0, 0};
    auto dist = [&](const std::string& d) {
      if (d == "beckmann") return (int)RTG_DIST_BECKMANN;
      if (d == "ggx") return (int)RTG_DIST_GGX;
      perr("bsdf " + *type + ": distribution '" + d + "' is not supported (beckmann, ggx)");
    };
    if (lt == "diffuse") {
      m.type = RTG_MAT_DIFFUSE;
      P.color("albedo", half, m.albedo);  // bsdf/Diffuse.cpp:15-17
    } else if (lt == "roughconductor") {
      m.type = RTG_MAT_ROUGH_CONDUCTOR;  // bsdf/RoughConductor.cpp:13-30
      m.alpha = (float)P.real("alpha", 0.2);
      m.distribution = dist(P.str("distribution", "beckmann"));
      conductor_ior(P, m);
      P.color("reflectanceTexture", one, m.albedo);
    } else if (lt == "conductor") {
      m.type = RTG_MAT_CONDUCTOR;  // bsdf/Conductor.cpp:12-19
      conductor_ior(P, m);
      P.color("reflectanceTexture", one, m.albedo);
    } else if (lt == "dielectric") {
      m.type = RTG_MAT_DIELECTRIC;  // bsdf/Dielectric.cpp:10-17
      m.eta_ext = (float)P.real("etaExt", 1.000277);
      m.eta_int = (float)P.real("etaInt", 1.5046);
      P.color("reflectanceTexture", one, m.albedo);
      P.color("transmittedTexture", one, m.transmitted);
    } else if (lt == "roughdielectric") {
      m.type = RTG_MAT_ROUGH_DIELECTRIC;  // bsdf/RoughDielectric.cpp:13-27
      m.alpha = (float)P.real("alpha", 0.2);
      m.eta_ext = (float)P.real("etaExt", 1.000277);
      m.eta_int = (float)P.real("etaInt", 1.5046);
      m.distribution = dist(P.str("distribution", "beckmann"));
      P.color("reflectanceTexture", one, m.albedo);
      P.color("transmittedTexture", one, m.transmitted);
    } else {
      perr("bsdf type " + *type + " is not rendered by the GPU path");
    }
    // image textures in place of constants (1-based in rtg_material)
    auto slot = [&](const char* k) { return P.tex.count(k) ? P.tex.at(k) + 1 : 0; };
    m.tex_albedo = lt == "diffuse" ? slot("albedo") : slot("reflectanceTexture");
    m.tex_alpha = (lt == "roughconductor" || lt == "roughdielectric") ? slot("alpha") : 0;
    m.tex_transmitted = (lt == "dielectric" || lt == "roughdielectric") ? slot("transmittedTexture") : 0;
    mats.push_back(m);
    return (int)mats.size() - 1;
  }

  // io/ObjLoader.cpp:97-289: positions and normals transformed by toWorld (normals by the inverse
  // transpose), quads split into (0,1,2) and (3,0,2), vertices deduplicated by (position, uv,
  // normal) index triple, negative indices counted back from the end of the file's lists.
  void obj(const std::string& path, const M4& T, Mesh& out) {
    std::ifstream is(path);
    if (!is) perr("cannot open mesh file " + path);
    M4 inv;
    if (!invert(T, inv)) perr("singular toWorld for " + path);
    std::vector<double> P, N, UV;
    struct VK { int p, t, n; };
    struct VH {
      size_t operator()(const VK& k) const { return ((size_t)k.p * 73856093u) ^ ((size_t)k.t * 19349663u) ^ ((size_t)k.n * 83492791u); }
    };
    struct VE {
      bool operator()(const VK& a, const VK& b) const { return a.p == b.p && a.t == b.t && a.n == b.n; }
    };
    std::unordered_map<VK, int, VH, VE> map;
    std::vector<VK> verts;
    std::string line;
    auto parse_v = [&](const std::string& tok) {
      VK k{1, 1, 1};
      std::vector<std::string> parts;
      std::string cur;
      for (char c : tok) {
        if (c == '/') { parts.push_back(cur); cur.clear(); } else cur += c;
      }
      parts.push_back(cur);
      if (parts.empty() || parts.size() > 3 || parts[0].empty()) perr("bad face vertex '" + tok + "' in " + path);
      k.p = std::atoi(parts[0].c_str());
      if (parts.size() > 1 && !parts[1].empty()) k.t = std::atoi(parts[1].c_str());
      if (parts.size() > 2 && !parts[2].empty()) k.n = std::atoi(parts[2].c_str());
      return k;
    };
    while (std::getline(is, line)) {
      std::istringstream sl(line);
      std::string type;
      sl >> type;
      if (type == "v") {
        double x, y, z;
        if (!(sl >> x >> y >> z)) perr("bad vertex line '" + line + "' in " + path);
        const double* m = T.m;
        P.push_back(m[0] * x + m[1] * y + m[2] * z + m[3]);
        P.push_back(m[4] * x + m[5] * y + m[6] * z + m[7]);
        P.push_back(m[8] * x + m[9] * y + m[10] * z + m[11]);
      } else if (type == "vn") {
        double x, y, z;
        if (!(sl >> x >> y >> z)) perr("bad normal line '" + line + "' in " + path);
        const double* m = inv.m;  // transpose of the inverse's 3x3
        N.push_back(m[0] * x + m[4] * y + m[8] * z);
        N.push_back(m[1] * x + m[5] * y + m[9] * z);
        N.push_back(m[2] * x + m[6] * y + m[10] * z);
      } else if (type == "vt") {
        double u, v;
        if (!(sl >> u)) perr("bad uv line '" + line + "' in " + path);
        if (!(sl >> v)) v = 0.0;
        UV.push_back(u);
        UV.push_back(v);
      } else if (type == "f") {
        std::string t[5];
        sl >> t[0] >> t[1] >> t[2] >> t[3] >> t[4];
        if (t[2].empty()) perr("face with fewer than 3 vertices in " + path);
        // a polygon with more than four vertices keeps its first four, as in the reference, which
        // only logs it (io/ObjLoader.cpp:176-177)
        VK f[6] = {parse_v(t[0]), parse_v(t[1]), parse_v(t[2])};
        int nv = 3;
        if (!t[3].empty()) {
          f[3] = parse_v(t[3]);
          f[4] = f[0];
          f[5] = f[2];
          nv = 6;
        }
        for (int i = 0; i < nv; i++) {
          auto it = map.find(f[i]);
          if (it == map.end()) {
            map[f[i]] = (int)verts.size();
            out.idx.push_back((int)verts.size());
            verts.push_back(f[i]);
          } else {
            out.idx.push_back(it->second);
          }
        }
      }
    }
    const int np = (int)P.size() / 3, nn = (int)N.size() / 3, nt = (int)UV.size() / 2;
    auto fix = [](int i, int n) { return i < 0 ? n + i + 1 : i; };
    for (const VK& k : verts) {
      int p = fix(k.p, np) - 1;
      if (p < 0 || p >= np) perr("face references a missing position in " + path);
      for (int c = 0; c < 3; c++) out.v.push_back((float)P[3 * p + c]);
    }
    if (nn) {
      for (const VK& k : verts) {
        int q = fix(k.n, nn) - 1;
        if (q < 0 || q >= nn) perr("face references a missing normal in " + path);
        for (int c = 0; c < 3; c++) out.n.push_back((float)N[3 * q + c]);
      }
    }
    if (nt) {
      for (const VK& k : verts) {
        int q = fix(k.t, nt) - 1;
        if (q < 0 || q >= nt) perr("face references a missing uv in " + path);
        for (int c = 0; c < 2; c++) out.uv.push_back((float)UV[2 * q + c]);
      }
    }
  }

  // geometricShape/Cube.cpp: the [-1,1]^3 cube, 24 vertices with per-face normals and UVs, two
  // triangles (0,1,2), (3,0,2) per face; positions and normals transformed by toWorld
  void cube(const M4& T, Mesh& out) {
    static const int pos[24][3] = {{1, -1, -1}, {1, -1, 1}, {-1, -1, 1}, {-1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                   {-1, 1, 1}, {1, 1, 1}, {1, -1, -1}, {1, 1, -1}, {1, 1, 1}, {1, -1, 1},
                                   {1, -1, 1}, {1, 1, 1}, {-1, 1, 1}, {-1, -1, 1}, {-1, -1, 1}, {-1, 1, 1},
                                   {-1, 1, -1}, {-1, -1, -1}, {1, 1, -1}, {1, -1, -1}, {-1, -1, -1}, {-1, 1, -1}};
    static const int nrm[6][3] = {{0, -1, 0}, {0, 1, 0}, {1, 0, 0}, {0, 0, 1}, {-1, 0, 0}, {0, 0, -1}};
    static const int uv[4][2] = {{0, 1}, {1, 1}, {1, 0}, {0, 0}};
    M4 inv;
    if (!invert(T, inv)) perr("Cube: singular toWorld");
    const double* m = T.m;
    const double* q = inv.m;
    for (int i = 0; i < 24; i++) {
      const double x = pos[i][0], y = pos[i][1], z = pos[i][2];
      out.v.push_back((float)(m[0] * x + m[1] * y + m[2] * z + m[3]));
      out.v.push_back((float)(m[4] * x + m[5] * y + m[6] * z + m[7]));
      out.v.push_back((float)(m[8] * x + m[9] * y + m[10] * z + m[11]));
      const double a = nrm[i / 4][0], b = nrm[i / 4][1], c = nrm[i / 4][2];
      out.n.push_back((float)(q[0] * a + q[4] * b + q[8] * c));
      out.n.push_back((float)(q[1] * a + q[5] * b + q[9] * c));
      out.n.push_back((float)(q[2] * a + q[6] * b + q[10] * c));
      out.uv.push_back((float)uv[i % 4][0]);
      out.uv.push_back((float)uv[i % 4][1]);
    }
    for (int f = 0; f < 6; f++) {
      const int b = 4 * f;
      const int t[6] = {b, b + 1, b + 2, b + 3, b, b + 2};
      for (int k : t) out.idx.push_back(k);
    }
  }

  // geometricShape/Plane.cpp: the square [-1,1]^2 at z = 0 of object space, normal +z; rendered as
  // two triangles spanning the same square
  void plane(const M4& T, Mesh& out) {
    static const int pos[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
    M4 inv;
    if (!invert(T, inv)) perr("Plane: singular toWorld");
    const double* m = T.m;
    const double* q = inv.m;
    for (int i = 0; i < 4; i++) {
      const double x = pos[i][0], y = pos[i][1];
      out.v.push_back((float)(m[0] * x + m[1] * y + m[3]));
      out.v.push_back((float)(m[4] * x + m[5] * y + m[7]));
      out.v.push_back((float)(m[8] * x + m[9] * y + m[11]));
      out.n.push_back((float)q[8]);
      out.n.push_back((float)q[9]);
      out.n.push_back((float)q[10]);
      out.uv.push_back(0.5f * (float)(x + 1));
      out.uv.push_back(0.5f * (float)(y + 1));
    }
    const int t[6] = {0, 1, 2, 0, 2, 3};
    for (int k : t) out.idx.push_back(k);
  }

  void object(const XNode& n) {
    const std::string* type = n.get("type");
    if (!type) perr("<" + n.name + "> without a type");
    const std::string lt = lower(*type);
    int mat = -1;
    const XNode* light = nullptr;
    for (const XNode& c : n.kids) {
      if (c.name == "bsdf") {
        if (mat >= 0) perr("shape " + *type + ": only one bsdf allowed");
        mat = bsdf(c);
      } else if (c.name == "light") {
        light = &c;
      } else if (c.name == "bssrdf" || c.name == "medium") {
        perr("shape " + *type + ": <" + c.name + "> is not rendered by the GPU path");
      }
    }
    if (mat < 0) {  // shapes without a BSDF get the reference's default diffuse (0.5)
      rtg_material m;
      std::memset(&m, 0, sizeof(m));
      m.type = RTG_MAT_DIFFUSE;
      m.albedo[0] = m.albedo[1] = m.albedo[2] = 0.5f;
      mats.push_back(m);
      mat = (int)mats.size() - 1;
    }
    Props P = props(n, "shape " + *type);
    const M4 T = P.transform("toWorld");
    if (lt == "sphere") {
      if (light) perr("shape Sphere: sphere emitters are not rendered by the GPU path");
      rtg_sphere_desc s;
      std::memset(&s, 0, sizeof(s));
      for (int i = 0; i < 16; i++) s.to_world[i] = (float)T.m[i];
      s.radius = (float)P.real("radius", 1.0);
      s.material = mat;
      spheres.push_back(s);
      return;
    }
    Mesh m;
    m.material = mat;
    if (lt == "mesh") {
      obj(resolve(P.str("path", "")), T, m);
    } else if (lt == "cube") {
      cube(T, m);
    } else if (lt == "plane") {
      plane(T, m);
    } else {
      perr("shape type " + *type + " is not rendered by the GPU path");
    }
    if (light) {
      const std::string ltype = type_of(*light);
      if (ltype != "meshlight" && ltype != "arealight") perr("shape light type must be MeshLight or AreaLight");
      Props L = props(*light, "light " + ltype);
      static const float zero[3] = {0, 0, 0};
      L.color("radiance", zero, m.emission);  // light/AreaLight.cpp:12
      m.emit = true;
    }
    meshes.push_back(std::move(m));
  }

  // Radiance RGBE (.hdr): new-style run-length encoded scanlines or flat pixels, -Y H +X W
  void read_hdr(const std::string& path, std::vector<float>& img, int& img_w, int& img_h) {
    std::ifstream f(path, std::ios::binary);
    if (!f) perr("cannot open image " + path);
    std::string line;
    bool fmt = false;
    while (std::getline(f, line)) {
      if (line.empty()) break;
      if (line.find("FORMAT=32-bit_rle_rgbe") != std::string::npos) fmt = true;
    }
    if (!fmt) perr(path + ": not a Radiance RGBE image (only .hdr images are supported)");
    std::getline(f, line);
    char ya[3], xa[3];
    int H = 0, W = 0;
    if (std::sscanf(line.c_str(), "%2s %d %2s %d", ya, &H, xa, &W) != 4 || std::string(ya) != "-Y" ||
        std::string(xa) != "+X" || W <= 0 || H <= 0)
      perr(path + ": unsupported RGBE orientation '" + line + "'");
    std::vector<unsigned char> row((size_t)W * 4);
    std::vector<float>& env = img;
    env.assign((size_t)W * H * 3, 0.f);
    for (int y = 0; y < H; y++) {
      unsigned char h[4];
      if (!f.read((char*)h, 4)) perr(path + ": truncated");
      if (W >= 8 && W < 32768 && h[0] == 2 && h[1] == 2 && ((h[2] << 8) | h[3]) == W) {
        for (int c = 0; c < 4; c++) {
          int x = 0;
          while (x < W) {
            unsigned char cnt;
            if (!f.read((char*)&cnt, 1)) perr(path + ": truncated");
            if (cnt > 128) {
              unsigned char v;
              if (!f.read((char*)&v, 1)) perr(path + ": truncated");
              cnt -= 128;
              if (x + cnt > W) perr(path + ": bad run length");
              for (int k = 0; k < cnt; k++) row[(size_t)(x++) * 4 + c] = v;
            } else {
              if (cnt == 0 || x + cnt > W) perr(path + ": bad run length");
              for (int k = 0; k < cnt; k++) {
                unsigned char v;
                if (!f.read((char*)&v, 1)) perr(path + ": truncated");
                row[(size_t)(x++) * 4 + c] = v;
              }
            }
          }
        }
      } else {  // flat
        std::memcpy(row.data(), h, 4);
        if (!f.read((char*)row.data() + 4, (std::streamsize)W * 4 - 4)) perr(path + ": truncated");
      }
      // the reference reads .hdr through stb_image's stbi_loadf (io/ImageLoader.cpp:80, gamma 1 for
      // HDR files at :45-48); its decode (ext/include/stb/stb_image.h:6830-6849) is mantissa x
      // 2^(exponent - 136), zero when the exponent byte is zero
      for (int x = 0; x < W; x++) {
        const unsigned char* p = &row[(size_t)x * 4];
        float s = p[3] ? std::ldexp(1.0f, (int)p[3] - (128 + 8)) : 0.f;
        for (int c = 0; c < 3; c++) env[((size_t)y * W + x) * 3 + c] = (float)p[c] * s;
      }
    }
    img_w = W;
    img_h = H;
  }

  void lights(const XNode& n) {
    for (const XNode& c : n.kids) {
      const std::string* type = c.get("type");
      if (!type) perr("light without a type");
      if (lower(*type) != "environmentlight") perr("light type " + *type + " is not rendered by the GPU path");
      Props P = props(c, "light EnvironmentLight");  // light/EnvironmentLight.cpp:8-73
      if (P.xf.count("toWorld")) {
        const M4 T = P.transform("toWorld");
        for (int i = 0; i < 16; i++)
          if (std::fabs(T.m[i] - M4::id().m[i]) > 1e-9) perr("EnvironmentLight: only the identity toWorld is supported");
      }
      if (P.real("gamma", 0.0) != 0.0 || P.real("offsetU", 0.0) != 0.0 || P.real("offsetV", 0.0) != 0.0 ||
          P.str("isProbe", "false") == "true")
        perr("EnvironmentLight: gamma, offsets and light probes are not supported");
      read_hdr(resolve(P.str("path", "")), env, env_w, env_h);
      env_scale = (float)P.real("scale", 1.0);
    }
  }

  void camera(const XNode& n) {
    if (type_of(n) != "noriperspective") perr("camera type must be NoriPerspective");
    M4 T = M4::id();
    for (const XNode& c : n.kids) {
      if (c.name == "transform") T = parse_transform(c);
      if (c.name == "filter") {
        // XMLLoader::handleReconstructionFilter (core/XMLLoader.cpp:425-438); parameters and
        // defaults of filter/ReconstructionFilter.cpp:13-19, GaussianFilter.cpp:21-27,
        // MitchellNetravaliFilter.cpp:15-21
        Props F = props(c, "filter");
        const std::string ft = type_of(c);
        rd.filter_radius[0] = (float)F.real("radiusX", 0.5);
        rd.filter_radius[1] = (float)F.real("radiusY", 0.5);
        if (ft == "boxfilter") {
          rd.filter = RTG_FILTER_BOX;
        } else if (ft == "gaussianfilter") {
          rd.filter = RTG_FILTER_GAUSSIAN;
          rd.filter_param[0] = (float)F.real("alpha", 1.0 / (2.0 * 0.5 * 0.5));
          rd.filter_param[1] = (float)F.real("omega", rd.filter_radius[0]);
        } else if (ft == "mitchellnetravalifilter") {
          rd.filter = RTG_FILTER_MITCHELL;
          rd.filter_param[0] = (float)F.real("B", 1.0 / 3);
          rd.filter_param[1] = (float)F.real("C", 1.0 / 3);
        } else {
          perr("camera filter type " + ft + " is not known");
        }
      }
    }
    Props P = props(n, "camera NoriPerspective");
    if (P.real("apertureRadius", 0.0) != 0.0) perr("camera: depth of field is not rendered by the GPU path");
    rd.fov = (float)P.real("fov", 30.0);  // camera/NoriPerspective.cpp:15-17, camera/Camera.cpp:11-14
    rd.near_clip = (float)P.real("nearClip", 1e-4);
    rd.far_clip = (float)P.real("farClip", 1e4);
    rd.width = P.integer("sizeX", 768);
    rd.height = P.integer("sizeY", 768);
    rd.use_camera_matrix = 1;
    for (int i = 0; i < 16; i++) rd.camera_to_world[i] = (float)T.m[i];
    have_camera = true;
  }

  void load(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) perr("cannot open scene file " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    size_t sl = path.find_last_of('/');
    dir = sl == std::string::npos ? "." : path.substr(0, sl);
    XNode root = XmlParser(text).parse();
    if (root.name != "scene") perr("root element must be <scene>");
    rd.spp = 1;
    rd.sampler = RTG_SAMPLER_RANDOM;
    rd.dimension = 4;
    rd.min_depth = 3;
    rd.max_depth = 1000;
    rd.seed = 1;
    // top-level references first so that objects may use them (the reference requires the same
    // order in the file; accepting either order is a superset)
    for (const XNode& c : root.kids) {
      if (c.name == "bsdf" && c.get("id")) ref_bsdf[*c.get("id")] = c;
      if (c.name == "transform" && c.get("id")) ref_xf[*c.get("id")] = parse_transform(c);
    }
    for (const XNode& c : root.kids) {
      if (c.name == "camera") {
        camera(c);
      } else if (c.name == "medium") {
        camera_medium(c);
      } else if (c.name == "integrator") {
        const std::string* t = c.get("type");
        const std::string ty = type_of(c);
        if (ty != "pathtracingmis" && ty != "volpathtracing")
          perr("integrator " + (t ? *t : std::string("?")) +
               " is not rendered by the GPU path (PathTracingMIS and VolPathTracing only)");
        // integrator/PathTracingMIS.cpp:12-20, VolPathTracing.cpp:9-17, Integrator.cpp:6-14
        Props P = props(c, "integrator " + (t ? *t : std::string("?")));
        rd.integrator = ty == "volpathtracing" ? RTG_INTEGRATOR_VOLPATH : RTG_INTEGRATOR_PATH_MIS;
        rd.min_depth = P.integer("minDepth", 3);
        rd.max_depth = P.integer("maxDepth", 1000);
        // integrator/Integrator.h:18-23 strategy names
        std::string s = P.str("strategy", "one_light_uniform");
        if (s == "one_light_uniform") strategy = RTG_LIGHTS_UNIFORM;
        else if (s == "one_light_weighted") strategy = RTG_LIGHTS_WEIGHTED;
        else if (s == "all_lights") strategy = RTG_LIGHTS_ALL;
        else perr("integrator strategy '" + s + "' is not known");
        have_integrator = true;
      } else if (c.name == "sampler") {
        Props P = props(c, "sampler");  // sampler/RandomSampler.cpp:10, sampler/StratifiedSampler.cpp:23-24
        if (type_of(c) == "randomsampler") rd.sampler = RTG_SAMPLER_RANDOM;
        else if (type_of(c) == "stratifiedsampler") rd.sampler = RTG_SAMPLER_STRATIFIED;
        else perr("sampler type must be RandomSampler or StratifiedSampler");
        rd.spp = P.integer("sampleNumber", 1);
        rd.dimension = P.integer("dimension", 4);
      } else if (c.name == "objects") {
        for (const XNode& o : c.kids) object(o);
      } else if (c.name == "lights") {
        lights(c);
      } else if (c.name == "parameters" || c.name == "bsdf" || c.name == "transform") {
        // threads, block size and output file do not change the image
      } else {
        perr("element <" + c.name + "> is not rendered by the GPU path");
      }
    }
    if (!have_camera) perr("scene has no camera");
  }

  // The scene's camera medium (XMLLoader::handleCameraMedium, core/XMLLoader.cpp:159-172):
  // Homogeneous (below), or Heterogeneous (medium/Heterogeneous.cpp:15-41: "albedo" default 0, "scale" default 1; as a
  // camera medium its "toWorld" and "centered" do not apply) over a GridVolume, with the
  // "phaseFunction" HGPhaseFunction ("g", default -0.8, HGPhaseFunction.cpp:7-10) or, by default,
  // IsotropicPhaseFunction (Medium.cpp:7-10; the same density 1/(4 pi) and uniform sampling as
  // HG with g = 0, which stands for it here).
  void camera_medium(const XNode& c) {
    const std::string* t = c.get("type");
    if (type_of(c) == "homogeneous") {
      // medium/Homogeneous.cpp:7-22: "sigmaA", "sigmaS" (default 0.15 each), "scale";
      // emission ("temperature" > 0) is not rendered
      Props P = props(c, "medium Homogeneous");
      static const float d15[3] = {0.15f, 0.15f, 0.15f};
      P.color("sigmaA", d15, medium_sigma_a);
      P.color("sigmaS", d15, medium_sigma_s);
      medium_scale = (float)P.real("scale", 1.0);
      if (P.real("temperature", 0.0) > 0.0) perr("medium Homogeneous: emission (temperature) is not rendered");
      medium_homogeneous = true;
      medium_g = phase(c);
      return;
    }
    if (type_of(c) != "heterogeneous")
      perr("medium " + (t ? *t : std::string("?")) +
           " is not rendered by the GPU path (Heterogeneous and Homogeneous only)");
    Props P = props(c, "medium Heterogeneous");
    static const float zero[3] = {0.f, 0.f, 0.f};
    P.color("albedo", zero, medium_albedo);
    medium_scale = (float)P.real("scale", 1.0);
    const XNode* vol = nullptr;
    for (const XNode& k : c.kids) {
      const std::string* nm = k.get("name");
      if (k.name == "volume" && nm && *nm == "volume") vol = &k;
    }
    if (!vol) perr("medium Heterogeneous: no <volume name=\"volume\">");
    if (type_of(*vol) != "gridvolume") perr("medium Heterogeneous: volume type must be GridVolume");
    Props V = props(*vol, "volume GridVolume");
    read_vol(resolve(V.str("path", "")));
    medium_g = phase(c);
  }

  // The medium's "phaseFunction" (Medium.cpp:7-10): HG g (default -0.8), or isotropic (g = 0)
  float phase(const XNode& c) {
    const XNode* ph = nullptr;
    for (const XNode& k : c.kids) {
      const std::string* nm = k.get("name");
      if (k.name == "phase" && nm && *nm == "phaseFunction") ph = &k;
    }
    if (!ph) return 0.f;
    const std::string pt = type_of(*ph);
    if (pt == "hgphasefunction") {
      Props H = props(*ph, "phase HGPhaseFunction");
      return (float)H.real("g", -0.8);
    }
    if (pt != "isotropicphasefunction") perr("phase function type must be HGPhaseFunction or IsotropicPhaseFunction");
    return 0.f;
  }

  // volume/GridVolume.cpp:7-64: "VOL", version 3, encoding, resolution x y z, channels, bounding
  // box (6 floats), 48 bytes in all; the reference then skips 4 more bytes before reading
  // x*y*z floats, so its grid starts at the file's second value, and its last value lies past
  // a standard file's end (left unset there; 0 here). Float32 single-channel files only.
  void read_vol(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) perr("cannot open volume file " + path);
    std::vector<char> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    if (b.size() < 48 || b[0] != 'V' || b[1] != 'O' || b[2] != 'L' || b[3] != 3)
      perr("volume file " + path + " is not a version 3 VOL file");
    int32_t hdr[5];
    std::memcpy(hdr, b.data() + 4, sizeof(hdr));  // encoding, x, y, z, channels
    if (hdr[0] != 1 || hdr[4] != 1) perr("volume file " + path + ": only float32 single-channel grids are supported");
    if (hdr[1] < 2 || hdr[2] < 2 || hdr[3] < 2) perr("volume file " + path + ": resolution below 2");
    const size_t n = (size_t)hdr[1] * hdr[2] * hdr[3];
    if (n > ((size_t)1 << 31)) perr("volume file " + path + ": grid too large");
    for (int k = 0; k < 3; k++) gres[k] = hdr[1 + k];
    grid.assign(n, 0.f);
    const size_t avail = b.size() > 52 ? (b.size() - 52) / 4 : 0;
    std::memcpy(grid.data(), b.data() + 52, std::min(n, avail) * 4);
  }
};

}  // namespace

extern "C" int rtg_spectrum_to_rgb(const float* lambda, const float* values, int32_t n, float* out_rgb) {
  if (!lambda || !values || !out_rgb || n < 2) return RTG_ERR_INVALID;
  spectrum_to_rgb(lambda, values, n, out_rgb);
  return RTG_OK;
}

extern "C" int rtg_scene_load_xml(rtg_ctx* ctx, const char* path, rtg_scene** out, rtg_render_desc* desc) {
  if (!ctx || !path || !out) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  try {
    Loader L;
    L.load(path);
    std::vector<rtg_mesh_desc> md(L.meshes.size());
    for (size_t i = 0; i < L.meshes.size(); i++) {
      Loader::Mesh& m = L.meshes[i];
      rtg_mesh_desc& d = md[i];
      std::memset(&d, 0, sizeof(d));
      d.vertices = m.v.data();
      d.num_vertices = (int32_t)(m.v.size() / 3);
      d.normals = m.n.empty() ? nullptr : m.n.data();
      d.uvs = m.uv.empty() ? nullptr : m.uv.data();
      d.indices = m.idx.data();
      d.num_triangles = (int32_t)(m.idx.size() / 3);
      d.material = m.material;
      d.emission = m.emit ? m.emission : nullptr;
    }
    rtg_scene_desc sd;
    std::memset(&sd, 0, sizeof(sd));
    sd.materials = L.mats.data();
    sd.num_materials = (int32_t)L.mats.size();
    sd.meshes = md.data();
    sd.num_meshes = (int32_t)md.size();
    sd.spheres = L.spheres.data();
    sd.num_spheres = (int32_t)L.spheres.size();
    sd.env_rgb = L.env.empty() ? nullptr : L.env.data();
    sd.env_width = L.env_w;
    sd.env_height = L.env_h;
    sd.env_scale = L.env_scale;
    sd.light_strategy = L.strategy;
    for (size_t i = 0; i < L.textures.size(); i++) L.textures[i].rgb = L.texdata[i].data();
    sd.textures = L.textures.empty() ? nullptr : L.textures.data();
    sd.num_textures = (int32_t)L.textures.size();
    if (L.medium_homogeneous) {
      sd.medium_homogeneous = 1;
      for (int k = 0; k < 3; k++) {
        sd.medium_sigma_a[k] = L.medium_sigma_a[k];
        sd.medium_sigma_s[k] = L.medium_sigma_s[k];
      }
      sd.medium_g = L.medium_g;
      sd.medium_scale = L.medium_scale;
    }
    if (!L.grid.empty()) {
      sd.medium_density = L.grid.data();
      for (int k = 0; k < 3; k++) {
        sd.medium_res[k] = L.gres[k];
        sd.medium_albedo[k] = L.medium_albedo[k];
      }
      sd.medium_g = L.medium_g;
      sd.medium_scale = L.medium_scale;
    }
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    *out = build_scene(ctx, &sd);
    if (desc) *desc = L.rd;
    return RTG_OK;
  } catch (const RtgError& e) {
    set_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_error(e.what());
    return RTG_ERR_PARSE;
  }
}
