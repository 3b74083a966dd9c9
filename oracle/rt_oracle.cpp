// CPU oracle: a plain C++17 restatement of the Iscandel/Raytracer MIS path-tracing hot path.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product (raytracer_amd/, include/) may include,
// link or call this file. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// use it, as the checker and as the timed CPU baseline.
//
// Every function cites the reference file:line it follows (paths relative to
// RaytracerNew2/src/). Arithmetic is IEEE float, no FMA contraction (built with
// -ffp-contract=off), and follows the reference's operation order. Where the reference's
// order is fixed by Eigen 3.4 internals (not vendored, absent here) this file uses:
//   * 3-term reductions (dot, squaredNorm, matrix*vector rows): a0 + (a1 + a2)
//     (Eigen's unrolled pairwise reduction); parity with Eigen itself is UNPINNED.
//   * normalized(): v / sqrt(squaredNorm) per component when squaredNorm > 0.
//   * 4x4 inverse: computed in double and rounded to float; parity with Eigen's float
//     inverse is UNPINNED (the camera and sphere matrices are inputs shared with the GPU path).
//
// Parity pins (which checks pin which function) are listed per function in the tests.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- constants
// core/BasicMath.h:16-21 (float branch)
static const float PI = 3.14159265358979323846f;
static const float INV_PI = 0.3183098861837906715377f;
static const float MAX_REAL = 10e7f;
static const float EPSILON = 1e-4f;
static const float ONE_MINUS_EPS = 0.99999994f;

// ---------------------------------------------------------------- vectors
struct V3 {
    float x, y, z;
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
struct V2 { float x, y; };
static inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
static inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
static inline V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
static inline V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
// Eigen unrolled reduction of 3 terms: a0 + (a1 + a2)
static inline float dot(V3 a, V3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
static inline V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float squaredNorm(V3 a) { return dot(a, a); }
static inline float norm(V3 a) { return std::sqrt(squaredNorm(a)); }
static inline V3 normalized(V3 a) {
    float z = squaredNorm(a);
    if (z > 0.f) return a / std::sqrt(z);
    return a;
}

// RGB colour, Color3 (core/Color.h:151-171,421-425)
struct C3 {
    float r, g, b;
};
static inline C3 c3(float v) { return C3{v, v, v}; }
static inline C3 c3(float r, float g, float b) { return C3{r, g, b}; }
static inline C3 operator+(C3 a, C3 b) { return c3(a.r + b.r, a.g + b.g, a.b + b.b); }
static inline C3 operator-(C3 a, C3 b) { return c3(a.r - b.r, a.g - b.g, a.b - b.b); }
static inline C3 operator*(C3 a, C3 b) { return c3(a.r * b.r, a.g * b.g, a.b * b.b); }
static inline C3 operator/(C3 a, C3 b) { return c3(a.r / b.r, a.g / b.g, a.b / b.b); }
static inline C3 operator*(C3 a, float s) { return c3(a.r * s, a.g * s, a.b * s); }
static inline C3 operator*(float s, C3 a) { return c3(s * a.r, s * a.g, s * a.b); }
static inline C3 operator/(C3 a, float s) { return c3(a.r / s, a.g / s, a.b / s); }
static inline C3& operator+=(C3& a, C3 b) { a = a + b; return a; }
static inline C3 operator+(C3 a, float s) { return c3(a.r + s, a.g + s, a.b + s); }
// Color.h:423-425
static inline float luminance(C3 c) { return 0.2126f * c.r + 0.7152f * c.g + 0.0722f * c.b; }
// Color.h:151-158
static inline float average(C3 c) {
    float s = 0.f;
    s += c.r; s += c.g; s += c.b;
    return s / 3.f;
}
// Color.h:163-171 (Eigen isZero(prec): |c_i| <= prec for all i)
static inline bool isZero(C3 c) {
    const float eps = 1e-5f;
    return std::fabs(c.r) <= eps && std::fabs(c.g) <= eps && std::fabs(c.b) <= eps;
}
static inline bool isNan(C3 c) { return std::isnan(c.r) || std::isnan(c.g) || std::isnan(c.b); }

// ---------------------------------------------------------------- transform (core/Transform.h:81-199)
struct M44 {
    float m[4][4];
};
static M44 identity44() {
    M44 r{};
    for (int i = 0; i < 4; i++) r.m[i][i] = 1.f;
    return r;
}
static M44 inverse44(const M44& a) {
    // Gauss-Jordan in double, rounded to float.
    double m[4][8];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            m[i][j] = a.m[i][j];
            m[i][j + 4] = (i == j) ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; c++) {
        int p = c;
        for (int r = c + 1; r < 4; r++)
            if (std::fabs(m[r][c]) > std::fabs(m[p][c])) p = r;
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
    M44 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) r.m[i][j] = (float)m[i][j + 4];
    return r;
}
// Transform.h:87-98 (point: explicit left-to-right sums, divide by w if w != 1 and != 0)
static inline V3 xformPoint(const M44& M, V3 p) {
    V3 o;
    o.x = M.m[0][0] * p.x + M.m[0][1] * p.y + M.m[0][2] * p.z + M.m[0][3];
    o.y = M.m[1][0] * p.x + M.m[1][1] * p.y + M.m[1][2] * p.z + M.m[1][3];
    o.z = M.m[2][0] * p.x + M.m[2][1] * p.y + M.m[2][2] * p.z + M.m[2][3];
    float w = M.m[3][0] * p.x + M.m[3][1] * p.y + M.m[3][2] * p.z + M.m[3][3];
    if (w != 1.f && w != 0.f) o = o / w;
    return o;
}
// Transform.h:109-116 (vector: top-left 3x3 times v, Eigen row reduction)
static inline V3 xformVector(const M44& M, V3 v) {
    return v3(M.m[0][0] * v.x + (M.m[0][1] * v.y + M.m[0][2] * v.z),
              M.m[1][0] * v.x + (M.m[1][1] * v.y + M.m[1][2] * v.z),
              M.m[2][0] * v.x + (M.m[2][1] * v.y + M.m[2][2] * v.z));
}
// Transform.h:176-184 (normal: transpose of inverse's 3x3 times n)
static inline V3 xformNormal(const M44& Inv, V3 n) {
    return v3(Inv.m[0][0] * n.x + (Inv.m[1][0] * n.y + Inv.m[2][0] * n.z),
              Inv.m[0][1] * n.x + (Inv.m[1][1] * n.y + Inv.m[2][1] * n.z),
              Inv.m[0][2] * n.x + (Inv.m[1][2] * n.y + Inv.m[2][2] * n.z));
}

// ---------------------------------------------------------------- ray (core/Ray.h:7-64)
struct Ray {
    V3 o, d, inv;
    float minT, maxT;
};
static inline Ray makeRay(V3 o, V3 d, float minT = EPSILON, float maxT = MAX_REAL) {
    Ray r;
    r.o = o;
    r.d = d;
    r.inv = v3(1.f / d.x, 1.f / d.y, 1.f / d.z);  // Ray.h:36-40 cwiseInverse
    r.minT = minT;
    r.maxT = maxT;
    return r;
}

// ---------------------------------------------------------------- bounding box (core/BoundingBox.h)
struct Box {
    V3 mn, mx;
};
static inline Box emptyBox() {  // BoundingBox.cpp:7-11
    return Box{v3(MAX_REAL, MAX_REAL, MAX_REAL), v3(-MAX_REAL, -MAX_REAL, -MAX_REAL)};
}
static inline Box unionBox(const Box& a, const Box& b) {  // BoundingBox.h:122-135
    Box r;
    r.mn = v3(std::min(a.mn.x, b.mn.x), std::min(a.mn.y, b.mn.y), std::min(a.mn.z, b.mn.z));
    r.mx = v3(std::max(a.mx.x, b.mx.x), std::max(a.mx.y, b.mx.y), std::max(a.mx.z, b.mx.z));
    return r;
}
static inline float surfaceValue(const Box& b) {  // BoundingBox.h:48-53
    V3 d = b.mx - b.mn;
    return 2.f * (d.x * d.y + d.y * d.z + d.x * d.z);
}
static inline V3 centroid(const Box& b) { return 0.5f * (b.mx + b.mn); }  // BoundingBox.h:68-76
// BoundingBox.h:154-183
static inline bool boxIntersect(const Box& b, const Ray& ray, float& nearT, float& farT) {
    nearT = -MAX_REAL;
    farT = MAX_REAL;
    for (int i = 0; i < 3; i++) {
        float origin = ray.o[i];
        float minVal = b.mn[i], maxVal = b.mx[i];
        if (ray.d[i] == 0) {
            if (origin < minVal || origin > maxVal) return false;
        } else {
            float t1 = (minVal - origin) * ray.inv[i];
            float t2 = (maxVal - origin) * ray.inv[i];
            if (t1 > t2) std::swap(t1, t2);
            nearT = std::max(t1, nearT);
            farT = std::min(t2, farT);
            if (!(nearT <= farT)) return false;
        }
    }
    return (ray.minT <= farT && ray.maxT >= nearT);
}

// ---------------------------------------------------------------- CDF (core/CDF.h:8-92)
struct CDF {
    std::vector<float> c{0.f};
    float sum = 0.f;
    void add(float pdf) {
        sum += pdf;
        c.push_back(sum);
    }
    void normalize() {
        if (sum != 0.f) {
            float factor = 1.f / sum;
            for (size_t i = 1; i < c.size(); i++) c[i] *= factor;
            c.back() = 1.f;
        }
    }
    int sample(float s) const {
        auto it = std::lower_bound(c.cbegin(), c.cend(), s);
        int index = std::max(0, (int)(it - c.cbegin() - 1));
        return std::min(index, (int)c.size() - 2);
    }
    int sample(float s, float& pdf) const {
        int index = sample(s);
        pdf = c[index + 1] - c[index];
        return index;
    }
    int sampleAndReuse(float& s) const {
        int index = sample(s);
        s = (s - c[index]) / (c[index + 1] - c[index]);
        return index;
    }
    int sampleAndReuse(float& s, float& pdf) const {
        int index = sampleAndReuse(s);
        pdf = c[index + 1] - c[index];
        return index;
    }
};

// ---------------------------------------------------------------- PCG32 (tools/Rng.h:68-141)
static const uint64_t PCG32_DEFAULT_STATE = 0x853c49e6748fea9bULL;
static const uint64_t PCG32_DEFAULT_STREAM = 0xda3e39cb94b95bdbULL;
static const uint64_t PCG32_MULT = 0x5851f42d4c957f2dULL;
struct Rng {
    uint64_t state = PCG32_DEFAULT_STATE, inc = PCG32_DEFAULT_STREAM;
    void seed(uint64_t initstate, uint64_t initseq = 1) {  // Rng.h:81-87
        state = 0U;
        inc = (initseq << 1u) | 1u;
        nextUInt();
        state += initstate;
        nextUInt();
    }
    uint32_t nextUInt() {  // Rng.h:90-96
        uint64_t oldstate = state;
        state = oldstate * PCG32_MULT + inc;
        uint32_t xorshifted = (uint32_t)(((oldstate >> 18u) ^ oldstate) >> 27u);
        uint32_t rot = (uint32_t)(oldstate >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((~rot + 1u) & 31));
    }
    uint32_t nextUInt(uint32_t bound) {  // Rng.h:99-128
        uint32_t threshold = (~bound + 1u) % bound;
        for (;;) {
            uint32_t r = nextUInt();
            if (r >= threshold) return r % bound;
        }
    }
    float nextReal() {  // Rng.h:132-141
        union {
            uint32_t u;
            float f;
        } x;
        x.u = (nextUInt() >> 9) | 0x3f800000UL;
        return x.f - 1.0f;
    }
};

// ---------------------------------------------------------------- samplers (sampler/*)
struct Sampler {
    virtual ~Sampler() {}
    virtual V2 next2D() = 0;
    virtual float next1D() = 0;
    virtual void seed(uint64_t s, uint64_t l) = 0;
    virtual void startPixel() {}
    virtual void advance() {}
    int number = 1;
};
// sampler/RandomSampler.cpp:28-45
struct RandomSampler : Sampler {
    Rng rng;
    V2 next2D() override {
        float a = rng.nextReal();
        float b = rng.nextReal();
        return V2{a, b};
    }
    float next1D() override { return rng.nextReal(); }
    void seed(uint64_t s, uint64_t l) override { rng.seed(s, l); }
};
// sampler/StratifiedSampler.cpp:18-124, StratifiedSampler.h:35-42
struct StratifiedSampler : Sampler {
    Rng rng;
    int dimension = 4, cur1D = 0, cur2D = 0, curSample = 0, sqrtDim = 1;
    std::vector<float> a1;  // [dim][number]
    std::vector<V2> a2;     // [dim][number]
    StratifiedSampler(int sampleNumber, int dim) {
        number = sampleNumber;
        dimension = dim;
        int i = 0;
        for (i = 0; i < number; i++)
            if (i * i >= number) break;
        if (i * i != number) number = i * i;
        sqrtDim = i;
        a1.assign((size_t)dimension * number, 0.f);
        a2.assign((size_t)dimension * number, V2{0.f, 0.f});
    }
    V2 next2D() override {
        if (cur2D < dimension) return a2[(size_t)(cur2D++) * number + curSample];
        float x = rng.nextReal();
        float y = rng.nextReal();
        return V2{x, y};
    }
    float next1D() override {
        if (cur1D < dimension) return a1[(size_t)(cur1D++) * number + curSample];
        return rng.nextReal();
    }
    void seed(uint64_t s, uint64_t l) override { rng.seed(s, l); }
    template <class T>
    void shuffle(std::vector<T>& arr, int dim) {
        for (int i = 0; i < number; i++) {
            int other = i + (int)rng.nextUInt((uint32_t)(number - i));
            std::swap(arr[(size_t)dim * number + i], arr[(size_t)dim * number + other]);
        }
    }
    void startPixel() override {
        curSample = 0;
        int cpt = 0;
        float dx = 1.f / sqrtDim;
        for (int i = 0; i < sqrtDim; i++)
            for (int j = 0; j < sqrtDim; j++) {
                for (int k = 0; k < dimension; k++) {
                    float jitterX = rng.nextReal();
                    float jitterY = rng.nextReal();
                    float x = std::min((i + jitterX) * dx, ONE_MINUS_EPS);
                    float y = std::min((j + jitterY) * dx, ONE_MINUS_EPS);
                    a2[(size_t)k * number + cpt] = V2{x, y};
                }
                cpt++;
            }
        dx = 1.f / number;
        for (int i = 0; i < number; i++)
            for (int k = 0; k < dimension; k++) {
                float val = std::min((i + rng.nextReal()) * dx, ONE_MINUS_EPS);
                a1[(size_t)k * number + i] = val;
            }
        for (int k = 0; k < dimension; k++) {
            shuffle(a1, k);
            shuffle(a2, k);
        }
    }
    void advance() override {
        curSample++;
        cur1D = 0;
        cur2D = 0;
    }
};

// ---------------------------------------------------------------- warps (core/Mapping.h)
// Mapping.h:110-140
static inline V2 squareToConcentricDisk(V2 s) {
    const float r1 = 2 * s.x - 1.f;
    const float r2 = 2 * s.y - 1.f;
    float rx, ry;
    if (r1 > -r2 && r1 > r2) {
        rx = r1;
        ry = (PI * r2) / (4 * r1);
    } else if (r1 < r2 && r1 > -r2) {
        rx = r2;
        ry = (PI / 4.f) * (2 - r1 / r2);
    } else if (r1 < -r2 && r1 < r2) {
        rx = -r1;
        ry = (PI / 4.f) * (4 + r2 / r1);
    } else {
        rx = -r2;
        ry = (PI / 4.f) * (6 - r1 / r2);
    }
    return V2{rx * std::cos(ry), rx * std::sin(ry)};
}
// Mapping.h:62-67
static inline V3 squareToCosineWeightedHemisphere(V2 s) {
    V2 t = squareToConcentricDisk(s);
    float z = std::sqrt(std::max(0.f, 1.f - t.x * t.x - t.y * t.y));
    return v3(t.x, t.y, z);
}
// Mapping.h:94-105
static inline V3 squareToTriangle(V2 s) {
    float r1 = s.x, r2 = s.y;
    float sqrtR1 = std::sqrt(r1);
    float alpha = 1 - sqrtR1;
    float beta = (1 - r2) * sqrtR1;
    float gamma = r2 * sqrtR1;
    return v3(alpha, beta, gamma);
}

// ---------------------------------------------------------------- spherical (core/Geometry.h:163-190)
static inline float sphericalPhiFromCartesian(V3 v) {
    float res = std::atan2(v.x, -v.z);
    return res < 0.f ? res + 2 * (float)3.14159265358979323846 : res;
}
static inline float sphericalThetaFromCartesian(V3 v) { return std::acos(v.y); }
static inline V3 cartesianFromSpherical(float theta, float phi) {
    float sinTheta = std::sin(theta);
    float cosTheta = std::cos(theta);
    float cosPhi = std::cos(phi);
    float sinPhi = std::sin(phi);
    return v3(sinPhi * sinTheta, cosTheta, -cosPhi * sinTheta);
}

// ---------------------------------------------------------------- frames (core/DifferentialGeometry.h)
struct Frame {
    V3 s, t, n;
};
// DifferentialGeometry.h:22-26,36-50
static inline Frame makeFrame(V3 n) {
    Frame f;
    f.n = n;
    V3 c;
    if (std::fabs(n.x) > std::fabs(n.y)) {
        float invLen = 1.0f / std::sqrt(n.x * n.x + n.z * n.z);
        c = v3(n.z * invLen, 0.0f, -n.x * invLen);
    } else {
        float invLen = 1.0f / std::sqrt(n.y * n.y + n.z * n.z);
        c = v3(0.0f, n.z * invLen, -n.y * invLen);
    }
    f.t = c;
    f.s = cross(c, n);
    return f;
}
// DifferentialGeometry.h:55-66
static inline V3 toLocal(const Frame& f, V3 v) { return v3(dot(v, f.s), dot(v, f.t), dot(v, f.n)); }
static inline V3 toWorld(const Frame& f, V3 v) { return f.s * v.x + f.t * v.y + f.n * v.z; }
static inline float tanTheta2(V3 v) {  // DifferentialGeometry.h:116-124
    float sin2 = 1.f - v.z * v.z;
    if (sin2 <= 0.f) return 0.f;
    return sin2 / (v.z * v.z);
}
static inline float sinTheta2(V3 v) { return 1.f - v.z * v.z; }
static inline float sinThetaF(V3 v) {  // DifferentialGeometry.h:95-101
    float s2 = sinTheta2(v);
    if (s2 < 0.f) return 0.f;
    return std::sqrt(s2);
}

// ---------------------------------------------------------------- Fresnel (core/Fresnel.h)
// Fresnel.h:30-49
static inline C3 fresnelConductor(C3 eta, C3 k, float cosI) {
    C3 etaPlusAbsorption = eta * eta + k * k;
    float cosI2 = cosI * cosI;
    C3 epaCos2 = etaPlusAbsorption * cosI2;
    C3 _2EtaCosI = 2 * eta * cosI;
    C3 rPar = (epaCos2 - _2EtaCosI + 1.f) / (epaCos2 + _2EtaCosI + 1.f);
    C3 rPerp = (etaPlusAbsorption - _2EtaCosI + cosI2) / (etaPlusAbsorption + _2EtaCosI + cosI2);
    return (rPar * rPar + rPerp * rPerp) * 0.5f;
}
// Fresnel.h:65-77
static inline float fresnelDielectric(float etaI, float etaT, float cosI, float cosT) {
    float etaTcosI = etaT * cosI;
    float etaIcosT = etaI * cosT;
    float etaIcosI = etaI * cosI;
    float etaTcosT = etaT * cosT;
    float rPar = (etaTcosI - etaIcosT) / (etaTcosI + etaIcosT);
    float rPerp = (etaIcosI - etaTcosT) / (etaIcosI + etaTcosT);
    return (rPar * rPar + rPerp * rPerp) * 0.5f;
}
// Fresnel.h:83-114 (Color version; all channels equal, kept scalar here)
static inline float estimateDielectric(float etaExt, float etaInt, float cosThetaI, float& etaI,
                                       float& etaT, float& relativeEta, float& cosThetaT) {
    etaI = etaExt;
    etaT = etaInt;
    bool isEntering = true;
    float fr = 0.f;
    if (cosThetaI <= 0.f) {
        isEntering = false;
        std::swap(etaI, etaT);
    }
    relativeEta = etaI / etaT;
    float sinThetaT = relativeEta * std::sqrt(std::max(0.f, 1 - cosThetaI * cosThetaI));
    cosThetaT = std::sqrt(std::max(0.f, 1 - sinThetaT * sinThetaT));
    if (sinThetaT >= 1.f) {
        cosThetaT = 0.f;
        fr = 1.f;
    } else {
        fr = fresnelDielectric(etaI, etaT, std::fabs(cosThetaI), cosThetaT);
        if (isEntering) cosThetaT = -cosThetaT;
    }
    return fr;
}
static inline float estimateDielectric(float etaExt, float etaInt, float cosThetaI) {
    float a, b, c, d;
    return estimateDielectric(etaExt, etaInt, cosThetaI, a, b, c, d);
}

// ---------------------------------------------------------------- microfacet (bsdf/MicrofacetDistribution.cpp)
enum { DIST_BECKMANN = 0, DIST_PHONG = 1, DIST_GGX = 2 };
// MicrofacetDistribution.cpp:29-52
static inline V3 mfSampleNormal(int type, V2 s, float alpha) {
    float theta;
    if (type == DIST_BECKMANN)
        theta = std::atan(std::sqrt(-alpha * alpha * std::log(1 - s.x)));
    else
        theta = std::atan(alpha * std::sqrt(s.x) / std::sqrt(1.f - s.x));
    float phi = 2 * PI * s.y;
    return v3(std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi), std::cos(theta));
}
// MicrofacetDistribution.cpp:54-87
static inline float mfD(int type, V3 wh, float alpha) {
    if (wh.z <= 0.f) return 0.f;
    if (type == DIST_BECKMANN) {
        float cosThetaH2 = wh.z * wh.z;
        float tanThetaH2 = sinTheta2(wh) / cosThetaH2;
        float alpha2 = alpha * alpha;
        return (std::exp(-tanThetaH2 / alpha2)) / (PI * alpha2 * cosThetaH2 * cosThetaH2);
    }
    float alpha2 = alpha * alpha;
    float cosThetaH2 = wh.z * wh.z;
    float tanThetaH2 = tanTheta2(wh);
    float smallSum = (alpha2 + tanThetaH2);
    return alpha2 / (PI * cosThetaH2 * cosThetaH2 * smallSum * smallSum);
}
// MicrofacetDistribution.cpp:105-140
static inline float mfG1(int type, V3 v, V3 m, float alpha) {
    if ((dot(v, m)) / (v.z) <= 0) return 0.f;
    if (type == DIST_BECKMANN) {
        float tanThetaV = sinThetaF(v) / v.z;
        if (tanThetaV == 0.f) return 1.f;
        float b = std::fabs(1.f / (alpha * tanThetaV));
        if (b < 1.6f) return (3.535f * b + 2.181f * b * b) / (1 + 2.276f * b + 2.577f * b * b);
        return 1.f;
    }
    float alpha2 = alpha * alpha;
    float tan2 = tanTheta2(v);
    return 2.f / (1 + std::sqrt(1 + alpha2 * tan2));
}
// MicrofacetDistribution.cpp:89-103
static inline float mfG(int type, V3 wi, V3 wo, V3 wh, float alpha) {
    return mfG1(type, wi, wh, alpha) * mfG1(type, wo, wh, alpha);
}
// MicrofacetDistribution.h:43-46
static inline float adjustedAlpha(float alpha, float cosThetaI) {
    return (1.2f - 0.2f * std::sqrt(std::fabs(cosThetaI))) * alpha;
}

// ---------------------------------------------------------------- BSDFs
// rough conductor / dielectric (RoughConductor, RoughDielectric) and the smooth, delta ones
// (Conductor, Dielectric); the values are rtg.h's RTG_MAT_*
enum { MAT_DIFFUSE = 0, MAT_CONDUCTOR = 1, MAT_DIELECTRIC = 2, MAT_SMOOTH_CONDUCTOR = 3, MAT_SMOOTH_DIELECTRIC = 4 };
// bsdf/BSDF.h:24-40
enum : uint32_t {
    BT_DELTA_REFLECTION = 0x2, BT_GLOSSY_REFLECTION = 0x4, BT_DELTA_TRANSMISSION = 0x8,
    BT_GLOSSY_TRANSMISSION = 0x10, BT_DIFFUSE = 0x200,
    BT_DELTA = BT_DELTA_REFLECTION | BT_DELTA_TRANSMISSION,
    BT_ALL = 0x1 | 0x2 | 0x4 | 0x8 | 0x4
};
struct Material {
    int type = MAT_DIFFUSE;
    C3 albedo = c3(0.5f);       // Diffuse albedo (Diffuse.cpp:11-18)
    C3 reflectance = c3(1.f);   // RoughConductor/RoughDielectric reflectanceTexture
    C3 transmitted = c3(1.f);   // RoughDielectric transmittedTexture
    float alphaRaw = 0.2f;      // ConstantTexture(alpha)
    float alpha = 0.2f;         // Color(alpha).average() as the reference evaluates it
    C3 eta = c3(0.f), k = c3(0.f);
    float etaInt = 1.5046f, etaExt = 1.000277f;
    int dist = DIST_BECKMANN;
    int tex[3] = {-1, -1, -1};  // image textures for albedo/reflectance, alpha, transmitted
};
struct BSDFInfo {  // bsdf/BSDF.h:82-139
    V3 wi, wo;
    uint32_t sampledType = BT_ALL;
    float relativeEta = 1.f;
    float pdf = -1.f;
};
static inline V3 reflectLocal(V3 v) { return v3(-v.x, -v.y, v.z); }  // BSDF.h:47
static inline V3 reflectN(V3 v, V3 n) { return 2 * dot(v, n) * n - v; }  // BSDF.h:49
static inline V3 refractN(V3 v, V3 n, float relEta, float cosT) {  // BSDF.h:53-56
    return n * (dot(v, n) * relEta + cosT) - v * relEta;
}

// ---- Diffuse (bsdf/Diffuse.cpp:28-106)
static C3 diffuseEval(const Material& m, const BSDFInfo& in) {
    if (in.wi.z <= 0.f || in.wo.z <= 0.f) return c3(0.f);
    return m.albedo * INV_PI;
}
static float diffusePdf(const Material&, const BSDFInfo& in) {
    if (in.wi.z <= 0.f || in.wo.z <= 0.f) return 0.f;
    float cosTheta = in.wo.z;
    return INV_PI * cosTheta;
}
static C3 diffuseSample(const Material& m, BSDFInfo& in, V2 s) {
    if (in.wi.z <= 0.f) return c3(0.f);
    in.sampledType = BT_DIFFUSE;
    in.wo = squareToCosineWeightedHemisphere(s);
    in.pdf = diffusePdf(m, in);
    return m.albedo;
}

// ---- RoughConductor (bsdf/RoughConductor.cpp:77-145)
static C3 conductorEval(const Material& m, const BSDFInfo& in) {
    float cosThetaI = in.wi.z, cosThetaO = in.wo.z;
    if (cosThetaI <= 0.f || cosThetaO <= 0.f) return c3(0.f);
    V3 wh = normalized(in.wi + in.wo);
    float D = mfD(m.dist, wh, m.alpha);
    float G = mfG(m.dist, in.wi, in.wo, wh, m.alpha);
    return m.reflectance * D * fresnelConductor(m.eta, m.k, cosThetaI) * G / (4 * cosThetaI * cosThetaO);
}
static float conductorPdf(const Material& m, const BSDFInfo& in) {
    float cosThetaI = in.wi.z, cosThetaO = in.wo.z;
    if (cosThetaI <= 0.f || cosThetaO <= 0.f) return 0.f;
    V3 wh = normalized(in.wi + in.wo);
    float cosThetaH = wh.z;
    float Jh = 1.f / (4 * std::fabs(dot(wh, in.wo)));
    float D = mfD(m.dist, wh, m.alpha);
    return D * cosThetaH * Jh;
}
static C3 conductorSample(const Material& m, BSDFInfo& in, V2 s) {
    if (in.wi.z <= 0.f) return c3(0.f);
    in.sampledType = BT_GLOSSY_REFLECTION;
    float alpha = adjustedAlpha(m.alpha, in.wi.z);
    V3 normal = mfSampleNormal(m.dist, s, alpha);
    in.wo = reflectN(in.wi, normal);
    float cosThetaO = in.wo.z;
    if (cosThetaO <= 0.f) return c3(0.f);
    in.pdf = conductorPdf(m, in);
    if (in.pdf <= 0) return c3(0.f);
    return conductorEval(m, in) * std::fabs(cosThetaO) / in.pdf;
}

// ---- RoughDielectric (bsdf/RoughDielectric.cpp:83-329, RoughDielectric.h:32-35)
static inline V3 halfDirTransmitted(V3 wi, V3 wo, float etaI, float etaT) {
    return normalized(-1.f * (etaI * wi + etaT * wo));
}
static inline float signf(float v) { return v >= 0.f ? 1.f : -1.f; }  // BasicMath.h:60-66
static C3 dielectricEvalReflection(const Material& m, const BSDFInfo& in) {  // :95-119
    float cosThetaI = in.wi.z, cosThetaO = in.wo.z;
    V3 wh = normalized(in.wi + in.wo);
    wh = signf(cosThetaI) * wh;
    float alpha = m.alpha;
    float F = estimateDielectric(m.etaExt, m.etaInt, dot(in.wi, wh));
    return m.reflectance * mfD(m.dist, wh, alpha) * c3(F) * mfG(m.dist, in.wi, in.wo, wh, alpha) /
           (4 * cosThetaI * cosThetaO);
}
static C3 dielectricEvalRefraction(const Material& m, const BSDFInfo& in) {  // :121-147
    float cosThetaI = in.wi.z, cosThetaO = in.wo.z;
    float etaI, etaT, relativeEta, cosThetaT;
    float fr = estimateDielectric(m.etaExt, m.etaInt, cosThetaI, etaI, etaT, relativeEta, cosThetaT);
    V3 wt = halfDirTransmitted(in.wi, in.wo, etaI, etaT);
    float alpha = m.alpha;
    float left = dot(in.wi, wt) * dot(in.wo, wt) / (cosThetaI * cosThetaO);
    C3 term = m.transmitted * left * etaT * etaT * ((-1.f * c3(fr)) + 1.f) *
              mfG(m.dist, in.wi, in.wo, wt, alpha) * mfD(m.dist, wt, alpha) * relativeEta * relativeEta;
    float denom = (etaI * (dot(in.wi, wt)) + etaT * dot(in.wo, wt));
    term = term / (denom * denom);
    return term;
}
static C3 dielectricEval(const Material& m, const BSDFInfo& in) {
    bool reflection = in.wi.z * in.wo.z > 0;
    return reflection ? dielectricEvalReflection(m, in) : dielectricEvalRefraction(m, in);
}
static float dielectricPdfRefraction(const Material& m, const BSDFInfo& in) {  // :262-278
    float cosThetaI = in.wi.z;
    float etaI, etaT, relativeEta, cosThetaT;
    estimateDielectric(m.etaExt, m.etaInt, cosThetaI, etaI, etaT, relativeEta, cosThetaT);
    V3 wt = halfDirTransmitted(in.wi, in.wo, etaI, etaT);
    float cosThetaWt = wt.z;
    float alpha = (1.2f - 0.2f * std::sqrt(std::fabs(cosThetaI))) * m.alpha;
    float halfDenom = (etaI * dot(in.wi, wt) + etaT * dot(in.wo, wt));
    float Jh = (etaT * etaT * dot(in.wo, wt)) / (halfDenom * halfDenom);
    return std::fabs(mfD(m.dist, wt, alpha) * cosThetaWt * Jh);
}
static float dielectricPdfReflection(const Material& m, const BSDFInfo& in) {  // :282-297
    float cosThetaI = in.wi.z;
    float alpha = (1.2f - 0.2f * std::sqrt(std::fabs(cosThetaI))) * m.alpha;
    V3 wh = normalized(in.wi + in.wo);
    wh = signf(cosThetaI) * wh;
    float cosThetaH = wh.z;
    float Jh = 1.f / (4 * dot(wh, in.wo));
    return std::fabs(mfD(m.dist, wh, alpha) * cosThetaH * Jh);
}
static float dielectricPdf(const Material& m, const BSDFInfo& in) {  // :301-329
    float cosThetaI = in.wi.z, cosThetaO = in.wo.z;
    bool reflection = cosThetaI * cosThetaO > 0;
    if (reflection) {
        V3 wh = normalized(in.wi + in.wo);
        wh = signf(cosThetaI) * wh;
        float fr = average(c3(estimateDielectric(m.etaExt, m.etaInt, dot(in.wi, wh))));
        return dielectricPdfReflection(m, in) * fr;
    }
    float etaI = m.etaExt, etaT = m.etaInt;
    if (cosThetaI < 0) std::swap(etaI, etaT);
    V3 wt = halfDirTransmitted(in.wi, in.wo, etaI, etaT);
    float fr = average(c3(estimateDielectric(m.etaExt, m.etaInt, dot(in.wi, wt))));
    return dielectricPdfRefraction(m, in) * (1.f - fr);
}
static C3 dielectricSample(const Material& m, BSDFInfo& in, V2 s) {  // :194-258
    float cosThetaI = in.wi.z;
    float alpha = adjustedAlpha(m.alpha, cosThetaI);
    V3 normal = mfSampleNormal(m.dist, s, alpha);
    float etaI, etaT, relativeEta, cosThetaT;
    float fr = average(c3(estimateDielectric(m.etaExt, m.etaInt, dot(in.wi, normal), etaI, etaT,
                                             relativeEta, cosThetaT)));
    if (s.x <= fr) {
        in.sampledType = BT_GLOSSY_REFLECTION;
        in.wo = reflectN(in.wi, normal);
        if (in.wi.z * in.wo.z <= 0) return c3(0.f);
        in.pdf = dielectricPdfReflection(m, in) * fr;
        if (in.pdf <= 0) return c3(0.f);
        V3 wh = normalized(in.wi + in.wo);
        wh = signf(cosThetaI) * wh;
        return c3(1.f) * dot(in.wi, normal) * mfG(m.dist, in.wi, in.wo, wh, alpha) / (cosThetaI * normal.z);
    }
    in.sampledType = BT_GLOSSY_TRANSMISSION;
    in.wo = refractN(in.wi, normal, relativeEta, cosThetaT);
    in.wo = normalized(in.wo);
    in.relativeEta = relativeEta;
    float cosThetaO = dot(in.wo, normal);
    if (in.wi.z * in.wo.z >= 0) return c3(0.f);
    in.pdf = dielectricPdfRefraction(m, in) * (1.f - fr);
    if (in.pdf <= 0) return c3(0.f);
    return dielectricEvalRefraction(m, in) * std::fabs(cosThetaO) / in.pdf;
}

// ---- Conductor (bsdf/Conductor.cpp:26-77). The integrator's infos carry the solid-angle measure
// (BSDF.h:95), so eval and pdf are 0 there; sample returns the mirror direction, pdf 1.
static C3 smoothConductorEval(const Material&, const BSDFInfo&) { return c3(0.f); }
static float smoothConductorPdf(const Material&, const BSDFInfo&) { return 0.f; }
static C3 smoothConductorSample(const Material& m, BSDFInfo& in) {
    if (in.wi.z <= 0.f) return c3(0.f);
    in.sampledType = BT_DELTA_REFLECTION;
    in.wo = reflectLocal(in.wi);
    in.pdf = 1.f;
    float cosI = in.wi.z;
    return m.reflectance * fresnelConductor(m.eta, m.k, cosI);
}
// ---- Dielectric (bsdf/Dielectric.cpp:28-104): eval and pdf 0; sample reflects with probability
// fr (weight: reflectance), else refracts (weight: transmitted * relEta^2); relativeEta is set
// for both outcomes.
static C3 smoothDielectricEval(const Material&, const BSDFInfo&) { return c3(0.f); }
static float smoothDielectricPdf(const Material&, const BSDFInfo&) { return 0.f; }
static C3 smoothDielectricSample(const Material& m, BSDFInfo& in, V2 s) {
    float etaI, etaT, rel, cosT;
    float fr = estimateDielectric(m.etaExt, m.etaInt, in.wi.z, etaI, etaT, rel, cosT);
    in.relativeEta = rel;
    if (s.x <= fr) {
        in.sampledType = BT_DELTA_REFLECTION;
        in.wo = reflectLocal(in.wi);
        in.pdf = fr;
        return m.reflectance;
    }
    in.sampledType = BT_DELTA_TRANSMISSION;
    in.wo = v3(rel * -in.wi.x, rel * -in.wi.y, cosT);  // BSDF.h:51
    in.pdf = 1.f - fr;
    return m.transmitted * rel * rel;
}

static C3 bsdfEval(const Material& m, const BSDFInfo& in) {
    switch (m.type) {
        case MAT_DIFFUSE: return diffuseEval(m, in);
        case MAT_CONDUCTOR: return conductorEval(m, in);
        case MAT_SMOOTH_CONDUCTOR: return smoothConductorEval(m, in);
        case MAT_SMOOTH_DIELECTRIC: return smoothDielectricEval(m, in);
        default: return dielectricEval(m, in);
    }
}
static float bsdfPdf(const Material& m, const BSDFInfo& in) {
    switch (m.type) {
        case MAT_DIFFUSE: return diffusePdf(m, in);
        case MAT_CONDUCTOR: return conductorPdf(m, in);
        case MAT_SMOOTH_CONDUCTOR: return smoothConductorPdf(m, in);
        case MAT_SMOOTH_DIELECTRIC: return smoothDielectricPdf(m, in);
        default: return dielectricPdf(m, in);
    }
}
static C3 bsdfSample(const Material& m, BSDFInfo& in, V2 s) {
    switch (m.type) {
        case MAT_DIFFUSE: return diffuseSample(m, in, s);
        case MAT_CONDUCTOR: return conductorSample(m, in, s);
        case MAT_SMOOTH_CONDUCTOR: return smoothConductorSample(m, in);
        case MAT_SMOOTH_DIELECTRIC: return smoothDielectricSample(m, in, s);
        default: return dielectricSample(m, in, s);
    }
}

// ---------------------------------------------------------------- scene primitives
struct Mesh {  // geometricShape/Mesh.{h,cpp}
    std::vector<V3> v, n;
    std::vector<V2> uv;
    std::vector<int> idx;     // 3 per triangle
    std::vector<int> triMat;  // material per triangle
    int light = -1;           // index into Scene::lights, or -1
    CDF cdf;
    float area = 0.f;
    int primBase = 0;         // global id of triangle 0
};
struct Sphere {  // geometricShape/Sphere.{h,cpp}
    M44 toWorld, inv;
    float radius = 1.f;
    int mat = 0;
    int primId = 0;
};
enum { LIGHT_MESH = 0, LIGHT_ENV = 1 };
struct Light {
    int type = LIGHT_MESH;
    int mesh = -1;
    C3 radiance = c3(0.f);
    // environment map (light/EnvironmentLight.cpp)
    int W = 0, H = 0;
    std::vector<C3> map;  // row-major y*W+x, unscaled
    float factor = 1.f;
    std::vector<CDF> rows;
    CDF marginal;
    std::vector<float> weight;
    float sphereRadius = 0.f;
    C3 getValue(int x, int y) const { return map[(size_t)y * W + x] * factor; }
};

struct Hit {  // core/Intersection.h:59-75
    float t = MAX_REAL;
    int prim = -1;
    V2 uv{0.f, 0.f};
    V3 p{0.f, 0.f, 0.f};
    Frame sh, tr;
};

// ---------------------------------------------------------------- BVH (accelStruct/Bvh.{h,cpp})
// Items are either triangles (in a mesh BVH) or objects (mesh BVHs and spheres, in the scene BVH).
struct LinearNode {  // Bvh.h:41-54
    Box box;
    int32_t axis = 0, first = 0, count = 0, right = -1;
};
struct TreeNode {  // Bvh.h:13-36
    Box box = emptyBox();
    std::unique_ptr<TreeNode> child[2];
    int first = 0, count = 0, axis = 0;
};
struct Bvh {
    std::vector<LinearNode> nodes;
    std::vector<int> items;  // item ids in leaf order
    int numberOfNodes = 0;
    // item geometry callbacks are given as arrays
    const std::vector<Box>* itemBox = nullptr;
    const std::vector<V3>* itemCentroid = nullptr;

    V3 wbCentroid(int it) const { return centroid((*itemBox)[it]); }

    // Bvh.cpp:51-79
    void build(const std::vector<Box>& boxes, const std::vector<V3>& cents) {
        itemBox = &boxes;
        itemCentroid = &cents;
        items.resize(boxes.size());
        for (size_t i = 0; i < boxes.size(); i++) items[i] = (int)i;
        Box box = emptyBox();
        for (size_t i = 0; i < boxes.size(); i++) box = unionBox(box, boxes[i]);
        TreeNode root;
        root.box = box;
        numberOfNodes = 0;
        buildBinNode(0, (int)items.size(), &root);
        nodes.assign(numberOfNodes, LinearNode());
        int num = 0;
        flatten(root, &num);
    }
    static int largestAxis(const Box& b) {  // Bvh.cpp:372-381
        V3 d = b.mx - b.mn;
        if (d.x > d.y && d.x > d.z) return 0;
        if (d.z > d.x && d.z > d.y) return 2;
        return 1;
    }
    // Bvh.cpp:83-246 (SAH branch: full sort over three axes)
    void buildNode(int start, int end, TreeNode* node) {
        numberOfNodes++;
        int number = end - start;
        std::vector<float> surfaces(number);
        float minCost = 1.f * number;
        int bestIndex = -1, bestAxis = -1;
        for (int axis = 0; axis < 3; axis++) {
            auto cmp = [&](int a, int b) { return wbCentroid(a)[axis] < wbCentroid(b)[axis]; };
            std::sort(items.begin() + start, items.begin() + end, cmp);
            Box box = emptyBox();
            for (int i = start, cpt = 0; i < end; i++, cpt++) {
                box = unionBox(box, (*itemBox)[items[i]]);
                surfaces[cpt] = surfaceValue(box);
            }
            if (axis == 0) node->box = box;
            float factor = 1.f / surfaceValue(node->box);
            Box rightBox = emptyBox();
            for (int j = number - 1; j >= 1; j--) {
                int index = start + j;
                rightBox = unionBox(rightBox, (*itemBox)[items[index]]);
                float leftArea = surfaces[j - 1];
                float rightArea = surfaceValue(rightBox);
                float leftNumber = (float)j;
                float rightNumber = (float)(number - j);
                float cost = 2 * 1.f + factor * (leftNumber * leftArea + rightNumber * rightArea);
                if (cost < minCost) {
                    minCost = cost;
                    bestIndex = index;
                    bestAxis = axis;
                }
            }
        }
        if (bestIndex == -1) {
            node->first = start;
            node->count = number;
        } else {
            auto cmp = [&](int a, int b) { return wbCentroid(a)[bestAxis] < wbCentroid(b)[bestAxis]; };
            std::sort(items.begin() + start, items.begin() + end, cmp);
            node->child[0].reset(new TreeNode);
            node->child[1].reset(new TreeNode);
            buildNode(start, bestIndex, node->child[0].get());
            buildNode(bestIndex, end, node->child[1].get());
        }
    }
    // Bvh.cpp:611-771
    void buildBinNode(int start, int end, TreeNode* node) {
        numberOfNodes++;
        int number = end - start;
        if (number < 32) {
            numberOfNodes--;
            buildNode(start, end, node);
            return;
        }
        const int BIN = 16;
        int axis = largestAxis(node->box);
        float mn = node->box.mn[axis];
        float mx = node->box.mx[axis];
        float binConstant = BIN / (mx - mn);
        node->axis = axis;
        Box binBox[BIN];
        int binCount[BIN];
        for (int i = 0; i < BIN; i++) {
            binBox[i] = emptyBox();
            binCount[i] = 0;
        }
        auto binOf = [&](int it) {
            return std::min(std::max((int)(((*itemCentroid)[it][axis] - mn) * binConstant), 0), BIN - 1);
        };
        for (int i = start; i < end; i++) {
            int index = binOf(items[i]);
            binBox[index] = unionBox(binBox[index], (*itemBox)[items[i]]);
            binCount[index]++;
        }
        float minCost = 1.f * number;
        int bestIndex = -1;
        Box leftBox[BIN];
        leftBox[0] = binBox[0];
        for (int i = 1; i < BIN; i++) {
            binCount[i] += binCount[i - 1];
            leftBox[i] = unionBox(leftBox[i - 1], binBox[i]);
        }
        float factor = 1.f / surfaceValue(node->box);
        Box rightBox = binBox[BIN - 1];
        Box bestRight = emptyBox();
        for (int i = BIN - 2; i >= 0; i--) {
            float leftNumber = (float)binCount[i];
            float rightNumber = (float)(number - binCount[i]);
            float cost = 2 * 1.f + factor * (leftNumber * surfaceValue(leftBox[i]) + rightNumber * surfaceValue(rightBox));
            if (cost < minCost) {
                minCost = cost;
                bestIndex = i;
                bestRight = rightBox;
            }
            rightBox = unionBox(rightBox, binBox[i]);
        }
        if (bestIndex == -1) {
            buildNode(start, end, node);
            return;
        }
        int leftCount = binCount[bestIndex];
        std::vector<int> tmp(number);
        int tl = 0, tr = leftCount;
        for (int i = start; i < end; i++) {
            if (binOf(items[i]) <= bestIndex)
                tmp[tl++] = items[i];
            else
                tmp[tr++] = items[i];
        }
        std::copy(tmp.begin(), tmp.end(), items.begin() + start);
        node->child[0].reset(new TreeNode);
        node->child[0]->box = leftBox[bestIndex];
        node->child[1].reset(new TreeNode);
        node->child[1]->box = bestRight;
        buildBinNode(start, start + leftCount, node->child[0].get());
        buildBinNode(start + leftCount, end, node->child[1].get());
    }
    // Bvh.cpp:587-607
    int flatten(TreeNode& node, int* current) {
        LinearNode& ln = nodes[*current];
        int offset = (*current)++;
        ln.box = node.box;
        if (!node.child[0] && !node.child[1]) {
            ln.first = node.first;
            ln.count = node.count;
        } else {
            ln.axis = node.axis;
            flatten(*node.child[0], current);
            int r = flatten(*node.child[1], current);
            nodes[offset].right = r;
        }
        return offset;
    }
};

// ---------------------------------------------------------------- image textures
// texture/ImageTexture.cpp:7-64 on the decoded texel array (the reference's ImageLoader output),
// times a ScaleTexture factor (texture/ScaleTexture.cpp:19-22)
struct Texture {
    int W = 0, H = 0;
    bool invertX = false, invertY = false;
    float scale = 1.f;       // ImageTexture "scale" (absolute value)
    float valueScale = 1.f;  // ScaleTexture "scale"
    std::vector<C3> px;      // row-major, H rows of W
    // BasicMath.h:74-90
    C3 interp2(float x, float y) const {
        float ax = x - (int)x;
        float ay = y - (int)y;
        int xMin = std::min(std::max((int)x, 0), W - 2);
        int yMin = std::min(std::max((int)y, 0), H - 2);
        int xMax = xMin + 1 < W ? xMin + 1 : xMin;
        xMax = std::min(std::max(xMax, 0), W - 1);
        int yMax = yMin + 1 < H ? yMin + 1 : yMin;
        yMax = std::min(std::max(yMax, 0), H - 1);
        auto A = [&](int xx, int yy) { return px[(size_t)yy * W + xx]; };
        return (1 - ax) * (1 - ay) * A(xMin, yMin) + (1 - ax) * ay * A(xMin, yMax) +
               ax * (1 - ay) * A(xMax, yMin) + ax * ay * A(xMax, yMax);
    }
    static float sign(float v) { return v >= 0.f ? 1.f : -1.f; }  // BasicMath.h:60-66
    // ImageTexture.cpp:27-64
    C3 eval(V2 uv) const {
        float x, y;
        if (!invertY)
            y = uv.y * H;
        else
            y = (int)uv.y + (sign(uv.y) * 1.f - (uv.y - (int)uv.y)) * H;
        if (!invertX)
            x = uv.x * (W - 1.f);
        else
            x = (int)uv.x + (sign(uv.x) * 1.f - (uv.x - (int)uv.x)) * W;
        x *= scale;
        y *= scale;
        if (x >= W)
            x = (int)x % W + (x - (int)x);
        else if (x < 0)
            x = W - std::abs((int)x) % W + (x - (int)x);
        if (y >= H)
            y = (int)y % H + (y - (int)y);
        else if (y < 0)
            y = H - std::abs((int)y) % H + (y - (int)y);
        return interp2(x, y) * valueScale;
    }
};

// ---------------------------------------------------------------- participating medium
// The camera medium (core/XMLLoader.cpp:159-172): medium/Heterogeneous.cpp over
// volume/GridVolume.cpp with medium/HGPhaseFunction.cpp. A camera medium never receives
// setOwnerBBox, so its world-to-array transform is the identity and its world box is [0,1]^3
// (Heterogeneous.cpp:15-41).
struct Medium {
    bool homogeneous = false;  // medium/Homogeneous.cpp (unbounded, chromatic coefficients)
    C3 sigA = c3(0.f), sigS = c3(0.f), sigT = c3(0.f);
    std::vector<float> data;  // x fastest, then y, then z
    int res[3] = {0, 0, 0};
    C3 albedo = c3(0.f);
    float g = 0.f, scale = 1.f, maxSigmaT = 0.f;
    bool present() const { return homogeneous || !data.empty(); }
    static C3 expC(C3 c) { return c3(std::exp(c.r), std::exp(c.g), std::exp(c.b)); }
    // Homogeneous.cpp:26-60: channel chosen uniformly, exponential distance with its extinction
    // (+EPSILON) from minT compared with maxT as an absolute distance, spectral-MIS weights
    bool sampleDistanceHomogeneous(const Ray& ray, Sampler& sampler, float& t, C3& weight) const {
        int channel = (int)(sampler.next1D() * 3);
        float sig = (channel == 0 ? sigT.r : channel == 1 ? sigT.g : sigT.b) + EPSILON;
        t = -std::log(1.f - sampler.next1D()) / sig;
        if (t > ray.maxT) {
            t = ray.maxT;
            float pdf = average(expC(-1.f * sigT * t));
            weight = expC(-1.f * sigT * (ray.maxT - ray.minT)) / pdf;
            return false;
        }
        float pdf = average(sigT * expC(-1.f * sigT * t));
        float len = t;
        t += ray.minT;
        weight = sigS * expC(-1.f * sigT * len) / pdf;
        return true;
    }
    // GridVolume.cpp:96-105 getMaxSigmaT x scale (Heterogeneous.cpp:40)
    void finalize() {
        float m = data[0];
        for (size_t i = 1; i < data.size(); i++)
            if (data[i] > m) m = data[i];
        maxSigmaT = m * scale;
    }
    // GridVolume.cpp:118-148 x Heterogeneous.cpp:403-406
    float lookupSigmaT(V3 p) const {
        float qx = p.x * (float)res[0], qy = p.y * (float)res[1], qz = p.z * (float)res[2];
        float fx = std::floor(qx), fy = std::floor(qy), fz = std::floor(qz);
        if (!(fx >= 0.f && fy >= 0.f && fz >= 0.f)) return 0.f;
        if (fx >= (float)(res[0] - 1) || fy >= (float)(res[1] - 1) || fz >= (float)(res[2] - 1)) return 0.f;
        size_t row = (size_t)res[0], slab = row * (size_t)res[1];
        size_t off = (size_t)fz * slab + (size_t)fy * row + (size_t)fx;
        const float d000 = data[off], d001 = data[off + 1], d010 = data[off + row], d011 = data[off + row + 1];
        const float d100 = data[off + slab], d101 = data[off + slab + 1], d110 = data[off + slab + row],
                    d111 = data[off + slab + row + 1];
        float w1x = qx - fx, w1y = qy - fy, w1z = qz - fz;
        float w0x = 1 - w1x, w0y = 1 - w1y, w0z = 1 - w1z;
        float v = ((d000 * w0x + d001 * w1x) * w0y + (d010 * w0x + d011 * w1x) * w1y) * w0z +
                  ((d100 * w0x + d101 * w1x) * w0y + (d110 * w0x + d111 * w1x) * w1y) * w1z;
        return v * scale;
    }
    // Heterogeneous.cpp:491-546 (delta tracking). The reference starts at ray.minT; its tentative
    // collisions before the box entry meet zero density and are all rejected, so starting at the
    // entry draws the same distribution. Bounded at 16384 tentative collisions.
    bool sampleDistance(const Ray& ray, Sampler& sampler, float& t, C3& weight) const {
        if (homogeneous) return sampleDistanceHomogeneous(ray, sampler, t, weight);
        Box box{v3(0, 0, 0), v3(1, 1, 1)};
        float nearT, farT;
        weight = c3(1.f);
        if (!boxIntersect(box, ray, nearT, farT)) {
            t = ray.maxT;
            return false;
        }
        float tMax = std::min(ray.maxT, farT);
        t = std::max(ray.minT, nearT);
        float sigma = 0.f;
        for (int it = 0;; it++) {
            if (it >= 16384) return false;
            t += -std::log(1.f - sampler.next1D()) / maxSigmaT;
            if (t > tMax) {
                t = tMax;
                return false;
            }
            sigma = lookupSigmaT(ray.o + t * ray.d);
            if (sampler.next1D() < sigma / maxSigmaT) break;
        }
        weight = sigma != 0.f ? albedo : c3(0.f);
        return sigma != 0.f;
    }
    // Homogeneous.cpp:26-30 (exp(-sigma_t (maxT - minT))) or the ratio tracking below, per channel
    C3 transmittanceC(const Ray& ray, Sampler& sampler) const {
        if (homogeneous) return expC(-1.f * sigT * (ray.maxT - ray.minT));
        return c3(transmittance(ray, sampler));
    }
    // Heterogeneous.cpp:419-489 (ratio tracking with roulette below 0.1; length maxT - minT)
    float transmittance(const Ray& ray, Sampler& sampler) const {
        float tMax = ray.maxT - ray.minT;
        float tr = 1.f, sampledT;
        float nearT, farT;
        Box box{v3(0, 0, 0), v3(1, 1, 1)};
        if (!boxIntersect(box, ray, nearT, farT)) return 1.f;
        tMax = std::min(tMax, farT);
        sampledT = std::max(nearT, ray.minT);
        int breaker = 0;
        while (true) {
            sampledT += -std::log(1.f - sampler.next1D()) / maxSigmaT;
            float proba = lookupSigmaT(ray.o + sampledT * ray.d);
            if (sampledT >= tMax) return tr;
            tr *= 1 - std::max(0.f, proba / maxSigmaT);
            const float threshold = .1f;
            if (tr < threshold) {
                float q = std::max(.05f, 1 - tr);
                if (sampler.next1D() < q) return 0.f;
                tr /= 1 - q;
            }
            if (breaker > 1024) return 0.f;
            breaker++;
        }
    }
    // HGPhaseFunction.cpp:61-65
    float hgEval(V3 wi, V3 wo) const {
        float cosTheta = dot(-wi, wo);
        return 0.07957747154594766788450883936693f * (1 - g * g) / std::pow(1 + g * g - 2 * g * cosTheta, 1.5f);
    }
    // HGPhaseFunction.cpp:27-45: the inversion as written, the direction around world +y
    // (Geometry.h:181-188) whatever wi is, weight 1
    V3 hgSample(V3 wi, V2 s, float& pdf) const {
        float cosTheta;
        if (g < EPSILON)
            cosTheta = 1.f - 2.f * s.x;
        else
            cosTheta = (1 + g * g - ((1 - g * g) * (1 - g * g) / (1 - g + 2 * g * s.x))) / (2 * g);
        float theta = std::acos(cosTheta);
        float phi = 2 * PI * s.y;
        V3 wo = cartesianFromSpherical(theta, phi);
        pdf = hgEval(wi, wo);
        return wo;
    }
};

// ---------------------------------------------------------------- scene
struct Scene {
    std::vector<Texture> textures;
    Medium medium;           // camera medium (empty: none)
    int integrator = 0;      // 0 PathTracingMIS, 1 VolPathTracing
    std::vector<Material> mats;
    std::vector<Mesh> meshes;
    std::vector<Sphere> spheres;
    std::vector<Light> lights;
    int envLight = -1;
    // integrator light-selection CDF (integrator/Integrator.cpp:24-38)
    CDF lightWeights;
    int strategy = 1;  // 0 all_lights, 1 one_light_uniform, 2 one_light_weighted (Integrator.h:18-22)
    // acceleration: one BVH per mesh + one scene BVH over objects (two-level, MeshFactory.h:123)
    std::vector<Bvh> meshBvh;
    std::vector<std::vector<Box>> meshTriBox;
    std::vector<std::vector<V3>> meshTriCent;
    Bvh sceneBvh;
    std::vector<Box> objBox;   // objects: meshes first, then spheres
    std::vector<V3> objCent;
    Box worldBox;
    int ntri = 0;
    // prim id -> (mesh, tri) lookup
    std::vector<int> primMesh;
    bool finalized = false;

    // ---- triangle (geometricShape/Triangle.cpp:46-89)
    bool triIntersect(const Mesh& m, int tri, const Ray& ray, float& t, V2& uv) const {
        const V3& p0 = m.v[m.idx[3 * tri]];
        const V3& p1 = m.v[m.idx[3 * tri + 1]];
        const V3& p2 = m.v[m.idx[3 * tri + 2]];
        V3 e1 = p1 - p0;
        V3 e2 = p2 - p0;
        V3 s = ray.o - p0;
        V3 s1 = cross(ray.d, e2);
        V3 s2 = cross(s, e1);
        float d = dot(s1, e1);
        if (std::fabs(d) < 1e-8) return false;
        float inv = 1 / d;
        uv.x = inv * dot(s1, s);
        if (uv.x < 0. || uv.x > 1.) return false;
        uv.y = inv * dot(s2, ray.d);
        if (uv.y < 0. || uv.x + uv.y > 1.) return false;
        t = inv * dot(s2, e2);
        if (t < ray.minT || t > ray.maxT) return false;
        return true;
    }
    // ---- sphere (geometricShape/Sphere.cpp:38-111)
    bool sphereIntersect(const Sphere& sp, const Ray& ray, float& t) const {
        V3 lo = xformPoint(sp.inv, ray.o);
        V3 ld = xformVector(sp.inv, ray.d);
        V3 s2o = lo - v3(0.f, 0.f, 0.f);
        float a = 1;
        float b = 2 * dot(ld, s2o);
        float c = dot(s2o, s2o) - sp.radius * sp.radius;
        float delta = b * b - 4 * a * c;
        if (delta < 0) return false;
        float t1 = (-b + std::sqrt(delta)) / (2 * a);
        float t2 = (-b - std::sqrt(delta)) / (2 * a);
        float tNear = t1, tFar = t2;
        if (tNear > tFar) std::swap(tNear, tFar);
        bool hit = false;
        if (tNear <= ray.minT) {
            if (tFar > ray.minT && tFar < ray.maxT) {
                t = tFar;
                hit = true;
            }
        } else if (tNear > ray.minT && tNear < ray.maxT) {
            hit = true;
            t = tNear;
        }
        return hit;
    }

    void finalize() {
        // per-mesh BVHs (MeshFactory.h:39-139) and area CDFs (Mesh.cpp:8-32)
        ntri = 0;
        primMesh.clear();
        meshBvh.assign(meshes.size(), Bvh());
        meshTriBox.assign(meshes.size(), {});
        meshTriCent.assign(meshes.size(), {});
        objBox.clear();
        objCent.clear();
        for (size_t mi = 0; mi < meshes.size(); mi++) {
            Mesh& m = meshes[mi];
            m.primBase = ntri;
            int nt = (int)m.idx.size() / 3;
            m.cdf = CDF();
            m.area = 0.f;
            auto& boxes = meshTriBox[mi];
            auto& cents = meshTriCent[mi];
            boxes.resize(nt);
            cents.resize(nt);
            for (int t = 0; t < nt; t++) {
                const V3& p0 = m.v[m.idx[3 * t]];
                const V3& p1 = m.v[m.idx[3 * t + 1]];
                const V3& p2 = m.v[m.idx[3 * t + 2]];
                float surface = 0.5f * norm(cross(p1 - p0, p2 - p0));  // Triangle.h:128-135
                m.cdf.add(surface);
                m.area += surface;
                Box b;  // Triangle.h:65-88
                b.mn = v3(std::min(std::min(p0.x, p1.x), p2.x), std::min(std::min(p0.y, p1.y), p2.y),
                          std::min(std::min(p0.z, p1.z), p2.z));
                b.mx = v3(std::max(std::max(p0.x, p1.x), p2.x), std::max(std::max(p0.y, p1.y), p2.y),
                          std::max(std::max(p0.z, p1.z), p2.z));
                boxes[t] = b;
                cents[t] = (1.f / 3) * (p0 + p1 + p2);  // Triangle.h:92-99
                primMesh.push_back((int)mi);
            }
            m.cdf.normalize();
            ntri += nt;
            meshBvh[mi].build(boxes, cents);
            Box rb = meshBvh[mi].nodes.empty() ? emptyBox() : meshBvh[mi].nodes[0].box;
            objBox.push_back(rb);
            objCent.push_back(centroid(rb));  // Bvh.h:178
        }
        for (size_t si = 0; si < spheres.size(); si++) {
            Sphere& sp = spheres[si];
            sp.primId = ntri + (int)si;
            // GeometricShape.h:87 + Transform.h:141-153 (8 transformed corners)
            float r = sp.radius;
            Box lb{v3(-r, -r, -r), v3(r, r, r)};
            Box wb = emptyBox();
            V3 corners[8] = {v3(lb.mn.x, lb.mn.y, lb.mn.z), v3(lb.mn.x, lb.mn.y, lb.mx.z),
                             v3(lb.mn.x, lb.mx.y, lb.mx.z), v3(lb.mx.x, lb.mx.y, lb.mx.z),
                             v3(lb.mx.x, lb.mn.y, lb.mn.z), v3(lb.mx.x, lb.mx.y, lb.mn.z),
                             v3(lb.mx.x, lb.mn.y, lb.mx.z), v3(lb.mn.x, lb.mx.y, lb.mn.z)};
            for (auto& c : corners) {
                V3 p = xformPoint(sp.toWorld, c);
                wb = unionBox(wb, Box{p, p});
            }
            objBox.push_back(wb);
            objCent.push_back(xformPoint(sp.toWorld, v3(0.f, 0.f, 0.f)));  // Sphere.h:74-77
        }
        sceneBvh.build(objBox, objCent);
        worldBox = sceneBvh.nodes.empty() ? emptyBox() : sceneBvh.nodes[0].box;
        // lights (EnvironmentLight.cpp:169-174)
        for (auto& L : lights)
            if (L.type == LIGHT_ENV) {
                V3 c = centroid(worldBox);
                L.sphereRadius = 2 * std::max(norm(worldBox.mn - c), norm(worldBox.mx - c));
            }
        // integrator light CDF (Integrator.cpp:24-38)
        lightWeights = CDF();
        float pdfUniform = 1.f / lights.size();
        for (auto& L : lights) {
            if (strategy == 2)
                lightWeights.add(luminance(lightPower(L)));
            else
                lightWeights.add(pdfUniform);
        }
        lightWeights.normalize();
        finalized = true;
    }
    C3 lightPower(const Light& L) const {
        if (L.type == LIGHT_MESH) return PI * L.radiance * meshes[L.mesh].area;  // MeshLight.cpp:29-33
        return 4 * PI * L.sphereRadius * L.sphereRadius * c3(1.f);             // EnvironmentLight.cpp:51-54
    }

    // ---- mesh BVH traversal (Bvh.cpp:392-481), items = triangles of one mesh
    bool meshIntersect(int mi, Ray& ray, Hit& hit, bool shadow) const {
        const Bvh& bvh = meshBvh[mi];
        const Mesh& m = meshes[mi];
        if (ray.minT == EPSILON)
            ray.minT = std::max(ray.minT, ray.minT * std::max(std::max(std::fabs(ray.o.x), std::fabs(ray.o.y)), std::fabs(ray.o.z)));
        if (bvh.nodes.empty() || ray.maxT < ray.minT) return false;
        int stack[64];
        int sp = 0, cur = 0;
        int dirIsNeg[3] = {ray.inv.x < 0, ray.inv.y < 0, ray.inv.z < 0};
        bool hasHit = false;
        while (true) {
            const LinearNode& node = bvh.nodes[cur];
            float nearT, farT;
            if (!boxIntersect(node.box, ray, nearT, farT)) {
                if (sp == 0) break;
                cur = stack[--sp];
                continue;
            }
            if (node.right != -1) {
                if (dirIsNeg[node.axis]) {
                    stack[sp++] = cur + 1;
                    cur = node.right;
                } else {
                    stack[sp++] = node.right;
                    cur++;
                }
            } else {
                for (int i = node.first; i < node.first + node.count; i++) {
                    int tri = bvh.items[i];
                    float t;
                    V2 uv;
                    if (triIntersect(m, tri, ray, t, uv)) {
                        hit.p = ray.o + t * ray.d;  // Primitive.cpp:19-39
                        hit.t = t;
                        hit.uv = uv;
                        hit.prim = m.primBase + tri;
                        if (shadow) return true;
                        ray.maxT = hit.t;
                        hasHit = true;
                    }
                }
                if (sp == 0) break;
                cur = stack[--sp];
            }
        }
        return hasHit;
    }
    // ---- scene BVH traversal (Bvh.cpp:392-481), items = objects
    bool intersect(const Ray& _ray, Hit& hit, bool shadow) const {
        Ray ray = _ray;
        if (ray.minT == EPSILON)
            ray.minT = std::max(ray.minT, ray.minT * std::max(std::max(std::fabs(ray.o.x), std::fabs(ray.o.y)), std::fabs(ray.o.z)));
        const Bvh& bvh = sceneBvh;
        if (bvh.nodes.empty() || ray.maxT < ray.minT) return false;
        int stack[64];
        int sp = 0, cur = 0;
        int dirIsNeg[3] = {ray.inv.x < 0, ray.inv.y < 0, ray.inv.z < 0};
        bool hasHit = false;
        int nm = (int)meshes.size();
        while (true) {
            const LinearNode& node = bvh.nodes[cur];
            float nearT, farT;
            if (!boxIntersect(node.box, ray, nearT, farT)) {
                if (sp == 0) break;
                cur = stack[--sp];
                continue;
            }
            if (node.right != -1) {
                if (dirIsNeg[node.axis]) {
                    stack[sp++] = cur + 1;
                    cur = node.right;
                } else {
                    stack[sp++] = node.right;
                    cur++;
                }
            } else {
                for (int i = node.first; i < node.first + node.count; i++) {
                    int obj = bvh.items[i];
                    bool h;
                    if (obj < nm) {
                        h = meshIntersect(obj, ray, hit, shadow);
                    } else {
                        const Sphere& spr = spheres[obj - nm];
                        float t;
                        h = sphereIntersect(spr, ray, t);
                        if (h) {
                            hit.p = ray.o + t * ray.d;
                            hit.t = t;
                            hit.uv = V2{0.f, 0.f};  // Sphere.cpp:100-104: uv unused by constant textures
                            hit.prim = spr.primId;
                        }
                    }
                    if (h) {
                        if (shadow) return true;
                        ray.maxT = hit.t;
                        hasHit = true;
                    }
                }
                if (sp == 0) break;
                cur = stack[--sp];
            }
        }
        if (hasHit && !shadow) diffGeom(hit);
        return hasHit;
    }
    // brute force closest hit over all primitives in global id order, same arithmetic and tie rule
    bool intersectBrute(const Ray& _ray, Hit& hit, bool shadow) const {
        Ray ray = _ray;
        if (ray.minT == EPSILON)
            ray.minT = std::max(ray.minT, ray.minT * std::max(std::max(std::fabs(ray.o.x), std::fabs(ray.o.y)), std::fabs(ray.o.z)));
        if (ray.maxT < ray.minT) return false;
        bool hasHit = false;
        for (size_t mi = 0; mi < meshes.size(); mi++) {
            const Mesh& m = meshes[mi];
            int nt = (int)m.idx.size() / 3;
            for (int tri = 0; tri < nt; tri++) {
                float t;
                V2 uv;
                if (triIntersect(m, tri, ray, t, uv)) {
                    hit.p = ray.o + t * ray.d;
                    hit.t = t;
                    hit.uv = uv;
                    hit.prim = m.primBase + tri;
                    if (shadow) return true;
                    ray.maxT = t;
                    hasHit = true;
                }
            }
        }
        for (const Sphere& spr : spheres) {
            float t;
            if (sphereIntersect(spr, ray, t)) {
                hit.p = ray.o + t * ray.d;
                hit.t = t;
                hit.uv = V2{0.f, 0.f};
                hit.prim = spr.primId;
                if (shadow) return true;
                ray.maxT = t;
                hasHit = true;
            }
        }
        if (hasHit && !shadow) diffGeom(hit);
        return hasHit;
    }
    // Triangle.cpp:130-215, Sphere.cpp:115-139
    void diffGeom(Hit& hit) const {
        if (hit.prim < ntri) {
            const Mesh& m = meshes[primMesh[hit.prim]];
            int tri = hit.prim - m.primBase;
            int i0 = m.idx[3 * tri], i1 = m.idx[3 * tri + 1], i2 = m.idx[3 * tri + 2];
            const V3 &p0 = m.v[i0], &p1 = m.v[i1], &p2 = m.v[i2];
            float b0 = 1 - (hit.uv.x + hit.uv.y);
            float b1 = hit.uv.x, b2 = hit.uv.y;
            hit.p = b0 * p0 + b1 * p1 + b2 * p2;
            if (m.uv.size() >= 3) {
                const V2 &u0 = m.uv[i0], &u1 = m.uv[i1], &u2 = m.uv[i2];
                hit.uv = V2{b0 * u0.x + b1 * u1.x + b2 * u2.x, b0 * u0.y + b1 * u1.y + b2 * u2.y};
            }
            hit.tr = makeFrame(normalized(cross(p1 - p0, p2 - p0)));
            if (!m.n.empty()) {
                V3 tmpN = b0 * m.n[i0] + b1 * m.n[i1] + b2 * m.n[i2];
                hit.sh = makeFrame(normalized(tmpN));
            } else {
                hit.sh = hit.tr;
            }
        } else {
            const Sphere& sp = spheres[hit.prim - ntri];
            V3 local = xformPoint(sp.inv, hit.p);
            V3 nrm = normalized(local - v3(0.f, 0.f, 0.f));  // Sphere.cpp:27-34
            hit.tr = makeFrame(normalized(xformNormal(sp.inv, nrm)));
            hit.sh = hit.tr;
        }
    }
    // Texture coordinates of a hit: a triangle's uv (barycentrics, or interpolated from its mesh's
    // uvs, Triangle.cpp:141-171); a sphere's are the spherical angles of the normalized world-space
    // hit point (Sphere.cpp:100-104)
    V2 textureUv(const Hit& hit) const {
        if (hit.prim < ntri) return hit.uv;
        V3 n = normalized(hit.p);
        return V2{sphericalPhiFromCartesian(n) / (2 * PI), sphericalThetaFromCartesian(n) / PI};
    }
    // The material at a hit with its textures evaluated (Diffuse.cpp:34,91, RoughConductor.cpp:87-90,
    // RoughDielectric.cpp:108-140, Conductor.cpp:42; alpha = the texture's average)
    Material materialAt(const Hit& hit) const {
        Material m = mats[primMaterial(hit.prim)];
        if (m.tex[0] < 0 && m.tex[1] < 0 && m.tex[2] < 0) return m;
        V2 uv = textureUv(hit);
        if (m.tex[0] >= 0) {
            C3 c = textures[m.tex[0]].eval(uv);
            if (m.type == MAT_DIFFUSE) m.albedo = c;
            else m.reflectance = c;
        }
        if (m.tex[1] >= 0) m.alpha = average(textures[m.tex[1]].eval(uv));
        if (m.tex[2] >= 0) m.transmitted = textures[m.tex[2]].eval(uv);
        return m;
    }
    int primMaterial(int prim) const {
        if (prim < ntri) {
            const Mesh& m = meshes[primMesh[prim]];
            return m.triMat[prim - m.primBase];
        }
        return spheres[prim - ntri].mat;
    }
    int primLight(int prim) const {
        if (prim < ntri) return meshes[primMesh[prim]].light;
        return -1;
    }

    // ---- lights
    struct LightSample {  // light/Light.h:15-46
        V3 point{0, 0, 0};
        C3 intensity = c3(0.f);
        V3 toLight{0, 0, 0};
        float distance = 0.f;
        V3 normal{0, 0, 0};
        float pdf = 0.f;
    };
    // Mesh.cpp:38-68 + Triangle.cpp:251-273
    void meshSample(const Mesh& m, V2 p, V3& sampled, V3& normal) const {
        V2 s = p;
        int index = m.cdf.sampleAndReuse(s.x);
        int i0 = m.idx[index * 3], i1 = m.idx[index * 3 + 1], i2 = m.idx[index * 3 + 2];
        const V3 &p0 = m.v[i0], &p1 = m.v[i1], &p2 = m.v[i2];
        V3 bc = squareToTriangle(s);
        sampled = p0 * bc.x + p1 * bc.y + p2 * bc.z;
        if (!m.n.empty()) {
            normal = normalized(bc.x * m.n[i0] + bc.y * m.n[i1] + bc.z * m.n[i2]);
        } else {
            normal = normalized(cross(p1 - p0, p2 - p0));
            normal = normalized(normal);
        }
    }
    // Mesh.cpp:70-93
    float meshPdf(const Mesh& m, V3 pFrom, V3 sampled, V3 normal) const {
        V3 lightToShape = pFrom - sampled;
        float distance = norm(lightToShape);
        lightToShape = normalized(lightToShape);
        float cosine = dot(normal, lightToShape);
        if (cosine > 0) return distance * distance / (m.area * cosine);
        return 0.f;
    }
    // EnvironmentLight.cpp:176-186
    void checkXY(const Light& L, float& x, float& y) const {
        if (x >= L.W)
            x = (int)x % L.W + (x - (int)x);
        else if (x < 0)
            x = L.W - std::abs((int)x) % L.W + (x - (int)x);
        if (y >= L.H)
            y = (int)y % L.H + (y - (int)y);
        else if (y < 0)
            y = L.H - std::abs((int)y) % L.H + (y - (int)y);
    }
    // core/BasicMath.h:74-90
    C3 interp2(const Light& L, float x, float y) const {
        float ax = x - (int)x;
        float ay = y - (int)y;
        int xMin = std::min(std::max((int)x, 0), L.W - 2);
        int yMin = std::min(std::max((int)y, 0), L.H - 2);
        int xMax = xMin + 1 < L.W ? xMin + 1 : xMin;
        xMax = std::min(std::max(xMax, 0), L.W - 1);
        int yMax = yMin + 1 < L.H ? yMin + 1 : yMin;
        yMax = std::min(std::max(yMax, 0), L.H - 1);
        auto A = [&](int xx, int yy) { return L.map[(size_t)yy * L.W + xx]; };
        return (1 - ax) * (1 - ay) * A(xMin, yMin) + (1 - ax) * ay * A(xMin, yMax) +
               ax * (1 - ay) * A(xMax, yMin) + ax * ay * A(xMax, yMax);
    }
    // EnvironmentLight.cpp:137-166 (identity toWorld)
    C3 envLe(const Light& L, V3 direction) const {
        V3 dir = normalized(direction);
        float fromPhi = (sphericalPhiFromCartesian(dir) / (2 * PI)) * L.W;
        float fromTheta = (sphericalThetaFromCartesian(dir) / PI) * L.H;
        checkXY(L, fromPhi, fromTheta);
        return interp2(L, fromPhi, fromTheta) * L.factor;
    }
    // EnvironmentLight.cpp:105-135
    float envPdf(const Light& L, V3 toLight) const {
        V3 dir = normalized(toLight);
        float fromPhi = (sphericalPhiFromCartesian(dir) / (2 * PI)) * L.W;
        float fromTheta = (sphericalThetaFromCartesian(dir) / PI) * L.H;
        checkXY(L, fromPhi, fromTheta);
        float sinTheta = std::sin(fromTheta);
        float normalizationFactor = L.weight[(int)fromTheta] / (L.marginal.sum * (2 * PI / L.W) * (PI / L.H));
        float lum = luminance(interp2(L, fromPhi, fromTheta) * L.factor);
        float pdf = lum * normalizationFactor;
        pdf /= std::max(std::fabs(sinTheta), EPSILON);
        return pdf;
    }
    // EnvironmentLight.cpp:56-103
    LightSample envSample(const Light& L, V3 pFrom, V2 s) const {
        LightSample ls;
        float pdfRow, pdfCol;
        int row = L.marginal.sample(s.x, pdfRow);
        int col = L.rows[row].sample(s.y, pdfCol);
        float rcol = col + 0.f * L.W;
        float rrow = row + 0.f * L.H;
        checkXY(L, rcol, rrow);
        row = (int)rrow;
        col = (int)rcol;
        float theta = PI * row / (float)L.H;
        float phi = 2 * PI * col / (float)L.W;
        float sinTheta = std::sin(theta);
        ls.toLight = cartesianFromSpherical(theta, phi);
        ls.distance = L.sphereRadius;
        ls.toLight = normalized(ls.toLight);
        ls.point = pFrom + L.sphereRadius * ls.toLight;
        ls.intensity = L.getValue(std::min(std::max(col, 0), L.W - 1), std::min(std::max(row, 0), L.H - 1));
        ls.pdf = luminance(ls.intensity) * L.weight[row] / (L.marginal.sum * (2 * PI / L.W) * (PI / L.H));
        ls.pdf /= std::max(std::fabs(sinTheta), EPSILON);
        return ls;
    }
    // MeshLight.cpp:32-64
    LightSample meshLightSample(const Light& L, V3 pFrom, V2 s) const {
        LightSample ls;
        const Mesh& m = meshes[L.mesh];
        meshSample(m, s, ls.point, ls.normal);
        ls.toLight = ls.point - pFrom;
        ls.distance = norm(ls.toLight);
        ls.toLight = ls.toLight / ls.distance;
        float cosine = dot(ls.normal, -ls.toLight);
        if (cosine <= 0) {
            ls.intensity = c3(0.f);
            ls.pdf = 0.f;
        } else {
            ls.intensity = L.radiance;
            ls.pdf = ls.distance * ls.distance / (m.area * cosine);
        }
        return ls;
    }
    LightSample lightSample(int li, V3 pFrom, V2 s) const {
        const Light& L = lights[li];
        return L.type == LIGHT_ENV ? envSample(L, pFrom, s) : meshLightSample(L, pFrom, s);
    }
    // AreaLight.cpp:57-61 / EnvironmentLight.cpp:105-135
    float lightPdf(int li, V3 pFrom, V3 point, V3 normal, V3 toLight) const {
        const Light& L = lights[li];
        if (L.type == LIGHT_ENV) return envPdf(L, toLight);
        return meshPdf(meshes[L.mesh], pFrom, point, normal);
    }
    // AreaLight.h:27-32 / MeshLight.h:21-26 (one-sided emission)
    C3 meshLe(const Light& L, V3 dir, V3 n) const {
        if (dot(dir, n) > 0.f) return L.radiance;
        return c3(0.f);
    }
};

// ---------------------------------------------------------------- integrator (integrator/*)
struct PathTracer {
    const Scene* scene;
    int minDepth = 3, maxDepth = 1000;
    uint64_t rays = 0;  // extension + shadow rays traced

    static float powerHeuristic(float a, float b) {  // Integrator.h:78-83
        float left = a * a;
        float right = b * b;
        return left / (left + right);
    }
    // Integrator.cpp:222-238
    C3 sampleOneLight(int li, V3 p, V2 s, Scene::LightSample& info) {
        info = scene->lightSample(li, p, s);
        Ray shadowRay = makeRay(p, info.toLight, EPSILON, info.distance - EPSILON);
        if (info.pdf > 0 && luminance(info.intensity) > 0) {
            Hit tmp;
            rays++;
            if (!scene->intersect(shadowRay, tmp, true)) return info.intensity / info.pdf;
        }
        return c3(0.f);
    }
    // Integrator.cpp:176-218
    C3 sampleLightDirect(V3 p, V2 _s, Scene::LightSample& info, int& li) {
        if (scene->lights.empty()) return c3(0.f);
        V2 s = _s;
        float lightWeight;
        li = scene->lightWeights.sampleAndReuse(s.x, lightWeight);
        C3 value = sampleOneLight(li, p, s, info);
        info.pdf *= lightWeight;
        value = value / lightWeight;
        return value;
    }
    // PathTracingMIS.cpp:27-263 (one-light strategies)
    C3 li(Sampler& sampler, const Ray& _ray) {
        C3 radiance = c3(0.f);
        C3 throughput = c3(1.f);
        float eta = 1.f;
        int depth = 0;
        Ray ray = _ray;
        bool emission = true;
        Hit hit;
        rays++;
        if (!scene->intersect(ray, hit, false)) {
            if (scene->envLight >= 0) radiance += scene->envLe(scene->lights[scene->envLight], ray.d);
            return radiance;
        }
        while (depth < maxDepth) {
            const Material bsdf = scene->materialAt(hit);
            int hitLight = scene->primLight(hit.prim);
            if (emission && hitLight >= 0)
                radiance += throughput * scene->meshLe(scene->lights[hitLight], -ray.d, hit.sh.n);
            // light sampling, :119-153
            if (scene->strategy == 0) {
                for (int li = 0; li < (int)scene->lights.size(); li++) {
                    Scene::LightSample info;
                    C3 value = sampleOneLight(li, hit.p, sampler.next2D(), info);
                    if (!isZero(value)) {
                        BSDFInfo bi;
                        bi.wi = toLocal(hit.sh, info.toLight);
                        bi.wo = toLocal(hit.sh, -ray.d);
                        float cosTheta = std::fabs(bi.wi.z);
                        float bsdfPDF = bsdfPdf(bsdf, bi);
                        float weight = powerHeuristic(info.pdf, bsdfPDF);
                        radiance += throughput * value * bsdfEval(bsdf, bi) * cosTheta * weight;
                    }
                }
            } else {
                Scene::LightSample info;
                int li = -1;
                C3 value = sampleLightDirect(hit.p, sampler.next2D(), info, li);
                if (!isZero(value)) {
                    BSDFInfo bi;
                    bi.wi = toLocal(hit.sh, info.toLight);
                    bi.wo = toLocal(hit.sh, -ray.d);
                    float cosTheta = std::fabs(bi.wi.z);
                    float bsdfPDF = bsdfPdf(bsdf, bi);
                    float weight = powerHeuristic(info.pdf, bsdfPDF);
                    radiance += throughput * value * bsdfEval(bsdf, bi) * cosTheta * weight;
                }
            }
            // BSDF sampling, :156-164
            BSDFInfo bi;
            bi.wi = toLocal(hit.sh, -ray.d);
            V2 s = sampler.next2D();
            C3 bsdfValue = bsdfSample(bsdf, bi, s);
            if (isZero(bsdfValue)) break;
            // :166-189
            Hit next;
            Ray reflected = makeRay(hit.p, toWorld(hit.sh, bi.wo));
            int caught = -1;
            C3 radianceLight = c3(0.f);
            rays++;
            bool hitNext = scene->intersect(reflected, next, false);
            if (hitNext) {
                int nl = scene->primLight(next.prim);
                if (nl >= 0) {
                    caught = nl;
                    radianceLight = scene->meshLe(scene->lights[nl], -reflected.d, next.sh.n);
                }
            } else {
                caught = scene->envLight;
                if (caught < 0) return radiance;
                radianceLight = scene->envLe(scene->lights[caught], reflected.d);
            }
            // :191-220
            if (caught >= 0) {
                float pdfLight = (bi.sampledType & BT_DELTA) ? 0.f
                                 : scene->lightPdf(caught, hit.p, next.p, next.sh.n, reflected.d);
                float weight = powerHeuristic(bi.pdf, pdfLight);
                radiance += throughput * radianceLight * bsdfValue * weight;
                if (caught == scene->envLight) return radiance;
            }
            // :222-253
            throughput = throughput * bsdfValue;
            eta *= bi.relativeEta;
            if (depth >= minDepth) {
                float stopVal = std::min(0.95f, luminance(throughput) * eta * eta);
                if (sampler.next1D() < stopVal)
                    throughput = throughput / stopVal;
                else
                    return radiance;
            }
            hit = next;
            ray = reflected;
            emission = false;
            depth++;
        }
        return radiance;
    }
};

// VolPathTracing::li (integrator/VolPathTracing.cpp:25-346) and its transmittance helpers
// (Integrator.cpp:81-170, VolPathTracing.cpp:348-423), for scenes whose only medium is the camera
// medium and whose shapes all carry a BSDF: no medium transitions, and the first surface on a
// shadow ray occludes it.
struct VolPathTracer {
    const Scene* scene;
    int minDepth = 3, maxDepth = 1000;
    uint64_t rays = 0;

    const Medium* medium() const { return scene->medium.present() ? &scene->medium : nullptr; }
    // Integrator.cpp:81-135
    C3 evalTransmittance(const Ray& shadowRay, Sampler& sampler) {
        Hit tmp;
        rays++;
        if (scene->intersect(shadowRay, tmp, true)) return c3(0.f);
        const Medium* m = medium();
        return m ? m->transmittanceC(shadowRay, sampler) : c3(1.f);
    }
    // Integrator.cpp:137-170
    C3 sampleAttenuatedLightDirect(V3 p, Sampler& sampler, Scene::LightSample& info) {
        if (scene->lights.empty()) return c3(0.f);
        V2 s = sampler.next2D();
        float lightWeight;
        int li = scene->lightWeights.sampleAndReuse(s.x, lightWeight);
        info = scene->lightSample(li, p, s);
        C3 value = c3(0.f);
        if (info.pdf > 0 && luminance(info.intensity) > 0) {
            Ray shadowRay = makeRay(p, info.toLight, EPSILON, info.distance - EPSILON);
            C3 tr = evalTransmittance(shadowRay, sampler);
            info.pdf *= lightWeight;
            value = tr * info.intensity / info.pdf;
        }
        return value;
    }
    // Mesh::pdf (Mesh.cpp:70-93) as VolPathTracing.cpp:152 calls it after a medium vertex: from the
    // point where the sampled ray met the light to the point light sampling drew. On a planar
    // emitter the two are coplanar and the cosine is 0 in exact arithmetic (pdf 0); rounding makes
    // the reference's sign arbitrary, so a cosine within 1e-4 of 0 is taken as 0.
    float mediumHitLightPdf(int li, V3 pFrom, V3 sampled, V3 normal) const {
        const Light& L = scene->lights[li];
        const Mesh& m = scene->meshes[L.mesh];
        V3 v = pFrom - sampled;
        float distance = norm(v);
        v = normalized(v);
        float cosine = dot(normal, v);
        if (cosine > 1e-4f) return distance * distance / (m.area * cosine);
        return 0.f;
    }
    static float maxCoeff(C3 c) { return std::max(c.r, std::max(c.g, c.b)); }

    C3 li(Sampler& sampler, const Ray& _ray) {
        const Medium* med = medium();
        C3 radiance = c3(0.f), throughput = c3(1.f);
        float eta = 1.f;
        int depth = 0;
        Ray ray = _ray;
        Hit intersection;  // keeps its t when an intersection fails (Bvh.cpp:392-480)
        rays++;
        bool wasIntersected = scene->intersect(ray, intersection, false);
        bool emission = true;
        const int env = scene->envLight;
        while (depth < maxDepth) {
            float t;
            C3 weightMedium;
            if (med && med->sampleDistance(makeRay(ray.o, ray.d, ray.minT, intersection.t), sampler, t, weightMedium)) {
                // :62-157
                throughput = throughput * weightMedium;
                V3 p = ray.o + t * ray.d;
                Scene::LightSample info;
                C3 value = sampleAttenuatedLightDirect(p, sampler, info);
                if (!isZero(value)) {
                    float pfPDF = med->hgEval(info.toLight, -ray.d);
                    float weight = PathTracer::powerHeuristic(info.pdf, pfPDF);
                    radiance += throughput * value * med->hgEval(info.toLight, -ray.d) * weight;
                }
                float pfPdf;
                V3 wo = med->hgSample(-ray.d, sampler.next2D(), pfPdf);
                ray = makeRay(p, wo, 0.f, ray.maxT);
                // evalTransmittanceFromSampledBSDF (:348-423)
                rays++;
                wasIntersected = scene->intersect(ray, intersection, false);
                C3 tr = med->transmittanceC(makeRay(ray.o, ray.d, ray.minT, wasIntersected ? intersection.t : ray.maxT),
                                            sampler);
                if (wasIntersected) {
                    int hl = scene->primLight(intersection.prim);
                    if (hl >= 0) {
                        C3 Le = scene->meshLe(scene->lights[hl], -ray.d, intersection.sh.n);
                        float pdfLight = mediumHitLightPdf(hl, intersection.p, info.point, info.normal);
                        float weight = PathTracer::powerHeuristic(pfPdf, pdfLight);
                        radiance += throughput * Le * tr * weight;
                    }
                }
            } else {
                // :159-326
                if (med) throughput = throughput * weightMedium;  // :160-165
                if (!wasIntersected) {
                    if (emission && env >= 0) radiance += throughput * scene->envLe(scene->lights[env], ray.d);
                    return radiance;
                }
                const Hit hit = intersection;
                if (scene->primMaterial(hit.prim) < 0) {
                    // :187-195: a shape without a BSDF is passed through, depth unchanged
                    ray = makeRay(hit.p, ray.d);
                    rays++;
                    wasIntersected = scene->intersect(ray, intersection, false);
                    continue;
                }
                const Material bsdf = scene->materialAt(hit);
                int hitLight = scene->primLight(hit.prim);
                if (emission && hitLight >= 0)
                    radiance += throughput * scene->meshLe(scene->lights[hitLight], -ray.d, hit.sh.n);
                Scene::LightSample info;
                C3 value = sampleAttenuatedLightDirect(hit.p, sampler, info);
                if (!isZero(value)) {
                    BSDFInfo bi;
                    bi.wi = toLocal(hit.sh, info.toLight);
                    bi.wo = toLocal(hit.sh, -ray.d);
                    float cosTheta = std::fabs(bi.wi.z);
                    float bsdfPDF = bsdfPdf(bsdf, bi);
                    float weight = PathTracer::powerHeuristic(info.pdf, bsdfPDF);
                    radiance += throughput * value * bsdfEval(bsdf, bi) * cosTheta * weight;
                }
                BSDFInfo bi;
                bi.wi = toLocal(hit.sh, -ray.d);
                C3 bsdfValue = bsdfSample(bsdf, bi, sampler.next2D());
                if (isZero(bsdfValue)) break;
                V3 pFrom = hit.p;
                ray = makeRay(hit.p, toWorld(hit.sh, bi.wo));
                rays++;
                wasIntersected = scene->intersect(ray, intersection, false);
                C3 tr = med ? med->transmittanceC(
                                  makeRay(ray.o, ray.d, ray.minT, wasIntersected ? intersection.t : ray.maxT), sampler)
                            : c3(1.f);
                int caught = -1;
                C3 radianceLight = c3(0.f);
                if (wasIntersected) {
                    int hl = scene->primLight(intersection.prim);
                    if (hl >= 0) {
                        caught = hl;
                        radianceLight = scene->meshLe(scene->lights[hl], -ray.d, intersection.sh.n);
                    }
                } else if (env >= 0) {
                    caught = env;
                    radianceLight = scene->envLe(scene->lights[env], ray.d);
                }
                if (caught >= 0) {
                    float pdfLight = (bi.sampledType & BT_DELTA)
                                         ? 0.f
                                         : scene->lightPdf(caught, pFrom, intersection.p, intersection.sh.n, ray.d);
                    float weight = PathTracer::powerHeuristic(bi.pdf, pdfLight);
                    radiance += throughput * radianceLight * bsdfValue * tr * weight;
                    if (caught == env && !med) return radiance;
                }
                throughput = throughput * bsdfValue;
                eta *= bi.relativeEta;
            }
            // :331-343
            if (depth >= minDepth) {
                float stopVal = std::min(0.95f, maxCoeff(throughput) * eta * eta);
                if (sampler.next1D() < stopVal)
                    throughput = throughput / stopVal;
                else
                    return radiance;
            }
            // :345 traces the same ray again; the result (and a failure's kept t) is unchanged
            emission = false;
            depth++;
        }
        return radiance;
    }
};

// ---------------------------------------------------------------- camera (camera/NoriPerspective.cpp:10-118)
struct Camera {
    int W = 0, H = 0;
    M44 sampleToCamera, camToWorld;
    float nearClip = 1e-4f, farClip = 1e4f;
    V3 invOut;
    Ray getRay(float px, float py) const {
        V3 nearP = xformPoint(sampleToCamera, v3(px * invOut.x, py * invOut.y, 0.0f));
        V3 d = normalized(nearP);
        float invZ = 1.0f / d.z;
        Ray r;
        r.o = xformPoint(camToWorld, v3(0.f, 0.f, 0.f));
        V3 wd = xformVector(camToWorld, d);
        r = makeRay(r.o, wd);
        r.minT = nearClip * invZ;
        r.maxT = farClip * invZ;
        return r;
    }
};
// camera/NoriPerspective.cpp:10-37 takes the XML's toWorld as camera-to-world: cam[12] = 1 with
// cam[13..28] (row-major) replaces the lookAt frame (a toWorld such as a lookAt with a mirroring scale)
static void cameraMatrix(Camera& c, const float* cam) {
    if (cam[12] != 1.f) return;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) c.camToWorld.m[i][j] = cam[13 + 4 * i + j];
}
// Transform.cpp:199-227 (fromLookAt) and NoriPerspective.cpp:10-37, in double, rounded to float
static void buildCamera(Camera& cam, const float* origin, const float* target, const float* up, float fov,
                        int W, int H, float nearC, float farC) {
    cam.W = W;
    cam.H = H;
    cam.nearClip = nearC;
    cam.farClip = farC;
    auto nrm = [](double v[3]) {
        double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        v[0] /= l; v[1] /= l; v[2] /= l;
    };
    auto crs = [](const double a[3], const double b[3], double o[3]) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    double dir[3] = {(double)target[0] - origin[0], (double)target[1] - origin[1], (double)target[2] - origin[2]};
    nrm(dir);
    double upn[3] = {up[0], up[1], up[2]};
    nrm(upn);
    double left[3];
    crs(upn, dir, left);
    nrm(left);
    double newUp[3];
    crs(dir, left, newUp);
    nrm(newUp);
    M44 c2w{};
    for (int i = 0; i < 3; i++) {
        c2w.m[i][0] = (float)left[i];
        c2w.m[i][1] = (float)newUp[i];
        c2w.m[i][2] = (float)dir[i];
        c2w.m[i][3] = origin[i];
    }
    c2w.m[3][3] = 1.f;
    cam.camToWorld = c2w;
    double aspect = W / (double)H;
    double recip = 1.0 / ((double)farC - nearC);
    double cot = 1.0 / std::tan((fov * 3.14159265358979323846 / 180.0) / 2.0);
    double P[4][4] = {{cot, 0, 0, 0}, {0, cot, 0, 0}, {0, 0, farC * recip, -(double)nearC * farC * recip}, {0, 0, 1, 0}};
    // S * T * P with S = diag(-0.5, -0.5*aspect, 1), T = translate(-1, -1/aspect, 0)
    double S[4][4] = {{-0.5, 0, 0, 0}, {0, -0.5 * aspect, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    double T[4][4] = {{1, 0, 0, -1.0}, {0, 1, 0, -1.0 / aspect}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    double ST[4][4] = {}, STP[4][4] = {};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 4; k++) ST[i][j] += S[i][k] * T[k][j];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 4; k++) STP[i][j] += ST[i][k] * P[k][j];
    // invert in double
    double m[4][8];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            m[i][j] = STP[i][j];
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
        for (int j = 0; j < 4; j++) cam.sampleToCamera.m[i][j] = (float)m[i][j + 4];
    cam.invOut = v3(1.f / W, 1.f / H, 0.f);
}

// ---------------------------------------------------------------- film (core/Screen.cpp:44-158, box filter)
struct Film {
    int W, H;
    std::vector<double> rgb;   // W*H*3
    std::vector<double> wgt;   // W*H
    std::mutex mtx;
};

// ---- reconstruction filters (filter/ReconstructionFilter.{h,cpp}, BoxFilter, GaussianFilter,
// MitchellNetravaliFilter): getValue tabulated at 32x32 points of [0, rx) x [0, ry), looked up
// by truncation (no interpolation)
struct Filter {
    int type = 0;  // 0 box, 1 gaussian, 2 mitchell
    float rx = 0.5f, ry = 0.5f, p0 = 0.f, p1 = 0.f;
    float lut[32 * 32];
    float gauss1(float v) const {  // GaussianFilter.cpp:40-43
        return std::max(0.f, std::exp(-p0 * v * v) - std::exp(-p0 * p1 * p1));
    }
    float mitchell1(float v) const {  // MitchellNetravaliFilter.cpp:44-59
        const float B = p0, C = p1;
        v = std::fabs(2.f * v);
        if (v > 1.f)
            return ((-B - 6 * C) * v * v * v + (6 * B + 30 * C) * v * v + (-12 * B - 48 * C) * v + 8 * B + 24 * C) *
                   (1.f / 6.f);
        return ((12 - 9 * B - 6 * C) * v * v * v + (-18 + 12 * B + 6 * C) * v * v + (6 - 2 * B)) * (1.f / 6.f);
    }
    float value(float x, float y) const {
        if (type == 1) return gauss1(x) * gauss1(y);
        if (type == 2) return mitchell1(x / rx) * mitchell1(y / ry);
        return 1.f;
    }
    void precompute() {  // ReconstructionFilter.cpp:26-39
        for (int x = 0; x < 32; x++)
            for (int y = 0; y < 32; y++) lut[x + y * 32] = value(x * rx / 32, y * ry / 32);
    }
    float operator()(float x, float y) const {  // ReconstructionFilter.h:53-62
        int ix = std::min((int)((32 * std::fabs(x)) / rx), 31);
        int iy = std::min((int)((32 * std::fabs(y)) / ry), 31);
        return lut[ix + iy * 32];
    }
};

struct RenderParams {
    int spp = 16, samplerType = 1 /*0 random, 1 stratified*/, dimension = 4;
    int minDepth = 3, maxDepth = 1000, threads = 1, block = 32;
    Filter filter;
};

// SampleJob.cpp:28-72 for one 32x32 block, merged into the film (Screen.cpp:136-158)
static void runBlock(const Scene& scene, const Camera& cam, const RenderParams& rp, int bx, int by,
                     Film& film, std::atomic<uint64_t>& rayCount, std::atomic<uint64_t>& nanCount,
                     uint64_t seedOffset) {
    std::unique_ptr<Sampler> sampler;
    if (rp.samplerType == 1)
        sampler.reset(new StratifiedSampler(rp.spp, rp.dimension));
    else {
        sampler.reset(new RandomSampler());
        sampler->number = rp.spp;
    }
    // Scene.cpp:89-90 (seedOffset 0 is the reference schedule; other values give independent renders)
    sampler->seed((uint64_t)bx + (seedOffset << 32), (uint64_t)by);
    PathTracer pt;
    pt.scene = &scene;
    pt.minDepth = rp.minDepth;
    pt.maxDepth = rp.maxDepth;
    VolPathTracer vpt;
    vpt.scene = &scene;
    vpt.minDepth = rp.minDepth;
    vpt.maxDepth = rp.maxDepth;
    int endX = std::min(cam.W, bx + rp.block), endY = std::min(cam.H, by + rp.block);
    int sx = endX - bx, sy = endY - by;
    const Filter& flt = rp.filter;
    int ovx = (int)std::ceil(flt.rx - 0.5f), ovy = (int)std::ceil(flt.ry - 0.5f);  // Screen.cpp:101-111
    int pw = sx + 2 * ovx, ph = sy + 2 * ovy;
    std::vector<double> sub((size_t)pw * ph * 3, 0.0), subw((size_t)pw * ph, 0.0);
    uint64_t nans = 0;
    for (int y = by; y < endY; y++)
        for (int x = bx; x < endX; x++) {
            sampler->startPixel();
            for (int i = 0; i < sampler->number; i++) {
                V2 s = sampler->next2D();
                float xx = (float)x + s.x;
                float yy = (float)y + s.y;
                Ray ray = cam.getRay(xx, yy);
                C3 col = scene.integrator == 1 ? vpt.li(*sampler, ray) : pt.li(*sampler, ray);
                // Screen.cpp:44-97
                if (isNan(col)) {
                    nans++;
                } else {
                    float fx = xx - 0.5f - (bx - ovx);
                    float fy = yy - 0.5f - (by - ovy);
                    int x0 = std::max(0, (int)std::ceil(fx - flt.rx));
                    int x1 = std::min(pw - 1, (int)std::floor(fx + flt.rx));
                    int y0 = std::max(0, (int)std::ceil(fy - flt.ry));
                    int y1 = std::min(ph - 1, (int)std::floor(fy + flt.ry));
                    for (int px = x0; px <= x1; px++)
                        for (int py = y0; py <= y1; py++) {
                            const float wgt = flt(px - fx, py - fy);
                            size_t k = (size_t)py * pw + px;
                            sub[3 * k] += col.r * wgt;
                            sub[3 * k + 1] += col.g * wgt;
                            sub[3 * k + 2] += col.b * wgt;
                            subw[k] += wgt;
                        }
                }
                sampler->advance();
            }
        }
    {
        std::lock_guard<std::mutex> lock(film.mtx);
        for (int py = 0; py < ph; py++)
            for (int px = 0; px < pw; px++) {
                int gx = bx - ovx + px, gy = by - ovy + py;
                if (gx < 0 || gy < 0 || gx >= film.W || gy >= film.H) continue;
                size_t k = (size_t)py * pw + px, g = (size_t)gy * film.W + gx;
                film.rgb[3 * g] += sub[3 * k];
                film.rgb[3 * g + 1] += sub[3 * k + 1];
                film.rgb[3 * g + 2] += sub[3 * k + 2];
                film.wgt[g] += subw[k];
            }
    }
    rayCount += pt.rays + vpt.rays;
    nanCount += nans;
}

}  // namespace orc

// ==================================================================== C API (ctypes)
using namespace orc;
extern "C" {

void* orc_scene_new() { return new Scene(); }
void orc_scene_free(void* s) { delete (Scene*)s; }

// params[20]: type, albedo/reflectance rgb (1-3), alpha (4), eta rgb (5-7), k rgb (8-10),
// etaInt (11), etaExt (12), distribution (13: 0 beckmann, 2 ggx), transmitted rgb (14-16)
int orc_add_material(void* sp, const float* p) {
    Scene* s = (Scene*)sp;
    Material m;
    m.type = (int)p[0];
    if (m.type == MAT_DIFFUSE)
        m.albedo = c3(p[1], p[2], p[3]);
    else
        m.reflectance = c3(p[1], p[2], p[3]);
    m.alphaRaw = p[4];
    m.alpha = average(c3(p[4]));  // ConstantTexture(real).eval(uv).average()
    m.eta = c3(p[5], p[6], p[7]);
    m.k = c3(p[8], p[9], p[10]);
    m.etaInt = p[11];
    m.etaExt = p[12];
    m.dist = (int)p[13];
    m.transmitted = c3(p[14], p[15], p[16]);
    for (int k = 0; k < 3; k++) m.tex[k] = (int)p[17 + k] - 1;  // 1-based, 0 = none
    s->mats.push_back(m);
    return (int)s->mats.size() - 1;
}

int orc_add_mesh(void* sp, const float* v, int nv, const float* n, const float* uv, const int* idx, int ntri,
                 const int* triMat, int mat, const float* emit) {
    Scene* s = (Scene*)sp;
    Mesh m;
    m.v.resize(nv);
    for (int i = 0; i < nv; i++) m.v[i] = v3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    if (n) {
        m.n.resize(nv);
        for (int i = 0; i < nv; i++) m.n[i] = v3(n[3 * i], n[3 * i + 1], n[3 * i + 2]);
    }
    if (uv) {
        m.uv.resize(nv);
        for (int i = 0; i < nv; i++) m.uv[i] = V2{uv[2 * i], uv[2 * i + 1]};
    }
    m.idx.assign(idx, idx + 3 * ntri);
    m.triMat.resize(ntri);
    for (int i = 0; i < ntri; i++) m.triMat[i] = triMat ? triMat[i] : mat;
    if (emit) {
        Light L;
        L.type = LIGHT_MESH;
        L.mesh = (int)s->meshes.size();
        L.radiance = c3(emit[0], emit[1], emit[2]);
        s->lights.push_back(L);
        m.light = (int)s->lights.size() - 1;
    }
    s->meshes.push_back(std::move(m));
    return (int)s->meshes.size() - 1;
}

// toWorld: 16 floats row-major
int orc_add_sphere(void* sp, const float* toWorld, float radius, int mat) {
    Scene* s = (Scene*)sp;
    Sphere sph;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) sph.toWorld.m[i][j] = toWorld[4 * i + j];
    sph.inv = inverse44(sph.toWorld);
    sph.radius = radius;
    sph.mat = mat;
    s->spheres.push_back(sph);
    return (int)s->spheres.size() - 1;
}

// EnvironmentLight.cpp:8-44 (identity toWorld, no offsets, lat-long)
// ImageTexture::eval of texture ti at n uv pairs -> n x 3
void orc_texture_eval(void* sp, int ti, const float* uv, int n, float* out) {
    const Texture& t = ((Scene*)sp)->textures[ti];
    for (int i = 0; i < n; i++) {
        C3 c = t.eval(V2{uv[2 * i], uv[2 * i + 1]});
        out[3 * i] = c.r; out[3 * i + 1] = c.g; out[3 * i + 2] = c.b;
    }
}
// The material at the closest hit of n rays (x 8) with its textures evaluated: out n x 8 =
// (albedo or reflectance rgb, alpha, transmitted rgb, texture u) ; v in uvout (n x 2 = u, v)
void orc_material_at(void* sp, const float* rays, int n, float* out, float* uvout) {
    const Scene* s = (const Scene*)sp;
    for (int i = 0; i < n; i++) {
        const float* r = rays + 8 * i;
        Hit hit;
        float* o = out + 8 * i;
        for (int k = 0; k < 8; k++) o[k] = 0.f;
        uvout[2 * i] = uvout[2 * i + 1] = 0.f;
        if (!s->intersect(makeRay(v3(r[0], r[1], r[2]), v3(r[3], r[4], r[5]), r[6], r[7]), hit, false)) continue;
        Material m = s->materialAt(hit);
        C3 a = m.type == MAT_DIFFUSE ? m.albedo : m.reflectance;
        o[0] = a.r; o[1] = a.g; o[2] = a.b; o[3] = m.alpha;
        o[4] = m.transmitted.r; o[5] = m.transmitted.g; o[6] = m.transmitted.b;
        V2 uv = s->textureUv(hit);
        uvout[2 * i] = uv.x;
        uvout[2 * i + 1] = uv.y;
    }
}

// Image texture: H x W x 3 texels, uv scale, inversion flags, value scale; returns its index
int orc_add_texture(void* sp, const float* rgb, int W, int H, float uvScale, int invertX, int invertY,
                    float valueScale) {
    Scene* s = (Scene*)sp;
    Texture t;
    t.W = W;
    t.H = H;
    t.scale = std::fabs(uvScale);
    t.invertX = invertX != 0;
    t.invertY = invertY != 0;
    t.valueScale = valueScale;
    t.px.resize((size_t)W * H);
    for (size_t i = 0; i < t.px.size(); i++) t.px[i] = c3(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
    s->textures.push_back(t);
    return (int)s->textures.size() - 1;
}

int orc_set_env(void* sp, const float* rgb, int W, int H, float scale) {
    Scene* s = (Scene*)sp;
    Light L;
    L.type = LIGHT_ENV;
    L.W = W;
    L.H = H;
    L.factor = scale;
    L.map.resize((size_t)W * H);
    for (size_t i = 0; i < (size_t)W * H; i++) L.map[i] = c3(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
    for (int i = 0; i < H; i++) {
        float weight = std::sin(PI * (i + 0.5f) / H);
        L.rows.push_back(CDF());
        for (int j = 0; j < W; j++) L.rows[i].add(luminance(L.getValue(j, i)));
        L.weight.push_back(weight);
        L.marginal.add(L.rows[i].sum * weight);
        L.rows[i].normalize();
    }
    L.marginal.normalize();
    s->lights.push_back(L);
    s->envLight = (int)s->lights.size() - 1;
    return s->envLight;
}

int orc_finalize(void* sp, int strategy) {
    Scene* s = (Scene*)sp;
    s->strategy = strategy;
    s->finalize();
    return s->ntri + (int)s->spheres.size();
}

// rays: n x 8 (ox oy oz dx dy dz minT maxT). brute: 1 = brute-force loop, 0 = two-level BVH.
// Outputs: prim (-1 on miss), t, uv (b1 b2 for triangles), p (hit point after diff-geom), n (shading normal)
void orc_trace(void* sp, const float* rays, int n, int anyhit, int brute, int* prim, float* t, float* uv,
               float* pout, float* nout) {
    Scene* s = (Scene*)sp;
    for (int i = 0; i < n; i++) {
        const float* r = rays + 8 * i;
        Ray ray = makeRay(v3(r[0], r[1], r[2]), v3(r[3], r[4], r[5]), r[6], r[7]);
        Hit h;
        bool hit = brute ? s->intersectBrute(ray, h, anyhit != 0) : s->intersect(ray, h, anyhit != 0);
        prim[i] = hit ? h.prim : -1;
        if (t) t[i] = hit ? h.t : 0.f;
        if (uv) {
            uv[2 * i] = hit ? h.uv.x : 0.f;
            uv[2 * i + 1] = hit ? h.uv.y : 0.f;
        }
        if (pout) {
            pout[3 * i] = hit ? h.p.x : 0.f;
            pout[3 * i + 1] = hit ? h.p.y : 0.f;
            pout[3 * i + 2] = hit ? h.p.z : 0.f;
        }
        if (nout) {
            nout[3 * i] = hit ? h.sh.n.x : 0.f;
            nout[3 * i + 1] = hit ? h.sh.n.y : 0.f;
            nout[3 * i + 2] = hit ? h.sh.n.z : 0.f;
        }
    }
}

void orc_bsdf_eval(void* sp, int mat, const float* wi, const float* wo, int n, float* f, float* pdf) {
    Scene* s = (Scene*)sp;
    const Material& m = s->mats[mat];
    for (int i = 0; i < n; i++) {
        BSDFInfo bi;
        bi.wi = v3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
        bi.wo = v3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
        C3 c = bsdfEval(m, bi);
        f[3 * i] = c.r;
        f[3 * i + 1] = c.g;
        f[3 * i + 2] = c.b;
        pdf[i] = bsdfPdf(m, bi);
    }
}

void orc_bsdf_sample(void* sp, int mat, const float* wi, const float* u, int n, float* wo, float* w, float* pdf,
                     float* relEta, int* type) {
    Scene* s = (Scene*)sp;
    const Material& m = s->mats[mat];
    for (int i = 0; i < n; i++) {
        BSDFInfo bi;
        bi.wi = v3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
        bi.wo = v3(0.f, 0.f, 0.f);
        C3 c = bsdfSample(m, bi, V2{u[2 * i], u[2 * i + 1]});
        wo[3 * i] = bi.wo.x;
        wo[3 * i + 1] = bi.wo.y;
        wo[3 * i + 2] = bi.wo.z;
        w[3 * i] = c.r;
        w[3 * i + 1] = c.g;
        w[3 * i + 2] = c.b;
        pdf[i] = bi.pdf;
        relEta[i] = bi.relativeEta;
        type[i] = (int)bi.sampledType;
    }
}

// microfacet pieces for closed-form tests
void orc_mf(int type, const float* wh, const float* alpha, int n, float* D, float* G1, const float* v) {
    for (int i = 0; i < n; i++) {
        V3 h = v3(wh[3 * i], wh[3 * i + 1], wh[3 * i + 2]);
        D[i] = mfD(type, h, alpha[i]);
        if (G1) G1[i] = mfG1(type, v3(v[3 * i], v[3 * i + 1], v[3 * i + 2]), h, alpha[i]);
    }
}
void orc_mf_sample(int type, const float* u, const float* alpha, int n, float* wh) {
    for (int i = 0; i < n; i++) {
        V3 h = mfSampleNormal(type, V2{u[2 * i], u[2 * i + 1]}, alpha[i]);
        wh[3 * i] = h.x;
        wh[3 * i + 1] = h.y;
        wh[3 * i + 2] = h.z;
    }
}
void orc_fresnel_conductor(const float* eta, const float* k, const float* cosI, int n, float* out) {
    for (int i = 0; i < n; i++) {
        C3 r = fresnelConductor(c3(eta[0], eta[1], eta[2]), c3(k[0], k[1], k[2]), cosI[i]);
        out[3 * i] = r.r;
        out[3 * i + 1] = r.g;
        out[3 * i + 2] = r.b;
    }
}
void orc_fresnel_dielectric(float etaExt, float etaInt, const float* cosI, int n, float* fr, float* cosT,
                            float* relEta) {
    for (int i = 0; i < n; i++) {
        float a, b, c, d;
        fr[i] = estimateDielectric(etaExt, etaInt, cosI[i], a, b, c, d);
        cosT[i] = d;
        relEta[i] = c;
    }
}
// warps: which 0 concentric disk (2 out), 1 cosine hemisphere (3), 2 square-to-triangle (3)
void orc_warp(int which, const float* u, int n, float* out) {
    for (int i = 0; i < n; i++) {
        V2 s{u[2 * i], u[2 * i + 1]};
        if (which == 0) {
            V2 d = squareToConcentricDisk(s);
            out[2 * i] = d.x;
            out[2 * i + 1] = d.y;
        } else {
            V3 r = which == 1 ? squareToCosineWeightedHemisphere(s) : squareToTriangle(s);
            out[3 * i] = r.x;
            out[3 * i + 1] = r.y;
            out[3 * i + 2] = r.z;
        }
    }
}
void orc_frame(const float* nrm, int n, float* s, float* t) {
    for (int i = 0; i < n; i++) {
        Frame f = makeFrame(v3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]));
        s[3 * i] = f.s.x; s[3 * i + 1] = f.s.y; s[3 * i + 2] = f.s.z;
        t[3 * i] = f.t.x; t[3 * i + 1] = f.t.y; t[3 * i + 2] = f.t.z;
    }
}

void orc_pcg32(uint64_t initstate, uint64_t initseq, int n, uint32_t* out) {
    Rng r;
    r.seed(initstate, initseq);
    for (int i = 0; i < n; i++) out[i] = r.nextUInt();
}
void orc_pcg32_real(uint64_t initstate, uint64_t initseq, int n, float* out) {
    Rng r;
    r.seed(initstate, initseq);
    for (int i = 0; i < n; i++) out[i] = r.nextReal();
}
// Stratified sampler dump for npix consecutive pixels of one block seeded (bx, by):
// out2d: npix * spp * ndraw2 * 2, out1d: npix * spp * ndraw1
int orc_stratified(int spp, int dim, uint64_t bx, uint64_t by, int npix, int ndraw2, int ndraw1, float* out2d,
                   float* out1d) {
    StratifiedSampler s(spp, dim);
    s.seed(bx, by);
    for (int p = 0; p < npix; p++) {
        s.startPixel();
        for (int i = 0; i < s.number; i++) {
            for (int k = 0; k < ndraw2; k++) {
                V2 v = s.next2D();
                size_t o = (((size_t)p * s.number + i) * ndraw2 + k) * 2;
                out2d[o] = v.x;
                out2d[o + 1] = v.y;
            }
            for (int k = 0; k < ndraw1; k++) out1d[((size_t)p * s.number + i) * ndraw1 + k] = s.next1D();
            s.advance();
        }
    }
    return s.number;
}
// CDF tap: builds CDF from w, samples u with sampleAndReuse: idx, remapped u, pdf
void orc_cdf(const float* w, int nw, const float* u, int n, int* idx, float* ur, float* pdf) {
    CDF c;
    for (int i = 0; i < nw; i++) c.add(w[i]);
    c.normalize();
    for (int i = 0; i < n; i++) {
        float s = u[i], p;
        idx[i] = c.sampleAndReuse(s, p);
        ur[i] = s;
        pdf[i] = p;
    }
}

// camera: origin3 target3 up3 fov near far; out rays n x 8 for pixel positions (px, py)
void orc_camera_rays(const float* cam, int W, int H, const float* pxy, int n, float* rays, float* mats) {
    Camera c;
    buildCamera(c, cam, cam + 3, cam + 6, cam[9], W, H, cam[10], cam[11]);
    cameraMatrix(c, cam);
    for (int i = 0; i < n; i++) {
        Ray r = c.getRay(pxy[2 * i], pxy[2 * i + 1]);
        float* o = rays + 8 * i;
        o[0] = r.o.x; o[1] = r.o.y; o[2] = r.o.z;
        o[3] = r.d.x; o[4] = r.d.y; o[5] = r.d.z;
        o[6] = r.minT; o[7] = r.maxT;
    }
    if (mats) {
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) {
                mats[4 * i + j] = c.sampleToCamera.m[i][j];
                mats[16 + 4 * i + j] = c.camToWorld.m[i][j];
            }
    }
}

// light taps: sample light li from points p with u; outputs toLight(3), dist, pdf(light-only), intensity(3), point(3), normal(3)
void orc_light_sample(void* sp, int li, const float* p, const float* u, int n, float* out) {
    Scene* s = (Scene*)sp;
    for (int i = 0; i < n; i++) {
        Scene::LightSample ls = s->lightSample(li, v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]), V2{u[2 * i], u[2 * i + 1]});
        float* o = out + 14 * i;
        o[0] = ls.toLight.x; o[1] = ls.toLight.y; o[2] = ls.toLight.z;
        o[3] = ls.distance; o[4] = ls.pdf;
        o[5] = ls.intensity.r; o[6] = ls.intensity.g; o[7] = ls.intensity.b;
        o[8] = ls.point.x; o[9] = ls.point.y; o[10] = ls.point.z;
        o[11] = ls.normal.x; o[12] = ls.normal.y; o[13] = ls.normal.z;
    }
}
// light pdf: pFrom(3), point(3), normal(3), dir(3) per query
void orc_light_pdf(void* sp, int li, const float* q, int n, float* pdf) {
    Scene* s = (Scene*)sp;
    for (int i = 0; i < n; i++) {
        const float* a = q + 12 * i;
        pdf[i] = s->lightPdf(li, v3(a[0], a[1], a[2]), v3(a[3], a[4], a[5]), v3(a[6], a[7], a[8]), v3(a[9], a[10], a[11]));
    }
}
void orc_env_le(void* sp, const float* dirs, int n, float* out) {
    Scene* s = (Scene*)sp;
    const Light& L = s->lights[s->envLight];
    for (int i = 0; i < n; i++) {
        C3 c = s->envLe(L, v3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
        out[3 * i] = c.r; out[3 * i + 1] = c.g; out[3 * i + 2] = c.b;
    }
}
// light selection CDF (cdf values, size nlights+1)
void orc_light_cdf(void* sp, float* out) {
    Scene* s = (Scene*)sp;
    for (size_t i = 0; i < s->lightWeights.c.size(); i++) out[i] = s->lightWeights.c[i];
}

// Single path with a caller-supplied sample stream (for estimator tests): u2 pairs and u1 values are
// consumed in order by a sampler that replays them. Returns L.
struct ReplaySampler : Sampler {
    const float* u2;
    const float* u1;
    int n2, n1, i2 = 0, i1 = 0;
    V2 next2D() override {
        V2 r{0.5f, 0.5f};
        if (i2 < n2) r = V2{u2[2 * i2], u2[2 * i2 + 1]};
        i2++;
        return r;
    }
    float next1D() override {
        float r = 0.5f;
        if (i1 < n1) r = u1[i1];
        i1++;
        return r;
    }
    void seed(uint64_t, uint64_t) override {}
};
void orc_li(void* sp, const float* ray, const float* u2, int n2, const float* u1, int n1, int minDepth,
            int maxDepth, float* L) {
    Scene* s = (Scene*)sp;
    ReplaySampler rs;
    rs.u2 = u2; rs.u1 = u1; rs.n2 = n2; rs.n1 = n1;
    PathTracer pt;
    pt.scene = s;
    pt.minDepth = minDepth;
    pt.maxDepth = maxDepth;
    Ray r = makeRay(v3(ray[0], ray[1], ray[2]), v3(ray[3], ray[4], ray[5]), ray[6], ray[7]);
    C3 c = pt.li(rs, r);
    L[0] = c.r; L[1] = c.g; L[2] = c.b;
}

// Full render with the reference schedule (Scene.cpp:83-101, JobManager.cpp:140-232): 32x32 blocks
// seeded (blockX, blockY), threads popping jobs LIFO. Accumulates in double.
// cam: origin3 target3 up3 fov near far. stats: [rays, nan samples]. Optional block range restricts work
// to blocks [b0, b1) in creation order (for bounded CPU baselines).
// Camera medium: density grid (x fastest), albedo rgb, HG g, scale (<= 0: 1).
int orc_set_medium(void* sp, const float* data, int rx, int ry, int rz, const float* albedo, float g, float scale) {
    Scene* s = (Scene*)sp;
    Medium& m = s->medium;
    m.res[0] = rx; m.res[1] = ry; m.res[2] = rz;
    m.data.assign(data, data + (size_t)rx * ry * rz);
    m.albedo = c3(albedo[0], albedo[1], albedo[2]);
    m.g = g;
    m.scale = scale > 0.f ? scale : 1.f;
    m.finalize();
    return 0;
}
// Homogeneous camera medium: sigmaA, sigmaS (rgb), HG g, scale (<= 0: 1)
int orc_set_medium_homogeneous(void* sp, const float* sigA, const float* sigS, float g, float scale) {
    Medium& m = ((Scene*)sp)->medium;
    m.homogeneous = true;
    float sc = scale > 0.f ? scale : 1.f;
    m.sigA = c3(sigA[0], sigA[1], sigA[2]);
    m.sigS = c3(sigS[0], sigS[1], sigS[2]);
    if (sc != 1.f) {  // Homogeneous.cpp:16-20
        m.sigA = m.sigA * sc;
        m.sigS = m.sigS * sc;
    }
    m.sigT = m.sigA + m.sigS;
    m.g = g;
    return 0;
}
// 0 PathTracingMIS, 1 VolPathTracing
void orc_set_integrator(void* sp, int type) { ((Scene*)sp)->integrator = type; }
// GridVolume::lookupSigmaT x scale at n points
void orc_medium_sigma(void* sp, const float* pts, int n, float* out) {
    const Medium& m = ((Scene*)sp)->medium;
    for (int i = 0; i < n; i++) out[i] = m.lookupSigmaT(v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
}
// Trackers on n rays (ox oy oz dx dy dz minT maxT) with K given 1D draws per ray; out n x 6.
// op 0 distance sampling: (scattered, t, draws used, weight rgb);
// op 1 transmittance: (tr (red channel), draws used, 0, tr rgb)
void orc_medium_track(void* sp, int op, const float* rays, int n, const float* u, int K, float* out) {
    const Medium& m = ((Scene*)sp)->medium;
    for (int i = 0; i < n; i++) {
        const float* r = rays + 8 * i;
        Ray ray = makeRay(v3(r[0], r[1], r[2]), v3(r[3], r[4], r[5]), r[6], r[7]);
        ReplaySampler smp;
        smp.u1 = u + (size_t)i * K;
        smp.n1 = K;
        smp.u2 = nullptr;
        smp.n2 = 0;
        float* o = out + 6 * i;
        if (op == 0) {
            float t;
            C3 w;
            bool sc = m.sampleDistance(ray, smp, t, w);
            o[0] = sc ? 1.f : 0.f;
            o[1] = t;
            o[2] = (float)smp.i1;
            o[3] = w.r; o[4] = w.g; o[5] = w.b;
        } else {
            C3 tr = m.transmittanceC(ray, smp);
            o[0] = tr.r;
            o[1] = (float)smp.i1;
            o[2] = 0.f;
            o[3] = tr.r; o[4] = tr.g; o[5] = tr.b;
        }
    }
}
// Henyey-Greenstein: op 0 eval(wi, wo) -> out[i]; op 1 sample(wi, u) -> out 4 per item (wo, pdf)
void orc_hg(float g, int op, const float* wi, const float* x, int n, float* out) {
    Medium m;
    m.g = g;
    for (int i = 0; i < n; i++) {
        V3 a = v3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
        if (op == 0) {
            out[i] = m.hgEval(a, v3(x[3 * i], x[3 * i + 1], x[3 * i + 2]));
        } else {
            float pdf;
            V3 wo = m.hgSample(a, V2{x[2 * i], x[2 * i + 1]}, pdf);
            out[4 * i] = wo.x; out[4 * i + 1] = wo.y; out[4 * i + 2] = wo.z; out[4 * i + 3] = pdf;
        }
    }
}

// core/main.cpp:185-290 tone mappers after Pixel::applyColor (core/Screen.cpp:8-15):
// mode 0 clamp, 1 gamma (sRGB, exposure), 2 global Reinhard (sRGB); out w*h*3 bytes
void orc_tonemap(const float* rgbw, int w, int h, int mode, float exposure, unsigned char* out) {
    const int n = w * h;
    auto color = [&](int i) {
        C3 c = c3(rgbw[4 * i], rgbw[4 * i + 1], rgbw[4 * i + 2]);
        float wt = rgbw[4 * i + 3];
        if (wt != 0.f) c = c / wt;
        return c;
    };
    auto toSRGB = [](float v) {  // Color.h:450-461
        if (v < 0.0031308) return 12.92f * v;
        return 1.055f * std::pow(v, 0.41666f) - 0.055f;
    };
    auto byte = [](float v) {  // thresholding(v * 255, 0, 255) as sf::Uint8 (NaN: 0)
        float x = v * 255;
        float c = x < 0.f ? 0.f : (x > 255.f ? 255.f : x);
        return (unsigned char)(c == c ? (int)c : 0);
    };
    float expo = (float)std::pow(2., (exposure - 0.5) * 20);  // main.cpp:193
    float maxLum = std::numeric_limits<float>::min();
    if (mode == 2)
        for (int i = 0; i < n; i++) maxLum = std::max(maxLum, luminance(color(i)));
    for (int i = 0; i < n; i++) {
        C3 c = color(i);
        float r, g, b;
        if (mode == 0) {
            r = c.r; g = c.g; b = c.b;
        } else if (mode == 1) {
            C3 e = c * expo;
            r = toSRGB(e.r); g = toSRGB(e.g); b = toSRGB(e.b);
        } else {
            float oldLum = luminance(c);
            float numerator = oldLum * (1.0 + (oldLum / (maxLum * maxLum)));
            float lum = numerator / (1.0f + oldLum);
            C3 f = c * lum / oldLum;
            r = toSRGB(f.r); g = toSRGB(f.g); b = toSRGB(f.b);
        }
        out[3 * i] = byte(r);
        out[3 * i + 1] = byte(g);
        out[3 * i + 2] = byte(b);
    }
}

// One table lookup of a filter (type, radius x, radius y, param 0, param 1).
float orc_filter_value(const float* filter, float x, float y) {
    Filter f;
    f.type = (int)filter[0];
    f.rx = filter[1];
    f.ry = filter[2];
    f.p0 = filter[3];
    f.p1 = filter[4];
    f.precompute();
    return f(x, y);
}

// filter: type, radius x, radius y, param 0, param 1 (NULL: box of radius 0.5)
int orc_render(void* sp, const float* cam, int W, int H, int spp, int samplerType, int minDepth, int maxDepth,
               int threads, int b0, int b1, int bstride, uint64_t seedOffset, double* rgb, double* wgt,
               uint64_t* stats, const float* filter) {
    Scene* s = (Scene*)sp;
    Camera c;
    buildCamera(c, cam, cam + 3, cam + 6, cam[9], W, H, cam[10], cam[11]);
    cameraMatrix(c, cam);
    RenderParams rp;
    rp.spp = spp;
    rp.samplerType = samplerType;
    rp.minDepth = minDepth;
    rp.maxDepth = maxDepth;
    rp.threads = std::max(1, threads);
    if (filter) {
        rp.filter.type = (int)filter[0];
        rp.filter.rx = filter[1];
        rp.filter.ry = filter[2];
        rp.filter.p0 = filter[3];
        rp.filter.p1 = filter[4];
    }
    rp.filter.precompute();
    Film film;
    film.W = W;
    film.H = H;
    film.rgb.assign((size_t)W * H * 3, 0.0);
    film.wgt.assign((size_t)W * H, 0.0);
    std::vector<std::pair<int, int>> jobs;
    for (int y = 0; y < H; y += rp.block)
        for (int x = 0; x < W; x += rp.block) jobs.push_back({x, y});
    if (b1 <= 0 || b1 > (int)jobs.size()) b1 = (int)jobs.size();
    if (b0 < 0) b0 = 0;
    if (bstride < 1) bstride = 1;
    std::vector<std::pair<int, int>> todo;
    for (int b = b0; b < b1; b += bstride) todo.push_back(jobs[b]);
    int ntodo = (int)todo.size();
    std::mutex jm;
    std::atomic<uint64_t> rays{0}, nans{0};
    auto worker = [&]() {
        for (;;) {
            std::pair<int, int> j;
            {
                std::lock_guard<std::mutex> lk(jm);
                if (todo.empty()) return;
                j = todo.back();
                todo.pop_back();
            }
            runBlock(*s, c, rp, j.first, j.second, film, rays, nans, seedOffset);
        }
    };
    std::vector<std::thread> th;
    for (int i = 0; i < rp.threads; i++) th.emplace_back(worker);
    for (auto& t : th) t.join();
    std::memcpy(rgb, film.rgb.data(), sizeof(double) * film.rgb.size());
    std::memcpy(wgt, film.wgt.data(), sizeof(double) * film.wgt.size());
    if (stats) {
        stats[0] = rays.load();
        stats[1] = nans.load();
    }
    return ntodo;
}

// BVH statistics (for tests of structure invariants): returns node count of scene BVH; depth via out
int orc_bvh_info(void* sp, int* maxDepthOut) {
    Scene* s = (Scene*)sp;
    int md = 0;
    for (auto& b : s->meshBvh) {
        // depth by walking
        std::vector<std::pair<int, int>> st{{0, 1}};
        while (!st.empty()) {
            auto [n, d] = st.back();
            st.pop_back();
            md = std::max(md, d);
            if (b.nodes.empty()) break;
            if (b.nodes[n].right != -1) {
                st.push_back({n + 1, d + 1});
                st.push_back({b.nodes[n].right, d + 1});
            }
        }
    }
    if (maxDepthOut) *maxDepthOut = md;
    return (int)s->sceneBvh.nodes.size();
}

}  // extern "C"
