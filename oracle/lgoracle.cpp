// LightGraph CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// An FP64 restatement of the reference's multiple-scattering path
// (asiroisvigneux/LightGraph, src/*.h) with the OpenVDB pieces it relies on
// restated from OpenVDB's public API (Ray, VolumeHDDA, BoxSampler,
// VolumeRayIntersector).  It is the checker for liblightgraph: only the tests,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and never as
// the thing measured or shipped.
//
// Substitutions against the reference (each is also made by the device path):
//   - random streams are Philox4x32-10 counter streams instead of std::mt19937
//     engines (RandGen.h), keyed so that no stream is shared between threads;
//     in the render, the primary ray of (pixel, AA sample) owns one stream
//     (camera jitter, primary steps), and each shadow ray batch of (pixel, AA
//     sample, shading point ordinal, light) owns a step stream and an env
//     stream, so shadow rays draw nothing from the primary ray's stream;
//   - points come from the density-weighted sampler instead of dart throwing
//     (LightGraph.h:206-274); the dart thrower is kept here as lgo_dart_throw
//     for the reference-mode CPU baseline;
//   - env-map CDFs are FP64 (Sampling.h:81-93 uses float);
//   - samples are taken at the index-space position of the ray, skipping the
//     index->world->index round trip of getWorldPos + wsSample (identical for
//     unit voxels at the origin).
//
// Parity status of each function is listed in the test modules that pin it;
// functions whose OpenVDB-defined semantics no test here can check against
// OpenVDB itself (HDDA span ends at node level, GridTransformer topology,
// computeVoxelSize) are marked "parity unpinned" where they are defined.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <random>
#include <thread>
#include <vector>

#include "lightgraph.h"

namespace {

// ------------------------------------------------------------------ Vec3 ----
struct Vec3 {
    double v[3];
    Vec3() : v{0, 0, 0} {}
    explicit Vec3(double s) : v{s, s, s} {}
    Vec3(double a, double b, double c) : v{a, b, c} {}
    double &operator[](int i) { return v[i]; }
    double operator[](int i) const { return v[i]; }
    Vec3 operator+(const Vec3 &o) const { return Vec3(v[0] + o[0], v[1] + o[1], v[2] + o[2]); }
    Vec3 operator-(const Vec3 &o) const { return Vec3(v[0] - o[0], v[1] - o[1], v[2] - o[2]); }
    Vec3 operator*(const Vec3 &o) const { return Vec3(v[0] * o[0], v[1] * o[1], v[2] * o[2]); }
    Vec3 operator*(double s) const { return Vec3(v[0] * s, v[1] * s, v[2] * s); }
    Vec3 operator/(double s) const { return Vec3(v[0] / s, v[1] / s, v[2] / s); }
    Vec3 operator/(const Vec3 &o) const { return Vec3(v[0] / o[0], v[1] / o[1], v[2] / o[2]); }
    Vec3 operator-() const { return Vec3(-v[0], -v[1], -v[2]); }
    Vec3 &operator+=(const Vec3 &o) {
        for (int i = 0; i < 3; ++i) v[i] += o[i];
        return *this;
    }
    Vec3 &operator*=(const Vec3 &o) {
        for (int i = 0; i < 3; ++i) v[i] *= o[i];
        return *this;
    }
    Vec3 &operator/=(const Vec3 &o) {
        for (int i = 0; i < 3; ++i) v[i] /= o[i];
        return *this;
    }
    double lengthSqr() const { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }
    double length() const { return std::sqrt(lengthSqr()); }
    double dot(const Vec3 &o) const { return v[0] * o[0] + v[1] * o[1] + v[2] * o[2]; }
    double sum() const { return v[0] + v[1] + v[2]; }
};

Vec3 expv(const Vec3 &a) { return Vec3(std::exp(a[0]), std::exp(a[1]), std::exp(a[2])); }
Vec3 maxComponent(const Vec3 &a, const Vec3 &b) {
    return Vec3(std::max(a[0], b[0]), std::max(a[1], b[1]), std::max(a[2], b[2]));
}

// ---------------------------------------------------------------- Philox ----
void philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// Sequential counter stream: draw d is word d&3 of block philox((a, b, d>>2, dom)).
struct CounterStream {
    uint32_t a, b, dom, key[2];
    uint64_t d = 0;
    uint32_t cache[4];
    CounterStream(uint32_t a_, uint32_t b_, uint32_t dom_, uint64_t seed) : a(a_), b(b_), dom(dom_) {
        key[0] = (uint32_t)seed;
        key[1] = (uint32_t)(seed >> 32);
    }
    float operator()() {
        if ((d & 3) == 0) {
            const uint32_t ctr[4] = {a, b, (uint32_t)(d >> 2), dom};
            philox4x32(ctr, key, cache);
        }
        return u01(cache[d++ & 3]);
    }
};

// ------------------------------------------------------------------ grid ----
// A dense float grid (x fastest) with the topology of an OpenVDB FloatGrid whose
// leaf nodes are the 8^3 bricks holding a voxel != 0.
struct Grid {
    int n[3];
    const float *v;
    double voxel, inv_voxel, origin[3];
    int nb[3];
    std::vector<uint8_t> leaf;
    int bmin[3], bmax[3];
    bool empty;

    Grid(const lg_grid_desc &d, const float *data) : v(data) {
        n[0] = d.nx;
        n[1] = d.ny;
        n[2] = d.nz;
        voxel = d.voxel_size;
        inv_voxel = 1.0 / d.voxel_size;
        for (int c = 0; c < 3; ++c) {
            origin[c] = d.origin[c];
            nb[c] = (n[c] + 7) / 8;
            bmin[c] = 1 << 30;
            bmax[c] = -(1 << 30);
        }
        leaf.assign((size_t)nb[0] * nb[1] * nb[2], 0);
        for (int z = 0; z < n[2]; ++z)
            for (int y = 0; y < n[1]; ++y)
                for (int x = 0; x < n[0]; ++x)
                    if (v[((size_t)z * n[1] + y) * n[0] + x] != 0.0f)
                        leaf[((size_t)(z >> 3) * nb[1] + (y >> 3)) * nb[0] + (x >> 3)] = 1;
        empty = true;
        for (int bz = 0; bz < nb[2]; ++bz)
            for (int by = 0; by < nb[1]; ++by)
                for (int bx = 0; bx < nb[0]; ++bx)
                    if (leaf[((size_t)bz * nb[1] + by) * nb[0] + bx]) {
                        empty = false;
                        const int b[3] = {bx, by, bz};
                        for (int c = 0; c < 3; ++c) {
                            bmin[c] = std::min(bmin[c], b[c]);
                            bmax[c] = std::max(bmax[c], b[c]);
                        }
                    }
    }
    float value(int x, int y, int z) const {
        if (x < 0 || y < 0 || z < 0 || x >= n[0] || y >= n[1] || z >= n[2]) return 0.0f;
        return v[((size_t)z * n[1] + y) * n[0] + x];
    }
    // probeConstNode<LeafNode>(ijk) for a DDA cell origin ijk
    bool hasLeaf(int x, int y, int z) const {
        const int bx = x >> 3, by = y >> 3, bz = z >> 3;
        if (bx < 0 || by < 0 || bz < 0 || bx >= nb[0] || by >= nb[1] || bz >= nb[2]) return false;
        return leaf[((size_t)bz * nb[1] + by) * nb[0] + bx] != 0;
    }
};

// OpenVDB tools::BoxSampler::sample on a FloatGrid: lerps take float operands
// and a double weight: a + float(double(b - a) * w).
float interpolate(float a, float b, double w) {
    const double temp = (double)(b - a) * w;
    return a + (float)temp;
}

float boxSample(const Grid &g, double x, double y, double z) {
    const int i = (int)std::floor(x), j = (int)std::floor(y), k = (int)std::floor(z);
    const double u = x - (double)i, vv = y - (double)j, w = z - (double)k;
    float d[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) d[a][b][c] = g.value(i + a, j + b, k + c);
    return interpolate(interpolate(interpolate(d[0][0][0], d[0][0][1], w), interpolate(d[0][1][0], d[0][1][1], w), vv),
                       interpolate(interpolate(d[1][0][0], d[1][0][1], w), interpolate(d[1][1][0], d[1][1][1], w), vv),
                       u);
}

float wsSample(const Grid &g, const Vec3 &p) {
    return boxSample(g, (p[0] - g.origin[0]) * g.inv_voxel, (p[1] - g.origin[1]) * g.inv_voxel,
                     (p[2] - g.origin[2]) * g.inv_voxel);
}

// -------------------------------------------------------------------- Ray ---
constexpr double kDelta = 1e-9;  // math::Delta<double>

struct Ray {
    Vec3 eye, dir, inv;
    double t0, t1;
    Ray(const Vec3 &e, const Vec3 &d, double a = kDelta, double b = DBL_MAX) : eye(e), dir(d), t0(a), t1(b) {
        for (int c = 0; c < 3; ++c) inv[c] = 1.0 / dir[c];
    }
    Vec3 at(double t) const { return eye + dir * t; }
    // Ray::worldToIndex -> applyInverseMap of a uniform scale(+translate) map
    Ray worldToIndex(const Grid &g) const {
        Vec3 e, d;
        for (int c = 0; c < 3; ++c) {
            e[c] = (eye[c] - g.origin[c]) * g.inv_voxel;
            d[c] = dir[c] * g.inv_voxel;
        }
        const double len = d.length();
        return Ray(e, d / len, len * t0, len * t1);
    }
    bool clip(const double lo[3], const double hi[3]) {
        double a0 = t0, a1 = t1;
        for (int i = 0; i < 3; ++i) {
            double a = lo[i], b = hi[i];
            if (!(std::fabs(dir[i] - 0.0) > 1e-15)) {
                if (eye[i] < a || eye[i] > b) return false;
            } else {
                a = (a - eye[i]) * inv[i];
                b = (b - eye[i]) * inv[i];
                if (a > b) std::swap(a, b);
                if (a > a0) a0 = a;
                if (b < a1) a1 = b;
                if (a0 > a1) return false;
            }
        }
        t0 = a0;
        t1 = a1;
        return true;
    }
};

// math::DDA<Ray, 3> (leaf level, 8^3 cells)
// math::DDA over 8^3 bricks.  The m-th plane crossing of axis a is timed in
// closed form, next0[a] + m * delta[a], where OpenVDB accumulates next += delta
// (the same plane; the closed form lets 64 lanes of a wave time 64 crossings at
// once on the device).
struct DDA {
    int voxel[3], step[3], m[3];
    double t0, t1, next0[3], next[3], delta[3];
    void init(const Ray &r) {
        t0 = r.t0;
        t1 = r.t1;
        const Vec3 pos = r.at(t0);
        for (int a = 0; a < 3; ++a) voxel[a] = ((int)std::floor(pos[a])) & ~7;
        for (int a = 0; a < 3; ++a) {
            m[a] = 0;
            if (r.dir[a] == 0.0) {
                step[a] = 0;
                next0[a] = DBL_MAX;
                delta[a] = DBL_MAX;
            } else if (r.inv[a] > 0) {
                step[a] = 8;
                next0[a] = t0 + ((double)(voxel[a] + 8) - pos[a]) * r.inv[a];
                delta[a] = step[a] * r.inv[a];
            } else {
                step[a] = -8;
                next0[a] = t0 + ((double)voxel[a] - pos[a]) * r.inv[a];
                delta[a] = step[a] * r.inv[a];
            }
            next[a] = next0[a];
        }
    }
    static int minIndex(const double v[3]) {
        static const int table[8] = {2, 1, 9, 1, 2, 9, 0, 0};
        const int key = ((v[0] < v[1]) << 2) + ((v[0] < v[2]) << 1) + (v[1] < v[2]);
        return table[key];
    }
    bool stepOnce() {
        const int a = minIndex(next);
        t0 = next[a];
        ++m[a];
        next[a] = next0[a] + (double)m[a] * delta[a];
        voxel[a] += step[a];
        return t0 <= t1;
    }
};

// VolumeRayIntersector<FloatGrid> with a leaf-level VolumeHDDA.  Parity unpinned
// against OpenVDB for span ends reached through internal-node levels (the
// reference's hierarchy computes the same boundary plane from another DDA).
struct Intersector {
    const Grid *g;
    Ray ray{Vec3(), Vec3(1, 0, 0)};
    double tmax = 0;
    double lo[3], hi[3];
    explicit Intersector(const Grid &grid) : g(&grid) {
        for (int c = 0; c < 3; ++c) {
            lo[c] = 8.0 * grid.bmin[c];
            hi[c] = 8.0 * (grid.bmax[c] + 1);
        }
    }
    bool hitGridBbox(const Ray &wRay) {
        if (g->empty) return false;
        ray = wRay.worldToIndex(*g);
        const bool hit = ray.clip(lo, hi);
        if (hit) tmax = ray.t1;
        return hit;
    }
    // VolumeHDDA::march over leaves
    void hdda(double &s0, double &s1) {
        s0 = s1 = -1.0;
        if (!((ray.t1 - ray.t0) > 1e-5)) return;  // Ray::valid(Delta<float>)
        DDA dda;
        dda.init(ray);
        do {
            if (g->hasLeaf(dda.voxel[0], dda.voxel[1], dda.voxel[2])) {
                if (s0 < 0) s0 = dda.t0;
            } else if (s0 >= 0) {
                s1 = dda.t0;
                if (s1 - s0 > kDelta) return;
                s0 = s1 = -1.0;
            }
        } while (dda.stepOnce());
        if (s0 >= 0) s1 = dda.t1;
    }
    bool march(double &t0, double &t1) {
        double s0, s1;
        hdda(s0, s1);
        if (s1 > 0) {
            ray.t0 = s1 + kDelta;
            ray.t1 = tmax;
        }
        t0 = s0;
        t1 = s1;
        return (s1 - s0) > kDelta;
    }
    Vec3 indexPos(double t) const { return ray.at(t); }
};

float sampleIndex(const Grid &g, const Vec3 &p) { return boxSample(g, p[0], p[1], p[2]); }

// ---------------------------------------------------------------- shading ---
struct Shading {
    const lg_shader *s;
    Vec3 ext() const { return Vec3(s->extinction[0], s->extinction[1], s->extinction[2]); }
    Vec3 albedo() const { return Vec3(s->albedo[0], s->albedo[1], s->albedo[2]); }
    // ColorRamp::getColor (Shader.h:63-92)
    Vec3 emissionColor(double x) const {
        const int nr = std::min(s->ramp_count, LG_MAX_RAMP);
        if (nr <= 0) return Vec3(0.0);
        const double hp = std::min(std::max((x - s->ramp_in) / (s->ramp_out - s->ramp_in), 0.0), 1.0);
        if (nr == 1) return Vec3(s->ramp_color[0][0], s->ramp_color[0][1], s->ramp_color[0][2]);
        size_t prev = 0, next = 0;
        for (int i = 0; i < nr; i++) {
            if (s->ramp_pos[i] <= hp) {
                prev = i;
            } else {
                next = i;
                break;
            }
        }
        const double range = s->ramp_pos[next] - s->ramp_pos[prev];
        const double blend = (hp - s->ramp_pos[prev]) / range;
        const Vec3 cp(s->ramp_color[prev][0], s->ramp_color[prev][1], s->ramp_color[prev][2]);
        const Vec3 cn(s->ramp_color[next][0], s->ramp_color[next][1], s->ramp_color[next][2]);
        return cp * (1.0 - blend) + cn * blend;
    }
};

// ----------------------------------------------------------------- lights ---
struct EnvMap {
    int w = 0, h = 0;
    const float *rgb = nullptr;
    std::vector<double> cond, marg;
};

double sample1d(const double *cdf, int count, double u, int *off) {
    const int size = count + 1;
    int first = 0, len = size;
    while (len > 0) {  // FindInterval (Sampling.h:60-72)
        int half = len >> 1, middle = first + half;
        if (cdf[middle] <= u) {
            first = middle + 1;
            len -= half + 1;
        } else
            len = half;
    }
    const int offset = std::min(std::max(first - 1, 0), size - 2);
    if (off) *off = offset;
    double du = u - cdf[offset];
    if ((cdf[offset + 1] - cdf[offset]) > 0) du /= (cdf[offset + 1] - cdf[offset]);
    return (offset + du) / count;
}

float sinF(float x) { return (float)std::sin((double)x); }
float cosF(float x) { return (float)std::cos((double)x); }

// EnvLight::illuminate (Light.h:192-231) for the two uniform draws r0, r1:
// the continuous sample (s0, s1) of the map's 2D distribution, the bilinear
// texel colour at (s0 w, s1 h) and the direction -(cos t sin p, cos p,
// sin t sin p) with t = 2 pi s0, p = pi s1, through the light's transform.
void envIlluminate(const lg_light &L, const EnvMap *env, float r0, float r1, Vec3 &lightDir, Vec3 &lightIntensity,
                   float *s0_out = nullptr, float *s1_out = nullptr) {
    const Vec3 color(L.color[0], L.color[1], L.color[2]);
    int v = 0;
    const float s1 = (float)sample1d(env->marg.data(), env->h, r1, &v);
    const float s0 = (float)sample1d(env->cond.data() + (size_t)v * (env->w + 1), env->w, r0, nullptr);
    if (s0_out) *s0_out = s0;
    if (s1_out) *s1_out = s1;
    int col_lo = (int)std::floor(s0 * env->w), row_lo = (int)std::floor(s1 * env->h);
    int col_hi = (int)std::ceil(s0 * env->w), row_hi = (int)std::ceil(s1 * env->h);
    const float col_ratio = s0 * env->w - (float)col_lo;
    const float row_ratio = s1 * env->h - (float)row_lo;
    col_lo = std::min(std::max(col_lo, 0), env->w - 1);
    col_hi = std::min(std::max(col_hi, 0), env->w - 1);
    row_lo = std::min(std::max(row_lo, 0), env->h - 1);
    row_hi = std::min(std::max(row_hi, 0), env->h - 1);
    const float theta = s0 * M_PI * 2.0f;
    const float phi = s1 * M_PI;
    const double sinPhi = sinF(phi);
    Vec3 sd((double)cosF(theta) * sinPhi, (double)cosF(phi), (double)sinF(theta) * sinPhi);
    const double *m = L.xform;
    Vec3 t(sd[0] * m[0] + sd[1] * m[4] + sd[2] * m[8] + m[12], sd[0] * m[1] + sd[1] * m[5] + sd[2] * m[9] + m[13],
           sd[0] * m[2] + sd[1] * m[6] + sd[2] * m[10] + m[14]);
    sd = t / t.length();
    auto px = [&](int r, int c) {
        const float *p = env->rgb + 3 * ((size_t)r * env->w + c);
        return Vec3(p[0], p[1], p[2]);
    };
    const Vec3 img = px(row_lo, col_lo) * (double)((1.0f - row_ratio) * (1.0f - col_ratio)) +
                     px(row_lo, col_hi) * (double)((1.0f - row_ratio) * col_ratio) +
                     px(row_hi, col_lo) * (double)(row_ratio * (1.0f - col_ratio)) +
                     px(row_hi, col_hi) * (double)(row_ratio * col_ratio);
    lightDir = -sd;
    lightIntensity = img * color * (double)L.intensity;
}

// Light::illuminate (Light.h:79-89, 118-126, 161-231)
void illuminate(const lg_light &L, const EnvMap *env, const Vec3 &hitPos, CounterStream *rng, Vec3 &lightDir,
                Vec3 &lightIntensity, double &distance) {
    const Vec3 color(L.color[0], L.color[1], L.color[2]);
    if (L.type == LG_LIGHT_DIRECTIONAL) {
        lightDir = Vec3(L.dir[0], L.dir[1], L.dir[2]);
        lightIntensity = color * (double)L.intensity;
        distance = DBL_MAX;
    } else if (L.type == LG_LIGHT_POINT) {
        lightDir = hitPos - Vec3(L.pos[0], L.pos[1], L.pos[2]);
        const float r2 = (float)lightDir.lengthSqr();
        distance = std::sqrt(r2);
        lightDir = lightDir / distance;
        lightIntensity = color * (double)L.intensity / (1.0 + 4.0 * M_PI * r2);
    } else {
        const float r0 = (*rng)();
        const float r1 = (*rng)();
        envIlluminate(L, env, r0, r1, lightDir, lightIntensity);
        distance = DBL_MAX;
    }
}

struct Scene {
    const Grid *density;
    const Grid *temperature;  // may be null
    Shading sh;
    const lg_march_params *mp;
    const lg_light *lights;
    int nlights;
    std::vector<EnvMap> envs;  // parallel to lights
};

// ----------------------------------------------------------- LightGraph -----
// buildGraph (LightGraph.h:287-318); cap <= 0: uncapped.
void buildGraph(const float *pts, int n, double radius, int cap, std::vector<int32_t> &edges) {
    const float rf = (float)radius;
    const float r2 = rf * rf;
    std::vector<uint32_t> cnt(n, 0);
    const uint32_t maxc = cap > 0 ? (uint32_t)cap : 0xffffffffu;
    edges.clear();
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) {
            if (cnt[i] >= maxc) break;
            if (cnt[j] >= maxc) continue;
            const Vec3 pi(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
            const Vec3 pj(pts[3 * j], pts[3 * j + 1], pts[3 * j + 2]);
            if (i < j && (pi - pj).lengthSqr() < r2) {
                edges.push_back(i);
                edges.push_back(j);
                cnt[i]++;
                cnt[j]++;
            }
        }
    }
}

// raymarchConnection for one edge (LightGraph.h:364-406); returns sTrans.
Vec3 marchEdge(const Scene &sc, const Vec3 &startPos, const Vec3 &endPos, bool *hit) {
    const Grid &g = *sc.density;
    const lg_shader &s = *sc.sh.s;
    const double h = sc.mp->step[1], cutoff = sc.mp->cutoff;
    Vec3 dir = endPos - startPos;
    const float dist = (float)dir.length();
    {  // Vec3::normalize(eps = 1e-7): multiply by the reciprocal length
        const double d = dir.length();
        if (!(std::fabs(d - 0.0) <= 1e-7)) dir = dir * (1.0 / d);
    }
    Ray sRay(startPos, dir);
    sRay.t1 = dist;  // setMaxTime
    Vec3 sTrans(1.0);
    Intersector shadow(g);
    *hit = shadow.hitGridBbox(sRay);
    if (!*hit) return Vec3(0.0);
    double sT0, sT1;
    const Vec3 ext = sc.sh.ext();
    while (shadow.march(sT0, sT1)) {
        bool sDropout = false;
        double sT = h * std::ceil(sT0 / h);
        for (; sT <= sT1; sT += h) {
            const double sDensity = std::min((double)sampleIndex(g, shadow.indexPos(sT)) * s.density_scale *
                                                 s.shadow_density_scale * s.scatter_density_scale,
                                             s.density_max);
            if (sDensity < cutoff && s.scatter_density_min < cutoff) continue;
            sTrans *= expv(ext * std::max(sDensity, s.scatter_density_min) * h);
            if (sTrans.lengthSqr() < cutoff) {
                sTrans = Vec3(0.0);
                sDropout = true;
                break;
            }
        }
        if (sDropout) break;
    }
    return sTrans;
}

// Shadow march toward a light (LightGraph.h:461-491): no floor, no jitter.
Vec3 shadowTransmittance(const Scene &sc, const Vec3 &p, const Vec3 &lightDir, double distance, bool *hit) {
    const Grid &g = *sc.density;
    const lg_shader &s = *sc.sh.s;
    const double h = sc.mp->step[1], cutoff = sc.mp->cutoff;
    Ray sRay(p, -lightDir, kDelta, distance);
    Vec3 sTrans(1.0);
    Intersector shadow(g);
    *hit = shadow.hitGridBbox(sRay);
    if (!*hit) return sTrans;
    double sT0, sT1;
    const Vec3 ext = sc.sh.ext();
    while (shadow.march(sT0, sT1)) {
        bool sDropout = false;
        double sT = h * std::ceil(sT0 / h);
        for (; sT <= sT1; sT += h) {
            const double sDensity = std::min((double)sampleIndex(g, shadow.indexPos(sT)) * s.density_scale *
                                                 s.shadow_density_scale * s.scatter_density_scale,
                                             s.density_max);
            if (sDensity < cutoff) continue;
            sTrans *= expv(ext * sDensity * h);
            if (sTrans.lengthSqr() < cutoff) {
                sTrans = Vec3(0.0);
                sDropout = true;
                break;
            }
        }
        if (sDropout) break;
    }
    return sTrans;
}

// computeLighting for one vertex (LightGraph.h:438-497)
Vec3 vertexLight(const Scene &sc, const Vec3 &p, uint32_t vertex, uint32_t solve, uint64_t env_seed) {
    Vec3 L(0.0);
    double pTemp = 0.0;
    if (sc.temperature) pTemp = (double)wsSample(*sc.temperature, p);
    const Vec3 emission = sc.sh.emissionColor(pTemp) * sc.sh.s->emission_scale;
    CounterStream env_rng(vertex, solve, 4u, env_seed);
    for (int j = 0; j < sc.nlights; j++) {
        Vec3 tmpLumi(0.0);
        uint64_t lightSamples = (uint64_t)sc.lights[j].samples;
        if (lightSamples > 1) lightSamples *= 10;
        for (size_t s = 0; s < lightSamples; s++) {
            Vec3 lightDir, lightIntensity;
            double distance;
            illuminate(sc.lights[j], &sc.envs[j], p, &env_rng, lightDir, lightIntensity, distance);
            bool hit = false;
            const Vec3 sTrans = shadowTransmittance(sc, p, lightDir, distance, &hit);
            if (!hit) continue;
            tmpLumi += lightIntensity * sTrans;
        }
        L += sc.sh.albedo() * (tmpLumi / (double)lightSamples);
    }
    L += emission;
    return L;
}

// floydAlgo (LightGraph.h:512-534) on a row-major N x N Vec3 array.
void floydMaxProduct(std::vector<Vec3> &graph, int N) {
    for (int k = 0; k < N; k++)
        for (int i = 0; i < N; i++)
            if (graph[(size_t)i * N + k].lengthSqr() > 0)
                for (int j = 0; j < N; j++)
                    if (graph[(size_t)k * N + j].lengthSqr() > 0)
                        graph[(size_t)i * N + j] =
                            maxComponent(graph[(size_t)i * N + j], graph[(size_t)i * N + k] * graph[(size_t)k * N + j]);
    for (int i = 0; i < N; i++) graph[(size_t)i * N + i] = Vec3(1.0);
}

// diffuseLight (LightGraph.h:554-579)
void diffuseLight(std::vector<Vec3> &P, const std::vector<Vec3> &L, std::vector<Vec3> &out, int N) {
    for (int j = 0; j < N; j++) {
        Vec3 accum(0.0);
        for (int i = 0; i < N; i++) accum += P[(size_t)i * N + j];
        accum = maxComponent(accum, Vec3(1.0));
        for (int i = 0; i < N; i++) P[(size_t)i * N + j] /= accum;
    }
    out.assign(N, Vec3(0.0));
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) out[i] += L[j] * P[(size_t)i * N + j];
}

Scene makeScene(const Grid *den, const Grid *tmp, const lg_shader *sh, const lg_march_params *mp,
                const lg_light *lights, int nl, const float *const *env_rgb, const int *env_w, const int *env_h);

}  // namespace

// ================================================================= C ABI =====
extern "C" {

void lgo_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox4x32(ctr, key, out); }

// libstdc++ std::mt19937 + uniform_real_distribution<float>(0,1), RandGen.h:16-44
void lgo_mt19937_floats(uint32_t seed, int64_t count, float *out) {
    std::mt19937 eng(seed);
    std::uniform_real_distribution<float> dist(0.0f, 1.0f);
    for (int64_t i = 0; i < count; ++i) out[i] = dist(eng);
}
uint32_t lgo_mt19937_nth(uint32_t seed, int64_t nth) {
    std::mt19937 eng(seed);
    uint32_t v = 0;
    for (int64_t i = 0; i < nth; ++i) v = eng();
    return v;
}

// BoxSampler at index-space points (count x 3 doubles).
void lgo_box_sample(const lg_grid_desc *d, const float *dense, int64_t count, const double *xyz, float *out) {
    Grid g(*d, dense);
    for (int64_t i = 0; i < count; ++i) out[i] = boxSample(g, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}

// Brick-level spans of a world ray: returns the number of spans written (<= cap),
// -1 on a bbox miss.
int32_t lgo_spans(const lg_grid_desc *d, const float *dense, const double eye[3], const double dir[3], double t0,
                  double t1, int32_t cap, double *out) {
    Grid g(*d, dense);
    Intersector is(g);
    if (!is.hitGridBbox(Ray(Vec3(eye[0], eye[1], eye[2]), Vec3(dir[0], dir[1], dir[2]), t0, t1))) return -1;
    int k = 0;
    double a, b;
    while (is.march(a, b)) {
        if (k < cap) {
            out[2 * k] = a;
            out[2 * k + 1] = b;
        }
        ++k;
    }
    return k;
}

// Density-weighted sampler (kernel (a) contract): voxels weighted by
// floor(rho*2^24), bricks in storage order (z, y, x), voxels in brick order.
int32_t lgo_sample_points(const lg_grid_desc *d, const float *dense, uint64_t seed, uint32_t solve, int32_t n,
                          float *out) {
    Grid g(*d, dense);
    auto q = [&](float f) -> uint64_t {
        if (!(f > 0.0f)) return 0;
        return (uint64_t)std::min((double)f * 16777216.0, 549755813888.0 * 16777216.0);
    };
    struct B {
        int bx, by, bz;
        uint64_t start;
    };
    std::vector<B> bricks;
    uint64_t total = 0;
    for (int bz = 0; bz < g.nb[2]; ++bz)
        for (int by = 0; by < g.nb[1]; ++by)
            for (int bx = 0; bx < g.nb[0]; ++bx) {
                if (!g.leaf[((size_t)bz * g.nb[1] + by) * g.nb[0] + bx]) continue;
                bricks.push_back({bx, by, bz, total});
                for (int l = 0; l < 512; ++l)
                    total += q(g.value(bx * 8 + (l & 7), by * 8 + ((l >> 3) & 7), bz * 8 + (l >> 6)));
            }
    if (total == 0) return -1;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int i = 0; i < n; ++i) {
        uint32_t r0[4], r1[4];
        const uint32_t c0[4] = {(uint32_t)i, solve, 0u, 1u}, c1[4] = {(uint32_t)i, solve, 1u, 1u};
        philox4x32(c0, key, r0);
        philox4x32(c1, key, r1);
        const uint64_t u = ((uint64_t)r0[1] << 32) | r0[0];
        const uint64_t target = (uint64_t)(((unsigned __int128)u * total) >> 64);
        size_t b = 0;
        for (size_t k = 0; k < bricks.size(); ++k) {
            if (bricks[k].start <= target) b = k;
            else break;
        }
        uint64_t local = target - bricks[b].start, cum = 0;
        int l = 511;
        for (int k = 0; k < 512; ++k) {
            cum += q(g.value(bricks[b].bx * 8 + (k & 7), bricks[b].by * 8 + ((k >> 3) & 7), bricks[b].bz * 8 + (k >> 6)));
            if (local < cum) {
                l = k;
                break;
            }
        }
        const int vx = bricks[b].bx * 8 + (l & 7), vy = bricks[b].by * 8 + ((l >> 3) & 7), vz = bricks[b].bz * 8 + (l >> 6);
        const double p[3] = {(double)vx - 0.5 + (double)u01(r0[2]), (double)vy - 0.5 + (double)u01(r0[3]),
                             (double)vz - 0.5 + (double)u01(r1[0])};
        for (int c = 0; c < 3; ++c) out[3 * i + c] = (float)(g.origin[c] + g.voxel * p[c]);
    }
    return 0;
}

int32_t lgo_build_edges(const float *pts, int32_t n, double radius, int32_t cap, int32_t *edges_out,
                        int32_t edge_cap) {
    std::vector<int32_t> e;
    buildGraph(pts, n, radius, cap, e);
    const int ne = (int)e.size() / 2;
    for (int k = 0; k < std::min(ne, edge_cap); ++k) {
        edges_out[2 * k] = e[2 * k];
        edges_out[2 * k + 1] = e[2 * k + 1];
    }
    return ne;
}

// Edge transmittance (n_edges x 3) and the optical depth it encodes
// (-ln(T_0)/sigma_0; +inf for a dropped or missed edge).
void lgo_march_edges(const lg_grid_desc *d, const float *dense, const lg_shader *sh, const lg_march_params *mp,
                     const float *pts, const int32_t *edges, int32_t ne, double *T_out, double *tau_out) {
    Grid g(*d, dense);
    Scene sc = makeScene(&g, nullptr, sh, mp, nullptr, 0, nullptr, nullptr, nullptr);
    for (int e = 0; e < ne; ++e) {
        const int i = edges[2 * e], j = edges[2 * e + 1];
        bool hit;
        const Vec3 T = marchEdge(sc, Vec3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]),
                                 Vec3(pts[3 * j], pts[3 * j + 1], pts[3 * j + 2]), &hit);
        for (int c = 0; c < 3; ++c) T_out[3 * e + c] = T[c];
        tau_out[e] = (T.lengthSqr() > 0) ? -std::log(T[0]) / -sh->extinction[0] : INFINITY;
    }
}

void lgo_vertex_light(const lg_grid_desc *d, const float *dense, const float *temp_dense, const lg_shader *sh,
                      const lg_march_params *mp, const lg_light *lights, int32_t nl, const float *const *env_rgb,
                      const int32_t *env_w, const int32_t *env_h, uint64_t env_seed, uint32_t solve,
                      const float *pts, int32_t n, double *L_out) {
    Grid g(*d, dense);
    std::unique_ptr<Grid> tg;
    if (temp_dense) tg.reset(new Grid(*d, temp_dense));
    Scene sc = makeScene(&g, tg.get(), sh, mp, lights, nl, env_rgb, env_w, env_h);
    for (int i = 0; i < n; ++i) {
        const Vec3 L = vertexLight(sc, Vec3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]), (uint32_t)i, solve, env_seed);
        for (int c = 0; c < 3; ++c) L_out[3 * i + c] = L[c];
    }
}

// Literal max-product Floyd-Warshall on Vec3 transmittance, in place (N x N x 3).
void lgo_floyd_maxprod(int32_t N, double *P) {
    std::vector<Vec3> g((size_t)N * N);
    for (size_t k = 0; k < g.size(); ++k) g[k] = Vec3(P[3 * k], P[3 * k + 1], P[3 * k + 2]);
    floydMaxProduct(g, N);
    for (size_t k = 0; k < g.size(); ++k)
        for (int c = 0; c < 3; ++c) P[3 * k + c] = g[k][c];
}

// Classical Floyd-Warshall in (min, +) with k ascending, FP64 or exact on
// integer-valued inputs; D is N x ld row-major, +inf = no edge, zero diagonal.
void lgo_floyd_minplus(int32_t N, int32_t ld, double *D) {
    for (int k = 0; k < N; ++k)
        for (int i = 0; i < N; ++i) {
            const double dik = D[(size_t)i * ld + k];
            if (!(dik < INFINITY)) continue;
            double *row = D + (size_t)i * ld;
            const double *rk = D + (size_t)k * ld;
            for (int j = 0; j < N; ++j) {
                const double v = dik + rk[j];
                if (v < row[j]) row[j] = v;
            }
        }
}

// Same in float32 (the device's arithmetic) with the classical order.
void lgo_floyd_minplus_f32(int32_t N, int32_t ld, float *D) {
    for (int k = 0; k < N; ++k)
        for (int i = 0; i < N; ++i) {
            const float dik = D[(size_t)i * ld + k];
            if (!(dik < INFINITY)) continue;
            float *row = D + (size_t)i * ld;
            const float *rk = D + (size_t)k * ld;
            for (int j = 0; j < N; ++j) {
                const float v = dik + rk[j];
                if (v < row[j]) row[j] = v;
            }
        }
}

// Canonical predecessors from a closed D (N x ld) and the edge list.
void lgo_predecessors(int32_t N, int32_t ld, const double *D, const int32_t *edges, const double *w, int32_t ne,
                      int32_t *pred) {
    std::vector<std::vector<std::pair<int, double>>> adj(N);
    for (int e = 0; e < ne; ++e) {
        adj[edges[2 * e + 1]].push_back({edges[2 * e], w[e]});
        adj[edges[2 * e]].push_back({edges[2 * e + 1], w[e]});
    }
    for (auto &a : adj) std::sort(a.begin(), a.end(), [](auto &x, auto &y) { return x.first < y.first; });
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            int best_k = -1;
            if (i == j)
                best_k = i;
            else if (D[(size_t)i * ld + j] < INFINITY) {
                double best = INFINITY;
                for (auto &kw : adj[j]) {
                    const double v = D[(size_t)i * ld + kw.first] + kw.second;
                    if (v < best) {
                        best = v;
                        best_k = kw.first;
                    }
                }
            }
            pred[(size_t)i * ld + j] = best_k;
        }
}

// diffuseLight on transmittance P (N x N x 3, diagonal already 1).
void lgo_diffuse(int32_t N, const double *P, const double *L, double *out) {
    std::vector<Vec3> g((size_t)N * N), l(N), o;
    for (size_t k = 0; k < g.size(); ++k) g[k] = Vec3(P[3 * k], P[3 * k + 1], P[3 * k + 2]);
    for (int k = 0; k < N; ++k) l[k] = Vec3(L[3 * k], L[3 * k + 1], L[3 * k + 2]);
    diffuseLight(g, l, o, N);
    for (int k = 0; k < N; ++k)
        for (int c = 0; c < 3; ++c) out[3 * k + c] = o[k][c];
}

// One full solve of the reference's LightGraph::operator() body on given points:
// edges -> march -> max-product FW -> vertex light -> diffuse.  Outputs: edges
// (cap), transmittance P (N x N x 3, optional), L, diffuse.  Returns the edge
// count.
int32_t lgo_solve_points(const lg_grid_desc *d, const float *dense, const float *temp_dense, const lg_shader *sh,
                         const lg_march_params *mp, const lg_light *lights, int32_t nl, const float *const *env_rgb,
                         const int32_t *env_w, const int32_t *env_h, uint64_t env_seed, uint32_t solve,
                         const float *pts, int32_t N, double radius, int32_t cap, int32_t *edges_out,
                         int32_t edge_cap, double *P_out, double *L_out, double *diffuse_out) {
    Grid g(*d, dense);
    std::unique_ptr<Grid> tg;
    if (temp_dense) tg.reset(new Grid(*d, temp_dense));
    Scene sc = makeScene(&g, tg.get(), sh, mp, lights, nl, env_rgb, env_w, env_h);
    std::vector<int32_t> e;
    buildGraph(pts, N, radius, cap, e);
    const int ne = (int)e.size() / 2;
    for (int k = 0; k < std::min(ne, edge_cap); ++k) {
        edges_out[2 * k] = e[2 * k];
        edges_out[2 * k + 1] = e[2 * k + 1];
    }
    std::vector<Vec3> P((size_t)N * N, Vec3(0.0));
    for (int k = 0; k < ne; ++k) {
        const int i = e[2 * k], j = e[2 * k + 1];
        bool hit;
        const Vec3 T = marchEdge(sc, Vec3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]),
                                 Vec3(pts[3 * j], pts[3 * j + 1], pts[3 * j + 2]), &hit);
        if (!hit) continue;
        P[(size_t)i * N + j] = T;
        P[(size_t)j * N + i] = T;
    }
    floydMaxProduct(P, N);
    if (P_out)
        for (size_t k = 0; k < P.size(); ++k)
            for (int c = 0; c < 3; ++c) P_out[3 * k + c] = P[k][c];
    std::vector<Vec3> L(N), out;
    for (int i = 0; i < N; ++i) L[i] = vertexLight(sc, Vec3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]), (uint32_t)i, solve, env_seed);
    diffuseLight(P, L, out, N);
    for (int i = 0; i < N; ++i)
        for (int c = 0; c < 3; ++c) {
            L_out[3 * i + c] = L[i][c];
            diffuse_out[3 * i + c] = out[i][c];
        }
    return ne;
}

// Reference-mode point scattering (LightGraph.h:206-274) with std::mt19937
// (RandGen.h) seeded 854123+iter, for the CPU baseline.  Returns the point count
// and updates *min_dist as the reference does.  hard_limit lifts the 1000 cap.
int32_t lgo_dart_throw(const lg_grid_desc *d, const float *dense, uint32_t iter, uint32_t scatter_count,
                       uint32_t target, float *min_dist, float *out, int32_t out_cap) {
    Grid g(*d, dense);
    std::mt19937 eng(854123u + iter);
    std::uniform_real_distribution<float> dist(0.0f, 1.0f);
    // world bbox of the active voxels at voxel resolution (main.cc:512), its size as Vec3s
    double lo[3];
    float sz[3];
    {
        int mn[3] = {1 << 30, 1 << 30, 1 << 30}, mx[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
        for (int z = 0; z < g.n[2]; ++z)
            for (int y = 0; y < g.n[1]; ++y)
                for (int x = 0; x < g.n[0]; ++x)
                    if (g.value(x, y, z) != 0.0f) {
                        const int p[3] = {x, y, z};
                        for (int c = 0; c < 3; ++c) {
                            mn[c] = std::min(mn[c], p[c]);
                            mx[c] = std::max(mx[c], p[c]);
                        }
                    }
        for (int c = 0; c < 3; ++c) {
            lo[c] = g.origin[c] + g.voxel * mn[c];
            sz[c] = (float)((g.origin[c] + g.voxel * mx[c]) - lo[c]);
        }
    }
    float minDistSquared = (*min_dist) * (*min_dist);
    float lastValid = minDistSquared;
    const uint32_t batch = std::max(1u, scatter_count / 100u);
    std::vector<Vec3> pc;
    uint32_t pcSize = 0;
    for (size_t n = 0; n < scatter_count; ++n) {
        float np[3];
        for (int c = 0; c < 3; ++c) np[c] = (float)((double)(dist(eng) * sz[c]) + lo[c]);
        const Vec3 pt(np[0], np[1], np[2]);
        if (pcSize > target) break;
        if ((n + 1) % batch == 0) {
            // PC_GROWTH_FUNC_POWER is a float constant (LightGraph.h:28): pow gets (double)0.17f
            const uint32_t expected = (uint32_t)(target * std::pow((double)n / (double)scatter_count, (double)0.17f));
            const double ratio = pcSize / (double)expected;
            minDistSquared *= ratio;
        }
        if (wsSample(g, pt) < 0.001f) continue;
        bool collision = false;
        for (size_t i = 0; i < pc.size(); i++)
            if ((pt - pc[i]).lengthSqr() < minDistSquared) {
                collision = true;
                break;
            }
        if (!collision) {
            pc.push_back(pt);
            lastValid = minDistSquared;
            pcSize++;
        }
    }
    *min_dist = std::sqrt(lastValid);
    const int cnt = (int)pc.size();
    for (int i = 0; i < std::min(cnt, out_cap); ++i)
        for (int c = 0; c < 3; ++c) out[3 * i + c] = (float)pc[i][c];
    return cnt;
}

}  // extern "C"

namespace {
Scene makeScene(const Grid *den, const Grid *tmp, const lg_shader *sh, const lg_march_params *mp,
                const lg_light *lights, int nl, const float *const *env_rgb, const int *env_w, const int *env_h) {
    Scene sc;
    sc.density = den;
    sc.temperature = tmp;
    sc.sh.s = sh;
    sc.mp = mp;
    sc.lights = lights;
    sc.nlights = nl;
    sc.envs.resize(std::max(nl, 0));
    for (int k = 0; k < nl; ++k) {
        if (lights[k].type != LG_LIGHT_ENV || !env_rgb || !env_rgb[k]) continue;
        EnvMap &E = sc.envs[k];
        E.w = env_w[k];
        E.h = env_h[k];
        E.rgb = env_rgb[k];
        // Light.h:165-176 weights, Distribution2D with FP64 prefix sums
        std::vector<float> Y((size_t)E.w * E.h);
        for (int i = 0; i < E.h; i++) {
            const float sinTheta = (float)std::sin(M_PI * (double)((float)i / (float)(E.h - 1)));
            for (int j = 0; j < E.w; j++) {
                const float *p = E.rgb + 3 * ((size_t)i * E.w + j);
                float y = (float)(0.2126 * (double)p[0] + 0.7152 * (double)p[1] + 0.0722 * (double)p[2]);
                y *= sinTheta;
                Y[(size_t)i * E.w + j] = y;
            }
        }
        E.cond.assign((size_t)E.h * (E.w + 1), 0.0);
        std::vector<double> rowint(E.h);
        for (int v = 0; v < E.h; ++v) {
            double *cdf = E.cond.data() + (size_t)v * (E.w + 1);
            cdf[0] = 0;
            for (int i = 1; i <= E.w; ++i) cdf[i] = cdf[i - 1] + (double)Y[(size_t)v * E.w + i - 1] / E.w;
            const double I = cdf[E.w];
            if (I == 0)
                for (int i = 1; i <= E.w; ++i) cdf[i] = (double)i / E.w;
            else
                for (int i = 1; i <= E.w; ++i) cdf[i] /= I;
            rowint[v] = I;
        }
        E.marg.assign(E.h + 1, 0.0);
        for (int i = 1; i <= E.h; ++i) E.marg[i] = E.marg[i - 1] + rowint[i - 1] / E.h;
        const double I = E.marg[E.h];
        if (I == 0)
            for (int i = 1; i <= E.h; ++i) E.marg[i] = (double)i / E.h;
        else
            for (int i = 1; i <= E.h; ++i) E.marg[i] /= I;
    }
    return sc;
}
}  // namespace

// ColorRamp::getColor (Shader.h:63-92) of the shader's temperature ramp at x.
extern "C" void lgo_ramp_color(const lg_shader *sh, double x, double out[3]) {
    const Vec3 c = Shading{sh}.emissionColor(x);
    for (int a = 0; a < 3; ++a) out[a] = c[a];
}

// Distribution1D::SampleContinuous (Sampling.h:104-120) over a CDF of count+1
// entries: the continuous sample in [0, 1) and the segment it fell in.
extern "C" double lgo_sample1d(const double *cdf, int32_t count, double u, int32_t *off) {
    int o = 0;
    const double x = sample1d(cdf, count, u, &o);
    if (off) *off = o;
    return x;
}

extern "C" void lgo_env_cdf(int32_t w, int32_t h, const float *rgb, double *cond, double *marg) {
    lg_light l{};
    l.type = LG_LIGHT_ENV;
    l.samples = 1;
    const float *maps[1] = {rgb};
    const int ws[1] = {w}, hs[1] = {h};
    Scene sc = makeScene(nullptr, nullptr, nullptr, nullptr, &l, 1, maps, ws, hs);
    memcpy(cond, sc.envs[0].cond.data(), sc.envs[0].cond.size() * 8);
    memcpy(marg, sc.envs[0].marg.data(), sc.envs[0].marg.size() * 8);
}

// EnvLight::illuminate for given draws (r0, r1): direction, intensity and the
// continuous sample (s0, s1) of the map's distribution.
extern "C" void lgo_env_illuminate(int32_t w, int32_t h, const float *rgb, const lg_light *L, float r0, float r1,
                                   double *dir, double *inten, float *s01) {
    const float *maps[1] = {rgb};
    const int ws[1] = {w}, hs[1] = {h};
    Scene sc = makeScene(nullptr, nullptr, nullptr, nullptr, L, 1, maps, ws, hs);
    Vec3 d, I;
    envIlluminate(*L, &sc.envs[0], r0, r1, d, I, &s01[0], &s01[1]);
    for (int c = 0; c < 3; ++c) {
        dir[c] = d[c];
        inten[c] = I[c];
    }
}

// ====================================================== splat and render =====
namespace {

struct Coarse {
    int n[3], i0[3];
    double cv;
    const float *rgb;  // n[0]*n[1]*n[2]*3
    Vec3 value(int x, int y, int z) const {
        x -= i0[0];
        y -= i0[1];
        z -= i0[2];
        if (x < 0 || y < 0 || z < 0 || x >= n[0] || y >= n[1] || z >= n[2]) return Vec3(0.0);
        const float *p = rgb + 3 * (((size_t)z * n[1] + y) * n[0] + x);
        return Vec3(p[0], p[1], p[2]);
    }
    // BoxSampler on a Vec3fGrid: per-channel float lerps with double weights
    Vec3 sample(const Vec3 &world) const {
        const double x = world[0] * (1.0 / cv), y = world[1] * (1.0 / cv), z = world[2] * (1.0 / cv);
        const int i = (int)std::floor(x), j = (int)std::floor(y), k = (int)std::floor(z);
        const double u = x - i, v = y - j, w = z - k;
        Vec3 d[2][2][2];
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int c = 0; c < 2; ++c) d[a][b][c] = value(i + a, j + b, k + c);
        Vec3 out;
        for (int ch = 0; ch < 3; ++ch) {
            auto f = [&](int a, int b, int c) { return (float)d[a][b][c][ch]; };
            out[ch] = interpolate(interpolate(interpolate(f(0, 0, 0), f(0, 0, 1), w), interpolate(f(0, 1, 0), f(0, 1, 1), w), v),
                                  interpolate(interpolate(f(1, 0, 0), f(1, 0, 1), w), interpolate(f(1, 1, 0), f(1, 1, 1), w), v),
                                  u);
        }
        return out;
    }
};

struct Camera {
    const lg_camera *c;
    // PerspectiveCamera::getRay (Renderer.h:216-229) with the two jitter draws
    Ray getRay(int i, int j, CounterStream &rng) const {
        const double *M = c->screen_to_world;
        const Vec3 eye(M[12], M[13], M[14]);
        const Vec3 axis(0.0 * M[0] + 0.0 * M[4] + -1.0 * M[8], 0.0 * M[1] + 0.0 * M[5] + -1.0 * M[9],
                        0.0 * M[2] + 0.0 * M[6] + -1.0 * M[10]);
        const double ri = (double)i + 0.5 + ((double)rng() - 0.5);
        const double rj = (double)j + 0.5 + ((double)rng() - 0.5);
        const Vec3 scr((2 * ri / double(c->width) - 1) * c->scale_width,
                       (1 - 2 * rj / double(c->height)) * c->scale_height, -1.0);
        Vec3 dir(scr[0] * M[0] + scr[1] * M[4] + scr[2] * M[8], scr[0] * M[1] + scr[1] * M[5] + scr[2] * M[9],
                 scr[0] * M[2] + scr[1] * M[6] + scr[2] * M[10]);
        const double d = dir.length();
        if (!(std::fabs(d - 0.0) <= 1e-7)) dir = dir * (1.0 / d);
        const double s = 1.0 / dir.dot(axis);
        return Ray(eye, dir, c->znear * s, c->zfar * s);
    }
};

}  // namespace

extern "C" {

// Coarse topology of ScatterGrid::buildScatterGrid (ScatterGrid.h:105-144) as
// restated by the library: footprint of f density voxels centred on the coarse
// voxel, then a 26-neighbour dilation.  Parity unpinned against OpenVDB's
// GridTransformer.
void lgo_scatter_topology(const lg_grid_desc *d, const float *dense, int32_t f, const int32_t n[3],
                          const int32_t i0[3], double cv, uint8_t *active) {
    Grid g(*d, dense);
    std::vector<uint8_t> raw((size_t)n[0] * n[1] * n[2], 0);
    for (int z = 0; z < n[2]; ++z)
        for (int y = 0; y < n[1]; ++y)
            for (int x = 0; x < n[0]; ++x) {
                const int C[3] = {x + i0[0], y + i0[1], z + i0[2]};
                int lo[3];
                for (int a = 0; a < 3; ++a) lo[a] = (int)std::floor(((double)C[a] * cv - g.origin[a]) * g.inv_voxel - 0.5 * f + 0.5);
                bool on = false;
                for (int dz = 0; dz < f && !on; ++dz)
                    for (int dy = 0; dy < f && !on; ++dy)
                        for (int dx = 0; dx < f && !on; ++dx) on = g.value(lo[0] + dx, lo[1] + dy, lo[2] + dz) != 0.0f;
                raw[((size_t)z * n[1] + y) * n[0] + x] = on;
            }
    for (int z = 0; z < n[2]; ++z)
        for (int y = 0; y < n[1]; ++y)
            for (int x = 0; x < n[0]; ++x) {
                bool on = false;
                for (int dz = -1; dz <= 1; ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            const int xx = x + dx, yy = y + dy, zz = z + dz;
                            if (xx < 0 || yy < 0 || zz < 0 || xx >= n[0] || yy >= n[1] || zz >= n[2]) continue;
                            on = on || raw[((size_t)zz * n[1] + yy) * n[0] + xx];
                        }
                active[((size_t)z * n[1] + y) * n[0] + x] = on;
            }
}

// ScatterGrid::operator() (ScatterGrid.h:61-98) over every active coarse voxel,
// by brute force over the points.
void lgo_splat(const float *pts, const float *rad, int64_t count, double radius, const int32_t n[3],
               const int32_t i0[3], double cv, const uint8_t *active, float *rgb_out) {
    const double sigmaSquared = -(radius * radius) / (2 * std::log(0.1));
    const size_t vox = (size_t)n[0] * n[1] * n[2];
    for (size_t k = 0; k < vox; ++k) {
        float *o = rgb_out + 3 * k;
        o[0] = o[1] = o[2] = 0.0f;
        if (!active[k]) continue;
        const int x = (int)(k % n[0]), y = (int)((k / n[0]) % n[1]), z = (int)(k / ((size_t)n[0] * n[1]));
        const Vec3 worldPos((x + i0[0]) * cv, (y + i0[1]) * cv, (z + i0[2]) * cv);
        Vec3 scatter(0.0);
        double totalWeigths = 0.0;
        for (int64_t p = 0; p < count; ++p) {
            const Vec3 pos(pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]);
            if (!((pos - worldPos).lengthSqr() < radius * radius)) continue;
            const Vec3 dd = worldPos - pos;
            const float fx = (float)dd[0], fy = (float)dd[1], fz = (float)dd[2];
            const float l2 = fx * fx + fy * fy + fz * fz;
            const double weigth = std::exp(-(1.0 / (2.0 * sigmaSquared)) * l2);
            totalWeigths += weigth;
            scatter += Vec3(rad[3 * p], rad[3 * p + 1], rad[3 * p + 2]) * weigth;
        }
        if (totalWeigths > 0.0) scatter = scatter / totalWeigths;
        for (int c = 0; c < 3; ++c) o[c] = (float)scatter[c];
    }
}

// VolumeRender::operator() (Renderer.h:506-693), rows split over threads.
// films: 4 consecutive width*height RGBA float arrays (rgb, direct, scatter,
// emission).  scatter_rgb may be NULL (no multiple scattering term).
// Streams: the primary ray of (pix, AA sample s) draws its camera jitter and
// its steps from (pix, s, 2; seed).  At the shading point with ordinal q (the
// q-th sample that passed the cutoff on that ray), light k draws its shadow
// steps from (pix, q, 5 | k << 8 | s << 12; seed) and its env-map samples from
// (pix, q, 3 | k << 8 | s << 12; env_seed), light samples in order.  The
// reference draws all of these from one engine per TBB range (Renderer.h:518).
// lgo_render_rows renders image rows [row0, row1) only (the pixels and streams
// those rows have in the full image); films then hold 4 consecutive
// width*(row1-row0) RGBA arrays.
void lgo_render_rows(const lg_grid_desc *d, const float *dense, const float *temp_dense, const float *scatter_rgb,
                     const int32_t sn[3], const int32_t si0[3], double scv, const lg_camera *cam, const lg_shader *sh,
                     const lg_light *lights, int32_t nl, const float *const *env_rgb, const int32_t *env_w,
                     const int32_t *env_h, const lg_march_params *mp, int32_t samples, uint64_t seed,
                     uint64_t env_seed, int32_t row0, int32_t row1, int32_t threads, float *films) {
    Grid g(*d, dense);
    std::unique_ptr<Grid> tg;
    if (temp_dense) tg.reset(new Grid(*d, temp_dense));
    Scene sc = makeScene(&g, tg.get(), sh, mp, lights, nl, env_rgb, env_w, env_h);
    Coarse coarse{};
    if (scatter_rgb) {
        for (int a = 0; a < 3; ++a) {
            coarse.n[a] = sn[a];
            coarse.i0[a] = si0[a];
        }
        coarse.cv = scv;
        coarse.rgb = scatter_rgb;
    }
    const Camera camera{cam};
    const int W = cam->width;
    row0 = std::max(row0, 0);
    row1 = std::min(row1, (int32_t)cam->height);
    if (row1 <= row0) return;
    const int H = row1 - row0;
    float *rgbF = films, *dirF = films + 4 * (size_t)W * H, *scaF = films + 8 * (size_t)W * H,
          *emiF = films + 12 * (size_t)W * H;
    const Vec3 One(1.0);
    const Vec3 invScatter = Vec3(1.0) / Vec3(sh->scattering[0], sh->scattering[1], sh->scattering[2]);
    const Vec3 ext = sc.sh.ext(), albedo = sc.sh.albedo();
    const double jitterAmp = 0.5, cutoff = mp->cutoff, h0 = mp->step[0], h1 = mp->step[1];
    auto rows = [&](int j0, int j1) {
        for (int j = j0; j < j1; ++j)
            for (int i = 0; i < W; ++i) {
                const uint32_t pix = (uint32_t)j * (uint32_t)W + (uint32_t)i;
                Vec3 pTransAA(0.0), pLumiAA(0.0), directLumiAA(0.0), scatterLumiAA(0.0), emitLumiAA(0.0);
                const double ratioAA = 1.0 / (double)samples;
                for (int sampleAA = 0; sampleAA < samples; sampleAA++) {
                    CounterStream rng(pix, (uint32_t)sampleAA, 2u, seed);
                    uint32_t ordinal = 0;  // shading points taken so far on this ray
                    const Ray pRay = camera.getRay(i, j, rng);
                    Vec3 pTrans(1.0);
                    Intersector primary(g), shadow(g);
                    if (!primary.hitGridBbox(pRay)) {
                        pTransAA += pTrans * ratioAA;
                        continue;
                    }
                    Vec3 pLumi(0.0), directLumi(0.0), scatterLumi(0.0), emitLumi(0.0);
                    double pT0, pT1;
                    while (primary.march(pT0, pT1)) {
                        bool pDropout = false;
                        double pStep = h0 * (1.0 + ((double)rng() - 0.5) * jitterAmp);
                        double pT = h0 * std::ceil(pT0 / h0);
                        for (; pT <= pT1; pT += pStep) {
                            const Vec3 ip = primary.indexPos(pT);
                            const Vec3 pPos(g.origin[0] + g.voxel * ip[0], g.origin[1] + g.voxel * ip[1],
                                            g.origin[2] + g.voxel * ip[2]);
                            const double pDensity = std::min((double)sampleIndex(g, ip) * sh->density_scale,
                                                             sh->density_max);
                            Vec3 pScatter(0.0);
                            if (scatter_rgb) pScatter = coarse.sample(pPos);
                            const Vec3 scatter = pScatter * invScatter *
                                                 std::pow(std::min(std::max(pDensity, 0.0), 1.0),
                                                          sh->scatter_density_mask_power) *
                                                 sh->scatter_scale;
                            const double pTemp = tg ? (double)sampleIndex(*tg, ip) : 0.0;
                            const Vec3 emission = sc.sh.emissionColor(pTemp) * sh->emission_scale;
                            if (pDensity < cutoff) continue;
                            const Vec3 dT = expv(ext * pDensity * pStep);
                            for (int k = 0; k < nl; k++) {
                                const uint32_t tag = ((uint32_t)sampleAA << 12) | ((uint32_t)k << 8);
                                CounterStream sRng(pix, ordinal, 5u | tag, seed);
                                CounterStream envRng(pix, ordinal, 3u | tag, env_seed);
                                Vec3 tmpLumi(0.0);
                                for (int s = 0; s < lights[k].samples; s++) {
                                    Vec3 sTrans(1.0), lightDir, lightIntensity;
                                    double distance;
                                    illuminate(lights[k], &sc.envs[k], pPos, &envRng, lightDir, lightIntensity, distance);
                                    const Ray sRay(pPos, -lightDir, kDelta, distance);
                                    if (!shadow.hitGridBbox(sRay)) continue;
                                    double sT0, sT1;
                                    while (shadow.march(sT0, sT1)) {
                                        bool sDropout = false;
                                        double sStep = h1 * (1.0 + ((double)sRng() - 0.5) * jitterAmp);
                                        double sT = h1 * std::ceil(sT0 / h1);
                                        for (; sT <= sT1; sT += sStep) {
                                            const double sDensity =
                                                std::min((double)sampleIndex(g, shadow.indexPos(sT)) *
                                                             sh->density_scale * sh->shadow_density_scale,
                                                         sh->density_max);
                                            if (sDensity < cutoff) continue;
                                            sTrans *= expv(ext * sDensity * sStep);
                                            if (sTrans.lengthSqr() < cutoff) {
                                                sTrans = Vec3(0.0);
                                                sDropout = true;
                                                break;
                                            }
                                            sStep = h1 * (1.0 + ((double)sRng() - 0.5) * jitterAmp);
                                        }
                                        if (sDropout) break;
                                    }
                                    tmpLumi += lightIntensity * sTrans;
                                }
                                const Vec3 tmpDirectLumi = albedo * (tmpLumi / (double)lights[k].samples) * pTrans * (One - dT);
                                pLumi += tmpDirectLumi;
                                directLumi += tmpDirectLumi;
                            }
                            const Vec3 tmpScatterLumi = scatter * pTrans * (One - dT);
                            const Vec3 tmpEmitLumi = emission * pTrans * (One - dT);
                            scatterLumi += tmpScatterLumi;
                            emitLumi += tmpEmitLumi;
                            pLumi += tmpScatterLumi + tmpEmitLumi;
                            ++ordinal;
                            pTrans *= dT;
                            if (pTrans.lengthSqr() < cutoff) {
                                pDropout = true;
                                break;
                            }
                            pStep = h0 * (1.0 + ((double)rng() - 0.5) * jitterAmp);
                        }
                        if (pDropout) break;
                    }
                    pLumiAA += pLumi * ratioAA;
                    directLumiAA += directLumi * ratioAA;
                    scatterLumiAA += scatterLumi * ratioAA;
                    emitLumiAA += emitLumi * ratioAA;
                    pTransAA += pTrans * ratioAA;
                }
                const size_t o = 4 * ((size_t)(j - row0) * W + (size_t)i);
                for (int c = 0; c < 3; ++c) {
                    rgbF[o + c] = (float)pLumiAA[c];
                    dirF[o + c] = (float)directLumiAA[c];
                    scaF[o + c] = (float)scatterLumiAA[c];
                    emiF[o + c] = (float)emitLumiAA[c];
                }
                rgbF[o + 3] = (float)(1.0f - pTransAA.sum() / 3.0f);
                dirF[o + 3] = scaF[o + 3] = emiF[o + 3] = 0.0f;
            }
    };
    if (threads <= 1) {
        rows(row0, row1);
    } else {
        // rows dealt round-robin: the cost of a row depends on what it crosses
        std::vector<std::thread> pool;
        for (int t = 0; t < threads && t < H; ++t)
            pool.emplace_back([&, t] {
                for (int j = row0 + t; j < row1; j += threads) rows(j, j + 1);
            });
        for (auto &th : pool) th.join();
    }
}

void lgo_render(const lg_grid_desc *d, const float *dense, const float *temp_dense, const float *scatter_rgb,
                const int32_t sn[3], const int32_t si0[3], double scv, const lg_camera *cam, const lg_shader *sh,
                const lg_light *lights, int32_t nl, const float *const *env_rgb, const int32_t *env_w,
                const int32_t *env_h, const lg_march_params *mp, int32_t samples, uint64_t seed, uint64_t env_seed,
                int32_t threads, float *films) {
    lgo_render_rows(d, dense, temp_dense, scatter_rgb, sn, si0, scv, cam, sh, lights, nl, env_rgb, env_w, env_h, mp,
                    samples, seed, env_seed, 0, cam->height, threads, films);
}

// computeVoxelSize restatement for the splat radius (see lg_splat_radius).
double lgo_point_voxel_size(const float *pts, int64_t count) {
    if (count <= 1) return (double)0.1f;
    int64_t target = std::max<int64_t>(count / 8, 1);
    Vec3 lo(1e300), hi(-1e300);
    for (int64_t i = 0; i < count; ++i)
        for (int c = 0; c < 3; ++c) {
            lo[c] = std::min(lo[c], (double)pts[3 * i + c]);
            hi[c] = std::max(hi[c], (double)pts[3 * i + c]);
        }
    double e[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    if (e[0] <= 1e-7 && e[1] <= 1e-7 && e[2] <= 1e-7) return (double)0.1f;
    double volume = e[0] * e[1] * e[2];
    if (volume <= 1e-7) {
        std::sort(e, e + 3);
        volume = (e[1] <= 1e-7) ? e[2] * e[2] * e[2] : e[2] * e[2] * e[1];
    }
    float vs = (float)std::pow(std::max(volume / (double)target, 3e-15), 1.0 / 3.0);
    double prevVolume = volume;
    int64_t prevCount = 0, vcount = 1;
    while (vcount > prevCount) {
        std::vector<uint64_t> keys((size_t)count);
        const double inv = 1.0 / (double)vs;
        for (int64_t i = 0; i < count; ++i) {
            uint64_t key = 0;
            for (int c = 0; c < 3; ++c) {
                const int64_t q = (int64_t)std::floor((double)pts[3 * i + c] * inv + 0.5);
                key = key * 2097152ull + (uint64_t)((q + 1048576) & 2097151);
            }
            keys[(size_t)i] = key;
        }
        std::sort(keys.begin(), keys.end());
        prevCount = vcount;
        vcount = (int64_t)(std::unique(keys.begin(), keys.end()) - keys.begin());
        volume = (double)(vs * vs * vs * (float)vcount);
        if (volume >= prevVolume) break;
        prevVolume = volume;
        const float prev = vs;
        const double vv = volume / (double)target;
        if (vv < 3e-15) {
            vs = prev;
            break;
        }
        vs = (float)std::pow(vv, 1.0 / 3.0);
        if (vs / prev > 0.9f) break;
    }
    for (int dd = 5; dd < 11; ++dd) {
        const double t = std::pow(10.0, dd);
        const float tv = (float)(std::floor((double)vs * t + 0.5) / t);
        if (tv != 0.0f) return (double)tv;
    }
    return (double)vs;
}

}  // extern "C"
