// Device-side building blocks shared by the march kernels: Philox streams, the
// dense-brick grid view, world->index rays, bbox clipping, the brick-span DDA and
// the trilinear sampler.
//
// Arithmetic contract: everything that decides *where* a sample is taken (ray
// set-up, clip, DDA span times, sample times, trilinear weights) runs in FP64 with
// the operation order of the reference's OpenVDB calls, and this file is compiled
// with -ffp-contract=off.  CDNA4 issues FP64 add/mul at the rate of unpacked FP32,
// so this costs little, and it makes every discrete decision of a march (span
// ends, t <= t1, density < cutoff) identical to an FP64 CPU evaluation.  Radiance
// and transmittance values are FP32.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace lgd {

// -------------------------------------------------------------- Philox ------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
// 1, 2, 3", SC'11): 10 rounds of the 4x32 S-box with multipliers 0xD2511F53 /
// 0xCD9E8D57 and Weyl key increments 0x9E3779B9 / 0xBB67AE85.
struct U4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
        const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// 24 high bits -> float in [0, 1), the range of uniform_real_distribution<float>(0,1).
__device__ __forceinline__ float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// A sequential stream: draw d of stream (a, b) is word d&3 of block
// philox((a, b, d>>2, domain), key).
struct Stream {
    uint32_t a, b, domain, k0, k1;
    uint32_t block;
    uint32_t used;  // words of the cached block already consumed (4 = empty)
    U4 cache;
    __device__ void init(uint32_t a_, uint32_t b_, uint32_t dom, uint64_t seed) {
        a = a_;
        b = b_;
        domain = dom;
        k0 = (uint32_t)seed;
        k1 = (uint32_t)(seed >> 32);
        block = 0;
        used = 4;
    }
    // position the stream at word w (as if w words had been drawn)
    __device__ void skip(uint32_t w) {
        block = w >> 2;
        used = 4;
        for (uint32_t r = 0; r < (w & 3u); ++r) (void)next();
    }
    __device__ float next() {
        if (used == 4) {
            cache = philox(U4{a, b, block, domain}, k0, k1);
            ++block;
            used = 0;
        }
        const uint32_t w = used == 0 ? cache.x : used == 1 ? cache.y : used == 2 ? cache.z : cache.w;
        ++used;
        return u01(w);
    }
};

// Stream domains (fourth counter word).  The render's shadow streams carry
// (AA sample << 12 | light << 8) above the domain byte.
enum : uint32_t { DOM_SAMPLE = 1, DOM_RENDER = 2, DOM_RENDER_ENV = 3, DOM_VERTEX_ENV = 4, DOM_RENDER_SHADOW = 5 };

// ----------------------------------------------------------- grid view ------
struct bf16_t {
    uint16_t v;
};

__device__ __forceinline__ float to_f32(float v) { return v; }
__device__ __forceinline__ float to_f32(_Float16 v) { return (float)v; }
__device__ __forceinline__ float to_f32(bf16_t v) { return __uint_as_float(((uint32_t)v.v) << 16); }

struct GridView {
    const void *data;
    const uint8_t *occ;
    int NBx, NBy, NBz;     // storage bricks per axis (with the 1-brick border)
    int nbx, nby, nbz;     // bricks per axis
    long long sy, sz;      // element stride of one brick step in y and z
    double bbmin[3];       // clip box, index space: [8*bmin, 8*(bmax+1)]
    double bbmax[3];
    double origin[3];
    double voxel, inv_voxel;
    int empty;             // 1: no active brick (every ray misses)
};

// Element offset of voxel (x,y,z); callers keep x in [-8, 8*nbx+7] etc.
__device__ __forceinline__ long long voxel_offset(const GridView &g, int x, int y, int z) {
    const long long brick = ((long long)((z >> 3) + 1) * g.NBy + ((y >> 3) + 1)) * g.NBx + ((x >> 3) + 1);
    return (brick << 9) + ((z & 7) << 6) + ((y & 7) << 3) + (x & 7);
}

// OpenVDB BoxSampler::trilinearInterpolation on a FloatGrid: each lerp is
// a + float(double(b - a) * w) with a, b float and w double.
__device__ __forceinline__ float box_lerp(float a, float b, double w) {
    const float d = b - a;
    const double t = (double)d * w;
    return a + (float)t;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// BoxSampler::sample at index-space position (x,y,z): background 0 outside the
// stored voxels.
template <typename T>
__device__ __forceinline__ float trilinear(const GridView &g, double x, double y, double z) {
    const double fx = floor(x), fy = floor(y), fz = floor(z);
    const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    const double u = x - (double)ix, v = y - (double)iy, w = z - (double)iz;
    // Out-of-range corners read the zero border: clamping keeps both taps of an
    // axis inside the border bricks, where every value is 0.
    const int cx = clampi(ix, -8, 8 * g.nbx + 6);
    const int cy = clampi(iy, -8, 8 * g.nby + 6);
    const int cz = clampi(iz, -8, 8 * g.nbz + 6);
    const T *d = static_cast<const T *>(g.data);
    const long long b = voxel_offset(g, cx, cy, cz);
    const long long dx = ((cx & 7) == 7) ? (512 - 7) : 1;
    const long long dy = ((cy & 7) == 7) ? (g.sy - 56) : 8;
    const long long dz = ((cz & 7) == 7) ? (g.sz - 448) : 64;
    const float d000 = to_f32(d[b]);
    const float d001 = to_f32(d[b + dz]);
    const float d010 = to_f32(d[b + dy]);
    const float d011 = to_f32(d[b + dy + dz]);
    const float d100 = to_f32(d[b + dx]);
    const float d101 = to_f32(d[b + dx + dz]);
    const float d110 = to_f32(d[b + dx + dy]);
    const float d111 = to_f32(d[b + dx + dy + dz]);
    const float c00 = box_lerp(d000, d001, w);
    const float c01 = box_lerp(d010, d011, w);
    const float c10 = box_lerp(d100, d101, w);
    const float c11 = box_lerp(d110, d111, w);
    const float c0 = box_lerp(c00, c01, v);
    const float c1 = box_lerp(c10, c11, v);
    return box_lerp(c0, c1, u);
}

__device__ __forceinline__ bool brick_exists(const GridView &g, int vx, int vy, int vz) {
    // vx.. are multiples of 8 (DDA cell origins)
    const int bx = (vx >> 3) + 1, by = (vy >> 3) + 1, bz = (vz >> 3) + 1;
    if (bx < 0 || by < 0 || bz < 0 || bx >= g.NBx || by >= g.NBy || bz >= g.NBz) return false;
    return g.occ[((long long)bz * g.NBy + by) * g.NBx + bx] != 0;
}

// ------------------------------------------------------------------ rays ----
// math::Ray<double> in index space plus the VolumeRayIntersector state.
struct Isect {
    double e[3];   // eye
    double d[3];   // unit direction
    double inv[3]; // 1/d
    double t0, t1; // current time span
    double tmax;   // end time after clipping
};

constexpr double kDeltaD = 1e-9;  // math::Delta<double>
constexpr double kDeltaF = 1e-5;  // math::Delta<float> (Ray::valid)
constexpr double kDblMax = 1.7976931348623157e308;

// Ray(eye_w, dir_w, t0, t1).worldToIndex(grid) followed by clip(mBBox)
// (VolumeRayIntersector::hitGridBbox, Renderer.h:341-343).  Returns false on a miss.
__device__ __forceinline__ bool isect_world_ray(const GridView &g, Isect &s, const double ew[3],
                                                const double dw[3], double t0w, double t1w) {
    if (g.empty) return false;
    double di[3];
    for (int c = 0; c < 3; ++c) {
        s.e[c] = (ew[c] - g.origin[c]) * g.inv_voxel;
        di[c] = dw[c] * g.inv_voxel;
    }
    const double len = sqrt(di[0] * di[0] + di[1] * di[1] + di[2] * di[2]);
    for (int c = 0; c < 3; ++c) {
        s.d[c] = di[c] / len;
        s.inv[c] = 1.0 / s.d[c];
    }
    double t0 = len * t0w, t1 = len * t1w;
    // Ray::intersects(bbox, t0, t1)
    for (int c = 0; c < 3; ++c) {
        double a = g.bbmin[c], b = g.bbmax[c];
        if (fabs(s.d[c]) <= 1e-15) {
            if (s.e[c] < a || s.e[c] > b) return false;
        } else {
            a = (a - s.e[c]) * s.inv[c];
            b = (b - s.e[c]) * s.inv[c];
            if (a > b) {
                const double tmp = a;
                a = b;
                b = tmp;
            }
            if (a > t0) t0 = a;
            if (b < t1) t1 = b;
            if (t0 > t1) return false;
        }
    }
    s.t0 = t0;
    s.t1 = t1;
    s.tmax = t1;
    return true;
}

// One VolumeRayIntersector::march (Renderer.h:345-369) over the brick level:
// the next span of consecutive existing 8^3 bricks along the ray.  Returns true
// with [t0, t1] when a span longer than Delta<double> was found.
__device__ __forceinline__ bool isect_march(const GridView &g, Isect &s, double &t0, double &t1) {
    double st0 = -1.0, st1 = -1.0;
    if (s.t1 - s.t0 > kDeltaF) {
        // DDA<Ray, 3>::init
        double mT0 = s.t0;
        const double mT1 = s.t1;
        double pos[3], next[3], delta[3];
        int vox[3], step[3];
        for (int c = 0; c < 3; ++c) {
            pos[c] = s.e[c] + s.d[c] * mT0;
            vox[c] = ((int)floor(pos[c])) & ~7;
        }
        for (int c = 0; c < 3; ++c) {
            if (s.d[c] == 0.0) {
                step[c] = 0;
                next[c] = kDblMax;
                delta[c] = kDblMax;
            } else if (s.inv[c] > 0) {
                step[c] = 8;
                next[c] = mT0 + ((double)(vox[c] + 8) - pos[c]) * s.inv[c];
                delta[c] = 8.0 * s.inv[c];
            } else {
                step[c] = -8;
                next[c] = mT0 + ((double)vox[c] - pos[c]) * s.inv[c];
                delta[c] = -8.0 * s.inv[c];
            }
        }
        for (;;) {
            if (brick_exists(g, vox[0], vox[1], vox[2])) {
                if (st0 < 0) st0 = mT0;
            } else if (st0 >= 0) {
                st1 = mT0;
                if (st1 - st0 > kDeltaD) break;
                st0 = st1 = -1.0;
            }
            // DDA::step with math::MinIndex
            const int key = ((next[0] < next[1]) << 2) + ((next[0] < next[2]) << 1) + (next[1] < next[2]);
            const int ax = (key >= 6) ? 0 : ((key == 1 || key == 3) ? 1 : 2);
            mT0 = next[ax];
            next[ax] += delta[ax];
            vox[ax] += step[ax];
            if (!(mT0 <= mT1)) {
                if (st0 >= 0) st1 = mT1;
                break;
            }
        }
    }
    if (st1 > 0) {
        s.t0 = st1 + kDeltaD;
        s.t1 = s.tmax;
    }
    t0 = st0;
    t1 = st1;
    return (st1 - st0) > kDeltaD;
}

// Threshold on optical depth equivalent to the reference's dropout test
// |T|^2 < cutoff with T_c = exp(-sigma_c * tau): dropout iff tau > tau_star.
// tau_star is solved on the host once per call (lg_dropout_tau).

}  // namespace lgd
