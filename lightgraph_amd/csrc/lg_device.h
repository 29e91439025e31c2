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
        // one 32 x 32 -> 64 product per multiplier (a single v_mad_u64_u32
        // instead of a mul_lo / mul_hi pair)
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
        const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
        const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
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

// Element e of a raw buffer resource over brick storage of type T (32-bit
// byte offsets from the uniform base).
template <typename T> __device__ __forceinline__ float buf_tap(__amdgpu_buffer_rsrc_t r, uint32_t e);
template <> __device__ __forceinline__ float buf_tap<float>(__amdgpu_buffer_rsrc_t r, uint32_t e) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, e * 4u, 0, 0));
}
template <> __device__ __forceinline__ float buf_tap<_Float16>(__amdgpu_buffer_rsrc_t r, uint32_t e) {
    return (float)__builtin_bit_cast(_Float16, __builtin_amdgcn_raw_buffer_load_b16(r, e * 2u, 0, 0));
}
template <> __device__ __forceinline__ float buf_tap<bf16_t>(__amdgpu_buffer_rsrc_t r, uint32_t e) {
    return __uint_as_float(((uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, e * 2u, 0, 0)) << 16);
}

struct GridView {
    const void *data;
    const uint8_t *occ;
    const uint8_t *skip;   // per storage brick: radius of the uniform box around it
    int NBx, NBy, NBz;     // storage bricks per axis (with the 1-brick border)
    int nbx, nby, nbz;     // bricks per axis
    long long sy, sz;      // element stride of one brick step in y and z
    double bbmin[3];       // clip box, index space: [8*bmin, 8*(bmax+1)]
    double bbmax[3];
    double origin[3];
    double voxel, inv_voxel;
    int empty;             // 1: no active brick (every ray misses)
    int off32;             // 1: the storage is below 4 GiB (taps use buffer loads)
    uint32_t data_bytes;   // storage bytes when off32
    uint8_t *touched;      // work counters only: per storage brick, set when a tap reads it
};

// Element offset of voxel (x,y,z); callers keep x in [-8, 8*nbx+7] etc.
__device__ __forceinline__ long long voxel_offset(const GridView &g, int x, int y, int z) {
    const long long brick = ((long long)((z >> 3) + 1) * g.NBy + ((y >> 3) + 1)) * g.NBx + ((x >> 3) + 1);
    return (brick << 9) + ((z & 7) << 6) + ((y & 7) << 3) + (x & 7);
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// A value every lane of the wave holds equally, moved to scalar registers (a
// wave-per-ray march keeps its ray's state there instead of in 64 VGPR copies).
__device__ __forceinline__ double wave_uniform(double x) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ int wave_uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// x of a source lane chosen by a DPP control word, both 32-bit halves moved by
// the row permute network (no LDS); lanes without a source, or outside
// row_mask, get 0.
template <int kCtrl, int kRowMask, typename V>
__device__ __forceinline__ V dpp_move(V x) {
    if constexpr (sizeof(V) == 8) {
        const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
        const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, kCtrl, kRowMask, 0xf, true);
        const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), kCtrl, kRowMask, 0xf, true);
        return __builtin_bit_cast(V, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
    } else {
        static_assert(sizeof(V) == 4, "32- or 64-bit values");
        return __builtin_bit_cast(
            V, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), kCtrl, kRowMask, 0xf, true));
    }
}

// Inclusive prefix sum of x over the 64 lanes in lane order: shifts by 1, 2, 4
// and 8 inside each row of 16 lanes, then the row totals carried across rows
// with the row broadcasts (lane 15 of a row to the next row, lane 31 to rows 2
// and 3).  Every step adds a value or an exact zero.
template <typename V>
__device__ __forceinline__ V wave_prefix_sum(V x) {
    V s = x;
    s += dpp_move<0x111, 0xf>(s);  // row_shr:1
    s += dpp_move<0x112, 0xf>(s);  // row_shr:2
    s += dpp_move<0x114, 0xf>(s);  // row_shr:4
    s += dpp_move<0x118, 0xf>(s);  // row_shr:8
    s += dpp_move<0x142, 0xa>(s);  // row_bcast:15 into rows 1 and 3
    s += dpp_move<0x143, 0xc>(s);  // row_bcast:31 into rows 2 and 3
    return s;
}

// Inclusive prefix maximum of non-negative ints over the 64 lanes in lane order
// (the same DPP steps; lanes without a source read 0, the identity here).
__device__ __forceinline__ int wave_prefix_max(int x) {
    int s = x;
    s = max(s, dpp_move<0x111, 0xf>(s));
    s = max(s, dpp_move<0x112, 0xf>(s));
    s = max(s, dpp_move<0x114, 0xf>(s));
    s = max(s, dpp_move<0x118, 0xf>(s));
    s = max(s, dpp_move<0x142, 0xa>(s));
    s = max(s, dpp_move<0x143, 0xc>(s));
    return s;
}

// Value of lane `lane` of the wave, in scalar registers.
template <typename V>
__device__ __forceinline__ V wave_lane64(V x, int lane) {
    static_assert(sizeof(V) == 8, "64-bit values");
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __builtin_bit_cast(V, ((unsigned long long)hi << 32) | lo);
}

// Work counters: flags the storage bricks that the 8 taps of a trilinear
// sample at index-space (x, y, z) read (clamped as trilinear clamps them).
__device__ __forceinline__ void mark_taps(const GridView &g, double x, double y, double z) {
    const int cx = clampi((int)floor(x), -8, 8 * g.nbx + 6);
    const int cy = clampi((int)floor(y), -8, 8 * g.nby + 6);
    const int cz = clampi((int)floor(z), -8, 8 * g.nbz + 6);
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const long long b = ((long long)(((cz + dz) >> 3) + 1) * g.NBy + (((cy + dy) >> 3) + 1)) * g.NBx +
                                    (((cx + dx) >> 3) + 1);
                g.touched[b] = 1;
            }
}

// OpenVDB BoxSampler::trilinearInterpolation on a FloatGrid: each lerp is
// a + float(double(b - a) * w) with a, b float and w double.
__device__ __forceinline__ float box_lerp(float a, float b, double w) {
    const float d = b - a;
    const double t = (double)d * w;
    return a + (float)t;
}


// BoxSampler::sample at index-space position (x,y,z): background 0 outside the
// stored voxels.  kBufferTaps: raw buffer loads with 32-bit offsets when the
// storage allows (faster for per-lane rays; the wave-per-ray light march,
// whose lanes tap neighbouring voxels, runs faster on 64-bit addresses).
template <typename T, bool kBufferTaps = true>
__device__ __forceinline__ float trilinear(const GridView &g, double x, double y, double z) {
    const double fx = floor(x), fy = floor(y), fz = floor(z);
    const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    // x - floor(x) is x - double(int(floor(x))) for |x| < 2^31, without the
    // conversions back
    const double u = x - fx, v = y - fy, w = z - fz;
    // Out-of-range corners read the zero border: clamping keeps both taps of an
    // axis inside the border bricks, where every value is 0.
    const int cx = clampi(ix, -8, 8 * g.nbx + 6);
    const int cy = clampi(iy, -8, 8 * g.nby + 6);
    const int cz = clampi(iz, -8, 8 * g.nbz + 6);
    float d000, d001, d010, d011, d100, d101, d110, d111;
    if (kBufferTaps && g.off32) {
        // Storage below 4 GiB: raw buffer loads with 32-bit offsets (stride 0,
        // num_records = storage bytes; 0x20000 is the gfx9 raw-buffer format word).
        const __amdgpu_buffer_rsrc_t r =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(g.data), (short)0, (int)g.data_bytes, 0x00020000);
        const uint32_t brick = ((uint32_t)((cz >> 3) + 1) * (uint32_t)g.NBy + (uint32_t)((cy >> 3) + 1)) *
                                   (uint32_t)g.NBx + (uint32_t)((cx >> 3) + 1);
        const uint32_t b = (brick << 9) + ((cz & 7) << 6) + ((cy & 7) << 3) + (cx & 7);
        const uint32_t dx = ((cx & 7) == 7) ? (512u - 7u) : 1u;
        const uint32_t dy = ((cy & 7) == 7) ? ((uint32_t)g.sy - 56u) : 8u;
        const uint32_t dz = ((cz & 7) == 7) ? ((uint32_t)g.sz - 448u) : 64u;
        auto at = [r](uint32_t e) { return buf_tap<T>(r, e); };
        d000 = at(b);
        d001 = at(b + dz);
        d010 = at(b + dy);
        d011 = at(b + dy + dz);
        d100 = at(b + dx);
        d101 = at(b + dx + dz);
        d110 = at(b + dx + dy);
        d111 = at(b + dx + dy + dz);
    } else {
        const T *d = static_cast<const T *>(g.data);
        const long long b = voxel_offset(g, cx, cy, cz);
        const long long dx = ((cx & 7) == 7) ? (512 - 7) : 1;
        const long long dy = ((cy & 7) == 7) ? (g.sy - 56) : 8;
        const long long dz = ((cz & 7) == 7) ? (g.sz - 448) : 64;
        d000 = to_f32(d[b]);
        d001 = to_f32(d[b + dz]);
        d010 = to_f32(d[b + dy]);
        d011 = to_f32(d[b + dy + dz]);
        d100 = to_f32(d[b + dx]);
        d101 = to_f32(d[b + dx + dz]);
        d110 = to_f32(d[b + dx + dy]);
        d111 = to_f32(d[b + dx + dy + dz]);
    }
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

// Per storage brick, the Chebyshev radius (in bricks) of the uniform box around
// it: an empty brick holds d in [1, kSkipCap], the distance to the nearest
// existing brick; an existing brick holds kSkipExists + d, d the distance to
// the nearest empty brick.  Every brick closer than d has the same state.
constexpr int kSkipCap = 16;
constexpr int kSkipExists = 128;
// Smallest d at which isect_march jumps over the box around a cell.
constexpr int kSkipMin = 2;

// Box code of the brick cell at (vx,vy,vz) (multiples of 8): cells outside the
// storage read 1 (empty, nothing known around them).
__device__ __forceinline__ int brick_skip(const GridView &g, int vx, int vy, int vz) {
    const int bx = (vx >> 3) + 1, by = (vy >> 3) + 1, bz = (vz >> 3) + 1;
    if (bx < 0 || by < 0 || bz < 0 || bx >= g.NBx || by >= g.NBy || bz >= g.NBz) return 1;
    return g.skip[((long long)bz * g.NBy + by) * g.NBx + bx];
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

// 1/x from v_rcp_f64 refined by two Newton steps (x finite and nonzero): within
// an ulp of the correctly rounded quotient, for a fraction of a division's cost.
__device__ __forceinline__ double rcp_newton(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// Ray(eye_w, dir_w, t0, t1).worldToIndex(grid) followed by clip(mBBox)
// (VolumeRayIntersector::hitGridBbox, Renderer.h:341-343).  Returns false on a miss.
// kRcp: the unit direction and its inverse from reciprocals (one per
// component and one of the length) instead of OpenVDB's four divisions; the
// times then move by at most an ulp, which the pixel tolerance absorbs.
template <bool kRcp = false>
__device__ __forceinline__ bool isect_world_ray(const GridView &g, Isect &s, const double ew[3],
                                                const double dw[3], double t0w, double t1w) {
    if (g.empty) return false;
    double di[3];
    for (int c = 0; c < 3; ++c) {
        s.e[c] = (ew[c] - g.origin[c]) * g.inv_voxel;
        di[c] = dw[c] * g.inv_voxel;
    }
    const double len = sqrt(di[0] * di[0] + di[1] * di[1] + di[2] * di[2]);
    if (kRcp) {
        const double rl = rcp_newton(len);
        for (int c = 0; c < 3; ++c) {
            s.d[c] = di[c] * rl;
            s.inv[c] = s.d[c] == 0.0 ? (signbit(s.d[c]) ? -INFINITY : INFINITY) : rcp_newton(s.d[c]);
        }
    } else {
        for (int c = 0; c < 3; ++c) {
            s.d[c] = di[c] / len;
            s.inv[c] = 1.0 / s.d[c];
        }
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

// DDA<Ray, 3> set-up at time t0 (math/DDA.h): the brick of the entry point, and
// per axis the step, the time of the first plane crossing and the time between
// crossings.  The m-th crossing of axis c is at next0[c] + m * delta[c]: a closed
// form (OpenVDB accumulates next += delta; the plane is the same), so that lanes
// can time different crossings independently.
__device__ __forceinline__ void dda_setup(const Isect &s, double t0, int vox[3], int step[3], double next0[3],
                                          double delta[3]) {
    double pos[3];
    for (int c = 0; c < 3; ++c) {
        pos[c] = s.e[c] + s.d[c] * t0;
        vox[c] = ((int)floor(pos[c])) & ~7;
    }
    for (int c = 0; c < 3; ++c) {
        if (s.d[c] == 0.0) {
            step[c] = 0;
            next0[c] = kDblMax;
            delta[c] = kDblMax;
        } else if (s.inv[c] > 0) {
            step[c] = 8;
            next0[c] = t0 + ((double)(vox[c] + 8) - pos[c]) * s.inv[c];
            delta[c] = 8.0 * s.inv[c];
        } else {
            step[c] = -8;
            next0[c] = t0 + ((double)vox[c] - pos[c]) * s.inv[c];
            delta[c] = -8.0 * s.inv[c];
        }
    }
}

// math::MinIndex: the axis whose crossing comes first (ties: z, then y, then x).
__device__ __forceinline__ int dda_min_axis(double n0, double n1, double n2) {
    const int key = ((n0 < n1) << 2) + ((n0 < n2) << 1) + (n1 < n2);
    return (key >= 6) ? 0 : ((key == 1 || key == 3) ? 1 : 2);
}

// One VolumeRayIntersector::march (Renderer.h:345-369) over the brick level:
// the next span of consecutive existing 8^3 bricks along the ray.  Returns true
// with [t0, t1] when a span longer than Delta<double> was found.
//
// The walk is DDA<Ray, 3>::step over the brick cells: crossing m of axis c is
// at next0[c] + m * delta[c] and MinIndex takes the first (ties z, y, x).
// Each iteration reads the box code of the current cell; when its uniform box
// has radius R = d - 1 >= kSkipMin - 1 the walk through the box only moves the
// DDA state (inside an empty box no span is open, inside an existing one the
// open span goes on), so every crossing up to and including the first that
// leaves the box is taken at once: along each axis that is crossing m + R, the
// exit is the first of the three in MinIndex order, and the other axes take
// the crossings that precede it (at most R each, counted by binary search on
// the closed form).  R = 0 is the single DDA step.  The crossings, cells and
// times are those of the cell-by-cell walk.  Every cell read has an entry time
// within the clip box, so it lies inside the storage (one border brick on each
// side): the brick index is kept as one linear offset, range-checked only as
// a guard.
__device__ __forceinline__ bool isect_march(const GridView &g, Isect &s, double &t0, double &t1) {
    double st0 = -1.0, st1 = -1.0;
    if (s.t1 - s.t0 > kDeltaF) {
        double mT0 = s.t0;
        const double mT1 = s.t1;
        double next0[3], delta[3];
        int vox[3], step[3];
        dda_setup(s, mT0, vox, step, next0, delta);
        const int sx = 1, sy = g.NBx, sz = g.NBx * g.NBy;
        const int nlin = sz * g.NBz;
        int lin = ((vox[2] >> 3) + 1) * sz + ((vox[1] >> 3) + 1) * sy + ((vox[0] >> 3) + 1);
        const int l0 = step[0] > 0 ? sx : (step[0] < 0 ? -sx : 0);
        const int l1 = step[1] > 0 ? sy : (step[1] < 0 ? -sy : 0);
        const int l2 = step[2] > 0 ? sz : (step[2] < 0 ? -sz : 0);
        int m0 = 0, m1 = 0, m2 = 0;
        for (;;) {
            const bool inside = (unsigned)lin < (unsigned)nlin;
            const int code = inside ? (int)g.skip[inside ? lin : 0] : 1;
            const bool exists = code >= kSkipExists;
            if (!exists && st0 >= 0 && mT0 - st0 > kDeltaD) {
                st1 = mT0;
                break;
            }
            st0 = exists ? (st0 < 0 ? mT0 : st0) : -1.0;
            const int d = exists ? code - kSkipExists : code;
            const int R = d >= kSkipMin ? d - 1 : 0;
            const double e0 = next0[0] + (double)(m0 + R) * delta[0];
            const double e1 = next0[1] + (double)(m1 + R) * delta[1];
            const double e2 = next0[2] + (double)(m2 + R) * delta[2];
            const int ax = dda_min_axis(e0, e1, e2);
            const double ex = ax == 0 ? e0 : (ax == 1 ? e1 : e2);
            mT0 = ex;
            if (!(ex <= mT1)) {  // the walk ends inside the box (or at this step)
                if (st0 >= 0) st1 = mT1;
                break;
            }
            int k0 = 0, k1 = 0, k2 = 0;
            if (R > 0) {
                auto taken = [&](int b, int m) {
                    int lo = 0, hi = R;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        const double c = next0[b] + (double)(m + mid) * delta[b];
                        if (c < ex || (c == ex && b > ax)) lo = mid + 1;
                        else hi = mid;
                    }
                    return lo;
                };
                k0 = ax == 0 ? 0 : taken(0, m0);
                k1 = ax == 1 ? 0 : taken(1, m1);
                k2 = ax == 2 ? 0 : taken(2, m2);
            }
            k0 += ax == 0 ? R + 1 : 0;
            k1 += ax == 1 ? R + 1 : 0;
            k2 += ax == 2 ? R + 1 : 0;
            m0 += k0;
            m1 += k1;
            m2 += k2;
            lin += __mul24(k0, l0) + __mul24(k1, l1) + __mul24(k2, l2);
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
