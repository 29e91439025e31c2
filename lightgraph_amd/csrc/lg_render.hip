// Splat (ScatterGrid.h), camera (Renderer.h:129-248), environment map CDFs
// (Light.h:132-176, Sampling.h:76-167) and the per-pixel march (kernel d,
// Renderer.h:506-693).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include <rocprim/device/device_scan.hpp>

#include "lg_internal.h"
#include "lg_lights.h"

namespace lg {

// ------------------------------------------------------------ env map ------
// Conditional CDFs, one lane per row, sequential FP64 prefix (Sampling.h:81-93
// with double accumulation).
__global__ void k_env_cond(const float *__restrict__ func, int w, int h, double *__restrict__ cond,
                           double *__restrict__ rowint) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= h) return;
    double *cdf = cond + (long long)v * (w + 1);
    const float *f = func + (long long)v * w;
    cdf[0] = 0.0;
    for (int i = 1; i <= w; ++i) cdf[i] = cdf[i - 1] + (double)f[i - 1] / (double)w;
    const double I = cdf[w];
    if (I == 0.0) {
        for (int i = 1; i <= w; ++i) cdf[i] = (double)i / (double)w;
    } else {
        for (int i = 1; i <= w; ++i) cdf[i] = cdf[i] / I;
    }
    rowint[v] = I;
}

__global__ void k_env_marg(const double *__restrict__ rowint, int h, double *__restrict__ marg,
                           double *__restrict__ total) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    marg[0] = 0.0;
    for (int i = 1; i <= h; ++i) marg[i] = marg[i - 1] + rowint[i - 1] / (double)h;
    const double I = marg[h];
    if (I == 0.0) {
        for (int i = 1; i <= h; ++i) marg[i] = (double)i / (double)h;
    } else {
        for (int i = 1; i <= h; ++i) marg[i] = marg[i] / I;
    }
    *total = I;
}

// ----------------------------------------------------------- scatter grid ----
struct CoarseView {
    int nx, ny, nz;
    int i0x, i0y, i0z;
    double cv, inv_cv;
    const float *rgb;
};

inline CoarseView coarse_view(const lg_scatter_grid *sg) {
    CoarseView c{};
    c.nx = sg->nx;
    c.ny = sg->ny;
    c.nz = sg->nz;
    c.i0x = sg->i0[0];
    c.i0y = sg->i0[1];
    c.i0z = sg->i0[2];
    c.cv = sg->voxel_size;
    c.inv_cv = 1.0 / sg->voxel_size;
    c.rgb = sg->d_rgb;
    return c;
}

__device__ __forceinline__ void coarse_tap(const CoarseView &c, int x, int y, int z, float v[3]) {
    x -= c.i0x;
    y -= c.i0y;
    z -= c.i0z;
    if (x < 0 || y < 0 || z < 0 || x >= c.nx || y >= c.ny || z >= c.nz) {
        v[0] = v[1] = v[2] = 0.0f;
        return;
    }
    const float *p = c.rgb + 3 * (((long long)z * c.ny + y) * c.nx + x);
    v[0] = p[0];
    v[1] = p[1];
    v[2] = p[2];
}

// BoxSampler on the Vec3f scatter grid at coarse index position (x, y, z).
__device__ __forceinline__ void coarse_sample(const CoarseView &c, double x, double y, double z, float out[3]) {
    const int ix = (int)floor(x), iy = (int)floor(y), iz = (int)floor(z);
    const double u = x - (double)ix, v = y - (double)iy, w = z - (double)iz;
    float d[2][2][2][3];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int e = 0; e < 2; ++e) coarse_tap(c, ix + a, iy + b, iz + e, d[a][b][e]);
    for (int ch = 0; ch < 3; ++ch) {
        const float c00 = lgd::box_lerp(d[0][0][0][ch], d[0][0][1][ch], w);
        const float c01 = lgd::box_lerp(d[0][1][0][ch], d[0][1][1][ch], w);
        const float c10 = lgd::box_lerp(d[1][0][0][ch], d[1][0][1][ch], w);
        const float c11 = lgd::box_lerp(d[1][1][0][ch], d[1][1][1][ch], w);
        out[ch] = lgd::box_lerp(lgd::box_lerp(c00, c01, v), lgd::box_lerp(c10, c11, v), u);
    }
}

// Coarse topology: a coarse voxel is active when a density voxel of its
// footprint (f voxels centred on it) is non-zero.
template <typename T>
__global__ void k_coarse_topology(lgd::GridView g, int nxd, int nyd, int nzd, int f, CoarseView c,
                                  uint8_t *__restrict__ act) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)c.nx * c.ny * c.nz;
    if (k >= total) return;
    const int x = (int)(k % c.nx), y = (int)((k / c.nx) % c.ny), z = (int)(k / ((long long)c.nx * c.ny));
    const int C[3] = {x + c.i0x, y + c.i0y, z + c.i0z};
    const int nd[3] = {nxd, nyd, nzd};
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        const double centre = ((double)C[a] * c.cv - g.origin[a]) * g.inv_voxel;
        lo[a] = (int)floor(centre - 0.5 * f + 0.5);
        hi[a] = lo[a] + f;
        lo[a] = lo[a] < 0 ? 0 : lo[a];
        hi[a] = hi[a] > nd[a] ? nd[a] : hi[a];
    }
    const T *d = static_cast<const T *>(g.data);
    uint8_t on = 0;
    for (int zz = lo[2]; zz < hi[2] && !on; ++zz)
        for (int yy = lo[1]; yy < hi[1] && !on; ++yy)
            for (int xx = lo[0]; xx < hi[0]; ++xx)
                if (lgd::to_f32(d[lgd::voxel_offset(g, xx, yy, zz)]) != 0.0f) {
                    on = 1;
                    break;
                }
    act[k] = on;
}

// dilateActiveValues(NN_FACE_EDGE_VERTEX, 1 iteration)
__global__ void k_dilate26(int nx, int ny, int nz, const uint8_t *__restrict__ in, uint8_t *__restrict__ out) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)nx * ny * nz;
    if (k >= total) return;
    const int x = (int)(k % nx), y = (int)((k / nx) % ny), z = (int)(k / ((long long)nx * ny));
    uint8_t on = 0;
    for (int dz = -1; dz <= 1 && !on; ++dz)
        for (int dy = -1; dy <= 1 && !on; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int xx = x + dx, yy = y + dy, zz = z + dz;
                if (xx < 0 || yy < 0 || zz < 0 || xx >= nx || yy >= ny || zz >= nz) continue;
                if (in[((long long)zz * ny + yy) * nx + xx]) {
                    on = 1;
                    break;
                }
            }
    out[k] = on;
}

// ------------------------------------------------------------------ splat ----
struct BinGrid {
    double lo[3];
    double inv_cs;
    int n[3];
};

__device__ __forceinline__ int bin_axis(const BinGrid &b, int a, double p) {
    int i = (int)floor((p - b.lo[a]) * b.inv_cs);
    return i < 0 ? 0 : (i >= b.n[a] ? b.n[a] - 1 : i);
}

__device__ __forceinline__ long long bin_of(const BinGrid &b, const float *p) {
    const int x = bin_axis(b, 0, p[0]), y = bin_axis(b, 1, p[1]), z = bin_axis(b, 2, p[2]);
    return ((long long)z * b.n[1] + y) * b.n[0] + x;
}

__global__ void k_bin_count(const float *__restrict__ pts, long long count, BinGrid b, int *__restrict__ cnt,
                            int *__restrict__ cell) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const long long c = bin_of(b, pts + 3 * i);
    cell[i] = (int)c;
    atomicAdd(&cnt[c], 1);
}

__global__ void k_bin_fill(long long count, const int *__restrict__ cell, const int *__restrict__ off,
                           int *__restrict__ cursor, int *__restrict__ idx) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int c = cell[i];
    idx[off[c] + atomicAdd(&cursor[c], 1)] = (int)i;
}

// ScatterGrid::operator() (ScatterGrid.h:61-98) for every active coarse voxel:
// Gaussian-weighted mean of the radiance of the points within radius r.
__global__ void k_splat(CoarseView c, const uint8_t *__restrict__ act, const float *__restrict__ pts,
                        const float *__restrict__ rad, BinGrid b, int reach, const int *__restrict__ off,
                        const int *__restrict__ idx, double r2, double inv2sig2, float *__restrict__ out) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)c.nx * c.ny * c.nz;
    if (k >= total) return;
    float *o = out + 3 * k;
    if (!act[k]) {
        o[0] = o[1] = o[2] = 0.0f;
        return;
    }
    const int x = (int)(k % c.nx), y = (int)((k / c.nx) % c.ny), z = (int)(k / ((long long)c.nx * c.ny));
    const double w[3] = {(double)(x + c.i0x) * c.cv, (double)(y + c.i0y) * c.cv, (double)(z + c.i0z) * c.cv};
    const int bx = bin_axis(b, 0, w[0]), by = bin_axis(b, 1, w[1]), bz = bin_axis(b, 2, w[2]);
    double tot = 0.0, s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int zz = max(0, bz - reach); zz <= min(b.n[2] - 1, bz + reach); ++zz)
        for (int yy = max(0, by - reach); yy <= min(b.n[1] - 1, by + reach); ++yy)
            for (int xx = max(0, bx - reach); xx <= min(b.n[0] - 1, bx + reach); ++xx) {
                const long long cell = ((long long)zz * b.n[1] + yy) * b.n[0] + xx;
                for (int q = off[cell]; q < off[cell + 1]; ++q) {
                    const int i = idx[q];
                    const double dx = w[0] - (double)pts[3 * i], dy = w[1] - (double)pts[3 * i + 1],
                                 dz = w[2] - (double)pts[3 * i + 2];
                    if (!(dx * dx + dy * dy + dz * dz < r2)) continue;
                    const float fx = (float)dx, fy = (float)dy, fz = (float)dz;
                    const float d2 = fx * fx + fy * fy + fz * fz;  // Vec3s::lengthSqr
                    const double wt = exp(-inv2sig2 * (double)d2);
                    tot += wt;
                    s0 += (double)rad[3 * i] * wt;
                    s1 += (double)rad[3 * i + 1] * wt;
                    s2 += (double)rad[3 * i + 2] * wt;
                }
            }
    if (tot > 0.0) {
        s0 /= tot;
        s1 /= tot;
        s2 /= tot;
    }
    o[0] = (float)s0;
    o[1] = (float)s1;
    o[2] = (float)s2;
}

// ------------------------------------------------------------------ render ---
struct RenderConsts {
    double ds, sds, dmax, cutoff, step0, step1, tau_star;
    double albedo[3], ext[3], inv_scat[3];
    double scatter_scale, mask_pow, emission_scale;
    double ramp_in, ramp_out;
    int ramp_n;
    double ramp_pos[LG_MAX_RAMP];
    double ramp_col[LG_MAX_RAMP][3];
    double M[16];  // screen to world
    double scale_w, scale_h, znear, zfar;
    int width, height, samples;
    double ratio_aa;
};

__device__ __forceinline__ void render_ramp(const RenderConsts &c, double x, double out[3]) {
    if (c.ramp_n == 0) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    double hp = (x - c.ramp_in) / (c.ramp_out - c.ramp_in);
    hp = (hp < 0.0) ? 0.0 : hp;
    hp = (1.0 < hp) ? 1.0 : hp;
    if (c.ramp_n == 1) {
        for (int k = 0; k < 3; ++k) out[k] = c.ramp_col[0][k];
        return;
    }
    int prev = 0, next = 0;
    for (int i = 0; i < c.ramp_n; ++i) {
        if (c.ramp_pos[i] <= hp) {
            prev = i;
        } else {
            next = i;
            break;
        }
    }
    const double range = c.ramp_pos[next] - c.ramp_pos[prev];
    const double blend = (hp - c.ramp_pos[prev]) / range;
    for (int k = 0; k < 3; ++k) out[k] = c.ramp_col[prev][k] * (1.0 - blend) + c.ramp_col[next][k] * blend;
}

// Kernel (d), the per-pixel march of Renderer.h:506-693, in phases over the
// (pixel, AA sample) rays, ray r = pixel * samples + sample:
//   count   one lane per ray marches the primary ray and counts its shading
//           points (samples that pass the cutoff, Renderer.h:595);
//   then per chunk of pixels:
//   fill    one lane per ray marches again and writes each shading point's
//           world position, ray and ordinal;
//   shadow  one lane per (light, shading point) runs that light's samples
//           (illuminate + shadow march, Renderer.h:609-643) and writes the sum
//           of intensity * transmittance;
//   combine one lane per ray marches again and shades every point in the
//           reference's order: direct term per light, scatter, emission
//           (Renderer.h:644-666), into the ray's four AOVs and transmittance;
//   aa      one lane per pixel sums its rays (Renderer.h:670-689).
// The shadow rays of a shading point draw from their own streams, so the
// primary march does not wait on them.  The points of each group of 64 rays
// (one wave of fill and combine) are stored together, ordinal-major: the k-th
// points of the group's rays in ray order, then the (k+1)-th, ...  A shadow wave
// then holds points at the same depth of neighbouring rays, whose shadow rays
// are alike, and reads and writes its points in order.
struct ShadePoint {
    double ip[3];  // index-space position (world: origin + voxel * ip)
    uint32_t ray;  // r = pixel * samples + sample
    uint32_t ord;  // ordinal of the shading point on its ray
};
// What the combine phase takes from the primary march at a shading point: the
// optical depth before the point, its density and the march step it was taken
// with (same slot as the ShadePoint).
struct ShadeExtra {
    double tau_before, pd, step;
};
struct RayOut {
    double lum[3], dl[3], sl[3], el[3];  // beauty, direct, scatter, emission
    double tr[3];                        // final transmittance
};

enum : int { RS_CAMERA_RAYS = 0, RS_PRIMARY_SAMPLES, RS_SHADING_POINTS, RS_SHADOW_RAYS, RS_SHADOW_SAMPLES,
             RS_SHADOW_LANE_SLOTS, RS_SHADOW_BRICKS, RS_COUNT };

__device__ __forceinline__ void render_stat(unsigned long long *stats, int counter, unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0 && v) stat_add(stats, counter, v);
}

// 64 x the largest v of the wave: the lane-iterations the wave spent on v
__device__ __forceinline__ void render_stat_wave_max(unsigned long long *stats, int counter, uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
    if ((threadIdx.x & 63) == 0 && v) stat_add(stats, counter, 64ull * v);
}

__device__ __forceinline__ uint32_t shadow_tag(uint32_t sample, int light) {
    return (sample << 12) | ((uint32_t)light << 8);
}

// PerspectiveCamera::getRay (Renderer.h:216-229) for pixel (i, j) with the
// jitter from rng, clipped to the grid (VolumeRayIntersector::setWorldRay).
__device__ __forceinline__ bool camera_ray(const lgd::GridView &g, const RenderConsts &c, int i, int j,
                                           lgd::Stream &rng, lgd::Isect &pr) {
    const double eye[3] = {c.M[12], c.M[13], c.M[14]};
    const double axis[3] = {0.0 * c.M[0] + 0.0 * c.M[4] + -1.0 * c.M[8], 0.0 * c.M[1] + 0.0 * c.M[5] + -1.0 * c.M[9],
                            0.0 * c.M[2] + 0.0 * c.M[6] + -1.0 * c.M[10]};
    const double ri = (double)i + 0.5 + ((double)rng.next() - 0.5);
    const double rj = (double)j + 0.5 + ((double)rng.next() - 0.5);
    const double sx = (2.0 * ri / (double)c.width - 1.0) * c.scale_w;
    const double sy = (1.0 - 2.0 * rj / (double)c.height) * c.scale_h;
    const double sz = -1.0;
    double dir[3];
    for (int a = 0; a < 3; ++a) dir[a] = sx * c.M[a] + sy * c.M[4 + a] + sz * c.M[8 + a];
    const double dl = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    if (!(fabs(dl - 0.0) <= 1e-7)) {
        const double inv = 1.0 / dl;
        for (int a = 0; a < 3; ++a) dir[a] = dir[a] * inv;
    }
    const double scale = 1.0 / (dir[0] * axis[0] + dir[1] * axis[1] + dir[2] * axis[2]);
    return lgd::isect_world_ray(g, pr, eye, dir, c.znear * scale, c.zfar * scale);
}

// The primary march (Renderer.h:548-687) without the shadow rays: f(ip, pd,
// step, tau) runs at every sample that passes the cutoff, tau being the optical
// depth before it; the step is redrawn at each span and after each such sample.
// Returns the final optical depth; *samples counts the trilinear samples.
template <typename T, typename F>
__device__ __forceinline__ double primary_march(const lgd::GridView &g, const RenderConsts &c, lgd::Isect &pr,
                                                lgd::Stream &rng, uint32_t *samples, F &&f) {
    double tau = 0.0;
    bool drop = false;
    double t0, t1;
    while (!drop && lgd::isect_march(g, pr, t0, t1)) {
        double step = c.step0 * (1.0 + ((double)rng.next() - 0.5) * 0.5);
        for (double t = c.step0 * ceil(t0 / c.step0); t <= t1; t += step) {
            const double ip[3] = {pr.e[0] + pr.d[0] * t, pr.e[1] + pr.d[1] * t, pr.e[2] + pr.d[2] * t};
            const float rho = lgd::trilinear<T>(g, ip[0], ip[1], ip[2]);
            ++*samples;
            double pd = (double)rho * c.ds;
            pd = (c.dmax < pd) ? c.dmax : pd;
            if (pd < c.cutoff) continue;
            f(ip, pd, step, tau);
            tau += pd * step;
            if (tau > c.tau_star) {
                drop = true;
                break;
            }
            step = c.step0 * (1.0 + ((double)rng.next() - 0.5) * 0.5);
        }
    }
    return tau;
}

__device__ __forceinline__ void ray_pixel(const RenderConsts &c, uint32_t r, uint32_t &pix, uint32_t &s, int &i,
                                          int &j) {
    pix = r / (uint32_t)c.samples;
    s = r - pix * (uint32_t)c.samples;
    j = (int)(pix / (uint32_t)c.width);
    i = (int)(pix - (uint32_t)j * (uint32_t)c.width);
}

// Phase count: shading points of every ray, and the camera-ray and
// primary-sample counters.
template <typename T>
__global__ void __launch_bounds__(256) k_render_count(lgd::GridView g, RenderConsts c, uint32_t k0, uint32_t k1,
                                                      uint32_t nrays, int *__restrict__ cnt,
                                                      unsigned long long *__restrict__ stats) {
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    uint32_t n = 0, smp = 0;
    if (r < nrays) {
        uint32_t pix, s;
        int i, j;
        ray_pixel(c, r, pix, s, i, j);
        lgd::Stream rng;
        rng.init(pix, s, lgd::DOM_RENDER, ((uint64_t)k1 << 32) | k0);
        lgd::Isect pr;
        if (camera_ray(g, c, i, j, rng, pr))
            (void)primary_march<T>(g, c, pr, rng, &smp, [&](const double *, double, double, double) { ++n; });
        cnt[r] = (int)n;
    }
    if (stats) {
        render_stat(stats, RS_CAMERA_RAYS, r < nrays ? 1 : 0);
        render_stat(stats, RS_PRIMARY_SAMPLES, smp);
    }
}

// Shading points of each block of 64 pixels.
__global__ void __launch_bounds__(256) k_render_block_sum(const int *__restrict__ cnt, uint32_t nrays,
                                                          uint32_t rays_per_block, uint32_t nblocks,
                                                          unsigned long long *__restrict__ bsum) {
    const uint32_t b = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (b >= nblocks) return;  // whole waves leave together
    const uint64_t r0 = (uint64_t)b * rays_per_block;
    const uint64_t r1 = std::min<uint64_t>(r0 + rays_per_block, nrays);
    unsigned long long v = 0;
    for (uint64_t r = r0 + lane; r < r1; r += 64) v += (unsigned long long)cnt[r];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) bsum[b] = v;
}

// Position of the k-th shading point of lane j's ray within its group of 64
// rays (counts c_i): S(k) + #{i < j : c_i > k}, S(k) = sum_i min(c_i, k), the
// points of ordinals below k.  A per-wave table holds S(k) and the ballot of
// c_i > k for k < kOrdTable; deeper ordinals are counted from the c_i.
constexpr int kOrdTable = 256;
struct OrdTable {
    unsigned long long mask[kOrdTable];
    int S[kOrdTable];
    int c[64];
};

// All 64 lanes of the wave call this with their count (0 for lanes past the
// last ray); the caller synchronises before ord_pos reads the table.
__device__ __forceinline__ void ord_table_build(OrdTable &t, int c) {
    const int lane = threadIdx.x & 63;
    int maxc = c;
    for (int o = 32; o > 0; o >>= 1) maxc = max(maxc, __shfl_xor(maxc, o, 64));
    t.c[lane] = c;
    const int kt = maxc < kOrdTable ? maxc : kOrdTable;
    int S = 0;
    for (int k = 0; k < kt; ++k) {
        const unsigned long long m = __ballot(c > k);
        if (lane == 0) {
            t.mask[k] = m;
            t.S[k] = S;
        }
        S += __popcll(m);
    }
}

__device__ __forceinline__ int ord_pos(const OrdTable &t, int k) {
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    if (k < kOrdTable) return t.S[k] + __popcll(t.mask[k] & below);
    int S = 0, rank = 0;
    for (int i = 0; i < 64; ++i) {
        const int ci = t.c[i];
        S += ci < k ? ci : k;
        rank += (i < lane && ci > k) ? 1 : 0;
    }
    return S + rank;
}

// Phase fill over rays [ray0, ray0 + nrays): the k-th shading point of ray r
// goes to group base + ord_pos(k), the group base being off[] of its first ray.
template <typename T>
__global__ void __launch_bounds__(256) k_render_fill(lgd::GridView g, RenderConsts c, uint32_t k0, uint32_t k1,
                                                     uint32_t ray0, uint32_t nrays, const int *__restrict__ cnt,
                                                     const int *__restrict__ off, ShadePoint *__restrict__ pts,
                                                     ShadeExtra *__restrict__ ext) {
    __shared__ OrdTable tab[4];
    const uint32_t rl = blockIdx.x * 256u + threadIdx.x;
    const uint32_t r = ray0 + rl;
    const bool valid = rl < nrays;
    const int qn = valid ? cnt[r] : 0;
    const int base = __shfl(valid ? off[rl] : 0, 0, 64);
    OrdTable &t = tab[threadIdx.x >> 6];
    ord_table_build(t, qn);
    __syncthreads();
    if (qn == 0) return;
    uint32_t pix, s;
    int i, j;
    ray_pixel(c, r, pix, s, i, j);
    lgd::Stream rng;
    rng.init(pix, s, lgd::DOM_RENDER, ((uint64_t)k1 << 32) | k0);
    lgd::Isect pr;
    if (!camera_ray(g, c, i, j, rng, pr)) return;
    uint32_t ord = 0, smp = 0;
    (void)primary_march<T>(g, c, pr, rng, &smp, [&](const double *ip, double pd, double step, double tau_before) {
        if ((int)ord < qn) {
            const int q = base + ord_pos(t, (int)ord);
            ShadePoint &p = pts[q];
            for (int k = 0; k < 3; ++k) p.ip[k] = ip[k];
            p.ray = r;
            p.ord = ord;
            ShadeExtra &e = ext[q];
            e.tau_before = tau_before;
            e.pd = pd;
            e.step = step;
        }
        ++ord;
    });
}

// Shadow march of the render (Renderer.h:615-643): jittered step redrawn at
// each span and after each sample that passed the cutoff.  Returns false when
// the ray misses the bbox; *tau_out receives the optical depth, *dropped the
// dropout (transmittance 0).
template <typename T, bool kCount>
__device__ __forceinline__ bool render_shadow(const lgd::GridView &g, const RenderConsts &c, const double p[3],
                                              const double ldir[3], double distance, lgd::Stream &rng,
                                              double *tau_out, bool *dropped, uint32_t *samples) {
    const double sd[3] = {-ldir[0], -ldir[1], -ldir[2]};
    lgd::Isect is;
    // per-ray set-up from reciprocals: 0.8% of the shadow phase, pixels unchanged
    if (!lgd::isect_world_ray<true>(g, is, p, sd, lgd::kDeltaD, distance)) return false;
    double tau = 0.0;
    double t0, t1;
    *dropped = false;
    while (lgd::isect_march(g, is, t0, t1)) {
        double step = c.step1 * (1.0 + ((double)rng.next() - 0.5) * 0.5);
        for (double t = c.step1 * ceil(t0 / c.step1); t <= t1; t += step) {
            const double x = is.e[0] + is.d[0] * t, y = is.e[1] + is.d[1] * t, z = is.e[2] + is.d[2] * t;
            if (kCount) lgd::mark_taps(g, x, y, z);
            const float rho = lgd::trilinear<T>(g, x, y, z);
            ++*samples;
            double dens = (double)rho * c.ds * c.sds;
            dens = (c.dmax < dens) ? c.dmax : dens;
            if (dens < c.cutoff) continue;
            tau += dens * step;
            if (tau > c.tau_star) {
                *dropped = true;
                *tau_out = tau;
                return true;
            }
            step = c.step1 * (1.0 + ((double)rng.next() - 0.5) * 0.5);
        }
    }
    *tau_out = tau;
    return true;
}

// Phase shadow: task t = light * P_pad + point (P_pad a multiple of 64, so a
// wave holds one light); out[3 t ..] = the sum over the light's samples of
// intensity * transmittance (Renderer.h:609-643).  The intensity of the sample
// being marched and the running sum wait in LDS during the march, which leaves
// registers for five waves per SIMD.
template <typename T, bool kCount>
__global__ void __launch_bounds__(256, 5) k_render_shadow(lgd::GridView g, RenderConsts c, lgl::DevLights L,
                                                       uint32_t k0, uint32_t k1, uint32_t ek0, uint32_t ek1,
                                                       const ShadePoint *__restrict__ pts, uint32_t P,
                                                       uint32_t P_pad, double *__restrict__ out,
                                                       unsigned long long *__restrict__ stats) {
    __shared__ double s_inten[3 * 256], s_sum[3 * 256];
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const int li = __builtin_amdgcn_readfirstlane((int)(t / P_pad));
    const uint32_t q = t - (uint32_t)li * P_pad;
    double *const inten = s_inten + 3 * threadIdx.x;
    double *const sum = s_sum + 3 * threadIdx.x;
    uint32_t rays = 0, smp = 0;
    if (li < L.n && q < P) {
        const lgl::DevLight &lt = L.l[li];
        const uint32_t ray = pts[q].ray, ord = pts[q].ord;
        const uint32_t pix = ray / (uint32_t)c.samples;
        const uint32_t tag = shadow_tag(ray - pix * (uint32_t)c.samples, li);
        lgd::Stream rng;
        rng.init(pix, ord, lgd::DOM_RENDER_SHADOW | tag, ((uint64_t)k1 << 32) | k0);
        for (int a = 0; a < 3; ++a) sum[a] = 0.0;
        for (int ls = 0; ls < lt.samples; ++ls) {
            // the position is re-read each light sample and kept opaque, so that
            // the compiler does not hoist the ray set-up's per-point values out
            // of this loop and hold them (in spills) across the march
            double ip[3] = {pts[q].ip[0], pts[q].ip[1], pts[q].ip[2]};
            asm volatile("" : "+v"(ip[0]), "+v"(ip[1]), "+v"(ip[2]));
            const double wp[3] = {g.origin[0] + g.voxel * ip[0], g.origin[1] + g.voxel * ip[1],
                                  g.origin[2] + g.voxel * ip[2]};
            double ldir[3], dist;
            {
                // an env sample draws 2 values: sample ls starts at draw 2 ls
                lgd::Stream env;
                env.init(pix, ord, lgd::DOM_RENDER_ENV | tag, ((uint64_t)ek1 << 32) | ek0);
                if (lt.type == LG_LIGHT_ENV && ls > 0) env.skip(2u * (uint32_t)ls);
                double in[3];
                lgl::illuminate<true>(lt, wp, env, ldir, in, dist);
                for (int a = 0; a < 3; ++a) inten[a] = in[a];
            }
            double tau_s;
            bool sdrop;
            ++rays;
            if (!render_shadow<T, kCount>(g, c, wp, ldir, dist, rng, &tau_s, &sdrop, &smp)) continue;
            if (sdrop) continue;
            for (int a = 0; a < 3; ++a) sum[a] += inten[a] * (double)expf((float)(c.ext[a] * tau_s));
        }
        double *o = out + 3ull * t;
        for (int a = 0; a < 3; ++a) o[a] = sum[a];
    }
    if (stats) {
        render_stat(stats, RS_SHADOW_RAYS, rays);
        render_stat(stats, RS_SHADOW_SAMPLES, smp);
        render_stat_wave_max(stats, RS_SHADOW_LANE_SLOTS, smp);
    }
}

// Phase combine over rays [ray0, ray0 + nrays) (Renderer.h:580-668): the
// shading points of each ray in march order, from the records the fill phase
// stored (no second march).  Per point: transmittance to the point and the
// absorbed fraction of its step, the light sums of the shadow phase, the
// scatter-grid and emission terms; the optical depth after the last point is
// the ray's (the march adds pd * step after each point and stops there at a
// dropout).
__global__ void __launch_bounds__(256) k_render_combine(lgd::GridView g, lgd::GridView tg, int has_temp,
                                                        CoarseView sc, int has_scatter, RenderConsts c,
                                                        lgl::DevLights L, uint32_t ray0, uint32_t nrays,
                                                        const int *__restrict__ cnt, const int *__restrict__ off,
                                                        const ShadePoint *__restrict__ pts,
                                                        const ShadeExtra *__restrict__ ext,
                                                        const double *__restrict__ sums, uint32_t P_pad,
                                                        RayOut *__restrict__ rout,
                                                        unsigned long long *__restrict__ stats) {
    __shared__ OrdTable tab[4];
    const uint32_t rl = blockIdx.x * 256u + threadIdx.x;
    const bool valid = rl < nrays;
    const int qn = valid ? cnt[ray0 + rl] : 0;
    const int base = __shfl(valid ? off[rl] : 0, 0, 64);
    OrdTable &tb = tab[threadIdx.x >> 6];
    ord_table_build(tb, qn);
    __syncthreads();
    if (valid) {
        RayOut o{};
        double tau = 0.0;
        for (int ord = 0; ord < qn; ++ord) {
            const uint32_t q = (uint32_t)(base + ord_pos(tb, ord));
            const ShadeExtra e = ext[q];
            const double ip[3] = {pts[q].ip[0], pts[q].ip[1], pts[q].ip[2]};
            const double pd = e.pd, step = e.step, tau_before = e.tau_before;
            const double wp[3] = {g.origin[0] + g.voxel * ip[0], g.origin[1] + g.voxel * ip[1],
                                  g.origin[2] + g.voxel * ip[2]};
            double tf[3], om[3];
            for (int k = 0; k < 3; ++k) {
                tf[k] = (double)expf((float)(c.ext[k] * tau_before));
                om[k] = (double)(-expm1f((float)(c.ext[k] * pd * step)));
            }
            for (int li = 0; li < L.n; ++li) {
                const double *tl = sums + 3ull * ((uint64_t)li * P_pad + q);
                double a[3];
                const double inv = L.l[li].inv_samples;
                if (inv != 0.0) {
                    // x * 2^-k is x / 2^k exactly (both correctly rounded)
                    for (int k = 0; k < 3; ++k) a[k] = tl[k] * inv;
                } else {
                    const double ns = (double)L.l[li].samples;
                    for (int k = 0; k < 3; ++k) a[k] = tl[k] / ns;
                }
                for (int k = 0; k < 3; ++k) {
                    const double v = c.albedo[k] * a[k] * tf[k] * om[k];
                    o.lum[k] += v;
                    o.dl[k] += v;
                }
            }
            double msc[3] = {0, 0, 0};
            if (has_scatter) {
                float sv[3];
                coarse_sample(sc, wp[0] * sc.inv_cv, wp[1] * sc.inv_cv, wp[2] * sc.inv_cv, sv);
                const double pc = pd < 0.0 ? 0.0 : (pd > 1.0 ? 1.0 : pd);
                const double mask = (double)powf((float)pc, (float)c.mask_pow);
                for (int k = 0; k < 3; ++k) msc[k] = (double)sv[k] * c.inv_scat[k] * mask * c.scatter_scale;
            }
            double emi[3] = {0, 0, 0};
            if (c.ramp_n > 0) {
                double temp = 0.0;
                if (has_temp) temp = (double)lgd::trilinear<float>(tg, ip[0], ip[1], ip[2]);
                render_ramp(c, temp, emi);
                for (int k = 0; k < 3; ++k) emi[k] *= c.emission_scale;
            }
            for (int k = 0; k < 3; ++k) {
                const double ts = msc[k] * tf[k] * om[k];
                const double te = emi[k] * tf[k] * om[k];
                o.sl[k] += ts;
                o.el[k] += te;
                o.lum[k] += ts + te;
            }
            tau = tau_before + pd * step;
        }
        for (int k = 0; k < 3; ++k) o.tr[k] = (double)expf((float)(c.ext[k] * tau));  // kept at dropout
        rout[rl] = o;
    }
    if (stats) render_stat(stats, RS_SHADING_POINTS, (unsigned long long)qn);
}

// Phase aa over pixels [pix0, pix0 + npix): AA samples summed in order.
__global__ void __launch_bounds__(256) k_render_aa(RenderConsts c, uint32_t pix0, uint32_t npix,
                                                   const RayOut *__restrict__ rout, float4 *__restrict__ rgb,
                                                   float4 *__restrict__ direct, float4 *__restrict__ scatter,
                                                   float4 *__restrict__ emission) {
    const uint32_t pl = blockIdx.x * 256u + threadIdx.x;
    if (pl >= npix) return;
    double lum[3] = {0, 0, 0}, dl[3] = {0, 0, 0}, sl[3] = {0, 0, 0}, el[3] = {0, 0, 0}, tr[3] = {0, 0, 0};
    const RayOut *ro = rout + (uint64_t)pl * (uint32_t)c.samples;
    for (int s = 0; s < c.samples; ++s) {
        const RayOut &o = ro[s];
        for (int k = 0; k < 3; ++k) {
            lum[k] += o.lum[k] * c.ratio_aa;
            dl[k] += o.dl[k] * c.ratio_aa;
            sl[k] += o.sl[k] * c.ratio_aa;
            el[k] += o.el[k] * c.ratio_aa;
            tr[k] += o.tr[k] * c.ratio_aa;
        }
    }
    const uint32_t pix = pix0 + pl;
    const float alpha = (float)(1.0f - (tr[0] + tr[1] + tr[2]) / 3.0f);
    if (rgb) rgb[pix] = make_float4((float)lum[0], (float)lum[1], (float)lum[2], alpha);
    if (direct) direct[pix] = make_float4((float)dl[0], (float)dl[1], (float)dl[2], 0.0f);
    if (scatter) scatter[pix] = make_float4((float)sl[0], (float)sl[1], (float)sl[2], 0.0f);
    if (emission) emission[pix] = make_float4((float)el[0], (float)el[1], (float)el[2], 0.0f);
}

// Render workspace.  Slot 0 holds the shading-point count of every ray and the
// per-block sums; slot 1 holds one chunk: ray offsets, scan storage, ray
// outputs, shading points and their light sums.  A chunk is a run of whole
// blocks of 64 pixels with at most kRenderRayCap rays and about
// kRenderPointBytes of shading-point data (a single block may exceed the latter).
constexpr uint32_t kRenderBlockPixels = 64;
constexpr uint64_t kRenderRayCap = 1u << 20;
constexpr uint64_t kRenderPointBytes = 2ull << 30;

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

struct RenderLayout {
    uint64_t npix, nrays, nblocks, rays_per_block, ray_cap, point_cap;
    size_t bsum_off, head_bytes;                          // slot 0
    size_t scan_off, scan_bytes, out_off, ray_area_bytes;  // slot 1, before the points
    size_t point_bytes;                                   // per shading point: record and light sums
};

inline RenderLayout render_layout(int w, int h, int samples, int n_lights) {
    RenderLayout l{};
    l.npix = (uint64_t)w * (uint64_t)h;
    l.nrays = l.npix * (uint64_t)samples;
    l.nblocks = (l.npix + kRenderBlockPixels - 1) / kRenderBlockPixels;
    l.rays_per_block = (uint64_t)kRenderBlockPixels * (uint64_t)samples;
    l.ray_cap = std::min(std::max(kRenderRayCap, l.rays_per_block), l.nblocks * l.rays_per_block);
    l.point_bytes = sizeof(ShadePoint) + sizeof(ShadeExtra) + 24ull * (uint64_t)n_lights;
    l.point_cap = kRenderPointBytes / l.point_bytes;
    l.bsum_off = align256(4 * l.nrays);
    l.head_bytes = l.bsum_off + align256(8 * l.nblocks);
    size_t sb = 0;
    (void)rocprim::exclusive_scan(nullptr, sb, (const int *)nullptr, (int *)nullptr, 0, (size_t)l.ray_cap,
                                  rocprim::plus<int>());
    l.scan_bytes = sb;
    l.scan_off = align256(4 * l.ray_cap);
    l.out_off = l.scan_off + align256(sb);
    l.ray_area_bytes = l.out_off + align256(sizeof(RayOut) * l.ray_cap);
    return l;
}

template <typename F>
static void for_dtype_r(int dtype, F &&f) {
    if (dtype == LG_F32)
        f(float{});
    else if (dtype == LG_F16)
        f(_Float16{});
    else
        f(lgd::bf16_t{});
}

}  // namespace lg

using namespace lg;

extern "C" {

int32_t lg_envmap_create(lg_ctx *ctx, int32_t w, int32_t h, const float *h_rgb, lg_envmap **out) {
    LG_CHECK_ARG(ctx && h_rgb && out && w >= 2 && h >= 2, "lg_envmap_create: bad argument");
    *out = nullptr;
    // luminance * sin(theta) weights in float (Light.h:165-176)
    std::vector<float> Y((size_t)w * h);
    for (int i = 0; i < h; ++i) {
        const float sin_theta = (float)std::sin(M_PI * (double)((float)i / (float)(h - 1)));
        for (int j = 0; j < w; ++j) {
            const float *p = h_rgb + 3 * ((size_t)i * w + j);
            float y = (float)(0.2126 * (double)p[0] + 0.7152 * (double)p[1] + 0.0722 * (double)p[2]);
            y *= sin_theta;
            Y[(size_t)i * w + j] = y;
        }
    }
    lg_envmap *e = new lg_envmap();
    e->width = w;
    e->height = h;
    float *func = nullptr;
    double *total = nullptr;
    int32_t rc = dalloc(ctx, (size_t)w * h * 3, &e->rgb);
    if (rc == LG_OK) rc = dalloc(ctx, (size_t)h * (w + 1), &e->cond);
    if (rc == LG_OK) rc = dalloc(ctx, (size_t)h + 1, &e->marg);
    if (rc == LG_OK) rc = dalloc(ctx, (size_t)h, &e->rowint);
    if (rc == LG_OK) rc = dalloc(ctx, (size_t)w * h, &func);
    if (rc == LG_OK) rc = dalloc(ctx, 1, &total);
    if (rc == LG_OK) {
        hipError_t err = hipMemcpyAsync(e->rgb, h_rgb, (size_t)w * h * 3 * 4, hipMemcpyHostToDevice, ctx->stream);
        if (err == hipSuccess) err = hipMemcpyAsync(func, Y.data(), Y.size() * 4, hipMemcpyHostToDevice, ctx->stream);
        if (err == hipSuccess) {
            k_env_cond<<<(h + 63) / 64, 64, 0, ctx->stream>>>(func, w, h, e->cond, e->rowint);
            k_env_marg<<<1, 64, 0, ctx->stream>>>(e->rowint, h, e->marg, total);
            err = hipGetLastError();
        }
        if (err == hipSuccess) err = hipMemcpyAsync(&e->total, total, 8, hipMemcpyDeviceToHost, ctx->stream);
        if (err == hipSuccess) err = hipStreamSynchronize(ctx->stream);
        if (err != hipSuccess) {
            set_error("lg_envmap_create: %s", hipGetErrorString(err));
            rc = LG_ERR_HIP;
        }
    }
    dfree(ctx, func);
    dfree(ctx, total);
    if (rc != LG_OK) {
        lg_envmap_destroy(ctx, e);
        return rc;
    }
    *out = e;
    return LG_OK;
}

int32_t lg_envmap_destroy(lg_ctx *ctx, lg_envmap *e) {
    if (!e) return LG_OK;
    dfree(ctx, e->rgb);
    dfree(ctx, e->cond);
    dfree(ctx, e->marg);
    dfree(ctx, e->rowint);
    delete e;
    return LG_OK;
}

int32_t lg_envmap_read_cdf(lg_ctx *ctx, const lg_envmap *e, double *cond, double *marg, double *rowint,
                           double *total) {
    LG_CHECK_ARG(ctx && e, "lg_envmap_read_cdf: NULL argument");
    if (cond) LG_HIP(hipMemcpyAsync(cond, e->cond, (size_t)e->height * (e->width + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (marg) LG_HIP(hipMemcpyAsync(marg, e->marg, ((size_t)e->height + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (rowint) LG_HIP(hipMemcpyAsync(rowint, e->rowint, (size_t)e->height * 8, hipMemcpyDeviceToHost, ctx->stream));
    LG_HIP(hipStreamSynchronize(ctx->stream));
    if (total) *total = e->total;
    return LG_OK;
}

int32_t lg_scatter_grid_create(lg_ctx *ctx, const lg_grid *density, int32_t res_factor, lg_scatter_grid *sg) {
    LG_CHECK_ARG(ctx && density && sg && res_factor >= 1, "lg_scatter_grid_create: bad argument");
    memset(sg, 0, sizeof(*sg));
    const lg_grid_desc &d = density->desc;
    const double cv = d.voxel_size * (double)res_factor;  // ScatterGrid.h:309-310
    const int nd[3] = {d.nx, d.ny, d.nz};
    int n[3], i0[3];
    for (int a = 0; a < 3; ++a) {
        const double wlo = d.origin[a] - 0.5 * d.voxel_size, whi = d.origin[a] + d.voxel_size * (nd[a] - 0.5);
        i0[a] = (int)floor(wlo / cv) - 2;
        n[a] = (int)ceil(whi / cv) + 2 - i0[a] + 1;
    }
    sg->nx = n[0];
    sg->ny = n[1];
    sg->nz = n[2];
    for (int a = 0; a < 3; ++a) sg->i0[a] = i0[a];
    sg->voxel_size = cv;
    const size_t vox = (size_t)n[0] * n[1] * n[2];
    uint8_t *tmp = nullptr;
    LG_TRY(dalloc(ctx, vox * 3, &sg->d_rgb));
    int32_t rc = dalloc(ctx, vox, &sg->d_active);
    if (rc == LG_OK) rc = dalloc(ctx, vox, &tmp);
    if (rc == LG_OK) {
        LG_HIP(hipMemsetAsync(sg->d_rgb, 0, vox * 12, ctx->stream));
        const CoarseView c = coarse_view(sg);
        const lgd::GridView g = make_view(density);
        const unsigned blocks = (unsigned)((vox + 255) / 256);
        for_dtype_r(d.dtype, [&](auto tag) {
            using T = decltype(tag);
            k_coarse_topology<T><<<blocks, 256, 0, ctx->stream>>>(g, d.nx, d.ny, d.nz, res_factor, c, tmp);
        });
        k_dilate26<<<blocks, 256, 0, ctx->stream>>>(n[0], n[1], n[2], tmp, sg->d_active);
        if (hipGetLastError() != hipSuccess) {
            set_error("scatter grid topology kernels failed to launch");
            rc = LG_ERR_HIP;
        }
    }
    dfree(ctx, tmp);
    if (rc != LG_OK) lg_scatter_grid_destroy(ctx, sg);
    return rc;
}

int32_t lg_scatter_grid_destroy(lg_ctx *ctx, lg_scatter_grid *sg) {
    if (!sg) return LG_OK;
    dfree(ctx, sg->d_rgb);
    dfree(ctx, sg->d_active);
    sg->d_rgb = nullptr;
    sg->d_active = nullptr;
    return LG_OK;
}

int32_t lg_splat(lg_ctx *ctx, const float *d_pts, const float *d_rad, int64_t count, double radius,
                 lg_scatter_grid *sg) {
    LG_CHECK_ARG(ctx && sg && sg->d_rgb && sg->d_active && count >= 0, "lg_splat: bad argument");
    LG_CHECK_ARG(count == 0 || (d_pts && d_rad), "lg_splat: NULL points");
    LG_CHECK_ARG(radius > 0.0 && count < (1ll << 31), "lg_splat: radius must be positive");
    const CoarseView c = coarse_view(sg);
    // bins of edge >= radius over the coarse lattice's world extent
    BinGrid b{};
    const double cs = std::max(radius, 1e-6);
    b.inv_cs = 1.0 / cs;
    const int ci0[3] = {c.i0x, c.i0y, c.i0z}, cn[3] = {c.nx, c.ny, c.nz};
    long long nbins = 1;
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = (double)ci0[a] * c.cv;
        const double extent = (double)(cn[a] - 1) * c.cv;
        b.n[a] = std::max(1, (int)ceil(extent / cs) + 1);
        nbins *= b.n[a];
    }
    LG_CHECK_ARG(nbins < (1ll << 28), "lg_splat: radius too small for the grid");
    int *cnt = nullptr, *off = nullptr, *cur = nullptr, *cell = nullptr, *idx = nullptr;
    int32_t rc = dalloc(ctx, (size_t)nbins + 1, &cnt);
    if (rc == LG_OK) rc = dalloc(ctx, (size_t)nbins + 1, &off);
    if (rc == LG_OK) rc = dalloc(ctx, (size_t)nbins + 1, &cur);
    if (rc == LG_OK) rc = dalloc(ctx, (size_t)std::max<int64_t>(count, 1), &cell);
    if (rc == LG_OK) rc = dalloc(ctx, (size_t)std::max<int64_t>(count, 1), &idx);
    if (rc == LG_OK) {
        hipEvent_t ev;
        ktimer_begin(ctx, "k_splat", &ev);
        (void)hipMemsetAsync(cnt, 0, sizeof(int) * ((size_t)nbins + 1), ctx->stream);
        (void)hipMemsetAsync(cur, 0, sizeof(int) * ((size_t)nbins + 1), ctx->stream);
        if (count > 0) k_bin_count<<<(unsigned)((count + 255) / 256), 256, 0, ctx->stream>>>(d_pts, count, b, cnt, cell);
        size_t bytes = 0;
        (void)rocprim::exclusive_scan(nullptr, bytes, cnt, off, 0, (size_t)nbins + 1, rocprim::plus<int>(), ctx->stream);
        void *tmp = nullptr;
        rc = scratch(ctx, bytes, &tmp);
        if (rc == LG_OK) {
            (void)rocprim::exclusive_scan(tmp, bytes, cnt, off, 0, (size_t)nbins + 1, rocprim::plus<int>(), ctx->stream);
            if (count > 0) k_bin_fill<<<(unsigned)((count + 255) / 256), 256, 0, ctx->stream>>>(count, cell, off, cur, idx);
            const long long vox = (long long)c.nx * c.ny * c.nz;
            const double r2 = radius * radius;
            // sigma^2 = -r^2 / (2 ln 0.1) (ScatterGrid.h:75), weight exp(-d^2 / (2 sigma^2))
            const double sigma2 = -(radius * radius) / (2.0 * log(0.1));
            const double inv2sig2 = 1.0 / (2.0 * sigma2);
            const int reach = (int)ceil(radius * b.inv_cs);
            k_splat<<<(unsigned)((vox + 127) / 128), 128, 0, ctx->stream>>>(c, sg->d_active, d_pts, d_rad, b, reach, off,
                                                                           idx, r2, inv2sig2, sg->d_rgb);
            if (hipGetLastError() != hipSuccess) {
                set_error("splat kernels failed to launch");
                rc = LG_ERR_HIP;
            }
        }
        ktimer_end(ctx, "k_splat", ev);
    }
    dfree(ctx, cnt);
    dfree(ctx, off);
    dfree(ctx, cur);
    dfree(ctx, cell);
    dfree(ctx, idx);
    return rc;
}

int32_t lg_render(lg_ctx *ctx, const lg_grid *density, const lg_scatter_grid *scatter, const lg_grid *temperature,
                  const lg_camera *camera, const lg_shader *sh, const lg_light *lights, int32_t n_lights,
                  const lg_march_params *mp, int32_t samples, uint64_t seed, uint64_t env_seed, const lg_films *films) {
    LG_CHECK_ARG(ctx && density && camera && sh && mp && films, "lg_render: NULL argument");
    LG_CHECK_ARG(camera->width > 0 && camera->height > 0, "lg_render: empty image");
    LG_CHECK_ARG(samples >= 1 && samples < (1 << 20), "lg_render: samples must be in [1, 2^20)");
    LG_CHECK_ARG((uint64_t)camera->width * (uint64_t)camera->height * (uint64_t)samples < (1ull << 32),
                 "lg_render: width * height * samples must be below 2^32");
    LG_CHECK_ARG(mp->step[0] > 0 && mp->step[1] > 0, "lg_render: steps must be positive");
    LG_CHECK_ARG(!temperature || temperature->desc.dtype == LG_F32, "temperature grid must be float32");
    lgl::DevLights L{};
    LG_TRY(lgl::make_dev_lights(lights, n_lights, &L));
    RenderConsts c{};
    c.ds = sh->density_scale;
    c.sds = sh->shadow_density_scale;
    c.dmax = sh->density_max;
    c.cutoff = mp->cutoff;
    c.step0 = mp->step[0];
    c.step1 = mp->step[1];
    c.tau_star = dropout_tau(sh->extinction, mp->cutoff);
    for (int a = 0; a < 3; ++a) {
        c.albedo[a] = sh->albedo[a];
        c.ext[a] = sh->extinction[a];
        c.inv_scat[a] = 1.0 / sh->scattering[a];  // Renderer.h:516
    }
    c.scatter_scale = sh->scatter_scale;
    c.mask_pow = sh->scatter_density_mask_power;
    c.emission_scale = sh->emission_scale;
    c.ramp_in = sh->ramp_in;
    c.ramp_out = sh->ramp_out;
    c.ramp_n = std::max(0, std::min(sh->ramp_count, LG_MAX_RAMP));
    for (int k = 0; k < c.ramp_n; ++k) {
        c.ramp_pos[k] = sh->ramp_pos[k];
        for (int a = 0; a < 3; ++a) c.ramp_col[k][a] = sh->ramp_color[k][a];
    }
    memcpy(c.M, camera->screen_to_world, sizeof(c.M));
    c.scale_w = camera->scale_width;
    c.scale_h = camera->scale_height;
    c.znear = camera->znear;
    c.zfar = camera->zfar;
    c.width = camera->width;
    c.height = camera->height;
    c.samples = samples;
    c.ratio_aa = 1.0 / (double)samples;
    lgd::GridView g = make_view(density);
    const lgd::GridView tg = temperature ? make_view(temperature) : lgd::GridView{};
    const CoarseView sc = scatter && scatter->d_rgb ? coarse_view(scatter) : CoarseView{};
    const int has_sc = (scatter && scatter->d_rgb) ? 1 : 0;
    const RenderLayout lay = render_layout(c.width, c.height, samples, L.n);
    const uint32_t nrays = (uint32_t)lay.nrays;
    ctx->render_chunks = 0;
    unsigned long long *stats = nullptr;
    if (ctx->timing) {
        if (!ctx->render_stats) LG_TRY(dalloc(ctx, (size_t)RS_COUNT * kStatSlots, &ctx->render_stats));
        stats = ctx->render_stats;
        LG_HIP(hipMemsetAsync(stats, 0, sizeof(unsigned long long) * RS_COUNT * kStatSlots, ctx->stream));
    }
    // distinct bricks read by each shadow launch (lg_ctx_set_brick_counting)
    lg_grid *touch = (stats && ctx->count_bricks) ? const_cast<lg_grid *>(density) : nullptr;
    if (touch) {
        LG_TRY(grid_touch_begin(ctx, touch));
        g.touched = touch->touched;
    }
    if (nrays == 0) return LG_OK;
    // phase count over the whole image, then the shading points of every block of pixels
    void *head = nullptr;
    LG_TRY(scratch(ctx, lay.head_bytes, &head, 0));
    int *cnt = (int *)head;
    unsigned long long *bsum = (unsigned long long *)((char *)head + lay.bsum_off);
    const uint32_t u0 = (uint32_t)seed, u1 = (uint32_t)(seed >> 32);
    const uint32_t e0 = (uint32_t)env_seed, e1 = (uint32_t)(env_seed >> 32);
    hipEvent_t ev;
    ktimer_begin(ctx, "k_render_count", &ev);
    for_dtype_r(density->desc.dtype, [&](auto tag) {
        using T = decltype(tag);
        k_render_count<T><<<(nrays + 255) / 256, 256, 0, ctx->stream>>>(g, c, u0, u1, nrays, cnt, stats);
    });
    LG_LAUNCHED("k_render_count");
    ktimer_end(ctx, "k_render_count", ev);
    const uint32_t nblocks = (uint32_t)lay.nblocks;
    k_render_block_sum<<<(nblocks + 3) / 4, 256, 0, ctx->stream>>>(cnt, nrays, (uint32_t)lay.rays_per_block, nblocks,
                                                                    bsum);
    LG_LAUNCHED("k_render_block_sum");
    std::vector<unsigned long long> h_bsum(nblocks);
    LG_HIP(hipMemcpyAsync(h_bsum.data(), bsum, sizeof(unsigned long long) * nblocks, hipMemcpyDeviceToHost,
                          ctx->stream));
    LG_HIP(hipStreamSynchronize(ctx->stream));
    // chunks of whole blocks within the ray and point caps
    struct Chunk {
        uint32_t b0, b1;
        uint64_t points;
    };
    std::vector<Chunk> chunks;
    uint64_t max_points = 0;
    {
        Chunk cur{0, 0, 0};
        for (uint32_t b = 0; b < nblocks; ++b) {
            const uint64_t rays_after = (uint64_t)(b + 1 - cur.b0) * lay.rays_per_block;
            if (cur.b1 > cur.b0 && (rays_after > lay.ray_cap || cur.points + h_bsum[b] > lay.point_cap)) {
                chunks.push_back(cur);
                cur = Chunk{b, b, 0};
            }
            cur.b1 = b + 1;
            cur.points += h_bsum[b];
        }
        chunks.push_back(cur);
        for (const Chunk &ch : chunks) max_points = std::max(max_points, ch.points);
    }
    LG_CHECK_ARG(max_points < (1ull << 31) / (uint64_t)std::max(L.n, 1),
                 "lg_render: %llu shading points in one block of %u pixels", (unsigned long long)max_points,
                 kRenderBlockPixels);
    const uint64_t pad_cap = (max_points + 63) & ~63ull;
    const size_t pts_off = lay.ray_area_bytes;
    const size_t ext_off = pts_off + align256(sizeof(ShadePoint) * max_points);
    const size_t sum_off = ext_off + align256(sizeof(ShadeExtra) * max_points);
    const size_t body_bytes = sum_off + align256(24ull * (uint64_t)L.n * pad_cap);
    // Several chunks alternate between two workspaces and the context's two
    // streams, so that one chunk's bandwidth-bound phases (fill, combine) run
    // beside the other's latency-bound shadow march.  Chunks share only
    // read-only inputs and write disjoint pixels.  With kernel timing on (per
    // phase times, brick counting) the chunks run in order on one stream.
    const bool two = chunks.size() > 1 && !ctx->timing && !touch;
    void *body0 = nullptr;
    LG_TRY(scratch(ctx, (two ? 2 : 1) * body_bytes, &body0, 1));
    struct SideJoin {
        lg_ctx *ctx;
        bool on;
        ~SideJoin() {
            if (on && hipEventRecord(ctx->ev_join, ctx->side) == hipSuccess)
                (void)hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0);
        }
    } join{ctx, false};
    if (two) {
        LG_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
        LG_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
        join.on = true;
    }
    for (size_t ci = 0; ci < chunks.size(); ++ci) {
        const Chunk &ch = chunks[ci];
        const bool alt = two && (ci & 1);
        hipStream_t st = alt ? ctx->side : ctx->stream;
        void *body = (char *)body0 + (alt ? body_bytes : 0);
        int *off = (int *)body;
        void *scan_tmp = (char *)body + lay.scan_off;
        RayOut *rout = (RayOut *)((char *)body + lay.out_off);
        ShadePoint *pts = (ShadePoint *)((char *)body + pts_off);
        ShadeExtra *ext = (ShadeExtra *)((char *)body + ext_off);
        double *sums = (double *)((char *)body + sum_off);
        const uint32_t pix0 = ch.b0 * kRenderBlockPixels;
        const uint32_t pix1 = std::min<uint64_t>((uint64_t)ch.b1 * kRenderBlockPixels, lay.npix);
        const uint32_t ray0 = pix0 * (uint32_t)samples, nr = (pix1 - pix0) * (uint32_t)samples;
        const uint32_t P = (uint32_t)ch.points, P_pad = (P + 63u) & ~63u;
        if (P > 0) {
            size_t tb = lay.scan_bytes;
            LG_HIP(rocprim::exclusive_scan(scan_tmp, tb, cnt + ray0, off, 0, (size_t)nr, rocprim::plus<int>(),
                                           st));
            ktimer_begin(ctx, "k_render_fill", &ev);
            for_dtype_r(density->desc.dtype, [&](auto tag) {
                using T = decltype(tag);
                k_render_fill<T><<<(nr + 255) / 256, 256, 0, st>>>(g, c, u0, u1, ray0, nr, cnt, off, pts,
                                                                            ext);
            });
            LG_LAUNCHED("k_render_fill");
            ktimer_end(ctx, "k_render_fill", ev);
            const uint64_t tasks = (uint64_t)L.n * P_pad;
            if (tasks > 0) {
                ktimer_begin(ctx, "k_render_shadow", &ev);
                for_dtype_r(density->desc.dtype, [&](auto tag) {
                    using T = decltype(tag);
                    if (touch)
                        k_render_shadow<T, true><<<(unsigned)((tasks + 255) / 256), 256, 0, st>>>(
                            g, c, L, u0, u1, e0, e1, pts, P, P_pad, sums, stats);
                    else
                        k_render_shadow<T, false><<<(unsigned)((tasks + 255) / 256), 256, 0, st>>>(
                            g, c, L, u0, u1, e0, e1, pts, P, P_pad, sums, stats);
                });
                LG_LAUNCHED("k_render_shadow");
                ktimer_end(ctx, "k_render_shadow", ev);
                if (touch) LG_TRY(grid_touch_count(ctx, touch, stats, RS_SHADOW_BRICKS));
            }
        }
        ktimer_begin(ctx, "k_render_combine", &ev);
        k_render_combine<<<(nr + 255) / 256, 256, 0, st>>>(g, tg, temperature ? 1 : 0, sc, has_sc, c, L,
                                                                    ray0, nr, cnt, off, pts, ext, sums, P_pad, rout,
                                                                    stats);
        LG_LAUNCHED("k_render_combine");
        ktimer_end(ctx, "k_render_combine", ev);
        k_render_aa<<<(pix1 - pix0 + 255) / 256, 256, 0, st>>>(
            c, pix0, pix1 - pix0, rout, (float4 *)films->d_rgb, (float4 *)films->d_direct, (float4 *)films->d_scatter,
            (float4 *)films->d_emission);
        LG_LAUNCHED("k_render_aa");
    }
    ctx->render_chunks = chunks.size();
    return LG_OK;
}

int32_t lg_render_get_stats(lg_ctx *ctx, lg_render_stats *out) {
    LG_CHECK_ARG(ctx && out, "lg_render_get_stats: NULL argument");
    memset(out, 0, sizeof(*out));
    out->chunks = ctx->render_chunks;
    if (!ctx->render_stats) return LG_OK;
    std::vector<unsigned long long> h((size_t)RS_COUNT * kStatSlots);
    LG_HIP(hipMemcpyAsync(h.data(), ctx->render_stats, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost,
                          ctx->stream));
    LG_HIP(hipStreamSynchronize(ctx->stream));
    uint64_t v[RS_COUNT] = {};
    for (int k = 0; k < RS_COUNT; ++k)
        for (int q = 0; q < kStatSlots; ++q) v[k] += h[(size_t)k * kStatSlots + q];
    out->camera_rays = v[RS_CAMERA_RAYS];
    out->primary_samples = v[RS_PRIMARY_SAMPLES];
    out->shading_points = v[RS_SHADING_POINTS];
    out->shadow_rays = v[RS_SHADOW_RAYS];
    out->shadow_samples = v[RS_SHADOW_SAMPLES];
    out->shadow_lane_slots = v[RS_SHADOW_LANE_SLOTS];
    out->shadow_bricks = v[RS_SHADOW_BRICKS];
    return LG_OK;
}

int32_t lg_render_workspace(int32_t width, int32_t height, int32_t samples, int32_t n_lights,
                            int64_t max_points_per_ray, uint64_t *bytes) {
    LG_CHECK_ARG(bytes && width > 0 && height > 0 && samples >= 1 && n_lights >= 0 && n_lights <= LG_MAX_LIGHTS,
                 "lg_render_workspace: bad argument");
    const RenderLayout l = render_layout(width, height, samples, n_lights);
    uint64_t pts = l.point_cap;
    if (max_points_per_ray > 0) {
        const uint64_t chunk_rays = std::max<uint64_t>(l.ray_cap, l.rays_per_block);
        pts = std::min<uint64_t>(pts, chunk_rays * (uint64_t)max_points_per_ray);
    }
    // more than one chunk (more rays than one chunk takes, or possibly more
    // points): lg_render keeps two chunk workspaces
    const bool two = l.nrays > l.ray_cap || max_points_per_ray <= 0 ||
                     (uint64_t)max_points_per_ray * l.nrays > l.point_cap;
    *bytes = l.head_bytes + (two ? 2 : 1) * (l.ray_area_bytes + align256(pts * l.point_bytes)) + 1024;
    return LG_OK;
}

int32_t lg_fw_autotune(lg_ctx *ctx, lg_fw_tuning *out) {
    LG_CHECK_ARG(ctx, "lg_fw_autotune: ctx is NULL");
    lg_fw_tuning t{};
    LG_TRY(fw_autotune(ctx, &t));
    ctx->fw_tune = t;
    ctx->fw_tune_set = true;
    if (out) *out = t;
    return LG_OK;
}

int32_t lg_fw_get_tuning(lg_ctx *ctx, lg_fw_tuning *out) {
    LG_CHECK_ARG(ctx && out, "lg_fw_get_tuning: NULL argument");
    if (!ctx->fw_tune_set) LG_TRY(fw_tuning_init(ctx));
    *out = ctx->fw_tune;
    return LG_OK;
}

int32_t lg_fw_set_tuning(lg_ctx *ctx, const lg_fw_tuning *t) {
    LG_CHECK_ARG(ctx && t && t->split_below >= 0 && (t->use_graph == 0 || t->use_graph == 1),
                 "lg_fw_set_tuning: bad argument");
    ctx->fw_tune = *t;
    ctx->fw_tune_set = true;
    return LG_OK;
}

int32_t lg_fw_solve_batched(lg_ctx *ctx, int32_t count, const int32_t *h_n, float *const *h_D, const int32_t *h_ld) {
    LG_CHECK_ARG(ctx && count >= 0 && (count == 0 || (h_n && h_D && h_ld)), "lg_fw_solve_batched: bad argument");
    std::vector<float *> Ds((size_t)count);
    std::vector<int> ns((size_t)count);
    for (int g = 0; g < count; ++g) {
        LG_CHECK_ARG(h_n[g] >= 0 && h_D[g] && h_ld[g] == fw_ld(h_n[g]), "lg_fw_solve_batched: graph %d: bad n/D/ld", g);
        Ds[g] = h_D[g];
        ns[g] = h_n[g];
    }
    return fw_solve_batched_impl(ctx, Ds, ns);
}

}  // extern "C"
