// Procedural density fields for the synthetic workloads.  The same functions run
// on the host (lg_scene_generate_host) and on the device (lg_grid_generate); with
// -ffp-contract=off and only +, *, floor and integer hashing they give the same
// bits on both.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define LGN_HD __host__ __device__ __forceinline__

namespace lgn {

// lowbias32 integer finaliser (C. Wellons' hash-prospector).
LGN_HD uint32_t mix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

LGN_HD uint32_t hash3(int x, int y, int z, uint32_t seed) {
    uint32_t h = seed ^ ((uint32_t)x * 0x8da6b343u) ^ ((uint32_t)y * 0xd8163841u) ^ ((uint32_t)z * 0xcb1ab31fu);
    return mix32(h);
}

// Gradient from Perlin's improved-noise set of 12 edge directions.
LGN_HD float grad(uint32_t h, float x, float y, float z) {
    const uint32_t g = h & 15u;
    const float u = g < 8u ? x : y;
    const float v = g < 4u ? y : ((g == 12u || g == 14u) ? x : z);
    return ((g & 1u) ? -u : u) + ((g & 2u) ? -v : v);
}

LGN_HD float fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }
LGN_HD float lerpf_(float a, float b, float t) { return a + t * (b - a); }
LGN_HD float floorf_(float x) { return __builtin_floorf(x); }

// 3D gradient noise, roughly in [-1, 1].
LGN_HD float gnoise(float x, float y, float z, uint32_t seed) {
    const float fx = floorf_(x), fy = floorf_(y), fz = floorf_(z);
    const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    const float rx = x - fx, ry = y - fy, rz = z - fz;
    const float u = fade(rx), v = fade(ry), w = fade(rz);
    const float n000 = grad(hash3(ix, iy, iz, seed), rx, ry, rz);
    const float n100 = grad(hash3(ix + 1, iy, iz, seed), rx - 1.0f, ry, rz);
    const float n010 = grad(hash3(ix, iy + 1, iz, seed), rx, ry - 1.0f, rz);
    const float n110 = grad(hash3(ix + 1, iy + 1, iz, seed), rx - 1.0f, ry - 1.0f, rz);
    const float n001 = grad(hash3(ix, iy, iz + 1, seed), rx, ry, rz - 1.0f);
    const float n101 = grad(hash3(ix + 1, iy, iz + 1, seed), rx - 1.0f, ry, rz - 1.0f);
    const float n011 = grad(hash3(ix, iy + 1, iz + 1, seed), rx, ry - 1.0f, rz - 1.0f);
    const float n111 = grad(hash3(ix + 1, iy + 1, iz + 1, seed), rx - 1.0f, ry - 1.0f, rz - 1.0f);
    const float x00 = lerpf_(n000, n100, u), x10 = lerpf_(n010, n110, u);
    const float x01 = lerpf_(n001, n101, u), x11 = lerpf_(n011, n111, u);
    return lerpf_(lerpf_(x00, x10, v), lerpf_(x01, x11, v), w);
}

// fBm with lacunarity 2 and gain 0.5, normalised by the amplitude sum, then
// remapped to [0, 1] as clamp(0.5 + fbm, 0, 1).  base_freq in cycles per voxel.
LGN_HD float fbm01(float x, float y, float z, int octaves, float base_freq, uint32_t seed) {
    float sum = 0.0f, amp = 1.0f, norm = 0.0f, f = base_freq;
    for (int o = 0; o < octaves; ++o) {
        const float off = 17.31f * (float)o;
        sum += amp * gnoise(x * f + off, y * f + off, z * f + off, seed + 1013u * (uint32_t)o);
        norm += amp;
        amp *= 0.5f;
        f *= 2.0f;
    }
    const float v = 0.5f + sum / norm;
    return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
}

LGN_HD float clamp01(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }

// Sphere of radius r centred at c with a 2-voxel smooth edge, times fBm.
LGN_HD float sphere_fbm(int x, int y, int z, float c, float r, int octaves, uint32_t seed) {
    const float dx = (float)x - c, dy = (float)y - c, dz = (float)z - c;
    const float d = __builtin_sqrtf(dx * dx + dy * dy + dz * dz);
    const float s = clamp01((r + 1.0f - d) * 0.5f);
    if (s == 0.0f) return 0.0f;
    return clamp01(s * fbm01((float)x, (float)y, (float)z, octaves, 1.0f / 16.0f, seed));
}

// Ellipsoid blob: centre, rotation rows (world->local) and inverse squared semi-axes.
struct Blob {
    float c[3];
    float r[9];
    float inv_a2[3];
};

// Soft union of ellipsoids: max over blobs of clamp(1 - q/2, 0, 1) with q the
// normalised squared radius, so a blob fades out at sqrt(2) times its semi-axes.
LGN_HD float blob_field(int x, int y, int z, const Blob *blobs, int nblobs) {
    float best = 0.0f;
    for (int b = 0; b < nblobs; ++b) {
        const Blob &B = blobs[b];
        const float px = (float)x - B.c[0], py = (float)y - B.c[1], pz = (float)z - B.c[2];
        const float lx = B.r[0] * px + B.r[1] * py + B.r[2] * pz;
        const float ly = B.r[3] * px + B.r[4] * py + B.r[5] * pz;
        const float lz = B.r[6] * px + B.r[7] * py + B.r[8] * pz;
        const float q = lx * lx * B.inv_a2[0] + ly * ly * B.inv_a2[1] + lz * lz * B.inv_a2[2];
        const float s = clamp01(1.0f - 0.5f * q);
        best = s > best ? s : best;
    }
    return best;
}

LGN_HD float blobs_raw(int x, int y, int z, const Blob *blobs, int nblobs, int octaves, uint32_t seed) {
    const float s = blob_field(x, y, z, blobs, nblobs);
    if (s == 0.0f) return 0.0f;
    return s * fbm01((float)x, (float)y, (float)z, octaves, 1.0f / 32.0f, seed);
}

// Threshold remap: rho = (f - t) / (1 - t) above t, 0 below.
LGN_HD float threshold_remap(float f, float t) {
    if (!(f > t)) return 0.0f;
    return clamp01((f - t) / (1.0f - t));
}

constexpr int kHistBins = 4096;
LGN_HD int hist_bin(float f) {
    int b = (int)(f * (float)kHistBins);
    return b < 0 ? 0 : (b >= kHistBins ? kHistBins - 1 : b);
}

// Plume: an fBm column rising along +y whose horizontal offset is advected by a
// curl-like swirl; density fades with height; temperature 0..1500 K peaks at the base.
LGN_HD void plume(int x, int y, int z, int n, int octaves, uint32_t seed, float *rho, float *temp) {
    const float fn = (float)n;
    const float h = (float)y / fn;                          // 0 bottom .. 1 top
    const float sw = 0.12f * fn * h;                        // swirl amplitude grows with height
    const float ang = 6.2831853f * 1.5f * h;
    // swirl offset from two noise channels (a divergence-free-looking advection)
    const float ox = sw * gnoise(0.5f, h * 3.0f, 0.25f, seed + 7u) + 0.05f * fn * (ang - 3.0f) * 0.1f;
    const float oz = sw * gnoise(0.75f, h * 3.0f, 0.5f, seed + 11u);
    const float cx = 0.5f * fn + ox, cz = 0.5f * fn + oz;
    const float dx = (float)x - cx, dz = (float)z - cz;
    const float rad = (0.06f + 0.22f * h) * fn;            // widening column
    const float d = __builtin_sqrtf(dx * dx + dz * dz);
    const float s = clamp01((rad - d) / (0.25f * rad + 1.0f));
    if (s == 0.0f || h < 0.03f) {
        *rho = 0.0f;
        *temp = 0.0f;
        return;
    }
    const float f = fbm01((float)x, (float)y - 40.0f * h * h * fn * 0.01f, (float)z, octaves, 4.0f / fn, seed);
    const float r = clamp01(s * f * 1.6f - 0.25f) * clamp01(1.15f - h);
    *rho = r;
    const float t = clamp01(1.0f - 1.6f * h) * clamp01(s * 1.5f) * f;
    *temp = r > 0.0f ? 1500.0f * t : 0.0f;
}

}  // namespace lgn
