// Procedural volumes for the synthetic workloads (no counterpart in the
// reference, which reads VDB files): generated straight into the brick storage on
// the device, with a host twin that writes a dense array.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "lg_internal.h"
#include "lg_noise.h"

namespace lg {

// Philox4x32-10 on the host (same rounds as lgd::philox).
static void philox_host(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

static double u01h(uint32_t x) { return (double)(x >> 8) * (1.0 / 16777216.0); }

// Ellipsoid parameters: stream (blob, 0, block, 5).
static std::vector<lgn::Blob> make_blobs(const lg_scene_spec &s) {
    std::vector<lgn::Blob> out((size_t)std::max(0, s.blobs));
    const uint32_t k0 = (uint32_t)s.seed, k1 = (uint32_t)(s.seed >> 32);
    for (int b = 0; b < s.blobs; ++b) {
        uint32_t r0[4] = {(uint32_t)b, 0u, 0u, 5u};
        uint32_t r1[4] = {(uint32_t)b, 0u, 1u, 5u};
        philox_host(r0, k0, k1);
        philox_host(r1, k0, k1);
        lgn::Blob &B = out[b];
        const double n = (double)s.n;
        for (int c = 0; c < 3; ++c) B.c[c] = (float)(u01h(r0[c]) * n);
        const double ax[3] = {s.axis_min + u01h(r0[3]) * (s.axis_max - s.axis_min),
                              s.axis_min + u01h(r1[0]) * (s.axis_max - s.axis_min),
                              s.axis_min + u01h(r1[1]) * (s.axis_max - s.axis_min)};
        for (int c = 0; c < 3; ++c) B.inv_a2[c] = (float)(1.0 / (ax[c] * ax[c]));
        // rotation Rz(c)*Ry(b)*Rx(a); rows map world offsets to the local frame
        const double a = u01h(r1[2]) * 6.283185307179586, bb = u01h(r1[3]) * 6.283185307179586,
                     cc = u01h(r0[2] ^ r1[3]) * 6.283185307179586;
        const double ca = cos(a), sa = sin(a), cb = cos(bb), sb = sin(bb), cg = cos(cc), sg = sin(cc);
        const double R[9] = {cg * cb, cg * sb * sa - sg * ca, cg * sb * ca + sg * sa,
                             sg * cb, sg * sb * sa + cg * ca, sg * sb * ca - cg * sa,
                             -sb,     cb * sa,                cb * ca};
        for (int i = 0; i < 9; ++i) B.r[i] = (float)R[i];
    }
    return out;
}

constexpr int kMaxBlobs = 64;  // lg_scene_spec.blobs is at most this

// One thread per voxel of the (unpadded) grid, written into storage brick layout.
// mode 0: sphere*fBm, 1: raw blob field (blobs: the context's device copy of the
// scene's ellipsoids), 2: plume density (+temperature).
__global__ void k_generate(int mode, int n, int octaves, uint32_t seed, float radius, const lgn::Blob *__restrict__ blobs,
                           int nblobs, int NBx, int NBy, float *__restrict__ out, float *__restrict__ temp_out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)n * n * n;
    if (i >= total) return;
    const int x = (int)(i % n), y = (int)((i / n) % n), z = (int)(i / ((long long)n * n));
    float v = 0.0f, t = 0.0f;
    if (mode == 0) {
        v = lgn::sphere_fbm(x, y, z, 0.5f * (float)(n - 1), radius, octaves, seed);
    } else if (mode == 1) {
        v = lgn::blobs_raw(x, y, z, blobs, nblobs, octaves, seed);
    } else {
        lgn::plume(x, y, z, n, octaves, seed, &v, &t);
    }
    const long long brick = ((long long)((z >> 3) + 1) * NBy + ((y >> 3) + 1)) * NBx + ((x >> 3) + 1);
    const long long off = (brick << 9) + ((z & 7) << 6) + ((y & 7) << 3) + (x & 7);
    out[off] = v;
    if (temp_out) temp_out[off] = t;
}

__global__ void k_histogram(const float *__restrict__ data, long long elems, unsigned long long *__restrict__ hist) {
    __shared__ unsigned int sh[lgn::kHistBins];
    for (int b = threadIdx.x; b < lgn::kHistBins; b += blockDim.x) sh[b] = 0;
    __syncthreads();
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < elems; i += (long long)gridDim.x * blockDim.x) {
        const float f = data[i];
        if (f > 0.0f) atomicAdd(&sh[lgn::hist_bin(f)], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < lgn::kHistBins; b += blockDim.x)
        if (sh[b]) atomicAdd(&hist[b], (unsigned long long)sh[b]);
}

__global__ void k_threshold(float *__restrict__ data, long long elems, float t) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < elems; i += (long long)gridDim.x * blockDim.x)
        data[i] = lgn::threshold_remap(data[i], t);
}

template <typename T>
__global__ void k_convert(const float *__restrict__ src, long long elems, T *__restrict__ dst) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < elems; i += (long long)gridDim.x * blockDim.x) {
        const float v = src[i];
        if constexpr (std::is_same<T, _Float16>::value) {
            dst[i] = (_Float16)v;
        } else {
            uint32_t u = __float_as_uint(v);
            u += 0x7fffu + ((u >> 16) & 1u);
            dst[i].v = (uint16_t)(u >> 16);
        }
    }
}

// Threshold bin k minimising |count(f >= k/bins) / total - target|.
static float pick_threshold(const std::vector<unsigned long long> &hist, double total, double target) {
    unsigned long long above = 0;
    int best_k = 0;
    double best_err = 1e300;
    for (int k = lgn::kHistBins - 1; k >= 0; --k) {
        above += hist[k];
        const double err = fabs((double)above / total - target);
        if (err < best_err) {
            best_err = err;
            best_k = k;
        }
    }
    return (float)best_k / (float)lgn::kHistBins;
}

static int32_t check_spec(const lg_scene_spec *s) {
    LG_CHECK_ARG(s, "scene: spec is NULL");
    LG_CHECK_ARG(s->n >= 8 && s->n <= 2048, "scene: side %d out of range [8, 2048]", s->n);
    LG_CHECK_ARG(s->kind >= LG_SCENE_SPHERE_FBM && s->kind <= LG_SCENE_PLUME, "scene: unknown kind %d", s->kind);
    LG_CHECK_ARG(s->octaves >= 1 && s->octaves <= 12, "scene: octaves out of range");
    LG_CHECK_ARG(s->kind != LG_SCENE_BLOBS_FBM || (s->blobs >= 1 && s->blobs <= 64), "scene: blobs must be in [1, 64]");
    static_assert(kMaxBlobs == 64, "the bound checked above");
    return LG_OK;
}

}  // namespace lg

using namespace lg;

extern "C" {

int32_t lg_grid_generate(lg_ctx *ctx, const lg_scene_spec *spec, lg_grid **density_out, lg_grid **temperature_out) {
    LG_CHECK_ARG(ctx && density_out, "lg_grid_generate: NULL argument");
    LG_TRY(check_spec(spec));
    *density_out = nullptr;
    if (temperature_out) *temperature_out = nullptr;
    const bool want_temp = temperature_out && spec->kind == LG_SCENE_PLUME && spec->with_temperature;
    lg_grid_desc d{};
    d.nx = d.ny = d.nz = spec->n;
    d.dtype = LG_F32;
    d.voxel_size = 1.0;
    lg_grid *g = nullptr, *t = nullptr;
    LG_TRY(grid_alloc(ctx, &d, &g));
    int32_t rc = LG_OK;
    if (want_temp) rc = grid_alloc(ctx, &d, &t);
    const int mode = spec->kind == LG_SCENE_SPHERE_FBM ? 0 : (spec->kind == LG_SCENE_BLOBS_FBM ? 1 : 2);
    std::vector<lgn::Blob> blobs;
    if (rc == LG_OK && mode == 1) {
        // each context keeps its own device copy: contexts on other threads
        // (the serving workers) generate their volumes at the same time
        blobs = make_blobs(*spec);
        if (!ctx->blobs_dev) rc = dmalloc(ctx, sizeof(lgn::Blob) * kMaxBlobs, &ctx->blobs_dev);
        if (rc == LG_OK && (blobs.size() > (size_t)kMaxBlobs ||
                            hipMemcpyAsync(ctx->blobs_dev, blobs.data(), blobs.size() * sizeof(lgn::Blob),
                                           hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
                            hipStreamSynchronize(ctx->stream) != hipSuccess)) {
            set_error("blob upload failed");
            rc = LG_ERR_HIP;
        }
    }
    const long long total = (long long)spec->n * spec->n * spec->n;
    const long long elems = (long long)grid_storage_elems(g);
    if (rc == LG_OK) {
        k_generate<<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(
            mode, spec->n, spec->octaves, (uint32_t)spec->seed, (float)spec->radius,
            (const lgn::Blob *)ctx->blobs_dev, (int)blobs.size(), g->NB[0], g->NB[1], (float *)g->data,
            t ? (float *)t->data : nullptr);
        if (hipGetLastError() != hipSuccess) {
            set_error("k_generate launch failed");
            rc = LG_ERR_HIP;
        }
    }
    if (rc == LG_OK && mode == 1 && spec->active_fraction > 0.0) {
        void *s = nullptr;
        rc = scratch(ctx, lgn::kHistBins * sizeof(unsigned long long), &s);
        if (rc == LG_OK) {
            unsigned long long *hist = (unsigned long long *)s;
            std::vector<unsigned long long> h(lgn::kHistBins);
            hipError_t e = hipMemsetAsync(hist, 0, lgn::kHistBins * sizeof(unsigned long long), ctx->stream);
            k_histogram<<<1024, 256, 0, ctx->stream>>>((const float *)g->data, elems, hist);
            if (e == hipSuccess) e = hipGetLastError();
            if (e == hipSuccess)
                e = hipMemcpyAsync(h.data(), hist, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                   ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) {
                set_error("histogram failed: %s", hipGetErrorString(e));
                rc = LG_ERR_HIP;
            } else {
                const float thr = pick_threshold(h, (double)total, spec->active_fraction);
                k_threshold<<<2048, 256, 0, ctx->stream>>>((float *)g->data, elems, thr);
                if (hipGetLastError() != hipSuccess) {
                    set_error("k_threshold launch failed");
                    rc = LG_ERR_HIP;
                }
            }
        }
    }
    if (rc == LG_OK && spec->dtype != LG_F32) {
        void *dst = nullptr;
        rc = dmalloc(ctx, (size_t)elems * 2, &dst);
        if (rc == LG_OK) {
            if (spec->dtype == LG_F16)
                k_convert<_Float16><<<2048, 256, 0, ctx->stream>>>((const float *)g->data, elems, (_Float16 *)dst);
            else
                k_convert<lgd::bf16_t><<<2048, 256, 0, ctx->stream>>>((const float *)g->data, elems, (lgd::bf16_t *)dst);
            if (hipGetLastError() != hipSuccess) {
                set_error("k_convert launch failed");
                rc = LG_ERR_HIP;
                dfree(ctx, dst);
            } else {
                dfree(ctx, g->data);
                g->data = dst;
                g->desc.dtype = spec->dtype;
                g->bytes = (uint64_t)elems * 2 + (uint64_t)g->NB[0] * g->NB[1] * g->NB[2];
            }
        }
    }
    if (rc == LG_OK) rc = grid_finalize(ctx, g);
    if (rc == LG_OK && t) rc = grid_finalize(ctx, t);
    if (rc != LG_OK) {
        lg_grid_destroy(ctx, g);
        lg_grid_destroy(ctx, t);
        return rc;
    }
    *density_out = g;
    if (temperature_out) *temperature_out = t;
    return LG_OK;
}

int32_t lg_scene_generate_host(const lg_scene_spec *spec, float *out, int32_t threads) {
    LG_CHECK_ARG(out, "lg_scene_generate_host: out is NULL");
    LG_TRY(check_spec(spec));
    const int n = spec->n;
    const long long total = (long long)n * n * n;
    const int mode = spec->kind == LG_SCENE_SPHERE_FBM ? 0 : (spec->kind == LG_SCENE_BLOBS_FBM ? 1 : 2);
    std::vector<lgn::Blob> blobs;
    if (mode == 1) blobs = make_blobs(*spec);
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    auto work = [&](int tid) {
        for (long long i = tid; i < total; i += threads) {
            const int x = (int)(i % n), y = (int)((i / n) % n), z = (int)(i / ((long long)n * n));
            float v = 0.0f, t = 0.0f;
            if (mode == 0)
                v = lgn::sphere_fbm(x, y, z, 0.5f * (float)(n - 1), (float)spec->radius, spec->octaves,
                                    (uint32_t)spec->seed);
            else if (mode == 1)
                v = lgn::blobs_raw(x, y, z, blobs.data(), (int)blobs.size(), spec->octaves, (uint32_t)spec->seed);
            else
                lgn::plume(x, y, z, n, spec->octaves, (uint32_t)spec->seed, &v, &t);
            out[i] = v;
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) pool.emplace_back(work, t);
    for (auto &th : pool) th.join();
    if (mode == 1 && spec->active_fraction > 0.0) {
        std::vector<unsigned long long> h(lgn::kHistBins, 0);
        for (long long i = 0; i < total; ++i)
            if (out[i] > 0.0f) ++h[lgn::hist_bin(out[i])];
        const float thr = pick_threshold(h, (double)total, spec->active_fraction);
        for (long long i = 0; i < total; ++i) out[i] = lgn::threshold_remap(out[i], thr);
    }
    return LG_OK;
}

}  // extern "C"
