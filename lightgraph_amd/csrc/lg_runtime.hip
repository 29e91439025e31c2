// Context, device memory, error reporting, timers and dense-brick grid creation.
#include <cstdarg>
#include <cstring>
#include <algorithm>

#include "lg_internal.h"
#include "lg_device.h"

namespace lg {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int32_t dmalloc(lg_ctx *ctx, size_t bytes, void **out) {
    *out = nullptr;
    if (bytes == 0) bytes = 256;
    if (ctx->budget && ctx->in_use + bytes > ctx->budget) {
        set_error("HBM budget exceeded: %llu + %llu > %llu bytes", (unsigned long long)ctx->in_use,
                  (unsigned long long)bytes, (unsigned long long)ctx->budget);
        return LG_ERR_BUDGET;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        return LG_ERR_NOMEM;
    }
    ctx->allocs[p] = bytes;
    ctx->in_use += bytes;
    *out = p;
    return LG_OK;
}

int32_t dfree(lg_ctx *ctx, void *p) {
    if (!p) return LG_OK;
    auto it = ctx->allocs.find(p);
    if (it == ctx->allocs.end()) {
        set_error("lg_free: pointer %p was not allocated by this context", p);
        return LG_ERR_INVALID;
    }
    // Work queued on the context stream may still use the buffer.
    LG_HIP(hipStreamSynchronize(ctx->stream));
    LG_HIP(hipFree(p));
    ctx->in_use -= it->second;
    ctx->allocs.erase(it);
    return LG_OK;
}

int32_t scratch(lg_ctx *ctx, size_t bytes, void **out, int slot) {
    void *&p = slot == 0 ? ctx->scratch : ctx->scratch1;
    size_t &have = slot == 0 ? ctx->scratch_bytes : ctx->scratch1_bytes;
    if (bytes > have) {
        if (p) LG_TRY(dfree(ctx, p));
        p = nullptr;
        have = 0;
        size_t want = std::max(bytes, (size_t)1 << 20);
        LG_TRY(dmalloc(ctx, want, &p));
        have = want;
    }
    *out = p;
    return LG_OK;
}

static hipEvent_t ktimer_event(lg_ctx *ctx) {
    if (!ctx->ktimer.pool.empty()) {
        hipEvent_t e = ctx->ktimer.pool.back();
        ctx->ktimer.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void ktimer_begin(lg_ctx *ctx, const char *name, hipEvent_t *a) {
    *a = nullptr;
    if (!ctx->timing) return;
    if (!ctx->timing_only.empty() && ctx->timing_only != name) return;
    *a = ktimer_event(ctx);
    if (*a) (void)hipEventRecord(*a, ctx->stream);
}

void ktimer_end(lg_ctx *ctx, const char *name, hipEvent_t a) {
    if (!ctx->timing || !a) return;
    hipEvent_t b = ktimer_event(ctx);
    if (!b) return;
    (void)hipEventRecord(b, ctx->stream);
    ctx->ktimer.pending.push_back({name, a, b});
}

static void ktimer_collect(lg_ctx *ctx) {
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &p : ctx->ktimer.pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            auto &t = ctx->ktimer.totals[p.name];
            t.first += 1;
            t.second += ms;
        }
        ctx->ktimer.pool.push_back(p.a);
        ctx->ktimer.pool.push_back(p.b);
    }
    ctx->ktimer.pending.clear();
}

// ------------------------------------------------------ grid conversion ----
struct GridDims {
    int nx, ny, nz;
    int NBx, NBy, NBz;
};

// Scatter one z-slab (8 voxel layers, x fastest) of a dense float volume into the
// brick storage, converting to the storage type.  One thread per voxel.
template <typename T>
__global__ void k_slab_to_bricks(const float *__restrict__ slab, int z0, int nzs, GridDims d,
                                 T *__restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per = (long long)d.nx * d.ny;
    if (i >= per * nzs) return;
    const int z = z0 + (int)(i / per);
    const int rem = (int)(i % per);
    const int y = rem / d.nx, x = rem % d.nx;
    const long long brick = ((long long)((z >> 3) + 1) * d.NBy + ((y >> 3) + 1)) * d.NBx + ((x >> 3) + 1);
    const long long off = (brick << 9) + ((z & 7) << 6) + ((y & 7) << 3) + (x & 7);
    const float v = slab[i];
    if constexpr (sizeof(T) == 4) {
        out[off] = v;
    } else if constexpr (std::is_same<T, _Float16>::value) {
        out[off] = (_Float16)v;
    } else {
        // round-to-nearest-even float -> bfloat16
        uint32_t u = __float_as_uint(v);
        if ((u & 0x7fffffffu) > 0x7f800000u) {
            out[off].v = (uint16_t)((u >> 16) | 0x40u);
        } else {
            u += 0x7fffu + ((u >> 16) & 1u);
            out[off].v = (uint16_t)(u >> 16);
        }
    }
}

// Per-brick occupancy and active-voxel count.  One 256-thread block per storage brick.
template <typename T>
__global__ void k_brick_occupancy(const T *__restrict__ data, long long nbricks, uint8_t *__restrict__ occ,
                                  unsigned long long *__restrict__ active_voxels) {
    const long long b = blockIdx.x;
    if (b >= nbricks) return;
    const T *p = data + (b << 9);
    int cnt = 0;
    for (int i = threadIdx.x; i < 512; i += blockDim.x) cnt += (lgd::to_f32(p[i]) != 0.0f) ? 1 : 0;
    __shared__ int s_cnt;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    if (cnt) atomicAdd(&s_cnt, cnt);
    __syncthreads();
    if (threadIdx.x == 0) {
        occ[b] = s_cnt ? 1 : 0;
        if (s_cnt) atomicAdd(active_voxels, (unsigned long long)s_cnt);
    }
}

// One axis of a max-norm distance transform over the storage bricks:
// out = min over |k| < kSkipCap of max(|k|, in at offset k along the axis),
// capped at kSkipCap (the max-norm distance separates into one pass per axis).
// Pass 0 reads the occupancy: 0 where occ == target, kSkipCap elsewhere.
// Neighbours outside the storage are left out (the border bricks inside it
// are empty and nearer).
__global__ void k_brick_dist(const uint8_t *__restrict__ in, uint8_t *__restrict__ out, int NBx, int NBy, int NBz,
                             int axis, int from_occ, int target) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nb = (long long)NBx * NBy * NBz;
    if (i >= nb) return;
    const int bx = (int)(i % NBx);
    const int by = (int)((i / NBx) % NBy);
    const int bz = (int)(i / ((long long)NBx * NBy));
    const int c = axis == 0 ? bx : (axis == 1 ? by : bz);
    const int n = axis == 0 ? NBx : (axis == 1 ? NBy : NBz);
    const long long stride = axis == 0 ? 1 : (axis == 1 ? (long long)NBx : (long long)NBx * NBy);
    int best = lgd::kSkipCap;
    for (int k = -(lgd::kSkipCap - 1); k < lgd::kSkipCap; ++k) {
        const int cc = c + k;
        if (cc < 0 || cc >= n) continue;
        const uint8_t v = in[i + k * stride];
        const int d = from_occ ? (((v != 0) == (target != 0)) ? 0 : lgd::kSkipCap) : (int)v;
        const int ak = k < 0 ? -k : k;
        const int m = d > ak ? d : ak;
        best = m < best ? m : best;
    }
    out[i] = (uint8_t)best;
}

// Box codes: an empty brick keeps its distance to the nearest existing brick,
// an existing one gets kSkipExists + its distance to the nearest empty brick.
__global__ void k_brick_code(const uint8_t *__restrict__ occ, const uint8_t *__restrict__ to_existing,
                             const uint8_t *__restrict__ to_empty, long long nb, uint8_t *__restrict__ code) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    code[i] = occ[i] ? (uint8_t)(lgd::kSkipExists + to_empty[i]) : to_existing[i];
}

// Brick storage back to a dense float volume (x fastest).
template <typename T>
__global__ void k_bricks_to_slab(const T *__restrict__ data, int z0, int nzs, GridDims d,
                                 float *__restrict__ slab) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per = (long long)d.nx * d.ny;
    if (i >= per * nzs) return;
    const int z = z0 + (int)(i / per);
    const int rem = (int)(i % per);
    const int y = rem / d.nx, x = rem % d.nx;
    const long long brick = ((long long)((z >> 3) + 1) * d.NBy + ((y >> 3) + 1)) * d.NBx + ((x >> 3) + 1);
    const long long off = (brick << 9) + ((z & 7) << 6) + ((y & 7) << 3) + (x & 7);
    slab[i] = lgd::to_f32(data[off]);
}

int32_t grid_alloc(lg_ctx *ctx, const lg_grid_desc *desc, lg_grid **out) {
    LG_CHECK_ARG(desc && desc->nx > 0 && desc->ny > 0 && desc->nz > 0, "grid: dimensions must be positive");
    LG_CHECK_ARG(desc->nx <= 16384 && desc->ny <= 16384 && desc->nz <= 16384, "grid: dimension too large");
    LG_CHECK_ARG(desc->dtype == LG_F32 || desc->dtype == LG_F16 || desc->dtype == LG_BF16, "grid: bad dtype");
    LG_CHECK_ARG(desc->voxel_size > 0.0, "grid: voxel size must be positive (uniform voxels)");
    lg_grid *g = new lg_grid();
    g->desc = *desc;
    const int n[3] = {desc->nx, desc->ny, desc->nz};
    for (int c = 0; c < 3; ++c) {
        g->nb[c] = (n[c] + 7) / 8;
        g->NB[c] = g->nb[c] + 2;
    }
    const size_t nbricks = (size_t)g->NB[0] * g->NB[1] * g->NB[2];
    const size_t bytes = grid_storage_elems(g) * dtype_bytes(desc->dtype);
    int32_t rc = dmalloc(ctx, bytes, &g->data);
    if (rc == LG_OK) rc = dmalloc(ctx, nbricks, (void **)&g->occ);
    if (rc != LG_OK) {
        dfree(ctx, g->data);
        dfree(ctx, g->occ);
        delete g;
        return rc;
    }
    g->bytes = bytes + nbricks;
    hipError_t e = hipMemsetAsync(g->data, 0, bytes, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(g->occ, 0, nbricks, ctx->stream);
    if (e != hipSuccess) {
        set_error("grid memset failed: %s", hipGetErrorString(e));
        dfree(ctx, g->data);
        dfree(ctx, g->occ);
        delete g;
        return LG_ERR_HIP;
    }
    *out = g;
    return LG_OK;
}

// Occupancy, active count and active bbox of a grid whose storage is filled.
int32_t grid_finalize(lg_ctx *ctx, lg_grid *g) {
    const long long nbricks = (long long)g->NB[0] * g->NB[1] * g->NB[2];
    unsigned long long *d_cnt = nullptr;
    void *s = nullptr;
    LG_TRY(scratch(ctx, 256 + 2 * (size_t)nbricks, &s));
    d_cnt = (unsigned long long *)s;
    uint8_t *d_a = (uint8_t *)s + 256, *d_b = d_a + nbricks;
    if (!g->skip) {
        LG_TRY(dmalloc(ctx, (size_t)nbricks, (void **)&g->skip));
        g->bytes += (uint64_t)nbricks;
    }
    LG_HIP(hipMemsetAsync(d_cnt, 0, sizeof(unsigned long long), ctx->stream));
    switch (g->desc.dtype) {
        case LG_F32:
            k_brick_occupancy<float><<<nbricks, 256, 0, ctx->stream>>>((const float *)g->data, nbricks, g->occ, d_cnt);
            break;
        case LG_F16:
            k_brick_occupancy<_Float16><<<nbricks, 256, 0, ctx->stream>>>((const _Float16 *)g->data, nbricks, g->occ, d_cnt);
            break;
        default:
            k_brick_occupancy<lgd::bf16_t><<<nbricks, 256, 0, ctx->stream>>>((const lgd::bf16_t *)g->data, nbricks, g->occ, d_cnt);
            break;
    }
    LG_LAUNCHED("k_brick_occupancy");
    {
        const unsigned blocks = (unsigned)((nbricks + 255) / 256);
        const int NBx = g->NB[0], NBy = g->NB[1], NBz = g->NB[2];
        // distance to the nearest existing brick into a, to the nearest empty one into b
        k_brick_dist<<<blocks, 256, 0, ctx->stream>>>(g->occ, d_a, NBx, NBy, NBz, 0, 1, 1);
        k_brick_dist<<<blocks, 256, 0, ctx->stream>>>(d_a, g->skip, NBx, NBy, NBz, 1, 0, 0);
        k_brick_dist<<<blocks, 256, 0, ctx->stream>>>(g->skip, d_a, NBx, NBy, NBz, 2, 0, 0);
        k_brick_dist<<<blocks, 256, 0, ctx->stream>>>(g->occ, d_b, NBx, NBy, NBz, 0, 1, 0);
        k_brick_dist<<<blocks, 256, 0, ctx->stream>>>(d_b, g->skip, NBx, NBy, NBz, 1, 0, 0);
        k_brick_dist<<<blocks, 256, 0, ctx->stream>>>(g->skip, d_b, NBx, NBy, NBz, 2, 0, 0);
        LG_LAUNCHED("k_brick_dist");
        k_brick_code<<<blocks, 256, 0, ctx->stream>>>(g->occ, d_a, d_b, nbricks, g->skip);
        LG_LAUNCHED("k_brick_code");
    }
    std::vector<uint8_t> occ((size_t)nbricks);
    unsigned long long cnt = 0;
    LG_HIP(hipMemcpyAsync(occ.data(), g->occ, nbricks, hipMemcpyDeviceToHost, ctx->stream));
    LG_HIP(hipMemcpyAsync(&cnt, d_cnt, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream));
    LG_HIP(hipStreamSynchronize(ctx->stream));
    g->active_voxels = cnt;
    g->active_bricks = 0;
    for (int c = 0; c < 3; ++c) {
        g->bmin[c] = 1 << 30;
        g->bmax[c] = -(1 << 30);
    }
    for (int bz = 0; bz < g->NB[2]; ++bz)
        for (int by = 0; by < g->NB[1]; ++by)
            for (int bx = 0; bx < g->NB[0]; ++bx) {
                if (!occ[((size_t)bz * g->NB[1] + by) * g->NB[0] + bx]) continue;
                ++g->active_bricks;
                const int b[3] = {bx - 1, by - 1, bz - 1};
                for (int c = 0; c < 3; ++c) {
                    g->bmin[c] = std::min(g->bmin[c], b[c]);
                    g->bmax[c] = std::max(g->bmax[c], b[c]);
                }
            }
    if (!g->active_bricks) {
        for (int c = 0; c < 3; ++c) {
            g->bmin[c] = 0;
            g->bmax[c] = -1;
        }
    }
    return LG_OK;
}

// ------------------------------------------------------ brick counters ----
__global__ void __launch_bounds__(256) k_touch_count(uint8_t *__restrict__ flags, size_t n,
                                                     unsigned long long *__restrict__ stats, int counter) {
    unsigned long long c = 0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        c += flags[i] != 0;
        flags[i] = 0;
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
    if ((threadIdx.x & 63) == 0 && c) stat_add(stats, counter, c);
}

__global__ void __launch_bounds__(256) k_touch_clear(uint8_t *__restrict__ flags, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) flags[i] = 0;
}

static size_t grid_storage_bricks(const lg_grid *g) { return (size_t)g->NB[0] * g->NB[1] * g->NB[2]; }

int32_t grid_touch_begin(lg_ctx *ctx, lg_grid *g) {
    const size_t n = grid_storage_bricks(g);
    if (!g->touched) LG_TRY(dalloc(ctx, n, &g->touched));
    k_touch_clear<<<(unsigned)std::min<size_t>((n + 255) / 256, 1024), 256, 0, ctx->stream>>>(g->touched, n);
    LG_LAUNCHED("k_touch_clear");
    return LG_OK;
}

int32_t grid_touch_count(lg_ctx *ctx, const lg_grid *g, unsigned long long *stats, int counter) {
    if (!g->touched) return LG_OK;
    const size_t n = grid_storage_bricks(g);
    k_touch_count<<<(unsigned)std::min<size_t>((n + 255) / 256, 1024), 256, 0, ctx->stream>>>(g->touched, n, stats,
                                                                                             counter);
    LG_LAUNCHED("k_touch_count");
    return LG_OK;
}

}  // namespace lg

using namespace lg;

// --------------------------------------------------------------- C ABI -------
extern "C" {

const char *lg_last_error(void) { return lg::g_err; }
int32_t lg_abi_version(void) { return LG_ABI_VERSION; }

int32_t lg_ctx_create(int32_t device, uint64_t hbm_budget_bytes, lg_ctx **out) {
    LG_CHECK_ARG(out, "lg_ctx_create: out is NULL");
    *out = nullptr;
    int count = 0;
    LG_HIP(hipGetDeviceCount(&count));
    LG_CHECK_ARG(device >= 0 && device < count, "lg_ctx_create: device %d not present (%d devices)", device, count);
    LG_HIP(hipSetDevice(device));
    lg_ctx *ctx = new lg_ctx();
    ctx->device = device;
    ctx->budget = hbm_budget_bytes;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        set_error("hipStreamCreate failed: %s", hipGetErrorString(e));
        return LG_ERR_HIP;
    }
    if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_light, hipEventDisableTiming) != hipSuccess) {
        if (ctx->side) (void)hipStreamDestroy(ctx->side);
        if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
        if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        set_error("lg_ctx_create: side stream or events could not be created");
        return LG_ERR_HIP;
    }
    for (auto &m : ctx->marks) (void)hipEventCreate(&m);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->cu_count = prop.multiProcessorCount;
    *out = ctx;
    return LG_OK;
}

int32_t lg_ctx_destroy(lg_ctx *ctx) {
    if (!ctx) return LG_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->side);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->fw_graph) (void)hipGraphExecDestroy(ctx->fw_graph);
    for (auto &kv : ctx->allocs) (void)hipFree(kv.first);
    ctx->allocs.clear();
    for (auto &m : ctx->marks)
        if (m) (void)hipEventDestroy(m);
    for (auto &p : ctx->ktimer.pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : ctx->ktimer.pool) (void)hipEventDestroy(e);
    destroy_stage_pipeline(ctx);
    (void)hipEventDestroy(ctx->ev_fork);
    (void)hipEventDestroy(ctx->ev_join);
    (void)hipEventDestroy(ctx->ev_light);
    (void)hipStreamDestroy(ctx->side);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return LG_OK;
}

int32_t lg_sync(lg_ctx *ctx) {
    LG_CHECK_ARG(ctx, "lg_sync: ctx is NULL");
    LG_HIP(hipStreamSynchronize(ctx->stream));
    return LG_OK;
}

int32_t lg_ctx_stream(lg_ctx *ctx, void **s) {
    LG_CHECK_ARG(ctx && s, "lg_ctx_stream: NULL argument");
    *s = (void *)ctx->stream;
    return LG_OK;
}

int32_t lg_ctx_bytes_in_use(lg_ctx *ctx, uint64_t *bytes) {
    LG_CHECK_ARG(ctx && bytes, "lg_ctx_bytes_in_use: NULL argument");
    *bytes = ctx->in_use;
    return LG_OK;
}

int32_t lg_ctx_set_kernel_timing(lg_ctx *ctx, int32_t enable) {
    LG_CHECK_ARG(ctx, "lg_ctx_set_kernel_timing: ctx is NULL");
    ctx->timing = enable != 0;
    return LG_OK;
}

int32_t lg_ctx_set_kernel_timing_only(lg_ctx *ctx, const char *kernel) {
    LG_CHECK_ARG(ctx, "lg_ctx_set_kernel_timing_only: ctx is NULL");
    ctx->timing_only = kernel ? kernel : "";
    return LG_OK;
}

int32_t lg_ctx_set_brick_counting(lg_ctx *ctx, int32_t enable) {
    LG_CHECK_ARG(ctx, "lg_ctx_set_brick_counting: ctx is NULL");
    ctx->count_bricks = enable != 0;
    return LG_OK;
}

int32_t lg_kernel_time(lg_ctx *ctx, const char *kernel, uint64_t *launches, double *total_ms) {
    LG_CHECK_ARG(ctx && kernel && launches && total_ms, "lg_kernel_time: NULL argument");
    ktimer_collect(ctx);
    auto it = ctx->ktimer.totals.find(kernel);
    *launches = it == ctx->ktimer.totals.end() ? 0 : it->second.first;
    *total_ms = it == ctx->ktimer.totals.end() ? 0.0 : it->second.second;
    return LG_OK;
}

int32_t lg_kernel_time_reset(lg_ctx *ctx) {
    LG_CHECK_ARG(ctx, "lg_kernel_time_reset: ctx is NULL");
    ktimer_collect(ctx);
    ctx->ktimer.totals.clear();
    return LG_OK;
}

int32_t lg_malloc(lg_ctx *ctx, size_t bytes, void **d_out) {
    LG_CHECK_ARG(ctx && d_out, "lg_malloc: NULL argument");
    return dmalloc(ctx, bytes, d_out);
}

int32_t lg_free(lg_ctx *ctx, void *d_ptr) {
    LG_CHECK_ARG(ctx, "lg_free: ctx is NULL");
    return dfree(ctx, d_ptr);
}

int32_t lg_memcpy_h2d(lg_ctx *ctx, void *d_dst, const void *h_src, size_t bytes) {
    LG_CHECK_ARG(ctx && d_dst && h_src, "lg_memcpy_h2d: NULL argument");
    LG_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    LG_HIP(hipStreamSynchronize(ctx->stream));
    return LG_OK;
}

int32_t lg_memcpy_d2h(lg_ctx *ctx, void *h_dst, const void *d_src, size_t bytes) {
    LG_CHECK_ARG(ctx && h_dst && d_src, "lg_memcpy_d2h: NULL argument");
    LG_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    LG_HIP(hipStreamSynchronize(ctx->stream));
    return LG_OK;
}

// Plain streaming copy, 16 B per lane per step: the reference point that HBM
// traffic counters are scaled against (its bytes are known exactly).
__global__ void __launch_bounds__(256) k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int32_t lg_copy_d(lg_ctx *ctx, void *d_dst, const void *d_src, size_t bytes) {
    LG_CHECK_ARG(ctx && d_dst && d_src && bytes % 16 == 0, "lg_copy_d: NULL pointer or size not a multiple of 16");
    LG_CHECK_ARG(((uintptr_t)d_dst | (uintptr_t)d_src) % 16 == 0, "lg_copy_d: pointers must be 16-byte aligned");
    if (bytes == 0) return LG_OK;
    const size_t n16 = bytes / 16;
    const unsigned blocks =
        (unsigned)std::max<size_t>(1, std::min<size_t>((n16 + 255) / 256, (size_t)std::max(ctx->cu_count, 1) * 16));
    hipEvent_t ev;
    ktimer_begin(ctx, "k_copy", &ev);
    k_copy<<<blocks, 256, 0, ctx->stream>>>((const uint4 *)d_src, (uint4 *)d_dst, n16);
    LG_LAUNCHED("k_copy");
    ktimer_end(ctx, "k_copy", ev);
    return LG_OK;
}

int32_t lg_memset_d(lg_ctx *ctx, void *d_dst, int32_t value, size_t bytes) {
    LG_CHECK_ARG(ctx && d_dst, "lg_memset_d: NULL argument");
    LG_HIP(hipMemsetAsync(d_dst, value, bytes, ctx->stream));
    return LG_OK;
}

int32_t lg_timer_mark(lg_ctx *ctx, int32_t slot) {
    LG_CHECK_ARG(ctx && slot >= 0 && slot < 16, "lg_timer_mark: bad slot");
    LG_HIP(hipEventRecord(ctx->marks[slot], ctx->stream));
    return LG_OK;
}

int32_t lg_timer_elapsed(lg_ctx *ctx, int32_t a, int32_t b, float *ms) {
    LG_CHECK_ARG(ctx && ms && a >= 0 && a < 16 && b >= 0 && b < 16, "lg_timer_elapsed: bad argument");
    LG_HIP(hipEventSynchronize(ctx->marks[b]));
    LG_HIP(hipEventElapsedTime(ms, ctx->marks[a], ctx->marks[b]));
    return LG_OK;
}

int32_t lg_grid_create(lg_ctx *ctx, const lg_grid_desc *desc, const float *h_dense, lg_grid **out) {
    LG_CHECK_ARG(ctx && desc && h_dense && out, "lg_grid_create: NULL argument");
    *out = nullptr;
    lg_grid *g = nullptr;
    LG_TRY(grid_alloc(ctx, desc, &g));
    GridDims d{desc->nx, desc->ny, desc->nz, g->NB[0], g->NB[1], g->NB[2]};
    const size_t slab_elems = (size_t)desc->nx * desc->ny * 8;
    float *slab = nullptr;
    int32_t rc = dalloc(ctx, slab_elems, &slab);
    if (rc != LG_OK) {
        lg_grid_destroy(ctx, g);
        return rc;
    }
    for (int z0 = 0; z0 < desc->nz && rc == LG_OK; z0 += 8) {
        const int nzs = std::min(8, desc->nz - z0);
        const size_t cnt = (size_t)desc->nx * desc->ny * nzs;
        hipError_t e = hipMemcpyAsync(slab, h_dense + (size_t)z0 * desc->nx * desc->ny, cnt * 4,
                                      hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) {
            set_error("grid upload failed: %s", hipGetErrorString(e));
            rc = LG_ERR_HIP;
            break;
        }
        const unsigned blocks = (unsigned)((cnt + 255) / 256);
        switch (desc->dtype) {
            case LG_F32:
                k_slab_to_bricks<float><<<blocks, 256, 0, ctx->stream>>>(slab, z0, nzs, d, (float *)g->data);
                break;
            case LG_F16:
                k_slab_to_bricks<_Float16><<<blocks, 256, 0, ctx->stream>>>(slab, z0, nzs, d, (_Float16 *)g->data);
                break;
            default:
                k_slab_to_bricks<lgd::bf16_t><<<blocks, 256, 0, ctx->stream>>>(slab, z0, nzs, d, (lgd::bf16_t *)g->data);
                break;
        }
        if (hipGetLastError() != hipSuccess) {
            set_error("k_slab_to_bricks launch failed");
            rc = LG_ERR_HIP;
        }
        // the slab buffer is reused: keep uploads and conversions ordered
        if (rc == LG_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) {
            set_error("grid upload sync failed");
            rc = LG_ERR_HIP;
        }
    }
    dfree(ctx, slab);
    if (rc == LG_OK) rc = grid_finalize(ctx, g);
    if (rc != LG_OK) {
        lg_grid_destroy(ctx, g);
        return rc;
    }
    *out = g;
    return LG_OK;
}

int32_t lg_grid_info_get(const lg_grid *g, lg_grid_info *info) {
    LG_CHECK_ARG(g && info, "lg_grid_info_get: NULL argument");
    memset(info, 0, sizeof(*info));
    info->desc = g->desc;
    for (int c = 0; c < 3; ++c) {
        info->bricks[c] = g->nb[c];
        info->bbox_min[c] = 8 * g->bmin[c];
        info->bbox_max[c] = 8 * (g->bmax[c] + 1);
    }
    info->active_bricks = g->active_bricks;
    info->active_voxels = g->active_voxels;
    info->device_bytes = g->bytes;
    return LG_OK;
}

int32_t lg_grid_read_dense(lg_ctx *ctx, const lg_grid *g, float *h_out) {
    LG_CHECK_ARG(ctx && g && h_out, "lg_grid_read_dense: NULL argument");
    const lg_grid_desc &desc = g->desc;
    GridDims d{desc.nx, desc.ny, desc.nz, g->NB[0], g->NB[1], g->NB[2]};
    const size_t slab_elems = (size_t)desc.nx * desc.ny * 8;
    float *slab = nullptr;
    LG_TRY(dalloc(ctx, slab_elems, &slab));
    int32_t rc = LG_OK;
    for (int z0 = 0; z0 < desc.nz; z0 += 8) {
        const int nzs = std::min(8, desc.nz - z0);
        const size_t cnt = (size_t)desc.nx * desc.ny * nzs;
        const unsigned blocks = (unsigned)((cnt + 255) / 256);
        switch (desc.dtype) {
            case LG_F32:
                k_bricks_to_slab<float><<<blocks, 256, 0, ctx->stream>>>((const float *)g->data, z0, nzs, d, slab);
                break;
            case LG_F16:
                k_bricks_to_slab<_Float16><<<blocks, 256, 0, ctx->stream>>>((const _Float16 *)g->data, z0, nzs, d, slab);
                break;
            default:
                k_bricks_to_slab<lgd::bf16_t><<<blocks, 256, 0, ctx->stream>>>((const lgd::bf16_t *)g->data, z0, nzs, d, slab);
                break;
        }
        hipError_t e = hipMemcpyAsync(h_out + (size_t)z0 * desc.nx * desc.ny, slab, cnt * 4, hipMemcpyDeviceToHost,
                                      ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            set_error("grid download failed: %s", hipGetErrorString(e));
            rc = LG_ERR_HIP;
            break;
        }
    }
    dfree(ctx, slab);
    return rc;
}

int32_t lg_grid_destroy(lg_ctx *ctx, lg_grid *g) {
    if (!g) return LG_OK;
    LG_CHECK_ARG(ctx, "lg_grid_destroy: ctx is NULL");
    dfree(ctx, g->data);
    dfree(ctx, g->occ);
    dfree(ctx, g->skip);
    dfree(ctx, g->active_list);
    dfree(ctx, g->mass_prefix);
    dfree(ctx, g->touched);
    delete g;
    return LG_OK;
}

}  // extern "C"
