// The reference's point sampler on the GPU: adaptive dart throwing
// (LightGraph::scatterPoints, LightGraph.h:206-274) with its own random stream
// (std::mt19937 seeded 854123 + iteration drawn through
// uniform_real_distribution<float>(0, 1), RandGen.h:16-44), one workgroup per
// iteration.  The accepted set, its order and the final minimum distance are
// those of the sequential loop:
//   - the MT19937 state lives in LDS and is advanced one twist (624 words) at a
//     time, split into three phases so that every word reads exactly the values
//     the sequential recurrence reads; 624 words are 208 candidates (3 draws each);
//   - within a run of candidates that share the squared minimum distance, the
//     density test and the collision test against the points accepted before the
//     run are independent per candidate and run in parallel; one wave then walks
//     the run in throw order, tests the survivors against the points accepted
//     within the run, accepts, and stops once more than `target` are accepted;
//   - a run ends at each rescale point n ((n + 1) % (scatter / 100) == 0), where
//     the squared distance is multiplied by accepted / expected(n); expected(n)
//     comes from the host, computed with std::pow exactly as the reference does.
// Also here: the active-voxel bounding box at voxel resolution
// (grid->evalActiveVoxelBoundingBox(), main.cc:512), the box the darts fall in.
#include <algorithm>
#include <cmath>
#include <vector>

#include "lg_internal.h"

namespace lg {
namespace {

constexpr int kDartThreads = 256;
constexpr int kDartMaxPoints = 2048;  // accepted points kept in LDS (the reference asserts < 1000)
constexpr int kCellG = 16;            // collision cells per axis
constexpr int kCells = kCellG * kCellG * kCellG;
constexpr int kMtN = 624;
constexpr int kChunk = kMtN / 3;      // candidates per twist
constexpr uint32_t kUpper = 0x80000000u, kLower = 0x7fffffffu, kMatrixA = 0x9908b0dfu;

struct DartConsts {
    double lo[3];     // mBbox.min() (world)
    float sz[3];      // Vec3s(mBbox.max() - mBbox.min())
    double cell_min;  // smallest collision cell: the longest box side / kCellG
    uint32_t scatter; // mScatterCount
    uint32_t target;  // TARGET_PC_SIZE
    uint32_t batch;   // mScatterCount / PC_GROWTH_NUMBER_OF_ESTIMATE
    int32_t out_cap;
};

__device__ __forceinline__ uint32_t mt_next_word(uint32_t cur, uint32_t nxt, uint32_t far) {
    const uint32_t y = (cur & kUpper) | (nxt & kLower);
    return far ^ (y >> 1) ^ ((y & 1u) ? kMatrixA : 0u);
}

// One MT19937 twist of the LDS state.  The sequential recurrence updates
// word k from words k, k+1 (still old) and k+397 mod 624 (new once k >= 227);
// the three phases [0,227), [227,454), [454,624) each read only words that
// phase has not written and that earlier phases have finished.
__device__ __forceinline__ void mt_twist(uint32_t *mt, int tid) {
    uint32_t v = 0;
    if (tid < 227) v = mt_next_word(mt[tid], mt[tid + 1], mt[tid + 397]);
    __syncthreads();
    if (tid < 227) mt[tid] = v;
    __syncthreads();
    if (tid < 227) v = mt_next_word(mt[227 + tid], mt[228 + tid], mt[tid]);
    __syncthreads();
    if (tid < 227) mt[227 + tid] = v;
    __syncthreads();
    if (tid < 170) {
        const int i = 454 + tid;
        v = mt_next_word(mt[i], i == kMtN - 1 ? mt[0] : mt[i + 1], mt[i - 227]);
    }
    __syncthreads();
    if (tid < 170) mt[454 + tid] = v;
    __syncthreads();
}

// Tempering, then uniform_real_distribution<float>(0, 1) as libstdc++ draws it
// from a 32-bit engine: float(x) / 2^32, folded below 1.
__device__ __forceinline__ float mt_uniform(uint32_t y) {
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    const float r = (float)y / 4294967296.0f;
    return r >= 1.0f ? 0x1.fffffep-1f : r;
}

// Collision cells of side h >= sqrt(msq) (1 + 2^-10): a point closer than the
// minimum distance differs from the candidate by less than h on every axis, so
// their floor((x - lo) / h) indices differ by at most one (the margin covers the
// rounding of the two quotients) and the 27 cells around the candidate hold it.
__device__ __forceinline__ double cell_side(double msq, double cell_min) {
    return fmax(sqrt(msq) * (1.0 + 0x1p-10), cell_min);  // fmax drops a NaN distance
}

__device__ __forceinline__ int cell_axis(float x, double lo, double h) {
    const double q = floor(((double)x - lo) / h);
    return q < 0.0 ? 0 : (q > (double)(kCellG - 1) ? kCellG - 1 : (int)q);
}

__device__ __forceinline__ int cell_of(const float *p, const DartConsts &c, double h) {
    return (cell_axis(p[2], c.lo[2], h) * kCellG + cell_axis(p[1], c.lo[1], h)) * kCellG + cell_axis(p[0], c.lo[0], h);
}

__device__ __forceinline__ bool near_any(const float *p, const float (*acc)[3], const int *head, const int *next,
                                         const DartConsts &c, double h, double msq) {
    const int cx = cell_axis(p[0], c.lo[0], h), cy = cell_axis(p[1], c.lo[1], h), cz = cell_axis(p[2], c.lo[2], h);
    const double x = p[0], y = p[1], z = p[2];
    for (int z0 = max(cz - 1, 0); z0 <= min(cz + 1, kCellG - 1); ++z0)
        for (int y0 = max(cy - 1, 0); y0 <= min(cy + 1, kCellG - 1); ++y0)
            for (int x0 = max(cx - 1, 0); x0 <= min(cx + 1, kCellG - 1); ++x0)
                for (int i = head[(z0 * kCellG + y0) * kCellG + x0]; i >= 0; i = next[i]) {
                    const double dx = x - (double)acc[i][0], dy = y - (double)acc[i][1], dz = z - (double)acc[i][2];
                    if (dx * dx + dy * dy + dz * dz < msq) return true;
                }
    return false;
}

__device__ __forceinline__ bool too_close(const float *p, const float (*acc)[3], int i0, int i1, double msq) {
    const double x = p[0], y = p[1], z = p[2];
    for (int i = i0; i < i1; ++i) {
        const double dx = x - (double)acc[i][0], dy = y - (double)acc[i][1], dz = z - (double)acc[i][2];
        if (dx * dx + dy * dy + dz * dz < msq) return true;
    }
    return false;
}

template <typename T>
__global__ void __launch_bounds__(kDartThreads) k_dart_throw(lgd::GridView g, DartConsts c,
                                                             const uint32_t *__restrict__ iters,
                                                             const uint32_t *__restrict__ expected,
                                                             float *__restrict__ min_dist, float *__restrict__ out,
                                                             int32_t *__restrict__ counts) {
    __shared__ uint32_t mt[kMtN];
    __shared__ float cand[kChunk][3];
    __shared__ float acc[kDartMaxPoints][3];
    __shared__ uint8_t ok[kChunk];  // passes the density test and misses every point accepted before its run
    __shared__ int head[kCells];    // per collision cell: last accepted point in it (-1: none)
    __shared__ int next[kDartMaxPoints];
    __shared__ int s_pc, s_stop;
    __shared__ float s_msq, s_last;
    __shared__ double s_h;          // side of the cells the accepted points are filed under
    const int s = (int)blockIdx.x, tid = (int)threadIdx.x, lane = tid & 63;
    if (tid == 0) {
        uint32_t x = 854123u + iters[s];  // RandGen.h:28-29
        mt[0] = x;
        for (int i = 1; i < kMtN; ++i) {
            x = 1812433253u * (x ^ (x >> 30)) + (uint32_t)i;
            mt[i] = x;
        }
        const float md = min_dist[s];
        s_msq = md * md;
        s_last = s_msq;
        s_pc = 0;
        s_stop = 0;
        s_h = -1.0;
    }
    __syncthreads();
    for (uint32_t n0 = 0; n0 < c.scatter && !s_stop; n0 += kChunk) {
        mt_twist(mt, tid);
        const int nk = (int)min((uint32_t)kChunk, c.scatter - n0);
        if (tid < nk) {
            // newPos[i] = mRng(iter) * bboxSize[i] + mBbox.min()[i] (LightGraph.h:225-227)
            for (int a = 0; a < 3; ++a)
                cand[tid][a] = (float)((double)(mt_uniform(mt[3 * tid + a]) * c.sz[a]) + c.lo[a]);
        }
        __syncthreads();
        int a = 0;
        while (a < nk) {
            // the run [a, b): a rescale point starts a run; the next one ends it
            const uint32_t na = n0 + (uint32_t)a;
            const uint32_t next_rescale = ((na + 1) / c.batch + 1) * c.batch - 1;  // first rescale point > na
            const int b = (int)min((uint32_t)nk, next_rescale - n0);
            if (tid == 0 && (na + 1) % c.batch == 0 && !((uint32_t)s_pc > c.target)) {
                const uint32_t e = expected[(na + 1) / c.batch - 1];
                const double ratio = (double)(uint32_t)s_pc / (double)e;  // LightGraph.h:240-245
                s_msq = (float)((double)s_msq * ratio);
            }
            __syncthreads();
            const int p0 = s_pc;
            const double msq = (double)s_msq;
            const double h = cell_side(msq, c.cell_min);
            if (h != s_h) {
                // the distance changed at a rescale point: file the accepted points again
                __syncthreads();
                for (int i = tid; i < kCells; i += kDartThreads) head[i] = -1;
                __syncthreads();
                for (int i = tid; i < p0; i += kDartThreads) next[i] = atomicExch(&head[cell_of(acc[i], c, h)], i);
                if (tid == 0) s_h = h;
                __syncthreads();
            }
            if (tid >= a && tid < b) {
                const float *p = cand[tid];
                const float rho = lgd::trilinear<T>(g, ((double)p[0] - g.origin[0]) * g.inv_voxel,
                                                    ((double)p[1] - g.origin[1]) * g.inv_voxel,
                                                    ((double)p[2] - g.origin[2]) * g.inv_voxel);
                ok[tid] = !(rho < 0.001f) && !near_any(p, acc, head, next, c, h, msq);  // PC_DENSITY_CUTOFF
            }
            __syncthreads();
            if (tid < 64) {
                // throw order: survivors against the points accepted in this run
                int pc = p0;
                bool stop = (uint32_t)pc > c.target;
                float last = s_last;
                for (int base = a; base < b && !stop; base += 64) {
                    const int k = base + lane;
                    uint64_t m = __ballot(k < b && ok[k]);
                    while (m && !stop) {
                        const int kk = base + __builtin_ctzll(m);
                        m &= m - 1;
                        // lanes split the points accepted earlier in this run
                        bool hit = false;
                        for (int i = p0 + lane; i < pc; i += 64) hit = hit || too_close(cand[kk], acc, i, i + 1, msq);
                        if (__ballot(hit) != 0ull) continue;
                        if (pc >= kDartMaxPoints) {
                            stop = true;  // reported as an error by the host (count above the LDS limit)
                            pc = kDartMaxPoints + 1;
                            break;
                        }
                        if (lane < 3) acc[pc][lane] = cand[kk][lane];
                        if (lane == 0) {
                            const int cell = cell_of(cand[kk], c, h);
                            next[pc] = head[cell];
                            head[cell] = pc;
                        }
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                        ++pc;
                        last = s_msq;
                        stop = (uint32_t)pc > c.target;  // the next throw breaks (LightGraph.h:232-238)
                    }
                }
                if (lane == 0) {
                    s_pc = pc;
                    s_last = last;
                    if (stop) s_stop = 1;
                }
            }
            __syncthreads();
            if (s_stop) break;
            a = b;
        }
        __syncthreads();
    }
    const int pc = s_pc;
    const int nw = min(pc, min(c.out_cap, kDartMaxPoints));
    for (int i = tid; i < 3 * nw; i += kDartThreads) out[(long long)s * c.out_cap * 3 + i] = acc[i / 3][i % 3];
    if (tid == 0) {
        counts[s] = pc;
        min_dist[s] = (float)sqrt((double)s_last);  // sqrt(lastValidminDistSquared), LightGraph.h:266
    }
}

// evalActiveVoxelBoundingBox at voxel resolution: one workgroup per stored
// brick that holds an active voxel; bb = {min x, y, z, -max x, -max y, -max z}.
template <typename T>
__global__ void __launch_bounds__(256) k_active_bbox(lgd::GridView g, const int *__restrict__ nvox, int *__restrict__ bb) {
    const long long brick = blockIdx.x;
    if (!g.occ[brick]) return;
    const int bx = (int)(brick % g.NBx) - 1, by = (int)((brick / g.NBx) % g.NBy) - 1,
              bz = (int)(brick / ((long long)g.NBx * g.NBy)) - 1;
    const T *d = static_cast<const T *>(g.data) + (brick << 9);
    int mn[3] = {1 << 30, 1 << 30, 1 << 30}, mx[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
    for (int v = threadIdx.x; v < 512; v += 256) {
        const int x = 8 * bx + (v & 7), y = 8 * by + ((v >> 3) & 7), z = 8 * bz + (v >> 6);
        if (x < nvox[0] && y < nvox[1] && z < nvox[2] && lgd::to_f32(d[v]) != 0.0f) {
            mn[0] = min(mn[0], x);
            mn[1] = min(mn[1], y);
            mn[2] = min(mn[2], z);
            mx[0] = max(mx[0], x);
            mx[1] = max(mx[1], y);
            mx[2] = max(mx[2], z);
        }
    }
    for (int a = 0; a < 3; ++a) {
        for (int o = 32; o > 0; o >>= 1) {
            mn[a] = min(mn[a], __shfl_xor(mn[a], o, 64));
            mx[a] = max(mx[a], __shfl_xor(mx[a], o, 64));
        }
    }
    if ((threadIdx.x & 63) == 0 && mn[0] <= mx[0])
        for (int a = 0; a < 3; ++a) {
            atomicMin(&bb[a], mn[a]);
            atomicMin(&bb[3 + a], -mx[a]);
        }
}

template <typename F>
void for_grid_type(int dtype, F &&f) {
    if (dtype == LG_F32)
        f(float{});
    else if (dtype == LG_F16)
        f(_Float16{});
    else
        f(lgd::bf16_t{});
}

}  // namespace
}  // namespace lg

using namespace lg;

extern "C" {

int32_t lg_grid_active_bbox(lg_ctx *ctx, const lg_grid *grid, int32_t vmin[3], int32_t vmax[3]) {
    LG_CHECK_ARG(ctx && grid && vmin && vmax, "lg_grid_active_bbox: NULL argument");
    const long long nbricks = (long long)grid->NB[0] * grid->NB[1] * grid->NB[2];
    void *s = nullptr;
    LG_TRY(scratch(ctx, 64, &s));
    int *d = (int *)s;
    int h[9] = {1 << 30, 1 << 30, 1 << 30, 1 << 30, 1 << 30, 1 << 30, grid->desc.nx, grid->desc.ny, grid->desc.nz};
    LG_HIP(hipMemcpyAsync(d, h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
    if (grid->active_bricks > 0) {
        const lgd::GridView v = make_view(grid);
        for_grid_type(grid->desc.dtype, [&](auto tag) {
            using T = decltype(tag);
            k_active_bbox<T><<<(unsigned)nbricks, 256, 0, ctx->stream>>>(v, d + 6, d);
        });
        LG_LAUNCHED("k_active_bbox");
    }
    LG_HIP(hipMemcpyAsync(h, d, 6 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    LG_HIP(hipStreamSynchronize(ctx->stream));
    for (int a = 0; a < 3; ++a) {
        vmin[a] = h[a];
        vmax[a] = -h[3 + a];
    }
    if (h[0] == (1 << 30)) {  // no active voxel: an empty box (min > max)
        for (int a = 0; a < 3; ++a) {
            vmin[a] = 0;
            vmax[a] = -1;
        }
    }
    return LG_OK;
}

int32_t lg_dart_throw(lg_ctx *ctx, const lg_grid *density, const double bbox_min[3], const double bbox_max[3],
                      int32_t count, const uint32_t *h_iters, uint32_t scatter_count, uint32_t target,
                      float *h_min_dist, float *d_points_out, int32_t out_cap, int32_t *h_counts) {
    LG_CHECK_ARG(ctx && density && bbox_min && bbox_max && h_iters && h_min_dist && d_points_out && h_counts,
                 "lg_dart_throw: NULL argument");
    LG_CHECK_ARG(count >= 1 && out_cap >= 1, "lg_dart_throw: count=%d, out_cap=%d must be >= 1", count, out_cap);
    LG_CHECK_ARG(scatter_count >= 100, "lg_dart_throw: scatter_count=%u must be >= 100 (batch = scatter / 100)",
                 scatter_count);
    LG_CHECK_ARG(target + 1 <= (uint32_t)kDartMaxPoints,
                 "lg_dart_throw: target=%u, at most %d points are kept per iteration", target, kDartMaxPoints - 1);
    DartConsts c{};
    for (int a = 0; a < 3; ++a) {
        c.lo[a] = bbox_min[a];
        c.sz[a] = (float)(bbox_max[a] - bbox_min[a]);
    }
    c.cell_min = std::max(std::max((double)c.sz[0], (double)c.sz[1]), (double)c.sz[2]) / kCellG;
    if (!(c.cell_min > 0.0)) c.cell_min = 1.0;  // a flat or empty box: one cell covers it
    c.scatter = scatter_count;
    c.target = target;
    c.batch = scatter_count / 100u;  // PC_GROWTH_NUMBER_OF_ESTIMATE
    c.out_cap = out_cap;
    // expected size at each rescale point n = k * batch - 1 (LightGraph.h:241-242)
    const uint32_t nres = scatter_count / c.batch;
    std::vector<uint32_t> up((size_t)nres + (size_t)count);
    for (uint32_t k = 1; k <= nres; ++k) {
        const size_t n = (size_t)k * c.batch - 1;
        // PC_GROWTH_FUNC_POWER is a float constant (LightGraph.h:28): pow gets (double)0.17f
        up[k - 1] = (uint32_t)(target * std::pow((double)n / (double)scatter_count, (double)0.17f));
    }
    for (int s = 0; s < count; ++s) up[nres + (size_t)s] = h_iters[s];
    const size_t up_bytes = up.size() * sizeof(uint32_t);
    const size_t f_bytes = (size_t)count * sizeof(float), n_bytes = (size_t)count * sizeof(int32_t);
    void *s = nullptr;
    LG_TRY(scratch(ctx, up_bytes + f_bytes + n_bytes + 512, &s));
    uint32_t *d_exp = (uint32_t *)s;
    uint32_t *d_iters = d_exp + nres;
    float *d_md = (float *)((char *)s + ((up_bytes + 255) / 256) * 256);
    int32_t *d_cnt = (int32_t *)(d_md + count);
    LG_HIP(hipMemcpyAsync(d_exp, up.data(), up_bytes, hipMemcpyHostToDevice, ctx->stream));
    LG_HIP(hipMemcpyAsync(d_md, h_min_dist, f_bytes, hipMemcpyHostToDevice, ctx->stream));
    const lgd::GridView v = make_view(density);
    for_grid_type(density->desc.dtype, [&](auto tag) {
        using T = decltype(tag);
        k_dart_throw<T><<<(unsigned)count, kDartThreads, 0, ctx->stream>>>(v, c, d_iters, d_exp, d_md, d_points_out,
                                                                         d_cnt);
    });
    LG_LAUNCHED("k_dart_throw");
    LG_HIP(hipMemcpyAsync(h_min_dist, d_md, f_bytes, hipMemcpyDeviceToHost, ctx->stream));
    LG_HIP(hipMemcpyAsync(h_counts, d_cnt, n_bytes, hipMemcpyDeviceToHost, ctx->stream));
    LG_HIP(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < count; ++k)
        if (h_counts[k] > kDartMaxPoints) {
            set_error("lg_dart_throw: iteration %d accepted more than %d points", k, kDartMaxPoints);
            return LG_ERR_OVERFLOW;
        }
    return LG_OK;
}

}  // extern "C"
