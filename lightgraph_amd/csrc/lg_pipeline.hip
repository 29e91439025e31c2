// The LightGraph solve pipeline for S >= 1 independent solves at once
// (LightGraph::operator(), LightGraph.h:158-196), organised around the
// connected components of each light graph:
//
//   sample (a) -> vertex light -> candidate CSR -> capped greedy selection ->
//   edge march (b) -> union-find over finite edges -> radix sort by component ->
//   per component: min-plus closure (c) + column-normalised diffusion.
//
// Optical depth between vertices of different components is +INF (the
// reference's zero transmittance), so closing each component separately gives
// the same matrix as closing the whole graph, with the same pivot order inside a
// component (members keep ascending vertex order).  Components of at most 128
// vertices are closed and diffused in LDS by one workgroup each, all in one
// launch; larger ones go through the blocked Floyd-Warshall kernels.
//
// The greedy degree cap (LightGraph.h:297-310) is order dependent, but only for
// edges touching a vertex whose candidate degree exceeds the cap ("risky"):
// every other candidate edge is always accepted.  Only rows holding such edges
// are swept sequentially, one wave per solve.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "lg_internal.h"
#include "lg_lights.h"
#include "lg_pipeline.h"
#include "lg_fw_group.h"

namespace lg {

// ------------------------------------------------------------- utilities ----
__device__ __forceinline__ int atomic_load_i32(int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_store_i32(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_solve_ids(const SolveDesc *__restrict__ sd, int S, int V, int *__restrict__ sid) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    int lo = 0, hi = S - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sd[mid].voff <= v)
            lo = mid;
        else
            hi = mid - 1;
    }
    sid[v] = lo;
}

// ------------------------------------------------------------ sampling ------
// one wave per point
template <typename T>
__global__ void __launch_bounds__(256) k_sample(const SolveDesc *__restrict__ sd, const int *__restrict__ sid, int V,
                                                uint32_t k0, uint32_t k1, float *__restrict__ pts) {
    const int v = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (v >= V) return;  // whole waves leave together
    const SolveDesc &d = sd[sid[v]];
    sample_point_wave<T>(d, v - d.voff, k0, k1, pts + 3 * (long long)v);
}

// ----------------------------------------------------- vertex lighting ------
template <typename T>
__global__ void k_light(const SolveDesc *__restrict__ sd, const int *__restrict__ sid, int V, LightConsts lc,
                        lgl::DevLights L, uint32_t ek0, uint32_t ek1, const float *__restrict__ pts,
                        float *__restrict__ out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const SolveDesc &d = sd[sid[v]];
    const int i = v - d.voff;
    const double p[3] = {(double)pts[3 * v], (double)pts[3 * v + 1], (double)pts[3 * v + 2]};
    double acc[3];
    vertex_light_one<T>(d.g, d.tg, d.has_temp, lc, L, ek0, ek1, (uint32_t)i, d.solve, p, acc);
    for (int k = 0; k < 3; ++k) out[3 * v + k] = (float)acc[k];
}

// Morton key of a vertex's position on a 32^3 lattice over its grid's clip box
// (15 bits): the order in which large calls trace their shadow rays.
__device__ __forceinline__ uint32_t morton_spread5(uint32_t x) {
    x &= 0x1fu;
    x = (x | (x << 8)) & 0x0000f00fu;
    x = (x | (x << 4)) & 0x000c30c3u;
    x = (x | (x << 2)) & 0x00249249u;
    return x;
}

__global__ void k_light_keys(const SolveDesc *__restrict__ sd, const int *__restrict__ sid, int V,
                             const float *__restrict__ pts, uint32_t *__restrict__ keys, int *__restrict__ vals) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const lgd::GridView &g = sd[sid[v]].g;
    uint32_t key = 0;
    for (int c = 0; c < 3; ++c) {
        const double f = floor(((double)pts[3 * v + c] - g.bbmin[c]) * 32.0 / (g.bbmax[c] - g.bbmin[c]));
        const uint32_t q = f > 0.0 ? (f < 31.0 ? (uint32_t)f : 31u) : 0u;
        key |= morton_spread5(q) << c;
    }
    keys[v] = key;
    vals[v] = v;
}

// Vertex lighting for exact steps: one wave per (vertex, light sample) shadow
// ray, then one lane per vertex sums the samples in the order of
// computeLighting (LightGraph.h:446-497).  Without `order` the waves take the
// rays vertex-major.  With it (large calls) they take them light-sample-major,
// the vertices of each sample in the given spatial order, so that concurrent
// rays share bricks in L2 (workgroups stay round-robin over the XCDs: giving
// each XCD one contiguous run of the order measured slower).  Which wave traces
// a ray does not change the ray or where its result goes.
template <typename T, bool kCount>
__global__ void __launch_bounds__(256) k_light_rays(const SolveDesc *__restrict__ sd, const int *__restrict__ sid,
                                                    int V, int nt, LightConsts lc, lgl::DevLights L, uint32_t ek0,
                                                    uint32_t ek1, const float *__restrict__ pts,
                                                    const int *__restrict__ order, double *__restrict__ ray,
                                                    unsigned long long *__restrict__ stats) {
    const long long w = (long long)blockIdx.x * (blockDim.x >> 6) + lgd::wave_uniform((int)(threadIdx.x >> 6));
    if (w >= (long long)V * nt) return;
    int v, t;
    if (order) {
        t = (int)(w / V);
        v = order[w - (long long)t * V];
    } else {
        v = (int)(w / nt);
        t = (int)(w - (long long)v * nt);
    }
    v = lgd::wave_uniform(v);
    const long long slot = (long long)v * nt + t;
    int li;
    uint32_t ew;
    light_task(L, t, &li, &ew);
    const SolveDesc &d = sd[sid[v]];
    const double p[3] = {(double)pts[3 * v], (double)pts[3 * v + 1], (double)pts[3 * v + 2]};
    lgd::Stream rng;
    rng.init((uint32_t)(v - d.voff), d.solve, lgd::DOM_VERTEX_ENV, ((uint64_t)ek1 << 32) | ek0);
    rng.skip(ew);
    double ldir[3], inten[3], distance;
    lgl::illuminate(L.l[li], p, rng, ldir, inten, distance);
    for (int k = 0; k < 3; ++k) inten[k] = lgd::wave_uniform(inten[k]);
    double tau;
    bool dropped;
    double r[3] = {0.0, 0.0, 0.0};
    int ns = 0;
    if (shadow_tau_cells<T, kCount>(d.g, lc, p, ldir, distance, &tau, &dropped, &ns) && !dropped)
        for (int k = 0; k < 3; ++k) r[k] = inten[k] * (double)expf((float)(lc.ext[k] * tau));
    if ((threadIdx.x & 63) == 0) {
        for (int k = 0; k < 3; ++k) ray[3 * slot + k] = r[k];
        if (stats) {
            stat_add(stats, ST_SHADOW_RAYS, 1ull);
            stat_add(stats, ST_SHADOW_SAMPLES, (unsigned long long)ns);
        }
    }
}

// The 16-lane form for large calls: DPP row r of the wave (lanes 16r..16r+15)
// traces ray 4 * wave + r, in the same order (vertex-major, or sample-major over
// `order`), with the same results as k_light_rays.  Run alone it is slower (a
// C2 solve's 16K rays fill only a quarter of the waves); beside the main
// stream's work of a large batch it leaves that work more of the GPU.
template <typename T, bool kCount>
__global__ void __launch_bounds__(256) k_light_rays16(const SolveDesc *__restrict__ sd, const int *__restrict__ sid,
                                                      int V, int nt, LightConsts lc, lgl::DevLights L, uint32_t ek0,
                                                      uint32_t ek1, const float *__restrict__ pts,
                                                      const int *__restrict__ order, double *__restrict__ ray,
                                                      unsigned long long *__restrict__ stats) {
    const long long w0 = ((long long)blockIdx.x * (blockDim.x >> 6) + lgd::wave_uniform((int)(threadIdx.x >> 6))) * 4;
    const long long total = (long long)V * nt;
    if (w0 >= total) return;  // whole waves leave together
    const int lane = threadIdx.x & 63;
    const long long w = w0 + (lane >> 4);
    const bool valid = w < total;
    double ldir[3] = {0.0, 0.0, 1.0}, inten[3] = {0.0, 0.0, 0.0}, distance = 0.0;
    double p[3] = {0.0, 0.0, 0.0};
    long long slot = 0;
    const SolveDesc *d = sd;
    if (valid) {
        int v, t;
        if (order) {
            t = (int)(w / V);
            v = order[w - (long long)t * V];
        } else {
            v = (int)(w / nt);
            t = (int)(w - (long long)v * nt);
        }
        slot = (long long)v * nt + t;
        int li;
        uint32_t ew;
        light_task(L, t, &li, &ew);
        d = &sd[sid[v]];
        for (int c = 0; c < 3; ++c) p[c] = (double)pts[3 * v + c];
        lgd::Stream rng;
        rng.init((uint32_t)(v - d->voff), d->solve, lgd::DOM_VERTEX_ENV, ((uint64_t)ek1 << 32) | ek0);
        rng.skip(ew);
        lgl::illuminate(L.l[li], p, rng, ldir, inten, distance);
    }
    double tau;
    bool hit, dropped;
    int ns = 0;
    shadow_tau_rows<T, kCount>(d->g, lc, p, ldir, distance, valid, &hit, &tau, &dropped, &ns);
    if (valid && (lane & 15) == 0) {
        double r[3] = {0.0, 0.0, 0.0};
        if (hit && !dropped)
            for (int k = 0; k < 3; ++k) r[k] = inten[k] * (double)expf((float)(lc.ext[k] * tau));
        for (int k = 0; k < 3; ++k) ray[3 * slot + k] = r[k];
        if (stats) {
            stat_add(stats, ST_SHADOW_RAYS, 1ull);
            stat_add(stats, ST_SHADOW_SAMPLES, (unsigned long long)ns);
        }
    }
}

__global__ void k_light_combine(const SolveDesc *__restrict__ sd, const int *__restrict__ sid, int V, int nt,
                                LightConsts lc, lgl::DevLights L, const float *__restrict__ pts,
                                const double *__restrict__ ray, float *__restrict__ out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const SolveDesc &d = sd[sid[v]];
    const double p[3] = {(double)pts[3 * v], (double)pts[3 * v + 1], (double)pts[3 * v + 2]};
    double emission[3];
    vertex_emission(d.tg, d.has_temp, lc, p, emission);
    const double *r = ray + 3 * (long long)v * nt;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int li = 0; li < L.n; ++li) {
        int ns = L.l[li].samples;
        if (ns > 1) ns *= 10;
        double lumi[3] = {0.0, 0.0, 0.0};
        for (int s = 0; s < ns; ++s, r += 3)
            for (int k = 0; k < 3; ++k) lumi[k] += r[k];
        for (int k = 0; k < 3; ++k) acc[k] += lc.albedo[k] * (lumi[k] / (double)ns);
    }
    for (int k = 0; k < 3; ++k) out[3 * v + k] = (float)(acc[k] + emission[k]);
}

// ---------------------------------------------------------- candidates ------

// ------------------------------------------------------- neighbour grid ----
// Candidates j > i with |p_i - p_j|^2 < r2 are found through a uniform grid of
// cells at least r wide per solve: points within r of each other sit in the same
// or adjacent cells (cell indices are clamped to the grid, which only merges
// cells, so adjacency still holds).  Points are radix-sorted by cell; the 3 cells
// of one x-run are consecutive in that order.
__device__ __forceinline__ int cell_axis(double p, double lo, double inv, int n) {
    const double f = floor((p - lo) * inv);
    if (!(f >= 0.0)) return 0;
    if (f > (double)(n - 1)) return n - 1;
    return (int)f;
}

// Block s: bounding box of solve s's points and its cell grid.
__device__ __forceinline__ void cell_setup_block(const SolveDesc &d, const float *__restrict__ pts, double cell_min,
                                                 CellDesc *__restrict__ out, CellDesc *__restrict__ out_lds) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int v = d.voff + threadIdx.x; v < d.voff + d.n; v += blockDim.x)
        for (int c = 0; c < 3; ++c) {
            lo[c] = fminf(lo[c], pts[3 * v + c]);
            hi[c] = fmaxf(hi[c], pts[3 * v + c]);
        }
    // min / max within each wave by shuffles, then across the waves in LDS
    __shared__ float s_lo[3][16], s_hi[3][16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            lo[c] = fminf(lo[c], __shfl_xor(lo[c], o, 64));
            hi[c] = fmaxf(hi[c], __shfl_xor(hi[c], o, 64));
        }
        if (lane == 0) {
            s_lo[c][w] = lo[c];
            s_hi[c][w] = hi[c];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c)
            for (int k = 1; k < nw; ++k) {
                s_lo[c][0] = fminf(s_lo[c][0], s_lo[c][k]);
                s_hi[c][0] = fmaxf(s_hi[c][0], s_hi[c][k]);
            }
    if (threadIdx.x == 0) {
        CellDesc c{};
        for (int a = 0; a < 3; ++a) {
            const double l = (double)s_lo[a][0], h = (double)s_hi[a][0];
            const double ext = h - l;
            if (!(ext >= 0.0) || !isfinite(ext) || !isfinite(l)) {
                c.lo[a] = 0.0;
                c.inv[a] = 0.0;
                c.n[a] = 1;
                continue;
            }
            const double cell = fmax(cell_min, ext / (double)d.cell_A);
            int n = (int)floor(ext / cell) + 1;
            if (n > d.cell_A + 1) n = d.cell_A + 1;
            if (n < 1) n = 1;
            c.lo[a] = l;
            c.inv[a] = 1.0 / cell;
            c.n[a] = n;
        }
        c.base = d.cell_base;
        *out = c;
        if (out_lds) *out_lds = c;
    }
}

__global__ void __launch_bounds__(1024) k_cell_setup(const SolveDesc *__restrict__ sd, const float *__restrict__ pts,
                                                     double cell_min, CellDesc *__restrict__ cd) {
    cell_setup_block(sd[blockIdx.x], pts, cell_min, cd + blockIdx.x, nullptr);
}

// The neighbour grid of a small call in one workgroup: a counting sort of the
// points by cell (LDS counters, ranks within a cell from the atomics), the
// cell ranges, and the sorted point list.  Points keep no order inside a cell:
// k_cand reads whole x-runs of cells and sorts each row itself.  With `setup`
// (a one-solve call) the workgroup first derives that solve's grid itself
// (cell_setup_block), which saves the k_cell_setup launch.
constexpr int kCellSortCells = 32768;  // LDS counters (128 KiB)
constexpr int kCellSortMaxV = 65536;   // ranks in 16 bits

__global__ void __launch_bounds__(1024) k_cell_sort_small(CellDesc *__restrict__ cd, const int *__restrict__ sid,
                                                          int V, const float *__restrict__ pts, int cells,
                                                          uint32_t *__restrict__ packed, int *__restrict__ cstart,
                                                          int *__restrict__ cend, int *__restrict__ svals,
                                                          int *__restrict__ zero_word, const SolveDesc *__restrict__ sd,
                                                          double cell_min, int setup) {
    __shared__ int cnt[kCellSortCells];
    __shared__ int wsum[16];
    __shared__ CellDesc s_cd;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (setup) {
        cell_setup_block(sd[0], pts, cell_min, cd, &s_cd);
        __syncthreads();
    }
    if (tid == 0) *zero_word = 0;  // the caller's count slot after the last vertex
    for (int k = tid; k < cells; k += 1024) cnt[k] = 0;
    __syncthreads();
    for (int v = tid; v < V; v += 1024) {
        const CellDesc c = setup ? s_cd : cd[sid[v]];
        const int x = cell_axis((double)pts[3 * v], c.lo[0], c.inv[0], c.n[0]);
        const int y = cell_axis((double)pts[3 * v + 1], c.lo[1], c.inv[1], c.n[1]);
        const int z = cell_axis((double)pts[3 * v + 2], c.lo[2], c.inv[2], c.n[2]);
        const int key = c.base + (z * c.n[1] + y) * c.n[0] + x;
        const int r = atomicAdd(&cnt[key], 1);
        packed[v] = ((uint32_t)key << 16) | (uint32_t)r;
    }
    __syncthreads();
    // exclusive scan of the counters in place: a run of `per` cells per thread
    const int per = (cells + 1023) / 1024;
    const int k0 = min(tid * per, cells), k1 = min(k0 + per, cells);
    int run = 0;
    for (int k = k0; k < k1; ++k) run += cnt[k];
    const int incl = lgd::wave_prefix_sum(run);
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int base = 0;
    for (int k = 0; k < w; ++k) base += wsum[k];
    int acc = base + incl - run;
    for (int k = k0; k < k1; ++k) {
        const int n = cnt[k];
        cnt[k] = acc;
        acc += n;
    }
    __syncthreads();
    for (int k = tid; k < cells; k += 1024) {
        const int a = cnt[k], b = k + 1 < cells ? cnt[k + 1] : V;
        cstart[k] = b > a ? a : 0;
        cend[k] = b > a ? b : 0;
    }
    for (int v = tid; v < V; v += 1024) {
        const uint32_t pk = packed[v];
        svals[cnt[pk >> 16] + (int)(pk & 0xffffu)] = v;
    }
}

__global__ void k_cell_keys(const CellDesc *__restrict__ cd, const int *__restrict__ sid, int V,
                            const float *__restrict__ pts, uint32_t *__restrict__ keys, int *__restrict__ vals) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const CellDesc c = cd[sid[v]];
    const int x = cell_axis((double)pts[3 * v], c.lo[0], c.inv[0], c.n[0]);
    const int y = cell_axis((double)pts[3 * v + 1], c.lo[1], c.inv[1], c.n[1]);
    const int z = cell_axis((double)pts[3 * v + 2], c.lo[2], c.inv[2], c.n[2]);
    keys[v] = (uint32_t)(c.base + (z * c.n[1] + y) * c.n[0] + x);
    vals[v] = v;
}

struct ZeroList {
    int *p[8];
    long long n[8];
    int count;
};

// Clears the regions of z with the whole grid (a kernel that also does other
// work clears buffers its own launch does not read).
__device__ __forceinline__ void grid_zero(const ZeroList &z) {
    const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (int r = 0; r < z.count; ++r)
        for (long long i = i0; i < z.n[r]; i += stride) z.p[r][i] = 0;
}

__global__ void k_zero(ZeroList z) { grid_zero(z); }

__global__ void k_cell_bounds(int V, const uint32_t *__restrict__ sk, int *__restrict__ cstart, int *__restrict__ cend) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= V) return;
    const uint32_t key = sk[k];
    if (k == 0 || sk[k - 1] != key) cstart[key] = k;
    if (k == V - 1 || sk[k + 1] != key) cend[key] = k + 1;
}

__device__ __forceinline__ double pdist2(const float *__restrict__ pts, int i, int j) {
    const double dx = (double)pts[3 * i + 0] - (double)pts[3 * j + 0];
    const double dy = (double)pts[3 * i + 1] - (double)pts[3 * j + 1];
    const double dz = (double)pts[3 * i + 2] - (double)pts[3 * j + 2];
    return dx * dx + dy * dy + dz * dz;
}

constexpr int kCandLds = 512;  // candidates per wave sorted in LDS; longer rows take the ordered scan
constexpr int kCandGroupedFromV = 65536;  // calls with this many vertices search with 4 lanes per vertex
// the fill pass of such calls: 4 lanes per vertex up to this mean row length,
// 16 up to the next, else a wave
constexpr double kCandFill4MaxRow = 3.0;
constexpr double kCandFill16MaxRow = 16.0;
constexpr int kLightSortFromV = 65536;    // calls with this many vertices trace their shadow rays in spatial order
constexpr int kLightRows16FromV = 65536;  // calls with this many vertices trace 4 shadow rays per wave

// G lanes per global vertex i (64 / G vertices per wave).  Small calls, where
// few vertices have to fill the GPU, take a wave per vertex; large ones count
// with 4 lanes (their rows are short: a wave per row leaves most lanes idle)
// and fill with 4, 16 or 64 by the mean row length.  FILL = false: count
// candidates into count[i]; FILL = true: write them, ascending, at off[i].
template <bool FILL, int G>
__global__ void __launch_bounds__(256) k_cand(const SolveDesc *__restrict__ sd, const CellDesc *__restrict__ cd,
                                              const int *__restrict__ sid, int V, const float *__restrict__ pts,
                                              double r2, const int *__restrict__ cstart, const int *__restrict__ cend,
                                              const int *__restrict__ svals, const int *__restrict__ off,
                                              int *__restrict__ count, int *__restrict__ fill, ZeroList zero) {
    static_assert(G == 64 || G == 32 || G == 16 || G == 8 || G == 4, "lanes per vertex");
    constexpr int GPW = 64 / G;          // vertices per wave
    constexpr int CAP = kCandLds / GPW;  // LDS slots per vertex
    __shared__ int buf[4][kCandLds];
    grid_zero(zero);  // the next stage's counters (not read here)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int grp = lane / G, gl = lane % G;
    const int i = (blockIdx.x * 4 + wv) * GPW + grp;
    if (__builtin_amdgcn_readfirstlane((blockIdx.x * 4 + wv) * GPW) >= V) return;  // the wave's first vertex
    const bool live = i < V;
    const uint64_t gmask = G == 64 ? ~0ull : (((1ull << G) - 1ull) << (grp * G));
    int rs[9], pre[10];
    pre[0] = 0;
    int total = 0;
    if (live) {
        const CellDesc c = cd[sid[i]];
        const int cx = cell_axis((double)pts[3 * i], c.lo[0], c.inv[0], c.n[0]);
        const int cy = cell_axis((double)pts[3 * i + 1], c.lo[1], c.inv[1], c.n[1]);
        const int cz = cell_axis((double)pts[3 * i + 2], c.lo[2], c.inv[2], c.n[2]);
        const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx + 1 < c.n[0] ? cx + 1 : c.n[0] - 1;
        // the 9 x-runs as position ranges of the sorted points
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            const int y = cy + (r % 3) - 1, z = cz + (r / 3) - 1;
            int s0 = 0x7fffffff, e0 = 0;
            if (y >= 0 && y < c.n[1] && z >= 0 && z < c.n[2]) {
                const int k0 = c.base + (z * c.n[1] + y) * c.n[0];
                for (int x = x0; x <= x1; ++x) {
                    const int e = cend[k0 + x];
                    if (e > 0) {
                        s0 = min(s0, cstart[k0 + x]);
                        e0 = max(e0, e);
                    }
                }
            }
            rs[r] = s0;
            pre[r + 1] = pre[r] + (e0 > 0 ? e0 - s0 : 0);
        }
        total = pre[9];
    }
    // the wave walks as many G-wide steps as its longest group needs
    int tmax = total;
    if (G < 64)
        for (int o = 32; o >= G; o >>= 1) tmax = max(tmax, __shfl_xor(tmax, o, 64));
    int cnt = 0;
    int *gbuf = buf[wv] + grp * CAP;
    for (int f0 = 0; f0 < tmax; f0 += G) {
        const int f = f0 + gl;
        bool hit = false;
        int j = -1;
        if (f < total) {
            int pos = 0;
#pragma unroll
            for (int r = 0; r < 9; ++r)
                if (f >= pre[r] && f < pre[r + 1]) pos = rs[r] + (f - pre[r]);
            j = svals[pos];
            hit = j > i && pdist2(pts, i, j) < r2;
        }
        const uint64_t m = __ballot(hit) & gmask;
        if (FILL && hit) {
            const int slot = cnt + __popcll(m & ((1ull << lane) - 1ull));
            if (slot < CAP) gbuf[slot] = j;
        }
        cnt += __popcll(m);
    }
    if (!FILL) {
        if (live && gl == 0) count[i] = cnt;
        return;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (!live) return;
    const int o = off[i];
    if (cnt <= CAP) {
        // rank of each candidate among the row's (distinct) candidates
        for (int e = gl; e < cnt; e += G) {
            const int val = gbuf[e];
            int rank = 0;
            for (int x = 0; x < cnt; ++x) rank += gbuf[x] < val;
            fill[o + rank] = val;
        }
        return;
    }
    // long row: ordered scan over the whole solve, G points at a time
    const SolveDesc &d = sd[sid[i]];
    const int end = d.voff + d.n;
    int w = 0;
    for (int base = i + 1; base < end; base += G) {
        const int jj = base + gl;
        const bool h = jj < end && pdist2(pts, i, jj) < r2;
        const uint64_t m = __ballot(h) & gmask;
        if (h) fill[o + w + __popcll(m & ((1ull << lane) - 1ull))] = jj;
        w += __popcll(m);
    }
}

__global__ void k_colcount(int V, const int *__restrict__ off, const int *__restrict__ cand, int *__restrict__ colcnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    for (int k = off[i]; k < off[i + 1]; ++k) atomicAdd(&colcnt[cand[k]], 1);
}

// A vertex is risky when its candidate degree (row plus column) exceeds the cap:
// only its rows can lose edges to the greedy sweep.
__device__ __forceinline__ bool vertex_risky(int v, int cap, const int *__restrict__ off,
                                             const int *__restrict__ colcnt) {
    return (off[v + 1] - off[v] + colcnt[v]) > cap;
}

// accepted = 1 for candidates with no risky endpoint; visit = rows the greedy
// sweep has to look at; risky[i] for the sweep.  Uncapped (cap <= 0):
// everything accepted.
__global__ void k_classify(int V, int cap, const int *__restrict__ off, const int *__restrict__ cand,
                           const int *__restrict__ colcnt, uint8_t *__restrict__ risky, int *__restrict__ accepted,
                           int *__restrict__ visit) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const int s = off[i], e = off[i + 1];
    if (cap <= 0) {
        for (int k = s; k < e; ++k) accepted[k] = 1;
        visit[i] = 0;
        return;
    }
    const bool ri = vertex_risky(i, cap, off, colcnt);
    risky[i] = ri ? 1 : 0;
    bool any = false;
    for (int k = s; k < e; ++k) {
        const bool rj = vertex_risky(cand[k], cap, off, colcnt);
        accepted[k] = (!ri && !rj) ? 1 : 0;
        any |= rj;
    }
    visit[i] = (e > s && (ri || any)) ? 1 : 0;
}

__global__ void k_compact_rows(int V, const int *__restrict__ flag, const int *__restrict__ pos, int *__restrict__ list) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    if (flag[i]) list[pos[i]] = i;
}

__device__ __forceinline__ int lower_bound_i32(const int *a, int n, int x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Orders the LDS accesses of the lanes of one wave (a one-wave workgroup needs no
// s_barrier, and __syncthreads would also wait for every outstanding global
// store, a full memory round trip per call).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sequential greedy over contested edges, one wave per solve.
__global__ void __launch_bounds__(64) k_greedy(const SolveDesc *__restrict__ sd, int cap, const int *__restrict__ off,
                                               const int *__restrict__ cand, const uint8_t *__restrict__ risky,
                                               const int *__restrict__ visit_list, const int *__restrict__ n_visit,
                                               int *__restrict__ accepted) {
    extern __shared__ uint8_t cnt[];
    const int s = blockIdx.x;
    const int lane = threadIdx.x;
    const int n = sd[s].n, v0 = sd[s].voff;
    for (int k = lane; k < n; k += 64) cnt[k] = 0;
    __syncthreads();
    const int nv = *n_visit;
    const int lo = lower_bound_i32(visit_list, nv, v0), hi = lower_bound_i32(visit_list, nv, v0 + n);
    for (int idx = lo; idx < hi; ++idx) {
        const int i = visit_list[idx];
        const bool ri = risky[i] != 0;
        int ci = ri ? (int)cnt[i - v0] : 0;
        if (ri && ci >= cap) continue;
        const int e = off[i + 1];
        for (int base = off[i]; base < e; base += 64) {
            const int k = base + lane;
            const bool valid = k < e;
            const int j = valid ? cand[k] : v0;
            const bool rj = valid && risky[j] != 0;
            const bool contested = valid && (ri || rj);
            const bool elig = contested && (!rj || (int)cnt[j - v0] < cap);
            bool sel;
            int taken = 0;
            if (ri) {
                const uint64_t m = __ballot(elig);
                const int need = cap - ci;
                sel = elig && __popcll(m & ((1ull << lane) - 1ull)) < need;
                taken = min((int)__popcll(m), need);
            } else {
                sel = elig;
            }
            if (sel) {
                accepted[k] = 1;
                if (rj) cnt[j - v0] = (uint8_t)(cnt[j - v0] + 1);
            }
            ci += taken;
            wave_sync();
            if (ri && ci >= cap) break;
        }
        if (ri && lane == 0) cnt[i - v0] = (uint8_t)ci;
        wave_sync();
    }
}

// The same sweep with the rows staged in LDS: per batch of up to 64 rows (and
// kGreedyBuf candidates), one wave loads the row headers, then every candidate
// and its risky flag with 32 loads in flight per lane, then runs the sequential
// decisions out of LDS.  Row order and every decision are those of k_greedy.
constexpr int kGreedyBuf = 2048;

__global__ void __launch_bounds__(64) k_greedy_staged(const SolveDesc *__restrict__ sd, int cap,
                                                      const int *__restrict__ off, const int *__restrict__ cand,
                                                      const uint8_t *__restrict__ risky,
                                                      const int *__restrict__ visit_list, const int *__restrict__ n_visit,
                                                      int *__restrict__ accepted) {
    extern __shared__ uint8_t cnt[];
    __shared__ uint32_t cbuf[kGreedyBuf];
    const int s = blockIdx.x;
    const int lane = threadIdx.x;
    const int n = sd[s].n, v0 = sd[s].voff;
    for (int k = lane; k < n; k += 64) cnt[k] = 0;
    const int nv = *n_visit;
    const int lo = lower_bound_i32(visit_list, nv, v0), hi = lower_bound_i32(visit_list, nv, v0 + n);
    __syncthreads();
    int idx = lo;
    while (idx < hi) {
        // row headers
        const bool have = idx + lane < hi;
        int i = 0, k0 = 0, len = 0, rf = 0;
        if (have) {
            i = visit_list[idx + lane];
            k0 = off[i];
            len = off[i + 1] - k0;
            rf = risky[i];
        }
        const int pre = lgd::wave_prefix_sum(len);  // inclusive scan
        const int pos = pre - len;
        const uint64_t fits = __ballot(have && pre <= kGreedyBuf);
        const int rows = __popcll(fits);  // rows form a prefix: pre grows with the lane
        if (rows == 0) {
            // a row longer than the buffer: the direct sweep for that row alone
            const int ii = __shfl(i, 0, 64);
            const bool ri = risky[ii] != 0;
            int ci = ri ? (int)cnt[ii - v0] : 0;
            if (!(ri && ci >= cap)) {
                const int e = off[ii + 1];
                for (int base = off[ii]; base < e; base += 64) {
                    const int k = base + lane;
                    const bool valid = k < e;
                    const int j = valid ? cand[k] : v0;
                    const bool rj = valid && risky[j] != 0;
                    const bool contested = valid && (ri || rj);
                    const bool elig = contested && (!rj || (int)cnt[j - v0] < cap);
                    bool sel;
                    int taken = 0;
                    if (ri) {
                        const uint64_t m = __ballot(elig);
                        const int need = cap - ci;
                        sel = elig && __popcll(m & ((1ull << lane) - 1ull)) < need;
                        taken = min((int)__popcll(m), need);
                    } else {
                        sel = elig;
                    }
                    if (sel) {
                        accepted[k] = 1;
                        if (rj) cnt[j - v0] = (uint8_t)(cnt[j - v0] + 1);
                    }
                    ci += taken;
                    wave_sync();
                    if (ri && ci >= cap) break;
                }
                if (ri && lane == 0) cnt[ii - v0] = (uint8_t)ci;
            }
            wave_sync();
            ++idx;
            continue;
        }
        const int total = __shfl(pre, rows - 1, 64);
        // row of every flat candidate position: mark row starts, then a running
        // max over the positions (rows start in increasing order)
        for (int q = lane; q < total; q += 64) cbuf[q] = 0u;
        wave_sync();
        if (lane < rows) cbuf[pos] = (uint32_t)lane;
        wave_sync();
        // candidates and their risky flags, 32 loads in flight per lane
        int jv[kGreedyBuf / 64];
        int carry = 0;
#pragma unroll
        for (int u = 0; u < kGreedyBuf / 64; ++u) {
            const int q = lane + 64 * u;
            int r = lgd::wave_prefix_max(q < total ? (int)cbuf[q] : 0);
            r = max(r, carry);
            carry = __shfl(r, 63, 64);
            const int kb = __shfl(k0, r, 64), pb = __shfl(pos, r, 64);  // all lanes active
            jv[u] = -1;
            if (q < total) jv[u] = cand[kb + (q - pb)];
        }
        wave_sync();
#pragma unroll
        for (int u = 0; u < kGreedyBuf / 64; ++u) {
            const int q = lane + 64 * u;
            if (q < total) cbuf[q] = (uint32_t)(jv[u] - v0) | ((risky[jv[u]] != 0) ? 0x80000000u : 0u);
        }
        wave_sync();
        // row headers stay in the lanes that loaded them (readlane); the first
        // candidate chunk of the next row is fetched while this row decides
        const int mr = i - v0;
        uint32_t wnext = lane < __builtin_amdgcn_readlane(len, 0) ? cbuf[__builtin_amdgcn_readlane(pos, 0) + lane] : 0u;
        for (int r = 0; r < rows; ++r) {
            const int il = __builtin_amdgcn_readlane(mr, r), p0 = __builtin_amdgcn_readlane(pos, r);
            const int ln = __builtin_amdgcn_readlane(len, r), kk0 = __builtin_amdgcn_readlane(k0, r);
            const bool ri = __builtin_amdgcn_readlane(rf, r) != 0;
            const uint32_t w0 = wnext;
            if (r + 1 < rows) {
                const int p1 = __builtin_amdgcn_readlane(pos, r + 1), l1 = __builtin_amdgcn_readlane(len, r + 1);
                wnext = lane < l1 ? cbuf[p1 + lane] : 0u;
            }
            int ci = ri ? (int)cnt[il] : 0;
            if (ri && ci >= cap) continue;
            for (int t0 = 0; t0 < ln; t0 += 64) {
                const int t = t0 + lane;
                const bool valid = t < ln;
                const uint32_t w = t0 == 0 ? w0 : (valid ? cbuf[p0 + t] : 0u);
                const int jl = (int)(w & 0x7fffffffu);
                const bool rj = valid && (w >> 31) != 0;
                const bool contested = valid && (ri || rj);
                const int cj = rj ? (int)cnt[jl] : 0;
                const bool elig = contested && (!rj || cj < cap);
                bool sel;
                int taken = 0;
                if (ri) {
                    const uint64_t m = __ballot(elig);
                    const int need = cap - ci;
                    sel = elig && __popcll(m & ((1ull << lane) - 1ull)) < need;
                    taken = min((int)__popcll(m), need);
                } else {
                    sel = elig;
                }
                if (sel) {
                    accepted[kk0 + t] = 1;
                    if (rj) cnt[jl] = (uint8_t)(cj + 1);
                }
                ci += taken;
                if (ri && ci >= cap) break;
            }
            if (ri && lane == 0) cnt[il] = (uint8_t)ci;
            wave_sync();
        }
        idx += rows;
    }
}

// ------------------------------------------- greedy cap, rows in parallel ----
// The greedy sweep's decisions for row i read and change only the degree
// counters of its capped ("risky") vertices: i itself and its risky candidates
// j.  Row i therefore only has to follow the earlier rows that touch one of
// those vertices; rows touching disjoint vertex sets may be decided at once.
// For a candidate slot k = (a, j) with j risky, rank[k] = the number of rows
// a' < a that also hold j (slots are row-major, so the slot order of a column
// is its row order); the last toucher of a risky vertex i is row i itself,
// after colcnt[i] earlier rows.
__global__ void k_col_fill(int K, const int *__restrict__ cand, const uint8_t *__restrict__ risky,
                           const int *__restrict__ coloff, int *__restrict__ fill, int *__restrict__ colslot) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int j = cand[k];
    if (risky[j]) colslot[coloff[j] + atomicAdd(&fill[j], 1)] = k;
}

// slot[k] = (j, rank[k] if j is risky else -1): what the greedy needs of a
// candidate, in one 8-byte load.  One thread per slot.
__global__ void k_col_rank(int K, const int *__restrict__ cand, const uint8_t *__restrict__ risky,
                           const int *__restrict__ coloff, const int *__restrict__ colcnt,
                           const int *__restrict__ colslot, int2 *__restrict__ slot) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int j = cand[k];
    int r = -1;
    if (risky[j]) {
        const int c0 = coloff[j], c1 = c0 + colcnt[j];
        r = 0;
        for (int c = c0; c < c1; ++c) r += colslot[c] < k ? 1 : 0;
    }
    slot[k] = make_int2(j, r);
}

// One workgroup of kGreedyWaves waves per solve; wave w decides rows w, w + W,
// ... of the solve's contested rows in order.  Each vertex has one LDS word:
// rows that have finished with it (<< 8) | its accepted degree.  A row waits
// (per 64-candidate chunk) until each risky candidate's word shows rank[k]
// finished rows, and, for a risky row vertex, colcnt[i]; it then decides
// exactly as k_greedy does and publishes each risky candidate's word (finished
// + 1, new degree) with a release store, accepted or not.  The earliest
// unfinished row never waits on an unfinished row, so the sweep always advances;
// a wait longer than kGreedySpin polls sets *err (the host then fails the
// solve) instead of spinning on.
constexpr int kGreedyWaves = 16;
constexpr int kGreedySpin = 1 << 22;

__device__ __forceinline__ uint32_t lds_acquire(uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_release(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Poll until every lane's condition holds.  Bounded: the first wave of the
// workgroup that polls kGreedySpin times sets *abort, and every wave then stops
// waiting at once (false: the decisions are void and the host fails the solve).
template <typename F>
__device__ __forceinline__ bool wave_wait(F &&ready, uint32_t *abort) {
    for (int it = 0; it < kGreedySpin; ++it) {
        if (__ballot(!ready()) == 0ull) return true;
        if (lds_acquire(abort) != 0u) return false;
        __builtin_amdgcn_s_sleep(1);
    }
    lds_release(abort, 1u);
    return false;
}

// A contested row as a wave holds it: the header in every lane, the first 64
// candidate slots one per lane (longer rows load the rest when they get there).
struct GreedyRow {
    int i, k0, k1, self_want;
    bool ri;
    int2 sl;
};

__device__ __forceinline__ GreedyRow greedy_row(int r, int hi, int lane, const int *__restrict__ visit_list,
                                                const int *__restrict__ off, const uint8_t *__restrict__ risky,
                                                const int *__restrict__ colcnt, const int2 *__restrict__ slot) {
    GreedyRow g{};
    g.k0 = g.k1 = 0;
    if (r >= hi) return g;
    g.i = visit_list[r];
    g.k0 = off[g.i];
    g.k1 = off[g.i + 1];
    g.ri = risky[g.i] != 0;
    g.self_want = colcnt[g.i];
    g.sl = g.k0 + lane < g.k1 ? slot[g.k0 + lane] : make_int2(0, -1);
    return g;
}

__global__ void __launch_bounds__(64 * kGreedyWaves) k_greedy_waves(
    const SolveDesc *__restrict__ sd, int cap, const int *__restrict__ off, const int2 *__restrict__ slot,
    const uint8_t *__restrict__ risky, const int *__restrict__ colcnt, const int *__restrict__ visit_list,
    const int *__restrict__ n_visit, int *__restrict__ accepted, int *__restrict__ err) {
    extern __shared__ uint32_t word[];
    __shared__ uint32_t abort;
    const int s = blockIdx.x;
    const int w = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    const int n = sd[s].n, v0 = sd[s].voff;
    for (int k = threadIdx.x; k < n; k += blockDim.x) word[k] = 0u;
    if (threadIdx.x == 0) abort = 0u;
    __syncthreads();
    const int nv = *n_visit;
    const int lo = lower_bound_i32(visit_list, nv, v0), hi = lower_bound_i32(visit_list, nv, v0 + n);
    bool ok = true;
    GreedyRow nxt = greedy_row(lo + w, hi, lane, visit_list, off, risky, colcnt, slot);
    for (int r = lo + w; r < hi; r += kGreedyWaves) {
        const GreedyRow row = nxt;
        // the wave's next row loads while this one waits and decides
        nxt = greedy_row(r + kGreedyWaves, hi, lane, visit_list, off, risky, colcnt, slot);
        const int il = row.i - v0;
        int ci = 0;
        if (row.ri) {
            const uint32_t want = (uint32_t)row.self_want;
            ok = ok && wave_wait([&] { return lane != 0 || (lds_acquire(&word[il]) >> 8) == want; }, &abort);
            ci = (int)(lds_acquire(&word[il]) & 0xffu);
        }
        for (int base = row.k0; base < row.k1; base += 64) {
            const int k = base + lane;
            const bool valid = k < row.k1;
            const int2 sl = base == row.k0 ? row.sl : (valid ? slot[k] : make_int2(0, -1));
            const bool rj = valid && sl.y >= 0;
            const int jl = sl.x - v0;
            const uint32_t want = rj ? (uint32_t)sl.y : 0u;
            ok = ok && wave_wait([&] { return !rj || (lds_acquire(&word[jl]) >> 8) == want; }, &abort);
            const uint32_t wj = rj ? lds_acquire(&word[jl]) : 0u;
            const int cj = (int)(wj & 0xffu);
            const bool contested = valid && (row.ri || rj);
            const bool open = !(row.ri && ci >= cap);
            const bool elig = open && contested && (!rj || cj < cap);
            bool sel;
            int taken = 0;
            if (row.ri) {
                const uint64_t m = __ballot(elig);
                const int need = cap - ci;
                sel = elig && __popcll(m & ((1ull << lane) - 1ull)) < need;
                taken = min((int)__popcll(m), need);
            } else {
                sel = elig;
            }
            if (sel) accepted[k] = 1;
            if (rj) lds_release(&word[jl], ((want + 1u) << 8) | (uint32_t)(sel ? cj + 1 : cj));
            if (open) ci += taken;
        }
        if (row.ri && lane == 0) lds_release(&word[il], (lds_acquire(&word[il]) & ~0xffu) | (uint32_t)ci);
    }
    if (!ok && lane == 0) atomicOr(err, 1);
}

// Edge list (i, j) in row-major order from the accepted candidates.
__global__ void k_edges(int V, const int *__restrict__ off, const int *__restrict__ cand,
                        const int *__restrict__ accepted, const int *__restrict__ epos, int *__restrict__ edges,
                        ZeroList zero) {
    grid_zero(zero);  // the component stage's counters (not read here)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    for (int k = off[i]; k < off[i + 1]; ++k)
        if (accepted[k]) {
            const int e = epos[k];
            edges[2 * e] = i;
            edges[2 * e + 1] = cand[k];
        }
}

// ---------------------------------------------------------- edge march ------
template <typename T, bool kCount>
__global__ void k_march(const SolveDesc *__restrict__ sd, const int *__restrict__ sid, MarchConsts mc,
                        const float *__restrict__ pts, const int *__restrict__ edges, int edge_limit,
                        const int *__restrict__ n_edges, float *__restrict__ w, unsigned long long *__restrict__ stats) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    int ns = 0;
    if (e < min(*n_edges, edge_limit)) {
        const int i = edges[2 * e], j = edges[2 * e + 1];
        w[e] = march_edge_one<T, kCount>(sd[sid[i]].g, mc, pts + 3 * (long long)i, pts + 3 * (long long)j, &ns);
    }
    if (stats) {  // one atomic per wave
        unsigned long long v = (unsigned long long)ns;
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if ((threadIdx.x & 63) == 0 && v) stat_add(stats, ST_EDGE_SAMPLES, v);
    }
}

// ------------------------------------------------ adjacency and components --
// Lock-free union-find over the finite edges of every solve of the call, one
// thread per edge: hook the larger root under the smaller with a CAS that only
// succeeds while it is still a root, so one pass joins every edge.  The final root
// of a component is its smallest vertex, whatever the interleaving.
__device__ __forceinline__ int uf_find(int *P, int x) {
    for (;;) {
        const int p = atomic_load_i32(&P[x]);
        if (p == x) return x;
        const int g = atomic_load_i32(&P[p]);
        if (g == p) return p;
        atomic_store_i32(&P[x], g);  // path halving; pointers only decrease
        x = g;
    }
}

// Joins the trees of a and b: the larger root is linked below the smaller.
__device__ __forceinline__ void uf_link(int *__restrict__ P, int a, int b) {
    for (;;) {
        a = uf_find(P, a);
        b = uf_find(P, b);
        if (a == b) return;
        if (a < b) {
            const int t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(&P[a], a, b) == a) return;
    }
}

// Degrees of the finite edges; thread v < V also starts the union-find forest
// (parent[v] = v) that k_adj_fill_union links.
__global__ void k_adj_deg(const int *__restrict__ edges, const float *__restrict__ w, const int *__restrict__ n_edges,
                          int *__restrict__ deg, int V, int *__restrict__ P) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < V) atomic_store_i32(&P[e], e);
    if (e >= *n_edges || !(w[e] < INFINITY)) return;
    atomicAdd(&deg[edges[2 * e]], 1);
    atomicAdd(&deg[edges[2 * e + 1]], 1);
}

// Adjacency lists of the finite edges, and each such edge's union in the
// union-find forest (parents started by k_adj_deg).
__global__ void k_adj_fill_union(const int *__restrict__ edges, const float *__restrict__ w,
                                 const int *__restrict__ n_edges, const int *__restrict__ aoff, int *__restrict__ cursor,
                                 int *__restrict__ nb, float *__restrict__ nw, int *__restrict__ P) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= *n_edges || !(w[e] < INFINITY)) return;
    const int i = edges[2 * e], j = edges[2 * e + 1];
    const float we = w[e];
    int p = aoff[i] + atomicAdd(&cursor[i], 1);
    nb[p] = j;
    nw[p] = we;
    p = aoff[j] + atomicAdd(&cursor[j], 1);
    nb[p] = i;
    nw[p] = we;
    uf_link(P, i, j);
}

// Each vertex's root (its component's smallest vertex) and its sort key
// (root, vertex).
__global__ void k_uf_root_keys(int V, int bits, int *__restrict__ P, int *__restrict__ root,
                               uint64_t *__restrict__ keys) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int r = uf_find(P, v);
    root[v] = r;
    keys[v] = ((uint64_t)r << bits) | (uint64_t)v;
}

__global__ void k_comp_flags(int V, int bits, const uint64_t *__restrict__ sorted, int *__restrict__ flag,
                             int *__restrict__ member) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= V) return;
    const uint64_t mask = (1ull << bits) - 1ull;
    member[k] = (int)(sorted[k] & mask);
    flag[k] = (k == 0 || (sorted[k] >> bits) != (sorted[k - 1] >> bits)) ? 1 : 0;
}

__global__ void k_comp_offsets(int V, const int *__restrict__ flag, const int *__restrict__ cid_excl,
                               int *__restrict__ comp_off, const int *__restrict__ n_comp) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) comp_off[*n_comp] = V;
    if (k >= V) return;
    if (flag[k]) comp_off[cid_excl[k]] = k;
}

// local index of every vertex inside its component, the (component, size) pairs
// of the components too large for the LDS kernels, and the lists of the other
// components by size class (2..32, 33..64, 65..128), and the isolated vertices,
// listed from the end of the 2..32 list (cls16[V - 1 - slot]: that list holds at
// most (V - n1) / 2 components, so the two never meet).  Reads no vertex light,
// so it need not wait for it.
__global__ void k_comp_local(int V, const SolveDesc *__restrict__ sd, const int *__restrict__ sid,
                             const int *__restrict__ flag, const int *__restrict__ cid_excl,
                             const int *__restrict__ comp_off, const int *__restrict__ member, int *__restrict__ local,
                             int *__restrict__ big_list, int *__restrict__ counters, int *__restrict__ cls16,
                             int *__restrict__ cls48, int *__restrict__ cls128, const int *__restrict__ n_comp,
                             const int *__restrict__ n_edges, const int *__restrict__ greedy_err) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) {  // for the host's one readback
        counters[8] = *n_comp;
        counters[9] = *n_edges;
        counters[10] = *greedy_err;
    }
    if (k >= V) return;
    const int c = cid_excl[k] + flag[k] - 1;
    local[member[k]] = k - comp_off[c];
    if (!flag[k]) return;
    const int m = comp_off[c + 1] - comp_off[c];
    if (m > kSmallComp) {
        const int slot = atomicAdd(&counters[0], 1);
        big_list[2 * slot] = c;
        big_list[2 * slot + 1] = m;
    } else if (m > 64) {
        cls128[atomicAdd(&counters[3], 1)] = c;
    } else if (m > 32) {
        cls48[atomicAdd(&counters[2], 1)] = c;
    } else if (m > 1) {
        cls16[atomicAdd(&counters[1], 1)] = c;
    } else {
        const int v = member[k];
        cls16[V - 1 - atomicAdd(&counters[4], 1)] = v;
        const SolveDesc &d = sd[sid[v]];
        const int li = v - d.voff;
        if (d.dense_D) d.dense_D[(long long)li * d.ld + li] = 0.0f;
        if (d.dense_pred) d.dense_pred[(long long)li * d.ld + li] = li;
    }
}

// ------------------------------------------------ breadth-first order ------
// Optional vertex order of the large components (LG_SOLVE_BFS_ORDER): members
// sorted by (breadth-first level from the component's smallest vertex, vertex
// index) instead of by vertex index.  The blocked closure of a geometric graph
// then meets few finite panel tiles ahead of the pivot row: the vertices of
// rows 0..k form a connected ball and only its rim reaches the later rows, so
// most remainder tiles hold +INF operands and are skipped.  Distances are the
// same shortest paths; the pivot order, and so the FP32 rounding of their
// sums, follows the new order.

// One workgroup per large component (slot b of big_list, counters[0] of them):
// level-synchronous breadth-first search over the finite-edge adjacency with
// frontier queues (qa, qb: V ints each; the component uses its member range
// [a0, a0 + m) of both).  A vertex is claimed by the CAS that moves it from
// unseen to L + 1, so it enters the next frontier once; the queue order varies
// between runs, the levels do not.  Every member is reachable from the first
// (the component's smallest vertex), so each frontier holds at most m vertices.
__global__ void __launch_bounds__(1024) k_bfs_levels(const int *__restrict__ counters,
                                                     const int *__restrict__ big_list,
                                                     const int *__restrict__ comp_off, const int *__restrict__ member,
                                                     const int *__restrict__ aoff, const int *__restrict__ anb,
                                                     int *__restrict__ level, int *__restrict__ qa,
                                                     int *__restrict__ qb) {
    const int b = (int)blockIdx.x;
    if (b >= counters[0]) return;
    const int c = big_list[2 * b], m = big_list[2 * b + 1];
    const int a0 = comp_off[c];
    constexpr int kUnseen = 0x7fffffff;
    __shared__ int n_next;
    for (int i = threadIdx.x; i < m; i += blockDim.x) level[member[a0 + i]] = (i == 0) ? 0 : kUnseen;
    __threadfence();  // the CAS below runs at device scope
    int *cur = qa + a0, *nxt = qb + a0;
    if (threadIdx.x == 0) {
        cur[0] = member[a0];
        n_next = 0;
    }
    __syncthreads();
    int n_cur = 1;
    for (int L = 0; n_cur > 0 && L < m; ++L) {
        for (int t = threadIdx.x; t < n_cur; t += blockDim.x) {
            const int v = cur[t];
            const int p1 = aoff[v + 1];
            for (int p = aoff[v]; p < p1; ++p) {
                const int u = anb[p];
                // most neighbours were reached earlier: read before claiming
                if (__hip_atomic_load(&level[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != kUnseen) continue;
                if (atomicCAS(&level[u], kUnseen, L + 1) == kUnseen) {
                    const int slot = atomicAdd(&n_next, 1);
                    if (slot < m) nxt[slot] = u;
                }
            }
        }
        __syncthreads();
        n_cur = n_next < m ? n_next : m;
        __syncthreads();
        if (threadIdx.x == 0) n_next = 0;
        int *t = cur;
        cur = nxt;
        nxt = t;
        __syncthreads();
    }
}

// Nested-dissection vertex order of the large components (LG_SOLVE_ND_ORDER):
// one workgroup per component splits its member list recursively.  A segment
// is cut at the midpoint of its points' longest bounding-box side; its left
// vertices with a neighbour on the right (in the same segment) form the
// separator; the segment is stably partitioned into [left - separator | right |
// separator] and the first two parts are split again down to kNdLeaf vertices.
// Separators come after both halves they separate, so while the blocked closure
// pivots through one half, the other half's rows stay +INF in the panel and
// their tiles are skipped.  Starting from ascending vertex order, every part
// stays ascending: the order is a function of the points and the graph only.
constexpr int kNdLeaf = 256;
constexpr int kNdStack = 64;

__device__ __forceinline__ float block_reduce_min(float v, float *red) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    float r = red[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) r = fminf(r, red[i]);
    return r;
}

__device__ __forceinline__ int block_excl_scan(int x, int *red, int *total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int inc = lgd::wave_prefix_sum(x);
    __syncthreads();
    if (lane == 63) red[w] = inc;
    __syncthreads();
    int before = 0, all = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
        if (i < w) before += red[i];
        all += red[i];
    }
    *total = all;
    return before + inc - x;
}

__global__ void __launch_bounds__(1024) k_nd_order(const int *__restrict__ counters, const int *__restrict__ big_list,
                                                   const int *__restrict__ comp_off, int *__restrict__ member,
                                                   int *__restrict__ local, const float *__restrict__ pts,
                                                   const int *__restrict__ aoff, const int *__restrict__ anb,
                                                   int *__restrict__ segtag, int *__restrict__ cls,
                                                   int *__restrict__ tmp) {
    const int b = (int)blockIdx.x;
    if (b >= counters[0]) return;
    const int c = big_list[2 * b], m = big_list[2 * b + 1];
    const int a0 = comp_off[c];
    int *perm = member + a0;
    int *scratch = tmp + a0;
    __shared__ float redf[16];
    __shared__ int redi[16];
    __shared__ int st_lo[kNdStack], st_hi[kNdStack];
    __shared__ int sp;
    const int tid = threadIdx.x, nt = blockDim.x;
    if (tid == 0) {
        st_lo[0] = 0;
        st_hi[0] = m;
        sp = 1;
    }
    __syncthreads();
    int segid = 0;
    while (true) {
        __syncthreads();
        if (sp == 0) break;
        const int lo = st_lo[sp - 1], hi = st_hi[sp - 1];
        __syncthreads();
        if (tid == 0) --sp;
        const int n = hi - lo;
        if (n <= kNdLeaf) continue;
        ++segid;
        // tag the segment's vertices, bounding box of their points
        float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int i = lo + tid; i < hi; i += nt) {
            const int v = perm[i];
            segtag[v] = segid;
            for (int a = 0; a < 3; ++a) {
                mn[a] = fminf(mn[a], pts[3 * v + a]);
                mx[a] = fmaxf(mx[a], pts[3 * v + a]);
            }
        }
        float bmn[3], bmx[3];
        for (int a = 0; a < 3; ++a) {
            bmn[a] = block_reduce_min(mn[a], redf);
            bmx[a] = -block_reduce_min(-mx[a], redf);
        }
        int ax = 0;
        for (int a = 1; a < 3; ++a)
            if (bmx[a] - bmn[a] > bmx[ax] - bmn[ax]) ax = a;
        const float mid = 0.5f * (bmn[ax] + bmx[ax]);
        // side of every vertex (0 left, 1 right), then the separator (2)
        int nright = 0;
        for (int i = lo + tid; i < hi; i += nt) {
            const int v = perm[i];
            const int sd = pts[3 * v + ax] < mid ? 0 : 1;
            cls[v] = sd;
            nright += sd;
        }
        __threadfence_block();
        int right_all = 0;
        (void)block_excl_scan(nright, redi, &right_all);
        if (right_all == 0 || right_all == n) continue;  // all on one side: a leaf
        __syncthreads();
        // separator: left vertices with a right neighbour in this segment get
        // class 2 (a left vertex turning 0 -> 2 leaves every "is right" test alone)
        for (int i = lo + tid; i < hi; i += nt) {
            const int v = perm[i];
            if (cls[v] != 0) continue;
            bool sep = false;
            for (int p = aoff[v]; p < aoff[v + 1] && !sep; ++p) {
                const int u = anb[p];
                sep = segtag[u] == segid && cls[u] == 1;
            }
            if (sep) cls[v] = 2;
        }
        __threadfence_block();
        __syncthreads();
        // stable partition [left - separator | right | separator] through scratch
        int counts[3] = {0, 0, 0};
        for (int i = lo + tid; i < hi; i += nt) ++counts[cls[perm[i]]];
        int tot[3];
        for (int q = 0; q < 3; ++q) (void)block_excl_scan(counts[q], redi, &tot[q]);
        int carry[3] = {0, 0, 0};
        const int start[3] = {lo, lo + tot[0], lo + tot[0] + tot[1]};
        for (int base = lo; base < hi; base += nt) {
            const int i = base + tid;
            const int q = i < hi ? cls[perm[i]] : -1;
            for (int qq = 0; qq < 3; ++qq) {
                int t = 0;
                const int off = block_excl_scan(q == qq ? 1 : 0, redi, &t);
                if (q == qq) scratch[start[qq] - lo + carry[qq] + off] = perm[i];
                carry[qq] += t;
            }
        }
        __syncthreads();
        for (int i = lo + tid; i < hi; i += nt) perm[i] = scratch[i - lo];
        __syncthreads();
        if (tid == 0) {
            if (tot[1] > kNdLeaf && sp < kNdStack) {
                st_lo[sp] = lo + tot[0];
                st_hi[sp] = lo + tot[0] + tot[1];
                ++sp;
            }
            if (tot[0] > kNdLeaf && sp < kNdStack) {
                st_lo[sp] = lo;
                st_hi[sp] = lo + tot[0];
                ++sp;
            }
        }
    }
    for (int i = tid; i < m; i += nt) local[perm[i]] = i;
}

// Sort keys (root, level, vertex); vertices of the other components keep level 0.
__global__ void k_keys_bfs(int V, int bits, const int *__restrict__ root, const int *__restrict__ level,
                           uint64_t *__restrict__ keys) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const uint64_t lmax = (1ull << bits) - 1ull;
    const uint64_t l = (uint64_t)(unsigned)level[v] < lmax ? (uint64_t)(unsigned)level[v] : lmax;
    keys[v] = ((((uint64_t)root[v] << bits) | l) << bits) | (uint64_t)v;
}

// member / local from the re-sorted keys; the components keep their ranges.
__global__ void k_member_local(int V, int bits, const uint64_t *__restrict__ sorted, const int *__restrict__ flag,
                               const int *__restrict__ cid_excl, const int *__restrict__ comp_off,
                               int *__restrict__ member, int *__restrict__ local) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= V) return;
    const int v = (int)(sorted[k] & ((1ull << bits) - 1ull));
    member[k] = v;
    local[v] = k - comp_off[cid_excl[k] + flag[k] - 1];
}

// ------------------------------------------------ small-component closure ----
// Sum of x over each group of CG consecutive lanes (a row of the closure's
// register tile), valid at the group's last lane.  Groups of 16 and 32 lanes
// are DPP rows: prefix sums within the rows (row_shr 1, 2, 4, 8), then for 32
// row 0's total added into row 1 (row_bcast 15), in a fixed order and without
// the LDS round trips of lane shuffles; smaller groups use shuffles.
template <int CG>
__device__ __forceinline__ double row_group_sum(double x) {
    if constexpr (CG == 16 || CG == 32) {
        x += lgd::dpp_move<0x111, 0xf>(x);
        x += lgd::dpp_move<0x112, 0xf>(x);
        x += lgd::dpp_move<0x114, 0xf>(x);
        x += lgd::dpp_move<0x118, 0xf>(x);
        if constexpr (CG == 32) x += lgd::dpp_move<0x142, 0xa>(x);
        return x;
    } else {
#pragma unroll
        for (int o = CG / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, CG);
        return x;
    }
}

// One workgroup per component of m <= 128 vertices: build the m x m optical
// depth matrix in LDS from the adjacency, close it pivot by pivot (ascending
// vertex order), then diffuse (LightGraph.h:554-579) and write the outputs.
// One workgroup of (MS/4)^2 threads per component of at most MS vertices.  The
// component's optical-depth matrix is built in LDS from the adjacency (padding
// vertices are isolated: +INF rows, zero diagonal), then closed in registers
// (4 x 4 per thread) in pivot groups like the pivot tile (fw_pivot_tile,
// lg_fw_group.h): pivots in ascending vertex order, bit-identical to the
// one-pivot loop.  Diffusion (LightGraph.h:554-587) follows from the registers:
// column sums = row sums (D is symmetric), each reduced over the row's lanes in
// a fixed order.
template <int MS>
__global__ void __launch_bounds__((MS / 4) * (MS / 4)) k_comp_small(
    const SolveDesc *__restrict__ sd, const int *__restrict__ sid, const int *__restrict__ cls,
    const int *__restrict__ comp_off, const int *__restrict__ member, const int *__restrict__ local,
    const int *__restrict__ aoff, const int *__restrict__ anb, const float *__restrict__ aw,
    const float *__restrict__ L, float e0, float e1, float e2, float *__restrict__ out,
    unsigned long long *__restrict__ stats, int ncls, const int *__restrict__ single_end, int nsingle) {
    constexpr int CG = MS / 4;  // column groups = lanes sharing a row group
    constexpr int NT = CG * CG;
    constexpr int KG = kPivotGroup;
    __shared__ float T[MS][MS + 1];
    __shared__ float P[KG][MS];
    __shared__ float G[2][KG][MS];  // alternating: see fw_pivot_tile
    __shared__ int mem[MS];
    __shared__ float Ln[MS][3];
    if ((int)blockIdx.x >= ncls) {
        // workgroups past the list: isolated vertices, whose P = [1] makes the
        // diffused radiance the vertex light itself (single_end[-1 - i])
        const int idx = ((int)blockIdx.x - ncls) * NT + (int)threadIdx.x;
        if (idx < nsingle) {
            const int v = single_end[-1 - idx];
            for (int q = 0; q < 3; ++q) out[3 * v + q] = L[3 * v + q];
        }
        return;
    }
    const int c = cls[blockIdx.x];
    const int a0 = comp_off[c];
    const int m = comp_off[c + 1] - a0;
    if (m > MS) return;
    const int tid = threadIdx.x;
    const int rg = tid / CG, cg = tid % CG;
    if (stats && tid == 0) {
        stat_add(stats, ST_SMALL_UPDATES, (unsigned long long)m * m * m);
        if (MS <= 64) stat_add(stats, MS <= 32 ? ST_SMALL_UPDATES_32 : ST_SMALL_UPDATES_64, (unsigned long long)m * m * m);
    }
    for (int a = tid; a < m; a += NT) {
        const int vm = member[a0 + a];
        mem[a] = vm;
        // the member's light, read early (the diffusion's first pass divides it)
        for (int q = 0; q < 3; ++q) Ln[a][q] = L[3 * vm + q];
    }
    for (int e = tid; e < MS * MS; e += NT) {
        const int i = e / MS, j = e - i * MS;
        T[i][j] = (i == j) ? 0.0f : INFINITY;
    }
    __syncthreads();
    {
        // NT / MS lanes per row walk its adjacency together: each lane's loads
        // (neighbour, its local index, weight) depend on one another, and a
        // row's edges are distinct, so the lanes write distinct entries
        constexpr int LPR = NT / MS;
        const int a = tid / LPR, l = tid % LPR;
        if (a < m) {
            const int v = mem[a];
            for (int p = aoff[v] + l; p < aoff[v + 1]; p += LPR) T[a][local[anb[p]]] = aw[p];
        }
    }
    __syncthreads();
    float v[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) v[a][b] = T[4 * rg + a][4 * cg + b];
    const int groups = (m + KG - 1) / KG;
    for (int g = 0; g < groups; ++g) {
        const int k = KG * g;
        if constexpr (MS == 128) {
            // the owners of rows k..k+KG-1 are wave g: it publishes them and
            // derives the group's pivot rows without a workgroup barrier
            if ((tid >> 6) == g) {
                const int r0 = 4 * (rg % (KG / 4));
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) P[r0 + a][4 * cg + b] = v[a][b];
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                pivot_group_rows_wave128<KG>(P, G[g & 1], k, tid & 63);
            }
            __syncthreads();
        } else {
            if (rg / (KG / 4) == g) {  // the owners of rows k..k+KG-1 publish them
                const int r0 = 4 * (rg % (KG / 4));
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) P[r0 + a][4 * cg + b] = v[a][b];
            }
            __syncthreads();
            if (tid < MS) pivot_group_rows<KG, MS>(P, G[g & 1], k, tid);
            __syncthreads();
        }
        pivot_group_update<KG, MS>(G[g & 1], rg, cg, v);
    }
    // diffusion, one channel at a time: the row's 16 transmittances stay in
    // registers between the two reductions over the row's lanes (column sums =
    // row sums, D being symmetric), each in a fixed order
    const float ex[3] = {e0, e1, e2};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        float tq[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) tq[a][b] = (4 * cg + b < m) ? trans_f(ex[q], v[a][b]) : 0.0f;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            double sq = 0.0;
#pragma unroll
            for (int b = 0; b < 4; ++b) sq += (double)tq[a][b];
            sq = row_group_sum<CG>(sq);
            const int i = 4 * rg + a;
            if (cg == CG - 1 && i < m) Ln[i][q] = (float)((double)Ln[i][q] / (sq < 1.0 ? 1.0 : sq));
        }
        __syncthreads();
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            double oq = 0.0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int j = 4 * cg + b;
                if (j < m) oq += (double)Ln[j][q] * (double)tq[a][b];
            }
            oq = row_group_sum<CG>(oq);
            const int i = 4 * rg + a;
            if (cg == CG - 1 && i < m) out[3 * mem[i] + q] = (float)oq;
        }
    }
    const SolveDesc &d = sd[sid[mem[0]]];
    if (!d.dense_D && !d.dense_pred) return;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) T[4 * rg + a][4 * cg + b] = v[a][b];
    __syncthreads();
    if (d.dense_D) {
        for (int e = tid; e < m * m; e += NT) {
            const int i = e / m, j = e - i * m;
            d.dense_D[(long long)(mem[i] - d.voff) * d.ld + (mem[j] - d.voff)] = T[i][j];
        }
    }
    if (d.dense_pred) {
        // canonical predecessor: smallest neighbour k of j minimising D[i][k] + w(k, j)
        for (int e = tid; e < m * m; e += NT) {
            const int i = e / m, j = e - i * m;
            int best_k = -1;
            if (i == j) {
                best_k = mem[i] - d.voff;
            } else if (T[i][j] < INFINITY) {
                float best = INFINITY;
                const int vj = mem[j];
                for (int p = aoff[vj]; p < aoff[vj + 1]; ++p) {
                    const int kv = anb[p];
                    const float val = T[i][local[kv]] + aw[p];
                    const int kl = kv - d.voff;
                    if (val < best || (val == best && kl < best_k)) {
                        best = val;
                        best_k = kl;
                    }
                }
            }
            d.dense_pred[(long long)(mem[i] - d.voff) * d.ld + (mem[j] - d.voff)] = best_k;
        }
    }
}

// ------------------------------------------------ large-component path ------
// Components above kSmallComp vertices are gathered into per-component matrices
// (one arena, ldm = 128-padded size) and closed together by the batched blocked
// Floyd-Warshall.  Every kernel below covers all large components of the call in
// one launch: block b handles row b of the concatenated row space and finds its
// component by binary search over the row prefix.
struct BigJob {
    long long toff;  // first element of the component's matrix in the arena
    int a0;          // first member in the sorted member list
    int m, ldm;
    int row0;        // first row of the component in the concatenated m-row space
    int pad0, pad1;
};

__device__ __forceinline__ int big_find(const int *__restrict__ prefix, int J, int b) {
    int lo = 0, hi = J - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// rows of ldm: +INF off the diagonal, 0 on it, then the component's edges.  The
// component's first member a0 = comp_off[comp[q]] is gathered here, and row 0
// records it in the job table for the diffusion kernels.
__global__ void __launch_bounds__(256) k_big_fill(BigJob *__restrict__ jobs, const int *__restrict__ ld_prefix,
                                                  int J, const int *__restrict__ comp, const int *__restrict__ comp_off,
                                                  const int *__restrict__ member, const int *__restrict__ local,
                                                  const int *__restrict__ aoff, const int *__restrict__ anb,
                                                  const float *__restrict__ aw, float *__restrict__ T) {
    const int q = big_find(ld_prefix, J, (int)blockIdx.x);
    BigJob jb = jobs[q];
    jb.a0 = comp_off[comp[q]];
    const int r = (int)blockIdx.x - ld_prefix[q];
    if (r == 0 && threadIdx.x == 0) jobs[q].a0 = jb.a0;
    float *row = T + jb.toff + (long long)r * jb.ldm;
    for (int j = threadIdx.x; j < jb.ldm; j += blockDim.x) row[j] = (j == r) ? 0.0f : INFINITY;
    __syncthreads();
    if (r >= jb.m) return;
    const int v = member[jb.a0 + r];
    for (int p = aoff[v] + threadIdx.x; p < aoff[v + 1]; p += blockDim.x) row[local[anb[p]]] = aw[p];
}

// Sums of three values over a 256-thread workgroup, at thread 0: each wave
// reduces by DPP prefix sums (its total at lane 63), then thread 0 adds the
// four waves' totals in order.  One barrier, a fixed order.
__device__ __forceinline__ void block256_sum3(double v[3], double (*red)[4]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double x = lgd::wave_prefix_sum(v[c]);
        if ((threadIdx.x & 63) == 63) red[c][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] = ((red[c][0] + red[c][1]) + red[c][2]) + red[c][3];
}

// per-row sums of exp(ext_c * D) over the component (D symmetric: = column
// sums), and from them the vertex's normalised light L_i / max(colsum_i, 1),
// the factor that every row of the diffusion multiplies column i by.
__global__ void __launch_bounds__(256) k_big_rowsum(const BigJob *__restrict__ jobs, const int *__restrict__ m_prefix,
                                                    int J, const float *__restrict__ T, const int *__restrict__ member,
                                                    const float *__restrict__ L, float e0, float e1, float e2,
                                                    float *__restrict__ ln) {
    const int q = big_find(m_prefix, J, (int)blockIdx.x);
    const BigJob jb = jobs[q];
    const int i = (int)blockIdx.x - m_prefix[q];
    const float *row = T + jb.toff + (long long)i * jb.ldm;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    for (int j = threadIdx.x; j < jb.m; j += blockDim.x) {
        const float d = row[j];
        a0 += trans_f(e0, d);
        a1 += trans_f(e1, d);
        a2 += trans_f(e2, d);
    }
    __shared__ double s[3][4];
    double a[3] = {(double)a0, (double)a1, (double)a2};
    block256_sum3(a, s);
    if (threadIdx.x == 0) {
        const int vi = member[jb.a0 + i];
        for (int c = 0; c < 3; ++c) {
            const double sc = a[c];
            ln[3 * (long long)(jb.row0 + i) + c] = (float)((double)L[3 * vi + c] / (sc < 1.0 ? 1.0 : sc));
        }
    }
}

// diffuse_i = sum_j ln_j * exp(ext * D_ij), scattered to the component's global
// vertex ids.
__global__ void __launch_bounds__(256) k_big_diffuse(const BigJob *__restrict__ jobs, const int *__restrict__ m_prefix,
                                                     int J, const float *__restrict__ T, const int *__restrict__ member,
                                                     const float *__restrict__ ln, float e0, float e1, float e2,
                                                     float *__restrict__ out) {
    const int q = big_find(m_prefix, J, (int)blockIdx.x);
    const BigJob jb = jobs[q];
    const int i = (int)blockIdx.x - m_prefix[q];
    const float *row = T + jb.toff + (long long)i * jb.ldm;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int j = threadIdx.x; j < jb.m; j += blockDim.x) {
        const float d = row[j];
        const float *lj = ln + 3 * (long long)(jb.row0 + j);
        a0 += (double)(lj[0] * trans_f(e0, d));
        a1 += (double)(lj[1] * trans_f(e1, d));
        a2 += (double)(lj[2] * trans_f(e2, d));
    }
    __shared__ double s[3][4];
    double a[3] = {(double)a0, (double)a1, (double)a2};
    block256_sum3(a, s);
    if (threadIdx.x == 0) {
        const int vi = member[jb.a0 + i];
        for (int c = 0; c < 3; ++c) out[3 * vi + c] = (float)a[c];
    }
}

// optional dense outputs of the solve that owns the component: D and canonical
// predecessors (smallest neighbour k of j minimising D_ik + w_kj).
__global__ void __launch_bounds__(256) k_big_dense(const BigJob *__restrict__ jobs, const int *__restrict__ m_prefix,
                                                   int J, const float *__restrict__ T, const SolveDesc *__restrict__ sd,
                                                   const int *__restrict__ sid, const int *__restrict__ member,
                                                   const int *__restrict__ local, const int *__restrict__ aoff,
                                                   const int *__restrict__ anb, const float *__restrict__ aw) {
    const int q = big_find(m_prefix, J, (int)blockIdx.x);
    const BigJob jb = jobs[q];
    const int i = (int)blockIdx.x - m_prefix[q];
    const SolveDesc &d = sd[sid[member[jb.a0]]];
    const float *row = T + jb.toff + (long long)i * jb.ldm;
    const int vi = member[jb.a0 + i] - d.voff;
    for (int j = threadIdx.x; j < jb.m; j += blockDim.x) {
        const int vj = member[jb.a0 + j];
        if (d.dense_D) d.dense_D[(long long)vi * d.ld + (vj - d.voff)] = row[j];
        if (!d.dense_pred) continue;
        int best_k = -1;
        if (i == j) {
            best_k = vi;
        } else if (row[j] < INFINITY) {
            float best = INFINITY;
            for (int p = aoff[vj]; p < aoff[vj + 1]; ++p) {
                const int kv = anb[p];
                const float val = row[local[kv]] + aw[p];
                const int kl = kv - d.voff;
                if (val < best || (val == best && kl < best_k)) {
                    best = val;
                    best_k = kl;
                }
            }
        }
        d.dense_pred[(long long)vi * d.ld + (vj - d.voff)] = best_k;
    }
}

__global__ void k_pred_init(int n, int ld, int *__restrict__ pred) {
    const long long total = (long long)n * ld;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(k / ld), j = (int)(k % ld);
        pred[k] = (i == j) ? i : -1;
    }
}

// ------------------------------------------------------------- driver -------
template <typename F>
static void for_dtype(int dtype, F &&f) {
    if (dtype == LG_F32)
        f(float{});
    else if (dtype == LG_F16)
        f(_Float16{});
    else
        f(lgd::bf16_t{});
}

template <typename T>
static int32_t grow(lg_ctx *ctx, T **p, size_t *cap, size_t want) {
    if (want <= *cap && *p) return LG_OK;
    dfree(ctx, *p);
    *p = nullptr;
    *cap = 0;
    const size_t n = std::max<size_t>(want + want / 4, 256);
    LG_TRY(dalloc(ctx, n, p));
    *cap = n;
    return LG_OK;
}

int32_t Pipeline::reserve_vertices(lg_ctx *ctx, int V, int S) {
    if (V <= v_cap && S <= s_cap) return LG_OK;
    release_vertex(ctx);
    v_cap = std::max(V, 1);
    s_cap = std::max(S, 1);
    int32_t rc = LG_OK;
    auto A = [&](auto **p, size_t n) {
        if (rc == LG_OK) rc = dalloc(ctx, n, p);
    };
    A(&sd, (size_t)s_cap);
    A(&sid, (size_t)v_cap);
    A(&rowcnt, (size_t)v_cap + 1);
    A(&off, (size_t)v_cap + 1);
    A(&colcnt, (size_t)v_cap);
    A(&risky, (size_t)v_cap);
    A(&visit, (size_t)v_cap + 1);
    A(&visit_pos, (size_t)v_cap + 1);
    A(&visit_list, (size_t)v_cap);
    A(&adeg, (size_t)v_cap + 1);
    A(&aoff, (size_t)v_cap + 1);
    A(&acur, (size_t)v_cap);
    A(&parent, (size_t)v_cap);
    A(&root, (size_t)v_cap);
    A(&keys, (size_t)v_cap);
    A(&keys_sorted, (size_t)v_cap);
    A(&flag, (size_t)v_cap + 1);
    A(&cid, (size_t)v_cap + 1);
    A(&member, (size_t)v_cap);
    A(&local, (size_t)v_cap);
    A(&level, (size_t)v_cap);
    A(&coloff, (size_t)v_cap + 1);
    A(&colfill, (size_t)v_cap);
    A(&greedy_err, 1);
    A(&comp_off, (size_t)v_cap + 1);
    // counters[0..15] then the (component, size) pairs of the large components:
    // at most v_cap / (kSmallComp + 1) of them
    A(&counters, 16 + 2 * ((size_t)v_cap / (kSmallComp + 1) + 1));
    big_list = counters ? counters + 16 : nullptr;
    A(&cd, (size_t)s_cap);
    A(&cls16, (size_t)v_cap);
    A(&cls48, (size_t)v_cap);
    A(&cls128, (size_t)v_cap);
    A(&ckeys, (size_t)v_cap);
    A(&ckeys_sorted, (size_t)v_cap);
    A(&cvals, (size_t)v_cap);
    A(&cvals_sorted, (size_t)v_cap);
    // rocPRIM temporary storage for scans and sorts over up to v_cap keys
    size_t b1 = 0, b2 = 0, b3 = 0;
    if (rc == LG_OK &&
        (rocprim::radix_sort_keys(nullptr, b1, keys, keys_sorted, (size_t)v_cap, 0, 64, ctx->stream) != hipSuccess ||
         rocprim::exclusive_scan(nullptr, b2, rowcnt, off, 0, (size_t)v_cap + 1, rocprim::plus<int>(), ctx->stream) !=
             hipSuccess ||
         rocprim::radix_sort_pairs(nullptr, b3, ckeys, ckeys_sorted, cvals, cvals_sorted, (size_t)v_cap, 0, 32,
                                   ctx->stream) != hipSuccess)) {
        set_error("rocPRIM temporary-size query failed");
        rc = LG_ERR_HIP;
    }
    tmp_vertex_bytes = std::max(std::max(b1, b2), b3);
    A((uint8_t **)&tmp_vertex, tmp_vertex_bytes + 256);
    return rc;
}

void Pipeline::release_vertex(lg_ctx *ctx) {
    void *ps[] = {sd,     sid,  rowcnt,      off,    colcnt,       risky, visit,     visit_pos, visit_list,
                  adeg,   aoff, acur,        parent, root,         keys,  keys_sorted, flag,    cid,
                  member, local, level, coloff, colfill, greedy_err, comp_off, counters, tmp_vertex, cd,
                  ckeys, ckeys_sorted,
                  cvals,  cvals_sorted, cls16, cls48, cls128};
    for (void *p : ps) dfree(ctx, p);
    cls16 = cls48 = cls128 = nullptr;
    sd = nullptr;
    cd = nullptr;
    ckeys = ckeys_sorted = nullptr;
    cvals = cvals_sorted = nullptr;
    sid = rowcnt = off = colcnt = visit = visit_pos = visit_list = adeg = aoff = acur = parent = root = nullptr;
    flag = cid = member = local = level = comp_off = big_list = counters = nullptr;
    coloff = colfill = greedy_err = nullptr;
    risky = nullptr;
    keys = keys_sorted = nullptr;
    tmp_vertex = nullptr;
    v_cap = s_cap = 0;
}

int32_t Pipeline::pinned_words(size_t n, int **out) {
    if (n > h_pin_cap) {
        if (h_pin) (void)hipHostFree(h_pin);
        h_pin = nullptr;
        h_pin_cap = 0;
        const size_t want = std::max<size_t>(n + n / 4, 256);
        LG_HIP(hipHostMalloc((void **)&h_pin, want * sizeof(int), hipHostMallocDefault));
        h_pin_cap = want;
    }
    *out = h_pin;
    return LG_OK;
}

void Pipeline::release(lg_ctx *ctx) {
    release_vertex(ctx);
    if (h_pin) (void)hipHostFree(h_pin);
    h_pin = nullptr;
    h_pin_cap = 0;
    void *ps[] = {cand,  accepted, epos, tmp_cand, edges_, w_,   anb,    aw,      big_T,
                  big_ln, big_meta, ray, cstart,   cend,   stats, colslot, slot_info, lkeys, tmp_light};
    for (void *p : ps) dfree(ctx, p);
    stats = nullptr;
    cstart = cend = nullptr;
    cells_cap = cend_cap = 0;
    cand = accepted = epos = edges_ = anb = colslot = nullptr;
    slot_info = nullptr;
    colslot_cap = slot_info_cap = 0;
    w_ = aw = big_T = nullptr;
    big_meta = nullptr;
    big_ln = nullptr;
    ray = nullptr;
    lkeys = nullptr;
    tmp_light = nullptr;
    lsort_cap = tmp_light_cap = 0;
    tmp_cand = nullptr;
    cand_cap = accepted_cap = epos_cap = tmp_cand_cap = edges_cap = w_cap = anb_cap = aw_cap = 0;
    bigT_cap = bigln_cap = bigmeta_cap = ray_cap = 0;
}

// Clears up to 8 int regions in one launch (a stage's counters and flags
// together instead of one fill dispatch each).
// Up to 8 int regions to clear (null or empty ones are skipped).
static int32_t zero_list(std::initializer_list<std::pair<int *, size_t>> regions, ZeroList *out, size_t *most_out) {
    ZeroList z{};
    size_t most = 0;
    for (const auto &r : regions) {
        if (!r.first || r.second == 0) continue;
        LG_CHECK_ARG(z.count < 8, "zero_ints: more than 8 regions");
        z.p[z.count] = r.first;
        z.n[z.count] = (long long)r.second;
        ++z.count;
        most = std::max(most, r.second);
    }
    *out = z;
    if (most_out) *most_out = most;
    return LG_OK;
}

static int32_t zero_ints(lg_ctx *ctx, std::initializer_list<std::pair<int *, size_t>> regions) {
    ZeroList z{};
    size_t most = 0;
    LG_TRY(zero_list(regions, &z, &most));
    if (z.count == 0) return LG_OK;
    const unsigned blocks = (unsigned)std::min<size_t>(1024, (most + 255) / 256);
    k_zero<<<blocks, 256, 0, ctx->stream>>>(z);
    LG_LAUNCHED("k_zero");
    return LG_OK;
}

// Exclusive scan of up to kSmallScan ints in one workgroup: thread t sums its
// run of ceil(n / 1024) elements, the run sums are scanned in LDS, then each
// thread writes its run.  One dispatch instead of a device-wide scan's two.
constexpr size_t kSmallScan = 65536;

constexpr int kScanStage = 12288;  // inputs up to this size are staged through LDS

__global__ void __launch_bounds__(1024) k_scan_small(const int *__restrict__ in, int *__restrict__ out, int n) {
    // thread t scans its run of consecutive elements; the run totals are
    // scanned within each wave by shuffles and across the 16 waves in LDS.
    // Inputs of at most kScanStage elements move through LDS with coalesced
    // loads and stores (the runs are strided across the lanes).
    __shared__ int buf[kScanStage];
    __shared__ int wsum[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const bool staged = n <= kScanStage;
    if (staged) {
        for (int i = t; i < n; i += 1024) buf[i] = in[i];
        __syncthreads();
    }
    const int per = (n + 1023) / 1024;
    const int b = t * per, e = min(b + per, n);
    int sum = 0;
    for (int i = b; i < e; ++i) sum += staged ? buf[i] : in[i];
    const int v = lgd::wave_prefix_sum(sum);
    if (lane == 63) wsum[w] = v;
    __syncthreads();
    if (w == 0) {
        const int x = lgd::wave_prefix_sum(lane < 16 ? wsum[lane] : 0);
        if (lane < 16) wsum[lane] = x;
    }
    __syncthreads();
    int run = v - sum + (w > 0 ? wsum[w - 1] : 0);
    if (staged) {
        for (int i = b; i < e; ++i) {
            const int x = buf[i];
            buf[i] = run;
            run += x;
        }
        __syncthreads();
        for (int i = t; i < n; i += 1024) out[i] = buf[i];
    } else {
        for (int i = b; i < e; ++i) {
            const int x = in[i];
            out[i] = run;
            run += x;
        }
    }
}

static int32_t scan_excl(lg_ctx *ctx, Pipeline &pl, const int *in, int *out, size_t n, bool cand_sized) {
    if (n <= kSmallScan) {
        if (n > 0) k_scan_small<<<1, 1024, 0, ctx->stream>>>(in, out, (int)n);
        LG_LAUNCHED("k_scan_small");
        return LG_OK;
    }
    size_t bytes = 0;
    LG_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, 0, n, rocprim::plus<int>(), ctx->stream));
    void *tmp = nullptr;
    if (!cand_sized && bytes <= pl.tmp_vertex_bytes) {
        tmp = pl.tmp_vertex;
    } else {
        LG_TRY(grow(ctx, (uint8_t **)&pl.tmp_cand, &pl.tmp_cand_cap, bytes));
        tmp = pl.tmp_cand;
    }
    LG_HIP(rocprim::exclusive_scan(tmp, bytes, in, out, 0, n, rocprim::plus<int>(), ctx->stream));
    return LG_OK;
}

int32_t pl_begin(lg_ctx *ctx, Pipeline &pl, const std::vector<SolveDesc> &descs, int *V_out, int *nmax_out) {
    const int S = (int)descs.size();
    LG_CHECK_ARG(S >= 1, "pipeline: no solve");
    int V = 0, nmax = 0;
    for (auto &d : descs) {
        LG_CHECK_ARG(d.n >= 1 && d.n <= 65536, "pipeline: n=%d outside [1, 65536]", d.n);
        LG_CHECK_ARG(d.voff == V, "pipeline: vertex offsets must be contiguous");
        LG_CHECK_ARG(d.dtype == descs[0].dtype, "pipeline: all volumes of a batch share one brick type");
        V += d.n;
        nmax = std::max(nmax, d.n);
    }
    LG_TRY(pl.reserve_vertices(ctx, V, S));
    // neighbour-grid cell budget: at most (A + 1)^3 cells per solve, A = cbrt(2n) + 1
    std::vector<SolveDesc> up(descs);
    long long cells = 0;
    for (auto &d : up) {
        const int A = (int)floor(cbrt(2.0 * d.n)) + 1;
        d.cell_A = A;
        d.cell_base = (int)cells;
        cells += (long long)(A + 1) * (A + 1) * (A + 1);
    }
    LG_CHECK_ARG(cells < (1ll << 31), "pipeline: neighbour grid too large (%lld cells)", cells);
    pl.total_cells = cells;
    LG_HIP(hipMemcpyAsync(pl.sd, up.data(), sizeof(SolveDesc) * S, hipMemcpyHostToDevice, ctx->stream));
    k_solve_ids<<<(unsigned)((V + 255) / 256), 256, 0, ctx->stream>>>(pl.sd, S, V, pl.sid);
    LG_LAUNCHED("k_solve_ids");
    *V_out = V;
    if (nmax_out) *nmax_out = nmax;
    return LG_OK;
}

int32_t pl_sample(lg_ctx *ctx, Pipeline &pl, int V, int dtype, uint64_t seed, float *pts) {
    hipEvent_t ev;
    ktimer_begin(ctx, "k_sample_points", &ev);
    for_dtype(dtype, [&](auto tag) {
        using T = decltype(tag);
        k_sample<T><<<(unsigned)((V + 3) / 4), 256, 0, ctx->stream>>>(pl.sd, pl.sid, V, (uint32_t)seed,
                                                                         (uint32_t)(seed >> 32), pts);
    });
    LG_LAUNCHED("k_sample");
    ktimer_end(ctx, "k_sample_points", ev);
    return LG_OK;
}

int32_t pl_light(lg_ctx *ctx, Pipeline &pl, int V, int dtype, const LightConsts &lc, const lgl::DevLights &lights,
                 uint64_t env_seed, const float *pts, float *L) {
    hipEvent_t ev;
    ktimer_begin(ctx, "k_vertex_light", &ev);
    int nt = 0;
    for (int li = 0; li < lights.n; ++li) nt += lights.l[li].samples > 1 ? lights.l[li].samples * 10 : 1;
    const uint32_t ek0 = (uint32_t)env_seed, ek1 = (uint32_t)(env_seed >> 32);
    if (lc.exact_steps) {
        const long long rays = (long long)V * nt;
        if (rays > 0) {
            LG_TRY(grow(ctx, &pl.ray, &pl.ray_cap, (size_t)rays * 3));
            const int *order = nullptr;
            if (V >= kLightSortFromV) {
                // spatial order of the rays: Morton keys of the vertices, sorted
                LG_TRY(grow(ctx, &pl.lkeys, &pl.lsort_cap, (size_t)V * 4));
                uint32_t *keys = pl.lkeys, *keys_sorted = pl.lkeys + pl.lsort_cap / 4;
                int *vals = (int *)(pl.lkeys + 2 * (pl.lsort_cap / 4)), *vals_sorted = (int *)(pl.lkeys + 3 * (pl.lsort_cap / 4));
                size_t sbytes = 0;
                LG_HIP(rocprim::radix_sort_pairs(nullptr, sbytes, keys, keys_sorted, vals, vals_sorted, (size_t)V, 0, 15,
                                                 ctx->stream));
                LG_TRY(grow(ctx, &pl.tmp_light, &pl.tmp_light_cap, sbytes + 256));
                k_light_keys<<<(unsigned)((V + 255) / 256), 256, 0, ctx->stream>>>(pl.sd, pl.sid, V, pts, keys, vals);
                LG_LAUNCHED("k_light_keys");
                LG_HIP(rocprim::radix_sort_pairs(pl.tmp_light, sbytes, keys, keys_sorted, vals, vals_sorted, (size_t)V, 0,
                                                 15, ctx->stream));
                order = vals_sorted;
            }
            hipEvent_t ev_r;
            ktimer_begin(ctx, "k_light_rays", &ev_r);
            for_dtype(dtype, [&](auto tag) {
                using T = decltype(tag);
                if (V >= kLightRows16FromV) {
                    const unsigned gb = (unsigned)((rays + 15) / 16);
                    if (pl.stats_on)
                        k_light_rays16<T, true><<<gb, 256, 0, ctx->stream>>>(pl.sd, pl.sid, V, nt, lc, lights, ek0,
                                                                           ek1, pts, order, pl.ray, pl.stats);
                    else
                        k_light_rays16<T, false><<<gb, 256, 0, ctx->stream>>>(pl.sd, pl.sid, V, nt, lc, lights, ek0,
                                                                            ek1, pts, order, pl.ray, nullptr);
                } else if (pl.stats_on)
                    k_light_rays<T, true><<<(unsigned)((rays + 3) / 4), 256, 0, ctx->stream>>>(
                        pl.sd, pl.sid, V, nt, lc, lights, ek0, ek1, pts, order, pl.ray, pl.stats);
                else
                    k_light_rays<T, false><<<(unsigned)((rays + 3) / 4), 256, 0, ctx->stream>>>(
                        pl.sd, pl.sid, V, nt, lc, lights, ek0, ek1, pts, order, pl.ray, nullptr);
            });
            LG_LAUNCHED("k_light_rays");
            ktimer_end(ctx, "k_light_rays", ev_r);
        }
        k_light_combine<<<(unsigned)((V + 255) / 256), 256, 0, ctx->stream>>>(pl.sd, pl.sid, V, nt, lc, lights, pts,
                                                                            pl.ray, L);
        LG_LAUNCHED("k_light_combine");
    } else {
        for_dtype(dtype, [&](auto tag) {
            using T = decltype(tag);
            k_light<T><<<(unsigned)((V + 127) / 128), 128, 0, ctx->stream>>>(pl.sd, pl.sid, V, lc, lights, ek0, ek1,
                                                                            pts, L);
        });
        LG_LAUNCHED("k_light");
    }
    ktimer_end(ctx, "k_vertex_light", ev);
    return LG_OK;
}

// Candidate CSR, greedy cap and the edge list (pl.edges_, count at pl.epos[Ktot]).
int32_t pl_graph(lg_ctx *ctx, Pipeline &pl, int V, int S, int nmax, const float *pts, double r2, int max_connect,
                 int *Ktot_out, const std::function<int32_t()> *after_grid) {
    hipStream_t st = ctx->stream;
    const unsigned vb = (unsigned)((V + 255) / 256);
    hipEvent_t ev;
    ktimer_begin(ctx, "k_graph", &ev);
    const unsigned wb = (unsigned)((V + 3) / 4);
    // neighbour grid: cells at least r wide (a hair wider so that rounding in the
    // cell index cannot separate two points closer than r by two cells)
    const double cell_min = sqrt(r2) * (1.0 + 1e-6) + 1e-30;
    hipEvent_t ev_c;
    ktimer_begin(ctx, "k_cand", &ev_c);
    LG_TRY(grow(ctx, &pl.cstart, &pl.cells_cap, (size_t)pl.total_cells));
    LG_TRY(grow(ctx, &pl.cend, &pl.cend_cap, (size_t)pl.total_cells));
    const bool small_sort = pl.total_cells <= kCellSortCells && V <= kCellSortMaxV;
    // one solve with the small sort: its workgroup derives the grid too
    const int fused_setup = (small_sort && S == 1) ? 1 : 0;
    if (!fused_setup) {
        k_cell_setup<<<S, 1024, 0, st>>>(pl.sd, pts, cell_min, pl.cd);
        LG_LAUNCHED("k_cell_setup");
    }
    if (small_sort) {
        // one workgroup: counting sort, cell ranges and sorted list in one launch
        k_cell_sort_small<<<1, 1024, 0, st>>>(pl.cd, pl.sid, V, pts, (int)pl.total_cells, pl.ckeys, pl.cstart,
                                              pl.cend, pl.cvals_sorted, pl.rowcnt + V, pl.sd, cell_min, fused_setup);
        LG_LAUNCHED("k_cell_sort_small");
    } else {
        k_cell_keys<<<vb, 256, 0, st>>>(pl.cd, pl.sid, V, pts, pl.ckeys, pl.cvals);
        LG_LAUNCHED("k_cell_keys");
        int kbits = 1;
        while ((1ll << kbits) < pl.total_cells) ++kbits;
        size_t sbytes = pl.tmp_vertex_bytes;
        LG_HIP(rocprim::radix_sort_pairs(pl.tmp_vertex, sbytes, pl.ckeys, pl.ckeys_sorted, pl.cvals,
                                         pl.cvals_sorted, (size_t)V, 0, kbits, st));
        LG_TRY(zero_ints(ctx, {{pl.cstart, (size_t)pl.total_cells}, {pl.cend, (size_t)pl.total_cells},
                               {pl.rowcnt + V, 1}}));
        k_cell_bounds<<<vb, 256, 0, st>>>(V, pl.ckeys_sorted, pl.cstart, pl.cend);
        LG_LAUNCHED("k_cell_bounds");
    }
    if (after_grid && *after_grid) LG_TRY((*after_grid)());
    // rows are short: large calls give each vertex 4 lanes, small ones a wave
    const bool grouped = V >= kCandGroupedFromV;
    const unsigned cb = grouped ? (unsigned)((V + 63) / 64) : wb;
    if (grouped)
        k_cand<false, 4><<<cb, 256, 0, st>>>(pl.sd, pl.cd, pl.sid, V, pts, r2, pl.cstart, pl.cend, pl.cvals_sorted,
                                              nullptr, pl.rowcnt, nullptr, ZeroList{});
    else
        k_cand<false, 64><<<cb, 256, 0, st>>>(pl.sd, pl.cd, pl.sid, V, pts, r2, pl.cstart, pl.cend, pl.cvals_sorted,
                                              nullptr, pl.rowcnt, nullptr, ZeroList{});
    LG_LAUNCHED("k_cand");
    LG_TRY(scan_excl(ctx, pl, pl.rowcnt, pl.off, (size_t)V + 1, false));
    int *hk = nullptr;
    LG_TRY(pl.pinned_words(16, &hk));
    LG_HIP(hipMemcpyAsync(hk, pl.off + V, sizeof(int), hipMemcpyDeviceToHost, st));
    LG_HIP(hipStreamSynchronize(st));
    const int Ktot = hk[0];
    LG_TRY(grow(ctx, &pl.cand, &pl.cand_cap, (size_t)Ktot + 1));
    LG_TRY(grow(ctx, &pl.accepted, &pl.accepted_cap, (size_t)Ktot + 1));
    LG_TRY(grow(ctx, &pl.epos, &pl.epos_cap, (size_t)Ktot + 1));
    LG_TRY(grow(ctx, &pl.edges_, &pl.edges_cap, (size_t)2 * Ktot + 2));
    LG_TRY(grow(ctx, &pl.w_, &pl.w_cap, (size_t)Ktot + 1));
    const int cap = (max_connect > 0 && max_connect < 255) ? max_connect : 0;
    // the fill pass also clears the counters of the cap stage
    ZeroList zl{};
    LG_TRY(zero_list({{cap > 0 ? pl.colcnt : nullptr, (size_t)V},
                      {pl.visit + V, 1},
                      {pl.accepted + Ktot, 1},
                      {cap > 0 ? pl.colfill : nullptr, (size_t)V},
                      {pl.greedy_err, 1}},
                     &zl, nullptr));
    // the fill ranks each row's candidates in LDS, O(row^2 / lanes) per lane:
    // long rows (dense volumes) take more lanes per vertex
    const double avg_row = (double)Ktot / (double)V;
    const int gf = !grouped ? 64 : (avg_row <= kCandFill4MaxRow ? 4 : (avg_row <= kCandFill16MaxRow ? 16 : 64));
    const unsigned fb = (unsigned)((V + 4 * (64 / gf) - 1) / (4 * (64 / gf)));
    if (gf == 4)
        k_cand<true, 4><<<fb, 256, 0, st>>>(pl.sd, pl.cd, pl.sid, V, pts, r2, pl.cstart, pl.cend, pl.cvals_sorted,
                                             pl.off, nullptr, pl.cand, zl);
    else if (gf == 16)
        k_cand<true, 16><<<fb, 256, 0, st>>>(pl.sd, pl.cd, pl.sid, V, pts, r2, pl.cstart, pl.cend, pl.cvals_sorted,
                                              pl.off, nullptr, pl.cand, zl);
    else
        k_cand<true, 64><<<fb, 256, 0, st>>>(pl.sd, pl.cd, pl.sid, V, pts, r2, pl.cstart, pl.cend, pl.cvals_sorted,
                                              pl.off, nullptr, pl.cand, zl);
    LG_LAUNCHED("k_cand");
    ktimer_end(ctx, "k_cand", ev_c);
    if (cap > 0) k_colcount<<<vb, 256, 0, st>>>(V, pl.off, pl.cand, pl.colcnt);
    k_classify<<<vb, 256, 0, st>>>(V, cap, pl.off, pl.cand, pl.colcnt, pl.risky, pl.accepted, pl.visit);
    LG_LAUNCHED("k_classify");
    if (cap > 0) {
        LG_TRY(scan_excl(ctx, pl, pl.visit, pl.visit_pos, (size_t)V + 1, false));
        k_compact_rows<<<vb, 256, 0, st>>>(V, pl.visit, pl.visit_pos, pl.visit_list);
        hipEvent_t ev_g;
        ktimer_begin(ctx, "k_greedy", &ev_g);
        const size_t word_bytes = (size_t)nmax * sizeof(uint32_t);
        // dense graphs (many contested rows) decide rows in parallel; sparse ones
        // touch few rows, and the single-wave sweep avoids the rank launches
        if (word_bytes <= 60 * 1024 && (long long)Ktot >= 4ll * V) {
            // rows in parallel: column ranks of the risky slots, then the waves
            LG_TRY(grow(ctx, &pl.colslot, &pl.colslot_cap, (size_t)Ktot + 1));
            LG_TRY(grow(ctx, &pl.slot_info, &pl.slot_info_cap, (size_t)Ktot + 1));
            LG_TRY(scan_excl(ctx, pl, pl.colcnt, pl.coloff, (size_t)V, false));
            const unsigned kb = (unsigned)((Ktot + 255) / 256 + 1);
            k_col_fill<<<kb, 256, 0, st>>>(Ktot, pl.cand, pl.risky, pl.coloff, pl.colfill, pl.colslot);
            k_col_rank<<<kb, 256, 0, st>>>(Ktot, pl.cand, pl.risky, pl.coloff, pl.colcnt, pl.colslot, pl.slot_info);
            k_greedy_waves<<<S, 64 * kGreedyWaves, word_bytes, st>>>(pl.sd, cap, pl.off, pl.slot_info, pl.risky,
                                                                     pl.colcnt, pl.visit_list, pl.visit_pos + V,
                                                                     pl.accepted, pl.greedy_err);
        } else if (nmax <= 48 * 1024)
            k_greedy_staged<<<S, 64, (size_t)nmax, st>>>(pl.sd, cap, pl.off, pl.cand, pl.risky, pl.visit_list,
                                                         pl.visit_pos + V, pl.accepted);
        else
            k_greedy<<<S, 64, (size_t)nmax, st>>>(pl.sd, cap, pl.off, pl.cand, pl.risky, pl.visit_list,
                                                  pl.visit_pos + V, pl.accepted);
        LG_LAUNCHED("k_greedy");
        ktimer_end(ctx, "k_greedy", ev_g);
    }
    LG_TRY(scan_excl(ctx, pl, pl.accepted, pl.epos, (size_t)Ktot + 1, true));
    // the edge list's launch also clears the component stage's counters
    ZeroList zc{};
    LG_TRY(zero_list({{pl.adeg, (size_t)V + 1}, {pl.acur, (size_t)V}, {pl.flag + V, 1}, {pl.counters, 16}}, &zc,
                     nullptr));
    k_edges<<<vb, 256, 0, st>>>(V, pl.off, pl.cand, pl.accepted, pl.epos, pl.edges_, zc);
    LG_LAUNCHED("k_edges");
    ktimer_end(ctx, "k_graph", ev);
    *Ktot_out = Ktot;
    return LG_OK;
}

int32_t pl_march(lg_ctx *ctx, Pipeline &pl, int dtype, const MarchConsts &mc, const float *pts, int Ktot) {
    hipEvent_t ev;
    ktimer_begin(ctx, "k_march_edges", &ev);
    if (Ktot > 0) {
        for_dtype(dtype, [&](auto tag) {
            using T = decltype(tag);
            if (pl.stats_on)
                k_march<T, true><<<(unsigned)((Ktot + 127) / 128), 128, 0, ctx->stream>>>(
                    pl.sd, pl.sid, mc, pts, pl.edges_, Ktot, pl.epos + Ktot, pl.w_, pl.stats);
            else
                k_march<T, false><<<(unsigned)((Ktot + 127) / 128), 128, 0, ctx->stream>>>(
                    pl.sd, pl.sid, mc, pts, pl.edges_, Ktot, pl.epos + Ktot, pl.w_, nullptr);
        });
        LG_LAUNCHED("k_march");
    }
    ktimer_end(ctx, "k_march_edges", ev);
    return LG_OK;
}

// Components, closure (c) and diffusion; dense outputs when the descriptors ask.
int32_t pl_components(lg_ctx *ctx, Pipeline &pl, const std::vector<SolveDesc> &descs, int V, int Ktot,
                      const double ext[3], const float *L, float *out, PipelineResult *res, hipEvent_t L_ready,
                      const float *pts) {
    hipStream_t st = ctx->stream;
    const unsigned vb = (unsigned)((V + 255) / 256);
    const int *d_nedges = pl.epos + Ktot;
    hipEvent_t ev;
    ktimer_begin(ctx, "k_components", &ev);
    // adeg, acur, flag[V] and the counters were cleared by pl_graph's k_edges launch
    const unsigned eb = (unsigned)((Ktot + 255) / 256 + 1);
    // adjacency (degrees, scan, lists) and the union-find forest ride the same
    // passes over the edges
    hipEvent_t ev_u;
    ktimer_begin(ctx, "k_union_find", &ev_u);
    k_adj_deg<<<std::max(eb, vb), 256, 0, st>>>(pl.edges_, pl.w_, d_nedges, pl.adeg, V, pl.parent);
    LG_TRY(scan_excl(ctx, pl, pl.adeg, pl.aoff, (size_t)V + 1, false));
    LG_TRY(grow(ctx, &pl.anb, &pl.anb_cap, (size_t)2 * Ktot + 2));
    LG_TRY(grow(ctx, &pl.aw, &pl.aw_cap, (size_t)2 * Ktot + 2));
    k_adj_fill_union<<<eb, 256, 0, st>>>(pl.edges_, pl.w_, d_nedges, pl.aoff, pl.acur, pl.anb, pl.aw, pl.parent);
    LG_LAUNCHED("k_adj_fill_union");
    ktimer_end(ctx, "k_union_find", ev_u);
    hipEvent_t ev_s;
    ktimer_begin(ctx, "k_comp_sort", &ev_s);
    int bits = 1;
    while ((1 << bits) < V) ++bits;
    k_uf_root_keys<<<vb, 256, 0, st>>>(V, bits, pl.parent, pl.root, pl.keys);
    size_t sbytes = pl.tmp_vertex_bytes;
    LG_HIP(rocprim::radix_sort_keys(pl.tmp_vertex, sbytes, pl.keys, pl.keys_sorted, (size_t)V, 0, 2 * bits, st));
    k_comp_flags<<<vb, 256, 0, st>>>(V, bits, pl.keys_sorted, pl.flag, pl.member);
    LG_TRY(scan_excl(ctx, pl, pl.flag, pl.cid, (size_t)V + 1, false));
    k_comp_offsets<<<vb, 256, 0, st>>>(V, pl.flag, pl.cid, pl.comp_off, pl.cid + V);
    k_comp_local<<<vb, 256, 0, st>>>(V, pl.sd, pl.sid, pl.flag, pl.cid, pl.comp_off, pl.member, pl.local, pl.big_list,
                                     pl.counters, pl.cls16, pl.cls48, pl.cls128, pl.cid + V, d_nedges, pl.greedy_err);
    LG_LAUNCHED("k_comp_local");
    const int max_big = V / (kSmallComp + 1);
    if (pl.nd_order && max_big > 0) {
        LG_CHECK_ARG(pts, "pl_components: nested-dissection order needs the points");
        // segment tags in level[], classes and the partition scratch in the key buffer
        int *qa = reinterpret_cast<int *>(pl.keys);
        k_nd_order<<<(unsigned)max_big, 1024, 0, st>>>(pl.counters, pl.big_list, pl.comp_off, pl.member, pl.local, pts,
                                                        pl.aoff, pl.anb, pl.level, qa, qa + V);
        LG_LAUNCHED("k_nd_order");
    } else if (pl.bfs_order && max_big > 0) {
        LG_CHECK_ARG(3 * bits <= 64, "breadth-first order: %d vertices in one call exceed 2^21", V);
        LG_TRY(zero_ints(ctx, {{pl.level, (size_t)V}}));
        // the frontier queues borrow the 64-bit key buffer (2V ints), refilled below
        int *qa = reinterpret_cast<int *>(pl.keys);
        k_bfs_levels<<<(unsigned)max_big, 1024, 0, st>>>(pl.counters, pl.big_list, pl.comp_off, pl.member, pl.aoff,
                                                          pl.anb, pl.level, qa, qa + V);
        k_keys_bfs<<<vb, 256, 0, st>>>(V, bits, pl.root, pl.level, pl.keys);
        sbytes = pl.tmp_vertex_bytes;
        LG_HIP(rocprim::radix_sort_keys(pl.tmp_vertex, sbytes, pl.keys, pl.keys_sorted, (size_t)V, 0, 3 * bits, st));
        k_member_local<<<vb, 256, 0, st>>>(V, bits, pl.keys_sorted, pl.flag, pl.cid, pl.comp_off, pl.member,
                                           pl.local);
        LG_LAUNCHED("k_member_local");
    }
    ktimer_end(ctx, "k_comp_sort", ev_s);
    // counts and the large-component pairs come back in one synchronisation
    // (counters[0..4], component count, edge count, then the large-component
    // pairs: big_list = counters + 16)
    const size_t nrb = (size_t)16 + 2 * (size_t)max_big;
    int *hrb = nullptr;
    LG_TRY(pl.pinned_words(nrb, &hrb));
    LG_HIP(hipMemcpyAsync(hrb, pl.counters, sizeof(int) * nrb, hipMemcpyDeviceToHost, st));
    LG_HIP(hipStreamSynchronize(st));
    const std::vector<int> rb(hrb, hrb + nrb);
    ktimer_end(ctx, "k_components", ev);
    if (rb[10] != 0) {
        set_error("pl_graph: a row of the parallel greedy cap waited past its bound");
        return LG_ERR_HIP;
    }
    const int C = rb[8], nbig = rb[0], n16 = rb[1], n48 = rb[2], n128 = rb[3], n1 = rb[4];
    res->n_edges_total = rb[9];
    res->n_components = C;
    res->n_big = nbig;
    res->Ktot = Ktot;
    if (nbig > max_big) {
        set_error("pl_components: %d large components exceed the bound %d", nbig, max_big);
        return LG_ERR_HIP;
    }

    // dense outputs (optional) start as +INF / -1 with the diagonal
    bool any_dense = false;
    for (auto &d : descs) {
        if (d.dense_D) LG_TRY(fw_init_impl(ctx, d.n, d.dense_D, d.ld));
        if (d.dense_pred)
            k_pred_init<<<(unsigned)std::min<long long>(((long long)d.n * d.ld + 255) / 256, 65536), 256, 0, st>>>(
                d.n, d.ld, d.dense_pred);
        any_dense = any_dense || d.dense_D || d.dense_pred;
    }

    // (c) + diffusion: small components, one launch per size class.  They share
    // no output with the large components (disjoint rows of out and of the dense
    // matrices), so when large ones follow they run on the side stream, beside
    // the blocked closure, and the main stream joins them before returning.
    const float e0 = (float)ext[0], e1 = (float)ext[1], e2 = (float)ext[2];
    const bool side = nbig > 0 && (n128 + n48 + n16 + n1) > 0;
    auto small_closures = [&]() -> int32_t {
        // the first readers of the vertex light: three kernels, one per size
        // class, each timed on its own (the isolated vertices ride the last)
        if (L_ready) LG_HIP(hipStreamWaitEvent(ctx->stream, L_ready, 0));
        if (n128 > 0) {
            ktimer_begin(ctx, "k_comp_small_128", &ev);
            k_comp_small<kSmallComp><<<n128, 1024, 0, ctx->stream>>>(
                pl.sd, pl.sid, pl.cls128, pl.comp_off, pl.member, pl.local, pl.aoff, pl.anb, pl.aw, L, e0, e1, e2, out,
                pl.stats_on ? pl.stats : nullptr, n128, nullptr, 0);
            LG_LAUNCHED("k_comp_small<128>");
            ktimer_end(ctx, "k_comp_small_128", ev);
        }
        if (n48 > 0) {
            ktimer_begin(ctx, "k_comp_small_64", &ev);
            k_comp_small<64><<<n48, 256, 0, ctx->stream>>>(
                pl.sd, pl.sid, pl.cls48, pl.comp_off, pl.member, pl.local, pl.aoff, pl.anb, pl.aw, L, e0, e1, e2, out,
                pl.stats_on ? pl.stats : nullptr, n48, nullptr, 0);
            LG_LAUNCHED("k_comp_small<64>");
            ktimer_end(ctx, "k_comp_small_64", ev);
        }
        if (n16 + n1 > 0) {
            ktimer_begin(ctx, "k_comp_small_32", &ev);
            k_comp_small<32><<<(unsigned)(n16 + (n1 + 63) / 64), 64, 0, ctx->stream>>>(
                pl.sd, pl.sid, pl.cls16, pl.comp_off, pl.member, pl.local, pl.aoff, pl.anb, pl.aw, L, e0, e1, e2, out,
                pl.stats_on ? pl.stats : nullptr, n16, pl.cls16 + V, n1);
            LG_LAUNCHED("k_comp_small<32>");
            ktimer_end(ctx, "k_comp_small_32", ev);
        }
        return LG_OK;
    };
    struct JoinGuard {
        lg_ctx *ctx;
        bool on;
        ~JoinGuard() {
            if (on) (void)hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0);
        }
    } join{ctx, false};
    if (side) {
        LG_HIP(hipEventRecord(ctx->ev_fork, st));
        LG_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
        std::swap(ctx->stream, ctx->side);
        const int32_t rc = small_closures();
        std::swap(ctx->stream, ctx->side);
        LG_HIP(hipEventRecord(ctx->ev_join, ctx->side));
        join.on = true;
        if (rc != LG_OK) return rc;
    } else {
        LG_TRY(small_closures());
    }
    if (nbig == 0) {
        res->edges = pl.edges_;
        res->w = pl.w_;
        res->epos = pl.epos;
        res->off = pl.off;
        return LG_OK;
    }

    // large components: one arena, batched blocked Floyd-Warshall, batched diffusion
    std::vector<std::pair<int, int>> big((size_t)nbig);
    for (int b = 0; b < nbig; ++b) big[b] = {rb[16 + 2 * b], rb[17 + 2 * b]};
    std::sort(big.begin(), big.end());  // deterministic order (slots come from atomics)
    std::vector<BigJob> jobs((size_t)nbig);
    std::vector<int> ld_prefix((size_t)nbig), m_prefix((size_t)nbig);
    std::vector<float *> mats((size_t)nbig);
    std::vector<int> ms((size_t)nbig);
    long long toff = 0;
    int rows = 0, mrows = 0;
    for (int b = 0; b < nbig; ++b) {
        const int m = big[b].second, ldm = fw_ld(m);
        jobs[b] = BigJob{toff, 0, m, ldm, mrows, 0, 0};
        ld_prefix[b] = rows;
        m_prefix[b] = mrows;
        rows += ldm;
        mrows += m;
        toff += (long long)ldm * ldm;
        ms[b] = m;
    }
    // a0 = comp_off[c]: gathered on the device (k_big_fill) from the component ids
    LG_TRY(grow(ctx, &pl.big_T, &pl.bigT_cap, (size_t)toff));
    LG_TRY(grow(ctx, &pl.big_ln, &pl.bigln_cap, (size_t)3 * mrows));
    const size_t jb_bytes = sizeof(BigJob) * nbig, pre_bytes = sizeof(int) * nbig;
    LG_TRY(grow(ctx, &pl.big_meta, &pl.bigmeta_cap, jb_bytes + 2 * pre_bytes + sizeof(int) * nbig + 256));
    BigJob *d_jobs = (BigJob *)pl.big_meta;
    int *d_ldp = (int *)(pl.big_meta + jb_bytes);
    int *d_mp = d_ldp + nbig;
    int *d_comp = d_mp + nbig;
    // one upload of the job table, the two prefixes and the component ids
    std::vector<uint8_t> up(jb_bytes + 3 * pre_bytes);
    memcpy(up.data(), jobs.data(), jb_bytes);
    memcpy(up.data() + jb_bytes, ld_prefix.data(), pre_bytes);
    memcpy(up.data() + jb_bytes + pre_bytes, m_prefix.data(), pre_bytes);
    int *up_comp = (int *)(up.data() + jb_bytes + 2 * pre_bytes);
    for (int b = 0; b < nbig; ++b) up_comp[b] = big[b].first;
    LG_HIP(hipMemcpyAsync(d_jobs, up.data(), up.size(), hipMemcpyHostToDevice, st));
    ktimer_begin(ctx, "k_big_fill", &ev);
    k_big_fill<<<rows, 256, 0, st>>>(d_jobs, d_ldp, nbig, d_comp, pl.comp_off, pl.member, pl.local, pl.aoff, pl.anb,
                                     pl.aw, pl.big_T);
    LG_LAUNCHED("k_big_fill");
    ktimer_end(ctx, "k_big_fill", ev);
    for (int b = 0; b < nbig; ++b) mats[b] = pl.big_T + jobs[b].toff;
    LG_TRY(fw_solve_batched_impl(ctx, mats, ms, pl.stats_on ? pl.stats : nullptr));
    ktimer_begin(ctx, "k_diffuse", &ev);
    if (L_ready) LG_HIP(hipStreamWaitEvent(st, L_ready, 0));  // the closure above needs no vertex light
    k_big_rowsum<<<mrows, 256, 0, st>>>(d_jobs, d_mp, nbig, pl.big_T, pl.member, L, e0, e1, e2, pl.big_ln);
    LG_LAUNCHED("k_big_rowsum");
    k_big_diffuse<<<mrows, 256, 0, st>>>(d_jobs, d_mp, nbig, pl.big_T, pl.member, pl.big_ln, e0, e1, e2, out);
    LG_LAUNCHED("k_big_diffuse");
    ktimer_end(ctx, "k_diffuse", ev);
    if (any_dense) {
        k_big_dense<<<mrows, 256, 0, st>>>(d_jobs, d_mp, nbig, pl.big_T, pl.sd, pl.sid, pl.member, pl.local, pl.aoff,
                                           pl.anb, pl.aw);
        LG_LAUNCHED("k_big_dense");
    }
    res->edges = pl.edges_;
    res->w = pl.w_;
    res->epos = pl.epos;
    res->off = pl.off;
    return LG_OK;
}

void destroy_stage_pipeline(lg_ctx *ctx) {
    delete static_cast<Pipeline *>(ctx->stage_pipeline);
    ctx->stage_pipeline = nullptr;
}

}  // namespace lg
