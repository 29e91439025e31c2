// Solve stages of one LightGraph iteration (LightGraph.h:158-196): point
// sampling, graph build, edge march, vertex lighting, diffusion, and the solver
// object that chains them with Floyd-Warshall on one stream.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "lg_internal.h"
#include "lg_lights.h"

namespace lg {

double dropout_tau(const double ext[3], double cutoff) {
    // f(tau) = sum_c exp(2*ext_c*tau) is non-increasing for ext_c <= 0.
    auto f = [&](double tau) {
        double s = 0.0;
        for (int c = 0; c < 3; ++c) s += exp(2.0 * ext[c] * tau);
        return s;
    };
    for (int c = 0; c < 3; ++c)
        if (!(ext[c] <= 0.0)) return INFINITY;  // a channel never decays: no dropout
    if (!(f(0.0) >= cutoff)) return -1.0;       // every first sample already drops
    int zero = 0;
    for (int c = 0; c < 3; ++c) zero += ext[c] == 0.0;
    if ((double)zero >= cutoff) return INFINITY;
    double lo = 0.0, hi = 1.0;
    while (f(hi) >= cutoff) {
        lo = hi;
        hi *= 2.0;
        if (hi > 1e300) return INFINITY;
    }
    // largest tau with f(tau) >= cutoff
    for (int it = 0; it < 2000; ++it) {
        const double mid = lo + 0.5 * (hi - lo);
        if (mid <= lo || mid >= hi) break;
        if (f(mid) >= cutoff)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// ------------------------------------------------------------ sampling ------
// Brick masses: q_v = floor(rho_v * 2^24) for rho_v > 0; one wave per active brick.
template <typename T>
__global__ void k_brick_mass(const T *__restrict__ data, const int32_t *__restrict__ list, int nact,
                             uint64_t *__restrict__ mass) {
    const int b = blockIdx.x;
    if (b >= nact) return;
    const T *p = data + ((long long)list[b] << 9);
    uint64_t s = 0;
    for (int i = threadIdx.x; i < 512; i += 64) {
        const float v = lgd::to_f32(p[i]);
        if (v > 0.0f) s += (uint64_t)fmin((double)v * 16777216.0, 549755813888.0 * 16777216.0);
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (threadIdx.x == 0) mass[b] = s;
}

int32_t grid_build_sampling_tables(lg_ctx *ctx, lg_grid *g) {
    if (g->mass_prefix) return LG_OK;
    LG_CHECK_ARG(g->active_bricks > 0, "sampling: the grid has no active voxel");
    const size_t nbricks = (size_t)g->NB[0] * g->NB[1] * g->NB[2];
    std::vector<uint8_t> occ(nbricks);
    LG_HIP(hipMemcpyAsync(occ.data(), g->occ, nbricks, hipMemcpyDeviceToHost, ctx->stream));
    LG_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<int32_t> list;
    list.reserve(g->active_bricks);
    for (size_t b = 0; b < nbricks; ++b)
        if (occ[b]) list.push_back((int32_t)b);
    const int nact = (int)list.size();
    int32_t *d_list = nullptr;
    uint64_t *d_mass = nullptr, *d_prefix = nullptr;
    LG_TRY(dalloc(ctx, nact, &d_list));
    int32_t rc = dalloc(ctx, nact, &d_mass);
    if (rc == LG_OK) rc = dalloc(ctx, (size_t)nact + 1, &d_prefix);
    if (rc != LG_OK) {
        dfree(ctx, d_list);
        dfree(ctx, d_mass);
        return rc;
    }
    LG_HIP(hipMemcpyAsync(d_list, list.data(), nact * 4, hipMemcpyHostToDevice, ctx->stream));
    switch (g->desc.dtype) {
        case LG_F32:
            k_brick_mass<float><<<nact, 64, 0, ctx->stream>>>((const float *)g->data, d_list, nact, d_mass);
            break;
        case LG_F16:
            k_brick_mass<_Float16><<<nact, 64, 0, ctx->stream>>>((const _Float16 *)g->data, d_list, nact, d_mass);
            break;
        default:
            k_brick_mass<lgd::bf16_t><<<nact, 64, 0, ctx->stream>>>((const lgd::bf16_t *)g->data, d_list, nact, d_mass);
            break;
    }
    LG_LAUNCHED("k_brick_mass");
    std::vector<uint64_t> mass(nact), prefix((size_t)nact + 1);
    LG_HIP(hipMemcpyAsync(mass.data(), d_mass, (size_t)nact * 8, hipMemcpyDeviceToHost, ctx->stream));
    LG_HIP(hipStreamSynchronize(ctx->stream));
    prefix[0] = 0;
    for (int b = 0; b < nact; ++b) prefix[b + 1] = prefix[b] + mass[b];
    LG_HIP(hipMemcpyAsync(d_prefix, prefix.data(), prefix.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    LG_HIP(hipStreamSynchronize(ctx->stream));
    dfree(ctx, d_mass);
    if (prefix[nact] == 0) {
        dfree(ctx, d_list);
        dfree(ctx, d_prefix);
        set_error("sampling: the grid holds no positive density");
        return LG_ERR_INVALID;
    }
    g->active_list = d_list;
    g->mass_prefix = d_prefix;
    g->total_mass = prefix[nact];
    return LG_OK;
}

// Kernel (a).  Point n of solve s: target = mulhi64(u64, Q) picks a brick by
// binary search of the mass prefix and a voxel by a scan inside the brick; the
// point is jittered uniformly in the voxel cell [v-0.5, v+0.5).
template <typename T>
__global__ void k_sample_points(lgd::GridView g, const int32_t *__restrict__ list,
                                const uint64_t *__restrict__ prefix, int nact, uint64_t Q, uint32_t k0, uint32_t k1,
                                uint32_t solve, int n, float *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const lgd::U4 r0 = lgd::philox(lgd::U4{(uint32_t)i, solve, 0u, lgd::DOM_SAMPLE}, k0, k1);
    const lgd::U4 r1 = lgd::philox(lgd::U4{(uint32_t)i, solve, 1u, lgd::DOM_SAMPLE}, k0, k1);
    const uint64_t u = ((uint64_t)r0.y << 32) | (uint64_t)r0.x;
    const uint64_t target = __umul64hi(u, Q);
    // largest b with prefix[b] <= target
    int lo = 0, hi = nact - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= target)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int sb = list[lo];
    uint64_t local = target - prefix[lo];
    const T *p = static_cast<const T *>(g.data) + ((long long)sb << 9);
    int v = 511;
    uint64_t cum = 0;
    for (int k = 0; k < 512; ++k) {
        const float f = lgd::to_f32(p[k]);
        if (f > 0.0f) cum += (uint64_t)fmin((double)f * 16777216.0, 549755813888.0 * 16777216.0);
        if (local < cum) {
            v = k;
            break;
        }
    }
    const int sbx = sb % g.NBx, sby = (sb / g.NBx) % g.NBy, sbz = sb / (g.NBx * g.NBy);
    const int vx = (sbx - 1) * 8 + (v & 7), vy = (sby - 1) * 8 + ((v >> 3) & 7), vz = (sbz - 1) * 8 + (v >> 6);
    const double px = (double)vx - 0.5 + (double)lgd::u01(r0.z);
    const double py = (double)vy - 0.5 + (double)lgd::u01(r0.w);
    const double pz = (double)vz - 0.5 + (double)lgd::u01(r1.x);
    out[3 * i + 0] = (float)(g.origin[0] + g.voxel * px);
    out[3 * i + 1] = (float)(g.origin[1] + g.voxel * py);
    out[3 * i + 2] = (float)(g.origin[2] + g.voxel * pz);
}

// --------------------------------------------------------- graph build ------
__device__ __forceinline__ double dist2(const float *__restrict__ pts, int i, int j) {
    const double dx = (double)pts[3 * i + 0] - (double)pts[3 * j + 0];
    const double dy = (double)pts[3 * i + 1] - (double)pts[3 * j + 1];
    const double dz = (double)pts[3 * i + 2] - (double)pts[3 * j + 2];
    return dx * dx + dy * dy + dz * dz;
}

// Candidates of row i: j > i with |p_i - p_j|^2 < r2, in ascending j.  One wave
// per row; fill == nullptr counts only.
__global__ void k_candidates(const float *__restrict__ pts, int n, double r2, const int32_t *__restrict__ off,
                             int32_t *__restrict__ count, int32_t *__restrict__ fill) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (i >= n) return;
    int c = 0;
    const int base_out = fill ? off[i] : 0;
    for (int base = i + 1; base < n; base += 64) {
        const int j = base + lane;
        const bool hit = j < n && dist2(pts, i, j) < r2;
        const uint64_t m = __ballot(hit);
        if (fill && hit) {
            const int pos = __popcll(m & ((1ull << lane) - 1ull));
            fill[base_out + c + pos] = j;
        }
        c += __popcll(m);
    }
    if (!fill && lane == 0) count[i] = c;
}

// Exclusive scan of n int32 (one block of 1024 threads); out has n+1 entries.
__global__ void k_scan_i32(const int32_t *__restrict__ in, int n, int32_t *__restrict__ out) {
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int s = t * per, e = min(n, s + per);
    int64_t acc = 0;
    for (int k = s; k < e; ++k) acc += in[k];
    part[t] = acc;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int64_t v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t run = t ? part[t - 1] : 0;
    for (int k = s; k < e; ++k) {
        out[k] = (int32_t)run;
        run += in[k];
    }
    if (t == 1023) out[n] = (int32_t)part[1023];
}

// Greedy capped selection (LightGraph.h:297-310), one wave for the whole graph:
// rows in ascending order; within a row the first (cap - cnt[i]) candidates j with
// cnt[j] < cap.  Edges are appended in the reference's row-major scan order.
__global__ void __launch_bounds__(64) k_greedy(int n, int cap, const int32_t *__restrict__ off,
                                               const int32_t *__restrict__ cand, int32_t *__restrict__ edges,
                                               int edge_cap, int32_t *__restrict__ n_edges) {
    extern __shared__ uint8_t cnt[];
    const int lane = threadIdx.x;
    for (int k = lane; k < n; k += 64) cnt[k] = 0;
    __syncthreads();
    int E = 0;
    bool overflow = false;
    for (int i = 0; i < n; ++i) {
        int ci = cnt[i];
        if (ci >= cap) continue;
        const int s = off[i], e = off[i + 1];
        for (int base = s; base < e; base += 64) {
            const int idx = base + lane;
            const int j = idx < e ? cand[idx] : -1;
            const bool elig = j >= 0 && cnt[j] < cap;
            const uint64_t m = __ballot(elig);
            const int need = cap - ci;
            const int pos = __popcll(m & ((1ull << lane) - 1ull));
            const bool sel = elig && pos < need;
            if (sel) {
                cnt[j] = (uint8_t)(cnt[j] + 1);
                const int slot = E + pos;
                if (slot < edge_cap) {
                    edges[2 * slot + 0] = i;
                    edges[2 * slot + 1] = j;
                }
            }
            const int taken = min(__popcll(m), need);
            E += taken;
            ci += taken;
            __syncthreads();
            if (ci >= cap) break;
        }
        if (lane == 0) cnt[i] = (uint8_t)ci;
        __syncthreads();
    }
    if (E > edge_cap) overflow = true;
    if (lane == 0) *n_edges = overflow ? -E : E;
}

// Uncapped: every candidate is an edge; the candidate CSR is the edge list.
__global__ void k_uncapped_edges(int n, const int32_t *__restrict__ off, const int32_t *__restrict__ cand,
                                 int32_t *__restrict__ edges, int edge_cap, int32_t *__restrict__ n_edges) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = off[i]; k < off[i + 1]; ++k) {
        if (k < edge_cap) {
            edges[2 * k + 0] = i;
            edges[2 * k + 1] = cand[k];
        }
    }
    if (i == n - 1) *n_edges = off[n] > edge_cap ? -off[n] : off[n];
}

// ------------------------------------------------------------ D init --------
__global__ void k_fw_init(int n, int ld, float *__restrict__ D) {
    const long long total = (long long)ld * ld;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(k / ld), j = (int)(k % ld);
        D[k] = (i == j) ? 0.0f : INFINITY;
    }
    (void)n;
}

// ---------------------------------------------------------- edge march ------
struct MarchConsts {
    double ds_sds_scds[3];  // densityScale, shadowDensityScale, scatterDensityScale
    double dmax, sdm, cutoff, h, tau_star;
};

// Kernel (b): one lane per edge (LightGraph.h:359-407).
template <typename T>
__global__ void k_march_edges(lgd::GridView g, MarchConsts mc, const float *__restrict__ pts,
                              const int32_t *__restrict__ edges, int edge_limit,
                              const int32_t *__restrict__ n_edges_dev, float *__restrict__ w,
                              float *__restrict__ D, int ld) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int ne = n_edges_dev ? min(abs(*n_edges_dev), edge_limit) : edge_limit;
    if (e >= ne) return;
    const int i = edges[2 * e], j = edges[2 * e + 1];
    const double sp[3] = {(double)pts[3 * i], (double)pts[3 * i + 1], (double)pts[3 * i + 2]};
    double dir[3] = {(double)pts[3 * j] - sp[0], (double)pts[3 * j + 1] - sp[1], (double)pts[3 * j + 2] - sp[2]};
    const double len = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    const float dist = (float)len;
    if (!(fabs(len - 0.0) <= 1e-7)) {  // Vec3::normalize(eps=1e-7): *this *= 1/len
        const double inv = 1.0 / len;
        for (int c = 0; c < 3; ++c) dir[c] = dir[c] * inv;
    }
    float tau_out = INFINITY;
    lgd::Isect s;
    if (lgd::isect_world_ray(g, s, sp, dir, lgd::kDeltaD, (double)dist)) {
        double tau = 0.0;
        bool dropped = false;
        double t0, t1;
        while (!dropped && lgd::isect_march(g, s, t0, t1)) {
            for (double t = mc.h * ceil(t0 / mc.h); t <= t1; t += mc.h) {
                const float rho = lgd::trilinear<T>(g, s.e[0] + s.d[0] * t, s.e[1] + s.d[1] * t, s.e[2] + s.d[2] * t);
                double dens = (double)rho * mc.ds_sds_scds[0] * mc.ds_sds_scds[1] * mc.ds_sds_scds[2];
                dens = (mc.dmax < dens) ? mc.dmax : dens;
                if (dens < mc.cutoff && mc.sdm < mc.cutoff) continue;
                const double m = (dens < mc.sdm) ? mc.sdm : dens;
                tau += m * mc.h;
                if (tau > mc.tau_star) {
                    dropped = true;
                    break;
                }
            }
        }
        tau_out = dropped ? INFINITY : (float)tau;
    }
    if (w) w[e] = tau_out;
    if (D) {
        D[(long long)i * ld + j] = tau_out;
        D[(long long)j * ld + i] = tau_out;
    }
}

// ----------------------------------------------------- vertex lighting ------
struct LightConsts {
    double ds_sds_scds[3];
    double dmax, cutoff, h, tau_star;
    double albedo[3], ext[3];
    double emission_scale;
    double ramp_in, ramp_out;
    int ramp_n;
    double ramp_pos[LG_MAX_RAMP];
    double ramp_col[LG_MAX_RAMP][3];
};

// ColorRamp::getColor (Shader.h:63-92) in FP64, NaN behaviour included.
__device__ __forceinline__ void ramp_color(const LightConsts &c, double x, double out[3]) {
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

// One lane per vertex (LightGraph.h:438-497).
template <typename T>
__global__ void k_vertex_light(lgd::GridView g, lgd::GridView tg, int has_temp, LightConsts lc, lgl::DevLights L,
                               uint32_t ek0, uint32_t ek1, uint32_t solve, const float *__restrict__ pts, int n,
                               float *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double p[3] = {(double)pts[3 * i], (double)pts[3 * i + 1], (double)pts[3 * i + 2]};
    double emission[3] = {0.0, 0.0, 0.0};
    if (lc.ramp_n > 0) {
        double temp = 0.0;
        if (has_temp) {
            const double x = (p[0] - tg.origin[0]) * tg.inv_voxel, y = (p[1] - tg.origin[1]) * tg.inv_voxel,
                         z = (p[2] - tg.origin[2]) * tg.inv_voxel;
            temp = (double)lgd::trilinear<float>(tg, x, y, z);
        }
        ramp_color(lc, temp, emission);
        for (int k = 0; k < 3; ++k) emission[k] = emission[k] * lc.emission_scale;
    }
    lgd::Stream env_rng;
    env_rng.init((uint32_t)i, solve, lgd::DOM_VERTEX_ENV, ((uint64_t)ek1 << 32) | ek0);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int li = 0; li < L.n; ++li) {
        const lgl::DevLight &lt = L.l[li];
        int samples = lt.samples;
        if (samples > 1) samples *= 10;  // LightGraph.h:450-452
        double lumi[3] = {0.0, 0.0, 0.0};
        for (int s = 0; s < samples; ++s) {
            double ldir[3], inten[3], distance;
            lgl::illuminate(lt, p, env_rng, ldir, inten, distance);
            const double sd[3] = {-ldir[0], -ldir[1], -ldir[2]};
            lgd::Isect is;
            if (!lgd::isect_world_ray(g, is, p, sd, lgd::kDeltaD, distance)) continue;
            double tau = 0.0;
            bool dropped = false;
            double t0, t1;
            while (!dropped && lgd::isect_march(g, is, t0, t1)) {
                for (double t = lc.h * ceil(t0 / lc.h); t <= t1; t += lc.h) {
                    const float rho =
                        lgd::trilinear<T>(g, is.e[0] + is.d[0] * t, is.e[1] + is.d[1] * t, is.e[2] + is.d[2] * t);
                    double dens = (double)rho * lc.ds_sds_scds[0] * lc.ds_sds_scds[1] * lc.ds_sds_scds[2];
                    dens = (lc.dmax < dens) ? lc.dmax : dens;
                    if (dens < lc.cutoff) continue;
                    tau += dens * lc.h;
                    if (tau > lc.tau_star) {
                        dropped = true;
                        break;
                    }
                }
            }
            if (dropped) continue;  // transmittance set to zero
            for (int k = 0; k < 3; ++k) lumi[k] += inten[k] * (double)expf((float)(lc.ext[k] * tau));
        }
        for (int k = 0; k < 3; ++k) acc[k] += lc.albedo[k] * (lumi[k] / (double)samples);
    }
    for (int k = 0; k < 3; ++k) out[3 * i + k] = (float)(acc[k] + emission[k]);
}

// ------------------------------------------------------------ diffusion ------
// Row sums of P_c = exp(ext_c * D) over the first n columns (= column sums, D is
// symmetric).  One block of 256 threads per row.
__global__ void k_rowsum(const float *__restrict__ D, int n, int ld, float e0, float e1, float e2,
                         double *__restrict__ sums) {
    const int i = blockIdx.x;
    const float *row = D + (long long)i * ld;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const float d = row[j];
        a0 += __expf(e0 * d);
        a1 += __expf(e1 * d);
        a2 += __expf(e2 * d);
    }
    __shared__ double s[3][256];
    s[0][threadIdx.x] = a0;
    s[1][threadIdx.x] = a1;
    s[2][threadIdx.x] = a2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int c = 0; c < 3; ++c) s[c][threadIdx.x] += s[c][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) sums[3 * i + c] = s[c][0];
}

__global__ void k_lnorm(const float *__restrict__ L, const double *__restrict__ sums, int n, float *__restrict__ Ln) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    for (int c = 0; c < 3; ++c) {
        const double s = sums[3 * j + c];
        Ln[3 * j + c] = (float)((double)L[3 * j + c] / (s < 1.0 ? 1.0 : s));
    }
}

__global__ void k_diffuse_out(const float *__restrict__ D, int n, int ld, float e0, float e1, float e2,
                              const float *__restrict__ Ln, float *__restrict__ out) {
    const int i = blockIdx.x;
    const float *row = D + (long long)i * ld;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const float d = row[j];
        a0 += Ln[3 * j + 0] * __expf(e0 * d);
        a1 += Ln[3 * j + 1] * __expf(e1 * d);
        a2 += Ln[3 * j + 2] * __expf(e2 * d);
    }
    __shared__ double s[3][256];
    s[0][threadIdx.x] = a0;
    s[1][threadIdx.x] = a1;
    s[2][threadIdx.x] = a2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int c = 0; c < 3; ++c) s[c][threadIdx.x] += s[c][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) out[3 * i + c] = (float)s[c][0];
}

// ------------------------------------------------------------- helpers ------
static MarchConsts march_consts(const lg_shader *sh, const lg_march_params *mp) {
    MarchConsts mc{};
    mc.ds_sds_scds[0] = sh->density_scale;
    mc.ds_sds_scds[1] = sh->shadow_density_scale;
    mc.ds_sds_scds[2] = sh->scatter_density_scale;
    mc.dmax = sh->density_max;
    mc.sdm = sh->scatter_density_min;
    mc.cutoff = mp->cutoff;
    mc.h = mp->step[1];
    mc.tau_star = dropout_tau(sh->extinction, mp->cutoff);
    return mc;
}

static LightConsts light_consts(const lg_shader *sh, const lg_march_params *mp) {
    LightConsts lc{};
    lc.ds_sds_scds[0] = sh->density_scale;
    lc.ds_sds_scds[1] = sh->shadow_density_scale;
    lc.ds_sds_scds[2] = sh->scatter_density_scale;
    lc.dmax = sh->density_max;
    lc.cutoff = mp->cutoff;
    lc.h = mp->step[1];
    lc.tau_star = dropout_tau(sh->extinction, mp->cutoff);
    for (int c = 0; c < 3; ++c) {
        lc.albedo[c] = sh->albedo[c];
        lc.ext[c] = sh->extinction[c];
    }
    lc.emission_scale = sh->emission_scale;
    lc.ramp_in = sh->ramp_in;
    lc.ramp_out = sh->ramp_out;
    lc.ramp_n = std::max(0, std::min(sh->ramp_count, LG_MAX_RAMP));
    for (int k = 0; k < lc.ramp_n; ++k) {
        lc.ramp_pos[k] = sh->ramp_pos[k];
        for (int c = 0; c < 3; ++c) lc.ramp_col[k][c] = sh->ramp_color[k][c];
    }
    return lc;
}

static int32_t check_params(const lg_shader *sh, const lg_march_params *mp) {
    LG_CHECK_ARG(sh && mp, "shader/params are NULL");
    LG_CHECK_ARG(mp->step[0] > 0.0 && mp->step[1] > 0.0, "march steps must be positive");
    return LG_OK;
}

int32_t launch_march_edges(lg_ctx *ctx, const lg_grid *grid, const lg_shader *sh, const lg_march_params *mp,
                           const float *d_pts, const int32_t *d_edges, int edge_limit, const int32_t *d_nedges,
                           float *d_w, float *d_D, int ld) {
    if (edge_limit <= 0) return LG_OK;
    const lgd::GridView g = make_view(grid);
    const MarchConsts mc = march_consts(sh, mp);
    const unsigned blocks = (unsigned)((edge_limit + 255) / 256);
    hipEvent_t ev;
    ktimer_begin(ctx, "k_march_edges", &ev);
    switch (grid->desc.dtype) {
        case LG_F32:
            k_march_edges<float><<<blocks, 256, 0, ctx->stream>>>(g, mc, d_pts, d_edges, edge_limit, d_nedges, d_w, d_D, ld);
            break;
        case LG_F16:
            k_march_edges<_Float16><<<blocks, 256, 0, ctx->stream>>>(g, mc, d_pts, d_edges, edge_limit, d_nedges, d_w, d_D, ld);
            break;
        default:
            k_march_edges<lgd::bf16_t><<<blocks, 256, 0, ctx->stream>>>(g, mc, d_pts, d_edges, edge_limit, d_nedges, d_w, d_D, ld);
            break;
    }
    LG_LAUNCHED("k_march_edges");
    ktimer_end(ctx, "k_march_edges", ev);
    return LG_OK;
}

int32_t launch_vertex_light(lg_ctx *ctx, const lg_grid *density, const lg_grid *temperature, const lg_shader *sh,
                            const lg_light *lights, int n_lights, const lg_march_params *mp, uint64_t env_seed,
                            uint32_t solve, const float *d_pts, int n, float *d_L) {
    if (n <= 0) return LG_OK;
    lgl::DevLights L{};
    LG_TRY(lgl::make_dev_lights(lights, n_lights, &L));
    const lgd::GridView g = make_view(density);
    lgd::GridView tg{};
    int has_temp = 0;
    if (temperature) {
        LG_CHECK_ARG(temperature->desc.dtype == LG_F32, "temperature grid must be float32");
        tg = make_view(temperature);
        has_temp = 1;
    }
    const LightConsts lc = light_consts(sh, mp);
    const unsigned blocks = (unsigned)((n + 127) / 128);
    hipEvent_t ev;
    ktimer_begin(ctx, "k_vertex_light", &ev);
    switch (density->desc.dtype) {
        case LG_F32:
            k_vertex_light<float><<<blocks, 128, 0, ctx->stream>>>(g, tg, has_temp, lc, L, (uint32_t)env_seed,
                                                                    (uint32_t)(env_seed >> 32), solve, d_pts, n, d_L);
            break;
        case LG_F16:
            k_vertex_light<_Float16><<<blocks, 128, 0, ctx->stream>>>(g, tg, has_temp, lc, L, (uint32_t)env_seed,
                                                                       (uint32_t)(env_seed >> 32), solve, d_pts, n, d_L);
            break;
        default:
            k_vertex_light<lgd::bf16_t><<<blocks, 128, 0, ctx->stream>>>(g, tg, has_temp, lc, L, (uint32_t)env_seed,
                                                                          (uint32_t)(env_seed >> 32), solve, d_pts, n, d_L);
            break;
    }
    LG_LAUNCHED("k_vertex_light");
    ktimer_end(ctx, "k_vertex_light", ev);
    return LG_OK;
}

int32_t launch_sample(lg_ctx *ctx, lg_grid *grid, uint64_t seed, uint32_t solve, int n, float *d_pts) {
    LG_TRY(grid_build_sampling_tables(ctx, grid));
    const lgd::GridView g = make_view(grid);
    const unsigned blocks = (unsigned)((n + 127) / 128);
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    hipEvent_t ev;
    ktimer_begin(ctx, "k_sample_points", &ev);
    switch (grid->desc.dtype) {
        case LG_F32:
            k_sample_points<float><<<blocks, 128, 0, ctx->stream>>>(g, grid->active_list, grid->mass_prefix,
                                                                     grid->active_bricks, grid->total_mass, k0, k1, solve, n, d_pts);
            break;
        case LG_F16:
            k_sample_points<_Float16><<<blocks, 128, 0, ctx->stream>>>(g, grid->active_list, grid->mass_prefix,
                                                                        grid->active_bricks, grid->total_mass, k0, k1, solve, n, d_pts);
            break;
        default:
            k_sample_points<lgd::bf16_t><<<blocks, 128, 0, ctx->stream>>>(g, grid->active_list, grid->mass_prefix,
                                                                           grid->active_bricks, grid->total_mass, k0, k1, solve, n, d_pts);
            break;
    }
    LG_LAUNCHED("k_sample_points");
    ktimer_end(ctx, "k_sample_points", ev);
    return LG_OK;
}

// Candidate CSR + greedy/uncapped edge list.  The candidate total is read back
// (one synchronisation) to size the candidate buffer.
int32_t build_edges_impl(lg_ctx *ctx, const float *d_pts, int n, double radius, int max_connect,
                         int32_t *d_edges, int edge_cap, int32_t *d_nedges, int32_t **cand_buf,
                         int64_t *cand_cap, int32_t *count_buf, int32_t *off_buf) {
    LG_CHECK_ARG(n >= 1 && n <= 65536, "graph: n=%d out of range [1, 65536]", n);
    const float rf = (float)radius;
    const float r2f = rf * rf;  // float, as searchRadiusConnectSquared (LightGraph.h:171)
    const double r2 = (double)r2f;
    const unsigned rows_blocks = (unsigned)((n + 3) / 4);
    hipEvent_t ev;
    ktimer_begin(ctx, "k_graph", &ev);
    k_candidates<<<rows_blocks, 256, 0, ctx->stream>>>(d_pts, n, r2, nullptr, count_buf, nullptr);
    LG_LAUNCHED("k_candidates");
    k_scan_i32<<<1, 1024, 0, ctx->stream>>>(count_buf, n, off_buf);
    LG_LAUNCHED("k_scan_i32");
    int32_t total = 0;
    LG_HIP(hipMemcpyAsync(&total, off_buf + n, 4, hipMemcpyDeviceToHost, ctx->stream));
    LG_HIP(hipStreamSynchronize(ctx->stream));
    if ((int64_t)total > *cand_cap) {
        dfree(ctx, *cand_buf);
        *cand_buf = nullptr;
        *cand_cap = 0;
        const int64_t want = std::max<int64_t>((int64_t)total + total / 4, 1024);
        LG_TRY(dalloc(ctx, (size_t)want, cand_buf));
        *cand_cap = want;
    }
    k_candidates<<<rows_blocks, 256, 0, ctx->stream>>>(d_pts, n, r2, off_buf, count_buf, *cand_buf);
    LG_LAUNCHED("k_candidates");
    if (max_connect > 0 && max_connect < 255) {
        k_greedy<<<1, 64, (size_t)n, ctx->stream>>>(n, max_connect, off_buf, *cand_buf, d_edges, edge_cap, d_nedges);
        LG_LAUNCHED("k_greedy");
    } else {
        k_uncapped_edges<<<(n + 255) / 256, 256, 0, ctx->stream>>>(n, off_buf, *cand_buf, d_edges, edge_cap, d_nedges);
        LG_LAUNCHED("k_uncapped_edges");
    }
    ktimer_end(ctx, "k_graph", ev);
    return LG_OK;
}

int32_t diffuse_impl(lg_ctx *ctx, int n, const float *d_D, int ld, const double sigma_t[3], const float *d_L,
                     float *d_out, double *d_sums, float *d_Ln) {
    if (n <= 0) return LG_OK;
    const float e0 = (float)(-sigma_t[0]), e1 = (float)(-sigma_t[1]), e2 = (float)(-sigma_t[2]);
    hipEvent_t ev;
    ktimer_begin(ctx, "k_diffuse", &ev);
    k_rowsum<<<n, 256, 0, ctx->stream>>>(d_D, n, ld, e0, e1, e2, d_sums);
    LG_LAUNCHED("k_rowsum");
    k_lnorm<<<(n + 255) / 256, 256, 0, ctx->stream>>>(d_L, d_sums, n, d_Ln);
    LG_LAUNCHED("k_lnorm");
    k_diffuse_out<<<n, 256, 0, ctx->stream>>>(d_D, n, ld, e0, e1, e2, d_Ln, d_out);
    LG_LAUNCHED("k_diffuse_out");
    ktimer_end(ctx, "k_diffuse", ev);
    return LG_OK;
}

int32_t fw_init_impl(lg_ctx *ctx, int n, float *d_D, int ld) {
    k_fw_init<<<std::min<long long>(((long long)ld * ld + 255) / 256, 65536), 256, 0, ctx->stream>>>(n, ld, d_D);
    LG_LAUNCHED("k_fw_init");
    return LG_OK;
}

}  // namespace lg

using namespace lg;

extern "C" {

int32_t lg_sample_points(lg_ctx *ctx, const lg_grid *grid, uint64_t seed, uint32_t solve, int32_t n, float *d_pts) {
    LG_CHECK_ARG(ctx && grid && d_pts, "lg_sample_points: NULL argument");
    LG_CHECK_ARG(n >= 0, "lg_sample_points: n < 0");
    if (n == 0) return LG_OK;
    return launch_sample(ctx, const_cast<lg_grid *>(grid), seed, solve, n, d_pts);
}

int32_t lg_build_edges(lg_ctx *ctx, const float *d_pts, int32_t n, double radius, int32_t max_connect,
                       int32_t *d_edges, int32_t edge_cap, int32_t *n_edges_out) {
    LG_CHECK_ARG(ctx && d_pts && d_edges && n_edges_out, "lg_build_edges: NULL argument");
    LG_CHECK_ARG(edge_cap >= 0, "lg_build_edges: edge_cap < 0");
    int32_t *count = nullptr, *off = nullptr, *nedges = nullptr, *cand = nullptr;
    int64_t cand_cap = 0;
    LG_TRY(dalloc(ctx, (size_t)n + 1, &count));
    int32_t rc = dalloc(ctx, (size_t)n + 1, &off);
    if (rc == LG_OK) rc = dalloc(ctx, 1, &nedges);
    if (rc == LG_OK)
        rc = build_edges_impl(ctx, d_pts, n, radius, max_connect, d_edges, edge_cap, nedges, &cand, &cand_cap, count, off);
    int32_t ne = 0;
    if (rc == LG_OK) {
        if (hipMemcpyAsync(&ne, nedges, 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
            set_error("lg_build_edges: readback failed");
            rc = LG_ERR_HIP;
        }
    }
    dfree(ctx, count);
    dfree(ctx, off);
    dfree(ctx, nedges);
    dfree(ctx, cand);
    if (rc != LG_OK) return rc;
    if (ne < 0) {
        *n_edges_out = -ne;
        set_error("lg_build_edges: %d edges exceed edge_cap %d", -ne, edge_cap);
        return LG_ERR_OVERFLOW;
    }
    *n_edges_out = ne;
    return LG_OK;
}

int32_t lg_march_edges(lg_ctx *ctx, const lg_grid *grid, const lg_shader *shader, const lg_march_params *params,
                       const float *d_pts, const int32_t *d_edges, int32_t n_edges, float *d_w, float *d_D,
                       int32_t ld) {
    LG_CHECK_ARG(ctx && grid && d_pts && (d_edges || n_edges == 0), "lg_march_edges: NULL argument");
    LG_TRY(check_params(shader, params));
    LG_CHECK_ARG(!d_D || ld > 0, "lg_march_edges: ld must be positive");
    return launch_march_edges(ctx, grid, shader, params, d_pts, d_edges, n_edges, nullptr, d_w, d_D, ld);
}

int32_t lg_vertex_light(lg_ctx *ctx, const lg_grid *density, const lg_grid *temperature, const lg_shader *shader,
                        const lg_light *lights, int32_t n_lights, const lg_march_params *params, uint64_t env_seed,
                        uint32_t solve, const float *d_pts, int32_t n, float *d_L_out) {
    LG_CHECK_ARG(ctx && density && d_pts && d_L_out, "lg_vertex_light: NULL argument");
    LG_TRY(check_params(shader, params));
    return launch_vertex_light(ctx, density, temperature, shader, lights, n_lights, params, env_seed, solve, d_pts, n,
                               d_L_out);
}

int32_t lg_fw_ld(int32_t n) { return n <= 0 ? 0 : fw_ld(n); }

int32_t lg_fw_init(lg_ctx *ctx, int32_t n, float *d_D, int32_t ld) {
    LG_CHECK_ARG(ctx && d_D && n >= 0 && ld >= n, "lg_fw_init: bad argument");
    return fw_init_impl(ctx, n, d_D, ld);
}

int32_t lg_diffuse(lg_ctx *ctx, int32_t n, const float *d_D, int32_t ld, const double sigma_t[3], const float *d_L,
                   float *d_out) {
    LG_CHECK_ARG(ctx && d_D && sigma_t && d_L && d_out && ld >= n, "lg_diffuse: bad argument");
    if (n == 0) return LG_OK;
    double *sums = nullptr;
    float *ln = nullptr;
    LG_TRY(dalloc(ctx, (size_t)n * 3, &sums));
    int32_t rc = dalloc(ctx, (size_t)n * 3, &ln);
    if (rc == LG_OK) rc = diffuse_impl(ctx, n, d_D, ld, sigma_t, d_L, d_out, sums, ln);
    dfree(ctx, sums);
    dfree(ctx, ln);
    return rc;
}

int32_t lg_solver_create(lg_ctx *ctx, int32_t n_max, int32_t edge_cap, int32_t with_pred, lg_solver **out) {
    LG_CHECK_ARG(ctx && out && n_max >= 1 && n_max <= 65536 && edge_cap >= 0, "lg_solver_create: bad argument");
    *out = nullptr;
    lg_solver *s = new lg_solver();
    s->n_max = n_max;
    s->ld_max = fw_ld(n_max);
    s->edge_cap = edge_cap ? edge_cap : n_max * 6 + 64;
    s->with_pred = with_pred != 0;
    int32_t rc = LG_OK;
    auto A = [&](auto **p, size_t count) {
        if (rc == LG_OK) rc = dalloc(ctx, count, p);
    };
    A(&s->buf.d_points, (size_t)n_max * 3);
    A(&s->buf.d_edges, (size_t)s->edge_cap * 2);
    A(&s->buf.d_weights, (size_t)s->edge_cap);
    A(&s->buf.d_D, (size_t)s->ld_max * s->ld_max);
    A(&s->buf.d_L, (size_t)n_max * 3);
    A(&s->buf.d_diffuse, (size_t)n_max * 3);
    if (s->with_pred) A(&s->buf.d_pred, (size_t)s->ld_max * s->ld_max);
    A(&s->cand_count, (size_t)n_max + 1);
    A(&s->cand_off, (size_t)n_max + 1);
    A(&s->d_nedges, 1);
    A(&s->d_colsum, (size_t)n_max * 3);
    A(&s->d_Lnorm, (size_t)n_max * 3);
    for (auto &e : s->ev)
        if (rc == LG_OK && hipEventCreate(&e) != hipSuccess) {
            set_error("hipEventCreate failed");
            rc = LG_ERR_HIP;
        }
    if (rc != LG_OK) {
        lg_solver_destroy(ctx, s);
        return rc;
    }
    *out = s;
    return LG_OK;
}

int32_t lg_solver_destroy(lg_ctx *ctx, lg_solver *s) {
    if (!s) return LG_OK;
    dfree(ctx, s->buf.d_points);
    dfree(ctx, s->buf.d_edges);
    dfree(ctx, s->buf.d_weights);
    dfree(ctx, s->buf.d_D);
    dfree(ctx, s->buf.d_L);
    dfree(ctx, s->buf.d_diffuse);
    dfree(ctx, s->buf.d_pred);
    dfree(ctx, s->cand_count);
    dfree(ctx, s->cand_off);
    dfree(ctx, s->cand);
    dfree(ctx, s->d_nedges);
    dfree(ctx, s->d_colsum);
    dfree(ctx, s->d_Lnorm);
    for (auto &e : s->ev)
        if (e) (void)hipEventDestroy(e);
    delete s;
    return LG_OK;
}

int32_t lg_solver_get(const lg_solver *s, lg_solver_buffers *out) {
    LG_CHECK_ARG(s && out, "lg_solver_get: NULL argument");
    *out = s->buf;
    return LG_OK;
}

int32_t lg_solver_timing(const lg_solver *s, lg_solve_timing *out) {
    LG_CHECK_ARG(s && out, "lg_solver_timing: NULL argument");
    lg_solve_timing t{};
    float ms[7] = {0};
    for (int k = 0; k < 6; ++k) {
        if (hipEventSynchronize(s->ev[k + 1]) != hipSuccess ||
            hipEventElapsedTime(&ms[k], s->ev[k], s->ev[k + 1]) != hipSuccess) {
            set_error("lg_solver_timing: no timed solve recorded");
            return LG_ERR_INVALID;
        }
    }
    (void)hipEventElapsedTime(&ms[6], s->ev[0], s->ev[6]);
    t.sample_ms = ms[0];
    t.graph_ms = ms[1];
    t.march_ms = ms[2];
    t.fw_ms = ms[3];
    t.light_ms = ms[4];
    t.diffuse_ms = ms[5];
    t.total_ms = ms[6];
    *out = t;
    return LG_OK;
}

int32_t lg_solve(lg_ctx *ctx, lg_solver *s, const lg_grid *density, const lg_grid *temperature,
                 const lg_shader *shader, const lg_light *lights, int32_t n_lights, const lg_march_params *params,
                 const lg_solve_params *sp) {
    LG_CHECK_ARG(ctx && s && density && sp, "lg_solve: NULL argument");
    LG_TRY(check_params(shader, params));
    const int n = sp->n;
    LG_CHECK_ARG(n >= 1 && n <= s->n_max, "lg_solve: n=%d outside [1, %d]", n, s->n_max);
    const int ld = fw_ld(n);
    const bool t = sp->timing != 0;
    s->buf.n = n;
    s->buf.ld = ld;
    s->buf.edge_cap = s->edge_cap;
    if (t) LG_HIP(hipEventRecord(s->ev[0], ctx->stream));
    if (!sp->external_points)
        LG_TRY(launch_sample(ctx, const_cast<lg_grid *>(density), sp->seed, sp->solve, n, s->buf.d_points));
    if (t) LG_HIP(hipEventRecord(s->ev[1], ctx->stream));
    LG_TRY(build_edges_impl(ctx, s->buf.d_points, n, sp->radius, sp->max_connect, s->buf.d_edges, s->edge_cap,
                            s->d_nedges, &s->cand, &s->cand_cap, s->cand_count, s->cand_off));
    if (t) LG_HIP(hipEventRecord(s->ev[2], ctx->stream));
    LG_TRY(fw_init_impl(ctx, n, s->buf.d_D, ld));
    LG_TRY(launch_march_edges(ctx, density, shader, params, s->buf.d_points, s->buf.d_edges, s->edge_cap, s->d_nedges,
                              s->buf.d_weights, s->buf.d_D, ld));
    if (t) LG_HIP(hipEventRecord(s->ev[3], ctx->stream));
    LG_TRY(fw_solve_impl(ctx, n, s->buf.d_D, ld));
    if (t) LG_HIP(hipEventRecord(s->ev[4], ctx->stream));
    LG_TRY(launch_vertex_light(ctx, density, temperature, shader, lights, n_lights, params, sp->env_seed, sp->solve,
                               s->buf.d_points, n, s->buf.d_L));
    if (t) LG_HIP(hipEventRecord(s->ev[5], ctx->stream));
    const double sigma[3] = {-shader->extinction[0], -shader->extinction[1], -shader->extinction[2]};
    LG_TRY(diffuse_impl(ctx, n, s->buf.d_D, ld, sigma, s->buf.d_L, s->buf.d_diffuse, s->d_colsum, s->d_Lnorm));
    if (t) LG_HIP(hipEventRecord(s->ev[6], ctx->stream));
    int32_t ne = 0;
    LG_HIP(hipMemcpyAsync(&ne, s->d_nedges, 4, hipMemcpyDeviceToHost, ctx->stream));
    LG_HIP(hipStreamSynchronize(ctx->stream));
    if (ne < 0) {
        set_error("lg_solve: %d edges exceed the solver's edge capacity %d", -ne, s->edge_cap);
        return LG_ERR_OVERFLOW;
    }
    s->buf.n_edges = ne;
    if (s->with_pred)
        LG_TRY(fw_pred_impl(ctx, n, s->buf.d_D, ld, s->buf.d_pred, s->buf.d_edges, s->buf.d_weights, ne));
    return LG_OK;
}

}  // extern "C"
