// C ABI of the solve stages and of one full LightGraph solve
// (LightGraph::operator(), LightGraph.h:158-196), over the pipeline in
// lg_pipeline.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "lg_internal.h"
#include "lg_lights.h"
#include "lg_pipeline.h"
#include <functional>

// Scratch of lg_diffuse_noalloc: the column sums (3 doubles per vertex) then the
// normalised light (3 floats per vertex) of up to LG_DIFFUSE_NOALLOC_MAX_N
// vertices, one copy per GPU for the life of the process.  Host code never
// passes `scratch` itself to a kernel: its device address comes from
// hipGetSymbolAddress, once per device.
static constexpr size_t kNoAllocSums = (size_t)LG_DIFFUSE_NOALLOC_MAX_N * 3 * sizeof(double);
static constexpr size_t kNoAllocBytes = kNoAllocSums + (size_t)LG_DIFFUSE_NOALLOC_MAX_N * 3 * sizeof(float);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

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
    if (!(f(0.0) >= cutoff)) return -1.0;       // the first sample already drops
    int zero = 0;
    for (int c = 0; c < 3; ++c) zero += ext[c] == 0.0;
    if ((double)zero >= cutoff) return INFINITY;
    double lo = 0.0, hi = 1.0;
    while (f(hi) >= cutoff) {
        lo = hi;
        hi *= 2.0;
        if (hi > 1e300) return INFINITY;
    }
    for (int it = 0; it < 2000; ++it) {  // largest tau with f(tau) >= cutoff
        const double mid = lo + 0.5 * (hi - lo);
        if (mid <= lo || mid >= hi) break;
        if (f(mid) >= cutoff)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// ------------------------------------------------------ sampling tables -----
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

SolveDesc make_desc(lg_grid *density, const lg_grid *temperature, int n, int voff, uint32_t solve) {
    SolveDesc d{};
    if (density) {
        d.g = make_view(density);
        d.active_list = density->active_list;
        d.mass_prefix = density->mass_prefix;
        d.Q = density->total_mass;
        d.nact = density->active_bricks;
        d.dtype = density->desc.dtype;
    }
    if (temperature) {
        d.tg = make_view(temperature);
        d.has_temp = 1;
    }
    d.n = n;
    d.voff = voff;
    d.solve = solve;
    return d;
}

// ------------------------------------------------------- march constants ----
bool step_is_exact(double h) {
    if (!(h > 0.0) || !std::isfinite(h)) return false;
    const double m = ldexp(h, 30);
    return m == floor(m) && h < 8388608.0;
}

MarchConsts march_consts(const lg_shader *sh, const lg_march_params *mp) {
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

LightConsts light_consts(const lg_shader *sh, const lg_march_params *mp) {
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
    lc.exact_steps = step_is_exact(lc.h) ? 1 : 0;
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

// ------------------------------------------------- single-stage kernels -----
template <typename T>
__global__ void k_march_list(lgd::GridView g, MarchConsts mc, const float *__restrict__ pts,
                             const int32_t *__restrict__ edges, int ne, float *__restrict__ w, float *__restrict__ D,
                             int ld) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int i = edges[2 * e], j = edges[2 * e + 1];
    const float tau = march_edge_one<T>(g, mc, pts + 3 * (long long)i, pts + 3 * (long long)j);
    if (w) w[e] = tau;
    if (D) {
        D[(long long)i * ld + j] = tau;
        D[(long long)j * ld + i] = tau;
    }
}

// ------------------------------------------------------------ D init --------
__global__ void k_fw_init(int ld, float *__restrict__ D) {
    const long long total = (long long)ld * ld;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(k / ld), j = (int)(k % ld);
        D[k] = (i == j) ? 0.0f : INFINITY;
    }
}

int32_t fw_init_impl(lg_ctx *ctx, int n, float *d_D, int ld) {
    (void)n;
    k_fw_init<<<(unsigned)std::min<long long>(((long long)ld * ld + 255) / 256, 65536), 256, 0, ctx->stream>>>(ld, d_D);
    LG_LAUNCHED("k_fw_init");
    return LG_OK;
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
        a0 += trans_f(e0, d);
        a1 += trans_f(e1, d);
        a2 += trans_f(e2, d);
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
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const float d = row[j];
        a0 += (double)(Ln[3 * j + 0] * trans_f(e0, d));
        a1 += (double)(Ln[3 * j + 1] * trans_f(e1, d));
        a2 += (double)(Ln[3 * j + 2] * trans_f(e2, d));
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

int32_t diffuse_impl(lg_ctx *ctx, int n, const float *d_D, int ld, const double sigma_t[3], const float *d_L,
                     float *d_out, double *d_sums, float *d_Ln, bool timed) {
    if (n <= 0) return LG_OK;
    const float e0 = (float)(-sigma_t[0]), e1 = (float)(-sigma_t[1]), e2 = (float)(-sigma_t[2]);
    hipEvent_t ev = nullptr;
    if (timed) ktimer_begin(ctx, "k_diffuse", &ev);
    k_rowsum<<<n, 256, 0, ctx->stream>>>(d_D, n, ld, e0, e1, e2, d_sums);
    LG_LAUNCHED("k_rowsum");
    k_lnorm<<<(n + 255) / 256, 256, 0, ctx->stream>>>(d_L, d_sums, n, d_Ln);
    LG_LAUNCHED("k_lnorm");
    k_diffuse_out<<<n, 256, 0, ctx->stream>>>(d_D, n, ld, e0, e1, e2, d_Ln, d_out);
    LG_LAUNCHED("k_diffuse_out");
    if (timed) ktimer_end(ctx, "k_diffuse", ev);
    return LG_OK;
}

static constexpr int kNoAllocDevices = 64;
static std::mutex g_noalloc_mu;
static void *g_noalloc_ptr[kNoAllocDevices];

// Device address of `scratch` on the context's device (looked up on the first
// call on that device, with the device current).
static int32_t noalloc_scratch(lg_ctx *ctx, unsigned char **out) {
    LG_CHECK_ARG(ctx->device >= 0 && ctx->device < kNoAllocDevices, "device %d outside the scratch table",
                 ctx->device);
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[ctx->device]) {
        int prev = 0;
        LG_HIP(hipGetDevice(&prev));
        LG_HIP(hipSetDevice(ctx->device));
        void *ptr = nullptr;
        const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(::scratch));
        (void)hipSetDevice(prev);
        if (err != hipSuccess || !ptr) {
            set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", ctx->device, hipGetErrorString(err));
            return LG_ERR_HIP;
        }
        g_noalloc_ptr[ctx->device] = ptr;
    }
    *out = static_cast<unsigned char *>(g_noalloc_ptr[ctx->device]);
    return LG_OK;
}

// Runs the pipeline over the solves in descs (vertex offsets voff, one shared
// shader / light set / march parameters) into the solver's buffers.
// With LG_SOLVE_STATS, `grids` are the call's distinct density grids, whose
// per-brick flags the descriptors' views point at: the light and the edge march
// run one after the other and each counts the distinct bricks its taps read.
static int32_t solve_descs(lg_ctx *ctx, lg_solver *s, const std::vector<SolveDesc> &descs, const lg_shader *shader,
                           const lg_light *lights, int32_t n_lights, const lg_march_params *params,
                           const lg_solve_params *sp, const std::vector<lg_grid *> &grids) {
    Pipeline &pl = *static_cast<Pipeline *>(s->pipeline);
    PipelineParams pp{};
    pp.pts = s->buf.d_points;
    pp.L = s->buf.d_L;
    pp.out = s->buf.d_diffuse;
    pp.seed = sp->seed;
    pp.env_seed = sp->env_seed;
    pp.external_points = sp->external_points;
    pp.max_connect = sp->max_connect;
    const float rf = (float)sp->radius;
    const float r2f = rf * rf;  // float, as searchRadiusConnectSquared (LightGraph.h:171)
    pp.r2 = (double)r2f;
    for (int c = 0; c < 3; ++c) pp.ext[c] = shader->extinction[c];
    pp.mc = march_consts(shader, params);
    pp.lc = light_consts(shader, params);
    LG_TRY(lgl::make_dev_lights(lights, n_lights, &pp.lights));
    const bool t = sp->timing != 0;
    const int dtype = descs[0].dtype;
    pl.stats_on = (sp->flags & LG_SOLVE_STATS) != 0;
    LG_CHECK_ARG((sp->flags & (LG_SOLVE_BFS_ORDER | LG_SOLVE_ND_ORDER)) != (LG_SOLVE_BFS_ORDER | LG_SOLVE_ND_ORDER),
                 "lg_solve: LG_SOLVE_BFS_ORDER and LG_SOLVE_ND_ORDER exclude each other");
    pl.bfs_order = (sp->flags & LG_SOLVE_BFS_ORDER) != 0;
    pl.nd_order = (sp->flags & LG_SOLVE_ND_ORDER) != 0;
    if (pl.stats_on) {
        if (!pl.stats) LG_TRY(dalloc(ctx, (size_t)ST_COUNT * kStatSlots, &pl.stats));
        LG_HIP(hipMemsetAsync(pl.stats, 0, sizeof(unsigned long long) * ST_COUNT * kStatSlots, ctx->stream));
    }
    int V = 0, nmax = 0, Ktot = 0;
    if (t) LG_HIP(hipEventRecord(s->ev[0], ctx->stream));
    LG_TRY(pl_begin(ctx, pl, descs, &V, &nmax));
    if (!pp.external_points) LG_TRY(pl_sample(ctx, pl, V, dtype, pp.seed, pp.pts));
    if (t) LG_HIP(hipEventRecord(s->ev[1], ctx->stream));
    // The vertex light needs only the points, and its first reader is the
    // component stage: it runs on the side stream beside the graph, edge march
    // and component sort, forked once the neighbour grid is queued (the grid's
    // one-workgroup sort of a small call runs alone, before the light's waves
    // take every CU).  With per-stage timing on, the stages run in sequence so
    // that each interval times one stage.
    const bool overlap = !t && !pl.stats_on;
    // every return after the fork leaves the main stream ordered after the light
    struct JoinGuard {
        lg_ctx *ctx;
        hipEvent_t ev;
        ~JoinGuard() {
            if (ev) (void)hipStreamWaitEvent(ctx->stream, ev, 0);
        }
    } join{ctx, nullptr};
    std::function<int32_t()> fork_light;
    if (overlap) {
        fork_light = [&]() -> int32_t {
            LG_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
            LG_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
            std::swap(ctx->stream, ctx->side);
            const int32_t rc = pl_light(ctx, pl, V, dtype, pp.lc, pp.lights, pp.env_seed, pp.pts, pp.L);
            std::swap(ctx->stream, ctx->side);
            LG_HIP(hipEventRecord(ctx->ev_light, ctx->side));
            join.ev = ctx->ev_light;
            return rc;
        };
    } else {
        LG_TRY(pl_light(ctx, pl, V, dtype, pp.lc, pp.lights, pp.env_seed, pp.pts, pp.L));
        if (pl.stats_on)
            for (lg_grid *g : grids) LG_TRY(grid_touch_count(ctx, g, pl.stats, ST_LIGHT_BRICKS));
    }
    if (t) LG_HIP(hipEventRecord(s->ev[2], ctx->stream));
    LG_TRY(pl_graph(ctx, pl, V, (int)descs.size(), nmax, pp.pts, pp.r2, pp.max_connect, &Ktot,
                    overlap ? &fork_light : nullptr));
    if (overlap && !join.ev) LG_TRY(fork_light());  // a graph stage that queued no grid
    if (t) LG_HIP(hipEventRecord(s->ev[3], ctx->stream));
    LG_TRY(pl_march(ctx, pl, dtype, pp.mc, pp.pts, Ktot));
    if (pl.stats_on)
        for (lg_grid *g : grids) LG_TRY(grid_touch_count(ctx, g, pl.stats, ST_EDGE_BRICKS));
    if (t) LG_HIP(hipEventRecord(s->ev[4], ctx->stream));
    PipelineResult res;
    LG_TRY(pl_components(ctx, pl, descs, V, Ktot, pp.ext, pp.L, pp.out, &res, join.ev, pp.pts));
    if (t) LG_HIP(hipEventRecord(s->ev[5], ctx->stream));
    s->timed = t;
    s->buf.n = V;
    s->buf.ld = descs.size() == 1 ? descs[0].ld : 0;
    s->buf.n_edges = res.n_edges_total;
    s->buf.n_components = res.n_components;
    s->n_big = res.n_big;
    s->stats_valid = pl.stats_on;
    s->buf.d_edges = res.edges;
    s->buf.d_weights = res.w;
    return LG_OK;
}

Pipeline &stage_pipeline(lg_ctx *ctx) {
    if (!ctx->stage_pipeline) ctx->stage_pipeline = new Pipeline();
    return *static_cast<Pipeline *>(ctx->stage_pipeline);
}

template <typename F>
static void for_dtype(int dtype, F &&f) {
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

int32_t lg_sample_points(lg_ctx *ctx, const lg_grid *grid, uint64_t seed, uint32_t solve, int32_t n, float *d_pts) {
    LG_CHECK_ARG(ctx && grid && d_pts, "lg_sample_points: NULL argument");
    LG_CHECK_ARG(n >= 0 && n <= 65536, "lg_sample_points: n=%d outside [0, 65536]", n);
    if (n == 0) return LG_OK;
    lg_grid *g = const_cast<lg_grid *>(grid);
    LG_TRY(grid_build_sampling_tables(ctx, g));
    Pipeline &pl = stage_pipeline(ctx);
    std::vector<SolveDesc> descs = {make_desc(g, nullptr, n, 0, solve)};
    int V = 0;
    LG_TRY(pl_begin(ctx, pl, descs, &V, nullptr));
    return pl_sample(ctx, pl, V, g->desc.dtype, seed, d_pts);
}

int32_t lg_build_edges(lg_ctx *ctx, const float *d_pts, int32_t n, double radius, int32_t max_connect,
                       int32_t *d_edges, int32_t edge_cap, int32_t *n_edges_out) {
    LG_CHECK_ARG(ctx && d_pts && d_edges && n_edges_out, "lg_build_edges: NULL argument");
    LG_CHECK_ARG(edge_cap >= 0 && n >= 1 && n <= 65536, "lg_build_edges: bad n or edge_cap");
    Pipeline &pl = stage_pipeline(ctx);
    std::vector<SolveDesc> descs = {make_desc(nullptr, nullptr, n, 0, 0)};
    int V = 0, nmax = 0, Ktot = 0;
    LG_TRY(pl_begin(ctx, pl, descs, &V, &nmax));
    const float rf = (float)radius;
    const float r2f = rf * rf;  // float, as searchRadiusConnectSquared (LightGraph.h:171)
    LG_TRY(pl_graph(ctx, pl, V, 1, nmax, d_pts, (double)r2f, max_connect, &Ktot));
    int32_t ne = 0, gerr = 0;
    LG_HIP(hipMemcpyAsync(&ne, pl.epos + Ktot, 4, hipMemcpyDeviceToHost, ctx->stream));
    LG_HIP(hipMemcpyAsync(&gerr, pl.greedy_err, 4, hipMemcpyDeviceToHost, ctx->stream));
    LG_HIP(hipStreamSynchronize(ctx->stream));
    if (gerr != 0) {
        set_error("lg_build_edges: a row of the parallel greedy cap waited past its bound");
        return LG_ERR_HIP;
    }
    *n_edges_out = ne;
    if (ne > edge_cap) {
        set_error("lg_build_edges: %d edges exceed edge_cap %d", ne, edge_cap);
        return LG_ERR_OVERFLOW;
    }
    if (ne > 0)
        LG_HIP(hipMemcpyAsync(d_edges, pl.edges_, (size_t)ne * 8, hipMemcpyDeviceToDevice, ctx->stream));
    return LG_OK;
}

int32_t lg_march_edges(lg_ctx *ctx, const lg_grid *grid, const lg_shader *shader, const lg_march_params *params,
                       const float *d_pts, const int32_t *d_edges, int32_t n_edges, float *d_w, float *d_D,
                       int32_t ld) {
    LG_CHECK_ARG(ctx && grid && d_pts && (d_edges || n_edges == 0), "lg_march_edges: NULL argument");
    LG_TRY(check_params(shader, params));
    LG_CHECK_ARG(!d_D || ld > 0, "lg_march_edges: ld must be positive");
    if (n_edges <= 0) return LG_OK;
    const lgd::GridView g = make_view(grid);
    const MarchConsts mc = march_consts(shader, params);
    hipEvent_t ev;
    ktimer_begin(ctx, "k_march_edges", &ev);
    for_dtype(grid->desc.dtype, [&](auto tag) {
        using T = decltype(tag);
        k_march_list<T><<<(unsigned)((n_edges + 127) / 128), 128, 0, ctx->stream>>>(g, mc, d_pts, d_edges, n_edges,
                                                                                     d_w, d_D, ld);
    });
    LG_LAUNCHED("k_march_list");
    ktimer_end(ctx, "k_march_edges", ev);
    return LG_OK;
}

int32_t lg_vertex_light(lg_ctx *ctx, const lg_grid *density, const lg_grid *temperature, const lg_shader *shader,
                        const lg_light *lights, int32_t n_lights, const lg_march_params *params, uint64_t env_seed,
                        uint32_t solve, const float *d_pts, int32_t n, float *d_L_out) {
    LG_CHECK_ARG(ctx && density && d_pts && d_L_out, "lg_vertex_light: NULL argument");
    LG_TRY(check_params(shader, params));
    if (n <= 0) return LG_OK;
    LG_CHECK_ARG(n <= 65536, "lg_vertex_light: n=%d above 65536", n);
    lgl::DevLights L{};
    LG_TRY(lgl::make_dev_lights(lights, n_lights, &L));
    LG_CHECK_ARG(!temperature || temperature->desc.dtype == LG_F32, "temperature grid must be float32");
    Pipeline &pl = stage_pipeline(ctx);
    int V = 0;
    LG_TRY(pl_begin(ctx, pl, {make_desc(const_cast<lg_grid *>(density), temperature, n, 0, solve)}, &V, nullptr));
    return pl_light(ctx, pl, V, density->desc.dtype, light_consts(shader, params), L, env_seed, d_pts, d_L_out);
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
    if (rc == LG_OK) rc = diffuse_impl(ctx, n, d_D, ld, sigma_t, d_L, d_out, sums, ln, true);
    dfree(ctx, sums);
    dfree(ctx, ln);
    return rc;
}

int32_t lg_diffuse_noalloc(lg_ctx *ctx, int32_t n, const float *d_D, int32_t ld, const double sigma_t[3],
                           const float *d_L, float *d_out) {
    LG_CHECK_ARG(ctx && d_D && sigma_t && d_L && d_out && ld >= n && n >= 0, "lg_diffuse_noalloc: bad argument");
    LG_CHECK_ARG(n <= LG_DIFFUSE_NOALLOC_MAX_N, "lg_diffuse_noalloc: n=%d above the scratch limit of %d vertices",
                 n, LG_DIFFUSE_NOALLOC_MAX_N);
    if (n == 0) return LG_OK;
    unsigned char *base = nullptr;
    LG_TRY(noalloc_scratch(ctx, &base));
    // k_rowsum writes every sum it reads back and k_lnorm every normalised
    // entry, so nothing left by an earlier call is read
    return diffuse_impl(ctx, n, d_D, ld, sigma_t, d_L, d_out, reinterpret_cast<double *>(base),
                        reinterpret_cast<float *>(base + kNoAllocSums), false);
}

int32_t lg_solver_create(lg_ctx *ctx, int32_t n_max, int32_t edge_cap, int32_t flags, lg_solver **out) {
    LG_CHECK_ARG(ctx && out && n_max >= 1 && edge_cap >= 0, "lg_solver_create: bad argument");
    LG_CHECK_ARG(n_max <= (1 << 24), "lg_solver_create: n_max %d above 2^24", n_max);
    LG_CHECK_ARG(!(flags & (LG_SOLVER_DENSE_D | LG_SOLVER_PRED)) || n_max <= 65536,
                 "lg_solver_create: dense outputs need n_max <= 65536");
    *out = nullptr;
    lg_solver *s = new lg_solver();
    s->n_max = n_max;
    s->ld_max = fw_ld(n_max);
    s->with_pred = (flags & LG_SOLVER_PRED) != 0;
    s->with_dense = (flags & (LG_SOLVER_DENSE_D | LG_SOLVER_PRED)) != 0;
    s->pipeline = new Pipeline();
    int32_t rc = LG_OK;
    auto A = [&](auto **p, size_t count) {
        if (rc == LG_OK) rc = dalloc(ctx, count, p);
    };
    A(&s->buf.d_points, (size_t)n_max * 3);
    A(&s->buf.d_L, (size_t)n_max * 3);
    A(&s->buf.d_diffuse, (size_t)n_max * 3);
    if (s->with_dense) A(&s->buf.d_D, (size_t)s->ld_max * s->ld_max);
    if (s->with_pred) A(&s->buf.d_pred, (size_t)s->ld_max * s->ld_max);
    for (auto &e : s->ev)
        if (rc == LG_OK && hipEventCreate(&e) != hipSuccess) {
            set_error("hipEventCreate failed");
            rc = LG_ERR_HIP;
        }
    if (rc == LG_OK) rc = static_cast<Pipeline *>(s->pipeline)->reserve_vertices(ctx, n_max, 1);
    if (rc != LG_OK) {
        lg_solver_destroy(ctx, s);
        return rc;
    }
    *out = s;
    return LG_OK;
}

int32_t lg_solver_stats(lg_ctx *ctx, const lg_solver *s, lg_solve_stats *out) {
    LG_CHECK_ARG(ctx && s && out, "lg_solver_stats: NULL argument");
    std::memset(out, 0, sizeof(*out));
    const Pipeline &pl = *static_cast<const Pipeline *>(s->pipeline);
    out->points = (uint64_t)s->buf.n;
    out->edges = (uint64_t)s->buf.n_edges;
    out->components = (uint64_t)s->buf.n_components;
    out->big_components = (uint64_t)s->n_big;
    if (!s->stats_valid || !pl.stats) return LG_OK;
    unsigned long long slots[ST_COUNT * kStatSlots], h[ST_COUNT] = {};
    LG_HIP(hipMemcpyAsync(slots, pl.stats, sizeof(slots), hipMemcpyDeviceToHost, ctx->stream));
    LG_HIP(hipStreamSynchronize(ctx->stream));
    for (int c = 0; c < ST_COUNT; ++c)
        for (int k = 0; k < kStatSlots; ++k) h[c] += slots[c * kStatSlots + k];
    out->edge_samples = h[ST_EDGE_SAMPLES];
    out->shadow_rays = h[ST_SHADOW_RAYS];
    out->shadow_samples = h[ST_SHADOW_SAMPLES];
    out->fw_updates = h[ST_FW_TILES];
    out->fw_panel_updates = h[ST_FW_PANEL_TILES];
    out->fw_pivot_updates = h[ST_FW_PIVOTS];
    out->light_bricks = h[ST_LIGHT_BRICKS];
    out->edge_bricks = h[ST_EDGE_BRICKS];
    out->small_updates = h[ST_SMALL_UPDATES];
    out->small_updates_64 = h[ST_SMALL_UPDATES_64];
    out->small_updates_32 = h[ST_SMALL_UPDATES_32];
    return LG_OK;
}

int32_t lg_solver_destroy(lg_ctx *ctx, lg_solver *s) {
    if (!s) return LG_OK;
    dfree(ctx, s->buf.d_points);
    dfree(ctx, s->buf.d_L);
    dfree(ctx, s->buf.d_diffuse);
    dfree(ctx, s->buf.d_D);
    dfree(ctx, s->buf.d_pred);
    if (s->pipeline) {
        static_cast<Pipeline *>(s->pipeline)->release(ctx);
        delete static_cast<Pipeline *>(s->pipeline);
    }
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
    LG_CHECK_ARG(s->timed, "lg_solver_timing: no timed solve recorded");
    float ms[6] = {0};
    for (int k = 0; k < 5; ++k) {
        LG_HIP(hipEventSynchronize(s->ev[k + 1]));
        LG_HIP(hipEventElapsedTime(&ms[k], s->ev[k], s->ev[k + 1]));
    }
    LG_HIP(hipEventElapsedTime(&ms[5], s->ev[0], s->ev[5]));
    lg_solve_timing t{};
    t.sample_ms = ms[0];
    t.light_ms = ms[1];
    t.graph_ms = ms[2];
    t.march_ms = ms[3];
    t.fw_ms = ms[4];  // components: closure + diffusion
    t.diffuse_ms = 0.0f;
    t.total_ms = ms[5];
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
    LG_CHECK_ARG(!temperature || temperature->desc.dtype == LG_F32, "temperature grid must be float32");
    const bool dense = (sp->flags & LG_SOLVE_DENSE) != 0;
    LG_CHECK_ARG(!dense || s->with_dense, "lg_solve: dense outputs need a solver created with LG_SOLVER_DENSE_D");
    LG_CHECK_ARG(!dense || fw_ld(n) <= s->ld_max, "lg_solve: dense outputs larger than the solver's");
    lg_grid *g = const_cast<lg_grid *>(density);
    LG_TRY(grid_build_sampling_tables(ctx, g));
    SolveDesc d = make_desc(g, temperature, n, 0, sp->solve);
    d.ld = fw_ld(n);
    d.dense_D = dense ? s->buf.d_D : nullptr;
    d.dense_pred = (dense && s->with_pred) ? s->buf.d_pred : nullptr;
    std::vector<lg_grid *> grids;
    if (sp->flags & LG_SOLVE_STATS) {
        LG_TRY(grid_touch_begin(ctx, g));
        d.g.touched = g->touched;
        grids.push_back(g);
    }
    return solve_descs(ctx, s, {d}, shader, lights, n_lights, params, sp, grids);
}

int32_t lg_solve_batch(lg_ctx *ctx, lg_solver *s, const lg_batch_item *items, int32_t count, const lg_shader *shader,
                       const lg_light *lights, int32_t n_lights, const lg_march_params *params,
                       const lg_solve_params *sp) {
    LG_CHECK_ARG(ctx && s && items && sp && count >= 1, "lg_solve_batch: bad argument");
    LG_TRY(check_params(shader, params));
    LG_CHECK_ARG((sp->flags & LG_SOLVE_DENSE) == 0, "lg_solve_batch: dense outputs are per-solve only (use lg_solve)");
    std::vector<SolveDesc> descs;
    descs.reserve((size_t)count);
    std::vector<lg_grid *> grids;
    long long V = 0;
    for (int k = 0; k < count; ++k) {
        const lg_batch_item &it = items[k];
        LG_CHECK_ARG(it.density && it.n >= 1, "lg_solve_batch: item %d: NULL density or n < 1", k);
        LG_CHECK_ARG(!it.temperature || it.temperature->desc.dtype == LG_F32,
                     "lg_solve_batch: item %d: temperature grid must be float32", k);
        LG_CHECK_ARG(it.density->desc.dtype == items[0].density->desc.dtype,
                     "lg_solve_batch: item %d: every density grid of a batch needs the same element type", k);
        lg_grid *g = const_cast<lg_grid *>(it.density);
        LG_TRY(grid_build_sampling_tables(ctx, g));
        SolveDesc d = make_desc(g, it.temperature, it.n, (int)V, it.solve);
        d.ld = fw_ld(it.n);
        if (sp->flags & LG_SOLVE_STATS) {
            if (std::find(grids.begin(), grids.end(), g) == grids.end()) {
                LG_TRY(grid_touch_begin(ctx, g));
                grids.push_back(g);
            }
            d.g.touched = g->touched;
        }
        descs.push_back(d);
        V += it.n;
        LG_CHECK_ARG(V <= s->n_max, "lg_solve_batch: %lld points exceed the solver's n_max %d", V, s->n_max);
    }
    return solve_descs(ctx, s, descs, shader, lights, n_lights, params, sp, grids);
}

}  // extern "C"
