// Shared declarations of the solve pipeline: per-solve descriptors, march and
// lighting device functions, and the reusable device workspace.
#pragma once

#include <functional>

#include <vector>

#include "lg_internal.h"
#include "lg_lights.h"

namespace lg {

constexpr int kSmallComp = 128;  // components up to this size are closed in LDS

// Per-solve descriptor (device array, one entry per solve of a batch).
struct SolveDesc {
    lgd::GridView g;              // density
    lgd::GridView tg;             // temperature (valid if has_temp)
    const int32_t *active_list;   // sampling tables of the density grid
    const uint64_t *mass_prefix;
    uint64_t Q;
    int nact;
    int n, voff;                  // vertices of this solve: [voff, voff + n)
    uint32_t solve;               // the reference's LightGraph iteration index
    int has_temp;
    int dtype;
    int ld;                       // leading dimension of the optional dense outputs
    float *dense_D;               // optional n x ld optical depth (solve-local indices)
    int32_t *dense_pred;          // optional n x ld canonical predecessors
    int cell_A;                   // neighbour grid: at most cell_A + 1 cells per axis
    int cell_base;                // first cell id of this solve in the batch's cell space
};

// Uniform neighbour grid of one solve (cells at least the edge radius wide).
struct CellDesc {
    double lo[3];
    double inv[3];
    int n[3];
    int base;
};

struct MarchConsts {
    double ds_sds_scds[3];  // densityScale, shadowDensityScale, scatterDensityScale
    double dmax, sdm, cutoff, h, tau_star;
};

struct LightConsts {
    double ds_sds_scds[3];
    double dmax, cutoff, h, tau_star;
    double albedo[3], ext[3];
    double emission_scale;
    double ramp_in, ramp_out;
    int ramp_n;
    double ramp_pos[LG_MAX_RAMP];
    double ramp_col[LG_MAX_RAMP][3];
    int exact_steps;  // 1: t0 + k*h is exact for every sample (see step_is_exact)
};

// True when h is a multiple of 2^-30 below 2^23: every multiple of h below 2^23
// is then a multiple of 2^-30 below 2^23, which FP64 holds exactly, so for sample
// times under 2^23 index units the sequential recurrence t += h and the direct
// form h*ceil(t0/h) + k*h give the same bits.
bool step_is_exact(double h);

// Transmittance exp(ext * tau) of one channel (ext = -sigma_t <= 0); +INF (no
// path) gives 0 also for a channel without extinction.
__device__ __forceinline__ float trans_f(float ext, float tau) { return tau < INFINITY ? expf(ext * tau) : 0.0f; }

MarchConsts march_consts(const lg_shader *sh, const lg_march_params *mp);
LightConsts light_consts(const lg_shader *sh, const lg_march_params *mp);

// Edge optical depth (raymarchConnection, LightGraph.h:364-406): +INF when the
// ray misses the grid bbox or the dropout |T|^2 < cutoff fires.
// kCount: also flag the bricks the taps read (g.touched, work counters).
template <typename T, bool kCount = false>
__device__ __forceinline__ float march_edge_one(const lgd::GridView &g, const MarchConsts &mc, const float *pi,
                                                const float *pj, int *nsamp = nullptr) {
    const double sp[3] = {(double)pi[0], (double)pi[1], (double)pi[2]};
    double dir[3] = {(double)pj[0] - sp[0], (double)pj[1] - sp[1], (double)pj[2] - sp[2]};
    const double len = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    const float dist = (float)len;
    if (!(fabs(len - 0.0) <= 1e-7)) {  // Vec3::normalize(eps=1e-7): *this *= 1/len
        const double inv = 1.0 / len;
        for (int c = 0; c < 3; ++c) dir[c] = dir[c] * inv;
    }
    lgd::Isect s;
    if (!lgd::isect_world_ray(g, s, sp, dir, lgd::kDeltaD, (double)dist)) return INFINITY;
    double tau = 0.0;
    double t0, t1;
    int count = 0;
    while (lgd::isect_march(g, s, t0, t1)) {
        for (double t = mc.h * ceil(t0 / mc.h); t <= t1; t += mc.h) {
            ++count;
            if (nsamp) *nsamp = count;
            const double x = s.e[0] + s.d[0] * t, y = s.e[1] + s.d[1] * t, z = s.e[2] + s.d[2] * t;
            if (kCount) lgd::mark_taps(g, x, y, z);
            const float rho = lgd::trilinear<T>(g, x, y, z);
            double dens = (double)rho * mc.ds_sds_scds[0] * mc.ds_sds_scds[1] * mc.ds_sds_scds[2];
            dens = (mc.dmax < dens) ? mc.dmax : dens;
            if (dens < mc.cutoff && mc.sdm < mc.cutoff) continue;
            const double m = (dens < mc.sdm) ? mc.sdm : dens;
            tau += m * mc.h;
            if (tau > mc.tau_star) return INFINITY;
        }
    }
    return (float)tau;
}

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

// Shadow optical depth toward a light (LightGraph.h:461-491: no floor, no
// jitter).  Returns false when the ray misses the grid bbox (the reference then
// adds nothing for this sample); *dropped reports the |T|^2 < cutoff dropout.
template <typename T>
__device__ __forceinline__ bool shadow_tau(const lgd::GridView &g, const LightConsts &lc, const double p[3],
                                           const double ldir[3], double distance, double *tau_out, bool *dropped) {
    const double sd[3] = {-ldir[0], -ldir[1], -ldir[2]};
    lgd::Isect is;
    if (!lgd::isect_world_ray(g, is, p, sd, lgd::kDeltaD, distance)) return false;
    double tau = 0.0;
    double t0, t1;
    *dropped = false;
    while (lgd::isect_march(g, is, t0, t1)) {
        for (double t = lc.h * ceil(t0 / lc.h); t <= t1; t += lc.h) {
            const double x = is.e[0] + is.d[0] * t, y = is.e[1] + is.d[1] * t, z = is.e[2] + is.d[2] * t;
            if (g.touched) lgd::mark_taps(g, x, y, z);
            const float rho = lgd::trilinear<T>(g, x, y, z);
            double dens = (double)rho * lc.ds_sds_scds[0] * lc.ds_sds_scds[1] * lc.ds_sds_scds[2];
            dens = (lc.dmax < dens) ? lc.dmax : dens;
            if (dens < lc.cutoff) continue;
            tau += dens * lc.h;
            if (tau > lc.tau_star) {
                *dropped = true;
                *tau_out = tau;
                return true;
            }
        }
    }
    *tau_out = tau;
    return true;
}

// Vertex-light march for steps that are multiples of 2^-30 (computeLighting's
// shadow ray, LightGraph.h:449-488).  Every sample of every span sits at
// t = k h, so sample k is taken exactly when the 8^3 brick holding its position
// exists: the spans are the runs of existing bricks along the ray.  (A sample
// exactly on a span's closing plane, a span shorter than Delta<double> and a
// span that opens within Delta<float> of the ray's end are the exceptions; each
// needs an exact tie between two FP64 times.)  The 64 lanes take 64 consecutive
// k, each loading its brick's box code together with its 8 taps: one memory
// round trip per 64 samples and no brick walk.  A batch whose samples all lie in
// empty bricks jumps over the empty box around its last sample.  The optical
// depth is summed in rows of 16 samples aligned to multiples of 16 in k: within
// a row by a prefix sum, rows in order.  shadow_tau_rows sums the same rows in
// the same order (skipped and idle samples add exact zeros), so both forms give
// every ray the same bits.
template <typename T, bool kCount>
__device__ __forceinline__ bool shadow_tau_cells(const lgd::GridView &g, const LightConsts &lc, const double p[3],
                                                 const double ldir[3], double distance, double *tau_out,
                                                 bool *dropped, int *nsamp) {
    const int lane = threadIdx.x & 63;
    const double sd[3] = {-ldir[0], -ldir[1], -ldir[2]};
    lgd::Isect is;
    if (!lgd::isect_world_ray(g, is, p, sd, lgd::kDeltaD, distance)) return false;
    for (int c = 0; c < 3; ++c) {  // the wave's ray: scalar registers
        is.e[c] = lgd::wave_uniform(is.e[c]);
        is.d[c] = lgd::wave_uniform(is.d[c]);
        is.inv[c] = lgd::wave_uniform(is.inv[c]);
    }
    is.t0 = lgd::wave_uniform(is.t0);
    is.tmax = lgd::wave_uniform(is.tmax);
    *dropped = false;
    double tau = 0.0;
    *tau_out = 0.0;
    if (!(is.tmax - is.t0 > lgd::kDeltaF)) return true;  // isect_march: no span on a ray this short
    const double h = lc.h;
    // steps start at multiples of 16 (samples before the first k are idle), so
    // that the 16-sample rows below are the same whatever the lanes per ray
    const double kfirst = lgd::wave_uniform(ceil(is.t0 / h));
    double kb = floor(kfirst * (1.0 / 16.0)) * 16.0;
    for (;;) {
        const double k = kb + (double)lane;
        const double t = k * h;
        const bool act = k >= kfirst && t <= is.tmax;
        const double px = is.e[0] + is.d[0] * t, py = is.e[1] + is.d[1] * t, pz = is.e[2] + is.d[2] * t;
        // positions up to the clipped end lie inside the storage (its border
        // included); the clamp only guards the index against rounding
        const int bx = lgd::clampi(((int)floor(px)) >> 3, -1, g.NBx - 2);
        const int by = lgd::clampi(((int)floor(py)) >> 3, -1, g.NBy - 2);
        const int bz = lgd::clampi(((int)floor(pz)) >> 3, -1, g.NBz - 2);
        const int code = act ? (int)g.skip[((long long)(bz + 1) * g.NBy + (by + 1)) * g.NBx + (bx + 1)] : 0;
        double x = 0.0;
        float rho = 0.0f;
        if (act) rho = lgd::trilinear<T, false>(g, px, py, pz);
        const bool take = act && code >= lgd::kSkipExists;
        if (take) {
            double dens = (double)rho * lc.ds_sds_scds[0] * lc.ds_sds_scds[1] * lc.ds_sds_scds[2];
            dens = (lc.dmax < dens) ? lc.dmax : dens;
            if (!(dens < lc.cutoff)) x = dens * h;
        }
        const unsigned long long tmask = __ballot(take);
        // optical depth in rows of 16 samples: a prefix sum within each row
        // (DPP row_shr 1, 2, 4, 8), the rows added in order
        double pre = x;
        pre += lgd::dpp_move<0x111, 0xf>(pre);
        pre += lgd::dpp_move<0x112, 0xf>(pre);
        pre += lgd::dpp_move<0x114, 0xf>(pre);
        pre += lgd::dpp_move<0x118, 0xf>(pre);
        double trow[5];
        trow[0] = tau;
#pragma unroll
        for (int r = 0; r < 4; ++r) trow[r + 1] = trow[r] + lgd::wave_lane64(pre, 16 * r + 15);
        const int row = lane >> 4;
        const double tbase = row == 0 ? trow[0] : (row == 1 ? trow[1] : (row == 2 ? trow[2] : trow[3]));
        const unsigned long long dmask = __ballot(take && tbase + pre > lc.tau_star);
        // the samples the sequential march takes: up to the one that drops the ray
        const unsigned long long upto = dmask ? (2ull << (__ffsll((long long)dmask) - 1)) - 1ull : ~0ull;
        *nsamp += __popcll(tmask & upto);
        if (kCount && take && ((upto >> lane) & 1ull)) lgd::mark_taps(g, px, py, pz);
        if (dmask != 0ull) {
            *dropped = true;
            *tau_out = tau;
            return true;
        }
        tau = trow[4];
        if (!((__ballot(act) >> 63) & 1ull)) break;  // the ray ends inside this batch
        double knext = kb + 64.0;
        if (tmask == 0ull) {
            // all 64 samples in empty bricks: the box of Chebyshev radius d - 1
            // bricks around the last one is empty (its code d >= 2)
            const int d = __builtin_amdgcn_readlane(code, 63);
            if (d >= lgd::kSkipMin) {
                const int cb[3] = {__builtin_amdgcn_readlane(bx, 63), __builtin_amdgcn_readlane(by, 63),
                                   __builtin_amdgcn_readlane(bz, 63)};
                double texit = lgd::kDblMax;
                for (int c = 0; c < 3; ++c) {
                    if (is.d[c] == 0.0) continue;
                    const double plane = 8.0 * (double)(is.d[c] > 0.0 ? cb[c] + d : cb[c] - d + 1);
                    const double tc = (plane - is.e[c]) * is.inv[c];
                    texit = tc < texit ? tc : texit;
                }
                // every sample before texit (less a rounding margin) is in the box
                const double kj = floor((texit - 1e-9 * (1.0 + fabs(texit))) / h);
                knext = kj > knext ? floor(kj * (1.0 / 16.0)) * 16.0 : knext;
            }
        }
        kb = lgd::wave_uniform(knext);
    }
    *tau_out = tau;
    return true;
}

// The same march with one 16-lane DPP row per ray (four rays per wave): each
// row takes 16 consecutive k per step, from a multiple of 16, and sums them as
// shadow_tau_cells sums its rows, so each ray's optical depth, dropout and
// sample count are bit-identical to that form's.  Rows whose ray has ended (or
// that hold no ray: live false on entry) keep running the wave's shuffles with
// no samples.
template <typename T, bool kCount>
__device__ __forceinline__ void shadow_tau_rows(const lgd::GridView &g, const LightConsts &lc, const double p[3],
                                                const double ldir[3], double distance, bool live, bool *hit,
                                                double *tau_out, bool *dropped, int *nsamp) {
    const int lane = threadIdx.x & 63, gl = lane & 15, row = lane >> 4;
    const double sd[3] = {-ldir[0], -ldir[1], -ldir[2]};
    lgd::Isect is;
    *hit = false;
    *dropped = false;
    *tau_out = 0.0;
    if (live) {
        *hit = lgd::isect_world_ray(g, is, p, sd, lgd::kDeltaD, distance);
        // isect_march: no span on a ray this short
        live = *hit && is.tmax - is.t0 > lgd::kDeltaF;
    }
    const double h = lc.h;
    const double kfirst = live ? ceil(is.t0 / h) : 0.0;
    double kb = floor(kfirst * (1.0 / 16.0)) * 16.0;  // rows aligned as in shadow_tau_cells
    double tau = 0.0;
    for (;;) {
        if (__ballot(live) == 0ull) break;
        const double k = kb + (double)gl;
        const double t = k * h;
        const bool act = live && k >= kfirst && t <= is.tmax;
        double px = 0.0, py = 0.0, pz = 0.0;
        int bx = 0, by = 0, bz = 0, code = 0;
        double x = 0.0;
        bool take = false;
        if (act) {
            px = is.e[0] + is.d[0] * t;
            py = is.e[1] + is.d[1] * t;
            pz = is.e[2] + is.d[2] * t;
            bx = lgd::clampi(((int)floor(px)) >> 3, -1, g.NBx - 2);
            by = lgd::clampi(((int)floor(py)) >> 3, -1, g.NBy - 2);
            bz = lgd::clampi(((int)floor(pz)) >> 3, -1, g.NBz - 2);
            code = (int)g.skip[((long long)(bz + 1) * g.NBy + (by + 1)) * g.NBx + (bx + 1)];
            const float rho = lgd::trilinear<T, false>(g, px, py, pz);
            take = code >= lgd::kSkipExists;
            if (take) {
                double dens = (double)rho * lc.ds_sds_scds[0] * lc.ds_sds_scds[1] * lc.ds_sds_scds[2];
                dens = (lc.dmax < dens) ? lc.dmax : dens;
                if (!(dens < lc.cutoff)) x = dens * h;
            }
        }
        double pre = x;
        pre += lgd::dpp_move<0x111, 0xf>(pre);  // row_shr:1
        pre += lgd::dpp_move<0x112, 0xf>(pre);  // row_shr:2
        pre += lgd::dpp_move<0x114, 0xf>(pre);  // row_shr:4
        pre += lgd::dpp_move<0x118, 0xf>(pre);  // row_shr:8
        const unsigned tm = (unsigned)(__ballot(take) >> (16 * row)) & 0xffffu;
        const unsigned dm = (unsigned)(__ballot(take && tau + pre > lc.tau_star) >> (16 * row)) & 0xffffu;
        const unsigned am = (unsigned)(__ballot(act) >> (16 * row)) & 0xffffu;
        const double last = __shfl(pre, 15, 16);
        const int lcode = __shfl(code, 15, 16), lbx = __shfl(bx, 15, 16), lby = __shfl(by, 15, 16),
                  lbz = __shfl(bz, 15, 16);
        if (!live) continue;
        // the samples the sequential march takes: up to the one that drops the ray
        const unsigned upto = dm ? (2u << (__ffs((int)dm) - 1)) - 1u : 0xffffu;
        *nsamp += __popc(tm & upto);
        if (kCount && take && ((upto >> gl) & 1u)) lgd::mark_taps(g, px, py, pz);
        if (dm != 0u) {
            *dropped = true;
            live = false;
            continue;
        }
        tau = tau + last;
        if (!((am >> 15) & 1u)) {  // the ray ends inside this step
            live = false;
            continue;
        }
        double knext = kb + 16.0;
        if (tm == 0u && lcode >= lgd::kSkipMin) {
            // all 16 samples in empty bricks: the box of Chebyshev radius d - 1
            // bricks around the last one is empty (its code d >= 2)
            const int cb[3] = {lbx, lby, lbz};
            double texit = lgd::kDblMax;
            for (int c = 0; c < 3; ++c) {
                if (is.d[c] == 0.0) continue;
                const double plane = 8.0 * (double)(is.d[c] > 0.0 ? cb[c] + lcode : cb[c] - lcode + 1);
                const double tc = (plane - is.e[c]) * is.inv[c];
                texit = tc < texit ? tc : texit;
            }
            // every sample before texit (less a rounding margin) is in the box
            const double kj = floor((texit - 1e-9 * (1.0 + fabs(texit))) / h);
            knext = kj > knext ? floor(kj * (1.0 / 16.0)) * 16.0 : knext;
        }
        kb = knext;
    }
    *tau_out = tau;
}

// Emission term of computeLighting (LightGraph.h:494-496).
__device__ __forceinline__ void vertex_emission(const lgd::GridView &tg, int has_temp, const LightConsts &lc,
                                                const double p[3], double emission[3]) {
    emission[0] = emission[1] = emission[2] = 0.0;
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
}

// Light sample t (0 <= t < total samples) of a light list: light index, sample
// within the light and env-stream words drawn before it (2 per env sample,
// LightGraph.h:449-452 scales multi-sample lights by 10).
__device__ __forceinline__ void light_task(const lgl::DevLights &L, int t, int *li_out, uint32_t *env_words) {
    uint32_t w = 0;
    int li = 0;
    for (; li < L.n; ++li) {
        int ns = L.l[li].samples;
        if (ns > 1) ns *= 10;
        if (t < ns) break;
        t -= ns;
        if (L.l[li].type == LG_LIGHT_ENV) w += 2u * (uint32_t)ns;
    }
    if (li < L.n && L.l[li].type == LG_LIGHT_ENV) w += 2u * (uint32_t)t;
    *li_out = li;
    *env_words = w;
}

// computeLighting for one vertex (LightGraph.h:438-497).
template <typename T>
__device__ __forceinline__ void vertex_light_one(const lgd::GridView &g, const lgd::GridView &tg, int has_temp,
                                                 const LightConsts &lc, const lgl::DevLights &L, uint32_t ek0,
                                                 uint32_t ek1, uint32_t vertex, uint32_t solve, const double p[3],
                                                 double out[3]) {
    double emission[3];
    vertex_emission(tg, has_temp, lc, p, emission);
    lgd::Stream env_rng;
    env_rng.init(vertex, solve, lgd::DOM_VERTEX_ENV, ((uint64_t)ek1 << 32) | ek0);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int li = 0; li < L.n; ++li) {
        const lgl::DevLight &lt = L.l[li];
        int samples = lt.samples;
        if (samples > 1) samples *= 10;  // LightGraph.h:450-452
        double lumi[3] = {0.0, 0.0, 0.0};
        for (int s = 0; s < samples; ++s) {
            double ldir[3], inten[3], distance;
            lgl::illuminate(lt, p, env_rng, ldir, inten, distance);
            double tau;
            bool dropped;
            if (!shadow_tau<T>(g, lc, p, ldir, distance, &tau, &dropped)) continue;
            if (dropped) continue;  // transmittance set to zero
            for (int k = 0; k < 3; ++k) lumi[k] += inten[k] * (double)expf((float)(lc.ext[k] * tau));
        }
        for (int k = 0; k < 3; ++k) acc[k] += lc.albedo[k] * (lumi[k] / (double)samples);
    }
    for (int k = 0; k < 3; ++k) out[k] = acc[k] + emission[k];
}

// Kernel (a) for point i of solve d: the density-weighted sampler, run by one
// wave (all 64 lanes call it; lane 0 writes).  The point's u64 target in
// [0, total mass) selects the last active brick whose mass prefix is <= it
// (a 64-ary search: the same index as a binary search), then the first voxel
// of the brick, in storage order, at which the running sum of the integer
// voxel masses floor(rho 2^24) exceeds the rest (8 steps of a wave prefix sum
// over 64 voxels: the same voxel as the voxel-by-voxel scan).
template <typename T>
__device__ __forceinline__ void sample_point_wave(const SolveDesc &d, int i, uint32_t k0, uint32_t k1, float *out) {
    const lgd::GridView &g = d.g;
    const int lane = threadIdx.x & 63;
    const lgd::U4 r0 = lgd::philox(lgd::U4{(uint32_t)i, d.solve, 0u, lgd::DOM_SAMPLE}, k0, k1);
    const lgd::U4 r1 = lgd::philox(lgd::U4{(uint32_t)i, d.solve, 1u, lgd::DOM_SAMPLE}, k0, k1);
    const uint64_t u = ((uint64_t)r0.y << 32) | (uint64_t)r0.x;
    const uint64_t target = __umul64hi(u, d.Q);
    // largest lo in [0, nact - 1] with mass_prefix[lo] <= target (mass_prefix[0] = 0)
    int lo = 0, hi = d.nact - 1;
    while (lo < hi) {
        // 64 probes at lo + 1 + k (hi - lo) / 64, k = 0..63: the last probe that
        // passes bounds the answer from below, the next one from above
        const long long span = (long long)(hi - lo);
        const int probe = lo + 1 + (int)(((long long)lane * span) / 64);
        const bool ok = probe <= hi && d.mass_prefix[probe] <= target;
        const unsigned long long m = __ballot(ok);
        if (m == 0ull) {
            hi = lo;  // the first probe, lo + 1, failed
        } else {
            const int last = 63 - __clzll((long long)m);
            const int plast = lo + 1 + (int)(((long long)last * span) / 64);
            const int pnext = last < 63 ? lo + 1 + (int)(((long long)(last + 1) * span) / 64) : hi + 1;
            lo = plast;
            hi = (pnext - 1 < hi) ? pnext - 1 : hi;
        }
    }
    const int sb = d.active_list[lo];
    const uint64_t local = target - d.mass_prefix[lo];
    const T *p = static_cast<const T *>(g.data) + ((long long)sb << 9);
    // the brick's 512 voxels in one round trip: 8 loads in flight per lane
    float fv[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) fv[c] = lgd::to_f32(p[64 * c + lane]);
    int v = 511;
    uint64_t carry = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const float f = fv[c];
        uint64_t q = (f > 0.0f) ? (uint64_t)fmin((double)f * 16777216.0, 549755813888.0 * 16777216.0) : 0ull;
        q = lgd::wave_prefix_sum(q);
        const unsigned long long hit = __ballot(local < carry + q);
        if (hit) {
            v = 64 * c + (__ffsll((long long)hit) - 1);
            break;
        }
        carry += lgd::wave_lane64(q, 63);
    }
    if (lane != 0) return;
    const int sbx = sb % g.NBx, sby = (sb / g.NBx) % g.NBy, sbz = sb / (g.NBx * g.NBy);
    const int vx = (sbx - 1) * 8 + (v & 7), vy = (sby - 1) * 8 + ((v >> 3) & 7), vz = (sbz - 1) * 8 + (v >> 6);
    const double px = (double)vx - 0.5 + (double)lgd::u01(r0.z);
    const double py = (double)vy - 0.5 + (double)lgd::u01(r0.w);
    const double pz = (double)vz - 0.5 + (double)lgd::u01(r1.x);
    out[0] = (float)(g.origin[0] + g.voxel * px);
    out[1] = (float)(g.origin[1] + g.voxel * py);
    out[2] = (float)(g.origin[2] + g.voxel * pz);
}

struct PipelineParams {
    float *pts;             // V x 3 points (input when external_points, else output)
    float *L;               // V x 3 vertex light (output)
    float *out;             // V x 3 diffused radiance (output)
    uint64_t seed, env_seed;
    int external_points;
    int max_connect;
    double r2;              // squared edge radius (float arithmetic of the reference)
    double ext[3];          // shader extinction (negative)
    MarchConsts mc;
    LightConsts lc;
    lgl::DevLights lights;
};

struct PipelineResult {
    int Ktot = 0, n_edges_total = 0, n_components = 0, n_big = 0;
    int *edges = nullptr;   // global (i, j), row-major, solve-major
    float *w = nullptr;
    int *epos = nullptr;    // candidate -> edge index (exclusive scan of accepted)
    int *off = nullptr;     // vertex -> first candidate
};

// Device workspace reused across calls; grows on demand.
struct Pipeline {
    int v_cap = 0, s_cap = 0;
    SolveDesc *sd = nullptr;
    int *sid = nullptr, *rowcnt = nullptr, *off = nullptr, *colcnt = nullptr;
    uint8_t *risky = nullptr;
    int *visit = nullptr, *visit_pos = nullptr, *visit_list = nullptr;
    int *adeg = nullptr, *aoff = nullptr, *acur = nullptr;
    int *parent = nullptr, *root = nullptr;
    uint64_t *keys = nullptr, *keys_sorted = nullptr;
    int *flag = nullptr, *cid = nullptr, *member = nullptr, *local = nullptr, *comp_off = nullptr;
    int *level = nullptr;   // breadth-first level of each vertex inside its component (bfs_order)
    int *coloff = nullptr, *colfill = nullptr;  // risky-column slot lists of the parallel greedy
    int *greedy_err = nullptr;                  // 1: a greedy wait exceeded its bound
    int *big_list = nullptr, *counters = nullptr;
    int *cls16 = nullptr, *cls48 = nullptr, *cls128 = nullptr;  // components of 2-32, 33-64, 65-128 vertices
    void *tmp_vertex = nullptr;
    size_t tmp_vertex_bytes = 0;
    // candidate / edge sized
    int *cand = nullptr, *accepted = nullptr, *epos = nullptr, *edges_ = nullptr, *anb = nullptr;
    int *colslot = nullptr;     // risky-column slot lists of the parallel greedy
    int2 *slot_info = nullptr;  // per candidate slot: (j, rank in j's column or -1)
    size_t colslot_cap = 0, slot_info_cap = 0;
    float *w_ = nullptr, *aw = nullptr;
    uint8_t *tmp_cand = nullptr;
    size_t cand_cap = 0, accepted_cap = 0, epos_cap = 0, tmp_cand_cap = 0, edges_cap = 0, w_cap = 0, anb_cap = 0,
           aw_cap = 0;
    // large components
    float *big_T = nullptr;
    float *big_ln = nullptr;  // normalised vertex light of the large components (3 per row)
    unsigned long long *stats = nullptr;  // ST_COUNT counters; kernels add to them when stats_on
    bool stats_on = false;
    bool bfs_order = false;  // large components closed in breadth-first vertex order
    bool nd_order = false;   // large components closed in nested-dissection vertex order
    double *ray = nullptr;  // per (vertex, light sample) shadowed intensity
    size_t ray_cap = 0;
    // pinned host words for the solve's two readbacks (candidate count, component
    // counts): a copy into pageable memory goes through a staging buffer
    int *h_pin = nullptr;
    size_t h_pin_cap = 0;
    int32_t pinned_words(size_t n, int **out);
    ~Pipeline() {
        if (h_pin) (void)hipHostFree(h_pin);
    }
    // spatial ray order of large calls: keys, sorted keys, vertices, sorted vertices
    // (four runs of lsort_cap / 4 words) and the sort's temporary storage
    uint32_t *lkeys = nullptr;
    size_t lsort_cap = 0;
    uint8_t *tmp_light = nullptr;
    size_t tmp_light_cap = 0;
    // neighbour grid
    CellDesc *cd = nullptr;
    uint32_t *ckeys = nullptr, *ckeys_sorted = nullptr;
    int *cvals = nullptr, *cvals_sorted = nullptr;
    int *cstart = nullptr, *cend = nullptr;
    size_t cells_cap = 0, cend_cap = 0;
    long long total_cells = 0;
    uint8_t *big_meta = nullptr;
    size_t bigT_cap = 0, bigln_cap = 0, bigmeta_cap = 0;

    int32_t reserve_vertices(lg_ctx *ctx, int V, int S);
    void release_vertex(lg_ctx *ctx);
    void release(lg_ctx *ctx);
};

// stages of run_pipeline
int32_t pl_begin(lg_ctx *ctx, Pipeline &pl, const std::vector<SolveDesc> &descs, int *V_out, int *nmax_out);
int32_t pl_sample(lg_ctx *ctx, Pipeline &pl, int V, int dtype, uint64_t seed, float *pts);
int32_t pl_light(lg_ctx *ctx, Pipeline &pl, int V, int dtype, const LightConsts &lc, const lgl::DevLights &lights,
                 uint64_t env_seed, const float *pts, float *L);
// after_grid (may be null) runs on the host right after the neighbour-grid
// launches are queued (the solve forks its vertex light there).
int32_t pl_graph(lg_ctx *ctx, Pipeline &pl, int V, int S, int nmax, const float *pts, double r2, int max_connect,
                 int *Ktot_out, const std::function<int32_t()> *after_grid = nullptr);
int32_t pl_march(lg_ctx *ctx, Pipeline &pl, int dtype, const MarchConsts &mc, const float *pts, int Ktot);
// L_ready: when not null, an event after which L is complete (each stream
// waits on it before its first kernel that reads L: the component search and
// the blocked closure of the large components run before it).  Runs after pl_graph on the
// same pipeline: pl_graph's edge-list launch clears this stage's counters.
int32_t pl_components(lg_ctx *ctx, Pipeline &pl, const std::vector<SolveDesc> &descs, int V, int Ktot,
                      const double ext[3], const float *L, float *out, PipelineResult *res,
                      hipEvent_t L_ready = nullptr, const float *pts = nullptr);

// helpers from lg_solve.hip
int32_t fw_init_impl(lg_ctx *ctx, int n, float *d_D, int ld);
int32_t diffuse_impl(lg_ctx *ctx, int n, const float *d_D, int ld, const double sigma_t[3], const float *d_L,
                     float *d_out, double *d_sums, float *d_Ln, bool timed);
SolveDesc make_desc(lg_grid *density, const lg_grid *temperature, int n, int voff, uint32_t solve);

}  // namespace lg
