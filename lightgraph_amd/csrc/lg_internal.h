// Internal declarations shared by the HIP translation units of liblightgraph.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "lightgraph.h"

namespace lg {

// ---------------------------------------------------------------- errors ----
void set_error(const char *fmt, ...);

#define LG_CHECK_ARG(cond, ...)                 \
    do {                                        \
        if (!(cond)) {                          \
            ::lg::set_error(__VA_ARGS__);       \
            return LG_ERR_INVALID;              \
        }                                       \
    } while (0)

#define LG_HIP(expr)                                                              \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess) {                                                   \
            ::lg::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                            __FILE__, __LINE__);                                  \
            return LG_ERR_HIP;                                                    \
        }                                                                         \
    } while (0)

#define LG_TRY(expr)                      \
    do {                                  \
        int32_t rc_ = (expr);             \
        if (rc_ != LG_OK) return rc_;     \
    } while (0)

// A launch check: reports a bad launch configuration immediately.
#define LG_LAUNCHED(name)                                                         \
    do {                                                                          \
        hipError_t e_ = hipGetLastError();                                        \
        if (e_ != hipSuccess) {                                                   \
            ::lg::set_error("launch of %s failed: %s", name, hipGetErrorString(e_)); \
            return LG_ERR_HIP;                                                    \
        }                                                                         \
    } while (0)

// ------------------------------------------------------------ kernel timer ----
struct KernelTimer {
    struct Pending {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::map<std::string, std::pair<uint64_t, double>> totals;
    std::vector<hipEvent_t> pool;
};

}  // namespace lg

// ---------------------------------------------------------------- context ----
struct lg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream for work that overlaps the main stream inside one call (a
    // solve's vertex light beside its graph stages, the small-component
    // closures beside the blocked one); always joined back into `stream` before
    // the call returns.  The events are re-recorded per fork: a stream wait
    // binds to the record current when the wait is enqueued.  ev_light marks
    // the end of a solve's vertex light alone (the readers of the light wait
    // for it), ev_join the end of everything forked.
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_light = nullptr;
    uint64_t budget = 0;         // 0 = unlimited
    uint64_t in_use = 0;
    std::map<void *, size_t> allocs;
    bool timing = false;
    bool count_bricks = false;  // renders under timing also count distinct shadow bricks
    std::string timing_only;  // non-empty: time only this kernel's launches
    lg::KernelTimer ktimer;
    hipEvent_t marks[16] = {};
    // scratch reused by entry points that need temporary device memory (slot 0),
    // and a second slot for a call that needs two live at once (the render)
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    void *scratch1 = nullptr;
    size_t scratch1_bytes = 0;
    // device copy of the last blob scene's ellipsoids (lg_grid_generate)
    void *blobs_dev = nullptr;
    // work counters of the last lg_render (kernel timing on), device u64 slots
    unsigned long long *render_stats = nullptr;
    uint64_t render_chunks = 0;
    // batched Floyd-Warshall choices (lg_fw_tuning) and the last captured schedule
    lg_fw_tuning fw_tune{};
    bool fw_tune_set = false;
    hipGraphExec_t fw_graph = nullptr;
    std::vector<uint64_t> fw_graph_key;
    int cu_count = 256;
    void *stage_pipeline = nullptr;  // lg::Pipeline used by the single-stage entry points
};

// Dense-brick grid.  Storage holds (nb+2)^3 bricks of 8^3 voxels: one empty brick
// of border on every side so that every trilinear tap of a sample inside the
// clip box, and every DDA cell within one brick of it, is addressable without
// bounds checks.  Voxel (x,y,z) lives in storage brick ((x>>3)+1, (y>>3)+1,
// (z>>3)+1) at local offset (z&7)*64 + (y&7)*8 + (x&7).
struct lg_grid {
    lg_grid_desc desc{};
    int nb[3] = {0, 0, 0};       // bricks per axis (ceil(n/8))
    int NB[3] = {0, 0, 0};       // storage bricks per axis (nb+2)
    void *data = nullptr;        // NB[0]*NB[1]*NB[2]*512 elements of desc.dtype
    uint8_t *occ = nullptr;      // NB product, 1 = brick exists (has an active voxel)
    uint8_t *skip = nullptr;     // NB product, box code (lgd::brick_skip): radius of the uniform box
    int active_bricks = 0;
    int bmin[3] = {0, 0, 0}, bmax[3] = {-1, -1, -1};  // active brick bbox (inclusive)
    uint64_t active_voxels = 0;
    uint64_t bytes = 0;
    // density-weighted sampling tables (built lazily by lg_sample_points)
    int32_t *active_list = nullptr;   // storage indices of active bricks, ascending
    uint64_t *mass_prefix = nullptr;  // exclusive prefix of brick masses (active_bricks+1)
    uint64_t total_mass = 0;
    // work counters: one byte per storage brick, set by the marches' taps while
    // a call counts its work (allocated on the first such call)
    uint8_t *touched = nullptr;
};

struct lg_envmap {
    int width = 0, height = 0;
    float *rgb = nullptr;        // width*height*3
    double *cond = nullptr;      // height*(width+1)
    double *marg = nullptr;      // height+1
    double *rowint = nullptr;    // height
    double *func_marg = nullptr; // height (marginal function = row integrals)
    double total = 0.0;
};

struct lg_solver {
    int n_max = 0, ld_max = 0;
    bool with_pred = false, with_dense = false, timed = false, stats_valid = false;
    int n_big = 0;
    lg_solver_buffers buf{};
    void *pipeline = nullptr;  // lg::Pipeline
    hipEvent_t ev[6] = {};
};

namespace lg {

// memory (budget-checked)
int32_t dmalloc(lg_ctx *ctx, size_t bytes, void **out);
int32_t dfree(lg_ctx *ctx, void *p);
template <typename T>
int32_t dalloc(lg_ctx *ctx, size_t count, T **out) {
    void *p = nullptr;
    int32_t rc = dmalloc(ctx, count * sizeof(T), &p);
    *out = static_cast<T *>(p);
    return rc;
}
// scratch that lives until the next call that asks for scratch of the same slot
int32_t scratch(lg_ctx *ctx, size_t bytes, void **out, int slot = 0);

// kernel timing hooks (no-ops unless lg_ctx_set_kernel_timing enabled it)
void ktimer_begin(lg_ctx *ctx, const char *name, hipEvent_t *a);
void ktimer_end(lg_ctx *ctx, const char *name, hipEvent_t a);

// frees the host side of ctx->stage_pipeline (device memory goes with the ctx)
void destroy_stage_pipeline(lg_ctx *ctx);

// grid helpers (lg_runtime.hip)
int32_t grid_alloc(lg_ctx *ctx, const lg_grid_desc *desc, lg_grid **out);
int32_t grid_finalize(lg_ctx *ctx, lg_grid *g);

// helpers implemented in the stage translation units
int32_t grid_build_sampling_tables(lg_ctx *ctx, lg_grid *g);
// Work counters of one solve call (LG_SOLVE_STATS), slots of a device u64 array.
// Each counter has kStatSlots slots picked by block index (spreads the atomics);
// the host sums them.
enum : int { ST_EDGE_SAMPLES = 0, ST_SHADOW_RAYS, ST_SHADOW_SAMPLES, ST_FW_TILES, ST_FW_PIVOTS, ST_SMALL_UPDATES,
             ST_FW_PANEL_TILES, ST_LIGHT_BRICKS, ST_EDGE_BRICKS, ST_SMALL_UPDATES_64, ST_SMALL_UPDATES_32,
             ST_COUNT = 12 };
constexpr int kStatSlots = 64;
__device__ __forceinline__ void stat_add(unsigned long long *stats, int counter, unsigned long long v) {
    atomicAdd(&stats[counter * kStatSlots + (blockIdx.x & (kStatSlots - 1))], v);
}

int32_t fw_solve_impl(lg_ctx *ctx, int32_t n, float *d_D, int32_t ld);
// Blocked Floyd-Warshall over many matrices of uneven size n[g] (ld = fw_ld(n[g]))
// with one pivot, one panel and one remainder launch per 128-pivot round.
int32_t fw_solve_batched_impl(lg_ctx *ctx, const std::vector<float *> &Ds, const std::vector<int> &ns,
                              unsigned long long *stats = nullptr);
// lg_fw_tuning of the context: set it from the per-device measurement (taken on
// first use), or measure again
int32_t fw_tuning_init(lg_ctx *ctx);
int32_t fw_autotune(lg_ctx *ctx, lg_fw_tuning *out);
int32_t fw_pred_impl(lg_ctx *ctx, int32_t n, const float *d_D, int32_t ld, int32_t *d_pred,
                     const int32_t *d_edges, const float *d_w, int32_t n_edges);

constexpr int FW_TILE = 128;
inline int fw_ld(int n) { return ((n + FW_TILE - 1) / FW_TILE) * FW_TILE; }

inline size_t grid_storage_elems(const lg_grid *g) {
    return (size_t)g->NB[0] * g->NB[1] * g->NB[2] * 512;
}
inline size_t dtype_bytes(int dtype) { return dtype == LG_F32 ? 4 : 2; }

}  // namespace lg

#include "lg_device.h"

namespace lg {

// Distinct-brick counters of the marches (LG_SOLVE_STATS, render statistics):
// grid_touch_begin allocates the grid's per-brick flags if needed and clears
// them; grid_touch_count adds the number of flagged bricks to counter `counter`
// of `stats` (kStatSlots slots per counter) and clears the flags again.
int32_t grid_touch_begin(lg_ctx *ctx, lg_grid *g);
int32_t grid_touch_count(lg_ctx *ctx, const lg_grid *g, unsigned long long *stats, int counter);

// Kernel-side view of a grid.
inline lgd::GridView make_view(const lg_grid *g) {
    lgd::GridView v{};
    v.data = g->data;
    v.occ = g->occ;
    v.skip = g->skip;
    v.NBx = g->NB[0];
    v.NBy = g->NB[1];
    v.NBz = g->NB[2];
    v.nbx = g->nb[0];
    v.nby = g->nb[1];
    v.nbz = g->nb[2];
    v.sy = (long long)g->NB[0] * 512;
    v.sz = (long long)g->NB[0] * g->NB[1] * 512;
    for (int c = 0; c < 3; ++c) {
        v.bbmin[c] = 8.0 * g->bmin[c];
        v.bbmax[c] = 8.0 * (g->bmax[c] + 1);
        v.origin[c] = g->desc.origin[c];
    }
    v.voxel = g->desc.voxel_size;
    v.inv_voxel = 1.0 / g->desc.voxel_size;
    v.empty = g->active_bricks == 0;
    const size_t data_bytes = grid_storage_elems(g) * dtype_bytes(g->desc.dtype);
    v.off32 = data_bytes < (1ull << 32);
    v.data_bytes = v.off32 ? (uint32_t)data_bytes : 0u;
    return v;
}

// Largest optical depth that survives the reference's dropout test
// |T|^2 < cutoff with T_c = exp(extinction_c * tau) (LightGraph.h:396,484,
// Renderer.h:635,660): dropout happens iff tau > tau_star.
double dropout_tau(const double extinction[3], double cutoff);

}  // namespace lg
