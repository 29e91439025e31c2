// A fake of the device half of the C ABI (include/lightgraph.h), for testing the
// host code above it on a machine without a GPU and under the address and
// undefined-behaviour sanitizers: "device" memory is host memory counted by a
// fake allocator with an optional byte limit, kernels are replaced by simple host
// fills, and failures can be injected.  Test infrastructure only; it links with
// lightgraph_amd/csrc/lg_serve.cpp, lg_host.cpp and the command-line front end.
//
// Injection (fake_set_*): grids generated for a scene seed fail with
// LG_ERR_NOMEM; device allocations fail once the fake device holds more than a
// limit; render calls sleep a few milliseconds per megapixel so that the
// server's overlap and batching can be observed.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "lightgraph.h"

namespace lg {
thread_local char g_err[512];
void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
}  // namespace lg

struct lg_ctx {
    int device;
};
struct lg_grid {
    lg_grid_desc desc{};
    void *data = nullptr;
    size_t bytes = 0;
};
struct lg_solver {
    int32_t n_max = 0, ld_max = 0;
    lg_solver_buffers buf{};
};
struct lg_envmap {
    int w = 0, h = 0;
};

namespace {
std::mutex g_mu;
std::map<void *, size_t> g_allocs;
size_t g_in_use = 0, g_peak = 0, g_limit = 0;
uint64_t g_fail_seed = ~0ull;
double g_render_ms_per_mpix = 0.0;
std::atomic<int> g_renders_running{0}, g_renders_peak{0};

void *fake_alloc(size_t bytes) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_limit && g_in_use + bytes > g_limit) return nullptr;
    void *p = std::calloc(1, bytes ? bytes : 1);
    if (!p) return nullptr;
    g_allocs[p] = bytes;
    g_in_use += bytes;
    if (g_in_use > g_peak) g_peak = g_in_use;
    return p;
}

void fake_free(void *p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_allocs.find(p);
    if (it == g_allocs.end()) {
        std::fprintf(stderr, "fake_abi: free of an unknown pointer\n");
        std::abort();
    }
    g_in_use -= it->second;
    g_allocs.erase(it);
    std::free(p);
}

int32_t nomem(const char *what) {
    lg::set_error("%s: out of fake device memory", what);
    return LG_ERR_NOMEM;
}
}  // namespace

extern "C" {

// ----------------------------------------------------------- injection ----
void fake_set_limit(uint64_t bytes) { g_limit = bytes; }
void fake_set_fail_seed(uint64_t seed) { g_fail_seed = seed; }
void fake_set_render_ms_per_mpix(double ms) { g_render_ms_per_mpix = ms; }
uint64_t fake_bytes_in_use(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_in_use;
}
uint64_t fake_live_allocations(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_allocs.size();
}
int32_t fake_renders_peak(void) { return g_renders_peak.load(); }

// ------------------------------------------------------------- context ----
const char *lg_last_error(void) { return lg::g_err; }
int32_t lg_abi_version(void) { return LG_ABI_VERSION; }
int32_t lg_ctx_create(int32_t device, uint64_t, lg_ctx **out) {
    if (!out) return LG_ERR_INVALID;
    *out = new lg_ctx{device};
    return LG_OK;
}
int32_t lg_ctx_destroy(lg_ctx *ctx) {
    delete ctx;
    return LG_OK;
}
int32_t lg_sync(lg_ctx *) { return LG_OK; }
int32_t lg_malloc(lg_ctx *, size_t bytes, void **out) {
    *out = fake_alloc(bytes);
    return *out ? LG_OK : nomem("lg_malloc");
}
int32_t lg_free(lg_ctx *, void *p) {
    fake_free(p);
    return LG_OK;
}
int32_t lg_memcpy_h2d(lg_ctx *, void *dst, const void *src, size_t bytes) {
    if (bytes) std::memcpy(dst, src, bytes);
    return LG_OK;
}
int32_t lg_memcpy_d2h(lg_ctx *, void *dst, const void *src, size_t bytes) {
    if (bytes) std::memcpy(dst, src, bytes);
    return LG_OK;
}

// --------------------------------------------------------------- grids ----
static int32_t make_grid(const lg_grid_desc &d, lg_grid **out) {
    lg_grid *g = new lg_grid();
    g->desc = d;
    g->bytes = (size_t)d.nx * d.ny * d.nz * 4;
    g->data = fake_alloc(g->bytes);
    if (!g->data) {
        delete g;
        return nomem("grid");
    }
    *out = g;
    return LG_OK;
}
int32_t lg_grid_create(lg_ctx *, const lg_grid_desc *desc, const float *h_dense, lg_grid **out) {
    if (!desc || !out) return LG_ERR_INVALID;
    const int32_t rc = make_grid(*desc, out);
    if (rc == LG_OK && h_dense) std::memcpy((*out)->data, h_dense, (*out)->bytes);
    return rc;
}
int32_t lg_grid_generate(lg_ctx *, const lg_scene_spec *spec, lg_grid **density, lg_grid **temperature) {
    if (!spec || !density) return LG_ERR_INVALID;
    if (spec->seed == g_fail_seed) return nomem("lg_grid_generate (injected)");
    lg_grid_desc d{};
    d.nx = d.ny = d.nz = spec->n;
    d.dtype = spec->dtype;
    d.voxel_size = 1.0;
    int32_t rc = make_grid(d, density);
    if (rc == LG_OK && temperature) {
        rc = make_grid(d, temperature);
        if (rc != LG_OK) {
            fake_free((*density)->data);
            delete *density;
            *density = nullptr;
        }
    }
    if (rc == LG_OK) {
        float *v = static_cast<float *>((*density)->data);
        for (size_t k = 0; k < (size_t)spec->n * spec->n * spec->n; k += 7) v[k] = 0.25f;
    }
    return rc;
}
int32_t lg_grid_destroy(lg_ctx *, lg_grid *g) {
    if (!g) return LG_OK;
    fake_free(g->data);
    delete g;
    return LG_OK;
}
int32_t lg_grid_info_get(const lg_grid *g, lg_grid_info *info) {
    std::memset(info, 0, sizeof(*info));
    info->desc = g->desc;
    const int32_t n[3] = {g->desc.nx, g->desc.ny, g->desc.nz};
    for (int a = 0; a < 3; ++a) {
        info->bricks[a] = (n[a] + 7) / 8;
        info->bbox_max[a] = 8 * info->bricks[a];
    }
    info->active_bricks = info->bricks[0] * info->bricks[1] * info->bricks[2];
    info->device_bytes = g->bytes;
    return LG_OK;
}
int32_t lg_grid_active_bbox(lg_ctx *, const lg_grid *g, int32_t vmin[3], int32_t vmax[3]) {
    const int32_t n[3] = {g->desc.nx, g->desc.ny, g->desc.nz};
    for (int a = 0; a < 3; ++a) {
        vmin[a] = 0;
        vmax[a] = n[a] - 1;
    }
    return LG_OK;
}

// ------------------------------------------------------- solves (fills) ----
int32_t lg_solver_create(lg_ctx *, int32_t n_max, int32_t, int32_t flags, lg_solver **out) {
    lg_solver *s = new lg_solver();
    s->n_max = n_max;
    s->buf.d_points = static_cast<float *>(fake_alloc((size_t)n_max * 12));
    s->buf.d_L = static_cast<float *>(fake_alloc((size_t)n_max * 12));
    s->buf.d_diffuse = static_cast<float *>(fake_alloc((size_t)n_max * 12));
    if (flags & LG_SOLVER_DENSE_D) {
        s->ld_max = ((n_max + 127) / 128) * 128;
        s->buf.d_D = static_cast<float *>(fake_alloc((size_t)s->ld_max * s->ld_max * 4));
    }
    if (!s->buf.d_points || !s->buf.d_L || !s->buf.d_diffuse || ((flags & LG_SOLVER_DENSE_D) && !s->buf.d_D)) {
        fake_free(s->buf.d_points);
        fake_free(s->buf.d_L);
        fake_free(s->buf.d_diffuse);
        fake_free(s->buf.d_D);
        delete s;
        return nomem("lg_solver_create");
    }
    *out = s;
    return LG_OK;
}
int32_t lg_solver_destroy(lg_ctx *, lg_solver *s) {
    if (!s) return LG_OK;
    fake_free(s->buf.d_points);
    fake_free(s->buf.d_L);
    fake_free(s->buf.d_diffuse);
    fake_free(s->buf.d_D);
    delete s;
    return LG_OK;
}
int32_t lg_solver_timing(const lg_solver *, lg_solve_timing *out) {
    std::memset(out, 0, sizeof(*out));
    out->sample_ms = out->graph_ms = out->march_ms = out->light_ms = out->fw_ms = 0.25f;
    out->total_ms = 1.25f;
    return LG_OK;
}
int32_t lg_solver_get(const lg_solver *s, lg_solver_buffers *out) {
    *out = s->buf;
    return LG_OK;
}
static void fill_points(lg_solver *s, const lg_grid *g, int32_t voff, int32_t n, uint32_t solve) {
    for (int32_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            const uint32_t h = (uint32_t)(i * 2654435761u + c * 40503u + solve * 97u);
            const int32_t n[3] = {g->desc.nx, g->desc.ny, g->desc.nz};
            s->buf.d_points[3 * (voff + i) + c] = (float)(h % 1000u) * 1e-3f * (float)n[c];
            s->buf.d_diffuse[3 * (voff + i) + c] = 0.5f;
            s->buf.d_L[3 * (voff + i) + c] = 1.0f;
        }
}
int32_t lg_solve(lg_ctx *, lg_solver *s, const lg_grid *density, const lg_grid *, const lg_shader *, const lg_light *,
                 int32_t, const lg_march_params *, const lg_solve_params *sp) {
    if (sp->n < 1 || sp->n > s->n_max) return LG_ERR_INVALID;
    fill_points(s, density, 0, sp->n, sp->solve);
    s->buf.n = sp->n;
    s->buf.ld = 0;
    if ((sp->flags & LG_SOLVE_DENSE) && s->buf.d_D) {  // isolated vertices: +INF off the diagonal
        s->buf.ld = ((sp->n + 127) / 128) * 128;
        for (int32_t i = 0; i < sp->n; ++i)
            for (int32_t j = 0; j < s->buf.ld; ++j) s->buf.d_D[(size_t)i * s->buf.ld + j] = i == j ? 0.0f : INFINITY;
    }
    s->buf.n_edges = 0;
    s->buf.n_components = sp->n;
    return LG_OK;
}
int32_t lg_solve_batch(lg_ctx *, lg_solver *s, const lg_batch_item *items, int32_t count, const lg_shader *,
                       const lg_light *, int32_t, const lg_march_params *, const lg_solve_params *) {
    int32_t V = 0;
    for (int k = 0; k < count; ++k) {
        if (V + items[k].n > s->n_max) return LG_ERR_INVALID;
        fill_points(s, items[k].density, V, items[k].n, items[k].solve);
        V += items[k].n;
    }
    s->buf.n = V;
    s->buf.n_edges = 0;
    s->buf.n_components = V;
    return LG_OK;
}
int32_t lg_dart_throw(lg_ctx *, const lg_grid *density, const double *, const double *, int32_t count,
                      const uint32_t *h_iters, uint32_t, uint32_t target, float *, float *d_points_out,
                      int32_t out_cap, int32_t *h_counts) {
    for (int32_t it = 0; it < count; ++it) {
        const int32_t n = std::min((int32_t)target, out_cap);
        for (int32_t i = 0; i < 3 * n; ++i)
            d_points_out[(size_t)it * out_cap * 3 + i] =
                (float)((i * 7919 + (int)h_iters[it] * 13) % 997) * 1e-3f * (float)density->desc.nx;
        h_counts[it] = n;
    }
    return LG_OK;
}

// ------------------------------------------------------ splat and render ----
int32_t lg_scatter_grid_create(lg_ctx *, const lg_grid *density, int32_t f, lg_scatter_grid *sg) {
    std::memset(sg, 0, sizeof(*sg));
    sg->nx = sg->ny = sg->nz = density->desc.nx / f + 3;
    sg->voxel_size = density->desc.voxel_size * f;
    const size_t n = (size_t)sg->nx * sg->ny * sg->nz;
    sg->d_rgb = static_cast<float *>(fake_alloc(n * 12));
    sg->d_active = static_cast<uint8_t *>(fake_alloc(n));
    if (!sg->d_rgb || !sg->d_active) {
        fake_free(sg->d_rgb);
        fake_free(sg->d_active);
        std::memset(sg, 0, sizeof(*sg));
        return nomem("lg_scatter_grid_create");
    }
    return LG_OK;
}
int32_t lg_scatter_grid_destroy(lg_ctx *, lg_scatter_grid *sg) {
    fake_free(sg->d_rgb);
    fake_free(sg->d_active);
    std::memset(sg, 0, sizeof(*sg));
    return LG_OK;
}
int32_t lg_splat(lg_ctx *, const float *, const float *d_rad, int64_t count, double radius, lg_scatter_grid *sg) {
    if (!sg || !sg->d_rgb || radius < 0) return LG_ERR_INVALID;
    float m = 0.f;
    for (int64_t i = 0; i < 3 * count; ++i) m = std::max(m, d_rad[i]);
    const size_t n = (size_t)sg->nx * sg->ny * sg->nz * 3;
    for (size_t k = 0; k < n; ++k) sg->d_rgb[k] = m;
    return LG_OK;
}
int32_t lg_render(lg_ctx *, const lg_grid *density, const lg_scatter_grid *, const lg_grid *, const lg_camera *cam,
                  const lg_shader *, const lg_light *, int32_t n_lights, const lg_march_params *, int32_t samples,
                  uint64_t seed, uint64_t, const lg_films *films) {
    if (!density || !cam || !films || samples < 1) return LG_ERR_INVALID;
    const int running = ++g_renders_running;
    int peak = g_renders_peak.load();
    while (running > peak && !g_renders_peak.compare_exchange_weak(peak, running)) {
    }
    const size_t npix = (size_t)cam->width * cam->height;
    if (g_render_ms_per_mpix > 0)
        std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(1e3 * g_render_ms_per_mpix * npix / 1e6) + 1));
    float *f[4] = {films->d_rgb, films->d_direct, films->d_scatter, films->d_emission};
    for (int k = 0; k < 4; ++k)
        for (size_t p = 0; p < npix; ++p)
            for (int c = 0; c < 4; ++c)
                f[k][4 * p + c] = (float)((seed + (uint64_t)n_lights + (uint64_t)k + (uint64_t)c) % 7) * 0.125f;
    --g_renders_running;
    return LG_OK;
}
int32_t lg_render_workspace(int32_t width, int32_t height, int32_t samples, int32_t n_lights, int64_t,
                            uint64_t *bytes) {
    if (!bytes || width < 1 || height < 1 || samples < 1 || n_lights < 0) return LG_ERR_INVALID;
    *bytes = (uint64_t)width * height * samples * 64;
    return LG_OK;
}
int32_t lg_envmap_create(lg_ctx *, int32_t w, int32_t h, const float *, lg_envmap **out) {
    *out = new lg_envmap{w, h};
    return LG_OK;
}
int32_t lg_envmap_destroy(lg_ctx *, lg_envmap *e) {
    delete e;
    return LG_OK;
}

}  // extern "C"
