// C++ host surface of the MI355X LightGraph path: the classes and call shapes of
// the reference's LightGraph.h / Renderer.h / ScatterGrid.h / Shader.h /
// SceneSettings.h, implemented over the C ABI of include/lightgraph.h.  OpenVDB
// types in the reference's signatures become DenseBrickGrid and Vec3d.
#ifndef LIGHTGRAPH_HOST_HPP_
#define LIGHTGRAPH_HOST_HPP_

#include <algorithm>
#include <cstdio>
#include <sstream>
#include <iostream>
#include <fstream>
#include <chrono>
#include <sys/stat.h>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "lightgraph.h"

namespace lightgraph {

struct Vec3d {
    double x = 0.0, y = 0.0, z = 0.0;
    Vec3d() = default;
    Vec3d(double a, double b, double c) : x(a), y(b), z(c) {}
    explicit Vec3d(double a) : x(a), y(a), z(a) {}
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    void to(double out[3]) const {
        out[0] = x;
        out[1] = y;
        out[2] = z;
    }
};

// Errors of the C ABI surface as exceptions (the reference throws OPENVDB_THROW
// ValueError, Renderer.h:288, main.cc:104,132).
struct Error : std::runtime_error {
    int32_t code;
    Error(const std::string &what, int32_t rc) : std::runtime_error(what), code(rc) {}
};

inline void check(int32_t rc, const char *what) {
    if (rc != LG_OK) throw Error(std::string(what) + ": " + lg_last_error(), rc);
}
inline void check_host(int32_t rc, const char *what) {
    if (rc != LG_OK) throw Error(std::string(what) + ": " + lg_host_last_error(), rc);
}

// One device and its stream.
class Context {
   public:
    explicit Context(int device = 0, uint64_t hbm_budget = 0) { check(lg_ctx_create(device, hbm_budget, &ctx_), "lg_ctx_create"); }
    ~Context() { lg_ctx_destroy(ctx_); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    lg_ctx *get() const { return ctx_; }
    void sync() const { check(lg_sync(ctx_), "lg_sync"); }

   private:
    lg_ctx *ctx_ = nullptr;
};

// Replaces openvdb::FloatGrid (+ its GridSampler): a dense-brick device grid.
class DenseBrickGrid {
   public:
    DenseBrickGrid(Context &ctx, lg_grid *g) : ctx_(&ctx), g_(g) {}
    DenseBrickGrid(DenseBrickGrid &&o) noexcept : ctx_(o.ctx_), g_(o.g_) { o.g_ = nullptr; }
    DenseBrickGrid &operator=(DenseBrickGrid &&o) noexcept {
        std::swap(ctx_, o.ctx_);
        std::swap(g_, o.g_);
        return *this;
    }
    ~DenseBrickGrid() {
        if (g_) lg_grid_destroy(ctx_->get(), g_);
    }
    // dense z-y-x float array of nz*ny*nx voxels (index (x,y,z) = voxel centre)
    static DenseBrickGrid fromDense(Context &ctx, int nx, int ny, int nz, const float *data, double voxel_size = 1.0,
                                    Vec3d origin = Vec3d(), int32_t dtype = LG_F32) {
        lg_grid_desc d{};
        d.nx = nx;
        d.ny = ny;
        d.nz = nz;
        d.voxel_size = voxel_size;
        origin.to(d.origin);
        d.dtype = dtype;
        lg_grid *g = nullptr;
        check(lg_grid_create(ctx.get(), &d, data, &g), "lg_grid_create");
        return DenseBrickGrid(ctx, g);
    }
    // procedural scene (density and, for the plume, temperature)
    static std::pair<DenseBrickGrid, std::unique_ptr<DenseBrickGrid>> procedural(Context &ctx,
                                                                                 const lg_scene_spec &spec) {
        lg_grid *g = nullptr, *t = nullptr;
        check(lg_grid_generate(ctx.get(), &spec, &g, spec.with_temperature ? &t : nullptr), "lg_grid_generate");
        std::unique_ptr<DenseBrickGrid> temp;
        if (t) temp.reset(new DenseBrickGrid(ctx, t));
        return {DenseBrickGrid(ctx, g), std::move(temp)};
    }
    lg_grid_info info() const {
        lg_grid_info i{};
        check(lg_grid_info_get(g_, &i), "lg_grid_info_get");
        return i;
    }
    const lg_grid *get() const { return g_; }
    Context &context() const { return *ctx_; }

   private:
    Context *ctx_;
    lg_grid *g_;
};

// Environment map + its FP64 sampling CDF (EnvLight, Light.h:132-231).
class EnvMap {
   public:
    EnvMap(Context &ctx, int width, int height, const float *rgb) : ctx_(&ctx) {
        check(lg_envmap_create(ctx.get(), width, height, rgb, &e_), "lg_envmap_create");
    }
    ~EnvMap() {
        if (e_) lg_envmap_destroy(ctx_->get(), e_);
    }
    EnvMap(const EnvMap &) = delete;
    EnvMap &operator=(const EnvMap &) = delete;
    const lg_envmap *get() const { return e_; }

   private:
    Context *ctx_;
    lg_envmap *e_ = nullptr;
};

// Shader (Shader.h:100-274): same setters and defaults.
class Shader {
   public:
    // the coefficient defaults of Shader::Shader (Shader.h:128-130)
    Shader() : absorption_(0.1), scattering_(1.5), volume_color_(1.0) {
        check_host(lg_shader_default(&s_), "lg_shader_default");
    }
    void setDensityScale(double v) { s_.density_scale = v; }
    void setDensityMax(double v) { s_.density_max = v; }
    void setShadowDensityScale(double v) { s_.shadow_density_scale = v; }
    void setScatterDensityScale(double v) { s_.scatter_density_scale = v; }
    void setScatterDensityMin(double v) { s_.scatter_density_min = v; }
    void setAbsorption(const Vec3d &v) {
        absorption_ = v;
        update();
    }
    void setScattering(const Vec3d &v) {
        scattering_ = v;
        update();
    }
    void setVolumeColor(const Vec3d &v) {
        volume_color_ = v;
        update();
    }
    void setEmissionScale(double v) { s_.emission_scale = v; }
    void setScatterScale(double v) { s_.scatter_scale = v; }
    void setScatterDensityMaskPower(double v) { s_.scatter_density_mask_power = v; }
    // parseColorRamp + setEmissionColorRamp (main.cc:280-319)
    void setEmissionColorRamp(const std::string &text) { check_host(lg_shader_parse_ramp(&s_, text.c_str()), "ramp"); }
    const lg_shader &params() const { return s_; }

   private:
    void update() {
        double a[3], s[3], c[3];
        absorption_.to(a);
        scattering_.to(s);
        volume_color_.to(c);
        check_host(lg_shader_set_coefficients(&s_, a, s, c), "lg_shader_set_coefficients");
    }
    lg_shader s_{};
    Vec3d absorption_, scattering_, volume_color_;
};

// SceneSettings (SceneSettings.h:21-63): the same fields and defaults; lights are
// C-ABI light records.  The last block holds this library's additions: graph size
// and radius (the reference derives them from its dart thrower) and the scene.
struct SceneSettings {
    std::string camera = "perspective";
    float aperture = 36.0f, focal = 50.0f, znear = 1.0e-3f, zfar = FLT_MAX;
    Vec3d camRotate = Vec3d(0.0), camTranslate = Vec3d(0.0);
    double camYRotate = 0.0;
    uint64_t samples = 1;
    std::vector<lg_light> light;
    double cutoff = 0.005;
    double step[2] = {1.0, 3.0};
    uint64_t width = 960, height = 540;
    std::string compression = "zip";
    size_t iter = 10;
    size_t scatterGridResFactor = 4;
    size_t ptsPerKernel = 3;
    int threads = 0;
    bool verbose = false;
    bool geoDump = false;
    std::string geoName = "volume";  // file-name stem of the geometry dumps (the vdb file's stem)
    // additions
    uint32_t points = 750;   // graph points per LightGraph iteration (TARGET_PC_SIZE, LightGraph.h:26)
    double radius = 0.0;     // edge radius in world units; 0: derived from the bbox (see LightGraph)
    int32_t maxConnect = 12; // GRAPH_MAX_CONNECT (main.cc:62); <= 0: uncapped
    uint64_t seed = 854123;  // RandGen base seed (RandGen.h:28-29)
    uint64_t envSeed = 4564446;
    // "density": the density-weighted sampler (points per iteration = `points`);
    // "dart": the reference's dart thrower (LightGraph::scatterPoints), target
    // `points`, radius 1.7 x the thrower's final minimum distance per iteration
    std::string sampler = "density";
    uint32_t scatterCount = 1000000;  // PC_SCATTER_COUNT (main.cc:61): throws per iteration

    lg_march_params march() const {
        lg_march_params p{};
        check_host(lg_march_params_default(&p), "lg_march_params_default");
        p.cutoff = cutoff;
        p.step[0] = step[0];
        p.step[1] = step[1];
        return p;
    }
};

// LightGraph (LightGraph.h:40-196): operator()(begin, end) runs the iterations
// [begin, end) as one batched solve and stores each iteration's points and
// diffused radiance in pointCloudGather[i] / diffusePointGather[i], the output
// slots of the reference's functor.
class LightGraph {
   public:
    LightGraph(const DenseBrickGrid &grid, const DenseBrickGrid *temperature, const SceneSettings &opts,
               const Shader &shader, std::vector<std::vector<Vec3d>> &pointCloudGather,
               std::vector<std::vector<Vec3d>> &diffusePointGather)
        : grid_(grid), temp_(temperature), opts_(opts), shader_(shader), points_(pointCloudGather),
          diffuse_(diffusePointGather) {}

    // Edge radius: the option, else CONNECT_SEARCH_RADIUS_MULT x MIN_DIST_MULT x the
    // mean side of the active bbox (LightGraph.h:24, main.cc:60,512-516), scaled by
    // (750 / points)^(1/3) so the mean degree stays near the reference's.
    double radius() const {
        if (opts_.radius > 0.0) return opts_.radius;
        const lg_grid_info i = grid_.info();
        const double side = ((i.bbox_max[0] - i.bbox_min[0]) + (i.bbox_max[1] - i.bbox_min[1]) +
                             (i.bbox_max[2] - i.bbox_min[2])) * i.desc.voxel_size / 3.0;
        return 1.7 * 0.13 * side * std::cbrt(750.0 / std::max<uint32_t>(opts_.points, 1));
    }

    // One iteration's dart-thrown points (world, acceptance order) and the
    // thrower's final minimum distance (mMinDistVec[i] after scatterPoints).
    struct DartSet {
        std::vector<float> pts;
        float min_dist = 0.0f;
        size_t size() const { return pts.size() / 3; }
        // searchRadiusConnect = mMinDistVec[i] * CONNECT_SEARCH_RADIUS_MULT (LightGraph.h:169)
        double radius() const { return (double)(min_dist * 1.7f); }
    };

    // LightGraph::scatterPoints for iterations [begin, end): the darts fall in
    // the active-voxel box (main.cc:512) and start from MIN_DIST_MULT x its mean
    // side (main.cc:514-518), in the reference's float arithmetic.
    std::vector<DartSet> dartThrow(size_t begin, size_t end) const {
        Context &ctx = grid_.context();
        int32_t lo[3], hi[3];
        check(lg_grid_active_bbox(ctx.get(), grid_.get(), lo, hi), "lg_grid_active_bbox");
        const lg_grid_info gi = grid_.info();
        double bmin[3], bmax[3];
        float side[3];
        for (int c = 0; c < 3; ++c) {
            bmin[c] = gi.desc.origin[c] + gi.desc.voxel_size * lo[c];
            bmax[c] = gi.desc.origin[c] + gi.desc.voxel_size * hi[c];
            side[c] = (float)(bmax[c] - bmin[c]);
        }
        const float avgLength = (side[0] + side[1] + side[2]) / 3.0f;
        const float minDist = 0.13f * avgLength;
        const size_t count = end - begin;
        const int32_t cap = (int32_t)opts_.points + 1;
        std::vector<uint32_t> iters(count);
        std::vector<float> md(count, minDist);
        std::vector<int32_t> counts(count, 0);
        for (size_t k = 0; k < count; ++k) iters[k] = (uint32_t)(begin + k);
        void *d = nullptr;
        check(lg_malloc(ctx.get(), count * (size_t)cap * 3 * sizeof(float), &d), "lg_malloc");
        std::unique_ptr<void, std::function<void(void *)>> dguard(d, [&ctx](void *p) { lg_free(ctx.get(), p); });
        check(lg_dart_throw(ctx.get(), grid_.get(), bmin, bmax, (int32_t)count, iters.data(), opts_.scatterCount,
                            opts_.points, md.data(), (float *)d, cap, counts.data()),
              "lg_dart_throw");
        std::vector<float> all(count * (size_t)cap * 3);
        check(lg_memcpy_d2h(ctx.get(), all.data(), d, all.size() * sizeof(float)), "d2h dart points");
        std::vector<DartSet> out(count);
        for (size_t k = 0; k < count; ++k) {
            const size_t n = (size_t)std::min(counts[k], cap);
            const float *p = all.data() + k * (size_t)cap * 3;
            out[k].pts.assign(p, p + 3 * n);
            out[k].min_dist = md[k];
            if (opts_.verbose) std::cout << "PointCloud Size: " << counts[k] << std::endl;
        }
        return out;
    }

    void operator()(size_t begin, size_t end) const {
        if (end <= begin) return;
        if (opts_.sampler == "dart") {
            solveDart(begin, end);
            return;
        }
        Context &ctx = grid_.context();
        const size_t count = end - begin;
        const size_t total = count * opts_.points;
        lg_solver *solver = nullptr;
        check(lg_solver_create(ctx.get(), (int32_t)total, 0, 0, &solver), "lg_solver_create");
        std::unique_ptr<lg_solver, std::function<void(lg_solver *)>> guard(
            solver, [&ctx](lg_solver *s) { lg_solver_destroy(ctx.get(), s); });
        std::vector<lg_batch_item> items(count);
        for (size_t k = 0; k < count; ++k)
            items[k] = lg_batch_item{grid_.get(), temp_ ? temp_->get() : nullptr, (int32_t)opts_.points,
                                     (uint32_t)(begin + k)};
        lg_solve_params sp{};
        sp.max_connect = opts_.maxConnect;
        sp.radius = radius();
        sp.seed = opts_.seed;
        sp.env_seed = opts_.envSeed;
        sp.timing = opts_.verbose ? 1 : 0;  // stage events (the stages then run one after another)
        const lg_march_params mp = opts_.march();
        check(lg_solve_batch(ctx.get(), solver, items.data(), (int32_t)count, &shader_.params(),
                             opts_.light.data(), (int32_t)opts_.light.size(), &mp, &sp),
              "lg_solve_batch");
        if (opts_.verbose) printStageTimes(solver, count);
        lg_solver_buffers b{};
        check(lg_solver_get(solver, &b), "lg_solver_get");
        std::vector<float> pts(total * 3), dif(total * 3);
        check(lg_memcpy_d2h(ctx.get(), pts.data(), b.d_points, pts.size() * 4), "d2h points");
        check(lg_memcpy_d2h(ctx.get(), dif.data(), b.d_diffuse, dif.size() * 4), "d2h diffuse");
        for (size_t k = 0; k < count; ++k) {
            std::vector<Vec3d> &P = points_[begin + k], &D = diffuse_[begin + k];
            P.resize(opts_.points);
            D.resize(opts_.points);
            for (uint32_t v = 0; v < opts_.points; ++v) {
                const size_t o = 3 * (k * opts_.points + v);
                P[v] = Vec3d(pts[o], pts[o + 1], pts[o + 2]);
                D[v] = Vec3d(dif[o], dif[o + 1], dif[o + 2]);
            }
        }
        if (opts_.geoDump)
            for (size_t k = 0; k < count; ++k) dumpGeo(begin + k);
    }

    // The per-stage lines of -v (LightGraph.h:268-272, 312-316, 412-416,
    // 499-503, 536-540, 581-585), once for the whole batch of solves: the
    // stages run as one launch sequence for every solve of the range.  The
    // shortest paths and the diffusion are one stage here (the closures of the
    // components carry the diffusion with them), reported on the first line.
    void printStageTimes(const lg_solver *solver, size_t count) const {
        lg_solve_timing t{};
        check(lg_solver_timing(solver, &t), "lg_solver_timing");
        const std::string of = count > 1 ? " (" + std::to_string(count) + " solves)" : "";
        std::cout << "Point Cloud Creation: " << t.sample_ms * 1e-3 << " sec" << of << std::endl;
        std::cout << "Graph Creation: " << t.graph_ms * 1e-3 << " sec" << of << std::endl;
        std::cout << "Raymarch Graph Edges: " << t.march_ms * 1e-3 << " sec" << of << std::endl;
        std::cout << "Vertex Lighting Calculation: " << t.light_ms * 1e-3 << " sec" << of << std::endl;
        std::cout << "Shortest Path Finding: " << (t.fw_ms + t.diffuse_ms) * 1e-3 << " sec" << of
                  << " (with the radiance diffusion)" << std::endl;
        std::cout << "Radiance Diffusion: 0 sec (done with the shortest paths)" << std::endl;
    }

    // LightGraph::outputDebugGeo (LightGraph.h:604-696): five comma-separated
    // text files per iteration under geoDump/: points, edges (i < j), the
    // all-pairs transmittance of each edge after diffusion's column
    // normalisation, vertex light and diffused radiance.  The iteration is solved
    // again alone with the dense optical-depth output (the batched solve gives
    // the same numbers); transmittance T_c = exp(extinction_c D) per channel,
    // normalised by max(1, column sum) as diffuseLight does (LightGraph.h:563-571).
    // With `dart`, the iteration's dart-thrown points and radius are used.
    void dumpGeo(size_t iter, const DartSet *dart = nullptr) const {
        const auto t0 = std::chrono::steady_clock::now();
        Context &ctx = grid_.context();
        const uint32_t n = dart ? (uint32_t)dart->size() : opts_.points;
        if (n == 0) return;
        lg_solver *solver = nullptr;
        check(lg_solver_create(ctx.get(), (int32_t)n, 0, LG_SOLVER_DENSE_D, &solver), "lg_solver_create");
        std::unique_ptr<lg_solver, std::function<void(lg_solver *)>> guard(
            solver, [&ctx](lg_solver *s) { lg_solver_destroy(ctx.get(), s); });
        lg_solve_params sp{};
        sp.n = (int32_t)n;
        sp.max_connect = opts_.maxConnect;
        sp.radius = dart ? dart->radius() : radius();
        sp.seed = opts_.seed;
        sp.env_seed = opts_.envSeed;
        sp.solve = (uint32_t)iter;
        sp.flags = LG_SOLVE_DENSE;
        if (dart) {
            lg_solver_buffers pb{};
            check(lg_solver_get(solver, &pb), "lg_solver_get");
            check(lg_memcpy_h2d(ctx.get(), pb.d_points, dart->pts.data(), dart->pts.size() * sizeof(float)),
                  "h2d points");
            sp.external_points = 1;
        }
        const lg_march_params mp = opts_.march();
        check(lg_solve(ctx.get(), solver, grid_.get(), temp_ ? temp_->get() : nullptr, &shader_.params(),
                       opts_.light.data(), (int32_t)opts_.light.size(), &mp, &sp),
              "lg_solve");
        lg_solver_buffers b{};
        check(lg_solver_get(solver, &b), "lg_solver_get");
        if (!b.d_D || b.ld < (int32_t)n || b.n != (int32_t)n)
            throw std::runtime_error("geoDump: the solve returned no dense optical depths");
        std::vector<float> pts((size_t)n * 3), L((size_t)n * 3), dif((size_t)n * 3), Dm((size_t)n * b.ld);
        std::vector<int32_t> edges((size_t)b.n_edges * 2);
        check(lg_memcpy_d2h(ctx.get(), pts.data(), b.d_points, pts.size() * 4), "d2h points");
        check(lg_memcpy_d2h(ctx.get(), L.data(), b.d_L, L.size() * 4), "d2h light");
        check(lg_memcpy_d2h(ctx.get(), dif.data(), b.d_diffuse, dif.size() * 4), "d2h diffuse");
        check(lg_memcpy_d2h(ctx.get(), Dm.data(), b.d_D, Dm.size() * 4), "d2h D");
        if (b.n_edges > 0) check(lg_memcpy_d2h(ctx.get(), edges.data(), b.d_edges, edges.size() * 4), "d2h edges");
        const double *ext = shader_.params().extinction;
        // column normalisation of the transmittance matrix (D is symmetric)
        std::vector<double> accum((size_t)n * 3, 0.0);
        for (uint32_t j = 0; j < n; ++j)
            for (uint32_t i = 0; i < n; ++i) {
                const float d = Dm[(size_t)i * b.ld + j];
                if (!(d < INFINITY)) continue;
                for (int c = 0; c < 3; ++c) accum[3 * (size_t)j + c] += std::exp(ext[c] * (double)d);
            }
        for (auto &a : accum) a = std::max(a, 1.0);
        struct stat st_dir {};
        if (stat("geoDump", &st_dir) != 0 && mkdir("geoDump", 0755) == 0) std::cout << "Success" << "\n";
        char tag[16];
        std::snprintf(tag, sizeof(tag), "%04zu", iter);
        auto path = [&](const char *kind) { return "geoDump/" + opts_.geoName + "_" + kind + "_" + tag + ".txt"; };
        auto vec3 = [&](const char *kind, const std::vector<float> &a) {
            std::ofstream f(path(kind));
            for (uint32_t v = 0; v < n; ++v)
                f << (double)a[3 * v] << "," << (double)a[3 * v + 1] << "," << (double)a[3 * v + 2] << "\n";
        };
        vec3("pc", pts);
        {
            std::ofstream fe(path("edge")), fm(path("marching"));
            for (int32_t e = 0; e < b.n_edges; ++e) {
                const int32_t i = edges[2 * e], j = edges[2 * e + 1];
                fe << i << "," << j << "\n";
                const float d = Dm[(size_t)i * b.ld + j];
                double t[3] = {0.0, 0.0, 0.0};
                if (d < INFINITY)
                    for (int c = 0; c < 3; ++c) t[c] = std::exp(ext[c] * (double)d) / accum[3 * (size_t)j + c];
                fm << t[0] << "," << t[1] << "," << t[2] << "\n";
            }
        }
        vec3("light", L);
        vec3("diffuse", dif);
        if (opts_.verbose)
            std::cout << "Geo Dump to Disk: "
                      << std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() << " sec"
                      << std::endl;
    }

   private:
    // The reference's per-iteration loop (LightGraph.h:161-194) on dart-thrown
    // points: each iteration is solved alone with its own point count and radius.
    void solveDart(size_t begin, size_t end) const {
        Context &ctx = grid_.context();
        const std::vector<DartSet> sets = dartThrow(begin, end);
        lg_solver *solver = nullptr;
        check(lg_solver_create(ctx.get(), (int32_t)opts_.points + 1, 0, 0, &solver), "lg_solver_create");
        std::unique_ptr<lg_solver, std::function<void(lg_solver *)>> guard(
            solver, [&ctx](lg_solver *s) { lg_solver_destroy(ctx.get(), s); });
        lg_solver_buffers b{};
        check(lg_solver_get(solver, &b), "lg_solver_get");
        const lg_march_params mp = opts_.march();
        for (size_t k = 0; k < sets.size(); ++k) {
            const DartSet &set = sets[k];
            const size_t n = set.size();
            std::vector<Vec3d> &P = points_[begin + k], &D = diffuse_[begin + k];
            P.resize(n);
            D.resize(n);
            if (n == 0) continue;
            check(lg_memcpy_h2d(ctx.get(), b.d_points, set.pts.data(), set.pts.size() * sizeof(float)), "h2d points");
            lg_solve_params sp{};
            sp.n = (int32_t)n;
            sp.max_connect = opts_.maxConnect;
            sp.radius = set.radius();
            sp.seed = opts_.seed;
            sp.env_seed = opts_.envSeed;
            sp.solve = (uint32_t)(begin + k);
            sp.external_points = 1;
            sp.timing = opts_.verbose ? 1 : 0;
            check(lg_solve(ctx.get(), solver, grid_.get(), temp_ ? temp_->get() : nullptr, &shader_.params(),
                           opts_.light.data(), (int32_t)opts_.light.size(), &mp, &sp),
                  "lg_solve");
            if (opts_.verbose) printStageTimes(solver, 1);
            std::vector<float> dif(3 * n);
            check(lg_memcpy_d2h(ctx.get(), dif.data(), b.d_diffuse, dif.size() * 4), "d2h diffuse");
            for (size_t v = 0; v < n; ++v) {
                P[v] = Vec3d(set.pts[3 * v], set.pts[3 * v + 1], set.pts[3 * v + 2]);
                D[v] = Vec3d(dif[3 * v], dif[3 * v + 1], dif[3 * v + 2]);
            }
            if (opts_.geoDump) dumpGeo(begin + k, &set);
        }
    }

    const DenseBrickGrid &grid_;
    const DenseBrickGrid *temp_;
    const SceneSettings &opts_;
    const Shader &shader_;
    std::vector<std::vector<Vec3d>> &points_;
    std::vector<std::vector<Vec3d>> &diffuse_;
};

// ScatterGrid (ScatterGrid.h): coarse topology from the density grid and the
// Gaussian splat of the gathered points.
class ScatterGrid {
   public:
    ScatterGrid(const DenseBrickGrid &density, const SceneSettings &opts) : ctx_(&density.context()) {
        check(lg_scatter_grid_create(ctx_->get(), density.get(), (int32_t)opts.scatterGridResFactor, &sg_),
              "lg_scatter_grid_create");
    }
    ~ScatterGrid() { lg_scatter_grid_destroy(ctx_->get(), &sg_); }
    ScatterGrid(const ScatterGrid &) = delete;
    ScatterGrid &operator=(const ScatterGrid &) = delete;
    // the splat of main.cc:571-586 with kernel count ptsPerKernel * iter
    void splat(const std::vector<Vec3d> &points, const std::vector<Vec3d> &radiance, size_t kernel_total) {
        const size_t n = points.size();
        if (n == 0) return;
        std::vector<float> p(3 * n), r(3 * n);
        for (size_t i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) {
                p[3 * i + c] = (float)points[i][c];
                r[3 * i + c] = (float)radiance[i][c];
            }
        double radius = 0.0, vs = 0.0;
        check(lg_splat_radius(p.data(), (int64_t)n, (int64_t)kernel_total, &radius, &vs), "lg_splat_radius");
        void *dp = nullptr, *dr = nullptr;
        check(lg_malloc(ctx_->get(), p.size() * 4, &dp), "lg_malloc");
        check(lg_malloc(ctx_->get(), r.size() * 4, &dr), "lg_malloc");
        check(lg_memcpy_h2d(ctx_->get(), dp, p.data(), p.size() * 4), "h2d");
        check(lg_memcpy_h2d(ctx_->get(), dr, r.data(), r.size() * 4), "h2d");
        const int32_t rc = lg_splat(ctx_->get(), (const float *)dp, (const float *)dr, (int64_t)n, radius, &sg_);
        lg_free(ctx_->get(), dp);
        lg_free(ctx_->get(), dr);
        check(rc, "lg_splat");
        radius_ = radius;
    }
    const lg_scatter_grid *get() const { return &sg_; }
    double radius() const { return radius_; }

   private:
    Context *ctx_;
    lg_scatter_grid sg_{};
    double radius_ = 0.0;
};

// Film (Renderer.h:68-121): host RGBA float image.
struct Film {
    struct RGBA {
        float r, g, b, a;
    };
    size_t w, h;
    std::vector<RGBA> px;
    Film(size_t width, size_t height) : w(width), h(height), px(width * height, RGBA{0, 0, 0, 0}) {}
    size_t width() const { return w; }
    size_t height() const { return h; }
    const RGBA *pixels() const { return px.data(); }
};

// VolumeRender (Renderer.h:474-693) with the PerspectiveCamera of main.cc:96-102.
class VolumeRender {
   public:
    VolumeRender(const DenseBrickGrid &density, const ScatterGrid *scatter, const DenseBrickGrid *temperature,
                 const SceneSettings &opts, const Shader &shader)
        : density_(density), scatter_(scatter), temp_(temperature), opts_(opts), shader_(shader) {}

    // renders the four AOVs: beauty, direct, scatter, emission
    void render(Film &rgb, Film &direct, Film &scatter, Film &emission) const {
        Context &ctx = density_.context();
        double rot[3], tr[3];
        opts_.camRotate.to(rot);
        opts_.camTranslate.to(tr);
        lg_camera cam{};
        check(lg_camera_setup(rot, tr, opts_.camYRotate, opts_.focal, opts_.aperture, opts_.znear, opts_.zfar,
                              (int32_t)opts_.width, (int32_t)opts_.height, &cam),
              "lg_camera_setup");
        const size_t bytes = opts_.width * opts_.height * 16;
        void *f[4] = {nullptr, nullptr, nullptr, nullptr};
        for (auto &p : f) check(lg_malloc(ctx.get(), bytes, &p), "lg_malloc");
        lg_films films{(float *)f[0], (float *)f[1], (float *)f[2], (float *)f[3]};
        const lg_march_params mp = opts_.march();
        int32_t rc = lg_render(ctx.get(), density_.get(), scatter_ ? scatter_->get() : nullptr,
                               temp_ ? temp_->get() : nullptr, &cam, &shader_.params(), opts_.light.data(),
                               (int32_t)opts_.light.size(), &mp, (int32_t)opts_.samples, 1, opts_.envSeed, &films);
        Film *out[4] = {&rgb, &direct, &scatter, &emission};
        for (int k = 0; k < 4 && rc == LG_OK; ++k) {
            *out[k] = Film(opts_.width, opts_.height);
            rc = lg_memcpy_d2h(ctx.get(), out[k]->px.data(), f[k], bytes);
        }
        for (auto p : f) lg_free(ctx.get(), p);
        check(rc, "lg_render");
    }

   private:
    const DenseBrickGrid &density_;
    const ScatterGrid *scatter_;
    const DenseBrickGrid *temp_;
    const SceneSettings &opts_;
    const Shader &shader_;
};

// EXR I/O (Image.h:21-92 for the 13-channel AOV file; env maps, Sampling.h:175-189):
// scanline files, NONE / ZIPS / ZIP compression, FLOAT (written) or HALF/FLOAT (read).
void saveEXR(const std::string &fname, const Film &rgb, const Film &direct, const Film &scatter,
             const Film &emission, const SceneSettings &opts);
// Reads the R, G, B channels of a scanline EXR into row-major RGB float (row 0 = top).
std::vector<float> readEXR_RGB(const std::string &fname, int *width, int *height);

}  // namespace lightgraph

#endif  // LIGHTGRAPH_HOST_HPP_
