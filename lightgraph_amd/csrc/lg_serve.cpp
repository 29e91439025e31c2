// Serving form: a request loop over render jobs (volume -> LightGraph solves ->
// splat -> render) with continuous batching and HBM-budget admission.
//
// * Jobs are queued on submit.  Worker threads, each with its own context and
//   HIP stream (so their work overlaps on the device), take a batch when they
//   are free: the oldest admissible job plus the queued jobs that can share its
//   solve launch (same brick type, shader, lights, march and graph parameters),
//   as long as the batch's memory fits the server budget.
// * A job is admissible when its estimated device memory fits what is left of
//   the budget; a worker that cannot fit the head job evicts its own idle cached
//   volumes first.  Jobs larger than the whole budget are refused at submit.
// * Volumes (and their coarse scatter topology) stay resident in a per-worker
//   cache keyed by the scene description, so repeated volumes are not rebuilt.
// * All solves of a batch run as one lg_solve_batch; splat and render follow per
//   job.  Every job's films equal those of the same job run alone: the solve is
//   per-graph arithmetic and the render is per job.
// This file is host code over the public C ABI only.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "lightgraph.h"

namespace lg {
void set_error(const char *fmt, ...);  // the library's thread-local error message (lg_runtime.hip)
}

struct lg_server;

namespace {

struct Job {
    uint64_t id = 0;
    lg_job spec{};
    float *films[4] = {nullptr, nullptr, nullptr, nullptr};
    uint64_t need = 0;  // reserved bytes while it runs (transient part)
    uint64_t volume_bytes = 0;
    int32_t status = LG_OK;
    std::string error;
    double t_submit = 0, t_admit = 0, t_start = 0, t_done = 0;
    int32_t batch = 0;
    bool done = false;
};

uint64_t volume_bytes(const lg_scene_spec &s) {
    const uint64_t nb = (uint64_t)((s.n + 7) / 8) + 2;
    const uint64_t elem = s.dtype == LG_F32 ? 4 : 2;
    uint64_t b = nb * nb * nb * (512 * elem + 1);
    if (s.with_temperature) b += nb * nb * nb * (512 * 4 + 1);
    return b;
}

// transient device memory of one job besides its volume (scatter grid, films,
// render workspace, solver workspace with room for a component holding a
// quarter of the points)
uint64_t job_bytes(const lg_job &j) {
    const uint64_t V = (uint64_t)std::max(j.points, 1) * (uint64_t)std::max(j.iterations, 1);
    const uint64_t nc = (uint64_t)(j.scene.n / std::max(j.scatter_res_factor, 1)) + 6;
    const uint64_t scatter = nc * nc * nc * 13;
    const uint64_t films = (uint64_t)j.camera.width * (uint64_t)j.camera.height * 64;
    // a camera ray crosses at most sqrt(3) n voxels; jittered steps are >= 0.75 h
    const double h = j.march.step[0] > 0 ? j.march.step[0] : 1.0;
    const int64_t per_ray = (int64_t)std::ceil(1.7320508075688772 * (double)j.scene.n / (0.75 * h)) + 2;
    uint64_t render = 0;
    if (lg_render_workspace(std::max(j.camera.width, 1), std::max(j.camera.height, 1), std::max(j.samples, 1),
                            std::min(std::max(j.n_lights, 0), LG_MAX_LIGHTS), per_ray, &render) != LG_OK)
        render = 0;
    const uint64_t big = V / 4 + 128;
    const uint64_t solver = V * 2048 + big * big * 4;
    return scatter + films + render + solver;
}

double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// jobs that can share one lg_solve_batch
bool same_solve_setup(const lg_job &a, const lg_job &b) {
    return a.scene.dtype == b.scene.dtype && a.solve_flags == b.solve_flags && a.max_connect == b.max_connect &&
           a.radius == b.radius &&
           a.seed == b.seed && a.env_seed == b.env_seed && a.n_lights == b.n_lights &&
           std::memcmp(&a.shader, &b.shader, sizeof(lg_shader)) == 0 &&
           std::memcmp(&a.march, &b.march, sizeof(lg_march_params)) == 0 &&
           std::memcmp(a.lights, b.lights, sizeof(lg_light) * (size_t)std::max(a.n_lights, 0)) == 0;
}

struct CachedVolume {
    lg_grid *density = nullptr, *temperature = nullptr;
    lg_scatter_grid sg{};
    int32_t sg_factor = 0;
    uint64_t bytes = 0;
    int users = 0;
    uint64_t last_use = 0;
};

}  // namespace

struct lg_server {
    int32_t device = 0;
    uint64_t budget = 0;
    int32_t max_batch = 8;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job *> queue;
    std::map<uint64_t, std::unique_ptr<Job>> jobs;
    uint64_t next_id = 1;
    uint64_t reserved = 0, peak_reserved = 0;
    uint64_t completed = 0, failed = 0, batches = 0, batched_jobs = 0, solves = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    std::string create_error;
};

namespace {

std::string key_of(const lg_scene_spec &s) { return std::string(reinterpret_cast<const char *>(&s), sizeof(s)); }

struct Worker {
    lg_server *srv;
    lg_ctx *ctx = nullptr;
    lg_solver *solver = nullptr;
    int32_t solver_n = 0;
    std::map<std::string, CachedVolume> cache;  // guarded by srv->mu for bytes/users
    uint64_t tick = 0;
    void *films[4] = {nullptr, nullptr, nullptr, nullptr};
    size_t film_bytes_f[4] = {0, 0, 0, 0};

    // called with srv->mu held: evict idle cached volumes until `need` fits
    bool make_room(uint64_t need) {
        while (srv->reserved + need > srv->budget) {
            auto victim = cache.end();
            for (auto it = cache.begin(); it != cache.end(); ++it)
                if (it->second.users == 0 && (victim == cache.end() || it->second.last_use < victim->second.last_use))
                    victim = it;
            if (victim == cache.end()) return false;
            CachedVolume &v = victim->second;
            if (v.sg.d_rgb) lg_scatter_grid_destroy(ctx, &v.sg);
            if (v.density) lg_grid_destroy(ctx, v.density);
            if (v.temperature) lg_grid_destroy(ctx, v.temperature);
            srv->reserved -= v.bytes;
            cache.erase(victim);
        }
        return true;
    }

    // bytes a job adds to the reservation on this worker (its volume counts once
    // per batch and not at all when it is already resident here)
    uint64_t cost(const Job &j, const std::vector<Job *> &batch) const {
        const std::string k = key_of(j.spec.scene);
        bool have = cache.find(k) != cache.end();
        for (const Job *b : batch) have = have || key_of(b->spec.scene) == k;
        return j.need + (have ? 0 : j.volume_bytes);
    }

    // takes the next batch (blocks); empty when the server stops
    std::vector<Job *> take() {
        std::unique_lock<std::mutex> lk(srv->mu);
        for (;;) {
            if (srv->stop) return {};
            if (!srv->queue.empty()) {
                Job *head = srv->queue.front();
                uint64_t total = cost(*head, {});
                if (srv->reserved + total <= srv->budget || make_room(total)) {
                    std::vector<Job *> batch = {head};
                    srv->queue.pop_front();
                    for (auto it = srv->queue.begin(); it != srv->queue.end() && (int)batch.size() < srv->max_batch;) {
                        Job *j = *it;
                        const uint64_t c = cost(*j, batch);
                        if (same_solve_setup(head->spec, j->spec) && srv->reserved + total + c <= srv->budget) {
                            batch.push_back(j);
                            total += c;
                            it = srv->queue.erase(it);
                        } else {
                            ++it;
                        }
                    }
                    srv->reserved += total;
                    srv->peak_reserved = std::max(srv->peak_reserved, srv->reserved);
                    const double t = now_s(srv->t0);
                    for (Job *j : batch) {
                        j->t_admit = t;
                        j->batch = (int32_t)batch.size();
                        // a volume built for this job is charged to the cache from here on
                        auto c = cache.find(key_of(j->spec.scene));
                        if (c == cache.end()) {
                            CachedVolume &v = cache[key_of(j->spec.scene)];
                            v.bytes = j->volume_bytes;
                        }
                        cache[key_of(j->spec.scene)].users++;
                        cache[key_of(j->spec.scene)].last_use = ++tick;
                    }
                    ++srv->batches;
                    srv->batched_jobs += batch.size();
                    return batch;
                }
            }
            srv->cv_work.wait(lk);
        }
    }

    void finish(Job *j, int32_t rc, const char *msg) {
        std::lock_guard<std::mutex> lk(srv->mu);
        j->status = rc;
        if (rc != LG_OK) j->error = msg ? msg : "error";
        j->t_done = now_s(srv->t0);
        j->done = true;
        srv->reserved -= j->need;
        auto it = cache.find(key_of(j->spec.scene));
        if (it != cache.end()) {
            it->second.users--;
            if (rc != LG_OK && !it->second.density && it->second.users == 0) {  // never built
                srv->reserved -= it->second.bytes;
                cache.erase(it);
            }
        }
        if (rc == LG_OK) {
            ++srv->completed;
            srv->solves += (uint64_t)std::max(j->spec.iterations, 1);
        } else {
            ++srv->failed;
        }
        srv->cv_done.notify_all();
        srv->cv_work.notify_all();
    }

    int32_t ensure_volume(const lg_job &spec, CachedVolume **out) {
        CachedVolume *v;
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            v = &cache[key_of(spec.scene)];
        }
        if (!v->density) {
            int32_t rc = lg_grid_generate(ctx, &spec.scene, &v->density,
                                          spec.scene.with_temperature ? &v->temperature : nullptr);
            if (rc != LG_OK) return rc;
        }
        if (!v->sg.d_rgb || v->sg_factor != spec.scatter_res_factor) {
            if (v->sg.d_rgb) lg_scatter_grid_destroy(ctx, &v->sg);
            int32_t rc = lg_scatter_grid_create(ctx, v->density, spec.scatter_res_factor, &v->sg);
            if (rc != LG_OK) return rc;
            v->sg_factor = spec.scatter_res_factor;
        }
        *out = v;
        return LG_OK;
    }

    template <typename T>
    int32_t grow(T **p, size_t *cap, size_t bytes) {
        if (bytes <= *cap && *p) return LG_OK;
        if (*p) lg_free(ctx, *p);
        *p = nullptr;
        *cap = 0;
        void *q = nullptr;
        int32_t rc = lg_malloc(ctx, bytes, &q);
        if (rc != LG_OK) return rc;
        *p = static_cast<T *>(q);
        *cap = bytes;
        return LG_OK;
    }

    void run(std::vector<Job *> &batch) {
        const double t = now_s(srv->t0);
        for (Job *j : batch) j->t_start = t;
        // volumes
        std::vector<CachedVolume *> vols(batch.size(), nullptr);
        std::vector<Job *> live;
        std::vector<CachedVolume *> live_vols;
        for (size_t k = 0; k < batch.size(); ++k) {
            const int32_t rc = ensure_volume(batch[k]->spec, &vols[k]);
            if (rc != LG_OK) {
                finish(batch[k], rc, lg_last_error());
                continue;
            }
            live.push_back(batch[k]);
            live_vols.push_back(vols[k]);
        }
        if (live.empty()) return;
        // one solve launch sequence for every iteration of every job
        std::vector<lg_batch_item> items;
        std::vector<size_t> first(live.size());
        int64_t total = 0;
        for (size_t k = 0; k < live.size(); ++k) {
            first[k] = (size_t)total;
            const lg_job &s = live[k]->spec;
            for (int32_t it = 0; it < std::max(s.iterations, 1); ++it) {
                items.push_back(lg_batch_item{live_vols[k]->density, live_vols[k]->temperature, s.points,
                                              (uint32_t)(s.solve_base + it)});
                total += s.points;
            }
        }
        int32_t rc = LG_OK;
        if (total > solver_n) {
            if (solver) lg_solver_destroy(ctx, solver);
            solver = nullptr;
            solver_n = 0;
            rc = lg_solver_create(ctx, (int32_t)total, 0, 0, &solver);
            if (rc == LG_OK) solver_n = (int32_t)total;
        }
        const lg_job &h = live[0]->spec;
        lg_solve_params sp{};
        sp.max_connect = h.max_connect;
        sp.radius = h.radius;
        sp.seed = h.seed;
        sp.env_seed = h.env_seed;
        sp.flags = h.solve_flags & (LG_SOLVE_BFS_ORDER | LG_SOLVE_ND_ORDER);
        if (rc == LG_OK)
            rc = lg_solve_batch(ctx, solver, items.data(), (int32_t)items.size(), &h.shader, h.lights, h.n_lights,
                                &h.march, &sp);
        lg_solver_buffers b{};
        if (rc == LG_OK) rc = lg_solver_get(solver, &b);
        if (rc != LG_OK) {
            const std::string msg = lg_last_error();
            for (Job *j : live) finish(j, rc, msg.c_str());
            return;
        }
        // splat + render per job
        for (size_t k = 0; k < live.size(); ++k) {
            Job *j = live[k];
            const lg_job &s = j->spec;
            const size_t n = (size_t)s.points * (size_t)std::max(s.iterations, 1);
            std::vector<float> pts(3 * n);
            rc = lg_memcpy_d2h(ctx, pts.data(), (const float *)b.d_points + 3 * first[k], pts.size() * 4);
            double radius = 0.0, vs = 0.0;
            if (rc == LG_OK)
                rc = lg_splat_radius(pts.data(), (int64_t)n,
                                     (int64_t)std::max(s.pts_per_kernel, 1) * std::max(s.iterations, 1), &radius, &vs);
            if (rc == LG_OK)
                rc = lg_splat(ctx, (const float *)b.d_points + 3 * first[k], (const float *)b.d_diffuse + 3 * first[k],
                              (int64_t)n, radius, &live_vols[k]->sg);
            const size_t fb = (size_t)s.camera.width * (size_t)s.camera.height * 16;
            for (int f = 0; f < 4 && rc == LG_OK; ++f) rc = grow(&films[f], &film_bytes_f[f], fb);
            lg_films fl{(float *)films[0], (float *)films[1], (float *)films[2], (float *)films[3]};
            if (rc == LG_OK)
                rc = lg_render(ctx, live_vols[k]->density, &live_vols[k]->sg, live_vols[k]->temperature, &s.camera,
                               &s.shader, s.lights, s.n_lights, &s.march, s.samples, s.render_seed, s.env_seed, &fl);
            for (int f = 0; f < 4 && rc == LG_OK; ++f)
                if (j->films[f]) rc = lg_memcpy_d2h(ctx, j->films[f], films[f], fb);
            finish(j, rc, rc == LG_OK ? nullptr : lg_last_error());
        }
    }

    void loop() {
        for (;;) {
            std::vector<Job *> batch = take();
            if (batch.empty()) break;
            run(batch);
        }
        // release everything this worker holds
        std::lock_guard<std::mutex> lk(srv->mu);
        for (auto &kv : cache) {
            CachedVolume &v = kv.second;
            if (v.sg.d_rgb) lg_scatter_grid_destroy(ctx, &v.sg);
            if (v.density) lg_grid_destroy(ctx, v.density);
            if (v.temperature) lg_grid_destroy(ctx, v.temperature);
            srv->reserved -= std::min(srv->reserved, v.bytes);
        }
        cache.clear();
    }
};

}  // namespace

extern "C" {

int32_t lg_server_create(int32_t device, uint64_t hbm_budget, int32_t workers, int32_t max_batch, lg_server **out) {
    if (!out || workers < 1 || workers > 16 || max_batch < 1 || hbm_budget == 0) {
        lg::set_error("lg_server_create: need 1..16 workers, max_batch >= 1 and a budget");
        return LG_ERR_INVALID;
    }
    *out = nullptr;
    lg_server *s = new lg_server();
    s->device = device;
    s->budget = hbm_budget;
    s->max_batch = max_batch;
    // contexts are created up front so that a bad device fails here
    std::vector<lg_ctx *> ctxs;
    for (int w = 0; w < workers; ++w) {
        lg_ctx *c = nullptr;
        const int32_t rc = lg_ctx_create(device, 0, &c);
        if (rc != LG_OK) {
            for (lg_ctx *x : ctxs) lg_ctx_destroy(x);
            delete s;
            return rc;
        }
        ctxs.push_back(c);
    }
    for (int w = 0; w < workers; ++w) {
        s->workers.emplace_back([s, c = ctxs[w]]() {
            Worker wk;
            wk.srv = s;
            wk.ctx = c;
            wk.loop();
            for (auto &f : wk.films)
                if (f) lg_free(c, f);
            if (wk.solver) lg_solver_destroy(c, wk.solver);
            lg_ctx_destroy(c);
        });
    }
    *out = s;
    return LG_OK;
}

int32_t lg_server_submit(lg_server *s, const lg_job *job, float *const h_films[4], uint64_t *id_out) {
    if (!s || !job || !id_out) {
        lg::set_error("lg_server_submit: NULL argument");
        return LG_ERR_INVALID;
    }
    if (job->points < 1 || job->iterations < 1 || job->camera.width < 1 || job->camera.height < 1 ||
        job->n_lights < 0 || job->n_lights > LG_MAX_LIGHTS || job->samples < 1 || job->scatter_res_factor < 1) {
        lg::set_error("lg_server_submit: bad job (points, iterations, image, lights, samples or scatter factor)");
        return LG_ERR_INVALID;
    }
    std::unique_ptr<Job> j(new Job());
    j->spec = *job;
    for (int f = 0; f < 4; ++f) j->films[f] = h_films ? h_films[f] : nullptr;
    j->volume_bytes = volume_bytes(job->scene);
    j->need = job_bytes(*job);
    if (j->need + j->volume_bytes > s->budget) {
        lg::set_error("lg_server_submit: the job needs about %llu bytes, more than the %llu byte budget",
                      (unsigned long long)(j->need + j->volume_bytes), (unsigned long long)s->budget);
        return LG_ERR_BUDGET;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return LG_ERR_INVALID;
    j->id = s->next_id++;
    j->t_submit = now_s(s->t0);
    *id_out = j->id;
    s->queue.push_back(j.get());
    s->jobs[j->id] = std::move(j);
    s->cv_work.notify_all();
    return LG_OK;
}

static void fill_result(const Job &j, lg_job_result *r) {
    r->status = j.status;
    r->t_submit = j.t_submit;
    r->t_admit = j.t_admit;
    r->t_start = j.t_start;
    r->t_done = j.t_done;
    r->batch = j.batch;
    r->reserved_bytes = j.need + j.volume_bytes;
}

int32_t lg_server_wait(lg_server *s, uint64_t id, lg_job_result *out) {
    if (!s || !out) return LG_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) {
        lg::set_error("lg_server_wait: unknown job %llu", (unsigned long long)id);
        return LG_ERR_INVALID;
    }
    Job &j = *it->second;
    s->cv_done.wait(lk, [&] { return j.done; });
    fill_result(j, out);
    s->jobs.erase(it);
    return LG_OK;
}

int32_t lg_server_poll(lg_server *s, uint64_t id, lg_job_result *out, int32_t *done) {
    if (!s || !out || !done) return LG_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return LG_ERR_INVALID;
    *done = it->second->done ? 1 : 0;
    if (*done) {
        fill_result(*it->second, out);
        s->jobs.erase(it);
    }
    return LG_OK;
}

int32_t lg_server_get_stats(lg_server *s, lg_server_stats *out) {
    if (!s || !out) return LG_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    out->queued = (uint64_t)s->queue.size();
    out->completed = s->completed;
    out->failed = s->failed;
    out->batches = s->batches;
    out->batched_jobs = s->batched_jobs;
    out->solves = s->solves;
    out->reserved_bytes = s->reserved;
    out->peak_reserved_bytes = s->peak_reserved;
    out->budget_bytes = s->budget;
    out->clock_s = now_s(s->t0);
    return LG_OK;
}

int32_t lg_server_destroy(lg_server *s) {
    if (!s) return LG_OK;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.wait(lk, [&] { return s->queue.empty(); });
        // wait for the running batches: every submitted job done
        s->cv_done.wait(lk, [&] {
            for (auto &kv : s->jobs)
                if (!kv.second->done) return false;
            return true;
        });
        s->stop = true;
        s->cv_work.notify_all();
    }
    for (auto &t : s->workers) t.join();
    delete s;
    return LG_OK;
}

}  // extern "C"
