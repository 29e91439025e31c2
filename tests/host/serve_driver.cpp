// Failure handling and admission of the serving loop (lightgraph_amd/csrc/lg_serve.cpp)
// on the CPU, over the fake device of fake_abi.cpp: HBM-budget admission,
// an injected allocation failure that fails one job and leaves the others
// complete, refusal of a job larger than the budget, continuous batching of
// jobs that arrive while the worker is busy, overlap across workers, and no
// device memory left behind.  Prints PASS/FAIL per scenario; exit status 0
// when every scenario passes.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "lightgraph.h"

extern "C" {
void fake_set_limit(uint64_t bytes);
void fake_set_fail_seed(uint64_t seed);
void fake_set_render_ms_per_mpix(double ms);
uint64_t fake_bytes_in_use(void);
uint64_t fake_live_allocations(void);
int32_t fake_renders_peak(void);
}

namespace {

int g_failures = 0;

void check(bool ok, const std::string &scenario, const std::string &what) {
    if (!ok) {
        ++g_failures;
        std::printf("FAIL %s: %s\n", scenario.c_str(), what.c_str());
    }
}

lg_job make_job(int side, uint64_t seed, int points, int w, int h, int iterations = 1) {
    lg_job j;
    std::memset(&j, 0, sizeof(j));
    j.scene.kind = LG_SCENE_BLOBS_FBM;
    j.scene.n = side;
    j.scene.seed = seed;
    j.scene.octaves = 6;
    j.scene.blobs = 12;
    j.scene.axis_min = side / 16.0;
    j.scene.axis_max = side / 4.0;
    j.scene.active_fraction = 0.3;
    j.scene.dtype = LG_F32;
    j.points = points;
    j.iterations = iterations;
    j.radius = 6.0;
    j.max_connect = 12;
    j.samples = 1;
    j.seed = 854123;
    j.env_seed = 4564446;
    j.render_seed = 1;
    j.pts_per_kernel = 3;
    j.scatter_res_factor = 4;
    lg_shader_default(&j.shader);
    lg_march_params_default(&j.march);
    const double rot[3] = {0, 0, 0}, tr[3] = {side / 2.0, side / 2.0, 2.6 * side};
    lg_camera_setup(rot, tr, 0.0, 50.0, 36.0, 1e-3, 1e30, w, h, &j.camera);
    j.n_lights = 1;
    lg_light_parse("dir:0:1,1,1:1:-45,30,0", &j.lights[0]);
    return j;
}

struct Films {
    std::vector<float> f[4];
    float *ptr[4];
    Films(int w, int h) {
        for (int k = 0; k < 4; ++k) {
            f[k].assign((size_t)w * h * 4, -1.0f);
            ptr[k] = f[k].data();
        }
    }
};

struct Run {
    std::vector<lg_job_result> res;
    std::vector<Films *> films;
    lg_server_stats st{};
    ~Run() {
        for (Films *f : films) delete f;
    }
};

// submits the jobs (all at once), waits for each, reads the statistics
int32_t run_jobs(lg_server *srv, const std::vector<lg_job> &jobs, Run &out) {
    std::vector<uint64_t> ids;
    for (const lg_job &j : jobs) {
        Films *f = new Films(j.camera.width, j.camera.height);
        out.films.push_back(f);
        uint64_t id = 0;
        const int32_t rc = lg_server_submit(srv, &j, f->ptr, &id);
        if (rc != LG_OK) return rc;
        ids.push_back(id);
    }
    for (uint64_t id : ids) {
        lg_job_result r{};
        const int32_t rc = lg_server_wait(srv, id, &r);
        if (rc != LG_OK) return rc;
        out.res.push_back(r);
    }
    return lg_server_get_stats(srv, &out.st);
}

void budget_admission() {
    const std::string S = "budget_admission";
    const int before = g_failures;
    fake_set_render_ms_per_mpix(5000.0);  // 64 x 48: 15 ms per render
    lg_server *probe = nullptr;
    check(lg_server_create(0, 1ull << 40, 1, 1, &probe) == LG_OK, S, "create probe");
    Run p;
    check(run_jobs(probe, {make_job(64, 11, 500, 64, 48)}, p) == LG_OK, S, "probe run");
    lg_server_destroy(probe);
    const uint64_t need = p.res.empty() ? 0 : p.res[0].reserved_bytes;
    check(need > 0, S, "a job reserves memory");
    // room for one job at a time: the others wait for memory, all complete
    lg_server *srv = nullptr;
    check(lg_server_create(0, need + need / 2, 2, 4, &srv) == LG_OK, S, "create");
    std::vector<lg_job> jobs;
    for (int k = 0; k < 4; ++k) jobs.push_back(make_job(64, 20 + k, 500, 64, 48));
    Run r;
    check(run_jobs(srv, jobs, r) == LG_OK, S, "run");
    for (const lg_job_result &x : r.res) check(x.status == LG_OK, S, "every job completes");
    check(r.st.completed == 4 && r.st.failed == 0, S, "stats count 4 completed");
    check(r.st.peak_reserved_bytes <= r.st.budget_bytes, S, "peak reservation within the budget");
    // serialised by the budget: each job started after the previous one finished
    std::vector<std::pair<double, double>> se;
    for (const lg_job_result &x : r.res) se.push_back({x.t_start, x.t_done});
    std::sort(se.begin(), se.end());
    for (size_t k = 1; k < se.size(); ++k) check(se[k].first >= se[k - 1].second - 1e-6, S, "runs serialised");
    // a job larger than the whole budget is refused at submit
    lg_job big = make_job(512, 5, 60000, 2048, 2048);
    Films bf(2048, 2048);
    uint64_t id = 0;
    check(lg_server_submit(srv, &big, bf.ptr, &id) == LG_ERR_BUDGET, S, "oversized job refused with LG_ERR_BUDGET");
    lg_server_destroy(srv);
    check(fake_live_allocations() == 0, S, "nothing left on the device after destroy");
    std::printf("%s %s\n", g_failures > before ? "FAIL" : "PASS", S.c_str());
}

void injected_failure() {
    const std::string S = "injected_failure";
    const int before = g_failures;
    fake_set_render_ms_per_mpix(0.0);
    const uint64_t kFail = 666;
    std::vector<lg_job> ok_jobs, mixed;
    for (uint64_t seed : {1, 2, 4, 5}) ok_jobs.push_back(make_job(48, seed, 300, 32, 32));
    mixed = {ok_jobs[0], ok_jobs[1], make_job(48, kFail, 300, 32, 32), ok_jobs[2], ok_jobs[3]};
    // reference: the four good jobs alone
    lg_server *ref = nullptr;
    check(lg_server_create(0, 1ull << 40, 2, 4, &ref) == LG_OK, S, "create reference");
    Run a;
    check(run_jobs(ref, ok_jobs, a) == LG_OK, S, "reference run");
    lg_server_destroy(ref);
    fake_set_fail_seed(kFail);
    lg_server *srv = nullptr;
    check(lg_server_create(0, 1ull << 40, 2, 4, &srv) == LG_OK, S, "create");
    Run r;
    check(run_jobs(srv, mixed, r) == LG_OK, S, "run");
    check(r.res.size() == 5, S, "five results");
    for (size_t k = 0; k < r.res.size(); ++k) {
        if (k == 2)
            check(r.res[k].status == LG_ERR_NOMEM, S, "the injected job reports LG_ERR_NOMEM");
        else
            check(r.res[k].status == LG_OK, S, "the other jobs complete");
    }
    check(r.st.failed == 1 && r.st.completed == 4, S, "stats: 1 failed, 4 completed");
    // the failed job's volume is not left charged: the same resident volumes as the reference
    check(r.st.reserved_bytes == a.st.reserved_bytes, S, "reservation equals the reference's");
    // the good jobs' films equal the reference's
    for (int k = 0, m = 0; k < 5; ++k) {
        if (k == 2) continue;
        for (int f = 0; f < 4; ++f) check(r.films[k]->f[f] == a.films[m]->f[f], S, "films equal alone and mixed");
        ++m;
    }
    lg_server_destroy(srv);
    fake_set_fail_seed(~0ull);
    check(fake_live_allocations() == 0, S, "nothing left on the device after destroy");
    // device memory that runs out inside a job: the job fails, the next one completes
    lg_server *s2 = nullptr;
    check(lg_server_create(0, 1ull << 40, 1, 1, &s2) == LG_OK, S, "create (limit)");
    fake_set_limit(fake_bytes_in_use() + 4096);
    Run lim;
    check(run_jobs(s2, {make_job(48, 77, 300, 32, 32)}, lim) == LG_OK, S, "run under the limit");
    check(!lim.res.empty() && lim.res[0].status == LG_ERR_NOMEM, S, "allocation failure reported as LG_ERR_NOMEM");
    fake_set_limit(0);
    Run after;
    check(run_jobs(s2, {make_job(48, 78, 300, 32, 32)}, after) == LG_OK, S, "run after the limit");
    check(!after.res.empty() && after.res[0].status == LG_OK, S, "the next job completes");
    lg_server_destroy(s2);
    check(fake_live_allocations() == 0, S, "nothing left after the allocation failure");
    std::printf("%s %s\n", g_failures > before ? "FAIL" : "PASS", S.c_str());
}

void continuous_batching() {
    const std::string S = "continuous_batching";
    const int before = g_failures;
    fake_set_render_ms_per_mpix(400.0);  // 512 x 512: ~100 ms; 32 x 32: ~0.4 ms
    lg_server *srv = nullptr;
    check(lg_server_create(0, 1ull << 40, 1, 8, &srv) == LG_OK, S, "create");
    // the first job occupies the only worker; the next three arrive meanwhile,
    // share a solve set-up and are taken as one batch
    lg_job slow = make_job(64, 90, 400, 512, 512);
    Films sf(512, 512);
    uint64_t sid = 0;
    check(lg_server_submit(srv, &slow, sf.ptr, &sid) == LG_OK, S, "submit slow");
    // wait until the worker has taken it
    for (int it = 0; it < 2000; ++it) {
        lg_server_stats st{};
        lg_server_get_stats(srv, &st);
        if (st.queued == 0) break;
        std::this_thread::sleep_for(std::chrono::milliseconds(1));
    }
    Run r;
    std::vector<lg_job> jobs = {make_job(48, 91, 300, 32, 32), make_job(48, 92, 300, 32, 32),
                                make_job(48, 93, 300, 32, 32)};
    check(run_jobs(srv, jobs, r) == LG_OK, S, "run");
    lg_job_result s0{};
    lg_server_wait(srv, sid, &s0);
    check(s0.status == LG_OK && s0.batch == 1, S, "the first job ran alone");
    for (const lg_job_result &x : r.res) {
        check(x.status == LG_OK, S, "batched jobs complete");
        check(x.batch == 3, S, "the three waiting jobs share one batch");
        check(x.t_start >= s0.t_done - 1e-6, S, "the batch starts when the worker is free");
    }
    lg_server_destroy(srv);
    // overlap: three workers render at the same time
    lg_server *s3 = nullptr;
    check(lg_server_create(0, 1ull << 40, 3, 1, &s3) == LG_OK, S, "create (3 workers)");
    std::vector<lg_job> six;
    for (int k = 0; k < 6; ++k) six.push_back(make_job(48, 100 + k, 300, 256, 256));
    Run o;
    check(run_jobs(s3, six, o) == LG_OK, S, "run (3 workers)");
    check(fake_renders_peak() >= 2, S, "renders of different workers overlap");
    lg_server_destroy(s3);
    fake_set_render_ms_per_mpix(0.0);
    check(fake_live_allocations() == 0, S, "nothing left on the device after destroy");
    std::printf("%s %s\n", g_failures > before ? "FAIL" : "PASS", S.c_str());
}

}  // namespace

int main() {
    budget_admission();
    injected_failure();
    continuous_batching();
    std::printf("%s\n", g_failures ? "SOME FAILED" : "ALL PASS");
    return g_failures ? 1 : 0;
}
