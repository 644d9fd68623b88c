// Job server over the batch C-ABI: simulation jobs arrive at any time, are
// admitted when their HBM footprint fits the budget, and advance together
// with the jobs already on the device (continuous batching).  No
// counterpart in the reference, which runs one simulation per process run
// (main.cpp:130-300); the server is a client of the public C-ABI only.
//
// One tick: admit queued jobs in arrival order while the budget and the job
// limit allow, rebuild the batch if the set of jobs changed, advance every
// active job by min(tick, smallest remaining step count) steps with one
// launch per stage, then retire the jobs that are done (their positions are
// copied to the host and their device state is freed).  A job's trajectory
// is bit-identical to running it alone, whatever it shares the device with.

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "../../include/bdk.h"

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Job {
    int64_t id = 0;
    bd_params p{};
    std::vector<int32_t> git;
    std::vector<double> rel;
    int64_t steps = 0, done = 0;
    uint64_t bytes = 0;
    bdk_t *h = nullptr;
    double t_submit = 0, t_admit = 0, t_done = 0;
};

}  // namespace

struct bdk_server {
    int device = 0;
    uint64_t budget = 0, used = 0;
    int max_jobs = 0;
    int64_t tick = 0;
    int64_t next_id = 1;
    std::deque<Job *> queue;
    std::vector<Job *> active;
    std::map<int64_t, Job *> finished;
    bdk_batch_t *batch = nullptr;
    bool dirty = false;
    int64_t particle_steps = 0;
    double busy_s = 0;
    std::string err;

    int fail(int code, const std::string &m) {
        err = m;
        return code;
    }
};

extern "C" {

int bdk_serve_create(int32_t device, uint64_t hbm_budget, int32_t max_jobs, int64_t tick_steps,
                     bdk_server_t **out) {
    if (!out || hbm_budget == 0 || max_jobs < 1 || tick_steps < 1) return BDK_ERR_ARG;
    bdk_server_t *s = new bdk_server();
    s->device = device;
    s->budget = hbm_budget;
    s->max_jobs = max_jobs;
    s->tick = tick_steps;
    *out = s;
    return BDK_OK;
}

void bdk_serve_destroy(bdk_server_t *s) {
    if (!s) return;
    if (s->batch) bdk_batch_destroy(s->batch);
    for (Job *j : s->active) {
        bdk_destroy(j->h);
        delete j;
    }
    for (Job *j : s->queue) delete j;
    for (auto &kv : s->finished) delete kv.second;
    delete s;
}

const char *bdk_serve_last_error(const bdk_server_t *s) { return s ? s->err.c_str() : ""; }

int bdk_serve_submit(bdk_server_t *s, const bd_params *p, const int32_t *git, const double *rel, int64_t steps,
                     int64_t *job_id) {
    if (!s || !p || !git || !rel || steps < 1) return BDK_ERR_ARG;
    uint64_t bytes = 0;
    int rc = bdk_footprint(p, &bytes);
    if (rc) return s->fail(rc, std::string("submit: ") + bdk_last_error(nullptr));
    if (bytes > s->budget) return s->fail(BDK_ERR_ARG, "submit: the job's footprint exceeds the HBM budget");
    if (p->max_travel < 1e300 || p->idgas_only || p->noise_mode == BDK_NOISE_UPLOAD)
        return s->fail(BDK_ERR_ARG, "submit: served jobs need fixed dt, capillary forces and device noise");
    Job *j = new Job();
    j->id = s->next_id++;
    j->p = *p;
    j->p.device = s->device;
    const size_t N = (size_t)p->n1 + p->n2;
    j->git.assign(git, git + 2 * N);
    j->rel.assign(rel, rel + 2 * N);
    j->steps = steps;
    j->bytes = bytes;
    j->t_submit = now_s();
    s->queue.push_back(j);
    if (job_id) *job_id = j->id;
    return BDK_OK;
}

// One tick; *n_active / *n_queued / *n_done receive the counts after it.
int bdk_serve_step(bdk_server_t *s, int32_t *n_active, int32_t *n_queued, int32_t *n_done) {
    if (!s) return BDK_ERR_ARG;
    // admission in arrival order
    while (!s->queue.empty() && (int)s->active.size() < s->max_jobs) {
        Job *j = s->queue.front();
        if (s->used + j->bytes > s->budget) break;
        bdk_t *h = nullptr;
        int rc = bdk_create(&j->p, &h);
        if (rc) return s->fail(rc, std::string("admit: ") + bdk_last_error(nullptr));
        const int n1 = j->p.n1, n2 = j->p.n2;
        rc = bdk_set_positions(h, 1, n1, j->git.data(), j->rel.data());
        if (!rc && n2) rc = bdk_set_positions(h, 2, n2, j->git.data() + 2 * (size_t)n1, j->rel.data() + 2 * (size_t)n1);
        if (rc) {
            std::string m = bdk_last_error(h);
            bdk_destroy(h);
            return s->fail(rc, "admit: " + m);
        }
        j->h = h;
        j->t_admit = now_s();
        s->used += j->bytes;
        s->queue.pop_front();
        s->active.push_back(j);
        s->dirty = true;
    }
    if (!s->active.empty()) {
        if (s->dirty) {
            if (s->batch) bdk_batch_destroy(s->batch);
            s->batch = nullptr;
            std::vector<bdk_t *> hs;
            for (Job *j : s->active) hs.push_back(j->h);
            int rc = bdk_batch_create(hs.data(), (int32_t)hs.size(), &s->batch);
            if (rc) return s->fail(rc, std::string("batch: ") + bdk_batch_last_error(nullptr));
            s->dirty = false;
        }
        int64_t k = s->tick;
        for (Job *j : s->active) k = std::min(k, j->steps - j->done);
        const double t0 = now_s();
        int rc = bdk_batch_run_steps(s->batch, k);
        if (!rc) rc = bdk_batch_sync(s->batch);
        if (rc) return s->fail(rc, std::string("tick: ") + bdk_batch_last_error(s->batch));
        s->busy_s += now_s() - t0;
        std::vector<Job *> keep;
        for (Job *j : s->active) {
            j->done += k;
            s->particle_steps += k * ((int64_t)j->p.n1 + j->p.n2);
            if (j->done < j->steps) {
                keep.push_back(j);
                continue;
            }
            const int n1 = j->p.n1, n2 = j->p.n2;
            rc = bdk_get_positions(j->h, 1, n1, j->git.data(), j->rel.data());
            if (!rc && n2)
                rc = bdk_get_positions(j->h, 2, n2, j->git.data() + 2 * (size_t)n1, j->rel.data() + 2 * (size_t)n1);
            if (rc) return s->fail(rc, std::string("retire: ") + bdk_last_error(j->h));
            j->t_done = now_s();
            s->finished[j->id] = j;
            s->used -= j->bytes;
            s->dirty = true;
        }
        if (s->dirty) {
            // the batch refers to the retiring handles: drop it before freeing them
            if (s->batch) bdk_batch_destroy(s->batch);
            s->batch = nullptr;
            for (Job *j : s->active)
                if (j->done >= j->steps) {
                    bdk_destroy(j->h);
                    j->h = nullptr;
                }
        }
        s->active.swap(keep);
    }
    if (n_active) *n_active = (int32_t)s->active.size();
    if (n_queued) *n_queued = (int32_t)s->queue.size();
    if (n_done) *n_done = (int32_t)s->finished.size();
    return BDK_OK;
}

// Result of a finished job: positions (species 1 first) and its times in
// seconds from submission: admitted after t_admit, finished after t_done.
int bdk_serve_result(bdk_server_t *s, int64_t job_id, int32_t *git, double *rel, double *t_admit, double *t_done) {
    if (!s) return BDK_ERR_ARG;
    auto it = s->finished.find(job_id);
    if (it == s->finished.end()) return s->fail(BDK_ERR_STATE, "job not finished (or unknown id)");
    const Job *j = it->second;
    if (git) std::memcpy(git, j->git.data(), j->git.size() * sizeof(int32_t));
    if (rel) std::memcpy(rel, j->rel.data(), j->rel.size() * sizeof(double));
    if (t_admit) *t_admit = j->t_admit - j->t_submit;
    if (t_done) *t_done = j->t_done - j->t_submit;
    return BDK_OK;
}

// Totals since creation: particle-steps done and seconds spent in ticks.
int bdk_serve_stats(const bdk_server_t *s, int64_t *particle_steps, double *busy_s, uint64_t *hbm_used) {
    if (!s) return BDK_ERR_ARG;
    if (particle_steps) *particle_steps = s->particle_steps;
    if (busy_s) *busy_s = s->busy_s;
    if (hbm_used) *hbm_used = s->used;
    return BDK_OK;
}

}  // extern "C"
