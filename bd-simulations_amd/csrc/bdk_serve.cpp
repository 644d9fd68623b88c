// Job server over the batch C-ABI: simulation jobs arrive at any time, are
// admitted when their HBM footprint fits the budget, and advance together
// with the jobs already on the device (continuous batching).  No
// counterpart in the reference, which runs one simulation per process run
// (main.cpp:130-300); the server is a client of the public C-ABI only.
//
// One tick: the jobs admitted during the previous tick join (the batches are
// rebuilt when the set of jobs changed, one batch per precision class, side
// by side on the device), every active job advances by min(tick, smallest
// remaining step count) steps with one launch per stage and class, queued
// jobs are admitted in arrival order while those steps run (as far as the
// budget and the job limit allow), and the jobs that are done retire (their
// positions are copied to the host and their device state is freed).  A job's trajectory
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
    int status = BDK_OK;  // a failed job: its error code (admission, run-away, read-back)
    std::string err;
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
    std::vector<Job *> pending;  // admitted (on the device, forces enqueued), stepping from the next tick
    std::map<int64_t, Job *> finished;
    // one batch per precision class of the active jobs (a batch needs one
    // precision); idx = the jobs' positions in `active`
    struct Group {
        bdk_batch_t *batch = nullptr;
        std::vector<int> idx;
    };
    std::vector<Group> groups;
    bool dirty = false;
    int64_t particle_steps = 0;
    double busy_s = 0;
    // host seconds per part of the ticks: admission (create + upload),
    // batch rebuilds, stepping (enqueue + wait), retirement (read-back + free)
    double t_admit = 0, t_batch = 0, t_run = 0, t_retire = 0;
    std::string err;

    int fail(int code, const std::string &m) {
        err = m;
        return code;
    }
    void drop_batches() {
        for (Group &g : groups)
            if (g.batch) bdk_batch_destroy(g.batch);
        groups.clear();
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
    s->drop_batches();
    for (Job *j : s->active) {
        bdk_destroy(j->h);
        delete j;
    }
    for (Job *j : s->pending) {
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
    if (p->idgas_only || p->noise_mode == BDK_NOISE_UPLOAD)
        return s->fail(BDK_ERR_ARG, "submit: served jobs need capillary forces and device noise");
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

// A job leaves the device (finished or failed): its device state is freed and
// it moves to the finished map.  The caller has already taken it out of the
// active set (or the queue); the batch that referred to it is gone.
static void retire(bdk_server_t *s, Job *j, int status, const std::string &err) {
    if (j->h) {
        bdk_destroy(j->h);
        j->h = nullptr;
        s->used -= j->bytes;
    }
    j->status = status;
    j->err = err;
    j->t_done = now_s();
    s->finished[j->id] = j;
}

// Admission in arrival order while the job limit and the HBM budget allow:
// each job's handle is created, its positions uploaded and its forces
// enqueued on its own stream (bdk_prepare), so this overlaps the batches'
// steps; the job starts stepping with the next tick.  A job that cannot be
// admitted is finished as failed.
static void admit(bdk_server_t *s) {
    while (!s->queue.empty() && (int)(s->active.size() + s->pending.size()) < s->max_jobs) {
        Job *j = s->queue.front();
        if (s->used + j->bytes > s->budget) break;
        s->queue.pop_front();
        j->t_admit = now_s();
        bdk_t *h = nullptr;
        int rc = bdk_create(&j->p, &h);
        std::string m;
        if (rc) {
            m = bdk_last_error(nullptr);
        } else {
            const int n1 = j->p.n1, n2 = j->p.n2;
            rc = bdk_set_positions(h, 1, n1, j->git.data(), j->rel.data());
            if (!rc && n2)
                rc = bdk_set_positions(h, 2, n2, j->git.data() + 2 * (size_t)n1, j->rel.data() + 2 * (size_t)n1);
            if (!rc) rc = bdk_prepare(h);
            if (rc) {
                m = bdk_last_error(h);
                bdk_destroy(h);
                h = nullptr;
            }
        }
        if (rc) {
            retire(s, j, rc, "admit: " + m);
            continue;
        }
        j->h = h;
        s->used += j->bytes;
        s->pending.push_back(j);
    }
}

// One tick; *n_active / *n_queued / *n_done receive the counts after it
// (n_active counts the admitted jobs, n_done the failed ones too).  Every
// job is at all times in exactly one of the queue, the admitted jobs
// (pending or active) and the finished map.  The tick: the jobs admitted
// during the previous tick join the active ones (the batches are rebuilt,
// one per precision class); the active jobs' steps are enqueued; queued
// jobs are admitted while the device runs them; then the tick waits,
// checks every job and retires the finished and the failed ones (a job
// whose particles ran away fails alone).  The tick itself fails only when
// the device fails.
int bdk_serve_step(bdk_server_t *s, int32_t *n_active, int32_t *n_queued, int32_t *n_done) {
    if (!s) return BDK_ERR_ARG;
    double tp = now_s();
    auto lap = [&](double &acc) {
        const double t = now_s();
        acc += t - tp;
        tp = t;
    };
    if (s->active.empty() && s->pending.empty()) admit(s);  // nothing runs to overlap with
    if (!s->pending.empty()) {
        s->active.insert(s->active.end(), s->pending.begin(), s->pending.end());
        s->pending.clear();
        s->dirty = true;
    }
    lap(s->t_admit);
    if (!s->active.empty()) {
        if (s->dirty) {
            s->drop_batches();
            std::map<int, std::vector<int>> cls;
            for (int i = 0; i < (int)s->active.size(); ++i) cls[s->active[i]->p.precision].push_back(i);
            for (auto &kv : cls) {
                bdk_server::Group g;
                g.idx = kv.second;
                std::vector<bdk_t *> hs;
                for (int i : g.idx) hs.push_back(s->active[i]->h);
                int rc = bdk_batch_create(hs.data(), (int32_t)hs.size(), &g.batch);
                if (rc) {
                    s->drop_batches();
                    return s->fail(rc, std::string("batch: ") + bdk_batch_last_error(nullptr));
                }
                s->groups.push_back(g);
            }
            s->dirty = false;
        }
        lap(s->t_batch);
        int64_t k = s->tick;
        for (Job *j : s->active) k = std::min(k, j->steps - j->done);
        const double t0 = now_s();
        // every class's steps are enqueued before the first wait: the batches
        // run side by side on their own streams
        for (bdk_server::Group &g : s->groups) {
            const int rc = bdk_batch_run_steps(g.batch, k);
            if (rc) return s->fail(rc, std::string("tick: ") + bdk_batch_last_error(g.batch));
        }
        // the steps are enqueued: they count whatever the members' state
        for (Job *j : s->active) {
            j->done += k;
            s->particle_steps += k * ((int64_t)j->p.n1 + j->p.n2);
        }
        lap(s->t_run);
        admit(s);  // beside the enqueued steps
        lap(s->t_admit);
        const int n = (int)s->active.size();
        std::vector<int32_t> st(n, BDK_OK);
        for (bdk_server::Group &g : s->groups) {
            std::vector<int32_t> gs(g.idx.size(), BDK_OK);
            const int rc = bdk_batch_status(g.batch, (int32_t)gs.size(), gs.data());
            if (rc) return s->fail(rc, std::string("tick: ") + bdk_batch_last_error(g.batch));
            for (size_t q = 0; q < g.idx.size(); ++q) st[g.idx[q]] = gs[q];
        }
        s->busy_s += now_s() - t0;
        lap(s->t_run);
        std::vector<Job *> keep, out;
        std::vector<int> out_rc;
        std::vector<std::string> out_err;
        for (int i = 0; i < n; ++i) {
            Job *j = s->active[i];
            if (st[i] == BDK_OK && j->done < j->steps) {
                keep.push_back(j);
                continue;
            }
            int jrc = st[i];
            std::string m = jrc ? "run: a particle position became non-finite or ran away" : "";
            if (!jrc) {
                const int n1 = j->p.n1, n2 = j->p.n2;
                jrc = bdk_get_positions(j->h, 1, n1, j->git.data(), j->rel.data());
                if (!jrc && n2)
                    jrc = bdk_get_positions(j->h, 2, n2, j->git.data() + 2 * (size_t)n1,
                                            j->rel.data() + 2 * (size_t)n1);
                if (jrc) m = std::string("retire: ") + bdk_last_error(j->h);
            }
            out.push_back(j);
            out_rc.push_back(jrc);
            out_err.push_back(m);
        }
        if (!out.empty()) {
            // the batches refer to the leaving handles: drop them before freeing those
            s->drop_batches();
            s->dirty = true;
            s->active.swap(keep);
            for (size_t i = 0; i < out.size(); ++i) retire(s, out[i], out_rc[i], out_err[i]);
        }
        lap(s->t_retire);
    }
    if (n_active) *n_active = (int32_t)(s->active.size() + s->pending.size());
    if (n_queued) *n_queued = (int32_t)s->queue.size();
    if (n_done) *n_done = (int32_t)s->finished.size();
    return BDK_OK;
}

// Result of a finished job: positions (species 1 first) and its times in
// seconds from submission: admitted after t_admit, finished after t_done.  A
// failed job returns its error code (positions not written, times written).
int bdk_serve_result(bdk_server_t *s, int64_t job_id, int32_t *git, double *rel, double *t_admit, double *t_done) {
    if (!s) return BDK_ERR_ARG;
    auto it = s->finished.find(job_id);
    if (it == s->finished.end()) return s->fail(BDK_ERR_STATE, "job not finished (or unknown id)");
    const Job *j = it->second;
    if (t_admit) *t_admit = j->t_admit - j->t_submit;
    if (t_done) *t_done = j->t_done - j->t_submit;
    if (j->status) return s->fail(j->status, "job " + std::to_string(job_id) + " failed: " + j->err);
    if (git) std::memcpy(git, j->git.data(), j->git.size() * sizeof(int32_t));
    if (rel) std::memcpy(rel, j->rel.data(), j->rel.size() * sizeof(double));
    return BDK_OK;
}

int bdk_serve_timing(const bdk_server_t *s, double *admit_s, double *batch_s, double *run_s, double *retire_s) {
    if (!s) return BDK_ERR_ARG;
    if (admit_s) *admit_s = s->t_admit;
    if (batch_s) *batch_s = s->t_batch;
    if (run_s) *run_s = s->t_run;
    if (retire_s) *retire_s = s->t_retire;
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
