// fourier_hip serving loop: FIFO admission by HBM footprint, one worker per
// active job, each job's volume resident on the device on its own stream,
// chunks of concurrently running jobs overlapping on the GPU.
#include "fourier_hip.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

struct Job {
    int id = 0;
    fh_job spec{};
    std::string sym, mode;  // owned copies of the parameter strings
    fh_job_status st{};
    bool cancel = false;
};

}  // namespace

struct fh_server {
    int device = 0;
    long long budget = 0;
    int max_active = 1;
    int chunk = 2048;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<int> pending;
    std::deque<std::unique_ptr<Job>> jobs;
    long long committed = 0;
    int running = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    std::chrono::steady_clock::time_point t0;
    std::string err;

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

// device bytes of a handle by inserter mode name (engine.cpp)
long long footprint_bytes_for_mode(int P, int nsym, int chunk, const char *mode_name);

static long long job_footprint(const fh_job &j, int chunk)
{
    const int P = j.params.nx;
    // the operator count from the library's own symmetry table (c, d, h, tet,
    // oct, icos); an unknown name costs one operator and fails in fh_create
    int nsym = fh_get_symmetries(j.params.sym ? j.params.sym : "c1", nullptr, 0);
    if (nsym < 1) nsym = 1;
    const int m = j.n < chunk ? (j.n > 0 ? j.n : 1) : chunk;
    long long fp = footprint_bytes_for_mode(P, nsym, m, j.params.mode);
    fp += 4LL * m * j.box * j.box;       // staging of host particles
    fp += 4LL * P * P * P + 8LL * P * P;  // finish(): real map + conjugate planes
    return fp;
}

static void run_job(fh_server *s, Job *job)
{
    fh_reconstructor *h = nullptr;
    int rc = fh_create(&job->spec.params, &h);
    if (rc == FH_OK) {
        fh_tuning t;
        fh_get_tuning(h, &t);
        t.chunk = s->chunk;
        rc = fh_set_tuning(h, &t);
    }
    if (rc == FH_OK) rc = fh_setup(h);
    const fh_job &j = job->spec;
    const fh_ctf *ctf = j.use_ctf ? &j.ctf : nullptr;
    for (int c0 = 0; rc == FH_OK && c0 < j.n; c0 += s->chunk) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (job->cancel) {
                rc = -100;
                break;
            }
        }
        const int m = std::min(s->chunk, j.n - c0);
        rc = fh_insert_batch(h, j.imgs + (size_t)c0 * j.box * j.box, m, j.box, j.xforms + 12 * (size_t)c0,
                             j.weights + c0, j.defocus ? j.defocus + c0 : nullptr, ctf, j.where);
        if (rc == FH_OK) {
            std::lock_guard<std::mutex> lk(s->mu);
            job->st.particles_done += m;
        }
    }
    if (rc == FH_OK && j.out_map) rc = fh_finish(h, 1, j.out_map, nullptr);
    else if (rc == FH_OK) rc = fh_sync(h);
    fh_destroy(h);
    std::lock_guard<std::mutex> lk(s->mu);
    s->committed -= job->st.footprint_bytes;
    s->running--;
    job->st.t_done = s->now();
    if (rc == -100) job->st.state = FH_JOB_CANCELLED;
    else if (rc != FH_OK) {
        job->st.state = FH_JOB_FAILED;
        job->st.error = rc;
    } else job->st.state = FH_JOB_DONE;
    s->cv.notify_all();
}

static void worker(fh_server *s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        // FIFO admission: the head of the queue waits until its footprint fits
        s->cv.wait(lk, [s] {
            if (s->stop) return true;
            if (s->pending.empty()) return false;
            const Job *j = s->jobs[s->pending.front()].get();
            return j->st.footprint_bytes > s->budget || j->cancel ||
                   s->committed + j->st.footprint_bytes <= s->budget;
        });
        if (s->stop) return;
        Job *job = s->jobs[s->pending.front()].get();
        s->pending.pop_front();
        if (job->cancel) {
            job->st.state = FH_JOB_CANCELLED;
            job->st.t_done = s->now();
            s->cv.notify_all();
            continue;
        }
        if (job->st.footprint_bytes > s->budget) {
            job->st.state = FH_JOB_REJECTED;
            job->st.error = FH_ERR_NOMEM;
            job->st.t_done = s->now();
            s->cv.notify_all();
            continue;
        }
        s->committed += job->st.footprint_bytes;
        s->running++;
        job->st.state = FH_JOB_RUNNING;
        job->st.t_admit = s->now();
        s->cv.notify_all();
        lk.unlock();
        run_job(s, job);
        lk.lock();
    }
}

extern "C" {

int fh_server_create(int device, long long hbm_budget_bytes, int max_active, int chunk, fh_server **out)
{
    if (!out) return FH_ERR_NULL;
    *out = nullptr;
    if (hbm_budget_bytes <= 0 || max_active < 1 || chunk < 1) return FH_ERR_INVALID_PARAM;
    if (hipSetDevice(device) != hipSuccess) return FH_ERR_DEVICE;
    fh_server *s = new fh_server();
    s->device = device;
    s->budget = hbm_budget_bytes;
    s->max_active = max_active;
    s->chunk = chunk;
    s->t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < max_active; i++) s->workers.emplace_back(worker, s);
    *out = s;
    return FH_OK;
}

int fh_server_submit(fh_server *s, const fh_job *job, int *job_id)
{
    if (!s || !job || !job_id) return FH_ERR_NULL;
    if (!job->imgs || !job->xforms || !job->weights || job->n < 0) {
        s->err = "NullPointerException: job inputs";
        return FH_ERR_NULL;
    }
    auto j = std::make_unique<Job>();
    j->spec = *job;
    j->sym = job->params.sym ? job->params.sym : "c1";
    j->mode = job->params.mode ? job->params.mode : "gauss_2";
    j->spec.params.sym = j->sym.c_str();
    j->spec.params.mode = j->mode.c_str();
    std::lock_guard<std::mutex> lk(s->mu);
    j->id = (int)s->jobs.size();
    j->st.state = FH_JOB_QUEUED;
    j->st.t_submit = s->now();
    j->st.footprint_bytes = job_footprint(j->spec, s->chunk);
    *job_id = j->id;
    s->pending.push_back(j->id);
    s->jobs.push_back(std::move(j));
    s->cv.notify_all();
    return FH_OK;
}

int fh_server_poll(fh_server *s, int job_id, fh_job_status *st)
{
    if (!s || !st) return FH_ERR_NULL;
    std::lock_guard<std::mutex> lk(s->mu);
    if (job_id < 0 || job_id >= (int)s->jobs.size()) return FH_ERR_INVALID_PARAM;
    *st = s->jobs[job_id]->st;
    return FH_OK;
}

int fh_server_cancel(fh_server *s, int job_id)
{
    if (!s) return FH_ERR_NULL;
    std::lock_guard<std::mutex> lk(s->mu);
    if (job_id < 0 || job_id >= (int)s->jobs.size()) return FH_ERR_INVALID_PARAM;
    s->jobs[job_id]->cancel = true;
    s->cv.notify_all();
    return FH_OK;
}

int fh_server_drain(fh_server *s)
{
    if (!s) return FH_ERR_NULL;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv.wait(lk, [s] {
        if (!s->pending.empty() || s->running > 0) return false;
        for (auto &j : s->jobs)
            if (j->st.state == FH_JOB_QUEUED || j->st.state == FH_JOB_RUNNING) return false;
        return true;
    });
    return FH_OK;
}

double fh_server_clock(fh_server *s) { return s ? s->now() : 0.0; }

int fh_server_destroy(fh_server *s)
{
    if (!s) return FH_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (int id : s->pending) s->jobs[id]->st.state = FH_JOB_CANCELLED;
        s->pending.clear();
        for (auto &j : s->jobs) j->cancel = true;  // running jobs stop at their next chunk
        s->stop = true;
        s->cv.notify_all();
    }
    for (auto &t : s->workers) t.join();
    delete s;
    return FH_OK;
}

const char *fh_server_last_error(const fh_server *s) { return s ? s->err.c_str() : ""; }

}  // extern "C"
