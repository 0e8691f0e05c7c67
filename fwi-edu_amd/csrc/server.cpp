// Serving loop: continuous batching of modeling and gradient jobs on one GPU.
//
// No counterpart in the reference (its drivers run one call at a time).  A
// server owns `slots` workers; each worker is a host thread bound to its own
// engine context (its own HIP stream, device buffers, launch-shape tuning and
// graph cache), so the kernels of jobs on different workers overlap on the
// device.  Jobs are admitted in arrival order as a worker and device memory
// become free: a job reserves the bytes its batches need (call_footprint),
// a later job may pass a waiting one only while that one has waited less
// than a second, and an idle worker gives its buffers back when they block
// a waiting job.  Each job runs exactly the code of iso_acoustic2d_propagation
// (propagate_core), whose result does not depend on the batch size or the
// launch shape, so a job's records and gradient are the same alone and in
// the mix.
#include <math.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "engine.hpp"

namespace fwiedu {

using Clock = std::chrono::steady_clock;

struct ServerJob {
    FwieduJob j;
    long id = 0;
    Clock::time_point t_submit, t_start, t_end;
    size_t fixed = 0, per_shot = 0;
    double misfit = 0.0;
    bool done = false;
};

struct Worker {
    std::thread th;
    Ctx *ctx = nullptr;
    size_t held = 0;      // device bytes its buffers hold
    size_t reserved = 0;  // reservation of the running job
    bool busy = false;
};

struct Server {
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<long> queue;
    std::map<long, ServerJob *> jobs;
    std::vector<Worker> workers;
    long next_id = 1;
    size_t budget = 0;
    int graphs = 1;
    int ckpt_bits = 32;
    bool stop = false;
    int busy_now = 0, busy_peak = 0;
    long completed = 0;
    std::vector<double> latencies;
};

static size_t job_reserve(const ServerJob &sj, size_t avail)
{
    const size_t whole = sj.fixed + sj.per_shot * (size_t)sj.j.nshot;
    if (whole <= avail) return whole;
    return avail;  // batches of fewer shots inside what is free
}

// Index in s.queue of the job worker w may start now, or -1.  Called locked.
static int pick_job(Server &s, int w, size_t &reserve)
{
    size_t others = 0;
    for (int o = 0; o < (int)s.workers.size(); ++o)
        if (o != w)
            others += s.workers[o].busy ? std::max(s.workers[o].held, s.workers[o].reserved)
                                        : s.workers[o].held;
    const size_t held = s.workers[w].held;
    if (others >= s.budget) return -1;
    const size_t avail = s.budget - others;
    const auto now = Clock::now();
    for (int q = 0; q < (int)s.queue.size(); ++q) {
        const ServerJob &sj = *s.jobs[s.queue[q]];
        const size_t need_min = sj.fixed + sj.per_shot;
        if (need_min <= avail) {
            reserve = std::max(held, job_reserve(sj, avail));
            if (reserve > avail) reserve = avail;
            return q;
        }
        // the oldest waiting job keeps its turn once it has waited a second
        if (q == 0 && std::chrono::duration<double>(now - sj.t_submit).count() > 1.0) return -1;
    }
    return -1;
}

static void run_job(Worker &wk, ServerJob &sj, size_t reserve)
{
    const FwieduJob &j = sj.j;
    Ctx &c = *wk.ctx;
    c.mem_cap = reserve > sj.fixed ? reserve - sj.fixed : sj.per_shot;
    ensure_ready();
    Call P;
    P.vp = j.vp;
    P.rho = j.rho;
    P.obs = j.obs;
    P.pre = j.pre;
    P.Nshot = j.nshot;
    P.nz = j.nz;
    P.nx = j.nx;
    P.dx = j.dx;
    P.dz = j.dz;
    P.dt = j.dt;
    P.nt = j.nt;
    P.pml = j.pml;
    P.fs = j.freesurface;
    P.freq = j.freq;
    P.filter = 2;
    P.precondition = j.precondition;
    P.problem = j.problem;
    P.grad = j.grad;
    P.misfit = (j.problem == 1) ? &sj.misfit : nullptr;
    propagate_core(P);
}

static void worker_main(Server *s, int w)
{
    Worker &wk = s->workers[w];
    bind_ctx(wk.ctx);
    wk.ctx->graphs = s->graphs;
    wk.ctx->ckpt_bits = s->ckpt_bits;
    // requests rarely repeat a geometry: take shapes tuned for similar ones,
    // and time a short list for a new one
    wk.ctx->reuse_similar = 1;
    wk.ctx->quick_tune = 1;
    for (;;) {
        ServerJob *sj = nullptr;
        size_t reserve = 0;
        {
            std::unique_lock<std::mutex> lk(s->m);
            for (;;) {
                if (s->stop) return;
                const int q = s->queue.empty() ? -1 : pick_job(*s, w, reserve);
                if (q >= 0) {
                    sj = s->jobs[s->queue[q]];
                    s->queue.erase(s->queue.begin() + q);
                    break;
                }
                if (!s->queue.empty() && wk.held > 0) {
                    // memory this idle worker holds may be what the queue waits for
                    lk.unlock();
                    sync();
                    release_buffers(*wk.ctx);
                    lk.lock();
                    wk.held = 0;
                    continue;
                }
                s->cv_work.wait_for(lk, std::chrono::milliseconds(50));
            }
            wk.busy = true;
            wk.reserved = reserve;
            s->busy_now++;
            s->busy_peak = std::max(s->busy_peak, s->busy_now);
            sj->t_start = Clock::now();
        }
        run_job(wk, *sj, reserve);
        {
            std::lock_guard<std::mutex> lk(s->m);
            wk.held = held_bytes(*wk.ctx);
            wk.busy = false;
            wk.reserved = 0;
            s->busy_now--;
            sj->t_end = Clock::now();
            sj->done = true;
            s->completed++;
            s->latencies.push_back(std::chrono::duration<double>(sj->t_end - sj->t_submit).count());
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

}  // namespace fwiedu

using namespace fwiedu;

extern "C" void *fwiedu_server_create(int slots, double mem_fraction, int graphs, int ckpt_bits)
{
    if (slots < 1 || slots > 16) fail("<error>: fwiedu_server_create: 1 to 16 slots (got %d)", slots);
    if (!(mem_fraction > 0.0 && mem_fraction <= 0.95))
        fail("<error>: fwiedu_server_create: memory fraction in (0, 0.95]");
    if (ckpt_bits != 32 && ckpt_bits != 16)
        fail("<error>: fwiedu_server_create: checkpoint bits 32 or 16");
    ensure_ready();
    size_t fr = 0, tot = 0;
    HIPCHECK(hipMemGetInfo(&fr, &tot));
    Server *s = new Server();
    s->budget = (size_t)((double)(fr + pool_idle_bytes()) * mem_fraction);
    s->graphs = graphs ? 1 : 0;
    s->ckpt_bits = ckpt_bits;
    s->workers.resize(slots);
    for (int w = 0; w < slots; ++w) s->workers[w].ctx = new Ctx();
    for (int w = 0; w < slots; ++w) s->workers[w].th = std::thread(worker_main, s, w);
    return s;
}

extern "C" long fwiedu_server_submit(void *server, const FwieduJob *job)
{
    Server *s = static_cast<Server *>(server);
    if (!s || !job) fail("<error>: fwiedu_server_submit: null argument");
    const FwieduJob &j = *job;
    check_common(j.nz, j.nx, j.dx, j.dz, j.dt, j.nt, j.pml, j.freesurface);
    if (j.problem != 0 && j.problem != 1)
        fail("<error>: fwiedu_server_submit: problem 0 (modeling) or 1 (gradient), got %d",
             j.problem);
    if (j.nshot <= 0 || !j.obs || !j.pre || !j.vp || !j.rho)
        fail("<error>: fwiedu_server_submit: missing shots, records or model");
    if (j.problem == 1 && (!j.grad || !j.grad[0]))
        fail("<error>: fwiedu_server_submit: a gradient job needs a gradient array");
    ServerJob *sj = new ServerJob();
    sj->j = j;
    Call P;
    P.vp = j.vp;
    P.rho = j.rho;
    P.obs = j.obs;
    P.pre = j.pre;
    P.Nshot = j.nshot;
    P.nz = j.nz;
    P.nx = j.nx;
    P.dx = j.dx;
    P.dz = j.dz;
    P.dt = j.dt;
    P.nt = j.nt;
    P.pml = j.pml;
    P.fs = j.freesurface;
    P.problem = j.problem;
    call_footprint(P, s->ckpt_bits == 16 && j.problem == 1, &sj->fixed, &sj->per_shot);
    if (sj->fixed + sj->per_shot > s->budget)
        fail("<error>: fwiedu_server_submit: one shot of this job needs %.2f GB, more than the "
             "server's %.2f GB",
             (double)(sj->fixed + sj->per_shot) / 1e9, (double)s->budget / 1e9);
    long id;
    {
        std::lock_guard<std::mutex> lk(s->m);
        id = s->next_id++;
        sj->id = id;
        sj->t_submit = Clock::now();
        s->jobs[id] = sj;
        s->queue.push_back(id);
    }
    s->cv_work.notify_all();
    return id;
}

extern "C" int fwiedu_server_done(void *server, long id)
{
    Server *s = static_cast<Server *>(server);
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->jobs.find(id);
    return (it != s->jobs.end() && it->second->done) ? 1 : 0;
}

extern "C" int fwiedu_server_wait(void *server, long id, double *latency_s, double *queue_s,
                                  double *misfit)
{
    Server *s = static_cast<Server *>(server);
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return -1;
    ServerJob *sj = it->second;
    s->cv_done.wait(lk, [&] { return sj->done; });
    if (latency_s) *latency_s = std::chrono::duration<double>(sj->t_end - sj->t_submit).count();
    if (queue_s) *queue_s = std::chrono::duration<double>(sj->t_start - sj->t_submit).count();
    if (misfit) *misfit = sj->misfit;
    s->jobs.erase(it);
    delete sj;
    return 0;
}

extern "C" double fwiedu_server_stat(void *server, const char *name)
{
    Server *s = static_cast<Server *>(server);
    std::lock_guard<std::mutex> lk(s->m);
    const std::string n(name ? name : "");
    if (n == "completed") return (double)s->completed;
    if (n == "peak_concurrency") return (double)s->busy_peak;
    if (n == "budget_bytes") return (double)s->budget;
    if (n == "queued") return (double)s->queue.size();
    if (n == "p50_s" || n == "p99_s" || n == "max_s") {
        if (s->latencies.empty()) return 0.0;
        std::vector<double> v = s->latencies;
        std::sort(v.begin(), v.end());
        const double q = n == "p50_s" ? 0.50 : n == "p99_s" ? 0.99 : 1.0;
        // nearest rank
        size_t k = (size_t)ceil(q * (double)v.size());
        if (k < 1) k = 1;
        return v[std::min(k, v.size()) - 1];
    }
    return -1.0;
}

extern "C" void fwiedu_server_destroy(void *server)
{
    Server *s = static_cast<Server *>(server);
    if (!s) return;
    {
        std::unique_lock<std::mutex> lk(s->m);
        // finish what was submitted
        s->cv_done.wait(lk, [&] { return s->queue.empty() && s->busy_now == 0; });
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (Worker &w : s->workers) w.th.join();
    for (Worker &w : s->workers) {
        bind_ctx(w.ctx);
        if (w.ctx->ready) {
            sync();
            release_buffers(*w.ctx);
            (void)hipStreamDestroy(w.ctx->st);
            (void)hipStreamDestroy(w.ctx->cst);
        }
        delete w.ctx;
    }
    bind_ctx(nullptr);
    for (auto &e : s->jobs) delete e.second;
    delete s;
}
