// engine.hip -- batched and served proving on one GPU.
//
// The reference proves one AirTable at a time on a global rayon pool
// (src/examples/poseidon2.rs:153-174).  Here an engine owns n_slots worker
// threads, each with its own context (HIP stream, memory pool, twiddle cache),
// and a FIFO of jobs.  A job is admitted when a slot is free and its device
// footprint fits the memory budget beside the footprints of the jobs already
// running; admitted jobs overlap on the device (separate streams, so one job's
// host round trips are hidden by the others' kernels).  Every job runs the
// same code as the single-proof path (wl_air_table_new ->
// wl_air_generate_witness -> wl_air_prove_resident), so its proof does not
// depend on what else runs.
#include <hip/hip_runtime.h>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>
#include "../../include/whirlaway_hip.h"

void wl_set_error(const char *msg);

namespace {

using clk = std::chrono::steady_clock;

struct Job {
    uint64_t id = 0;
    wl_job spec{};
    uint64_t footprint = 0;
    clk::time_point t_submit, t_start, t_done;
    bool done = false;
    int status = WL_OK;
    std::string err;
    std::vector<uint32_t> proof;
};

double ms_between(clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

struct wl_engine {
    int device = 0;
    int n_slots = 0;
    uint64_t budget = 0;
    std::vector<wl_ctx *> ctxs;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Job>> queue;
    std::map<uint64_t, std::shared_ptr<Job>> jobs;
    uint64_t reserved = 0, peak_reserved = 0;
    int running = 0, peak_running = 0;
    uint64_t completed = 0, next_id = 1;
    bool stop = false;
    int waiters = 0;  // threads inside wl_engine_wait (destroy waits for them to leave)

    void worker(int slot);
};

namespace {

// Run one job on a context; fills job->proof or job->err.
void run_job(wl_ctx *ctx, Job &j, std::vector<uint32_t> &buf) {
    const wl_job &s = j.spec;
    wl_table *t = wl_air_table_new(ctx, s.air_rc, s.log_n, s.settings.univariate_skips);
    if (!t) { j.status = WL_ERR_DEVICE; j.err = wl_last_error(); return; }
    wl_witness *w = wl_air_generate_witness(t, s.inputs);
    if (!w) {
        j.status = WL_ERR_DEVICE;
        j.err = wl_last_error();
        wl_air_table_free(t);
        return;
    }
    size_t len = 0;
    int rc = wl_air_prove_resident(t, &s.settings, s.p16, s.p24, w, buf.data(), buf.size(), &len);
    if (rc == WL_ERR_BUFFER) {
        buf.resize(len);
        rc = wl_air_prove_resident(t, &s.settings, s.p16, s.p24, w, buf.data(), buf.size(), &len);
    }
    if (rc == WL_OK) j.proof.assign(buf.begin(), buf.begin() + len);
    else { j.status = rc; j.err = wl_last_error(); }
    wl_witness_free(w);
    wl_air_table_free(t);
}

}  // namespace

void wl_engine::worker(int slot) {
    (void)hipSetDevice(device);
    wl_ctx *ctx = ctxs[slot];
    std::vector<uint32_t> buf((size_t)1 << 18);
    for (;;) {
        std::shared_ptr<Job> j;
        {
            std::unique_lock<std::mutex> lk(mu);
            // FIFO: the head of the queue runs as soon as its footprint fits
            cv_work.wait(lk, [&] { return stop || (!queue.empty() && reserved + queue.front()->footprint <= budget); });
            if (queue.empty() && stop) return;
            if (stop && !(reserved + queue.front()->footprint <= budget)) return;
            j = queue.front();
            queue.pop_front();
            reserved += j->footprint;
            if (reserved > peak_reserved) peak_reserved = reserved;
            running++;
            if (running > peak_running) peak_running = running;
            j->t_start = clk::now();
        }
        cv_work.notify_all();  // the next head may fit too
        run_job(ctx, *j, buf);
        {
            std::lock_guard<std::mutex> lk(mu);
            j->t_done = clk::now();
            j->done = true;
            reserved -= j->footprint;
            running--;
            completed++;
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }
}

extern "C" {

uint64_t wl_job_footprint(uint32_t log_n, const wl_air_settings *settings) {
    // Device bytes a proof of 2^log_n rows can hold: the sum of the prover's buffers
    // over all phases (an upper bound of the live peak, since the phases' buffers
    // do not all coexist), plus a margin.  C = 165 committed columns, v = n + 8.
    wl_air_settings d{128, 7, 4, 1, 4, 5, WL_CAPACITY_BOUND};
    const wl_air_settings &s = settings ? *settings : d;
    const double C = 165.0, N = std::ldexp(1.0, (int)log_n);
    const int v = (int)log_n + 8, f0 = (int)s.folding_factor_first, f = (int)s.folding_factor;
    const int rate = (int)s.log_inv_rate, rs0 = (int)s.rs_domain_initial_reduction;
    const int k = (int)s.univariate_skips;
    const double K = std::ldexp(1.0, k), M = N / K;
    const double packed = std::ldexp(4.0, v);
    // first codeword: the non-zero columns (of 2^f0) times 2^(v + rate - f0) rows, its
    // encoding scratch of the same size, and the Merkle heap (64 B per node)
    const double nnz = std::ceil(C * N / std::ldexp(1.0, v - f0));
    const double rows0 = std::ldexp(1.0, v + rate - f0);
    const double commit = 2.0 * nnz * rows0 * 4.0 + rows0 * 64.0;
    // zerocheck: round-0 extended rows (skips <= 4), folded EF columns, eq tables
    const double zc = (k <= 4 ? (2.0 * K - 2.0) * C * M * 4.0 : 0.0) + C * M * 32.0 + 2.0 * M * 32.0;
    // inner sums and inner sumcheck: batched, mixed and matrix tables (EF)
    const double inner = 4.0 * N * 32.0;
    // WHIR opening: f, W and the column split of f (2^(v - f0) EF each), the first
    // round's codeword (8 << f columns of 2^(v + rate - rs0 - f) rows) with scratch
    // and tree, and the eq tables of the initial rounds' constraints
    const double m1 = std::ldexp(1.0, v - f0), rows1 = std::ldexp(1.0, v + rate - rs0 - f);
    const double whir = 3.0 * m1 * 32.0 + 2.0 * std::ldexp(8.0, f) * rows1 * 4.0 + rows1 * 64.0 + 4.0 * m1 * 32.0;
    return (uint64_t)(packed + commit + zc + inner + whir) + ((uint64_t)64 << 20);
}

wl_engine *wl_engine_create(int device, int n_slots, uint64_t mem_budget_bytes) {
    if (n_slots < 1 || n_slots > 16) { wl_set_error("n_slots must be in [1, 16]"); return nullptr; }
    auto *e = new wl_engine();
    e->device = device;
    e->n_slots = n_slots;
    if (mem_budget_bytes == 0) {
        size_t fr = 0, tot = 0;
        if (hipSetDevice(device) != hipSuccess || hipMemGetInfo(&fr, &tot) != hipSuccess) {
            wl_set_error("cannot query device memory");
            delete e;
            return nullptr;
        }
        mem_budget_bytes = (uint64_t)(fr * 0.8);
    }
    e->budget = mem_budget_bytes;
    for (int i = 0; i < n_slots; i++) {
        wl_ctx *c = wl_create(device);
        // slots enqueue each sumcheck round once its challenge is known: no kernel
        // spins on the GPU for a host value, so nothing else in the process can
        // stall a slot's proof (same throughput as device waits, measured)
        if (c) wl_set_device_waits(c, 0);
        if (!c) {
            for (wl_ctx *x : e->ctxs) wl_destroy(x);
            delete e;
            return nullptr;
        }
        e->ctxs.push_back(c);
    }
    for (int i = 0; i < n_slots; i++) e->workers.emplace_back([e, i] { e->worker(i); });
    return e;
}

void wl_engine_destroy(wl_engine *e) {
    if (!e) return;
    {
        std::lock_guard<std::mutex> lk(e->mu);
        e->stop = true;
        // jobs not yet admitted complete with an error, so their waiters return
        for (auto &j : e->queue) {
            j->status = WL_ERR_ARG;
            j->err = "engine stopped before the job was admitted";
            j->t_start = j->t_done = clk::now();
            j->done = true;
        }
        e->queue.clear();
    }
    e->cv_work.notify_all();
    e->cv_done.notify_all();
    for (auto &t : e->workers) t.join();
    {
        // running jobs have completed and notified; wait until every waiter has left
        std::unique_lock<std::mutex> lk(e->mu);
        e->cv_done.notify_all();
        e->cv_done.wait(lk, [&] { return e->waiters == 0; });
    }
    for (wl_ctx *c : e->ctxs) wl_destroy(c);
    delete e;
}

int wl_engine_submit(wl_engine *e, const wl_job *job, uint64_t *ticket) {
    if (!e || !job || !ticket || !job->air_rc || !job->inputs || !job->p16 || !job->p24) {
        wl_set_error("null argument");
        return WL_ERR_ARG;
    }
    if (job->log_n > 24) {
        wl_set_error("unsupported log_n");
        return WL_ERR_ARG;
    }
    if (wl_air_settings_validate(job->log_n, 165, &job->settings) != WL_OK) return WL_ERR_ARG;
    auto j = std::make_shared<Job>();
    j->spec = *job;
    j->footprint = wl_job_footprint(job->log_n, &job->settings);
    if (j->footprint > e->budget) {
        wl_set_error("job footprint exceeds the engine's memory budget");
        return WL_ERR_ARG;
    }
    {
        std::lock_guard<std::mutex> lk(e->mu);
        if (e->stop) { wl_set_error("engine stopped"); return WL_ERR_ARG; }
        j->id = e->next_id++;
        j->t_submit = clk::now();
        e->jobs[j->id] = j;
        e->queue.push_back(j);
    }
    e->cv_work.notify_all();
    *ticket = j->id;
    return WL_OK;
}

int wl_engine_poll(wl_engine *e, uint64_t ticket) {
    if (!e) return -1;
    std::lock_guard<std::mutex> lk(e->mu);
    auto it = e->jobs.find(ticket);
    if (it == e->jobs.end()) return -1;
    return it->second->done ? 1 : 0;
}

int wl_engine_wait(wl_engine *e, uint64_t ticket, uint32_t *proof_out, size_t proof_cap, size_t *proof_len,
                   double *times_ms) {
    if (!e) { wl_set_error("null engine"); return WL_ERR_ARG; }
    std::shared_ptr<Job> j;
    // every access to the engine happens under its lock, before `waiters` drops:
    // wl_engine_destroy frees the engine once no waiter is left
    std::unique_lock<std::mutex> lk(e->mu);
    auto it = e->jobs.find(ticket);
    if (it == e->jobs.end()) { wl_set_error("unknown ticket"); return WL_ERR_ARG; }
    j = it->second;
    e->waiters++;
    e->cv_done.wait(lk, [&] { return j->done; });
    int rc = j->status;
    if (rc == WL_OK && proof_out && proof_cap < j->proof.size()) rc = WL_ERR_BUFFER;  // the ticket stays
    else e->jobs.erase(ticket);
    e->waiters--;
    if (e->waiters == 0 && e->stop) e->cv_done.notify_all();  // destroy may be waiting for this
    lk.unlock();
    if (proof_len) *proof_len = j->proof.size();
    if (times_ms) {
        times_ms[0] = ms_between(j->t_submit, j->t_start);  // queued
        times_ms[1] = ms_between(j->t_start, j->t_done);    // running
        times_ms[2] = ms_between(j->t_submit, j->t_done);   // latency
    }
    if (rc == WL_ERR_BUFFER) { wl_set_error("proof buffer too small"); return rc; }
    if (rc != WL_OK) { wl_set_error(j->err.c_str()); return rc; }
    if (proof_out) memcpy(proof_out, j->proof.data(), j->proof.size() * 4);
    return WL_OK;
}

void wl_engine_stats(wl_engine *e, uint64_t *out4) {
    if (!e || !out4) return;
    std::lock_guard<std::mutex> lk(e->mu);
    out4[0] = e->completed;
    out4[1] = (uint64_t)e->peak_running;
    out4[2] = e->peak_reserved;
    out4[3] = e->budget;
}

}  // extern "C"
