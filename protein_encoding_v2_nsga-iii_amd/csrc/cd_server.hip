// cd_server.hip -- the served form: continuous batching of codon-design jobs.
//
// Three threads keep the GPU's tick stream busy:
//   admitter   takes requests in arrival order whenever their device bytes fit
//              the budget beside the jobs that hold memory (and a slot is
//              free), creates each job and enqueues its initialisation on the
//              job's own stream, and hands it over with an event;
//   worker     advances every resident job one generation per tick: one
//              hipGraph captured per resident set (captured again when the set
//              changes), at most two ticks in flight; a job whose
//              initialisation event has completed joins at the next tick, a
//              job leaves after the tick that completes its generation budget;
//   collector  waits for a finished job's last tick, copies its final pool and
//              ranks to host memory and frees its device memory, which may
//              let the admitter admit waiting requests.
// Jobs are independent (each is keyed by its own Philox seed and keeps its
// own state), so a job's result does not depend on what else is resident or
// when it was admitted.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "cd_host.h"

namespace {

using Clock = std::chrono::steady_clock;

// A job description with owned copies of every array it points to.
struct Request {
    cd_job_desc desc{};
    std::vector<int8_t> aa;
    std::string codons;
    std::vector<float> weights, cps;
    int32_t generations = 0;
    int64_t bytes = 0;  // device bytes the job holds while admitted

    void own(const cd_job_desc &d) {
        desc = d;
        aa.assign(d.aa, d.aa + d.L);
        codons.assign(d.tables.codons, 192);
        weights.assign(d.tables.weights, d.tables.weights + 64);
        cps.assign(d.tables.cps, d.tables.cps + 61 * 61);
        desc.aa = aa.data();
        desc.tables.codons = codons.c_str();
        desc.tables.weights = weights.data();
        desc.tables.cps = cps.data();
    }
};

struct Entry {
    Request req;
    int state = CD_SERVER_QUEUED;
    cd_job *job = nullptr;
    hipEvent_t ready = nullptr;  // initialisation done (admitter -> worker)
    hipEvent_t last = nullptr;   // last tick done (worker -> collector)
    int gens_done = 0;
    Clock::time_point t_submit, t_admit, t_done;
    std::vector<int8_t> genome, idx;
    std::vector<float> obj;
    std::vector<int32_t> pql, sorted, front_of;
    std::string error;
};

double seconds(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

}  // namespace

struct cd_server {
    std::mutex mu;
    // held by the worker while it captures a tick, and by the other threads
    // around the HIP calls that create, initialise, collect or destroy a job
    // (stream and event creation, waits and destruction, stream-ordered
    // allocation): such calls from another thread can end a capture in
    // progress
    std::mutex capture_mu;
    std::condition_variable cv_admit, cv_work, cv_collect, cv_done;
    std::thread admitter, worker, collector;
    bool stop = false;
    int64_t budget = 0, used = 0;
    int32_t max_jobs = 0, admitted = 0;  // admitted: jobs holding device memory
    int64_t next_ticket = 1;
    std::deque<int64_t> queue;        // tickets waiting for admission, oldest first
    std::deque<Entry *> ready, done;  // admitter -> worker, worker -> collector
    std::map<int64_t, std::unique_ptr<Entry>> entries;
    int32_t resident_n = 0;
    int64_t ticks = 0, designs = 0;
    double tick_ms = 0.0;  // GPU time of the ticks (events around each launch)
    double busy_s = 0.0;   // worker wall time with jobs resident
    int64_t rebuilds = 0;  // resident-set changes (lists uploaded, tick captured again)
    double rebuild_s = 0.0;
    double launch_s = 0.0;  // worker time launching ticks
    double wait_s = 0.0;    // worker time waiting for the tick two back
    hipStream_t stream = nullptr;
    // device work lists of the resident set, grown on demand
    cd::DevJob *d_jobs = nullptr;
    int2 *d_off = nullptr, *d_asc = nullptr;
    int4 *d_dom = nullptr;
    uint32_t *d_slist = nullptr;  // changed CDS of a tick (stem_list_entries)
    int32_t *d_sctl = nullptr;    // their counts (STEM_MAX_GROUPS), zero between ticks
    size_t cap_jobs = 0, cap_off = 0, cap_dom = 0, cap_asc = 0, cap_slist = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::string fatal;        // a device error that stopped the server
    std::string fatal_stage;  // the step of the tick's capture that failed, if it was one (worker only)
};

namespace {

template <class T>
hipError_t ensure(T *&p, size_t &cap, size_t need) {
    if (need <= cap) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t c = std::max(need, (size_t)64);
    c += c / 2;
    hipError_t e = hipMalloc(&p, sizeof(T) * c);
    if (e == hipSuccess) cap = c;
    return e;
}

// Copy the final pool of a finished job into its entry.
int collect(Entry *E) {
    cd_job *J = E->job;
    const size_t P = 2 * (size_t)E->req.desc.N, m = E->req.desc.m, L = E->req.desc.L;
    E->genome.resize(P * m * L);
    E->obj.resize(P * 6);
    E->idx.resize(P * 12);
    E->pql.resize(P * 3);
    E->sorted.resize(E->req.desc.N);
    E->front_of.resize(P);
    return cd_job_read(J, E->genome.data(), E->obj.data(), E->idx.data(), E->pql.data(), E->sorted.data(),
                       E->front_of.data(), nullptr);
}

// Release a job's device memory and end its request.
void finish(cd_server *S, Entry *E, int state, const std::string &err) {
    {
        std::lock_guard<std::mutex> cap(S->capture_mu);
        if (E->job) {
            cd_job_destroy(E->job);
            E->job = nullptr;
        }
        if (E->ready) (void)hipEventDestroy(E->ready);
        if (E->last) (void)hipEventDestroy(E->last);
        E->ready = E->last = nullptr;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    E->state = state;
    E->error = err;
    E->t_done = Clock::now();
    S->used -= E->req.bytes;
    S->admitted -= 1;
    S->cv_done.notify_all();
    S->cv_admit.notify_all();
}

// Stop serving after a device error: every request that has not ended fails.
void fail_all(cd_server *S, const std::string &msg) {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->fatal.empty()) S->fatal = msg;
    for (auto &kv : S->entries) {
        Entry *E = kv.second.get();
        if (E->state == CD_SERVER_QUEUED || E->state == CD_SERVER_RUNNING) {
            E->state = CD_SERVER_FAILED;
            E->error = msg;
        }
    }
    S->queue.clear();
    S->stop = true;
    S->cv_done.notify_all();
    S->cv_admit.notify_all();
    S->cv_work.notify_all();
    S->cv_collect.notify_all();
}

void admit_loop(cd_server *S) {
    for (;;) {
        Entry *E = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            // FIFO admission by device-memory budget and slot count; a request
            // never exceeds the budget alone (checked at submission)
            S->cv_admit.wait(lk, [&] {
                if (S->stop) return true;
                if (S->queue.empty() || S->admitted >= S->max_jobs) return false;
                return S->used + S->entries[S->queue.front()]->req.bytes <= S->budget;
            });
            if (S->stop) return;
            E = S->entries[S->queue.front()].get();
            S->queue.pop_front();
            S->used += E->req.bytes;
            S->admitted += 1;
        }
        int rc = CD_OK;
        {
            std::lock_guard<std::mutex> cap(S->capture_mu);
            rc = cd_job_create(&E->req.desc, &E->job);
            if (rc == CD_OK) rc = cd_job_init(E->job);  // enqueued on the job's own stream
            if (rc == CD_OK && hipEventCreateWithFlags(&E->ready, hipEventDisableTiming) != hipSuccess)
                rc = CD_E_DEVICE;
            if (rc == CD_OK && hipEventRecord(E->ready, E->job->stream) != hipSuccess) rc = CD_E_DEVICE;
        }
        if (rc != CD_OK) {
            finish(S, E, CD_SERVER_FAILED, cd_last_error());
            continue;
        }
        std::lock_guard<std::mutex> lk(S->mu);
        E->t_admit = Clock::now();
        E->state = CD_SERVER_RUNNING;
        S->ready.push_back(E);
        S->cv_work.notify_all();
    }
}

void collect_loop(cd_server *S) {
    for (;;) {
        Entry *E = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_collect.wait(lk, [&] { return S->stop || !S->done.empty(); });
            if (S->done.empty()) return;  // stopping with nothing left to collect
            E = S->done.front();
            S->done.pop_front();
        }
        int rc = CD_OK;
        {
            // E->last was recorded on the tick stream, which the worker may be
            // capturing: waiting on it then would end the capture
            std::lock_guard<std::mutex> cap(S->capture_mu);
            if (E->last && hipEventSynchronize(E->last) != hipSuccess) rc = CD_E_DEVICE;
            if (rc == CD_OK) rc = collect(E);
        }
        finish(S, E, rc == CD_OK ? CD_SERVER_DONE : CD_SERVER_FAILED, rc == CD_OK ? "" : cd_last_error());
    }
}

// Upload the work lists of `res` and capture one tick's launch sequence.
hipError_t rebuild(cd_server *S, const std::vector<Entry *> &res) {
    std::vector<cd_job *> jobs;
    for (Entry *E : res) jobs.push_back(E->job);
    cd::Shapes sh = cd::SHAPES_BATCH;  // the tuned shapes of this kind of resident set, if any
    (void)cdhost::cached_shapes(cdhost::work_signature(jobs.data(), (int)jobs.size()), &sh);
    cdhost::WorkLists wl;
    cdhost::build_lists(jobs.data(), (int)jobs.size(), sh, wl);
    hipError_t e = hipStreamSynchronize(S->stream);  // no tick in flight reads the old lists or graph
    if (e == hipSuccess) e = ensure(S->d_jobs, S->cap_jobs, wl.descs.size());
    if (e == hipSuccess) e = ensure(S->d_off, S->cap_off, wl.off.size());
    if (e == hipSuccess) e = ensure(S->d_dom, S->cap_dom, wl.dom.size());
    if (e == hipSuccess) e = ensure(S->d_asc, S->cap_asc, wl.asc.size());
    if (e == hipSuccess) e = ensure(S->d_slist, S->cap_slist, cd::stem_list_entries((int)wl.off.size(), wl.mg));
    if (e == hipSuccess)
        e = hipMemcpy(S->d_jobs, wl.descs.data(), sizeof(cd::DevJob) * wl.descs.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(S->d_off, wl.off.data(), sizeof(int2) * wl.off.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(S->d_dom, wl.dom.data(), sizeof(int4) * wl.dom.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(S->d_asc, wl.asc.data(), sizeof(int2) * wl.asc.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = cdhost::set_lds_limits(wl);
    if (e != hipSuccess) return e;
    const cdhost::Launch L{S->d_jobs,   (int)wl.descs.size(), S->d_off,      (int)wl.off.size(), S->d_dom,
                           (int)wl.dom.size(), S->d_asc,      (int)wl.asc.size(), wl.lds,         wl.stem_lds,
                           wl.max_pool,  wl.max_refs,        wl.mg,               wl.stem_parts,      S->d_slist,     S->d_sctl,
                           wl.n_cds,     wl.n_small,          wl.lds_small,       sh};
    if (S->gexec) (void)hipGraphExecDestroy(S->gexec);
    if (S->graph) (void)hipGraphDestroy(S->graph);
    S->gexec = nullptr;
    S->graph = nullptr;
    // job creation, initialisation, collection and destruction on the other
    // threads wait for the capture (capture_mu)
    std::lock_guard<std::mutex> cap(S->capture_mu);
    e = hipStreamBeginCapture(S->stream, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) {
        S->fatal_stage = "begin capture";
        return e;
    }
    const hipError_t ec = cdhost::enqueue_generation(L, S->stream, nullptr);
    e = hipStreamEndCapture(S->stream, &S->graph);
    if (ec != hipSuccess) {
        S->fatal_stage = "capture";
        return ec;
    }
    if (e != hipSuccess) {
        S->fatal_stage = "end capture";
        return e;
    }
    e = hipGraphInstantiate(&S->gexec, S->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) S->fatal_stage = "graph instantiation";
    return e;
}

void work_loop(cd_server *S) {
    std::vector<Entry *> res;  // resident, in the order they joined
    bool dirty = true;
    hipEvent_t ev[2][2] = {};  // per in-flight slot: start and end of a tick
    for (auto &p : ev)
        for (auto &x : p)
            if (hipEventCreate(&x) != hipSuccess) {
                fail_all(S, "serving loop: event creation failed");
                return;
            }
    int slot = 0;
    bool inflight[2] = {false, false};
    auto retire = [&](int k) -> hipError_t {  // wait for the tick in slot k and account its GPU time
        if (!inflight[k]) return hipSuccess;
        hipError_t e = hipEventSynchronize(ev[k][1]);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[k][0], ev[k][1]);
        inflight[k] = false;
        std::lock_guard<std::mutex> lk(S->mu);
        S->tick_ms += ms;
        return e;
    };
    for (;;) {
        Clock::time_point t0;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] { return S->stop || !res.empty() || !S->ready.empty(); });
            if (S->stop) break;
            t0 = Clock::now();
            // jobs whose initialisation has completed join the resident set;
            // with nothing resident, the worker waits for the oldest one
            for (auto it = S->ready.begin(); it != S->ready.end();) {
                hipError_t q = hipEventQuery((*it)->ready);
                if (q == hipErrorNotReady && !res.empty()) {
                    ++it;
                    continue;
                }
                Entry *E = *it;
                it = S->ready.erase(it);
                if (q == hipErrorNotReady) {
                    lk.unlock();
                    q = hipEventSynchronize(E->ready);
                    lk.lock();
                }
                if (q != hipSuccess) {
                    lk.unlock();
                    fail_all(S, std::string("serving loop: ") + hipGetErrorString(q));
                    lk.lock();
                    break;
                }
                if (E->req.generations == 0) {  // nothing to run: straight to the collector
                    S->done.push_back(E);
                    S->cv_collect.notify_all();
                    continue;
                }
                res.push_back(E);
                dirty = true;
            }
            if (S->stop) break;
            S->resident_n = (int32_t)res.size();
        }
        if (res.empty()) continue;
        hipError_t e = hipSuccess;
        if (dirty) {
            const Clock::time_point r0 = Clock::now();
            for (int k = 0; k < 2 && e == hipSuccess; ++k) e = retire(k);
            if (e == hipSuccess) e = rebuild(S, res);
            dirty = false;
            std::lock_guard<std::mutex> lk(S->mu);
            S->rebuilds += 1;
            S->rebuild_s += std::chrono::duration<double>(Clock::now() - r0).count();
        }
        const Clock::time_point w0 = Clock::now();
        if (e == hipSuccess) e = retire(slot);  // at most two ticks in flight
        const Clock::time_point w1 = Clock::now();
        if (e == hipSuccess) e = hipEventRecord(ev[slot][0], S->stream);
        if (e == hipSuccess) e = hipGraphLaunch(S->gexec, S->stream);
        if (e == hipSuccess) e = hipEventRecord(ev[slot][1], S->stream);
        const Clock::time_point w2 = Clock::now();
        if (e != hipSuccess) {  // a device error ends every request and the server
            fail_all(S, std::string("serving loop: ") + (S->fatal_stage.empty() ? "" : S->fatal_stage + ": ") +
                            hipGetErrorString(e));
            break;
        }
        inflight[slot] = true;
        slot ^= 1;
        int64_t designs = 0;
        std::vector<Entry *> keep, leaving;
        for (Entry *E : res) {
            designs += E->req.desc.N;
            if (++E->gens_done < E->req.generations)
                keep.push_back(E);
            else
                leaving.push_back(E);
        }
        for (Entry *E : leaving) {  // the collector waits for this tick, then reads the pool
            if (hipEventCreateWithFlags(&E->last, hipEventDisableTiming) != hipSuccess ||
                hipEventRecord(E->last, S->stream) != hipSuccess)
                e = hipErrorUnknown;
            dirty = true;
        }
        res.swap(keep);
        {
            std::lock_guard<std::mutex> lk(S->mu);
            for (Entry *E : leaving) S->done.push_back(E);
            if (!leaving.empty()) S->cv_collect.notify_all();
            S->ticks += 1;
            S->designs += designs;
            S->resident_n = (int32_t)res.size();
            S->busy_s += std::chrono::duration<double>(Clock::now() - t0).count();
            S->wait_s += std::chrono::duration<double>(w1 - w0).count();
            S->launch_s += std::chrono::duration<double>(w2 - w1).count();
        }
        if (e != hipSuccess) {
            fail_all(S, "serving loop: event record failed");
            break;
        }
    }
    for (int k = 0; k < 2; ++k) (void)retire(k);
    // stopping: the resident jobs end unfinished
    for (Entry *E : res) finish(S, E, CD_SERVER_FAILED, "server stopped");
    for (auto &p : ev)
        for (auto &x : p) (void)hipEventDestroy(x);
}

}  // namespace

extern "C" {

int cd_server_create(int64_t hbm_budget_bytes, int32_t max_jobs, cd_server **out) {
    if (!out || hbm_budget_bytes <= 0 || max_jobs <= 0) return cdhost::fail(CD_E_INVALID, "bad server arguments");
    cd_server *S = new cd_server();
    S->budget = hbm_budget_bytes;
    S->max_jobs = max_jobs;
    hipError_t e = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&S->d_sctl, sizeof(int32_t) * cd::STEM_MAX_GROUPS);
    if (e == hipSuccess) e = hipMemset(S->d_sctl, 0, sizeof(int32_t) * cd::STEM_MAX_GROUPS);
    if (e != hipSuccess) {
        if (S->stream) (void)hipStreamDestroy(S->stream);
        if (S->d_sctl) (void)hipFree(S->d_sctl);
        delete S;
        return cdhost::fail(CD_E_DEVICE, std::string("server stream: ") + hipGetErrorString(e));
    }
    S->admitter = std::thread(admit_loop, S);
    S->worker = std::thread(work_loop, S);
    S->collector = std::thread(collect_loop, S);
    *out = S;
    return CD_OK;
}

int cd_server_submit(cd_server *S, const cd_job_desc *desc, int32_t generations, int64_t *ticket) {
    if (!S || !desc || !ticket || generations < 0) return cdhost::fail(CD_E_INVALID, "bad submit arguments");
    if (!desc->aa || !desc->tables.codons || !desc->tables.weights || !desc->tables.cps || desc->L < 3)
        return cdhost::fail(CD_E_INVALID, "incomplete job description");
    const int64_t bytes = cd_job_device_bytes(desc->L, desc->m, desc->N, desc->H);
    if (bytes <= 0) return cdhost::fail(CD_E_INVALID, "job sizes out of range");
    std::unique_ptr<Entry> E(new Entry());
    E->req.own(*desc);
    E->req.generations = generations;
    E->req.bytes = bytes;
    E->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(S->mu);
    if (!S->fatal.empty()) return cdhost::fail(CD_E_DEVICE, S->fatal);
    if (bytes > S->budget) return cdhost::fail(CD_E_BUDGET, "the job alone exceeds the server's memory budget");
    const int64_t t = S->next_ticket++;
    S->entries[t] = std::move(E);
    S->queue.push_back(t);
    *ticket = t;
    S->cv_admit.notify_all();
    return CD_OK;
}

int cd_server_wait(cd_server *S, int64_t ticket, double timeout_s, int32_t *state) {
    if (!S || !state) return cdhost::fail(CD_E_INVALID, "bad wait arguments");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->entries.find(ticket);
    if (it == S->entries.end()) return cdhost::fail(CD_E_INVALID, "unknown ticket");
    Entry *E = it->second.get();
    auto ended = [&] { return E->state == CD_SERVER_DONE || E->state == CD_SERVER_FAILED; };
    if (timeout_s < 0)
        S->cv_done.wait(lk, ended);
    else
        S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ended);
    *state = E->state;
    return CD_OK;
}

int cd_server_result(cd_server *S, int64_t ticket, int8_t *genome, float *obj, int8_t *idx, int32_t *pql,
                     int32_t *sorted, int32_t *front_of, double *times) {
    if (!S) return cdhost::fail(CD_E_INVALID, "null server");
    std::unique_ptr<Entry> E;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->entries.find(ticket);
        if (it == S->entries.end()) return cdhost::fail(CD_E_INVALID, "unknown ticket");
        const int st = it->second->state;
        if (st == CD_SERVER_FAILED) {
            if (it->second->job) return cdhost::fail(CD_E_STATE, "job failed and is still being released");
            const std::string err = it->second->error;
            S->entries.erase(it);
            return cdhost::fail(CD_E_DEVICE, "job failed: " + err);
        }
        if (st != CD_SERVER_DONE) return cdhost::fail(CD_E_STATE, "job not finished");
        E = std::move(it->second);
        S->entries.erase(it);
    }
    auto put = [](auto *dst, const auto &v) {
        if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    put(genome, E->genome);
    put(obj, E->obj);
    put(idx, E->idx);
    put(pql, E->pql);
    put(sorted, E->sorted);
    put(front_of, E->front_of);
    if (times) {
        times[0] = seconds(E->t_submit, E->t_admit);  // queued
        times[1] = seconds(E->t_admit, E->t_done);    // resident
        times[2] = seconds(E->t_submit, E->t_done);   // time to front
    }
    return CD_OK;
}

int cd_server_stats(cd_server *S, int64_t *stats) {
    if (!S || !stats) return cdhost::fail(CD_E_INVALID, "bad stats arguments");
    std::lock_guard<std::mutex> lk(S->mu);
    stats[0] = (int64_t)S->queue.size();
    stats[1] = S->resident_n;
    stats[2] = S->used;
    stats[3] = S->ticks;
    stats[4] = S->designs;
    stats[5] = (int64_t)(S->tick_ms * 1000.0);
    stats[6] = (int64_t)(S->busy_s * 1e6);
    stats[7] = S->rebuilds;
    stats[8] = (int64_t)(S->rebuild_s * 1e6);
    stats[9] = (int64_t)(S->launch_s * 1e6);
    stats[10] = (int64_t)(S->wait_s * 1e6);
    return CD_OK;
}

int cd_server_destroy(cd_server *S) {
    if (!S) return CD_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv_admit.notify_all();
        S->cv_work.notify_all();
    }
    if (S->admitter.joinable()) S->admitter.join();
    if (S->worker.joinable()) S->worker.join();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->cv_collect.notify_all();
    }
    if (S->collector.joinable()) S->collector.join();
    for (auto &kv : S->entries) {  // admitted but never resident, or never collected
        Entry *E = kv.second.get();
        if (E->job) cd_job_destroy(E->job);
        if (E->ready) (void)hipEventDestroy(E->ready);
        if (E->last) (void)hipEventDestroy(E->last);
    }
    if (S->gexec) (void)hipGraphExecDestroy(S->gexec);
    if (S->graph) (void)hipGraphDestroy(S->graph);
    if (S->stream) (void)hipStreamDestroy(S->stream);
    if (S->d_jobs) (void)hipFree(S->d_jobs);
    if (S->d_off) (void)hipFree(S->d_off);
    if (S->d_dom) (void)hipFree(S->d_dom);
    if (S->d_asc) (void)hipFree(S->d_asc);
    if (S->d_slist) (void)hipFree(S->d_slist);
    if (S->d_sctl) (void)hipFree(S->d_sctl);
    delete S;
    return CD_OK;
}

}  // extern "C"
