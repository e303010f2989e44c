// cd_server.hip -- the served form: continuous batching of codon-design jobs.
//
// Requests are queued in arrival order.  A worker thread admits the head of
// the queue whenever the device bytes of the resident jobs plus the new
// job's fit the budget (and a slot is free), creates and initialises it, and
// advances every resident job by one generation per tick with one launch
// sequence over all of them (the batched kernels).  A job leaves at the end
// of the tick that completes its generation budget: its final pool and ranks
// are copied to host memory and its device memory is freed, which may admit
// waiting requests at the next tick.  Jobs are independent (each is keyed by
// its own Philox seed and keeps its own state), so a job's result does not
// depend on what else is resident.
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
    int64_t bytes = 0;  // device bytes the job holds while resident

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
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    bool stop = false;
    int64_t budget = 0, used = 0;
    int32_t max_jobs = 0;
    int64_t next_ticket = 1;
    std::deque<int64_t> queue;  // tickets waiting for admission, oldest first
    std::map<int64_t, std::unique_ptr<Entry>> entries;
    int32_t resident_n = 0;
    int64_t ticks = 0, designs = 0;
    hipStream_t stream = nullptr;
    // device work lists of the resident set, grown on demand
    cd::DevJob *d_jobs = nullptr;
    int2 *d_off = nullptr, *d_asc = nullptr;
    int4 *d_dom = nullptr;
    uint32_t *d_slist = nullptr;  // changed CDS of a tick
    int32_t *d_sctl = nullptr;    // its count, zero between ticks
    size_t cap_jobs = 0, cap_off = 0, cap_dom = 0, cap_asc = 0, cap_slist = 0;
    std::string fatal;  // a device error that stopped the worker
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

// Upload the work lists of `res` and return the launch description.
hipError_t rebuild(cd_server *S, const std::vector<Entry *> &res, cdhost::WorkLists &wl, cdhost::Launch &L) {
    std::vector<cd_job *> jobs;
    for (Entry *E : res) jobs.push_back(E->job);
    cd::Shapes sh = cd::SHAPES_BATCH;  // the tuned shapes of this kind of resident set, if any
    (void)cdhost::cached_shapes(cdhost::work_signature(jobs.data(), (int)jobs.size()), &sh);
    cdhost::build_lists(jobs.data(), (int)jobs.size(), sh, wl);
    hipError_t e = ensure(S->d_jobs, S->cap_jobs, wl.descs.size());
    if (e == hipSuccess) e = ensure(S->d_off, S->cap_off, wl.off.size());
    if (e == hipSuccess) e = ensure(S->d_dom, S->cap_dom, wl.dom.size());
    if (e == hipSuccess) e = ensure(S->d_asc, S->cap_asc, wl.asc.size());
    if (e == hipSuccess) e = ensure(S->d_slist, S->cap_slist, (size_t)wl.n_cds);
    if (e == hipSuccess)
        e = hipMemcpyAsync(S->d_jobs, wl.descs.data(), sizeof(cd::DevJob) * wl.descs.size(), hipMemcpyHostToDevice,
                           S->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(S->d_off, wl.off.data(), sizeof(int2) * wl.off.size(), hipMemcpyHostToDevice, S->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(S->d_dom, wl.dom.data(), sizeof(int4) * wl.dom.size(), hipMemcpyHostToDevice, S->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(S->d_asc, wl.asc.data(), sizeof(int2) * wl.asc.size(), hipMemcpyHostToDevice, S->stream);
    if (e == hipSuccess) e = cdhost::set_lds_limits(wl);
    L = cdhost::Launch{S->d_jobs,   (int)wl.descs.size(), S->d_off,      (int)wl.off.size(), S->d_dom,
                       (int)wl.dom.size(), S->d_asc,      (int)wl.asc.size(), wl.lds,         wl.stem_lds,
                       wl.niche_lds, wl.mg,               wl.stem_parts,      S->d_slist,     S->d_sctl,
                       wl.n_cds,     wl.n_small,          wl.lds_small,       sh};
    return e;
}

void finish(cd_server *S, Entry *E, int state, const std::string &err) {
    if (E->job) {
        cd_job_destroy(E->job);
        E->job = nullptr;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    E->state = state;
    E->error = err;
    E->t_done = Clock::now();
    S->used -= E->req.bytes;
    S->cv_done.notify_all();
}

void run(cd_server *S) {
    std::vector<Entry *> res;  // resident, in admission order
    cdhost::WorkLists wl;
    cdhost::Launch launch{};
    bool dirty = true;
    for (;;) {
        std::vector<Entry *> admit;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] { return S->stop || !res.empty() || !S->queue.empty(); });
            if (S->stop) break;
            // FIFO admission by device-memory budget and slot count; a request
            // never exceeds the budget alone (checked at submission)
            while (!S->queue.empty() && (int)(res.size() + admit.size()) < S->max_jobs) {
                Entry *E = S->entries[S->queue.front()].get();
                if (S->used + E->req.bytes > S->budget) break;
                S->used += E->req.bytes;
                S->queue.pop_front();
                admit.push_back(E);
            }
        }
        for (Entry *E : admit) {
            int rc = cd_job_create(&E->req.desc, &E->job);
            if (rc == CD_OK) rc = cd_job_init(E->job);
            if (rc == CD_OK) rc = cd_job_sync(E->job);
            if (rc == CD_OK && E->req.generations == 0) rc = collect(E);
            if (rc != CD_OK) {
                finish(S, E, CD_SERVER_FAILED, cd_last_error());
                continue;
            }
            {
                std::lock_guard<std::mutex> lk(S->mu);
                E->t_admit = Clock::now();
                E->state = CD_SERVER_RUNNING;
            }
            if (E->req.generations == 0) {
                finish(S, E, CD_SERVER_DONE, "");
                continue;
            }
            res.push_back(E);
            dirty = true;
        }
        if (res.empty()) continue;
        hipError_t e = hipSuccess;
        if (dirty) {
            e = rebuild(S, res, wl, launch);
            dirty = false;
        }
        if (e == hipSuccess) e = cdhost::enqueue_generation(launch, S->stream, nullptr);
        if (e == hipSuccess) e = hipStreamSynchronize(S->stream);
        if (e != hipSuccess) {  // a device error ends every resident job and the worker
            const std::string msg = std::string("serving loop: ") + hipGetErrorString(e);
            for (Entry *E : res) finish(S, E, CD_SERVER_FAILED, msg);
            std::lock_guard<std::mutex> lk(S->mu);
            S->fatal = msg;
            for (int64_t t : S->queue) {
                S->entries[t]->state = CD_SERVER_FAILED;
                S->entries[t]->error = msg;
            }
            S->queue.clear();
            S->resident_n = 0;
            S->cv_done.notify_all();
            return;
        }
        int64_t designs = 0;
        for (Entry *E : res) designs += E->req.desc.N;
        std::vector<Entry *> keep;
        for (Entry *E : res) {
            if (++E->gens_done < E->req.generations) {
                keep.push_back(E);
                continue;
            }
            const int rc = collect(E);
            finish(S, E, rc == CD_OK ? CD_SERVER_DONE : CD_SERVER_FAILED, rc == CD_OK ? "" : cd_last_error());
            dirty = true;
        }
        res.swap(keep);
        std::lock_guard<std::mutex> lk(S->mu);
        S->ticks += 1;
        S->designs += designs;
        S->resident_n = (int32_t)res.size();
    }
    for (Entry *E : res) finish(S, E, CD_SERVER_FAILED, "server destroyed");
}

}  // namespace

extern "C" {

int cd_server_create(int64_t hbm_budget_bytes, int32_t max_jobs, cd_server **out) {
    if (!out || hbm_budget_bytes <= 0 || max_jobs <= 0) return cdhost::fail(CD_E_INVALID, "bad server arguments");
    cd_server *S = new cd_server();
    S->budget = hbm_budget_bytes;
    S->max_jobs = max_jobs;
    hipError_t e = hipStreamCreate(&S->stream);
    if (e == hipSuccess) e = hipMalloc(&S->d_sctl, 16);
    if (e == hipSuccess) e = hipMemset(S->d_sctl, 0, 16);
    if (e != hipSuccess) {
        if (S->stream) (void)hipStreamDestroy(S->stream);
        if (S->d_sctl) (void)hipFree(S->d_sctl);
        delete S;
        return cdhost::fail(CD_E_DEVICE, std::string("server stream: ") + hipGetErrorString(e));
    }
    S->worker = std::thread(run, S);
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
    S->cv_work.notify_one();
    return CD_OK;
}

int cd_server_wait(cd_server *S, int64_t ticket, double timeout_s, int32_t *state) {
    if (!S || !state) return cdhost::fail(CD_E_INVALID, "bad wait arguments");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->entries.find(ticket);
    if (it == S->entries.end()) return cdhost::fail(CD_E_INVALID, "unknown ticket");
    Entry *E = it->second.get();
    auto ready = [&] { return E->state == CD_SERVER_DONE || E->state == CD_SERVER_FAILED; };
    if (timeout_s < 0)
        S->cv_done.wait(lk, ready);
    else
        S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready);
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
    return CD_OK;
}

int cd_server_destroy(cd_server *S) {
    if (!S) return CD_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv_work.notify_all();
    }
    if (S->worker.joinable()) S->worker.join();
    for (auto &kv : S->entries)
        if (kv.second->job) cd_job_destroy(kv.second->job);
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
