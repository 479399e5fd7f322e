// server.cpp -- the serving loop (rj_server_* in include/rasterjoin.h).
//
// Clients submit self-contained queries from any thread; one server thread
// per server admits them in arrival order while device memory allows (the
// memory model: the label maps, grid indexes and results a query would add,
// against the budget minus what queries in flight hold), prepares their
// polygon structures, and runs them as fused passes. Two pass slots form a
// pipeline: the pass of one group of queries runs on the device while the
// thread admits and prepares the next group and launches its pass behind it,
// so queries that arrive while others are on the device join the next pass
// (continuous batching). Each query's results are the ones it gets alone: the
// fused pass computes every query's sums independently and exactly.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "engine.hpp"

using namespace rj;

namespace {

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace

struct rj_server {
    rj_ctx* ctx = nullptr;
    rj_table* t = nullptr;
    int max_batch = kBatchMax;
    size_t max_pending = 65536;
    int64_t saved_limit = 0;

    struct Req {
        QuerySpec spec;
        rj_result* out = nullptr;
        double t_submit = 0, t_done = 0;
        int status = RJ_OK;
        bool done = false;
    };
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::map<uint64_t, Req> reqs;
    std::deque<uint64_t> queue;
    uint64_t next_id = 1;
    bool stop = false;
    std::thread th;

    // statistics
    int64_t submitted = 0, completed = 0, failed = 0, batches = 0, batched = 0;
    std::vector<double> lat;
    double first_submit = -1, last_done = 0;

    // a group of admitted queries and the passes it still has on the device
    struct Job {
        std::vector<uint64_t> ids;
        std::vector<QuerySpec> specs;
        std::vector<QueryOut> outs;
        std::vector<int> st;
        int passes_left = 0;
        bool unpinned = false;
    };
    std::shared_ptr<Job> slot_job[2];
    int next_slot = 0;

    void complete(Job& J, int i, int status);
    void release(Job& J);
    bool admit(std::vector<uint64_t>* ids);
    void finish_slot(int s);
    void run_job(std::shared_ptr<Job> J);
    void loop();
};

void rj_server::complete(Job& J, int i, int status)
{
    std::lock_guard<std::mutex> lk(m);
    auto it = reqs.find(J.ids[i]);
    if (it == reqs.end()) return;
    Req& r = it->second;
    r.status = status;
    if (r.out) {
        r.out->status = status;
        if (status == RJ_OK) rj_fill_result(J.outs[i], r.spec.P->ps.np, r.spec.fn, r.spec.agg_type, r.out);
    }
    r.t_done = now_ms();
    r.done = true;
    if (status == RJ_OK) {
        completed++;
        lat.push_back(r.t_done - r.t_submit);
        last_done = r.t_done;
    } else {
        failed++;
    }
    cv_done.notify_all();
}

void rj_server::release(Job& J)
{
    if (J.unpinned) return;
    for (auto& s : J.specs) s.P->pins--;
    J.unpinned = true;
}

// Admission: queries in arrival order while their estimated device bytes fit
// what the budget leaves (structures of collections no query in flight uses
// count as free: they can be evicted). The first query of an admission is
// always taken unless it cannot fit even an empty device.
bool rj_server::admit(std::vector<uint64_t>* ids)
{
    ids->clear();
    const bool limited = ctx->budget->limit > 0;
    // cached structures of collections no query in flight or in this
    // admission uses can be evicted: they count as free
    auto evictable = [](const rj_polyset* P) {
        return (int64_t)P->device_bytes() - (int64_t)(P->d_xy.bytes + P->d_off.bytes + P->d_prev.bytes);
    };
    int64_t avail = INT64_MAX;
    if (limited) {
        avail = ctx->budget->limit - ctx->budget->used;
        for (rj_polyset* P : ctx->polysets)
            if (P->pins == 0) avail += evictable(P);
    }
    std::vector<rj_polyset*> used;
    std::lock_guard<std::mutex> lk(m);
    int64_t need = 0;
    while (!queue.empty() && (int)ids->size() < max_batch) {
        const uint64_t id = queue.front();
        Req& r = reqs.at(id);
        const int64_t est = ctx->estimate_bytes(r.spec);
        if (limited && est > ctx->budget->limit) {
            // can never fit: fail it here
            queue.pop_front();
            r.status = RJ_ERR_NOMEM;
            if (r.out) r.out->status = RJ_ERR_NOMEM;
            r.done = true;
            r.t_done = now_ms();
            failed++;
            cv_done.notify_all();
            continue;
        }
        int64_t lose = 0;  // evictable bytes this query would pin
        rj_polyset* P = r.spec.P;
        if (limited && P->pins == 0 && std::find(used.begin(), used.end(), P) == used.end()) lose = evictable(P);
        if (limited && need + est > avail - lose && (!ids->empty() || slot_job[0] || slot_job[1])) break;
        if (std::find(used.begin(), used.end(), P) == used.end()) used.push_back(P);
        avail -= lose;
        need += est;
        ids->push_back(id);
        queue.pop_front();
    }
    return !ids->empty();
}

void rj_server::finish_slot(int s)
{
    std::shared_ptr<Job> J = slot_job[s];
    if (!J) return;
    slot_job[s].reset();
    const std::vector<int> members = ctx->slots[s].members;
    const int rc = ctx->batch_finish(s, t, J->specs.data(), J->outs.data());
    for (int i : members) complete(*J, i, rc ? rc : J->st[i]);
    if (--J->passes_left == 0) release(*J);
}

// Prepare a job and launch its passes; the last pass is left running.
void rj_server::run_job(std::shared_ptr<Job> J)
{
    const int n = (int)J->specs.size();
    J->outs.assign(n, QueryOut());
    std::vector<int> fused, alone;
    int rc = ctx->batch_prepare(t, J->specs.data(), n, &fused, &alone, &J->st);
    if (rc) {
        for (int i = 0; i < n; i++) complete(*J, i, rc);
        J->unpinned = true;  // batch_prepare released its pins on failure
        return;
    }
    for (int i = 0; i < n; i++)
        if (J->st[i] != RJ_OK) complete(*J, i, J->st[i]);
    // passes of at most max_batch queries (and border-queue entries)
    std::vector<std::vector<int>> passes;
    ctx->group_passes(t, J->specs.data(), fused, max_batch, &passes);
    J->passes_left = (int)passes.size() + 1;  // + 1 until the queries that run alone are done
    for (auto& p : passes) {
        const int s = next_slot;
        next_slot ^= 1;
        finish_slot(s);
        rc = ctx->batch_launch(s, t, J->specs.data(), p);
        if (rc) {
            for (int i : p) complete(*J, i, rc);
            J->passes_left--;
            continue;
        }
        slot_job[s] = J;
        {
            std::lock_guard<std::mutex> lk(m);
            batches++;
            batched += (int64_t)p.size();
        }
    }
    for (int i : alone) {
        if (J->st[i] != RJ_OK) continue;
        const int r2 = ctx->run_single(t, J->specs[i], &J->outs[i]);
        complete(*J, i, r2);
    }
    if (--J->passes_left == 0) release(*J);
}

void rj_server::loop()
{
    (void)hipSetDevice(ctx->device);
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(m);
            const bool inflight = slot_job[0] || slot_job[1];
            if (!inflight) cv_work.wait(lk, [&] { return stop || !queue.empty(); });
            if (stop && queue.empty() && !inflight) break;
        }
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        std::vector<uint64_t> ids;
        if (admit(&ids)) {
            auto J = std::make_shared<Job>();
            J->ids = ids;
            {
                std::lock_guard<std::mutex> lk2(m);
                for (uint64_t id : ids) J->specs.push_back(reqs.at(id).spec);
            }
            run_job(J);
        } else {
            // nothing admissible: let the oldest pass finish (frees memory,
            // completes queries), then look again
            const int s = next_slot;
            if (slot_job[s]) finish_slot(s);
            else finish_slot(s ^ 1);
        }
    }
}

extern "C" {

int rj_server_create(rj_ctx* ctx, rj_table* t, const rj_server_config* cfg, rj_server** out)
{
    if (!ctx || !t || !out) return RJ_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    rj_server* s = new rj_server();
    s->ctx = ctx;
    s->t = t;
    if (cfg) {
        if (cfg->max_batch < 0 || cfg->max_batch > kBatchMax || cfg->max_pending < 0 || cfg->mem_mb < 0) {
            delete s;
            return ctx->fail(RJ_ERR_ARG, "bad server configuration");
        }
        if (cfg->max_batch > 0) s->max_batch = cfg->max_batch;
        if (cfg->max_pending > 0) s->max_pending = (size_t)cfg->max_pending;
    }
    s->saved_limit = ctx->budget->limit;
    if (cfg && cfg->mem_mb > 0) ctx->budget->limit = cfg->mem_mb * 1024LL * 1024LL;
    s->th = std::thread([s] { s->loop(); });
    *out = s;
    return RJ_OK;
}

int rj_server_submit(rj_server* s, const rj_query* q, rj_result* out, uint64_t* ticket)
{
    if (!s || !q || !ticket) return RJ_ERR_ARG;
    const double t0 = now_ms();
    QuerySpec spec;
    std::string why;
    int rc = spec_check(s->ctx, s->t, q, &spec, &why);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(s->m);
    if (s->stop) return RJ_ERR_STATE;
    if (s->queue.size() >= s->max_pending) return RJ_ERR_LIMIT;
    const uint64_t id = s->next_id++;
    rj_server::Req& r = s->reqs[id];
    r.spec = std::move(spec);
    r.out = out;
    r.t_submit = t0;
    if (out) out->status = 1;  // pending
    s->queue.push_back(id);
    s->submitted++;
    if (s->first_submit < 0) s->first_submit = t0;
    *ticket = id;
    s->cv_work.notify_one();
    return RJ_OK;
}

int rj_server_poll(rj_server* s, uint64_t ticket, int32_t wait_ms)
{
    if (!s) return RJ_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return RJ_ERR_ARG;
    if (!it->second.done && wait_ms > 0)
        s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] {
            auto j = s->reqs.find(ticket);
            return j == s->reqs.end() || j->second.done;
        });
    it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return RJ_ERR_ARG;
    if (!it->second.done) return 0;
    const int st = it->second.status;
    s->reqs.erase(it);
    return st == RJ_OK ? 1 : st;
}

int rj_server_stats_get(rj_server* s, rj_server_stats* st)
{
    if (!s || !st) return RJ_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->m);
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->failed = s->failed;
    st->batches = s->batches;
    st->mean_batch = s->batches ? (double)s->batched / (double)s->batches : 0.0;
    std::vector<double> l = s->lat;
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) {
        if (l.empty()) return 0.0;
        size_t k = (size_t)std::ceil(p * (double)l.size()) - 1;
        return l[std::min(k, l.size() - 1)];
    };
    st->p50_ms = pct(0.50);
    st->p99_ms = pct(0.99);
    st->max_ms = l.empty() ? 0.0 : l.back();
    const double span = s->last_done - s->first_submit;
    st->qps = (span > 0 && !l.empty()) ? (double)l.size() / (span * 1e-3) : 0.0;
    return RJ_OK;
}

int rj_server_stats_reset(rj_server* s)
{
    if (!s) return RJ_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->m);
    s->lat.clear();
    s->first_submit = -1;
    s->last_done = 0;
    s->batches = s->batched = 0;
    return RJ_OK;
}

int rj_server_destroy(rj_server* s)
{
    if (!s) return RJ_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->th.joinable()) s->th.join();
    {
        std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
        s->ctx->budget->limit = s->saved_limit;
    }
    delete s;
    return RJ_OK;
}

}  // extern "C"
