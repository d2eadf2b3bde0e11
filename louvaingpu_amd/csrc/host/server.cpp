// server.cpp - continuous-batching request loop over the graph-set engine
// (include/louvain_serve.h).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "louvain_serve.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t id = 0;
    int32_t n = 0;
    int64_t e = 0;
    const int32_t *src = nullptr, *dst = nullptr;  // caller's arrays, valid until lvs_wait
    const void *w = nullptr;                       // raw 32-bit weights (int32 or float32)
    int wtype = LV_WEIGHT_INT32;
    double gamma = 1.0;
    uint64_t footprint = 0;
    bool done = false;
    bool claimed = false;  // an lvs_wait is waiting for it (a second one is refused)
    lvs_result res{};
    std::vector<int32_t> comm;
};

// HBM a request's graph needs in a set: the engine's arena sizing per edge and
// vertex (200 B/edge, 320 B/vertex).
uint64_t footprint(int32_t n, int64_t e) { return 200ull * (uint64_t)e + 320ull * (uint64_t)n; }

}  // namespace

struct lvs_server {
    lv_config cfg{};
    uint64_t budget = 0;
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queue;
    std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
    std::vector<std::thread> threads;
    std::vector<lv_ctx *> ctxs;
    int64_t next_id = 1;
    bool stop = false;
    lvs_stats stats{};

    double now() const
    {
        return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    }

    // Takes the waiting requests that fit the budget, in arrival order, that
    // share the first one's weight type and resolution (a graph set has one of
    // each).  Called with mu held and a non-empty queue; lvs_submit refuses a
    // request whose footprint exceeds the budget, so the head always fits.
    std::vector<Request *> take_set()
    {
        std::vector<Request *> set;
        uint64_t used = 0;
        const Request *head = queue.front();
        for (auto it = queue.begin(); it != queue.end();) {
            Request *r = *it;
            const bool same = r->wtype == head->wtype && r->gamma == head->gamma;
            if (same && used + r->footprint <= budget) {
                set.push_back(r);
                used += r->footprint;
                it = queue.erase(it);
            } else {
                ++it;
            }
            if (used >= budget) break;
        }
        return set;
    }

    // Runs the requests as one graph set and fills their results.  A failed set
    // is re-run one request at a time, so an invalid or oversized request gets
    // its own status and never fails the others (components are independent:
    // a request's result does not depend on the set it runs in).
    void run_set(lv_ctx *ctx, std::vector<Request *> &set)
    {
        const double t_start = now();
        const int rc = run_members(ctx, set.data(), (int32_t)set.size());
        if (rc != LV_OK && set.size() > 1)
            for (Request *r : set) (void)run_members(ctx, &r, 1);
        const int32_t ng = (int32_t)set.size();
        for (Request *r : set) {
            r->res.batch_size = ng;
            r->res.t_start_ms = t_start;
        }
        std::lock_guard<std::mutex> lk(mu);
        const double t_done = now();
        for (Request *r : set) {
            r->res.t_done_ms = t_done;
            r->done = true;
        }
        stats.completed += ng;
        stats.sets += 1;
        stats.max_set = std::max(stats.max_set, ng);
        cv_done.notify_all();
    }

    // One graph-set run of set[0..ng); every member's status is the set's, or
    // its own result query's.  Returns the set's status.
    int run_members(lv_ctx *ctx, Request *const *set, int32_t ng)
    {
        std::vector<int32_t> ns((size_t)ng);
        std::vector<int64_t> es((size_t)ng);
        std::vector<const int32_t *> ps((size_t)ng), pd((size_t)ng);
        std::vector<const void *> pw((size_t)ng);
        for (int32_t g = 0; g < ng; ++g) {
            ns[g] = set[g]->n;
            es[g] = set[g]->e;
            ps[g] = set[g]->src;
            pd[g] = set[g]->dst;
            pw[g] = set[g]->w;
        }
        const double gamma = set[0]->gamma;
        int rc = lv_batch_load_graphs(ctx, ng, ns.data(), es.data(), ps.data(), pd.data(),
                                      pw.data(), set[0]->wtype, 1, &gamma);
        if (rc == LV_OK) rc = lv_batch_run(ctx);
        for (int32_t g = 0; g < ng; ++g) {
            Request *r = set[g];
            r->res.status = rc;
            if (rc != LV_OK) continue;
            r->comm.resize((size_t)r->n);
            int rr = r->n ? lv_batch_final_assignment(ctx, g, r->comm.data()) : LV_OK;
            int32_t levels = 0;
            if (rr == LV_OK) rr = lv_batch_num_levels(ctx, g, &levels);
            lv_level_info info{};
            if (rr == LV_OK && levels > 0) rr = lv_batch_level_info(ctx, g, levels - 1, &info);
            r->res.status = rr;
            r->res.levels = levels;
            r->res.q = info.q;
            int32_t mx = -1;
            for (int32_t v : r->comm) mx = std::max(mx, v);
            r->res.communities = mx + 1;
        }
        return rc;
    }

    void worker(lv_ctx *ctx)
    {
        for (;;) {
            std::vector<Request *> set;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || !queue.empty(); });
                if (queue.empty()) return;  // stopping, nothing left
                set = take_set();
            }
            run_set(ctx, set);
        }
    }
};

extern "C" {

int lvs_create(lvs_server **out, const lv_config *cfg, int32_t workers, uint64_t budget_bytes)
{
    if (!out || workers < 1 || workers > 16) return LV_ERR_ARG;
    *out = nullptr;
    lvs_server *s = new (std::nothrow) lvs_server();
    if (!s) return LV_ERR_NOMEM;
    if (cfg)
        s->cfg = *cfg;
    else
        lv_config_default(&s->cfg);
    s->budget = budget_bytes ? budget_bytes : (32ull << 30);
    s->t0 = Clock::now();
    lv_config c = s->cfg;
    c.arena_bytes = s->budget + (1ull << 30);
    for (int32_t i = 0; i < workers; ++i) {
        lv_ctx *ctx = nullptr;
        int rc = lv_create(&ctx, &c);
        if (rc != LV_OK) {
            for (lv_ctx *x : s->ctxs) lv_destroy(x);
            delete s;
            return rc;
        }
        s->ctxs.push_back(ctx);
    }
    s->stats.workers = workers;
    for (lv_ctx *ctx : s->ctxs) s->threads.emplace_back([s, ctx] { s->worker(ctx); });
    *out = s;
    return LV_OK;
}

int lvs_submit(lvs_server *s, int32_t n, int64_t e, const int32_t *src, const int32_t *dst,
               const void *w, int wtype, double gamma, int64_t *id)
{
    if (!s || !id || n < 0 || e < 0 || (e > 0 && (!src || !dst || !w)) || !(gamma >= 0.0) ||
        (wtype != LV_WEIGHT_INT32 && wtype != LV_WEIGHT_FLOAT32))
        return LV_ERR_ARG;
    std::unique_ptr<Request> r(new (std::nothrow) Request());
    if (!r) return LV_ERR_NOMEM;
    r->n = n;
    r->e = e;
    r->wtype = wtype;
    r->gamma = gamma;
    r->footprint = footprint(n, e);
    if (r->footprint > s->budget) return LV_ERR_NOMEM;  // could never be admitted
    r->src = src;
    r->dst = dst;
    r->w = w;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return LV_ERR_ARG;
    r->id = s->next_id++;
    r->res.t_submit_ms = s->now();
    *id = r->id;
    s->queue.push_back(r.get());
    s->reqs.emplace(r->id, std::move(r));
    s->stats.submitted += 1;
    s->cv_work.notify_one();
    return LV_OK;
}

int lvs_wait(lvs_server *s, int64_t id, int32_t *comm, lvs_result *res)
{
    if (!s) return LV_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return LV_ERR_ARG;
    Request *r = it->second.get();
    if (r->claimed) return LV_ERR_ARG;  // another lvs_wait owns this request
    r->claimed = true;
    s->cv_done.wait(lk, [&] { return r->done; });
    if (comm && r->res.status == LV_OK && r->n)
        std::memcpy(comm, r->comm.data(), sizeof(int32_t) * (size_t)r->n);
    if (res) *res = r->res;
    const int st = r->res.status;
    s->reqs.erase(it);
    return st;
}

int lvs_get_stats(lvs_server *s, lvs_stats *out)
{
    if (!s || !out) return LV_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    return LV_OK;
}

double lvs_now_ms(lvs_server *s) { return s ? s->now() : 0.0; }

int lvs_destroy(lvs_server *s)
{
    if (!s) return LV_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->threads) t.join();
    for (lv_ctx *ctx : s->ctxs) lv_destroy(ctx);
    delete s;
    return LV_OK;
}

}  // extern "C"
