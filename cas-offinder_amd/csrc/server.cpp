// Serving loop: continuous batching of queued queries over the resident genome.
//
// Requests (one or more queries each) wait in a FIFO.  Each worker owns a
// lane (its own HIP stream and buffers, hit buffers sized at start from the
// memory budget); an idle worker admits requests from the head of the queue
// while the merged batch stays within max_batch_queries, its estimated hits
// fit the lane's hit buffers, and the in-flight estimate of key-table and
// hit bytes of all running batches stays within the budget; it runs one
// batched search for all admitted queries and splits the records back per
// request.  Requests that arrive while a batch is on the GPU are admitted by
// the next free worker, so batches on different lanes overlap on the device.
// The hit estimate of a request is its compare-key count times the running
// average of hits per compare key of the batches served so far; a batch that
// finds more hits than its lane holds grows the lane's buffers and reruns
// (counted in cof_server_stats).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "engine.hpp"

using namespace cof;

namespace {

using clk = std::chrono::steady_clock;

struct Request {
    uint64_t ticket = 0;
    std::vector<std::shared_ptr<Query>> qs;
    clk::time_point t_submit;
    uint64_t keys = 0;         // compare keys of the request's queries
    uint64_t table_bytes = 0;  // estimated key / bulge table bytes
    cof_result *result = nullptr;
    double latency_ms = 0;
    int status = 1;  // 1 pending, 0 done, < 0 error
    std::string err;
};

// Device bytes per hit in a lane: emitted and sorted sort keys (8 + 8) and
// mismatch counts (2 + 2).
constexpr uint64_t kLaneBytesPerHit = 20;
constexpr uint64_t kMinLaneHits = 1ull << 20, kMaxLaneHits = 1ull << 26;

// Table bytes of a query: its key table, or for bulge queries a record and a
// variant block per guide (entries of its splits plus the family headers).
uint64_t table_bytes(const Query &q) {
    uint64_t b = q.key_words.size() * 4ull;
    for (const BulgeGuide &g : q.bulge) b += (g.ent.size() + 128) * 4ull;
    return b;
}

}  // namespace

struct cof_server {
    cof_ctx *ctx = nullptr;
    cof_server_cfg cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> pending;
    std::map<uint64_t, Request *> requests;
    uint64_t next_ticket = 1;
    uint64_t inflight_bytes = 0;
    uint64_t lane_hits = 0;          // hit capacity of each lane
    double hits_per_key = 1024.0;    // running average over served batches
    cof_server_stats st{};
    bool stopping = false;
    std::vector<std::unique_ptr<Lane>> lanes;
    std::vector<std::thread> workers;

    uint64_t est_hits(const Request &r) const { return (uint64_t)(r.keys * hits_per_key) + 1; }
    void worker(Lane *L);
};

void cof_server::worker(Lane *L) {
    (void)hipSetDevice(ctx->device);
    for (;;) {
        std::vector<Request *> batch;
        uint64_t bytes = 0, hits = 0, keys = 0;
        size_t nq = 0;
        {
            std::unique_lock<std::mutex> lk(mu);
            for (;;) {
                if (!pending.empty()) {
                    while (!pending.empty()) {
                        Request *r = pending.front();
                        const bool first = batch.empty();
                        if (!first && nq + r->qs.size() > cfg.max_batch_queries) break;
                        if (nq + r->qs.size() > 255) break;
                        const uint64_t h = est_hits(*r), rb = r->table_bytes + h * kLaneBytesPerHit;
                        // the first request of a batch is admitted whatever its size when
                        // nothing else is in flight (a larger one grows the lane)
                        const bool fits = hits + h <= lane_hits && inflight_bytes + bytes + rb <= cfg.mem_budget_bytes;
                        if (!fits && !(first && inflight_bytes == 0)) {
                            if (first) st.admission_waits++;  // the head waits for running batches
                            break;
                        }
                        pending.pop_front();
                        batch.push_back(r);
                        bytes += rb;
                        hits += h;
                        keys += r->keys;
                        nq += r->qs.size();
                    }
                    if (!batch.empty()) break;
                }
                if (stopping && pending.empty()) return;
                cv_work.wait(lk);
            }
            inflight_bytes += bytes;
            st.peak_inflight_bytes = std::max(st.peak_inflight_bytes, inflight_bytes);
        }
        std::vector<std::shared_ptr<Query>> qs;
        for (Request *r : batch) qs.insert(qs.end(), r->qs.begin(), r->qs.end());
        cof_result *res = nullptr;
        std::string err;
        const uint64_t cap_before = L->d_hkey.n;
        const auto ts0 = clk::now();
        const int rc = run_search(*L, ctx->dg, ctx->slab_positions, qs, COF_MODE_BATCHED, &ctx->tuner, &res, err);
        const auto ts1 = clk::now();
        const bool grew = L->d_hkey.n > cap_before;
        const uint64_t found = rc == COF_OK ? res->hits.size() : 0;
        // split the merged result per request, query indices rebased to 0 (a
        // batch of one request keeps the merged result); each request's
        // records are copied by host threads, query run by query run
        std::vector<cof_result *> parts(batch.size(), nullptr);
        if (rc == COF_OK && batch.size() == 1) {
            res->genome_serial = ctx->genome_serial;  // reloads are refused while the server runs
            parts[0] = res;
            res = nullptr;
        } else if (rc == COF_OK) {
            size_t q0 = 0;
            for (size_t i = 0; i < batch.size(); i++) {
                auto *p = new cof_result();
                p->genome_serial = ctx->genome_serial;
                const size_t nqi = batch[i]->qs.size();
                p->queries = batch[i]->qs;
                p->qbegin.assign(nqi, 0);
                p->qend.assign(nqi, 0);
                size_t at = 0;
                for (size_t j = 0; j < nqi; j++) {
                    p->qbegin[j] = at;
                    at += res->qend[q0 + j] - res->qbegin[q0 + j];
                    p->qend[j] = at;
                }
                p->hits.resize(at);
                cof_hit *dst = p->hits.data();
                const cof_hit *src = res->hits.data();
                const std::vector<size_t> &b = p->qbegin, &e = p->qend;
                const size_t *sb = res->qbegin.data() + q0;
                parallel_for(at, 1 << 16, [&](uint64_t x0, uint64_t x1) {
                    size_t j = (size_t)(std::upper_bound(e.begin(), e.end(), (size_t)x0) - e.begin());
                    for (uint64_t x = x0; x < x1;) {
                        const uint64_t end = std::min<uint64_t>(x1, e[j]);
                        for (; x < end; x++) {
                            dst[x] = src[sb[j] + (x - b[j])];
                            dst[x].query = (uint8_t)j;
                        }
                        j++;
                    }
                });
                parts[i] = p;
                q0 += nqi;
            }
        }
        delete res;
        const auto now = clk::now();
        const cof_stats bst = L->stats;
        {
            std::lock_guard<std::mutex> lk(mu);
            for (size_t i = 0; i < batch.size(); i++) {
                Request *r = batch[i];
                r->result = parts[i];
                r->status = rc == COF_OK ? 0 : rc;
                r->err = err;
                r->latency_ms = std::chrono::duration<double, std::milli>(now - r->t_submit).count();
            }
            inflight_bytes -= bytes;
            st.batches++;
            st.requests += batch.size();
            st.queries += nq;
            st.hits += found;
            st.lane_grows += grew ? 1 : 0;
            st.launches += bst.n_launches;
            st.t_search_ms += std::chrono::duration<double, std::milli>(ts1 - ts0).count();
            st.t_plan_ms += bst.t_plan_ms;
            st.t_kernel_ms += bst.t_fused_ms;
            st.t_split_ms += std::chrono::duration<double, std::milli>(now - ts1).count();
            if (rc == COF_OK && keys)  // running average of hits per key, weighted by keys
                hits_per_key = std::max(1.0, 0.7 * hits_per_key + 0.3 * ((double)found / (double)keys));
            st.hits_per_key = hits_per_key;
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }
}

extern "C" {

int cof_server_start(cof_ctx *ctx, const cof_server_cfg *cfg, cof_server **out) {
    if (!ctx || !out) return COF_E_INVALID;
    *out = nullptr;
    if (ctx->device < 0) return ctx->fail(COF_E_UNSUPPORTED, "Host-only context cannot serve.");
    if (!ctx->genome_ready) return ctx->fail(COF_E_INVALID, "No genome loaded.");
    auto *s = new cof_server();
    s->ctx = ctx;
    s->cfg.mem_budget_bytes = cfg && cfg->mem_budget_bytes ? cfg->mem_budget_bytes : (8ull << 30);
    s->cfg.max_batch_queries = cfg && cfg->max_batch_queries ? std::min<uint32_t>(cfg->max_batch_queries, 255) : 64;
    (void)hipSetDevice(ctx->device);
    const int nlanes = 2;
    // each lane's hit buffers: its share of the budget, in [1 M, 64 M] hits
    s->lane_hits = std::min(kMaxLaneHits, std::max(kMinLaneHits, s->cfg.mem_budget_bytes / nlanes / kLaneBytesPerHit));
    s->st.lane_hits = s->lane_hits;
    for (int i = 0; i < nlanes; i++) {
        auto L = std::make_unique<Lane>();
        L->graphs = false;  // merged batches differ from call to call: nothing to replay
        L->trials = false;  // the tuner file's (nearest) decisions, no trials
        const uint64_t n = s->lane_hits;
        if (L->init() != hipSuccess || L->d_hkey.reserve(n) != hipSuccess || L->d_hmm.reserve(n) != hipSuccess ||
            L->d_hkey2.reserve(n) != hipSuccess || L->d_hmm2.reserve(n) != hipSuccess ||
            L->h_hkey.reserve(n) != hipSuccess || L->h_hmm.reserve(n) != hipSuccess) {
            for (auto &x : s->lanes) x->release();
            L->release();
            delete s;
            return ctx->fail(COF_E_DEVICE, "cannot create serving streams");
        }
        s->lanes.push_back(std::move(L));
    }
    ctx->servers++;
    for (auto &L : s->lanes) s->workers.emplace_back([s, p = L.get()] { s->worker(p); });
    *out = s;
    return COF_OK;
}

int cof_server_submit(cof_server *s, cof_query *const *queries, int n_queries, uint64_t *ticket) {
    if (!s || !queries || n_queries <= 0 || n_queries > 255 || !ticket) return COF_E_INVALID;
    auto *r = new Request();
    for (int i = 0; i < n_queries; i++) {
        if (!queries[i]) { delete r; return COF_E_INVALID; }
        r->qs.push_back(queries[i]->q);
        r->keys += queries[i]->q->keys.size();
        r->table_bytes += table_bytes(*queries[i]->q);
    }
    r->t_submit = clk::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopping) { delete r; return COF_E_INVALID; }
        r->ticket = s->next_ticket++;
        s->requests[r->ticket] = r;
        s->pending.push_back(r);
        *ticket = r->ticket;
    }
    s->cv_work.notify_one();
    return COF_OK;
}

int cof_server_poll(cof_server *s, uint64_t ticket, int wait, cof_result **out, double *latency_ms) {
    if (!s || !out) return COF_E_INVALID;
    *out = nullptr;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->requests.find(ticket);
    if (it == s->requests.end()) return COF_E_INVALID;
    Request *r = it->second;
    if (wait) s->cv_done.wait(lk, [r] { return r->status != 1; });
    if (r->status == 1) return 1;
    const int st = r->status;
    if (st == 0) *out = r->result;
    if (latency_ms) *latency_ms = r->latency_ms;
    if (st < 0) s->ctx->err = r->err;
    s->requests.erase(it);
    delete r;
    return st;
}

int cof_server_stats_get(cof_server *s, cof_server_stats *out) {
    if (!s || !out) return COF_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    return COF_OK;
}

int cof_server_stop(cof_server *s) {
    if (!s) return COF_E_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    for (auto &L : s->lanes) L->release();
    for (auto &kv : s->requests) {
        delete kv.second->result;
        delete kv.second;
    }
    s->ctx->servers--;
    delete s;
    return COF_OK;
}

}  // extern "C"
