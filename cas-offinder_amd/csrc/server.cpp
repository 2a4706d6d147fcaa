// Serving loop: continuous batching of queued queries over the resident genome.
//
// Requests (one or more queries each) wait in a FIFO.  Each worker owns a
// lane (its own HIP stream and buffers); an idle worker admits requests from
// the head of the queue while the merged batch stays within max_batch_queries
// and the in-flight estimate of key-table and hit-buffer bytes stays within
// the memory budget, runs one batched search for all admitted queries, and
// splits the records back per request.  Requests that arrive while a batch is
// on the GPU are admitted by the next free worker, so batches on different
// lanes overlap on the device.
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
    uint64_t est_bytes = 0;
    cof_result *result = nullptr;
    double latency_ms = 0;
    int status = 1;  // 1 pending, 0 done, < 0 error
    std::string err;
};

// Admission estimate: the query's key table plus a hit allowance of 4096
// records per compare key, at 20 device bytes per hit (8-byte sort key and
// 2-byte mismatch count, emitted and sorted copies).
uint64_t estimate_bytes(const Query &q) {
    return q.key_words.size() * 4ull + q.keys.size() * 4096ull * 20ull;
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
    bool stopping = false;
    std::vector<std::unique_ptr<Lane>> lanes;
    std::vector<std::thread> workers;

    void worker(Lane *L);
};

void cof_server::worker(Lane *L) {
    (void)hipSetDevice(ctx->device);
    for (;;) {
        std::vector<Request *> batch;
        uint64_t bytes = 0;
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
                        // a request that alone exceeds the budget runs when nothing else is in flight
                        const bool fits = inflight_bytes + bytes + r->est_bytes <= cfg.mem_budget_bytes;
                        if (!fits && !(first && inflight_bytes == 0)) break;
                        pending.pop_front();
                        batch.push_back(r);
                        bytes += r->est_bytes;
                        nq += r->qs.size();
                    }
                    if (!batch.empty()) break;
                }
                if (stopping && pending.empty()) return;
                cv_work.wait(lk);
            }
            inflight_bytes += bytes;
        }
        std::vector<std::shared_ptr<Query>> qs;
        for (Request *r : batch) qs.insert(qs.end(), r->qs.begin(), r->qs.end());
        cof_result *res = nullptr;
        std::string err;
        const int rc = run_search(*L, ctx->dg, ctx->slab_positions, qs, COF_MODE_BATCHED, &ctx->tuner, &res, err);
        // split the merged result per request, query indices rebased to 0
        std::vector<cof_result *> parts(batch.size(), nullptr);
        if (rc == COF_OK) {
            size_t q0 = 0;
            for (size_t i = 0; i < batch.size(); i++) {
                auto *p = new cof_result();
                p->genome_serial = ctx->genome_serial;  // reloads are refused while the server runs
                const size_t nqi = batch[i]->qs.size();
                p->queries = batch[i]->qs;
                p->qbegin.assign(nqi + 1, 0);
                const size_t h0 = res->qbegin[q0];
                p->hits.resize(res->qbegin[q0 + nqi] - h0);
                for (size_t j = 0; j < nqi; j++) {
                    p->qbegin[j] = res->qbegin[q0 + j] - h0;
                    for (size_t h = res->qbegin[q0 + j]; h < res->qbegin[q0 + j + 1]; h++) {
                        p->hits[h - h0] = res->hits[h];
                        p->hits[h - h0].query = (uint8_t)j;
                    }
                }
                p->qbegin[nqi] = p->hits.size();
                parts[i] = p;
                q0 += nqi;
            }
        }
        delete res;
        const auto now = clk::now();
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
    for (int i = 0; i < nlanes; i++) {
        auto L = std::make_unique<Lane>();
        L->graphs = false;  // merged batches differ from call to call: nothing to replay
        L->trials = false;  // tuned variants of the context are used, untuned ones default
        // hit buffers sized up front: a first large request then does not pay
        // for an overflow rerun
        constexpr uint64_t kLaneHits = 16ull << 20;
        if (L->init() != hipSuccess || L->d_hkey.reserve(kLaneHits) != hipSuccess ||
            L->d_hmm.reserve(kLaneHits) != hipSuccess || L->d_hkey2.reserve(kLaneHits) != hipSuccess ||
            L->d_hmm2.reserve(kLaneHits) != hipSuccess || L->h_hkey.reserve(kLaneHits) != hipSuccess ||
            L->h_hmm.reserve(kLaneHits) != hipSuccess) {
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
        r->est_bytes += estimate_bytes(*queries[i]->q);
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
