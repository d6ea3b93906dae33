// Serving loop: SpGEMM (C = A*B) and AMG Galerkin (C = P^T (A P)) requests
// of mixed sizes, continuously batched onto one GPU.
//
// Operands are registered once and stay resident on the device (a Galerkin
// operator also keeps its on-device transpose P^T).  A request is one or two
// product steps (Galerkin: AP, then P^T (AP) with AP kept resident between
// the steps).  Worker threads, each with its own context (HIP stream), take
// the waiting steps in arrival order and admit a step when its memory bound
// from the symbolic phase (tr_symbolic) fits the remaining budget; the
// admitted steps run as one batch (tr_batch_create, one task per step), so
// steps that arrived while other batches were on the GPU start as soon as a
// worker and memory are free, and batches of different workers overlap on
// the device.  The reference has no serving layer; this is the request loop
// the single-GPU path needs around its multiply.
#include "tr_internal.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <array>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

struct Operand {
    tr_csx_t m{};        // borrowed device CSR
    tr_csx_t mt{};       // transpose (owned by the server's admin context), built on first Galerkin use
    bool has_t = false;
};

struct Batch;

struct Request {
    int64_t id = 0;
    int32_t kind = TR_REQ_SPGEMM;
    bool discard = false;        // drop the result at completion
    int32_t a = 0, b = 0;        // operand ids
    int step = 0;                // Galerkin: 0 = AP, 1 = P^T AP
    Clock::time_point arrival;
    Clock::time_point done_at;
    bool done = false;
    int status = TR_OK;
    std::string error;
    size_t bound = 0;            // admitted bytes of the current step
    int oom_retries = 0;         // times the step was put back after a failed allocation
    std::shared_ptr<Batch> partial;  // AP of a Galerkin request
    int32_t partial_task = 0;
    std::shared_ptr<Batch> result;   // final product
    int32_t result_task = 0;
};

// A finished batch; stays alive while a result or partial refers to it.
// While it runs it counts its admitted bound against the budget; once it has
// finished, only the device bytes it keeps (its results) stay counted, and
// they are returned to the budget when it is destroyed.
// Each worker's pool reserves the worker's share of the budget once and
// places every batch's blocks in it (tr::Arena): a hipMalloc larger than
// what the runtime has cached from earlier frees takes about 14 ms per GB
// with the GPU idle, and batch sizes change all the time.  Blocks that do not
// fit the reservation come from the pool's cache, whose size (in use +
// cached) stays within the share.  An allocation that still fails trims
// every worker's pool and retries (worker_loop).
struct Batch {
    tr_batch_t b = nullptr;
    tr_ctx_t ctx = nullptr;
    std::mutex *ctx_mu = nullptr;
    std::atomic<size_t> *in_use = nullptr;
    size_t bound = 0;
    size_t cache_cap = 0;
    ~Batch() {
        if (b) {
            std::lock_guard<std::mutex> lk(*ctx_mu);
            tr_batch_destroy(b);
            ctx->pool.trim_to(cache_cap);
        }
        if (in_use) in_use->fetch_sub(bound);
    }
};

struct Worker {
    tr_ctx_t ctx = nullptr;
    std::mutex mu;  // serialises the worker's context between its thread and releases
    std::thread th;
};

}  // namespace

struct tr_server {
    int device = 0;
    size_t budget = 0;
    int32_t max_batch = 16;
    std::vector<std::unique_ptr<Worker>> workers;
    tr_ctx_t admin = nullptr;  // registration, transposes
    std::mutex admin_mu;

    std::mutex mu;
    std::condition_variable cv, done_cv;
    std::deque<Request *> queue;  // steps waiting for admission
    std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
    std::vector<Operand> ops;
    int64_t next_id = 1;
    std::atomic<size_t> in_use{0};
    int running = 0;  // batches admitted and not yet finished (s->mu)
    bool stopping = false;
    // symbolic memory bound per (kind, operand a, operand b, step): operands
    // are resident and products deterministic, so a bound is computed once
    std::map<std::array<int64_t, 4>, size_t> bounds;
    int64_t completed = 0, batches = 0, flops_done = 0, failed = 0;
    std::vector<double> lat_ms;
    Clock::time_point t0 = Clock::now();
    Clock::time_point last_done = Clock::now();
};

namespace {

void finish_step(tr_server *s, Request *r, const std::shared_ptr<Batch> &bt, int32_t task, int status,
                 const std::string &err) {
    // called with s->mu held
    r->bound = 0;
    if (status != TR_OK) {
        r->status = status;
        r->error = err;
        r->partial.reset();
        r->done = true;
        r->done_at = Clock::now();
        ++s->failed;
        s->done_cv.notify_all();
        return;
    }
    if (r->kind == TR_REQ_GALERKIN && r->step == 0) {
        r->partial = bt;
        r->partial_task = task;
        r->step = 1;
        s->queue.push_front(r);  // second step goes ahead of newer arrivals
        s->cv.notify_all();
        return;
    }
    if (!r->discard) {
        r->result = bt;
        r->result_task = task;
    }
    r->partial.reset();
    r->done = true;
    r->done_at = Clock::now();
    ++s->completed;
    s->lat_ms.push_back(std::chrono::duration<double, std::milli>(r->done_at - r->arrival).count());
    s->last_done = r->done_at;
    s->done_cv.notify_all();
}

// operands of a request's current step (s->mu held)
void step_operands(tr_server *s, Request *r, tr_csx_t &A, tr_csx_t &B) {
    if (r->kind == TR_REQ_SPGEMM) {
        A = s->ops[r->a].m;
        B = s->ops[r->b].m;
    } else if (r->step == 0) {
        A = s->ops[r->a].m;  // A
        B = s->ops[r->b].m;  // P
    } else {
        A = s->ops[r->b].mt;  // P^T
        tr_batch_result(r->partial->b, r->partial_task, &B);
    }
}

void worker_loop(tr_server *s, Worker *w) {
    for (;;) {
        std::vector<Request *> take;
        std::vector<tr_csx_t> As, Bs;
        size_t admitted = 0;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->stopping || !s->queue.empty(); });
            if (s->stopping) {
                // requests still waiting will not run: fail them so waiters return
                while (!s->queue.empty()) {
                    Request *q = s->queue.front();
                    s->queue.pop_front();
                    finish_step(s, q, nullptr, 0, TR_ERR_INVALID, "server stopped before the request ran");
                }
                return;
            }
            // admission in arrival order while the symbolic bound fits; a step
            // whose bound is not known yet is taken out, its bound computed
            // outside the server lock, and put back at the front
            while (!s->queue.empty() && (int32_t)take.size() < s->max_batch) {
                Request *r = s->queue.front();
                const std::array<int64_t, 4> key{r->kind, r->a, r->b, r->step};
                auto it = s->bounds.find(key);
                if (it == s->bounds.end()) {
                    if (!take.empty()) break;  // run what is admitted; bound it next time
                    tr_csx_t A, B;
                    step_operands(s, r, A, B);
                    s->queue.pop_front();
                    lk.unlock();
                    int64_t nnz = 0;
                    size_t bytes = 0;
                    int rc;
                    std::string err;
                    {
                        std::lock_guard<std::mutex> ck(w->mu);
                        rc = tr_symbolic(w->ctx, &A, &B, &nnz, &bytes);
                        if (rc != TR_OK) err = tr_last_error(w->ctx);
                    }
                    lk.lock();
                    if (rc != TR_OK) {
                        finish_step(s, r, nullptr, 0, rc, err);
                        continue;
                    }
                    s->bounds[key] = bytes;
                    s->queue.push_front(r);
                    continue;
                }
                const size_t bytes = it->second;
                if (bytes > s->budget) {
                    s->queue.pop_front();
                    finish_step(s, r, nullptr, 0, TR_ERR_OOM, "request exceeds the serving memory budget");
                    continue;
                }
                if (s->in_use.load() + admitted + bytes > s->budget) {
                    // A Galerkin continuation is the only way to free its own
                    // AP: with no batch running on any worker, nothing else can
                    // return memory, so it runs now (an allocation that then
                    // fails ends it with TR_ERR_OOM instead of waiting forever).
                    // Anything else waits for memory to free up.
                    const bool continuation = r->kind == TR_REQ_GALERKIN && r->step == 1;
                    if (!(continuation && take.empty() && s->running == 0)) break;
                }
                tr_csx_t A, B;
                step_operands(s, r, A, B);
                s->queue.pop_front();
                r->bound = bytes;
                admitted += bytes;
                take.push_back(r);
                As.push_back(A);
                Bs.push_back(B);
            }
            if (take.empty()) {
                // nothing admissible now: wait for a completion or new work
                s->done_cv.wait_for(lk, std::chrono::milliseconds(2));
                continue;
            }
            s->in_use.fetch_add(admitted);
            ++s->running;
        }
        // run the batch (flops of first steps counted for the request)
        auto bt = std::make_shared<Batch>();
        bt->ctx = w->ctx;
        bt->ctx_mu = &w->mu;
        bt->in_use = &s->in_use;
        bt->bound = admitted;
        bt->cache_cap = s->budget / s->workers.size();
        int rc;
        std::string err;
        int64_t fl = 0;
#ifdef TR_SERVER_TRACE
        const auto tb0 = Clock::now();
        const size_t nm0 = w->ctx->pool.n_malloc, nf0 = w->ctx->pool.n_free;
#endif
#ifdef TR_SERVER_TRACE
        double t_lock = 0, t_create = 0, t_sync = 0;
        size_t mem_free = 0, mem_total = 0;
        (void)hipMemGetInfo(&mem_free, &mem_total);
#endif
        for (int attempt = 0; attempt < 2; ++attempt) {
            {
#ifdef TR_SERVER_TRACE
                const auto ta = Clock::now();
#endif
                std::lock_guard<std::mutex> ck(w->mu);
#ifdef TR_SERVER_TRACE
                const auto tb = Clock::now();
#endif
                rc = tr::batch_create(w->ctx, (int32_t)take.size(), As.data(), Bs.data(), nullptr,
                                      TR_BATCH_NO_GRAPH | tr::kBatchSingleRun, false, &bt->b);
#ifdef TR_SERVER_TRACE
                const auto tc = Clock::now();
#endif
                if (rc == TR_OK) rc = tr_ctx_synchronize(w->ctx);
#ifdef TR_SERVER_TRACE
                const auto td = Clock::now();
                t_lock += std::chrono::duration<double, std::milli>(tb - ta).count();
                t_create += std::chrono::duration<double, std::milli>(tc - tb).count();
                t_sync += std::chrono::duration<double, std::milli>(td - tc).count();
#endif
                if (rc == TR_OK) tr_batch_stats(bt->b, &fl, nullptr, nullptr);
                if (rc != TR_OK) err = tr_last_error(w->ctx);
            }
            if (rc != TR_ERR_OOM || attempt) break;
            // device memory held in the workers' pool caches: return it and retry once
            for (auto &ow : s->workers) {
                std::lock_guard<std::mutex> ck(ow->mu);
                (void)tr_pool_trim(ow->ctx);
            }
        }
#ifdef TR_SERVER_TRACE
        fprintf(stderr,
                "[srv] worker %p tasks %zu bound %.1f MB flops %.3e batch %.2f ms (lock %.2f create %.2f sync %.2f) "
                "mallocs %zu frees %zu rc %d free %.1f GB pool %.1f+%.1f GB\n",
                (void *)w, take.size(), admitted / 1e6, (double)fl,
                std::chrono::duration<double, std::milli>(Clock::now() - tb0).count(), t_lock, t_create, t_sync,
                w->ctx->pool.n_malloc - nm0, w->ctx->pool.n_free - nf0, rc, mem_free / 1e9,
                w->ctx->pool.in_use() / 1e9, w->ctx->pool.cached() / 1e9);
#endif
        if (rc == TR_OK) {
            // from now on the batch holds only its results
            size_t keep;
            {
                std::lock_guard<std::mutex> ck(w->mu);
                keep = tr::batch_retained_bytes(bt->b);
            }
            if (keep < bt->bound) {
                s->in_use.fetch_sub(bt->bound - keep);
                bt->bound = keep;
            }
        }
        std::lock_guard<std::mutex> lk(s->mu);
        --s->running;
        ++s->batches;
        s->flops_done += fl;
        if (rc == TR_ERR_OOM && !s->stopping) {
            // put the steps back (in their order, ahead of newer arrivals) to be
            // admitted again, in smaller batches as memory allows; a step that
            // already failed to allocate once fails for good
            for (size_t t = take.size(); t-- > 0;) {
                Request *r = take[t];
                if (r->oom_retries++ == 0 && take.size() > 1) {
                    r->bound = 0;
                    s->queue.push_front(r);
                } else {
                    finish_step(s, r, nullptr, 0, rc, err);
                }
            }
            s->cv.notify_all();
            s->done_cv.notify_all();
            continue;
        }
        for (size_t t = 0; t < take.size(); ++t)
            finish_step(s, take[t], rc == TR_OK ? bt : nullptr, (int32_t)t, rc, err);
        s->done_cv.notify_all();
    }
}

}  // namespace

extern "C" {

int tr_server_create(int device, size_t mem_budget, int32_t workers, int32_t max_batch, tr_server_t *out) {
    if (!out || workers < 1 || max_batch < 1 || mem_budget == 0) return TR_ERR_INVALID;
    *out = nullptr;
    auto *s = new tr_server();
    s->device = device;
    s->budget = mem_budget;
    s->max_batch = max_batch;
    int rc = tr_ctx_create(&s->admin, device, nullptr);
    if (rc != TR_OK) {
        delete s;
        return rc;
    }
    for (int i = 0; i < workers; ++i) {
        auto w = std::make_unique<Worker>();
        rc = tr_ctx_create(&w->ctx, device, nullptr);
        if (rc != TR_OK) {
            for (auto &x : s->workers) tr_ctx_destroy(x->ctx);
            tr_ctx_destroy(s->admin);
            delete s;
            return rc;
        }
        // the worker's share of the budget, reserved once; blocks beyond it
        // (if the reservation fails or is fragmented) come from the cache
        (void)w->ctx->pool.reserve(mem_budget / (size_t)workers);
        w->ctx->pool.set_cache_cap(mem_budget / (size_t)workers);
        s->workers.push_back(std::move(w));
    }
    for (auto &w : s->workers) w->th = std::thread(worker_loop, s, w.get());
    *out = s;
    return TR_OK;
}

int tr_server_register(tr_server_t s, const tr_csx_t *M, int32_t *op_id) {
    if (!s || !M || !op_id || M->majordim != TR_ROWS || !M->ptr_vec) return TR_ERR_INVALID;
    {
        // validated once here: the server's batches do not check operands again
        std::lock_guard<std::mutex> ak(s->admin_mu);
        int32_t st = 0;
        int rc = tr_csx_validate(s->admin, M, &st);
        if (rc != TR_OK) return rc;
        if (st != 0) return TR_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    Operand o;
    o.m = *M;
    s->ops.push_back(o);
    *op_id = (int32_t)s->ops.size() - 1;
    return TR_OK;
}

int tr_server_submit(tr_server_t s, int32_t kind_flags, int32_t op_a, int32_t op_b, int64_t *req_id) {
    if (!s || !req_id) return TR_ERR_INVALID;
    const int32_t kind = kind_flags & 0xff;
    if (kind != TR_REQ_SPGEMM && kind != TR_REQ_GALERKIN) return TR_ERR_INVALID;
    if (kind_flags & ~(0xff | TR_REQ_DISCARD)) return TR_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (op_a < 0 || op_b < 0 || op_a >= (int32_t)s->ops.size() || op_b >= (int32_t)s->ops.size())
            return TR_ERR_INVALID;
    }
    if (kind == TR_REQ_GALERKIN) {
        // the operator's transpose stays resident: built once, on first use
        std::lock_guard<std::mutex> ak(s->admin_mu);
        tr_csx_t P;
        bool need;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            need = !s->ops[op_b].has_t;
            P = s->ops[op_b].m;
        }
        if (need) {
            tr_csx_t Pt{};
            int rc = tr_csx_transpose(s->admin, &P, &Pt);
            if (rc == TR_OK) rc = tr_ctx_synchronize(s->admin);
            if (rc != TR_OK) return rc;
            std::lock_guard<std::mutex> lk(s->mu);
            s->ops[op_b].mt = Pt;
            s->ops[op_b].has_t = true;
        }
    }
    auto r = std::make_unique<Request>();
    r->kind = kind;
    r->discard = (kind_flags & TR_REQ_DISCARD) != 0;
    r->a = op_a;
    r->b = op_b;
    r->arrival = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    r->id = s->next_id++;
    *req_id = r->id;
    s->queue.push_back(r.get());
    s->reqs.emplace(r->id, std::move(r));
    s->cv.notify_one();
    return TR_OK;
}

int tr_server_wait(tr_server_t s, int64_t req_id, tr_csx_t *C, double *latency_ms) {
    if (!s) return TR_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(req_id);
    if (it == s->reqs.end()) return TR_ERR_INVALID;
    Request *r = it->second.get();
    s->done_cv.wait(lk, [&] { return r->done; });
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->done_at - r->arrival).count();
    if (r->status != TR_OK) return r->status;
    if (C) {
        if (!r->result) return TR_ERR_INVALID;  // submitted with TR_REQ_DISCARD
        return tr_batch_result(r->result->b, r->result_task, C);
    }
    return TR_OK;
}

// Drop a finished request and its device result.
int tr_server_release(tr_server_t s, int64_t req_id) {
    if (!s) return TR_ERR_INVALID;
    std::unique_ptr<Request> r;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(req_id);
        if (it == s->reqs.end() || !it->second->done) return TR_ERR_INVALID;
        r = std::move(it->second);
        s->reqs.erase(it);
    }
    r.reset();  // may destroy the batch (takes the worker's context lock)
    return TR_OK;
}

int tr_server_stats(tr_server_t s, int64_t *completed, int64_t *batches, double *p50_ms, double *p99_ms,
                    double *gflops) {
    if (!s) return TR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    std::vector<double> v = s->lat_ms;
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) -> double {
        if (v.empty()) return 0.0;
        size_t i = (size_t)std::min<double>((double)v.size() - 1, q * (double)(v.size() - 1) + 0.5);
        return v[i];
    };
    if (completed) *completed = s->completed;
    if (batches) *batches = s->batches;
    if (p50_ms) *p50_ms = pct(0.50);
    if (p99_ms) *p99_ms = pct(0.99);
    if (gflops) {
        const double secs = std::chrono::duration<double>(s->last_done - s->t0).count();
        *gflops = secs > 0 ? 2.0 * (double)s->flops_done / secs / 1e9 : 0.0;
    }
    return TR_OK;
}

// Restart the throughput clock and latency record (e.g. after warm-up).
int tr_server_reset_stats(tr_server_t s) {
    if (!s) return TR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    s->lat_ms.clear();
    s->completed = s->batches = s->flops_done = s->failed = 0;
    s->t0 = s->last_done = Clock::now();
    return TR_OK;
}

int tr_server_destroy(tr_server_t s) {
    if (!s) return TR_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
        s->cv.notify_all();
    }
    for (auto &w : s->workers)
        if (w->th.joinable()) w->th.join();
    s->reqs.clear();  // drops results (batches) while the worker contexts live
    for (auto &o : s->ops)
        if (o.has_t) tr_csx_free(s->admin, &o.mt);
    for (auto &w : s->workers) tr_ctx_destroy(w->ctx);
    tr_ctx_destroy(s->admin);
    delete s;
    return TR_OK;
}

}  // extern "C"
