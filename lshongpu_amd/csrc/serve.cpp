// serve.cpp -- a request loop over a resident index (no counterpart in the
// reference, whose every call builds its tables and resets the device,
// cudaHelpers.cu:12-14).
//
// Continuous batching: callers submit requests (a block of host queries and
// its k) from any thread; a launch thread takes every queued request that
// fits the next batch -- at most max_batch_queries queries and a device
// memory budget -- packs them into one search at the largest rounded k among
// them, and each request receives its rows (the first calculateK(k) results
// of each: the top k by (distance, ID) does not depend on the other queries
// of the batch, nor on a larger k).  Four batch slots take turns, each with
// its own pinned and device buffers.  Where the index has a captured form
// for the batch (the SimHash index with pivot cells, rounded k <= 128, at
// most 4096 queries), the batch is enqueued on one of the index's batch
// slots with no host wait, so up to four batches are on the device
// together; otherwise one search runs at a time, and the next batch's query
// copy runs on a copy stream beside it.  A delivery thread takes the
// batches in launch order: it waits for each, redoes by an ordinary search
// the queries a captured batch's select flagged (or the whole batch, when
// its cells' second pass was too dense), copies the rows out to the
// requests and frees the slot, while the launch thread packs and launches
// the next.  A request's latency is from its submission to the delivery of
// its results.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.hpp"

namespace {

using lshgpu::guarded;

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
    int64_t id = 0;
    const float* q = nullptr;
    int n = 0;
    int k = 1;
    lsh_point_t* out = nullptr;
    double t_submit = 0, t_done = 0;
    int status = LSH_OK;
    std::string error;
    bool done = false;
};

constexpr int kSlots = 4;

struct Slot {
    float* hq = nullptr;          // pinned packed queries
    lsh_point_t* hres = nullptr;  // pinned results
    uint32_t* hflags = nullptr;   // pinned: a captured batch's query flags and status words
    lshgpu::DevBuf<float> dq;
    lshgpu::DevBuf<lsh_point_t> dres;
    hipEvent_t ev_in = nullptr, ev_done = nullptr, ev_out = nullptr;
    std::vector<Request*> reqs;
    int kq = 0;
    int total = 0;
    bool busy = false;   // admitted, until delivered (guarded by the server's mutex)
    bool graph = false;  // enqueued through the index's captured form
    int status = LSH_OK;
    std::string error;
};

}  // namespace

struct lsh_server {
    lsh_index_t* ix = nullptr;
    int max_batch = 4096;
    int64_t budget = 0;        // device bytes the in-flight batches may use
    int64_t per_query = 0;     // device bytes one query of a batch may use
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_slot, cv_inflight;
    std::deque<Request*> queue;
    std::deque<int> inflight;  // launched slots, oldest first
    bool launcher_done = false;
    std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
    int64_t next_id = 1;
    bool stop = false;
    Slot slot[kSlots];
    hipStream_t s_comp = nullptr, s_copy = nullptr;
    std::thread worker, deliverer;
    // statistics
    int64_t batches = 0, queries = 0;

    void run();
    void run_delivery();
    void launch(Slot& sl);
    void redo_flagged(Slot& sl);
    void deliver(Slot& sl);
};

void lsh_server::launch(Slot& sl) {
    const int d = ix->d;
    int total = 0;
    for (Request* r : sl.reqs) {
        std::memcpy(sl.hq + (size_t)total * d, r->q, sizeof(float) * (size_t)r->n * d);
        total += r->n;
    }
    sl.total = total;
    sl.graph = false;
    sl.status = LSH_OK;
    sl.error.clear();
    try {
        {
            std::lock_guard<std::mutex> lk(ix->mu);
            sl.graph = ix->enqueue_batch(sl.hq, false, total, sl.kq, sl.hres, false, sl.hflags, nullptr, sl.ev_out);
        }
        if (sl.graph) return;
        LSH_HIP(hipMemcpyAsync(sl.dq.get(), sl.hq, sizeof(float) * (size_t)total * d, hipMemcpyHostToDevice, s_copy));
        LSH_HIP(hipEventRecord(sl.ev_in, s_copy));
        LSH_HIP(hipStreamWaitEvent(s_comp, sl.ev_in, 0));
        {
            std::lock_guard<std::mutex> lk(ix->mu);
            ix->search(sl.dq.get(), total, sl.kq, true, sl.dres.get(), true, s_comp);
        }
        LSH_HIP(hipEventRecord(sl.ev_done, s_comp));
        LSH_HIP(hipStreamWaitEvent(s_copy, sl.ev_done, 0));
        LSH_HIP(hipMemcpyAsync(sl.hres, sl.dres.get(), sizeof(lsh_point_t) * (size_t)total * sl.kq,
                               hipMemcpyDeviceToHost, s_copy));
        LSH_HIP(hipEventRecord(sl.ev_out, s_copy));
    } catch (const lshgpu::Error& e) {
        sl.status = e.code;
        sl.error = e.what();
    } catch (const std::exception& e) {
        sl.status = LSH_ERR_INTERNAL;
        sl.error = e.what();
    }
}

// A captured batch's queries that its select left incomplete, or all of them
// when the cells' second pass was too dense (the index then stops using the
// captured form), searched again by an ordinary search whose results
// replace the batch's.
void lsh_server::redo_flagged(Slot& sl) {
    const int d = ix->d;
    const int n = sl.total;
    std::lock_guard<std::mutex> lk(ix->mu);
    if (sl.hflags[n]) {
        ix->cells_dense = true;
        ix->batch_redone += n;
        ix->search(sl.hq, n, sl.kq, false, sl.hres, false, nullptr);
        return;
    }
    std::vector<int> bad;
    for (int i = 0; i < n; ++i)
        if (sl.hflags[i]) bad.push_back(i);
    if (bad.empty()) return;
    const int m = (int)bad.size();
    std::vector<float> q((size_t)m * d);
    std::vector<lsh_point_t> r((size_t)m * sl.kq);
    for (int i = 0; i < m; ++i) std::memcpy(q.data() + (size_t)i * d, sl.hq + (size_t)bad[i] * d, sizeof(float) * d);
    ix->search(q.data(), m, sl.kq, false, r.data(), false, nullptr);
    for (int i = 0; i < m; ++i)
        std::memcpy(sl.hres + (size_t)bad[i] * sl.kq, r.data() + (size_t)i * sl.kq, sizeof(lsh_point_t) * sl.kq);
    ix->batch_redone += m;
}

void lsh_server::deliver(Slot& sl) {
    if (sl.status == LSH_OK) {
        const hipError_t e = hipEventSynchronize(sl.ev_out);
        if (e != hipSuccess) {
            sl.status = LSH_ERR_DEVICE;
            sl.error = std::string("hipEventSynchronize: ") + hipGetErrorString(e);
        }
    }
    if (sl.status == LSH_OK && sl.graph) {
        try {
            redo_flagged(sl);
        } catch (const lshgpu::Error& e) {
            sl.status = e.code;
            sl.error = e.what();
        } catch (const std::exception& e) {
            sl.status = LSH_ERR_INTERNAL;
            sl.error = e.what();
        }
    }
    int row = 0;
    const double t = now_ms();
    for (Request* r : sl.reqs) {
        const int kr = lshgpu::calculate_k(r->k);
        if (sl.status == LSH_OK)
            for (int i = 0; i < r->n; ++i)
                std::memcpy(r->out + (size_t)i * kr, sl.hres + (size_t)(row + i) * sl.kq, sizeof(lsh_point_t) * kr);
        row += r->n;
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        for (Request* r : sl.reqs) {
            r->status = sl.status;
            r->error = sl.error;
            r->t_done = t;
            r->done = true;
        }
        ++batches;
        queries += row;
    }
    cv_done.notify_all();
    sl.reqs.clear();
}

void lsh_server::run() {
    (void)hipSetDevice(ix->device);
    for (;;) {
        int cur = 0;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !queue.empty(); });
            if (queue.empty() && stop) break;
            // A free slot (with all four on the device, the oldest one's
            // delivery frees it); requests keep queueing meanwhile.
            cv_slot.wait(lk, [&] {
                for (const Slot& x : slot)
                    if (!x.busy) return true;
                return false;
            });
            while (slot[cur].busy) ++cur;
            // Admission: whole requests in arrival order while the batch
            // stays within max_batch queries and the memory budget, counted
            // for two batches at this batch's size.
            Slot& sl = slot[cur];
            int total = 0, kq = 0;
            while (!queue.empty()) {
                Request* r = queue.front();
                const int64_t need = (int64_t)(total + r->n) * per_query * 2;
                if (total > 0 && (total + r->n > max_batch || need > budget)) break;
                queue.pop_front();
                sl.reqs.push_back(r);
                total += r->n;
                kq = std::max(kq, lshgpu::calculate_k(r->k));
            }
            sl.kq = kq;
            sl.busy = true;
        }
        launch(slot[cur]);
        {
            std::lock_guard<std::mutex> lk(mu);
            inflight.push_back(cur);
        }
        cv_inflight.notify_one();
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        launcher_done = true;
    }
    cv_inflight.notify_one();
}

void lsh_server::run_delivery() {
    (void)hipSetDevice(ix->device);
    for (;;) {
        int cur;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_inflight.wait(lk, [&] { return !inflight.empty() || launcher_done; });
            if (inflight.empty()) break;
            cur = inflight.front();
            inflight.pop_front();
        }
        deliver(slot[cur]);
        {
            std::lock_guard<std::mutex> lk(mu);
            slot[cur].busy = false;
        }
        cv_slot.notify_one();
    }
}

extern "C" {

void lsh_server_config_default(lsh_server_config_t* cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->max_batch_queries = 4096;
    cfg->max_query_size = 2048;
    cfg->memory_budget_bytes = 0;
}

int lsh_server_create(lsh_index_t* index, const lsh_server_config_t* cfg_in, lsh_server_t** out) {
    return guarded([&] {
        LSH_REQUIRE(index && out, LSH_ERR_INVALID, "null pointer");
        *out = nullptr;
        lsh_server_config_t cfg;
        lsh_server_config_default(&cfg);
        if (cfg_in) cfg = *cfg_in;
        LSH_REQUIRE(cfg.max_batch_queries >= 1 && cfg.max_query_size >= 1 &&
                        cfg.max_query_size <= cfg.max_batch_queries,
                    LSH_ERR_INVALID, "need 1 <= max_query_size <= max_batch_queries");
        LSH_HIP(hipSetDevice(index->device));
        auto sv = std::make_unique<lsh_server>();
        sv->ix = index;
        sv->max_batch = cfg.max_batch_queries;
        // Device bytes a query of a batch may take: its copy, its results at
        // k = 128, and the search's per-query state (the top-C keys with
        // their append space, at most (2C + 512) 8-byte keys, plus lists).
        const int C = std::min<int64_t>(index->candidates, std::max<int64_t>(index->N, 1));
        sv->per_query = (int64_t)sizeof(float) * index->d * 2 + (int64_t)sizeof(lsh_point_t) * 128 * 2 +
                        (int64_t)8 * (3 * C + 768) + 4096;
        size_t free_b = 0, total_b = 0;
        LSH_HIP(hipMemGetInfo(&free_b, &total_b));
        const int64_t avail = (int64_t)(free_b / 2);
        sv->budget = cfg.memory_budget_bytes > 0 ? std::min<int64_t>(cfg.memory_budget_bytes, avail) : avail;
        LSH_REQUIRE(sv->budget >= (int64_t)cfg.max_query_size * sv->per_query * 2, LSH_ERR_NOMEM,
                    "the memory budget does not admit one request of max_query_size queries");
        // Buffers for the largest batch the budget admits.
        const int cap = (int)std::min<int64_t>(cfg.max_batch_queries, sv->budget / (2 * sv->per_query));
        sv->max_batch = std::max(cap, cfg.max_query_size);
        for (Slot& sl : sv->slot) {
            void* p = nullptr;
            LSH_HIP(hipHostMalloc(&p, sizeof(float) * (size_t)sv->max_batch * index->d, hipHostMallocDefault));
            sl.hq = static_cast<float*>(p);
            LSH_HIP(hipHostMalloc(&p, sizeof(lsh_point_t) * (size_t)sv->max_batch * 128, hipHostMallocDefault));
            sl.hres = static_cast<lsh_point_t*>(p);
            LSH_HIP(hipHostMalloc(&p, sizeof(uint32_t) * ((size_t)sv->max_batch + 3), hipHostMallocDefault));
            sl.hflags = static_cast<uint32_t*>(p);
            sl.dq.alloc((size_t)sv->max_batch * index->d);
            sl.dres.alloc((size_t)sv->max_batch * 128);
            LSH_HIP(hipEventCreateWithFlags(&sl.ev_in, hipEventDisableTiming));
            LSH_HIP(hipEventCreateWithFlags(&sl.ev_done, hipEventDisableTiming));
            LSH_HIP(hipEventCreateWithFlags(&sl.ev_out, hipEventDisableTiming));
        }
        LSH_HIP(hipStreamCreateWithFlags(&sv->s_comp, hipStreamNonBlocking));
        LSH_HIP(hipStreamCreateWithFlags(&sv->s_copy, hipStreamNonBlocking));
        lsh_server* raw = sv.get();
        sv->deliverer = std::thread([raw] { raw->run_delivery(); });
        try {
            sv->worker = std::thread([raw] { raw->run(); });
        } catch (...) {
            {
                std::lock_guard<std::mutex> lk(raw->mu);
                raw->launcher_done = true;
            }
            raw->cv_inflight.notify_one();
            raw->deliverer.join();
            throw;
        }
        *out = sv.release();
    });
}

int lsh_server_submit(lsh_server_t* sv, const float* queries, int n_query, int k, lsh_point_t* results,
                      int64_t* ticket) {
    return guarded([&] {
        LSH_REQUIRE(sv && ticket, LSH_ERR_INVALID, "null pointer");
        LSH_REQUIRE(n_query >= 1 && queries && results, LSH_ERR_INVALID, "a request needs queries and results");
        LSH_REQUIRE(k >= 1 && k <= 128, LSH_ERR_UNSUPPORTED, "served requests take k in 1..128");
        LSH_REQUIRE(n_query <= sv->max_batch, LSH_ERR_UNSUPPORTED,
                    "a request larger than the largest batch the memory budget admits");
        auto r = std::make_unique<Request>();
        r->q = queries;
        r->n = n_query;
        r->k = k;
        r->out = results;
        r->t_submit = now_ms();
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            LSH_REQUIRE(!sv->stop, LSH_ERR_INVALID, "the server is shutting down");
            r->id = sv->next_id++;
            *ticket = r->id;
            sv->queue.push_back(r.get());
            sv->reqs.emplace(r->id, std::move(r));
        }
        sv->cv_work.notify_one();
    });
}

int lsh_server_wait(lsh_server_t* sv, int64_t ticket, double* latency_ms) {
    int st = LSH_OK;
    std::string err;
    const int rc = guarded([&] {
        LSH_REQUIRE(sv, LSH_ERR_INVALID, "null pointer");
        std::unique_lock<std::mutex> lk(sv->mu);
        auto it = sv->reqs.find(ticket);
        LSH_REQUIRE(it != sv->reqs.end(), LSH_ERR_INVALID, "unknown ticket");
        Request* r = it->second.get();
        sv->cv_done.wait(lk, [&] { return r->done; });
        if (latency_ms) *latency_ms = r->t_done - r->t_submit;
        st = r->status;
        err = r->error;
        sv->reqs.erase(it);
    });
    if (rc != LSH_OK) return rc;
    if (st != LSH_OK) lshgpu::last_error() = err;
    return st;
}

int lsh_server_stats(lsh_server_t* sv, int64_t* batches, int64_t* queries, int* max_batch) {
    return guarded([&] {
        LSH_REQUIRE(sv, LSH_ERR_INVALID, "null pointer");
        std::lock_guard<std::mutex> lk(sv->mu);
        if (batches) *batches = sv->batches;
        if (queries) *queries = sv->queries;
        if (max_batch) *max_batch = sv->max_batch;
    });
}

int lsh_server_destroy(lsh_server_t* sv) {
    return guarded([&] {
        if (!sv) return;
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->stop = true;
        }
        sv->cv_work.notify_all();
        if (sv->worker.joinable()) sv->worker.join();
        if (sv->deliverer.joinable()) sv->deliverer.join();
        (void)hipSetDevice(sv->ix->device);
        for (Slot& sl : sv->slot) {
            if (sl.hq) (void)hipHostFree(sl.hq);
            if (sl.hres) (void)hipHostFree(sl.hres);
            if (sl.hflags) (void)hipHostFree(sl.hflags);
            for (hipEvent_t e : {sl.ev_in, sl.ev_done, sl.ev_out})
                if (e) (void)hipEventDestroy(e);
        }
        if (sv->s_comp) (void)hipStreamDestroy(sv->s_comp);
        if (sv->s_copy) (void)hipStreamDestroy(sv->s_copy);
        delete sv;
    });
}

}  // extern "C"
