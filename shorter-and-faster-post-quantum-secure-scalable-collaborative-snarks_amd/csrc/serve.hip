// serve.hip -- a device-resident proof server: continuous batching of plain
// proofs under a fixed HBM budget.
//
// Clients submit requests (device pointers to f and g, the size, a
// transcript start) from any thread.  A scheduler thread owns a context of
// its own and an arena of `budget` bytes of device memory; every request's
// working memory (codeword, tree, sumcheck buffers, state) and every pass's
// tables and records come from it.  Each tick the scheduler
//   1. retires finished passes: unpacks their proofs, frees their memory;
//   2. admits queued requests in arrival order while their working memory
//      fits the arena (the first that does not fit waits, and those behind
//      it wait too, so large requests are not starved by small ones);
//   3. launches the admitted requests as one batched pass (the kernels of
//      the batch API) on a free slot, a stream pair of its own.
// Passes on different slots overlap on the device, so requests that arrive
// while others are on the GPU start as soon as memory and a slot free up.
// A request's proof is the plain prover's for the same inputs and transcript.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>
#include "ctx.h"

using gf::fe;
using Clock = std::chrono::steady_clock;

namespace {

constexpr int SRV_MAX_PER_PASS = 256;
constexpr size_t SRV_TAB_CAP = 8u << 20;     // pinned staging per slot
constexpr size_t SRV_OUT_CAP = 1u << 20;

// First-fit allocator over one device block, 256-byte granules.
class Arena {
  public:
    bool init(size_t bytes) {
        size_ = bytes / 256 * 256;
        if (hipMalloc(&base_, size_) != hipSuccess) return false;
        free_[0] = size_;
        return true;
    }
    void destroy() {
        if (base_) (void)hipFree(base_);
        base_ = nullptr;
    }
    void *alloc(size_t bytes) {
        bytes = align256(bytes ? bytes : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > bytes) free_[off + bytes] = len - bytes;
            used_ += bytes;
            if (used_ > peak_) peak_ = used_;
            return base_ + off;
        }
        return nullptr;
    }
    void release(void *p, size_t bytes) {
        if (!p) return;
        bytes = align256(bytes ? bytes : 1);
        size_t off = (uint8_t *)p - base_, len = bytes;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + len == next->first) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                off = prev->first;
                len += prev->second;
                free_.erase(prev);
            }
        }
        free_[off] = len;
        used_ -= bytes;
    }
    size_t size() const { return size_; }
    size_t peak() const { return peak_; }
    size_t used() const { return used_; }

  private:
    uint8_t *base_ = nullptr;
    size_t size_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;     // offset -> length
};

struct Request {
    uint64_t id = 0;
    const fe *f = nullptr, *g = nullptr;
    uint64_t n = 0;
    fe prev{0, 0};
    Clock::time_point t_submit, t_done;
    void *work = nullptr;
    size_t work_bytes = 0;
    fe *cw = nullptr;                   // caller's buffers for the commitment, or null
    uint8_t *nodes = nullptr;
    std::vector<uint8_t> proof;
    uint64_t prev_out[2] = {0, 0};
    int status = COSNARK_OK;
    bool done = false;
    bool claimed = false;               // a wait has taken the ticket
};

struct Slot {
    hipStream_t s1 = nullptr, s2 = nullptr;
    hipEvent_t fork = nullptr, join = nullptr, done = nullptr;
    bool busy = false;
    std::vector<Request *> reqs;
    BatchCore core;
    void *shared = nullptr;             // arena: transcript starts + records
    size_t shared_bytes = 0;
    void *tables = nullptr;             // arena: launch tables
    size_t tables_bytes = 0;
    uint8_t *tab_host = nullptr, *out_host = nullptr;   // pinned staging
    fe *prev_host = nullptr;
    fe *tmp = nullptr;                  // arena: NTT temporary
    uint64_t tmp_elems = 0;
};

}  // namespace

struct cosnark_server {
    cosnark_ctx *ctx = nullptr;
    int rate = 4, max_log_n = 0;
    Arena arena;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queue;
    std::unordered_map<uint64_t, Request *> all;
    uint64_t next_id = 1;
    std::thread worker;
    std::atomic<bool> stop{false};
    uint64_t passes = 0, completed = 0, inflight = 0, max_inflight = 0;
};

namespace {

void retire(cosnark_server *S, Slot &sl, hipError_t q) {
    const size_t k = sl.reqs.size();
    std::vector<uint8_t *> dst(k);
    std::vector<cosnark_transcript> ts(k);
    std::vector<int32_t> errs(k, 0);
    for (size_t i = 0; i < k; ++i) {
        sl.reqs[i]->proof.resize(cosnark_plain_proof_bytes(sl.reqs[i]->n));
        dst[i] = sl.reqs[i]->proof.data();
    }
    if (q == hipSuccess) batch_unpack(sl.core, sl.out_host, dst.data(), ts.data(), errs.data());
    const auto now = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        for (size_t i = 0; i < k; ++i) {
            Request *r = sl.reqs[i];
            r->status = q != hipSuccess ? COSNARK_E_HIP : errs[i] ? COSNARK_E_CHECK : COSNARK_OK;
            r->prev_out[0] = ts[i].prev[0];
            r->prev_out[1] = ts[i].prev[1];
            r->t_done = now;
            r->done = true;
            S->arena.release(r->work, r->work_bytes);
            r->work = nullptr;
        }
        S->arena.release(sl.shared, sl.shared_bytes);
        S->arena.release(sl.tables, sl.tables_bytes);
        sl.shared = sl.tables = nullptr;
        S->inflight -= k;
        S->completed += k;
    }
    sl.reqs.clear();
    sl.busy = false;
    S->cv_done.notify_all();
}

// requeue at the front, in order, and give their memory back (mu held)
void requeue(cosnark_server *S, std::vector<Request *> &take) {
    for (auto it = take.rbegin(); it != take.rend(); ++it) {
        S->arena.release((*it)->work, (*it)->work_bytes);
        (*it)->work = nullptr;
        S->queue.push_front(*it);
    }
    take.clear();
}

// admit and launch one pass on `sl`; false when nothing was launched
bool launch(cosnark_server *S, Slot &sl) {
    std::vector<Request *> take;
    size_t out_bytes = 0;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        while (!S->queue.empty() && take.size() < (size_t)SRV_MAX_PER_PASS) {
            Request *r = S->queue.front();
            const size_t rec = batch_record_bytes(r->n);
            if (out_bytes + rec > SRV_OUT_CAP) break;
            void *w = S->arena.alloc(r->work_bytes);
            if (!w) break;
            r->work = w;
            out_bytes += rec;
            take.push_back(r);
            S->queue.pop_front();
        }
        if (take.empty()) return false;
        // transcript starts and records
        const size_t prev_b = align256(sizeof(fe) * take.size());
        sl.shared_bytes = prev_b + out_bytes;
        sl.shared = S->arena.alloc(sl.shared_bytes);
        if (!sl.shared) {
            requeue(S, take);
            return false;
        }
    }
    const int k = (int)take.size();
    std::vector<BatchProof> pr(k);
    for (int i = 0; i < k; ++i) {
        pr[i] = {take[i]->f, take[i]->g, take[i]->n, take[i]->work, take[i]->cw, take[i]->nodes};
        sl.prev_host[i] = take[i]->prev;
    }
    fe *prev_dev = (fe *)sl.shared;
    uint8_t *out_dev = (uint8_t *)sl.shared + align256(sizeof(fe) * k);
    int rc = batch_core_build(S->ctx, k, pr.data(), S->rate, prev_dev, out_dev, sl.core);
    const size_t tb = sl.core.tb.host.size();
    if (rc == COSNARK_OK && tb <= SRV_TAB_CAP) {
        std::lock_guard<std::mutex> lk(S->mu);
        sl.tables_bytes = tb;
        sl.tables = S->arena.alloc(tb);
    }
    if (rc != COSNARK_OK || tb > SRV_TAB_CAP || !sl.tables) {
        // no room for the tables (or no plan): fail nothing, retry later
        std::lock_guard<std::mutex> lk(S->mu);
        S->arena.release(sl.shared, sl.shared_bytes);
        sl.shared = nullptr;
        if (rc != COSNARK_OK || tb > SRV_TAB_CAP) {
            for (Request *r : take) {
                S->arena.release(r->work, r->work_bytes);
                r->work = nullptr;
                r->status = rc != COSNARK_OK ? rc : COSNARK_E_RANGE;
                r->t_done = Clock::now();
                r->done = true;
                S->completed++;
            }
            S->cv_done.notify_all();
        } else {
            requeue(S, take);
        }
        return false;
    }
    memcpy(sl.tab_host, sl.core.tb.host.data(), tb);
    hipError_t e = hipMemcpyAsync(sl.tables, sl.tab_host, tb, hipMemcpyHostToDevice, sl.s1);
    if (e == hipSuccess)
        rc = batch_core_enqueue(S->ctx, sl.core, (const uint8_t *)sl.tables, sl.prev_host, sl.out_host, sl.tmp,
                                sl.tmp_elems, sl.s1, sl.s2, sl.fork, sl.join);
    else
        rc = COSNARK_E_HIP;
    if (rc == COSNARK_OK && hipEventRecord(sl.done, sl.s1) != hipSuccess) rc = COSNARK_E_HIP;
    sl.reqs = take;
    sl.busy = true;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->inflight += k;
        if (S->inflight > S->max_inflight) S->max_inflight = S->inflight;
        S->passes++;
    }
    if (rc != COSNARK_OK) {
        (void)hipStreamSynchronize(sl.s1);
        (void)hipStreamSynchronize(sl.s2);
        retire(S, sl, hipErrorUnknown);
    }
    return true;
}

void serve_loop(cosnark_server *S) {
    (void)hipSetDevice(S->ctx->device);
    for (;;) {
        bool progressed = false;
        for (auto &sl : S->slots) {
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipErrorNotReady) continue;
            retire(S, sl, q);
            progressed = true;
        }
        for (auto &sl : S->slots) {
            if (sl.busy) continue;
            if (launch(S, sl)) progressed = true;
            break;
        }
        if (progressed) continue;
        std::unique_lock<std::mutex> lk(S->mu);
        bool busy = false;
        for (auto &sl : S->slots) busy = busy || sl.busy;
        if (S->stop.load() && S->queue.empty() && !busy) break;
        // idle: wake on a submission; poll in-flight passes every 20 us
        S->cv_work.wait_for(lk, std::chrono::microseconds(busy ? 20 : 1000));
    }
}

void server_release(cosnark_server *S) {
    for (auto &sl : S->slots) {
        if (sl.s1) (void)hipStreamSynchronize(sl.s1);
        if (sl.s2) (void)hipStreamSynchronize(sl.s2);
        for (hipEvent_t e : {sl.fork, sl.join, sl.done})
            if (e) (void)hipEventDestroy(e);
        if (sl.s1) (void)hipStreamDestroy(sl.s1);
        if (sl.s2) (void)hipStreamDestroy(sl.s2);
        if (sl.tab_host) (void)hipHostFree(sl.tab_host);
        if (sl.out_host) (void)hipHostFree(sl.out_host);
        if (sl.prev_host) (void)hipHostFree(sl.prev_host);
    }
    S->arena.destroy();
    for (auto &kv : S->all) delete kv.second;
    if (S->ctx) cosnark_destroy(S->ctx);
    delete S;
}

}  // namespace

extern "C" {

int cosnark_server_create(int device, uint64_t budget, int rate, int max_log_n, int nslots, cosnark_server **out) {
    if (!out) return COSNARK_E_ARG;
    *out = nullptr;
    if (rate < 1 || !is_pow2((uint64_t)rate) || max_log_n < 2 || nslots < 1 || nslots > 8 ||
        ((uint64_t)rate << max_log_n) > (1ull << 28))
        return COSNARK_E_ARG;
    auto *S = new cosnark_server;
    S->rate = rate;
    S->max_log_n = max_log_n;
    int rc = cosnark_create(device, &S->ctx);
    if (rc) {
        S->ctx = nullptr;
        server_release(S);
        return rc;
    }
    const uint64_t tmp_elems = (uint64_t)rate << max_log_n;
    const size_t need = nslots * align256(sizeof(fe) * tmp_elems) + batch_work_bytes(1ull << max_log_n, rate) +
                        (2u << 20);
    if (budget < need || !S->arena.init(budget)) {
        server_release(S);
        return budget < need ? COSNARK_E_ARG : COSNARK_E_NOMEM;
    }
    S->slots.resize(nslots);
    for (auto &sl : S->slots) {
        sl.tmp_elems = tmp_elems;
        sl.tmp = (fe *)S->arena.alloc(sizeof(fe) * tmp_elems);
        if (!sl.tmp || hipStreamCreateWithFlags(&sl.s1, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&sl.s2, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&sl.join, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc(&sl.tab_host, SRV_TAB_CAP, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc(&sl.out_host, SRV_OUT_CAP, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc(&sl.prev_host, sizeof(fe) * SRV_MAX_PER_PASS, hipHostMallocDefault) != hipSuccess) {
            server_release(S);
            return COSNARK_E_NOMEM;
        }
    }
    S->worker = std::thread(serve_loop, S);
    *out = S;
    return COSNARK_OK;
}

int cosnark_server_submit(cosnark_server *S, const uint64_t *d_f, const uint64_t *d_g, uint64_t n,
                          const uint64_t prev[2], uint64_t *ticket) {
    return cosnark_server_submit_ex(S, d_f, d_g, n, prev, nullptr, nullptr, ticket);
}

int cosnark_server_submit_ex(cosnark_server *S, const uint64_t *d_f, const uint64_t *d_g, uint64_t n,
                             const uint64_t prev[2], uint64_t *d_codeword, uint8_t *d_nodes, uint64_t *ticket) {
    if (!S || !d_f || !d_g || !ticket) return COSNARK_E_ARG;
    if ((d_codeword == nullptr) != (d_nodes == nullptr)) return COSNARK_E_ARG;
    if (!is_pow2(n) || n < 4 || n * (uint64_t)S->rate < 8) return COSNARK_E_ARG;
    if (n > (1ull << S->max_log_n)) return COSNARK_E_RANGE;
    auto *r = new Request;
    r->f = (const fe *)d_f;
    r->g = (const fe *)d_g;
    r->n = n;
    if (prev) r->prev = gf::mk(prev[0], prev[1]);
    r->cw = (fe *)d_codeword;
    r->nodes = d_nodes;
    r->work_bytes = batch_work_bytes(n, S->rate, d_codeword == nullptr);
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop.load()) {
            delete r;
            return COSNARK_E_ARG;
        }
        r->id = S->next_id++;
        S->all[r->id] = r;
        S->queue.push_back(r);
        *ticket = r->id;
    }
    S->cv_work.notify_one();
    return COSNARK_OK;
}

int cosnark_server_poll(cosnark_server *S, uint64_t ticket, int *done) {
    if (!S || !done) return COSNARK_E_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(ticket);
    if (it == S->all.end()) return COSNARK_E_ARG;
    *done = it->second->done ? 1 : 0;
    return COSNARK_OK;
}

int cosnark_server_wait(cosnark_server *S, uint64_t ticket, uint8_t *h_proof, uint64_t prev_out[2],
                        double *latency_ms) {
    if (!S) return COSNARK_E_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(ticket);
    if (it == S->all.end()) return COSNARK_E_ARG;
    Request *r = it->second;
    // one wait per ticket: a second waiter would read the request after the
    // first one deleted it
    if (r->claimed) return COSNARK_E_ARG;
    r->claimed = true;
    S->cv_done.wait(lk, [&] { return r->done; });
    if (h_proof && r->status == COSNARK_OK) memcpy(h_proof, r->proof.data(), r->proof.size());
    if (prev_out) {
        prev_out[0] = r->prev_out[0];
        prev_out[1] = r->prev_out[1];
    }
    if (latency_ms) latency_ms[0] = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    const int status = r->status;
    S->all.erase(it);
    delete r;
    return status;
}

int cosnark_server_stats(cosnark_server *S, uint64_t out[6]) {
    if (!S || !out) return COSNARK_E_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    out[0] = S->passes;
    out[1] = S->completed;
    out[2] = S->max_inflight;
    out[3] = S->arena.peak();
    out[4] = S->arena.size();
    out[5] = S->queue.size();
    return COSNARK_OK;
}

void cosnark_server_destroy(cosnark_server *S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop.store(true);
    }
    S->cv_work.notify_all();
    if (S->worker.joinable()) S->worker.join();
    (void)hipSetDevice(S->ctx ? S->ctx->device : 0);
    server_release(S);
}

}  // extern "C"
