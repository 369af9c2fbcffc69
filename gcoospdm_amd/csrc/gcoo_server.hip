// gcoo_server.hip -- continuous-batching request loop for SpDM jobs.
//
// No counterpart in the reference (a single-process benchmark that runs one
// multiply at a time, src/main.cu:46-81).  The server keeps operands resident
// on the device (registered GCOO handles and their dense D), accepts requests
// "C = S_op * D_op" from any thread, and runs a scheduler thread that
//   * admits queued requests in arrival order while their output fits the HBM
//     budget: outputs are carved first-fit out of one device arena of
//     `hbm_budget_bytes`, so the bytes in flight never exceed the budget (a
//     request larger than the whole arena is rejected at submit);
//   * whenever one of `slots` streams is idle, launches every admitted, not
//     yet running request (up to `max_batch_jobs`) as one persistent batch on
//     it -- requests that arrive while other batches are on the GPU go into
//     the next batch on another stream, so their work overlaps on the device;
//   * retires a batch when its completion event fires, timestamps its
//     requests (arrival -> completion latency) and wakes their waiters.
// Results stay in the arena until the client releases the request.  Every
// result is the plain kernels' fmaf chain, bit-identical alone or batched.
#include "gcoo_internal.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

struct Operand {
    gcoo_handle h;
    const void *D;
    int dtype;
    int N;
    int rows, vec;
};

struct Request {
    uint64_t id;
    int op;
    size_t offset = 0, bytes = 0;
    Clock::time_point arrival, done;
    int state = 0;  // 0 queued, 1 admitted, 2 running, 3 done, 4 released
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t event = nullptr;
    gcoo_batch_t batch = nullptr;
    std::vector<uint64_t> reqs;
    bool busy = false;
};

// First-fit allocator over [0, size) with 256-byte granularity.
class Arena {
public:
    void reset(size_t size)
    {
        free_.clear();
        if (size) free_[0] = size;
        used_ = peak_ = 0;
    }
    bool alloc(size_t bytes, size_t *off)
    {
        bytes = (bytes + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes;
                const size_t start = it->first + bytes;
                free_.erase(it);
                if (rest) free_[start] = rest;
                used_ += bytes;
                peak_ = std::max(peak_, used_);
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes)
    {
        bytes = (bytes + 255) & ~(size_t)255;
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }
    void reset_peak() { peak_ = used_; }

private:
    std::map<size_t, size_t> free_;
    size_t used_ = 0, peak_ = 0;
};

}  // namespace

using namespace gcoo;

struct gcoo_server {
    gcoo_server_config cfg{};
    int device = 0;
    char *arenaBase = nullptr;
    Arena arena;
    std::vector<Operand> ops;
    std::unordered_map<uint64_t, Request> reqs;
    std::deque<uint64_t> queue;     // waiting for admission (FIFO)
    std::deque<uint64_t> admitted;  // output allocated, not yet launched
    std::vector<Slot> slots;
    uint64_t nextId = 1;
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::thread worker;
    bool stop = false;
    // bumped by every event that can make admission or launch possible
    // (submit, release, stop); the idle scheduler sleeps until it changes
    uint64_t wakeups = 0;
    int workerError = 0;
    char workerMsg[256] = "";
    // statistics
    std::vector<double> latMs;
    uint64_t batches = 0, jobsRun = 0;
    Clock::time_point firstArrival, lastDone;
    bool anyArrival = false;
};

namespace {

size_t out_bytes(const gcoo_server *s, int op)
{
    const Operand &o = s->ops[op];
    return (size_t)o.h->M * (size_t)o.N * sizeof(float);
}

// Called with the lock held: admit queued requests in order while they fit.
void admit(gcoo_server *s)
{
    while (!s->queue.empty()) {
        Request &r = s->reqs[s->queue.front()];
        size_t off;
        if (!s->arena.alloc(r.bytes, &off)) break;
        r.offset = off;
        r.state = 1;
        s->admitted.push_back(r.id);
        s->queue.pop_front();
    }
}

// Called with the lock held: launch admitted requests on an idle slot.
int launch(gcoo_server *s, Slot &slot)
{
    int rc = gcoo_batch_clear(slot.batch);
    if (rc) return rc;
    slot.reqs.clear();
    while (!s->admitted.empty() && (int)slot.reqs.size() < s->cfg.max_batch_jobs) {
        Request &r = s->reqs[s->admitted.front()];
        const Operand &o = s->ops[r.op];
        float *C = reinterpret_cast<float *>(s->arenaBase + r.offset);
        rc = gcoo_batch_add(slot.batch, o.h, o.D, o.dtype, o.N, C, o.rows, o.vec, nullptr);
        if (rc) return rc;
        r.state = 2;
        slot.reqs.push_back(r.id);
        s->admitted.pop_front();
    }
    if (slot.reqs.empty()) return GCOO_OK;
    if ((rc = batch_commit_async(slot.batch, slot.stream))) return rc;
    if ((rc = gcoo_batch_run_async(slot.batch, slot.stream))) return rc;
    GCOO_HIP_TRY(hipEventRecord(slot.event, slot.stream));
    slot.busy = true;
    ++s->batches;
    s->jobsRun += slot.reqs.size();
    return GCOO_OK;
}

void worker_main(gcoo_server *s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        bool progressed = false;
        // retire finished batches
        for (Slot &slot : s->slots) {
            if (!slot.busy) continue;
            const hipError_t q = hipEventQuery(slot.event);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) {
                s->workerError = GCOO_ERR_HIP;
                snprintf(s->workerMsg, sizeof(s->workerMsg), "batch failed: %s",
                         hipGetErrorString(q));
            }
            const auto now = Clock::now();
            for (uint64_t id : slot.reqs) {
                Request &r = s->reqs[id];
                r.done = now;
                r.state = 3;
                s->latMs.push_back(std::chrono::duration<double, std::milli>(now - r.arrival).count());
            }
            s->lastDone = now;
            slot.reqs.clear();
            slot.busy = false;
            progressed = true;
            s->cvDone.notify_all();
        }
        admit(s);
        for (Slot &slot : s->slots) {
            if (slot.busy || s->admitted.empty()) continue;
            const int rc = launch(s, slot);
            if (rc) {
                s->workerError = rc;
                snprintf(s->workerMsg, sizeof(s->workerMsg), "launch failed: %s",
                         gcoo_last_error());
                s->stop = true;
            }
            progressed = true;
        }
        const bool anyBusy = std::any_of(s->slots.begin(), s->slots.end(),
                                         [](const Slot &x) { return x.busy; });
        if (s->stop && !anyBusy) break;
        if (!progressed) {
            if (anyBusy) {
                // poll completion events without holding the lock
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
                lk.lock();
            } else {
                // Nothing runs and nothing could be admitted or launched:
                // wait for a submit, a release or stop.  (Waiting on "queue
                // non-empty" instead would spin with the lock held while a
                // queued request waits for budget.)
                const uint64_t seen = s->wakeups;
                s->cvWork.wait(lk, [&] { return s->stop || s->wakeups != seen; });
            }
        }
    }
    s->cvDone.notify_all();
}

double percentile(std::vector<double> v, double p)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = p * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

}  // namespace

extern "C" {

int gcoo_server_create(const gcoo_server_config *cfg, gcoo_server_t *out)
{
    if (!cfg || !out) return set_error(GCOO_ERR_INVALID, "null argument");
    *out = nullptr;
    if (cfg->hbm_budget_bytes == 0 || cfg->slots < 1 || cfg->slots > 8 || cfg->max_batch_jobs < 1)
        return set_error(GCOO_ERR_INVALID, "budget > 0, 1 <= slots <= 8, max_batch_jobs >= 1");
    auto *s = new (std::nothrow) gcoo_server;
    if (!s) return set_error(GCOO_ERR_ALLOC, "server");
    s->cfg = *cfg;
    (void)hipGetDevice(&s->device);
    if (hipMalloc(&s->arenaBase, cfg->hbm_budget_bytes) != hipSuccess) {
        (void)hipGetLastError();
        delete s;
        return set_error(GCOO_ERR_ALLOC, "cannot allocate the %zu-byte output arena",
                         cfg->hbm_budget_bytes);
    }
    s->arena.reset(cfg->hbm_budget_bytes);
    s->slots.resize(cfg->slots);
    for (Slot &slot : s->slots) {
        if (hipStreamCreateWithFlags(&slot.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&slot.event, hipEventDisableTiming) != hipSuccess ||
            gcoo_batch_create(&slot.batch) != GCOO_OK) {
            (void)hipGetLastError();
            gcoo_server_destroy(s);
            return set_error(GCOO_ERR_HIP, "cannot create server streams");
        }
    }
    s->worker = std::thread(worker_main, s);
    *out = s;
    return GCOO_OK;
}

int gcoo_server_destroy(gcoo_server_t s)
{
    if (!s) return GCOO_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->queue.clear();
        ++s->wakeups;
    }
    s->cvWork.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (Slot &slot : s->slots) {
        if (slot.stream) (void)hipStreamSynchronize(slot.stream);
        if (slot.batch) gcoo_batch_destroy(slot.batch);
        if (slot.event) (void)hipEventDestroy(slot.event);
        if (slot.stream) (void)hipStreamDestroy(slot.stream);
    }
    if (s->arenaBase) (void)hipFree(s->arenaBase);
    delete s;
    return GCOO_OK;
}

int gcoo_server_register(gcoo_server_t s, gcoo_handle h, const void *D, int dtype, int N,
                         int rows, int vec, int *op_id)
{
    if (!s || !h || !op_id) return set_error(GCOO_ERR_INVALID, "null argument");
    if (!h->converted) return set_error(GCOO_ERR_STATE, "matrix not converted");
    if (dtype != 0 && dtype != 1) return set_error(GCOO_ERR_INVALID, "dtype %d", dtype);
    if (N < 0) return set_error(GCOO_ERR_INVALID, "N %d", N);
    if (dtype == 1) {
        // derive the resident BF16 values now, so that batches on any slot
        // stream can read them
        int rc = ensure_seg16(h, nullptr);
        if (rc) return rc;
        GCOO_HIP_TRY(hipStreamSynchronize(nullptr));
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->ops.push_back(Operand{h, D, dtype, N, rows, vec});
    *op_id = (int)s->ops.size() - 1;
    return GCOO_OK;
}

int gcoo_server_submit(gcoo_server_t s, int op_id, uint64_t *req_id)
{
    if (!s || !req_id) return set_error(GCOO_ERR_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (op_id < 0 || op_id >= (int)s->ops.size()) return set_error(GCOO_ERR_INVALID, "op %d", op_id);
    if (s->workerError) return set_error(s->workerError, "%s", s->workerMsg);
    if (s->stop) return set_error(GCOO_ERR_STATE, "server stopping");
    const size_t bytes = std::max<size_t>(out_bytes(s, op_id), 256);
    if (bytes > s->cfg.hbm_budget_bytes)
        return set_error(GCOO_ERR_BUDGET, "request needs %zu bytes, budget is %zu", bytes,
                         s->cfg.hbm_budget_bytes);
    Request r;
    r.id = s->nextId++;
    r.op = op_id;
    r.bytes = bytes;
    r.arrival = Clock::now();
    if (!s->anyArrival) {
        s->firstArrival = r.arrival;
        s->anyArrival = true;
    }
    s->reqs.emplace(r.id, r);
    s->queue.push_back(r.id);
    ++s->wakeups;
    *req_id = r.id;
    lk.unlock();
    s->cvWork.notify_one();
    return GCOO_OK;
}

int gcoo_server_wait(gcoo_server_t s, uint64_t req_id, double *latency_ms)
{
    if (!s) return set_error(GCOO_ERR_INVALID, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(req_id);
    if (it == s->reqs.end()) return set_error(GCOO_ERR_INVALID, "unknown request %llu",
                                              (unsigned long long)req_id);
    s->cvDone.wait(lk, [&] { return it->second.state >= 3 || s->workerError || s->stop; });
    if (it->second.state < 3)
        return set_error(s->workerError ? s->workerError : GCOO_ERR_STATE, "%s",
                         s->workerError ? s->workerMsg : "server stopped");
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(it->second.done - it->second.arrival)
                          .count();
    return GCOO_OK;
}

int gcoo_server_result(gcoo_server_t s, uint64_t req_id, const float **C, int *M, int *N)
{
    if (!s) return set_error(GCOO_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(req_id);
    if (it == s->reqs.end() || it->second.state != 3)
        return set_error(GCOO_ERR_STATE, "request %llu not completed", (unsigned long long)req_id);
    const Operand &o = s->ops[it->second.op];
    if (C) *C = reinterpret_cast<const float *>(s->arenaBase + it->second.offset);
    if (M) *M = o.h->M;
    if (N) *N = o.N;
    return GCOO_OK;
}

int gcoo_server_release(gcoo_server_t s, uint64_t req_id)
{
    if (!s) return set_error(GCOO_ERR_INVALID, "null server");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(req_id);
        if (it == s->reqs.end() || it->second.state != 3)
            return set_error(GCOO_ERR_STATE, "request %llu not completed",
                             (unsigned long long)req_id);
        s->arena.release(it->second.offset, it->second.bytes);
        s->reqs.erase(it);
        ++s->wakeups;
    }
    s->cvWork.notify_one();
    return GCOO_OK;
}

int gcoo_server_stats(gcoo_server_t s, gcoo_server_stats_t *st)
{
    if (!s || !st) return set_error(GCOO_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    st->completed = s->latMs.size();
    st->batches = s->batches;
    st->p50_ms = percentile(s->latMs, 0.50);
    st->p99_ms = percentile(s->latMs, 0.99);
    st->max_ms = s->latMs.empty() ? 0.0 : *std::max_element(s->latMs.begin(), s->latMs.end());
    double sum = 0;
    for (double x : s->latMs) sum += x;
    st->mean_ms = s->latMs.empty() ? 0.0 : sum / (double)s->latMs.size();
    st->wall_s = s->anyArrival
                     ? std::chrono::duration<double>(s->lastDone - s->firstArrival).count()
                     : 0.0;
    st->peak_bytes = s->arena.peak();
    st->budget_bytes = s->cfg.hbm_budget_bytes;
    return GCOO_OK;
}

int gcoo_server_reset_stats(gcoo_server_t s)
{
    if (!s) return set_error(GCOO_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->latMs.clear();
    s->batches = s->jobsRun = 0;
    s->anyArrival = false;
    s->arena.reset_peak();
    return GCOO_OK;
}

}  // extern "C"
