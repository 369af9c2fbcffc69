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
//     the next batch on another stream, so their work overlaps on the device.
//     With max_batch_jobs == 1 it is a plain FIFO: one stream, each request
//     one launch of the plain multiply, up to kPlainDepth launches queued.
// A waiter thread per stream blocks on its oldest launch's completion event,
// then timestamps that launch's requests (arrival -> completion latency),
// wakes their waiters and the scheduler.
// Results stay in the arena until the client releases the request.  Every
// result is the plain kernels' fmaf chain, bit-identical alone or batched.
#include "gcoo_internal.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
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

enum ReqState { kQueued = 0, kAdmitted = 1, kRunning = 2, kDone = 3, kFailed = 4 };

struct Request {
    uint64_t id;
    int op;
    size_t offset = 0, bytes = 0;
    bool allocated = false;  // holds arena bytes [offset, offset + bytes)
    Clock::time_point arrival, done;
    int state = kQueued;
    int err = GCOO_OK;       // error code of a failed request
};

// One launch in flight on a slot's stream: its completion event and requests.
struct Launch {
    hipEvent_t event = nullptr;
    std::vector<uint64_t> reqs;
};

// A stream with its launches in flight (oldest first).  Batched mode keeps
// one launch per slot in flight (the slot's batch object holds its job
// table); plain mode keeps up to kPlainDepth on the one slot.  A waiter
// thread per slot blocks on the oldest launch's event and retires it.
struct Slot {
    hipStream_t stream = nullptr;
    gcoo_batch_t batch = nullptr;
    std::vector<hipEvent_t> freeEvents;
    std::deque<Launch> inflight;
    std::condition_variable cv;  // a launch was queued, or the server stops
    std::thread waiter;
};

constexpr int kPlainDepth = 64;

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
    bool plain = false;  // max_batch_jobs == 1: one request per plain launch
    int depth = 1;       // launches in flight per slot
    char *arenaBase = nullptr;
    Arena arena;
    std::vector<Operand> ops;
    std::unordered_map<uint64_t, Request> reqs;
    std::deque<uint64_t> queue;     // waiting for admission (FIFO)
    std::deque<uint64_t> admitted;  // output allocated, not yet launched
    std::vector<std::unique_ptr<Slot>> slots;
    int injectFailure = 0;  // testing hook: the next launch fails with this code
    uint64_t nextId = 1;
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::thread worker;
    bool stop = false;
    // bumped by every event that can make admission or launch possible
    // (submit, release, a retired launch, stop); the scheduler sleeps until
    // it changes
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

// Called with the lock held: record the first failure and fail every request
// not yet launched, so that their waiters return the error; nothing is
// launched afterwards.
void fail_pending(gcoo_server *s, int code, const char *msg)
{
    if (!s->workerError) {
        s->workerError = code;
        snprintf(s->workerMsg, sizeof(s->workerMsg), "%s", msg);
    }
    for (std::deque<uint64_t> *q : {&s->queue, &s->admitted}) {
        for (uint64_t id : *q) {
            Request &r = s->reqs[id];
            r.state = kFailed;
            r.err = s->workerError;
        }
        q->clear();
    }
    ++s->wakeups;
    s->cvDone.notify_all();
    s->cvWork.notify_all();
}

// Called with the lock held: admit queued requests in order while they fit.
void admit(gcoo_server *s)
{
    while (!s->queue.empty()) {
        Request &r = s->reqs[s->queue.front()];
        size_t off;
        if (!s->arena.alloc(r.bytes, &off)) break;
        r.offset = off;
        r.allocated = true;
        r.state = kAdmitted;
        s->admitted.push_back(r.id);
        s->queue.pop_front();
    }
}

// Called with the lock held: launch admitted requests on a slot with room --
// in batched mode all of them (up to max_batch_jobs) as one persistent
// batch, in plain mode the oldest one through the plain multiply.
int launch(gcoo_server *s, Slot &slot)
{
    Launch L;
    int rc = GCOO_OK;
    if (s->injectFailure) {
        // testing hook: fail the oldest admitted request as a failed launch
        // would, without enqueuing anything
        rc = set_error(s->injectFailure, "injected launch failure");
        s->injectFailure = 0;
        L.reqs.push_back(s->admitted.front());
        s->admitted.pop_front();
    } else if (s->plain) {
        Request &r = s->reqs[s->admitted.front()];
        const Operand &o = s->ops[r.op];
        float *C = reinterpret_cast<float *>(s->arenaBase + r.offset);
        rc = o.dtype == 0
                 ? gcoo_spdm_f32_async(o.h, static_cast<const float *>(o.D), o.h->K, o.N, C,
                                       slot.stream)
                 : gcoo_spdm_bf16_async(o.h, static_cast<const uint16_t *>(o.D), o.h->K, o.N, C,
                                        slot.stream);
        L.reqs.push_back(r.id);
        s->admitted.pop_front();
    } else {
        if ((rc = gcoo_batch_clear(slot.batch))) return rc;
        while (!s->admitted.empty() && (int)L.reqs.size() < s->cfg.max_batch_jobs) {
            Request &r = s->reqs[s->admitted.front()];
            const Operand &o = s->ops[r.op];
            float *C = reinterpret_cast<float *>(s->arenaBase + r.offset);
            L.reqs.push_back(r.id);
            s->admitted.pop_front();
            if ((rc = gcoo_batch_add(slot.batch, o.h, o.D, o.dtype, o.N, C, o.rows, o.vec,
                                     nullptr)))
                break;
        }
        if (!rc) rc = batch_commit_async(slot.batch, slot.stream);
        if (!rc) rc = gcoo_batch_run_async(slot.batch, slot.stream);
    }
    if (!rc) {
        L.event = slot.freeEvents.back();
        if (hipEventRecord(L.event, slot.stream) != hipSuccess) {
            (void)hipGetLastError();
            rc = set_error(GCOO_ERR_HIP, "hipEventRecord failed");
        }
    }
    if (rc) {
        // the launch's requests fail with the error of the call that failed
        for (uint64_t id : L.reqs) {
            Request &r = s->reqs[id];
            r.state = kFailed;
            r.err = rc;
        }
        return rc;
    }
    slot.freeEvents.pop_back();
    for (uint64_t id : L.reqs) s->reqs[id].state = kRunning;
    ++s->batches;
    s->jobsRun += L.reqs.size();
    slot.inflight.push_back(std::move(L));
    slot.cv.notify_one();
    return GCOO_OK;
}

void worker_main(gcoo_server *s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        const uint64_t seen = s->wakeups;
        if (!s->workerError) {
            admit(s);
            for (auto &sp : s->slots) {
                Slot &slot = *sp;
                while (!s->workerError && !s->admitted.empty() &&
                       (int)slot.inflight.size() < s->depth) {
                    const int rc = launch(s, slot);
                    if (rc) {
                        char msg[256];
                        snprintf(msg, sizeof(msg), "launch failed: %s", gcoo_last_error());
                        fail_pending(s, rc, msg);
                    }
                }
            }
        }
        if (s->stop) break;
        s->cvWork.wait(lk, [&] { return s->stop || s->wakeups != seen; });
    }
}

// Blocks on the oldest launch's event of its slot (without the lock), then
// retires it: timestamps its requests, or fails them when the launch failed.
void waiter_main(gcoo_server *s, Slot *slot)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        slot->cv.wait(lk, [&] { return s->stop || !slot->inflight.empty(); });
        if (slot->inflight.empty()) break;  // stopping, nothing left in flight
        const hipEvent_t ev = slot->inflight.front().event;
        lk.unlock();
        const hipError_t e = hipEventSynchronize(ev);
        const auto now = Clock::now();
        lk.lock();
        Launch L = std::move(slot->inflight.front());
        slot->inflight.pop_front();
        slot->freeEvents.push_back(L.event);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            char msg[256];
            snprintf(msg, sizeof(msg), "batch failed: %s", hipGetErrorString(e));
            for (uint64_t id : L.reqs) {
                Request &r = s->reqs[id];
                r.state = kFailed;
                r.err = GCOO_ERR_HIP;
            }
            fail_pending(s, GCOO_ERR_HIP, msg);
        } else {
            for (uint64_t id : L.reqs) {
                Request &r = s->reqs[id];
                r.done = now;
                r.state = kDone;
                s->latMs.push_back(
                    std::chrono::duration<double, std::milli>(now - r.arrival).count());
            }
            s->lastDone = now;
        }
        ++s->wakeups;
        s->cvWork.notify_one();
        s->cvDone.notify_all();
    }
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
    // plain FIFO: one stream, requests launched one by one in arrival order
    s->plain = cfg->max_batch_jobs == 1;
    s->depth = s->plain ? kPlainDepth : 1;
    const int nSlots = s->plain ? 1 : cfg->slots;
    (void)hipGetDevice(&s->device);
    if (hipMalloc(&s->arenaBase, cfg->hbm_budget_bytes) != hipSuccess) {
        (void)hipGetLastError();
        delete s;
        return set_error(GCOO_ERR_ALLOC, "cannot allocate the %zu-byte output arena",
                         cfg->hbm_budget_bytes);
    }
    s->arena.reset(cfg->hbm_budget_bytes);
    for (int i = 0; i < nSlots; ++i) {
        s->slots.emplace_back(new (std::nothrow) Slot);
        Slot *slot = s->slots.back().get();
        bool ok = slot != nullptr &&
                  hipStreamCreateWithFlags(&slot->stream, hipStreamNonBlocking) == hipSuccess &&
                  (s->plain || (gcoo_batch_create(&slot->batch) == GCOO_OK &&
                                gcoo_batch_set_panel_mode(slot->batch, 0, 0.0, 0.0) == GCOO_OK));
        for (int e = 0; ok && e < s->depth; ++e) {
            hipEvent_t ev = nullptr;
            ok = hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
            if (ok) slot->freeEvents.push_back(ev);
        }
        if (!ok) {
            (void)hipGetLastError();
            gcoo_server_destroy(s);
            return set_error(GCOO_ERR_HIP, "cannot create server streams");
        }
    }
    for (auto &slot : s->slots) slot->waiter = std::thread(waiter_main, s, slot.get());
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
        ++s->wakeups;
    }
    s->cvWork.notify_all();
    if (s->worker.joinable()) s->worker.join();
    // the waiters retire what is still in flight, then exit
    for (auto &slot : s->slots) {
        if (!slot) continue;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            slot->cv.notify_all();
        }
        if (slot->waiter.joinable()) slot->waiter.join();
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->cvDone.notify_all();
    }
    for (auto &slot : s->slots) {
        if (!slot) continue;
        if (slot->stream) (void)hipStreamSynchronize(slot->stream);
        if (slot->batch) gcoo_batch_destroy(slot->batch);
        for (hipEvent_t ev : slot->freeEvents) (void)hipEventDestroy(ev);
        if (slot->stream) (void)hipStreamDestroy(slot->stream);
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
    // the request is looked up again after every wake-up: submits from other
    // threads may rehash the map while this one waits
    auto state = [&]() -> int {
        auto it = s->reqs.find(req_id);
        return it == s->reqs.end() ? -1 : it->second.state;
    };
    if (state() < 0)
        return set_error(GCOO_ERR_INVALID, "unknown request %llu", (unsigned long long)req_id);
    s->cvDone.wait(lk, [&] {
        const int st = state();
        return st < 0 || st >= kDone || s->stop;
    });
    auto it = s->reqs.find(req_id);
    if (it == s->reqs.end())
        return set_error(GCOO_ERR_STATE, "request %llu was released while waiting",
                         (unsigned long long)req_id);
    if (it->second.state == kFailed)
        return set_error(it->second.err, "request %llu failed: %s", (unsigned long long)req_id,
                         s->workerMsg);
    if (it->second.state != kDone) return set_error(GCOO_ERR_STATE, "server stopped");
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
    if (it == s->reqs.end()) return set_error(GCOO_ERR_STATE, "unknown request %llu",
                                              (unsigned long long)req_id);
    if (it->second.state == kFailed)
        return set_error(it->second.err, "request %llu failed: %s", (unsigned long long)req_id,
                         s->workerMsg);
    if (it->second.state != kDone)
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
        if (it == s->reqs.end() ||
            (it->second.state != kDone && it->second.state != kFailed))
            return set_error(GCOO_ERR_STATE, "request %llu not completed",
                             (unsigned long long)req_id);
        if (it->second.allocated) s->arena.release(it->second.offset, it->second.bytes);
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

int gcoo_test_server_inject_failure(gcoo_server_t s, int code)
{
    if (!s) return set_error(GCOO_ERR_INVALID, "null server");
    if (code <= 0) return set_error(GCOO_ERR_INVALID, "code %d (a GCOO_ERR_* value)", code);
    std::lock_guard<std::mutex> lk(s->mu);
    s->injectFailure = code;
    return GCOO_OK;
}

}  // extern "C"
