// Request server: continuous batching of forward / inverse transforms of mixed
// bandwidth on one device, on the batched mixed-precision pipeline.
//
// Requests name caller-owned device buffers (complex64 grids of (2B)^3 values,
// packed complex64 coefficient sets).  A dispatcher thread keeps up to `slots`
// batches in flight, one per slot, each slot with its own stream, pointer
// tables and work buffers; batches on different slots overlap on the device.
// Whenever a slot is free it takes the oldest waiting request, gathers the
// waiting requests of the same (direction, bandwidth) in arrival order up to
// the slot's capacity for that bandwidth, and launches them as one batch.
// Requests that arrive while others run are admitted as slots free up.
//
// Launch: the first batch of each (direction, bandwidth, size) on a slot is
// captured into a HIP graph (pointer-table upload + pipeline kernels), later
// ones replay it -- one launch instead of eight, which is what bounds the
// dispatcher on small-bandwidth batches.  Slot streams do not block on the
// null stream; each batch instead waits on an event recorded on the null
// stream at launch, so inputs written there before submission are ordered
// before the batch.  The capture holds no memset nodes.
//
// HBM admission: the server's own device memory -- the resident FP32 tables
// of every served bandwidth plus the slots' work buffers -- stays within the
// budget given at creation; a slot holds work buffers of budget_work / slots
// bytes, which bounds the batch size per bandwidth (two complex64 grids per
// signal), and a bandwidth whose single signal does not fit is refused at
// creation.  Results do not depend on which requests share a batch.
#include "pfs_internal.h"
#include "pfsoft_amd.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

using namespace pfs;

namespace {

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
    int kind = 0, B = 0;
    void* in = nullptr;
    void* out = nullptr;
    double t_submit = 0, t_done = 0;
    int status = -1;  // -1 waiting / running, else a pfs_status
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    hipEvent_t host_order = nullptr;  // recorded on the null stream before each batch
    std::unordered_map<uint64_t, hipGraphExec_t> graphs;  // (kind, B, n) -> replayable batch
    float2* work[2] = {nullptr, nullptr};
    float2** host_ptrs = nullptr;  // pinned: 2 * max_batch device addresses
    float2** dev_ptrs = nullptr;
    std::vector<uint64_t> reqs;
    bool busy = false;
};

}  // namespace

struct pfs_server {
    pfs_context* ctx = nullptr;
    int device = 0;
    int max_batch = 0;
    bool use_graphs = true;
    size_t slot_bytes = 0;
    size_t device_bytes = 0;  // tables + work buffers
    std::vector<int> bandwidths;
    std::vector<Slot> slots;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> pending;
    std::unordered_map<uint64_t, Request> reqs;
    std::vector<uint64_t> finished;  // completed since the last collect
    uint64_t next_id = 1;
    bool stop = false;
    std::thread dispatcher;
    std::string error;

    // statistics
    uint64_t n_batches = 0, n_done = 0, n_signals = 0, n_graphs = 0;
    int max_inflight = 0;
};

namespace {

int cap_for(const pfs_server* s, int B)
{
    const size_t per = 2 * batch_work_bytes(B, 1);
    return (int)std::min<size_t>((size_t)s->max_batch, s->slot_bytes / per);
}

bool served(const pfs_server* s, int B)
{
    return std::find(s->bandwidths.begin(), s->bandwidths.end(), B) != s->bandwidths.end();
}

// Called with s->mu held.  Marks the requests of every finished slot done.
void retire(pfs_server* s)
{
    for (Slot& sl : s->slots) {
        if (!sl.busy) continue;
        hipError_t q = hipEventQuery(sl.done);
        if (q == hipErrorNotReady) continue;
        const double t = now_s();
        const int st = q == hipSuccess ? PFS_OK : PFS_ERR_DEVICE;
        if (q != hipSuccess) s->error = std::string("batch failed on the device: ") + hipGetErrorString(q);
        for (uint64_t id : sl.reqs) {
            Request& r = s->reqs[id];
            r.t_done = t;
            r.status = st;
            s->finished.push_back(id);
            ++s->n_done;
        }
        sl.reqs.clear();
        sl.busy = false;
        s->cv_done.notify_all();
    }
}

// Called with s->mu held.  Launches one batch on slot sl; false if nothing waits.
bool launch(pfs_server* s, Slot& sl)
{
    if (s->pending.empty()) return false;
    const Request& head = s->reqs[s->pending.front()];
    const int kind = head.kind, B = head.B;
    const int cap = cap_for(s, B);
    std::vector<uint64_t> take;
    for (auto it = s->pending.begin(); it != s->pending.end() && (int)take.size() < cap;) {
        const Request& r = s->reqs[*it];
        if (r.kind == kind && r.B == B) {
            take.push_back(*it);
            it = s->pending.erase(it);
        } else {
            ++it;
        }
    }
    const int n = (int)take.size();
    for (int i = 0; i < n; ++i) {
        const Request& r = s->reqs[take[i]];
        sl.host_ptrs[i] = (float2*)r.in;
        sl.host_ptrs[s->max_batch + i] = (float2*)r.out;
    }
    int rc = PFS_OK;
    hipError_t e = hipSuccess;
    auto enqueue = [&]() -> int {
        e = hipMemcpyAsync(sl.dev_ptrs, sl.host_ptrs, 2 * s->max_batch * sizeof(float2*), hipMemcpyHostToDevice,
                           sl.stream);
        if (e != hipSuccess) {
            s->error = std::string("pointer table upload failed: ") + hipGetErrorString(e);
            return PFS_ERR_DEVICE;
        }
        const SigSet in{nullptr, 0, sl.dev_ptrs}, out{nullptr, 0, sl.dev_ptrs + s->max_batch};
        const int r = batch_run(s->ctx, kind == PFS_REQ_FORWARD, B, n, cap, in, out, sl.work[0], sl.work[1], sl.stream);
        if (r) s->error = error_text();
        return r;
    };
    // order the batch after the null-stream work enqueued before it (inputs)
    if ((e = hipEventRecord(sl.host_order, nullptr)) != hipSuccess ||
        (e = hipStreamWaitEvent(sl.stream, sl.host_order, 0)) != hipSuccess) {
        rc = PFS_ERR_DEVICE;
        s->error = std::string("ordering after the null stream failed: ") + hipGetErrorString(e);
    } else if (s->use_graphs) {
        const uint64_t key = ((uint64_t)kind << 48) | ((uint64_t)B << 24) | (uint64_t)n;
        auto it = sl.graphs.find(key);
        if (it == sl.graphs.end()) {
            hipGraph_t g = nullptr;
            hipGraphExec_t x = nullptr;
            if ((e = hipStreamBeginCapture(sl.stream, hipStreamCaptureModeThreadLocal)) != hipSuccess) {
                rc = PFS_ERR_DEVICE;
            } else {
                rc = enqueue();
                const hipError_t e2 = hipStreamEndCapture(sl.stream, &g);
                if (!rc && e2 != hipSuccess) {
                    e = e2;
                    rc = PFS_ERR_DEVICE;
                }
                if (!rc && (e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0)) != hipSuccess) rc = PFS_ERR_DEVICE;
                if (g) (void)hipGraphDestroy(g);
            }
            if (rc == PFS_ERR_DEVICE && s->error.empty())
                s->error = std::string("graph capture failed: ") + hipGetErrorString(e);
            if (!rc) {
                it = sl.graphs.emplace(key, x).first;
                ++s->n_graphs;
            }
        }
        if (!rc && (e = hipGraphLaunch(it->second, sl.stream)) != hipSuccess) {
            rc = PFS_ERR_DEVICE;
            s->error = std::string("graph launch failed: ") + hipGetErrorString(e);
        }
    } else {
        rc = enqueue();
    }
    if (rc == PFS_OK && (e = hipEventRecord(sl.done, sl.stream)) != hipSuccess) {
        rc = PFS_ERR_DEVICE;
        s->error = std::string("event record failed: ") + hipGetErrorString(e);
    }
    if (rc) {
        // nothing of this batch is pending on the device any more once the stream drains
        (void)hipStreamSynchronize(sl.stream);
        const double t = now_s();
        for (uint64_t id : take) {
            Request& r = s->reqs[id];
            r.t_done = t;
            r.status = rc;
            s->finished.push_back(id);
            ++s->n_done;
        }
        s->cv_done.notify_all();
        return true;
    }
    sl.reqs = std::move(take);
    sl.busy = true;
    ++s->n_batches;
    s->n_signals += n;
    int inflight = 0;
    for (const Slot& x : s->slots) inflight += x.busy;
    s->max_inflight = std::max(s->max_inflight, inflight);
    return true;
}

void run(pfs_server* s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        retire(s);
        bool busy = false;
        for (Slot& sl : s->slots) {
            if (!sl.busy) launch(s, sl);
            busy |= sl.busy;
        }
        if (s->stop && !busy && s->pending.empty()) break;
        if (busy) {
            // a batch is on the device: poll its completion at a fine interval
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            lk.lock();
        } else if (s->pending.empty()) {
            s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty(); });
        }
    }
}

void release(pfs_server* s)
{
    for (Slot& sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        for (float2* w : sl.work)
            if (w) (void)pfs_device_free(s->ctx, w);
        if (sl.dev_ptrs) (void)pfs_device_free(s->ctx, sl.dev_ptrs);
        if (sl.host_ptrs) (void)hipHostFree(sl.host_ptrs);
        for (auto& kv : sl.graphs) (void)hipGraphExecDestroy(kv.second);
        sl.graphs.clear();
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.host_order) (void)hipEventDestroy(sl.host_order);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    s->slots.clear();
}

}  // namespace

extern "C" {

int pfs_server_create(pfs_context* ctx, const int* bandwidths, int nbw, int slots, int max_batch,
                      size_t hbm_budget_bytes, pfs_server** out)
{
    if (!ctx || !bandwidths || nbw < 1 || slots < 1 || max_batch < 1 || !out)
        return set_error(PFS_ERR_ARG, "pfs_server_create: bad argument");
    *out = nullptr;
    pfs_server* s = new pfs_server();
    s->ctx = ctx;
    s->device = context_device(ctx);
    s->max_batch = max_batch;
    long graphs = 1;
    (void)pfs_tune_get(ctx, "server_graphs", &graphs);
    s->use_graphs = graphs != 0;
    int rc = PFS_OK;
    if (hipSetDevice(s->device) != hipSuccess) rc = set_error(PFS_ERR_DEVICE, "pfs_server_create: hipSetDevice failed");
    // resident tables of every served bandwidth
    size_t tables = 0;
    for (int i = 0; i < nbw && !rc; ++i) {
        const int B = bandwidths[i];
        if (served(s, B)) continue;
        s->bandwidths.push_back(B);
        rc = batch_warm(ctx, B);
        const size_t N = 2 * (size_t)B;
        tables += 2 * (size_t)batch_table_rows(B) * N * sizeof(float);
    }
    int maxB = 0;
    for (int B : s->bandwidths) maxB = std::max(maxB, B);
    if (!rc) {
        // work buffers: what the budget leaves after the tables, split over the slots,
        // and no more than a full batch of the largest bandwidth needs
        const size_t want = 2 * batch_work_bytes(maxB, max_batch);
        size_t per_slot = want;
        if (hbm_budget_bytes) {
            if (hbm_budget_bytes <= tables)
                rc = set_error(PFS_ERR_BUDGET, "pfs_server_create: the budget does not cover the resident tables");
            else
                per_slot = std::min(want, (hbm_budget_bytes - tables) / slots);
        }
        s->slot_bytes = per_slot;
        for (int B : s->bandwidths)
            if (!rc && cap_for(s, B) < 1) {
                char msg[160];
                snprintf(msg, sizeof(msg), "pfs_server_create: the budget leaves no room for one B = %d signal per slot", B);
                rc = set_error(PFS_ERR_BUDGET, msg);
            }
        s->device_bytes = tables + per_slot * slots;
    }
    long autotune = 1;
    (void)pfs_tune_get(ctx, "server_autotune", &autotune);
    for (size_t i = 0; i < s->bandwidths.size() && !rc && autotune; ++i) {
        const int B = s->bandwidths[i], cap = cap_for(s, B);
        if (!batch_is_tuned(ctx, B, cap)) rc = pfs_batch_autotune(ctx, B, cap);
    }
    for (int i = 0; i < slots && !rc; ++i) {
        s->slots.emplace_back();
        Slot& sl = s->slots.back();
        const size_t half = (s->slot_bytes / 2) & ~(size_t)255;
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&sl.host_order, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void**)&sl.host_ptrs, 2 * max_batch * sizeof(float2*), hipHostMallocDefault) != hipSuccess) {
            rc = set_error(PFS_ERR_DEVICE, "pfs_server_create: stream, event or pinned buffer creation failed");
            break;
        }
        rc = pfs_device_alloc(ctx, 2 * max_batch * sizeof(float2*), (void**)&sl.dev_ptrs);
        if (!rc) rc = pfs_device_alloc(ctx, half, (void**)&sl.work[0]);
        if (!rc) rc = pfs_device_alloc(ctx, half, (void**)&sl.work[1]);
    }
    if (rc) {
        release(s);
        delete s;
        return rc;
    }
    s->dispatcher = std::thread(run, s);
    *out = s;
    return PFS_OK;
}

int pfs_server_destroy(pfs_server* s)
{
    if (!s) return PFS_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->dispatcher.joinable()) s->dispatcher.join();
    release(s);
    delete s;
    return PFS_OK;
}

int pfs_server_submit(pfs_server* s, int kind, int B, const void* in_dev, void* out_dev, uint64_t* ticket)
{
    if (!s || !in_dev || !out_dev || !ticket) return set_error(PFS_ERR_ARG, "pfs_server_submit: null argument");
    if (kind != PFS_REQ_FORWARD && kind != PFS_REQ_INVERSE)
        return set_error(PFS_ERR_ARG, "pfs_server_submit: kind must be PFS_REQ_FORWARD or PFS_REQ_INVERSE");
    if (!served(s, B)) return set_error(PFS_ERR_UNSUPPORTED, "pfs_server_submit: bandwidth not served by this server");
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) return set_error(PFS_ERR_ARG, "pfs_server_submit: server is shutting down");
        const uint64_t id = s->next_id++;
        Request& r = s->reqs[id];
        r.kind = kind;
        r.B = B;
        r.in = const_cast<void*>(in_dev);
        r.out = out_dev;
        r.t_submit = now_s();
        s->pending.push_back(id);
        *ticket = id;
    }
    s->cv_work.notify_one();
    return PFS_OK;
}

int pfs_server_wait(pfs_server* s, uint64_t ticket, double timeout_s, double* latency_s)
{
    if (!s) return set_error(PFS_ERR_ARG, "pfs_server_wait: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return set_error(PFS_ERR_ARG, "pfs_server_wait: unknown or already collected ticket");
    const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s < 0 ? 1e9 : timeout_s);
    if (!s->cv_done.wait_until(lk, until, [&] { return s->reqs[ticket].status >= 0; }))
        return set_error(PFS_ERR_TIMEOUT, "pfs_server_wait: timed out; the request is still queued or running");
    const Request& r = s->reqs[ticket];
    if (latency_s) *latency_s = r.t_done - r.t_submit;
    const int st = r.status;
    // the ticket is consumed: forget the request and its entry in the list
    // pfs_server_collect would report
    s->reqs.erase(ticket);
    auto f = std::find(s->finished.begin(), s->finished.end(), ticket);
    if (f != s->finished.end()) s->finished.erase(f);
    if (st) {
        const std::string msg = "request failed: " + s->error;
        lk.unlock();
        return set_error(st, msg.c_str());
    }
    return PFS_OK;
}

int pfs_server_collect(pfs_server* s, uint64_t* tickets, double* latency_s, int* status, int cap, int* n_out)
{
    if (!s || !n_out || cap < 0 || (cap > 0 && (!tickets || !latency_s || !status)))
        return set_error(PFS_ERR_ARG, "pfs_server_collect: bad argument");
    std::lock_guard<std::mutex> g(s->mu);
    const int n = std::min<int>(cap, (int)s->finished.size());
    for (int i = 0; i < n; ++i) {
        const uint64_t id = s->finished[i];
        const Request& r = s->reqs[id];
        tickets[i] = id;
        latency_s[i] = r.t_done - r.t_submit;
        status[i] = r.status;
        s->reqs.erase(id);
    }
    s->finished.erase(s->finished.begin(), s->finished.begin() + n);
    *n_out = n;
    return PFS_OK;
}

int pfs_server_drain(pfs_server* s, double timeout_s)
{
    if (!s) return set_error(PFS_ERR_ARG, "pfs_server_drain: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s < 0 ? 1e9 : timeout_s);
    auto idle = [&] {
        if (!s->pending.empty()) return false;
        for (const Slot& sl : s->slots)
            if (sl.busy) return false;
        return true;
    };
    if (!s->cv_done.wait_until(lk, until, idle)) return set_error(PFS_ERR_TIMEOUT, "pfs_server_drain: timed out");
    return PFS_OK;
}

int pfs_server_stats(pfs_server* s, pfs_server_stats_t* out)
{
    if (!s || !out) return set_error(PFS_ERR_ARG, "pfs_server_stats: null argument");
    std::lock_guard<std::mutex> g(s->mu);
    out->completed = s->n_done;
    out->batches = s->n_batches;
    out->signals = s->n_signals;
    out->pending = s->pending.size();
    out->max_inflight_batches = s->max_inflight;
    out->device_bytes = s->device_bytes;
    out->slot_bytes = s->slot_bytes;
    out->graphs = s->n_graphs;
    out->tracked = s->reqs.size();
    return PFS_OK;
}

int pfs_server_capacity(pfs_server* s, int B, int* signals_per_batch)
{
    if (!s || !signals_per_batch) return set_error(PFS_ERR_ARG, "pfs_server_capacity: null argument");
    if (!served(s, B)) return set_error(PFS_ERR_UNSUPPORTED, "pfs_server_capacity: bandwidth not served");
    *signals_per_batch = cap_for(s, B);
    return PFS_OK;
}

}  // extern "C"
