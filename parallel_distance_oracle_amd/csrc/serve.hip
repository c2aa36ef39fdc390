// serve.hip -- continuous-batching query service over a resident index.
//
// query_distance (includes/pado_unw_unv_para.201912311607.parallel_bp_labeling.h:
// 2283-2416) answers one pair per call.  Here requests with any number of
// pairs are admitted into a pinned staging ring (admission = a free request
// slot + ring space, i.e. device staging memory), a dispatcher thread packs the
// requests queued at that moment into one launch per stream (up to `streams`
// launches in flight) of a staging kernel (the launch's pairs read from the
// mapped ring into device memory) and the label-merge kernel (one wave per
// pair, grid-stride; a persistent work-counter variant on request), and a
// completer thread marks requests done in launch order.  Weighted indexes are
// served the same way with u32 answers.  A request that arrives while launches run
// joins the next launch; its pairs overlap on the device with the running
// ones.  Every answer depends only on its own pair, so a request's result is
// the same alone and in any mix.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "internal.hpp"

namespace pado {
pado_gpu_status launch_query_any(const pado_gpu_index *idx, bool rank_space, const uint32_t *qa, const uint32_t *qb,
                                 uint64_t k, void *out, bool ref_compat, hipStream_t st,
                                 unsigned long long *work, uint32_t blocks);
uint32_t query_ws_blocks_per_cu(bool weighted);
}

using namespace pado;

namespace {

// Copies a launch's k pairs from the mapped host ring into device memory.  The
// runtime's host-to-device copy occasionally spent several milliseconds of
// host time inside the call (about once in 10^4 launches), holding back the
// launch behind it; a kernel goes through the stream's queue like every other
// launch.  Each thread reads four ids of each array before it writes, so
// that a wave keeps eight reads over the host link in flight.
constexpr uint32_t kStageUnroll = 4;

__global__ __launch_bounds__(256) void k_stage_pairs(const uint32_t *__restrict__ ha,
                                                     const uint32_t *__restrict__ hb, uint32_t *__restrict__ da,
                                                     uint32_t *__restrict__ db, uint64_t k)
{
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x * kStageUnroll;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x * kStageUnroll + threadIdx.x; base < k; base += step) {
        uint32_t va[kStageUnroll], vb[kStageUnroll];
#pragma unroll
        for (uint32_t j = 0; j < kStageUnroll; ++j) {
            const uint64_t i = base + (uint64_t)j * blockDim.x;
            if (i < k) {
                va[j] = ha[i];
                vb[j] = hb[i];
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < kStageUnroll; ++j) {
            const uint64_t i = base + (uint64_t)j * blockDim.x;
            if (i < k) {
                da[i] = va[j];
                db[i] = vb[j];
            }
        }
    }
}

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
    uint64_t id = 0;
    uint64_t start = 0;  // monotonic ring position of the first pair
    uint64_t k = 0;
    // 0 queued, 1 launched, 2 done, 3 consumed; written under the server
    // lock, read without it by a waiter that spins before it blocks.
    std::atomic<int> state{0};
    bool claimed = false;  // a waiter is copying the answers out
    bool failed = false;   // its launch failed: no answers
    bool cancelled = false;  // answers are not wanted: released once no launch reads its slot
    uint32_t slot = 0;       // resident mode: completion-flag slot
    double t_submit = 0, t_done = 0;
};

struct Launch {
    std::vector<Request *> reqs;
    int stream = 0;
    bool failed = false;        // a launch call returned an error
    bool end_recorded = false;  // the completion event is on the stream
};

}  // namespace

struct pado_gpu_server {
    pado_gpu_ctx *ctx = nullptr;
    pado_gpu_index *idx = nullptr;
    pado_gpu_serve_cfg cfg{};
    uint64_t cap = 0;               // ring capacity in pairs
    uint64_t head = 0, tail = 0;    // monotonic ring positions (released / reserved)
    uint32_t es = 1;                // answer bytes: 1 (u8 hops) or 4 (u32 weighted)
    uint32_t ws_blocks = 0;         // persistent kernel grid (workgroups)
    uint32_t *h_a = nullptr, *h_b = nullptr;
    const uint32_t *m_a = nullptr, *m_b = nullptr;  // the same ring as mapped for the device
    uint8_t *h_out = nullptr;
    DevBuf<uint32_t> d_a, d_b;
    DevBuf<uint8_t> d_out;
    DevBuf<unsigned long long> work;  // one work counter per stream
    std::vector<hipStream_t> streams;
    std::vector<int> stream_busy;
    std::vector<hipEvent_t> ev_start, ev_end;  // per stream, reused launch after launch
    uint64_t pending_pairs = 0;     // pairs admitted but not launched
    double max_wait_ms = 0.2;       // coalescing window while launches are in flight
    uint64_t min_launch = 0;        // pairs that start a launch at once
    std::deque<Request *> live;     // ring order, not yet released
    std::deque<Request *> pending;  // admitted, not launched
    std::unordered_map<uint64_t, Request *> by_id;
    std::deque<Launch> inflight;    // issue order
    uint64_t next_id = 1;
    bool stopping = false;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_launch;
    std::condition_variable cv_space;  // a request released its slot and ring space
    std::thread dispatcher, completer;
    pado_gpu_serve_stats st{};
    std::string error;
    // Resident mode (PADO_GPU_SERVE_RESIDENT): no dispatcher; submit
    // publishes descriptors to the resident kernel, the completer polls the
    // completion flags and relaunches the kernel after it exits.
    bool resident = false;
    ServeCtl *h_ctl = nullptr;
    ServeDesc *h_desc = nullptr;
    uint64_t *h_flag = nullptr;
    uint32_t dcap = 0;                       // descriptors per queue ring
    uint64_t qpub[kServeQueues] = {};        // descriptors published per queue
    uint32_t qnext = 0;                      // next queue (round robin over descriptors)
    std::vector<uint32_t> free_slots;
    std::deque<Request *> rinflight;         // published, not done
    DevBuf<unsigned long long> rstate;
    DevBuf<uint32_t> rdone;
    ServeKernelArgs rargs{};
    uint32_t rblocks = 0;
    hipStream_t rstream = nullptr;
    hipEvent_t rexit = nullptr;
    bool rlive = false;                      // a resident launch may still run
};

namespace {

// Pairs of [start, start+k) must not wrap: a request that would cross the
// physical end starts at the next lap.
uint64_t place(uint64_t tail, uint64_t k, uint64_t cap)
{
    uint64_t phys = tail % cap;
    if (phys + k > cap) return tail + (cap - phys);
    return tail;
}

// With nothing live the ring is empty: the next request starts at physical
// offset 0 (the next lap's start), so any request of up to `cap` pairs fits.
// Restarting at the current physical offset would refuse a request larger
// than that offset forever (it would be placed on the next lap, past head +
// cap, and nothing live could move head).
void rewind_if_empty(pado_gpu_server *s)
{
    if (!s->live.empty()) return;
    const uint64_t lap = (s->tail + s->cap - 1) / s->cap * s->cap;
    s->head = s->tail = lap;
}

void release_consumed(pado_gpu_server *s)
{
    while (!s->live.empty() && s->live.front()->state == 3) {
        Request *r = s->live.front();
        s->live.pop_front();
        s->by_id.erase(r->id);
        s->head = r->start + r->k;
        delete r;
    }
    rewind_if_empty(s);
}

// Continuous batching policy: with the device idle a launch starts at once
// (lowest latency); while launches are in flight, pending requests are held
// until they fill min_launch pairs or the oldest has waited max_wait_ms, so
// that launches stay large enough to amortise their fixed cost.
void dispatcher_loop(pado_gpu_server *s)
{
    (void)hipSetDevice(s->ctx->device);
    const uint64_t max_launch = s->cfg.max_launch_pairs;
    for (;;) {
        std::unique_lock<std::mutex> lk(s->mu);
        for (;;) {
            if (s->pending.empty()) {
                if (s->stopping) break;
                s->cv_work.wait(lk);
                continue;
            }
            const size_t busy = (size_t)std::count(s->stream_busy.begin(), s->stream_busy.end(), 1);
            if (busy == s->stream_busy.size()) {
                s->cv_work.wait(lk);
                continue;
            }
            if (busy == 0 || s->stopping || s->pending_pairs >= s->min_launch) break;
            const double age = now_ms() - s->pending.front()->t_submit;
            if (age >= s->max_wait_ms) break;
            s->cv_work.wait_for(lk, std::chrono::microseconds((int64_t)((s->max_wait_ms - age) * 1000.0) + 1));
        }
        if (s->stopping && s->pending.empty()) break;
        int si = (int)(std::find(s->stream_busy.begin(), s->stream_busy.end(), 0) - s->stream_busy.begin());
        // Pack consecutive pending requests that are physically contiguous.  A
        // request that ends exactly at the ring's end is followed, at the same
        // monotonic position, by one at physical offset 0 of the next lap:
        // the launch stops at the physical end, not only at a placement gap.
        Launch L;
        L.stream = si;
        uint64_t p0 = s->pending.front()->start, p1 = p0;
        while (!s->pending.empty()) {
            Request *r = s->pending.front();
            if (r->start != p1 && !L.reqs.empty()) break;              // placement gap
            if (!L.reqs.empty() && (r->start + r->k - p0) > max_launch) break;
            if ((p0 % s->cap) + (r->start + r->k - p0) > s->cap) break;  // crosses the physical end
            L.reqs.push_back(r);
            r->state = 1;
            p1 = r->start + r->k;
            s->pending_pairs -= r->k;
            s->pending.pop_front();
        }
        s->stream_busy[si] = 1;
        lk.unlock();

        const uint64_t phys = p0 % s->cap, k = p1 - p0;
        hipStream_t st = s->streams[si];
        const bool in_ring = phys + k <= s->cap;  // holds by the packing above; never launch past the ring
        std::string why;
        if (!in_ring) why = "serving launch crosses the staging ring's end";
        if (hipEventRecord(s->ev_start[si], st) != hipSuccess) why = "event record failed";
        if (k && in_ring) {
            const uint32_t sgrid = (uint32_t)std::min<uint64_t>((k + 1023) / 1024, 1024);
            hipLaunchKernelGGL(k_stage_pairs, dim3(sgrid), dim3(256), 0, st, s->m_a + phys, s->m_b + phys,
                               s->d_a.p + phys, s->d_b.p + phys, k);
            const bool grid_stride = (s->cfg.flags & PADO_GPU_SERVE_PERSISTENT) == 0;
            pado_gpu_status qs = launch_query_any(s->idx, false, s->d_a.p + phys, s->d_b.p + phys, k,
                                                  s->d_out.p + phys * s->es,
                                                  (s->cfg.flags & PADO_GPU_QUERY_REF_COMPAT) != 0, st,
                                                  grid_stride ? nullptr : s->work.p + si, s->ws_blocks);
            if (qs != PADO_GPU_OK) why = pado_gpu_last_error();
            if (hipMemcpyAsync(s->h_out + phys * s->es, s->d_out.p + phys * s->es, k * s->es,
                               hipMemcpyDeviceToHost, st) != hipSuccess)
                why = "answer copy failed";
        }
        // The end event is recorded whatever failed, so the completer retires
        // this launch (and fails its requests) in order.
        const hipError_t rec = hipEventRecord(s->ev_end[si], st);
        if (rec != hipSuccess) why = std::string("event record failed: ") + hipGetErrorString(rec);
        if (!why.empty()) {
            L.failed = true;
            std::lock_guard<std::mutex> g(s->mu);
            if (s->error.empty()) s->error = why;
        }
        L.end_recorded = rec == hipSuccess;
        lk.lock();
        s->inflight.push_back(std::move(L));
        s->st.launches += 1;
        lk.unlock();
        s->cv_launch.notify_all();
    }
    s->cv_launch.notify_all();
}

void completer_loop(pado_gpu_server *s)
{
    (void)hipSetDevice(s->ctx->device);
    for (;;) {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_launch.wait(lk, [&] { return !s->inflight.empty() || (s->stopping && s->pending.empty()); });
        if (s->inflight.empty()) {
            if (s->stopping && s->pending.empty()) {
                // The dispatcher may still be issuing its last launch.
                bool any_busy = std::find(s->stream_busy.begin(), s->stream_busy.end(), 1) != s->stream_busy.end();
                if (!any_busy) break;
                s->cv_launch.wait_for(lk, std::chrono::milliseconds(1));
            }
            continue;
        }
        const int si = s->inflight.front().stream;
        const bool recorded = s->inflight.front().end_recorded;
        hipEvent_t ev0 = s->ev_start[si], ev1 = s->ev_end[si];
        lk.unlock();
        // Poll rather than block: a blocking event wait can resume
        // milliseconds late, which would land in the latency tail.  Any
        // result other than not-ready / success fails the launch's requests;
        // without a recorded end event the stream is drained instead, so no
        // copy of this launch is still pending when its ring space is reused.
        hipError_t q = hipSuccess;
        if (recorded) {
            while ((q = hipEventQuery(ev1)) == hipErrorNotReady) std::this_thread::yield();
        } else {
            q = hipStreamSynchronize(s->streams[si]);
        }
        float ms = 0.f;
        if (q == hipSuccess && recorded) (void)hipEventElapsedTime(&ms, ev0, ev1);
        const double t = now_ms();
        lk.lock();
        Launch L = std::move(s->inflight.front());
        s->inflight.pop_front();
        if (q != hipSuccess) {
            L.failed = true;
            if (s->error.empty()) s->error = std::string("serving launch failed: ") + hipGetErrorString(q);
        }
        bool release = false;
        for (Request *r : L.reqs) {
            r->state = 2;
            r->failed = L.failed;
            r->t_done = t;
            s->st.requests_done += 1;
            s->st.pairs_done += r->k;
            if (r->cancelled) {
                r->state = 3;  // nobody will wait for it
                release = true;
            }
        }
        if (release) {
            release_consumed(s);
            s->cv_space.notify_all();
        }
        s->st.busy_ms += ms;
        s->stream_busy[L.stream] = 0;
        lk.unlock();
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

// ---- resident mode ---------------------------------------------------------

// Enqueues a resident launch unless one may still run (mu held).
void resident_launch(pado_gpu_server *s)
{
    if (s->rlive || !s->error.empty()) return;
    pado_gpu_status e = launch_serve_resident(s->idx, s->rargs, s->rblocks, s->rstream);
    hipError_t r = e == PADO_GPU_OK ? hipEventRecord(s->rexit, s->rstream) : hipErrorLaunchFailure;
    if (e != PADO_GPU_OK || r != hipSuccess) {
        s->error = e != PADO_GPU_OK ? std::string(pado_gpu_last_error())
                                    : std::string("event record failed: ") + hipGetErrorString(r);
        return;
    }
    s->rlive = true;
    s->st.launches += 1;
}

// Cuts request r into descriptors, spreads them over the queues and raises
// the queues' published counts (mu held; r's pairs are in the ring).
void resident_publish(pado_gpu_server *s, Request *r)
{
    uint32_t touched = 0;
    for (uint64_t o = 0; o < r->k; o += kServeGroup) {
        const uint32_t q = s->qnext;
        s->qnext = (s->qnext + 1) % kServeQueues;
        ServeDesc &d = s->h_desc[(uint64_t)q * s->dcap + s->qpub[q] % s->dcap];
        d.pos = r->start + o;
        d.len = (uint32_t)std::min<uint64_t>(kServeGroup, r->k - o);
        d.slot = r->slot;
        d.seq = r->id;
        d.k = (uint32_t)r->k;
        s->qpub[q] += 1;
        touched |= 1u << q;
    }
    // The descriptors (and the pairs copied in before) are visible before
    // any count that covers them.
    std::atomic_thread_fence(std::memory_order_release);
    for (uint32_t q = 0; q < kServeQueues; ++q)
        if (touched >> q & 1u) __atomic_store_n(&s->h_ctl->pub[q], s->qpub[q], __ATOMIC_RELEASE);
    r->state = 1;
    s->rinflight.push_back(r);
    resident_launch(s);
}

void resident_finish(pado_gpu_server *s, Request *r, double t, bool failed, bool &release)
{
    r->state = 2;
    r->failed = failed;
    r->t_done = t;
    s->free_slots.push_back(r->slot);
    s->st.requests_done += 1;
    s->st.pairs_done += r->k;
    if (r->cancelled) {
        r->state = 3;  // nobody will wait for it
        release = true;
    }
}

// Completion: polls the flags of published requests and keeps a resident
// launch running while any is outstanding.  The flags are read without the
// lock (a snapshot of the oldest outstanding requests -- they finish roughly
// in publication order), so submitters and waiters are not held up by the
// polling; the lock is taken only to retire finished requests and to check
// the launch.
void resident_completer_loop(pado_gpu_server *s)
{
    (void)hipSetDevice(s->ctx->device);
    constexpr size_t kScan = 256;
    struct Snap {
        Request *r;
        uint32_t slot;
        uint64_t id;
    };
    std::vector<Snap> snap;
    std::vector<Request *> finished;
    snap.reserve(kScan);
    uint32_t idle_rounds = 0;
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        if (s->rinflight.empty()) {
            if (s->stopping) break;
            s->cv_launch.wait(lk);
            continue;
        }
        snap.clear();
        for (size_t i = 0; i < s->rinflight.size() && i < kScan; ++i) {
            Request *r = s->rinflight[i];
            snap.push_back({r, r->slot, r->id});
        }
        lk.unlock();
        finished.clear();
        for (const Snap &x : snap)
            if (__atomic_load_n(&s->h_flag[x.slot], __ATOMIC_ACQUIRE) == x.id) finished.push_back(x.r);
        const double t = now_ms();
        lk.lock();
        bool progressed = false, release = false;
        if (!finished.empty()) {
            // Only this thread removes entries, so every finished request is
            // still outstanding; erase them in one pass.
            for (Request *r : finished) resident_finish(s, r, t, false, release);
            size_t w = 0;
            for (size_t i = 0; i < s->rinflight.size(); ++i)
                if (s->rinflight[i]->state == 1) s->rinflight[w++] = s->rinflight[i];
            s->rinflight.resize(w);
            progressed = true;
            idle_rounds = 0;
        }
        // The launch's state: every round while nothing finishes for a while
        // (it may have left), otherwise every 64th.
        if (s->rlive && (++idle_rounds > 64 || finished.empty())) {
            idle_rounds = 0;
            const hipError_t q = hipEventQuery(s->rexit);
            if (q == hipSuccess) {
                s->rlive = false;
            } else if (q != hipErrorNotReady) {
                // The launch failed: no outstanding request gets answers.
                if (s->error.empty()) s->error = std::string("resident serving launch failed: ") + hipGetErrorString(q);
                s->rlive = false;
                for (Request *r : s->rinflight) resident_finish(s, r, t, true, release);
                s->rinflight.clear();
                progressed = true;
            }
        }
        if (!s->rlive && !s->rinflight.empty()) {
            if (s->error.empty()) {
                resident_launch(s);
            } else {
                for (Request *r : s->rinflight) resident_finish(s, r, t, true, release);
                s->rinflight.clear();
                progressed = true;
            }
        }
        if (release) release_consumed(s);
        if (progressed) {
            lk.unlock();
            s->cv_done.notify_all();
            s->cv_space.notify_all();
            lk.lock();
        } else {
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
        }
    }
}

// Blocks until request `id` is done, copies its answers, releases its slot.
// The copy runs outside the lock (the slot is not reused before release): a
// caller buffer that page-faults on first touch holds back only this caller,
// not admission, dispatch or completion.
pado_gpu_status wait_request(pado_gpu_server *s, uint64_t id, void *out, double *t_submit, double *t_done,
                             int64_t timeout_us = -1)
{
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->by_id.find(id);
    if (it == s->by_id.end() || it->second->state == 3 || it->second->claimed || it->second->cancelled) {
        set_error("unknown request id");
        return PADO_GPU_ERR_ARG;
    }
    Request *r = it->second;
    r->claimed = true;
    const double t_wait0 = now_ms();
    if (r->state < 2 && timeout_us != 0) {
        // Spin (without the lock) for up to 200 us or the timeout before
        // blocking: a condition-variable wake-up costs tens of microseconds,
        // which at hundreds of thousands of requests per second would bound
        // the service.  The request stays live: only this waiter releases it.
        lk.unlock();
        const double spin_ms = timeout_us < 0 ? 0.2 : std::min(0.2, timeout_us * 1e-3);
        const double t_end = now_ms() + spin_ms;
        while (r->state.load(std::memory_order_acquire) < 2 && now_ms() < t_end) std::this_thread::yield();
        lk.lock();
    }
    if (timeout_us < 0) {
        s->cv_done.wait(lk, [&] { return r->state >= 2; });
    } else if (!s->cv_done.wait_for(lk,
                                    std::chrono::microseconds(std::max<int64_t>(
                                        0, timeout_us - (int64_t)((now_ms() - t_wait0) * 1e3))),
                                    [&] { return r->state >= 2; })) {
        // Not done in time: the request stays live; the caller may wait again
        // or cancel it.
        r->claimed = false;
        set_error("request not done within the timeout");
        return PADO_GPU_ERR_TIMEOUT;
    }
    if (r->failed) {
        // No answers: the request is released (its ring space and slot are
        // not leaked) and the server's first error is reported.
        const std::string msg = s->error.empty() ? std::string("serving launch failed") : s->error;
        r->state = 3;
        release_consumed(s);
        lk.unlock();
        s->cv_work.notify_all();
        s->cv_space.notify_all();
        set_error(msg);
        return PADO_GPU_ERR_HIP;
    }
    lk.unlock();
    if (r->k && out) std::memcpy(out, s->h_out + (r->start % s->cap) * s->es, r->k * s->es);
    lk.lock();
    *t_submit = r->t_submit;
    *t_done = r->t_done;
    r->state = 3;
    release_consumed(s);
    lk.unlock();
    s->cv_work.notify_all();
    s->cv_space.notify_all();
    return PADO_GPU_OK;
}

// Resident-mode resources: control words, descriptor rings and completion
// flags in mapped host memory, queue state and per-slot counts on the
// device, the launch stream and its exit event, the kernel arguments.
pado_gpu_status resident_setup(pado_gpu_server *s)
{
    const unsigned mf = hipHostMallocMapped | hipHostMallocCoherent;
    const uint32_t nreq = s->cfg.max_requests;
    // In-flight descriptors: at most cap / kServeGroup + max_requests, dealt
    // round robin, so each queue holds at most an eighth of that plus one.
    s->dcap = (uint32_t)((s->cap / kServeGroup + nreq) / kServeQueues + 64);
    if (hipHostMalloc((void **)&s->h_ctl, sizeof(ServeCtl), mf) != hipSuccess ||
        hipHostMalloc((void **)&s->h_desc, sizeof(ServeDesc) * kServeQueues * s->dcap, mf) != hipSuccess ||
        hipHostMalloc((void **)&s->h_flag, sizeof(uint64_t) * nreq, mf) != hipSuccess) {
        set_error("pinned control allocation failed");
        return PADO_GPU_ERR_NOMEM;
    }
    std::memset(s->h_ctl, 0, sizeof(ServeCtl));
    std::memset(s->h_flag, 0, sizeof(uint64_t) * nreq);
    ServeKernelArgs &A = s->rargs;
    void *m_out = nullptr, *m_ctl = nullptr, *m_desc = nullptr, *m_flag = nullptr;
    if (hipHostGetDevicePointer(&m_out, s->h_out, 0) != hipSuccess ||
        hipHostGetDevicePointer(&m_ctl, s->h_ctl, 0) != hipSuccess ||
        hipHostGetDevicePointer(&m_desc, s->h_desc, 0) != hipSuccess ||
        hipHostGetDevicePointer(&m_flag, s->h_flag, 0) != hipSuccess) {
        set_error("control memory is not mapped for the device");
        return PADO_GPU_ERR_HIP;
    }
    PADO_TRY(s->rstate.alloc(kServeStateWords));
    PADO_TRY(s->rdone.alloc(nreq));
    PADO_HIP(hipStreamCreateWithFlags(&s->rstream, hipStreamNonBlocking));
    PADO_HIP(hipEventCreateWithFlags(&s->rexit, hipEventDisableTiming));
    PADO_HIP(hipMemsetAsync(s->rstate.p, 0, sizeof(unsigned long long) * kServeStateWords, s->rstream));
    PADO_HIP(hipMemsetAsync(s->rdone.p, 0, sizeof(uint32_t) * nreq, s->rstream));
    PADO_HIP(hipStreamSynchronize(s->rstream));
    s->free_slots.resize(nreq);
    for (uint32_t i = 0; i < nreq; ++i) s->free_slots[i] = nreq - 1 - i;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s->ctx->device) != hipSuccess || khz <= 0)
        khz = 100000;  // the 100 MHz constant clock of CDNA devices
    A.ra = s->m_a;
    A.rb = s->m_b;
    A.out = m_out;
    A.cap = s->cap;
    A.ctl = static_cast<const ServeCtl *>(m_ctl);
    A.desc = static_cast<const ServeDesc *>(m_desc);
    A.dcap = s->dcap;
    A.flags = static_cast<uint64_t *>(m_flag);
    A.state = s->rstate.p;
    A.rdone = s->rdone.p;
    // A launch leaves after 20 ms without work, and drains and leaves after
    // 100 ms in any case (the completer relaunches it while work is
    // outstanding), so device-wide synchronisation elsewhere in the process
    // never waits long behind it.
    A.idle_ticks = (uint64_t)khz * 20;
    A.max_ticks = (uint64_t)khz * 100;
    A.ref_compat = (!s->idx->wide && (s->cfg.flags & PADO_GPU_QUERY_REF_COMPAT) != 0) ? 1 : 0;
    // Workgroup b serves queue b % kServeQueues: every queue needs at least
    // one workgroup of workers.
    s->rblocks = serve_resident_blocks_per_cu(s->idx) * (s->ctx->num_cu > 0 ? (uint32_t)s->ctx->num_cu : 256u);
    s->rblocks = std::max<uint32_t>(s->rblocks, 2 * kServeQueues);
    return PADO_GPU_OK;
}

// Admission: ring space and a request slot (plus a completion-flag slot in
// resident mode) for k pairs.  The request is not yet visible to the
// dispatcher or the kernel.
pado_gpu_status submit_reserve(pado_gpu_server *s, uint64_t k, Request **out)
{
    if (k > s->cap || (s->resident && k > 0xFFFFFFFFull)) {
        set_error("request larger than the staging ring");
        return PADO_GPU_ERR_ARG;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->stopping) {
        set_error("server stopping");
        return PADO_GPU_ERR_STATE;
    }
    if (!s->error.empty()) {
        // A launch failed: the device state is suspect, admit nothing more.
        set_error("server failed: " + s->error);
        return PADO_GPU_ERR_STATE;
    }
    rewind_if_empty(s);
    const uint64_t start = place(s->tail, k, s->cap);
    if (s->by_id.size() >= s->cfg.max_requests || start + k - s->head > s->cap ||
        (s->resident && s->free_slots.empty())) {
        s->st.rejected += 1;
        set_error("no free request slot or staging memory");
        return PADO_GPU_ERR_BUSY;
    }
    auto *r = new Request();
    r->id = s->next_id++;
    r->start = start;
    r->k = k;
    if (s->resident) {
        r->slot = s->free_slots.back();
        s->free_slots.pop_back();
    }
    r->t_submit = now_ms();
    s->tail = start + k;
    s->live.push_back(r);
    s->by_id[r->id] = r;
    *out = r;
    return PADO_GPU_OK;
}

// The request's pairs into its ring space (no lock: the space is its own).
void submit_copy(pado_gpu_server *s, const Request *r, const uint32_t *a, const uint32_t *b)
{
    const uint64_t phys = r->start % s->cap;
    if (r->k) {
        std::memcpy(s->h_a + phys, a, r->k * 4);
        std::memcpy(s->h_b + phys, b, r->k * 4);
    }
}

// Makes a reserved, filled request visible: queued for the dispatcher, or
// published to the resident kernel.
void submit_commit(pado_gpu_server *s, Request *r)
{
    std::unique_lock<std::mutex> lk(s->mu);
    if (r->cancelled) {
        // Cancelled while its pairs were copied in: never launched.
        r->state = 3;
        if (s->resident) s->free_slots.push_back(r->slot);
        release_consumed(s);
        lk.unlock();
        s->cv_space.notify_all();
        return;
    }
    if (s->resident) {
        if (r->k == 0) {
            // Nothing to answer: done at once.
            bool release = false;
            resident_finish(s, r, now_ms(), false, release);
            lk.unlock();
            s->cv_done.notify_all();
            return;
        }
        resident_publish(s, r);
        lk.unlock();
        s->cv_launch.notify_one();
        return;
    }
    s->pending.push_back(r);
    s->pending_pairs += r->k;
    lk.unlock();
    s->cv_work.notify_one();
}

}  // namespace

extern "C" {

pado_gpu_status pado_gpu_serve_start(pado_gpu_ctx *ctx, pado_gpu_index *idx, const pado_gpu_serve_cfg *cfg,
                                     pado_gpu_server **out)
{
    clear_error();
    if (!ctx || !idx || !out) {
        set_error("null argument");
        return PADO_GPU_ERR_ARG;
    }
    *out = nullptr;
    if (!idx->ordered) {
        set_error("serving needs an ordered index (call pado_gpu_index_order first)");
        return PADO_GPU_ERR_STATE;
    }
    PADO_HIP(hipSetDevice(ctx->device));
    auto *s = new pado_gpu_server();
    s->ctx = ctx;
    s->idx = idx;
    if (cfg) s->cfg = *cfg;
    if (!s->cfg.max_launch_pairs) s->cfg.max_launch_pairs = 4ull << 20;
    if (!s->cfg.max_requests) s->cfg.max_requests = 4096;
    if (!s->cfg.streams) s->cfg.streams = 3;
    s->es = idx->weighted ? 4u : (idx->wide ? 2u : 1u);
    // Persistent kernel: as many resident workgroups of 256 threads per
    // compute unit as the kernel's occupancy allows.
    s->ws_blocks = query_ws_blocks_per_cu(idx->weighted) * (ctx->num_cu > 0 ? (uint32_t)ctx->num_cu : 256u);
    uint64_t cap = s->cfg.staging_pairs;
    if (!cap) {
        // Admission by free device memory: a quarter of what is free, 8 bytes
        // of ids plus the answer per staged pair, at most 16M pairs.
        size_t fr = 0, tot = 0;
        PADO_HIP(hipMemGetInfo(&fr, &tot));
        cap = std::min<uint64_t>(16ull << 20, (uint64_t)(fr / 4) / (8 + s->es));
    }
    if (cap < 1024) cap = 1024;
    s->cap = cap;
    s->st.staging_pairs = cap;
    s->resident = (s->cfg.flags & PADO_GPU_SERVE_RESIDENT) != 0;
    auto fail = [&](pado_gpu_status code) {
        if (s->h_a) (void)hipHostFree(s->h_a);
        if (s->h_b) (void)hipHostFree(s->h_b);
        if (s->h_out) (void)hipHostFree(s->h_out);
        if (s->h_ctl) (void)hipHostFree(s->h_ctl);
        if (s->h_desc) (void)hipHostFree(s->h_desc);
        if (s->h_flag) (void)hipHostFree(s->h_flag);
        if (s->rstream) (void)hipStreamDestroy(s->rstream);
        if (s->rexit) (void)hipEventDestroy(s->rexit);
        for (auto q : s->streams) (void)hipStreamDestroy(q);
        for (auto e : s->ev_start) (void)hipEventDestroy(e);
        for (auto e : s->ev_end) (void)hipEventDestroy(e);
        delete s;
        return code;
    };
    // The id rings are read by k_stage_pairs over the host link: coherent
    // (not cached on the device), since every lap rewrites the same addresses.
    const unsigned ring_flags = hipHostMallocMapped | hipHostMallocCoherent;
    // In resident mode the kernel also writes the answers straight into the
    // (then mapped) answer ring.
    if (hipHostMalloc((void **)&s->h_a, cap * 4, ring_flags) != hipSuccess ||
        hipHostMalloc((void **)&s->h_b, cap * 4, ring_flags) != hipSuccess ||
        hipHostMalloc((void **)&s->h_out, cap * s->es, s->resident ? ring_flags : hipHostMallocDefault) !=
            hipSuccess) {
        set_error("pinned staging allocation failed");
        return fail(PADO_GPU_ERR_NOMEM);
    }
    if (hipHostGetDevicePointer((void **)&s->m_a, s->h_a, 0) != hipSuccess ||
        hipHostGetDevicePointer((void **)&s->m_b, s->h_b, 0) != hipSuccess) {
        set_error("staging ring is not mapped for the device");
        return fail(PADO_GPU_ERR_HIP);
    }
    if (s->resident) {
        pado_gpu_status rs = resident_setup(s);
        if (rs != PADO_GPU_OK) return fail(rs);
        s->completer = std::thread(resident_completer_loop, s);
        *out = s;
        return PADO_GPU_OK;
    }
    pado_gpu_status a1 = s->d_a.alloc(cap), a2 = s->d_b.alloc(cap), a3 = s->d_out.alloc(cap * s->es);
    pado_gpu_status a4 = s->work.alloc(s->cfg.streams);
    if (a1 != PADO_GPU_OK || a2 != PADO_GPU_OK || a3 != PADO_GPU_OK || a4 != PADO_GPU_OK)
        return fail(PADO_GPU_ERR_NOMEM);
    for (uint32_t i = 0; i < s->cfg.streams; ++i) {
        hipStream_t q = nullptr;
        if (hipStreamCreateWithFlags(&q, hipStreamNonBlocking) != hipSuccess) {
            set_error("stream creation failed");
            return fail(PADO_GPU_ERR_HIP);
        }
        s->streams.push_back(q);
        s->stream_busy.push_back(0);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
            set_error("event creation failed");
            return fail(PADO_GPU_ERR_HIP);
        }
        s->ev_start.push_back(e0);
        s->ev_end.push_back(e1);
    }
    s->min_launch = s->cfg.min_launch_pairs ? s->cfg.min_launch_pairs : (256u << 10);
    s->max_wait_ms = (s->cfg.max_wait_us ? s->cfg.max_wait_us : 200u) * 1e-3;
    s->dispatcher = std::thread(dispatcher_loop, s);
    s->completer = std::thread(completer_loop, s);
    *out = s;
    return PADO_GPU_OK;
}

pado_gpu_status pado_gpu_serve_submit(pado_gpu_server *s, const uint32_t *a, const uint32_t *b, uint64_t k,
                                      uint64_t *req_id)
{
    clear_error();
    if (!s || !req_id || (k && (!a || !b))) {
        set_error("null argument");
        return PADO_GPU_ERR_ARG;
    }
    Request *r = nullptr;
    PADO_TRY(submit_reserve(s, k, &r));
    *req_id = r->id;
    // The slot is reserved; the pairs are copied in without the lock and the
    // request becomes visible to the dispatcher (or the resident kernel)
    // only after the copy.
    submit_copy(s, r, a, b);
    submit_commit(s, r);
    return PADO_GPU_OK;
}

pado_gpu_status pado_gpu_serve_cancel(pado_gpu_server *s, uint64_t req_id)
{
    clear_error();
    if (!s) return PADO_GPU_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->by_id.find(req_id);
    if (it == s->by_id.end() || it->second->state == 3 || it->second->cancelled) {
        set_error("unknown request id");
        return PADO_GPU_ERR_ARG;
    }
    Request *r = it->second;
    if (r->claimed) {
        set_error("a wait on this request is in progress");
        return PADO_GPU_ERR_STATE;
    }
    r->cancelled = true;
    if (r->state == 0) {
        auto p = std::find(s->pending.begin(), s->pending.end(), r);
        if (p != s->pending.end()) {
            // Queued, not launched: dropped now.
            s->pending.erase(p);
            s->pending_pairs -= r->k;
            r->state = 3;
        }
        // Otherwise submit is still copying its pairs in and drops it after.
    } else if (r->state == 2) {
        r->state = 3;  // done: its answers are discarded
    }
    // A launched request (state 1) is released by the completer when its
    // launch finishes.
    release_consumed(s);
    lk.unlock();
    s->cv_space.notify_all();
    s->cv_work.notify_all();
    return PADO_GPU_OK;
}

pado_gpu_status pado_gpu_serve_poll(pado_gpu_server *s, uint64_t req_id, int *done)
{
    if (!s || !done) return PADO_GPU_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->by_id.find(req_id);
    if (it == s->by_id.end()) {
        set_error("unknown request id");
        return PADO_GPU_ERR_ARG;
    }
    *done = it->second->state >= 2;
    return PADO_GPU_OK;
}

pado_gpu_status pado_gpu_serve_wait(pado_gpu_server *s, uint64_t req_id, void *out, double *latency_ms)
{
    clear_error();
    if (!s) return PADO_GPU_ERR_ARG;
    double t_submit = 0, t_done = 0;
    PADO_TRY(wait_request(s, req_id, out, &t_submit, &t_done));
    if (latency_ms) *latency_ms = t_done - t_submit;
    return PADO_GPU_OK;
}

pado_gpu_status pado_gpu_serve_wait_for(pado_gpu_server *s, uint64_t req_id, void *out, double *latency_ms,
                                        uint64_t timeout_us)
{
    clear_error();
    if (!s) return PADO_GPU_ERR_ARG;
    double t_submit = 0, t_done = 0;
    // Beyond about 11 days the wait is unbounded (and the clock arithmetic safe).
    const int64_t t = timeout_us >= (1ull << 40) ? -1 : (int64_t)timeout_us;
    PADO_TRY(wait_request(s, req_id, out, &t_submit, &t_done, t));
    if (latency_ms) *latency_ms = t_done - t_submit;
    return PADO_GPU_OK;
}

pado_gpu_status pado_gpu_serve_replay(pado_gpu_server *s, const uint32_t *a, const uint32_t *b,
                                      const uint64_t *req_off, uint32_t num_req, const double *arrival_s,
                                      double *latency_ms, void *answers, double *wall_s)
{
    clear_error();
    if (!s || !req_off || !arrival_s || !latency_ms || (req_off[num_req] && (!a || !b))) {
        set_error("null argument");
        return PADO_GPU_ERR_ARG;
    }
    for (uint32_t i = 0; i < num_req; ++i)
        if (req_off[i + 1] < req_off[i] || (i && arrival_s[i] < arrival_s[i - 1])) {
            set_error("request offsets and arrival times must be non-decreasing");
            return PADO_GPU_ERR_ARG;
        }
    std::vector<std::atomic<uint64_t>> ids(num_req);
    for (auto &x : ids) x.store(0);
    std::atomic<int> failed{0};
    std::string sub_error;
    const double t0 = now_ms() + 1.0;
    // Open-loop submitters: request i enters at t0 + arrival_s[i] whatever
    // happened to earlier requests; a refused admission is retried.  In
    // resident mode several threads share the requests (i mod kSub), so the
    // copies of the pairs into the ring run in parallel, while admissions
    // stay in arrival order (a request reserves its space only after the one
    // before it).  The dispatcher of the launch mode packs requests that
    // become pending in ring order, so there one thread submits.
    const uint32_t kSub = s->resident ? 4u : 1u;
    std::atomic<uint32_t> turn{0};
    std::mutex err_mu;
    auto submitter = [&](uint32_t first) {
        for (uint32_t i = first; i < num_req; i += kSub) {
            const double due = t0 + arrival_s[i] * 1e3;
            while (now_ms() < due) {
                if (failed.load()) return;  // the waiter gave up: stop submitting
                if (due - now_ms() > 0.2) std::this_thread::yield();
            }
            while (turn.load(std::memory_order_acquire) != i) {
                if (failed.load()) return;
                std::this_thread::yield();
            }
            const uint64_t k = req_off[i + 1] - req_off[i];
            Request *r = nullptr;
            for (;;) {
                if (failed.load()) return;
                pado_gpu_status st = submit_reserve(s, k, &r);
                if (st == PADO_GPU_OK) break;
                if (st != PADO_GPU_ERR_BUSY) {
                    {
                        std::lock_guard<std::mutex> g(err_mu);
                        sub_error = pado_gpu_last_error();
                    }
                    failed.store(1);
                    return;
                }
                // Refused: wait for a release instead of retrying at once, so
                // that retries do not contend for the lock the completer
                // needs (under overload that cut the served rate).
                std::unique_lock<std::mutex> lk(s->mu);
                s->cv_space.wait_for(lk, std::chrono::microseconds(200));
            }
            turn.store(i + 1, std::memory_order_release);
            ids[i].store(r->id);
            submit_copy(s, r, a + req_off[i], b + req_off[i]);
            submit_commit(s, r);
        }
    };
    std::vector<std::thread> subs;
    for (uint32_t t = 0; t < kSub; ++t) subs.emplace_back(submitter, t);
    pado_gpu_status result = PADO_GPU_OK;
    double last_done = t0;
    for (uint32_t i = 0; i < num_req; ++i) {
        uint64_t id = 0;
        while ((id = ids[i].load()) == 0 && !failed.load()) std::this_thread::yield();
        if (id == 0) break;
        double t_submit = 0, t_done = 0;
        uint8_t *dst = answers ? (uint8_t *)answers + req_off[i] * s->es : nullptr;
        pado_gpu_status st = wait_request(s, id, dst, &t_submit, &t_done);
        if (st != PADO_GPU_OK) {
            result = st;
            failed.store(1);
            break;
        }
        latency_ms[i] = t_done - (t0 + arrival_s[i] * 1e3);
        if (t_done > last_done) last_done = t_done;
    }
    for (auto &t : subs) t.join();
    if (failed.load() && result == PADO_GPU_OK) {
        set_error(sub_error.empty() ? std::string("replay failed") : sub_error);
        result = PADO_GPU_ERR_HIP;
    }
    if (wall_s) *wall_s = (last_done - t0) * 1e-3;
    return result;
}

pado_gpu_status pado_gpu_serve_get_stats(pado_gpu_server *s, pado_gpu_serve_stats *st)
{
    if (!s || !st) return PADO_GPU_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    *st = s->st;
    return PADO_GPU_OK;
}

void pado_gpu_serve_stop(pado_gpu_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    s->cv_launch.notify_all();
    if (s->dispatcher.joinable()) s->dispatcher.join();
    s->cv_launch.notify_all();
    if (s->completer.joinable()) s->completer.join();
    if (s->resident) {
        // Every published request is done: the resident launch (if any)
        // drains nothing more and leaves.
        __atomic_store_n(&s->h_ctl->stop, 1u, __ATOMIC_RELEASE);
        (void)hipStreamSynchronize(s->rstream);
        (void)hipStreamDestroy(s->rstream);
        (void)hipEventDestroy(s->rexit);
        (void)hipHostFree(s->h_ctl);
        (void)hipHostFree(s->h_desc);
        (void)hipHostFree(s->h_flag);
    }
    for (auto q : s->streams) {
        (void)hipStreamSynchronize(q);
        (void)hipStreamDestroy(q);
    }
    for (auto e : s->ev_start) (void)hipEventDestroy(e);
    for (auto e : s->ev_end) (void)hipEventDestroy(e);
    for (Request *r : s->live) delete r;
    (void)hipHostFree(s->h_a);
    (void)hipHostFree(s->h_b);
    (void)hipHostFree(s->h_out);
    delete s;
}

}  // extern "C"
