// Request server: field evaluations and GS-PAT solves served continuously on
// one GPU. No reference counterpart (the reference evaluates one call at a
// time on the caller's thread, capi/src/calculator.rs:53-117).
//
//   * Transducer arrays are registered once and stay resident (packed sources).
//   * Admission by HBM budget: every request (or GS-PAT batch) takes its inputs
//     and outputs from a device arena of a fixed size, allocated once; a request
//     that does not fit waits (FIFO) until completions free space. Results are
//     staged through a pinned host arena the same way.
//   * Continuous batching: one dispatcher thread takes everything that has
//     arrived each cycle; GS-PAT requests on the same array become one batched
//     launch, field requests are spread over K streams so that small grids run
//     side by side; completions are polled through events, so new requests are
//     admitted while earlier ones are still on the GPU.
//   * Each request's result is produced by the same kernels, with the same
//     per-request reduction order, as when it runs alone.
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "afh_holo.hpp"
#include "afh_host.hpp"
#include "afh_internal.hpp"

namespace afh {

static double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over one block, 256-byte granularity.
struct Arena {
    size_t size = 0;
    std::map<size_t, size_t> free_;  // offset -> bytes
    void init(size_t bytes) {
        size = bytes;
        free_.clear();
        free_[0] = bytes;
    }
    static size_t round(size_t b) { return (std::max<size_t>(b, 1) + 255) & ~(size_t)255; }
    bool alloc(size_t bytes, size_t* off) {
        bytes = round(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes, o = it->first;
                free_.erase(it);
                if (rest) free_[o + bytes] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = round(bytes);
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct SrvArray {
    int source_type;
    uint32_t n, np;
    float* packed = nullptr;  // device
};

enum ReqKind { kReqField = 0, kReqGSPAT = 1 };
enum ReqState { kPending = 0, kRunning = 1, kDone = 2 };

struct Request {
    uint64_t ticket = 0;
    int kind = 0;
    uint32_t array_id = 0;
    const SrvArray* arr = nullptr;  // taken under the server lock at submission
    // field
    GridArea grid;
    bool scatter = false;
    std::vector<float> points;
    int field_type = 0;
    uint64_t n_points = 0;
    // GS-PAT
    std::vector<float> foci, amps;
    uint32_t m = 0;
    // outputs (caller memory)
    float* out = nullptr;
    float* out2 = nullptr;
    size_t out_bytes = 0;   // bytes copied to `out` (and again to `out2` for GS-PAT)
    size_t pin_off = 0;     // this request's results in the pinned arena
    size_t pin_off2 = 0;    // GS-PAT: its phases
    double t_submit = 0, t_done = 0;
    std::atomic<int> state{kPending};  // written by the dispatcher and copy threads, read by AFS_Wait
    int32_t status = AFH_OK;
    bool direct = false;               // field result DMA'd straight into `out` (pinned caller memory)
    std::atomic<int> parts_left{0};    // copy jobs of this request still to run
};

// A completed launch's pinned staging, released once the results of all its
// requests have been copied out.
struct PinHold {
    size_t off = 0, bytes = 0;
    std::atomic<int> left{0};
};

// Results of one request to copy from pinned staging to the caller's memory.
struct CopyJob {
    Request* r = nullptr;
    std::shared_ptr<PinHold> hold;
    size_t off = 0, bytes = 0;  // the part of the result this job copies
    bool second = false;        // GS-PAT: the phases (out2) rather than the amplitudes
};

// Results larger than this are copied by several copy threads in parts.
constexpr size_t kCopyPart = 1 << 20;

// One launch in flight: a field request or a GS-PAT batch.
struct Flight {
    std::vector<Request*> reqs;
    hipEvent_t ev = nullptr;
    size_t dev_off = 0, dev_bytes = 0, pin_off = 0, pin_bytes = 0;
    int stream = 0;
    double work = 0;  // source-point terms, for the stream choice
    bool gspat = false;
    int slot = 0;  // GS-PAT lane
};

struct Server {
    AFS_Config cfg;
    char* dev_base = nullptr;
    char* pin_base = nullptr;
    Arena dev, pin;
    std::vector<hipStream_t> streams;
    std::vector<double> outstanding;  // queued terms per stream
    std::vector<std::unique_ptr<Context>> ctxs;
    // arrays are added while serving: the worker reads them through the
    // pointers requests took at submission, which a push_back does not move
    std::vector<std::unique_ptr<SrvArray>> arrays;
    std::mutex mu;
    std::mutex add_mu;  // serializes AFS_AddArray (an array's id picks its GS-PAT context)
    std::condition_variable cv_submit, cv_done;
    std::deque<Request*> pending;
    std::unordered_map<uint64_t, Request*> by_ticket;
    std::deque<Flight> flights;
    // The copies out of pinned staging run on a few threads of their own: one
    // thread copying every result (several GB/s under load) would throttle
    // launching and, by holding staging space, admission.
    static constexpr int kCopiers = 4;
    std::mutex copy_mu, pin_mu;  // pin_mu guards the pinned arena
    std::condition_variable cv_copy;
    std::deque<CopyJob> copies;
    std::vector<std::thread> copiers;
    bool copy_stop = false;
    std::vector<double> latencies[2];  // by request kind
    uint64_t next_ticket = 1, completed = 0;
    bool stop = false;
    std::thread worker;
    int next_stream = 0;
    int device = 0;

    ~Server() {
        {
            std::lock_guard<std::mutex> g(mu);
            stop = true;
        }
        cv_submit.notify_all();
        if (worker.joinable()) worker.join();
        {
            std::lock_guard<std::mutex> g(copy_mu);
            copy_stop = true;
        }
        cv_copy.notify_all();
        for (auto& t : copiers)
            if (t.joinable()) t.join();
        for (auto s : streams) (void)hipStreamSynchronize(s);
        for (auto s : gextra) (void)hipStreamSynchronize(s);
        ctxs.clear();
        gextra_ctx.clear();
        for (auto s : streams) (void)hipStreamDestroy(s);
        for (auto s : gextra) (void)hipStreamDestroy(s);
        for (auto& a : arrays)
            if (a->packed) (void)hipFree(a->packed);
        if (dev_base) (void)hipFree(dev_base);
        if (pin_base) (void)hipHostFree(pin_base);
        for (auto& kv : by_ticket) delete kv.second;  // the worker drained `pending` first
    }

    // GS-PAT batches run on lanes, each a stream with its own context
    // (workspace): lane 0 is stream 0, the others are streams of their own (on
    // the CU partition when the configuration reserves one), so a solve never
    // waits behind a large field evaluation. Array a owns lanes_per_array lanes
    // from lanes_per_array * a on (mod the lane count): a batch of the array
    // starts on an idle one, so with two lanes its next batch runs beside the
    // previous one, and the batches of different arrays run side by side.
    // Without a CU partition there are 6 lanes, 2 per array; with one, 3 lanes,
    // 1 per array: every CU-masked stream holds a hardware queue of its own,
    // and when each lane also had 4 auxiliary streams, 6 lanes' 30 queues
    // beside the field streams' were measured to stall the field requests for
    // milliseconds. A field request goes to the field stream with the least
    // queued work (all streams when there is only one).
    static constexpr int kGspatLanes = 6;
    int gspat_inflight[kGspatLanes] = {};  // GS-PAT batches on the GPU, per lane
    int lanes_per_array = 2;
    std::vector<hipStream_t> gextra;       // lanes 1 .. gspat_lanes() - 1
    std::vector<std::unique_ptr<Context>> gextra_ctx;
    int gspat_lanes() const { return 1 + (int)gextra.size(); }
    int array_lane(uint32_t array_id, int k) const {
        return (int)(((uint64_t)lanes_per_array * array_id + (uint64_t)(k % lanes_per_array)) % gspat_lanes());
    }
    hipStream_t lane_stream(int lane) { return lane == 0 ? streams[0] : gextra[lane - 1]; }
    Context* lane_ctx(int lane) { return lane == 0 ? ctxs[0].get() : gextra_ctx[lane - 1].get(); }
    // an idle lane of the array, or -1
    int idle_lane(uint32_t array_id) const {
        for (int k = 0; k < lanes_per_array; ++k)
            if (gspat_inflight[array_lane(array_id, k)] == 0) return array_lane(array_id, k);
        return -1;
    }
    int field_stream_index() {
        if (streams.size() == 1) return 0;
        int best = 1;
        for (int i = 2; i < (int)streams.size(); ++i)
            if (outstanding[i] < outstanding[best]) best = i;
        return best;
    }
    int stream_index(hipStream_t s) const {
        for (size_t i = 0; i < streams.size(); ++i)
            if (streams[i] == s) return (int)i;
        return 0;
    }

    // Launches what fits, in arrival order: field requests one launch each (a
    // field request that does not fit holds back the later ones, so large
    // requests are not starved); GS-PAT requests grouped per array into batched
    // launches of at most max_gspat_batch instances. A new GS-PAT batch on a
    // GS-PAT stream starts only when that stream's previous batch has
    // completed: requests that arrive in the meantime wait and form the next,
    // larger batch (a batch costs a few latency-bound launches whatever its
    // size, so under load this raises the solve rate; at low load a request
    // still starts at once), while the other streams' arrays go ahead.
    void dispatch(const std::vector<Request*>& todo, std::vector<Request*>& admitted) {
        std::map<uint32_t, std::vector<Request*>> groups;
        for (Request* r : todo) {
            if (r->kind == kReqGSPAT) {
                if (idle_lane(r->array_id) >= 0) groups[r->array_id].push_back(r);
                continue;
            }
            if (!launch_field(r)) break;
            admitted.push_back(r);
        }
        for (auto& kv : groups) {
            std::vector<Request*>& rs = kv.second;
            for (size_t start = 0; start < rs.size();) {
                // the largest batch that fits the arena now (one request always
                // fits an empty arena: checked at submission)
                const int lane = idle_lane(kv.first);
                if (lane < 0) break;  // both lanes busy: the rest waits for the next batch
                size_t cnt = std::min<size_t>(rs.size() - start, cfg.max_gspat_batch);
                while (cnt > 0 && !launch_gspat(kv.first, lane, &rs[start], cnt)) cnt /= 2;
                if (cnt == 0) break;
                admitted.insert(admitted.end(), rs.begin() + start, rs.begin() + start + cnt);
                start += cnt;
            }
        }
    }

    bool launch_field(Request* r) {
        const SrvArray& A = *r->arr;
        const size_t per = r->field_type == AFC_FIELD_COMPLEX ? 2 : 1;
        const size_t out_b = r->n_points * per * sizeof(float);
        const size_t pts_b = r->scatter ? r->n_points * 3 * sizeof(float) : 0;
        Flight f;
        f.dev_bytes = Arena::round(out_b) + Arena::round(pts_b);  // out | points
        // pinned staging: the result (unless it goes straight to the caller's
        // pinned buffer), then the scatter points
        const size_t stage_out = r->direct ? 0 : out_b;
        f.pin_bytes = stage_out + pts_b;
        if (!dev.alloc(f.dev_bytes, &f.dev_off)) return false;
        if (f.pin_bytes && !pin_alloc(f.pin_bytes, &f.pin_off)) {
            dev.release(f.dev_off, f.dev_bytes);
            return false;
        }
        f.stream = field_stream_index();
        f.work = (double)r->n_points * A.n;
        outstanding[f.stream] += f.work;
        hipStream_t st = streams[f.stream];
        char* d = dev_base + f.dev_off;
        char* h = pin_base + f.pin_off;
        const float* pts_dev = nullptr;
        if (r->scatter) {
            memcpy(h + stage_out, r->points.data(), pts_b);
            (void)hipMemcpyAsync(d + Arena::round(out_b), h + stage_out, pts_b, hipMemcpyHostToDevice, st);
            pts_dev = reinterpret_cast<const float*>(d + Arena::round(out_b));
        }
        FieldArgs a = field_args_for(r->scatter ? nullptr : &r->grid, pts_dev, r->n_points);
        a.sources = A.packed;
        a.n_sources = A.np;
        a.out = reinterpret_cast<float*>(d);
        hipError_t e = afh::launch_field(a, A.source_type, r->scatter ? kScatter : kGrid, r->field_type, kFast, st);
        if (e == hipSuccess) e = hipMemcpyAsync(r->direct ? static_cast<void*>(r->out) : h, d, out_b, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&f.ev, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(f.ev, st);
        r->out_bytes = out_b;
        r->pin_off = f.pin_off;
        f.reqs.push_back(r);
        if (e != hipSuccess) {
            r->status = AFH_ERR_HIP;
            (void)hipStreamSynchronize(st);
        }
        r->state = kRunning;
        flights.push_back(f);
        return true;
    }

    bool launch_gspat(uint32_t array_id, int lane, Request** rs, size_t cnt) {
        const SrvArray& A = *rs[0]->arr;  // every request of the group is on array_id
        size_t sum_m = 0;
        for (size_t i = 0; i < cnt; ++i) sum_m += rs[i]->m;
        const size_t foci_b = sum_m * 3 * sizeof(float), amps_b = sum_m * sizeof(float);
        const size_t out_b = cnt * A.n * sizeof(float);
        const size_t o_amps = Arena::round(foci_b), o_oa = o_amps + Arena::round(amps_b);
        const size_t o_op = o_oa + Arena::round(out_b);
        Flight f;
        f.dev_bytes = o_op + out_b;
        f.pin_bytes = f.dev_bytes;
        if (!dev.alloc(f.dev_bytes, &f.dev_off)) return false;
        if (!pin_alloc(f.pin_bytes, &f.pin_off)) {
            dev.release(f.dev_off, f.dev_bytes);
            return false;
        }
        hipStream_t st = lane_stream(lane);
        Context* ctx = lane_ctx(lane);
        char* d = dev_base + f.dev_off;
        char* h = pin_base + f.pin_off;
        std::vector<uint32_t> offs(cnt + 1, 0);
        size_t pos = 0;
        for (size_t i = 0; i < cnt; ++i) {
            memcpy(h + pos * 3 * sizeof(float), rs[i]->foci.data(), rs[i]->m * 3 * sizeof(float));
            memcpy(h + o_amps + pos * sizeof(float), rs[i]->amps.data(), rs[i]->m * sizeof(float));
            pos += rs[i]->m;
            offs[i + 1] = (uint32_t)pos;
        }
        hipError_t e = hipMemcpyAsync(d, h, o_oa, hipMemcpyHostToDevice, st);
        HoloBatch hb;
        hb.sources = A.packed;
        hb.n = A.n;
        hb.n_packed = A.np;
        hb.source_type = A.source_type;
        hb.foci = reinterpret_cast<const float*>(d);
        hb.amps = reinterpret_cast<const float*>(d + o_amps);
        hb.offsets_host = offs.data();
        hb.batch = (uint32_t)cnt;
        hb.algo = kAlgoGSPAT;
        hb.precision = cfg.gspat_precision;
        hb.iterations = cfg.gspat_iterations;
        hb.out_amp = reinterpret_cast<float*>(d + o_oa);
        hb.out_phase = reinterpret_cast<float*>(d + o_op);
        int32_t rc = AFH_OK;
        if (e == hipSuccess) {
            std::lock_guard<std::mutex> g(ctx->mu);
            rc = holo_solve(ctx, hb, st);
        }
        if (e == hipSuccess && rc == AFH_OK)
            e = hipMemcpyAsync(h + o_oa, d + o_oa, o_op + out_b - o_oa, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&f.ev, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(f.ev, st);
        for (size_t i = 0; i < cnt; ++i) {
            Request* r = rs[i];
            r->out_bytes = A.n * sizeof(float);
            r->pin_off = f.pin_off + o_oa + i * A.n * sizeof(float);
            r->pin_off2 = f.pin_off + o_op + i * A.n * sizeof(float);
            r->status = (e != hipSuccess || rc != AFH_OK) ? (rc != AFH_OK ? rc : AFH_ERR_HIP) : AFH_OK;
            r->state = kRunning;
            f.reqs.push_back(r);
        }
        if (e != hipSuccess || rc != AFH_OK) (void)hipStreamSynchronize(st);
        f.gspat = true;
        f.slot = lane;
        ++gspat_inflight[lane];
        flights.push_back(f);
        return true;
    }

    bool pin_alloc(size_t bytes, size_t* off) {
        std::lock_guard<std::mutex> g(pin_mu);
        return pin.alloc(bytes, off);
    }

    // Hands finished flights to the copy threads (caller holds no lock);
    // returns how many flights finished.
    int reap() {
        int n = 0;
        std::vector<CopyJob> jobs;
        for (auto it = flights.begin(); it != flights.end();) {
            const hipError_t q = it->ev ? hipEventQuery(it->ev) : hipSuccess;
            if (q == hipErrorNotReady) {
                ++it;
                continue;
            }
            auto hold = std::make_shared<PinHold>();
            hold->off = it->pin_off;
            hold->bytes = it->pin_bytes;
            const size_t first = jobs.size();
            for (Request* r : it->reqs) {
                if (q != hipSuccess && r->status == AFH_OK) r->status = AFH_ERR_HIP;
                // a large field result is copied in parts by several copy
                // threads; a GS-PAT result (amplitudes and phases), a result
                // already in the caller's memory or a failed request is one job
                const bool parts = r->status == AFH_OK && r->kind == kReqField && !r->direct;
                const size_t bytes = parts ? r->out_bytes : 0;
                const size_t n = parts ? std::max<size_t>(1, (bytes + kCopyPart - 1) / kCopyPart) : 1;
                r->parts_left = (int)n;
                for (size_t p = 0; p < n; ++p) {
                    CopyJob j;
                    j.r = r;
                    j.hold = hold;
                    j.off = p * kCopyPart;
                    j.bytes = parts ? std::min(kCopyPart, bytes - j.off) : 0;
                    jobs.push_back(j);
                }
            }
            hold->left = (int)(jobs.size() - first);
            if (it->ev) (void)hipEventDestroy(it->ev);
            if (it->gspat) --gspat_inflight[it->slot];
            outstanding[it->stream] -= it->work;
            dev.release(it->dev_off, it->dev_bytes);
            it = flights.erase(it);
            ++n;
        }
        if (!jobs.empty()) {
            {
                std::lock_guard<std::mutex> g(copy_mu);
                copies.insert(copies.end(), jobs.begin(), jobs.end());
            }
            cv_copy.notify_all();
        }
        return n;
    }

    // Copy thread: results to the caller's buffers, then the request is done.
    void copy_loop() {
        for (;;) {
            CopyJob j;
            {
                std::unique_lock<std::mutex> g(copy_mu);
                cv_copy.wait(g, [&] { return copy_stop || !copies.empty(); });
                if (copies.empty()) return;
                j = copies.front();
                copies.pop_front();
            }
            Request* r = j.r;
            if (r->status == AFH_OK) {
                if (r->kind == kReqGSPAT) {
                    memcpy(r->out, pin_base + r->pin_off, r->out_bytes);
                    memcpy(r->out2, pin_base + r->pin_off2, r->out_bytes);
                } else if (j.bytes) {
                    memcpy(reinterpret_cast<char*>(r->out) + j.off, pin_base + r->pin_off + j.off, j.bytes);
                }
            }
            if (--j.hold->left == 0 && j.hold->bytes) {
                std::lock_guard<std::mutex> g(pin_mu);
                pin.release(j.hold->off, j.hold->bytes);
            }
            // the request completes with its last part
            if (--r->parts_left > 0) continue;
            const double t = now_ms();
            {
                std::lock_guard<std::mutex> g(mu);
                r->t_done = t;
                r->state = kDone;
                latencies[r->kind].push_back(t - r->t_submit);
                ++completed;
            }
            cv_done.notify_all();
        }
    }

    // The worker. Submissions go to `pending` under mu; the worker moves them
    // to its own backlog (in arrival order) and works on that without the
    // lock, so a submitter never waits behind a scan of the queued requests.
    void run() {
        (void)hipSetDevice(device);
        std::deque<Request*> backlog;  // submitted and not yet launched, in arrival order
        std::vector<Request*> todo;
        for (;;) {
            {
                std::unique_lock<std::mutex> g(mu);
                if (pending.empty() && backlog.empty() && flights.empty()) {
                    cv_submit.wait_for(g, std::chrono::milliseconds(2), [&] { return stop || !pending.empty(); });
                }
                if (stop && pending.empty() && backlog.empty() && flights.empty()) return;
                if (backlog.empty()) {
                    backlog.swap(pending);
                } else {
                    backlog.insert(backlog.end(), pending.begin(), pending.end());
                    pending.clear();
                }
            }
            std::vector<Request*> admitted;
            if (!backlog.empty()) {
                todo.assign(backlog.begin(), backlog.end());
                dispatch(todo, admitted);
            }
            if (!admitted.empty()) {
                // admitted requests are no longer kPending; keep the rest in order
                std::deque<Request*> keep;
                for (Request* r : backlog)
                    if (r->state == kPending) keep.push_back(r);
                backlog.swap(keep);
            }
            if (reap() == 0 && admitted.empty()) std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
    }
};

}  // namespace afh

using namespace afh;

extern "C" {

int32_t AFS_Create(const AFS_Config* cfg, void** out) {
    clear_error();
    if (!cfg || !out || cfg->n_streams == 0 || cfg->n_streams > 16 || cfg->max_gspat_batch == 0 ||
        cfg->hbm_budget_bytes < 4096 || cfg->pinned_budget_bytes < 4096 ||
        cfg->gspat_precision < AFH_HOLO_FP32 || cfg->gspat_precision > AFH_HOLO_MIXED_SPLIT || cfg->gspat_graphs > 1) {
        set_error("AFS_Create: bad configuration");
        return AFH_ERR_ARG;
    }
    *out = nullptr;
    if (!device().ensure()) return AFH_ERR_NO_DEVICE;
    std::unique_ptr<Server> s(new Server());
    s->cfg = *cfg;
    (void)hipGetDevice(&s->device);
    if (!hip_ok(hipMalloc(&s->dev_base, cfg->hbm_budget_bytes), "AFS_Create: hipMalloc (HBM budget)")) return AFH_ERR_HIP;
    if (!hip_ok(hipHostMalloc(&s->pin_base, cfg->pinned_budget_bytes, hipHostMallocDefault),
                "AFS_Create: hipHostMalloc"))
        return AFH_ERR_HIP;
    s->dev.init(cfg->hbm_budget_bytes);
    s->pin.init(cfg->pinned_budget_bytes);
    // CU partition: the GS-PAT stream (0) on the first gspat_cus CUs, the field
    // streams on the others; plain streams when no partition is asked for or
    // the runtime refuses the masks
    int n_cu = 0;
    (void)hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, s->device);
    const bool part = cfg->gspat_cus > 0 && cfg->n_streams > 1 && n_cu > (int)cfg->gspat_cus && n_cu <= 1024;
    std::vector<uint32_t> gmask;  // the GS-PAT partition (empty: all CUs)
    auto make_stream = [&](bool gspat, hipStream_t* st) {
        bool made = false;
        if (part) {
            std::vector<uint32_t> mask((n_cu + 31) / 32, 0u);
            for (int c = 0; c < n_cu; ++c) {
                const bool mine = gspat ? c < (int)cfg->gspat_cus : c >= (int)cfg->gspat_cus;
                if (mine) mask[c / 32] |= 1u << (c % 32);
            }
            made = hipExtStreamCreateWithCUMask(st, (uint32_t)mask.size(), mask.data()) == hipSuccess;
            if (!made) (void)hipGetLastError();
            if (made && gspat) gmask = mask;
        }
        return made || hip_ok(hipStreamCreateWithFlags(st, hipStreamNonBlocking), "AFS_Create: stream");
    };
    for (uint32_t i = 0; i < cfg->n_streams; ++i) {
        hipStream_t st = nullptr;
        if (!make_stream(i == 0, &st)) return AFH_ERR_HIP;
        s->streams.push_back(st);
        s->outstanding.push_back(0.0);
        s->ctxs.emplace_back(new Context());
    }
    s->lanes_per_array = part ? 1 : 2;
    const int lanes = part ? Server::kGspatLanes / 2 : Server::kGspatLanes;
    for (int i = 1; i < lanes; ++i) {
        hipStream_t st = nullptr;
        if (!make_stream(true, &st)) return AFH_ERR_HIP;
        s->gextra.push_back(st);
        s->gextra_ctx.emplace_back(new Context());
    }
    s->ctxs[0]->cu_mask = gmask;
    for (auto& c : s->gextra_ctx) c->cu_mask = gmask;
    for (int i = 0; i < s->gspat_lanes(); ++i) s->lane_ctx(i)->graphs = cfg->gspat_graphs != 0;
    // the GS-PAT staging for a full batch is reserved now rather than grown
    // while serving; the workspace, which depends on the array size, when an
    // array is added
    for (int i = 0; i < s->gspat_lanes(); ++i)
        if (!s->lane_ctx(i)->staging.reserve((2 * (size_t)cfg->max_gspat_batch + 2) * 4 + 16)) return AFH_ERR_HIP;
    if (!dir_cells_device()) {
        set_error("AFS_Create: directivity table");
        return AFH_ERR_HIP;
    }
    Server* raw = s.get();
    s->worker = std::thread([raw] { raw->run(); });
    for (int i = 0; i < Server::kCopiers; ++i) s->copiers.emplace_back([raw] { raw->copy_loop(); });
    *out = s.release();
    return AFH_OK;
}

void AFS_Destroy(void* srv) { delete static_cast<Server*>(srv); }

int32_t AFS_AddArray(void* srv, const void* sources, const float* wavenums, const float* attens, uint64_t n,
                     int32_t source_type, uint32_t* array_id) {
    clear_error();
    Server* s = static_cast<Server*>(srv);
    if (!s || !sources || !wavenums || !attens || !array_id || n == 0 || n > 0xFFFFFFF ||
        (source_type != kSphere && source_type != kT4010A1)) {
        set_error("AFS_AddArray: bad argument");
        return AFH_ERR_ARG;
    }
    std::lock_guard<std::mutex> add_lock(s->add_mu);
    SrvArray A;
    A.source_type = source_type;
    A.n = (uint32_t)n;
    A.np = (uint32_t)packed_count(n);
    std::vector<float> host((size_t)A.np * AFH_SOURCE_STRIDE);
    pack_sources(static_cast<const float*>(sources), wavenums, attens, n, source_type, host.data());
    if (!hip_ok(hipMalloc(&A.packed, host.size() * sizeof(float)), "AFS_AddArray: hipMalloc")) return AFH_ERR_HIP;
    if (!hip_ok(hipMemcpy(A.packed, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice),
                "AFS_AddArray: upload")) {
        (void)hipFree(A.packed);
        return AFH_ERR_HIP;
    }
    {
        // the GS-PAT workspace for a full batch of 128-focus instances on this
        // array, reserved now rather than grown while serving (growing frees the
        // old block, so work queued on the GS-PAT stream must finish first)
        uint32_t next_id;
        {
            std::lock_guard<std::mutex> g(s->mu);
            next_id = (uint32_t)s->arrays.size();
        }
        const size_t need = gspat_workspace_bound(A.np, s->cfg.max_gspat_batch);
        for (int k = 0; k < s->lanes_per_array; ++k) {
            const int lane = s->array_lane(next_id, k);
            Context* ctx = s->lane_ctx(lane);
            std::lock_guard<std::mutex> g(ctx->mu);
            if (need > ctx->work.cap &&
                (!hip_ok(hipStreamSynchronize(s->lane_stream(lane)), "AFS_AddArray: hipStreamSynchronize") ||
                 !ctx->work.reserve(need))) {
                (void)hipFree(A.packed);
                return AFH_ERR_HIP;
            }
        }
    }
    std::lock_guard<std::mutex> g(s->mu);
    *array_id = (uint32_t)s->arrays.size();
    s->arrays.emplace_back(new SrvArray(A));
    return AFH_OK;
}

// True when [p, p + bytes) lies inside one pinned host allocation (HIP
// page-locked memory), which the DMA engine can write directly.
static bool pinned_range(const void* p, size_t bytes) {
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();  // pageable memory is not an error here
        return false;
    }
    if (at.type != hipMemoryTypeHost || !at.devicePointer) return false;
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, at.devicePointer) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    const char* b = static_cast<const char*>(base);
    const char* d = static_cast<const char*>(at.devicePointer);
    return d >= b && (size_t)(d - b) <= size && bytes <= size - (size_t)(d - b);
}

static int32_t enqueue(Server* s, Request* r, uint64_t* ticket) {
    {
        std::lock_guard<std::mutex> g(s->mu);
        r->ticket = s->next_ticket++;
        r->t_submit = now_ms();
        s->by_ticket[r->ticket] = r;
        s->pending.push_back(r);
    }
    s->cv_submit.notify_one();
    *ticket = r->ticket;
    return AFH_OK;
}

int32_t AFS_SubmitField(void* srv, uint32_t array_id, const AFH_PointSet* pts, int32_t field_type, float* out_host,
                        uint64_t* ticket) {
    clear_error();
    Server* s = static_cast<Server*>(srv);
    if (!s || !pts || !out_host || !ticket || field_type < 0 || field_type > 2) {
        set_error("AFS_SubmitField: bad argument");
        return AFH_ERR_ARG;
    }
    std::unique_ptr<Request> r(new Request());
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (array_id >= s->arrays.size()) {
            set_error("AFS_SubmitField: unknown array");
            return AFH_ERR_ARG;
        }
        r->arr = s->arrays[array_id].get();
    }
    r->kind = kReqField;
    r->array_id = array_id;
    r->field_type = field_type;
    r->out = out_host;
    if (pts->kind == 0) {
        GridArea& g = r->grid;
        for (int k = 0; k < 3; ++k) {
            if (pts->order[k] < 0 || pts->order[k] > 2) {
                set_error("AFS_SubmitField: bad grid order");
                return AFH_ERR_ARG;
            }
            g.order[k] = pts->order[k];
            g.bounds[pts->order[k]] = pts->count[k];
            g.origin[k] = pts->origin[k];
        }
        if (g.order[0] == g.order[1] || g.order[1] == g.order[2] || g.order[0] == g.order[2]) {
            set_error("AFS_SubmitField: grid order must be a permutation");
            return AFH_ERR_ARG;
        }
        g.res = pts->resolution;
        r->n_points = (uint64_t)pts->count[0] * pts->count[1] * pts->count[2];
    } else if (pts->kind == 1 && (pts->points || pts->n_points == 0)) {
        r->scatter = true;
        r->n_points = pts->n_points;
        r->points.assign(pts->points, pts->points + 3 * pts->n_points);  // host points
    } else {
        set_error("AFS_SubmitField: bad point set");
        return AFH_ERR_ARG;
    }
    const size_t per = field_type == AFC_FIELD_COMPLEX ? 2 : 1;
    const uint64_t need = r->n_points * per * sizeof(float) + (r->scatter ? r->n_points * 12 : 0) + 512;
    if (r->n_points == 0 || r->n_points > kMaxPointsPerCall || need > s->cfg.hbm_budget_bytes ||
        need > s->cfg.pinned_budget_bytes) {
        set_error("AFS_SubmitField: request does not fit the HBM / pinned budget (or has no points)");
        return AFH_ERR_BUDGET;
    }
    r->direct = pinned_range(out_host, r->n_points * per * sizeof(float));
    return enqueue(s, r.release(), ticket);
}

int32_t AFS_SubmitGSPAT(void* srv, uint32_t array_id, const float* foci, const float* amps, uint32_t m,
                        float* out_amp, float* out_phase, uint64_t* ticket) {
    clear_error();
    Server* s = static_cast<Server*>(srv);
    if (!s || !foci || !amps || !out_amp || !out_phase || !ticket || m == 0 || m > 128) {
        set_error("AFS_SubmitGSPAT: bad argument (1 to 128 foci)");
        return AFH_ERR_ARG;
    }
    std::unique_ptr<Request> r(new Request());
    uint32_t n;
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (array_id >= s->arrays.size()) {
            set_error("AFS_SubmitGSPAT: unknown array");
            return AFH_ERR_ARG;
        }
        r->arr = s->arrays[array_id].get();
        n = r->arr->n;
    }
    const uint64_t need = (uint64_t)m * 16 + 2ull * n * 4 + 1024;
    if (need > s->cfg.hbm_budget_bytes || need > s->cfg.pinned_budget_bytes) {
        set_error("AFS_SubmitGSPAT: request does not fit the budget");
        return AFH_ERR_BUDGET;
    }
    r->kind = kReqGSPAT;
    r->array_id = array_id;
    r->m = m;
    r->foci.assign(foci, foci + 3 * (size_t)m);
    r->amps.assign(amps, amps + m);
    r->out = out_amp;
    r->out2 = out_phase;
    return enqueue(s, r.release(), ticket);
}

int32_t AFS_SubmitGSPATBatch(void* srv, uint32_t count, const uint32_t* array_ids, const float* const* foci,
                             const float* const* amps, const uint32_t* m, float* const* out_amp,
                             float* const* out_phase, uint64_t* tickets) {
    clear_error();
    Server* s = static_cast<Server*>(srv);
    if (!s || (count > 0 && (!array_ids || !foci || !amps || !m || !out_amp || !out_phase || !tickets))) {
        set_error("AFS_SubmitGSPATBatch: bad argument");
        return AFH_ERR_ARG;
    }
    std::vector<std::unique_ptr<Request>> rs(count);
    {
        std::lock_guard<std::mutex> g(s->mu);
        for (uint32_t i = 0; i < count; ++i) {
            if (!foci[i] || !amps[i] || !out_amp[i] || !out_phase[i] || m[i] == 0 || m[i] > 128) {
                set_error("AFS_SubmitGSPATBatch: bad argument in request " + std::to_string(i) + " (1 to 128 foci)");
                return AFH_ERR_ARG;
            }
            if (array_ids[i] >= s->arrays.size()) {
                set_error("AFS_SubmitGSPATBatch: unknown array in request " + std::to_string(i));
                return AFH_ERR_ARG;
            }
            rs[i].reset(new Request());
            rs[i]->arr = s->arrays[array_ids[i]].get();
        }
    }
    for (uint32_t i = 0; i < count; ++i) {
        Request* r = rs[i].get();
        const uint64_t need = (uint64_t)m[i] * 16 + 2ull * r->arr->n * 4 + 1024;
        if (need > s->cfg.hbm_budget_bytes || need > s->cfg.pinned_budget_bytes) {
            set_error("AFS_SubmitGSPATBatch: request " + std::to_string(i) + " does not fit the budget");
            return AFH_ERR_BUDGET;
        }
        r->kind = kReqGSPAT;
        r->array_id = array_ids[i];
        r->m = m[i];
        r->foci.assign(foci[i], foci[i] + 3 * (size_t)m[i]);
        r->amps.assign(amps[i], amps[i] + m[i]);
        r->out = out_amp[i];
        r->out2 = out_phase[i];
    }
    {
        // one lock and one wake-up for the whole batch, in the given order
        std::lock_guard<std::mutex> g(s->mu);
        const double t = now_ms();
        for (uint32_t i = 0; i < count; ++i) {
            Request* r = rs[i].release();
            r->ticket = s->next_ticket++;
            r->t_submit = t;
            s->by_ticket[r->ticket] = r;
            s->pending.push_back(r);
            tickets[i] = r->ticket;
        }
    }
    if (count > 0) s->cv_submit.notify_one();
    return AFH_OK;
}

int32_t AFS_Wait(void* srv, uint64_t ticket, double* latency_ms) {
    clear_error();
    Server* s = static_cast<Server*>(srv);
    if (!s) return AFH_ERR_ARG;
    std::unique_lock<std::mutex> g(s->mu);
    auto it = s->by_ticket.find(ticket);
    if (it == s->by_ticket.end()) {
        set_error("AFS_Wait: unknown ticket");
        return AFH_ERR_ARG;
    }
    Request* r = it->second;
    s->cv_done.wait(g, [&] { return r->state == kDone; });
    if (latency_ms) *latency_ms = r->t_done - r->t_submit;
    const int32_t st = r->status;
    s->by_ticket.erase(it);
    delete r;
    return st;
}

int32_t AFS_Stats(void* srv, int32_t kind, double* p50_ms, double* p99_ms, uint64_t* completed, int32_t reset) {
    Server* s = static_cast<Server*>(srv);
    if (!s || kind < -1 || kind > 1) return AFH_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    std::vector<double> v;
    for (int k = 0; k < 2; ++k)
        if (kind < 0 || kind == k) v.insert(v.end(), s->latencies[k].begin(), s->latencies[k].end());
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) { return v.empty() ? 0.0 : v[std::min(v.size() - 1, (size_t)(q * (v.size() - 1) + 0.5))]; };
    if (p50_ms) *p50_ms = pct(0.50);
    if (p99_ms) *p99_ms = pct(0.99);
    if (completed) *completed = v.size();
    if (reset) {
        s->latencies[0].clear();
        s->latencies[1].clear();
        s->completed = 0;
    }
    return AFH_OK;
}

int32_t AFS_GraphCounters(void* srv, uint64_t* launches, uint64_t* builds) {
    Server* s = static_cast<Server*>(srv);
    if (!s || !launches || !builds) return AFH_ERR_ARG;
    *launches = *builds = 0;
    for (int i = 0; i < s->gspat_lanes(); ++i) {
        Context* c = s->lane_ctx(i);
        std::lock_guard<std::mutex> g(c->mu);
        *launches += c->graph_launches;
        *builds += c->graph_builds;
    }
    return AFH_OK;
}

}  // extern "C"
