// Serving loop: resident grids, HBM-budget admission and continuous batching.
//
// Requests are queued FIFO. A dispatcher thread owns `streams` launch slots (one
// HIP stream each, default priority): whenever a slot is free it packs queued
// requests whose outputs fit in the remaining budget into ONE batched launch
// (wx_trace_batch), so requests that arrive while other launches run on the GPU
// join the next launch and overlap with them on the device. One completion
// thread per stream waits for its launches in order and timestamps the requests.
// Output buffers come from one device arena allocated at creation and
// sub-allocated first-fit, so serving never calls hipMalloc.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "wx_internal.hpp"

using wx::set_error;

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over [0, cap) in 256-byte units.
class Arena {
  public:
    void init(uint64_t cap) {
        free_.clear();
        if (cap) free_[0] = cap;
    }
    static uint64_t round(uint64_t n) { return (n + 255u) & ~uint64_t(255); }
    bool alloc(uint64_t n, uint64_t *off) {
        n = round(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            const uint64_t rest = it->second - n, start = it->first + n;
            free_.erase(it);
            if (rest) free_[start] = rest;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t n) {
        n = round(n);
        auto it = free_.emplace(off, n).first;
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

  private:
    std::map<uint64_t, uint64_t> free_;
};

enum JobState { Q_QUEUED, Q_LAUNCHED, Q_DONE, Q_CANCELLED };

struct Job {
    wx_serve_request req;
    uint64_t bytes = 0, off = 0;
    JobState state = Q_QUEUED;
    double t_submit = 0, t_launch = 0, t_done = 0;
    uint32_t batch_n = 0;
    int status = WX_OK;
};

struct Batch {
    std::vector<uint64_t> ids;
    hipEvent_t ev = nullptr;
    int status = WX_OK;
};

struct Slot {
    hipStream_t stream = nullptr;
    std::deque<Batch> inflight;  // at most one launched batch at a time
};

constexpr size_t kMaxBatchGrids = 64;               // grids one batched launch takes
constexpr uint64_t kMaxBatchSlots = 0x7FFFFFFFull;  // ray slots one batched launch takes

uint64_t job_bytes(const wx_serve_request &r) {
    const uint64_t px = (uint64_t)r.width * r.height;
    return px * 32u * ((r.flags & WX_TRACE_SHADOW) ? 2u : 1u);
}

// Latencies in log-spaced bins: bin k (1 <= k < kBins-1) holds [kLo * r^(k-1),
// kLo * r^k) ms with r = 1.01; bin 0 everything below kLo, the last bin everything
// above. Fixed size, so the record never grows; a percentile is read to within
// one bin (1 %).
constexpr double kLatLoMs = 1e-3, kLatRatio = 1.01;
constexpr size_t kLatBins = 2086;  // up to ~1e6 ms

struct LatencyHist {
    uint64_t count[kLatBins] = {};
    uint64_t n = 0;
    double sum = 0.0;

    void add(double ms) {
        size_t k = 0;
        if (ms >= kLatLoMs) k = std::min(kLatBins - 1, (size_t)(std::log(ms / kLatLoMs) / std::log(kLatRatio)) + 1u);
        count[k]++;
        n++;
        sum += ms;
    }
    // The geometric centre of the bin holding the q-quantile.
    double percentile(double q) const {
        if (!n) return 0.0;
        const uint64_t target = std::min<uint64_t>(n - 1u, (uint64_t)(q * (double)(n - 1u)));
        uint64_t seen = 0;
        for (size_t k = 0; k < kLatBins; ++k) {
            seen += count[k];
            if (seen > target) {
                if (k == 0) return kLatLoMs;
                return kLatLoMs * std::pow(kLatRatio, (double)k - 0.5);
            }
        }
        return kLatLoMs * std::pow(kLatRatio, (double)kLatBins);
    }
};

}  // namespace

struct wx_server {
    wx_server_config cfg{};
    bool dry = false;
    int device = 0;
    std::vector<wx_grid *> grids;
    char *arena_dev = nullptr;
    Arena arena;
    uint64_t in_use = 0, peak = 0;
    std::vector<Slot> slots;

    std::mutex mu;
    std::condition_variable cv_dispatch, cv_done, cv_slot;
    std::deque<uint64_t> queue;
    std::unordered_map<uint64_t, Job> jobs;
    uint64_t next_id = 1;
    bool stop = false;

    wx_server_stats st{};
    LatencyHist lat;
    double batch_req_sum = 0;

    std::thread dispatcher;
    std::vector<std::thread> completers;

    void dispatch_loop();
    void complete_loop(size_t k);
};

void wx_server::dispatch_loop() {
    if (!dry) (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        if (stop) {  // shutting down: requests not yet launched are cancelled
            for (uint64_t id : queue) {
                jobs[id].state = Q_CANCELLED;
                st.cancelled++;
            }
            queue.clear();
            cv_done.notify_all();
            cv_slot.notify_all();
            break;
        }
        size_t free_slot = slots.size();
        for (size_t k = 0; k < slots.size(); ++k)
            if (slots[k].inflight.empty()) {
                free_slot = k;
                break;
            }
        if (queue.empty() || free_slot == slots.size()) {
            cv_dispatch.wait(lk);
            continue;
        }
        // pack: FIFO while the outputs fit and the batch limits allow -- at most
        // max_batch_requests requests, max_batch_rays primary rays, 64 distinct grids
        // and 2^31 ray slots (64 per 8x8 tile), the limits of one batched launch
        std::vector<uint64_t> ids;
        std::vector<uint32_t> seen_grids;
        uint64_t rays = 0, ray_slots = 0;
        while (!queue.empty() && ids.size() < cfg.max_batch_requests) {
            Job &j = jobs[queue.front()];
            const uint64_t r = (uint64_t)j.req.width * j.req.height;
            if (!ids.empty() && cfg.max_batch_rays && rays + r > cfg.max_batch_rays) break;
            const uint64_t sl = (uint64_t)((j.req.width + 7u) / 8u) * ((j.req.height + 7u) / 8u) * 64u;
            if (ray_slots + sl > kMaxBatchSlots) break;  // one request alone always fits (65536^2 pixels)
            const bool new_grid =
                std::find(seen_grids.begin(), seen_grids.end(), j.req.grid) == seen_grids.end();
            if (new_grid && seen_grids.size() >= kMaxBatchGrids) break;
            uint64_t off = 0;
            if (!arena.alloc(j.bytes, &off)) break;  // wait for memory (keeps FIFO order)
            j.off = off;
            in_use += Arena::round(j.bytes);
            peak = std::max(peak, in_use);
            ids.push_back(queue.front());
            queue.pop_front();
            rays += r;
            ray_slots += sl;
            if (new_grid) seen_grids.push_back(j.req.grid);
        }
        if (ids.empty()) {  // the head request waits for memory to be released
            cv_dispatch.wait(lk);
            continue;
        }
        // grid ids of this batch -> compact launch-local indices
        std::vector<wx_grid *> bg;
        std::unordered_map<uint32_t, uint32_t> gmap;
        std::vector<wx_request> reqs(ids.size());
        const double t = now_s();
        for (size_t i = 0; i < ids.size(); ++i) {
            Job &j = jobs[ids[i]];
            j.state = Q_LAUNCHED;
            j.t_launch = t;
            j.batch_n = (uint32_t)ids.size();
            auto it = gmap.find(j.req.grid);
            if (it == gmap.end()) {
                it = gmap.emplace(j.req.grid, (uint32_t)bg.size()).first;
                bg.push_back(grids[j.req.grid]);
            }
            wx_request &r = reqs[i];
            std::memset(&r, 0, sizeof r);
            r.kind = WX_REQ_CAMERA;
            r.grid = it->second;
            r.width = j.req.width;
            r.height = j.req.height;
            r.flags = j.req.flags;
            r.pixel_offset = j.req.pixel_offset;
            r.camera = j.req.camera;
            r.light[0] = j.req.light[0], r.light[1] = j.req.light[1], r.light[2] = j.req.light[2];
            const uint64_t px = (uint64_t)j.req.width * j.req.height;
            r.out = reinterpret_cast<wx_hit *>(arena_dev + j.off);
            r.shadow_out = (j.req.flags & WX_TRACE_SHADOW) ? reinterpret_cast<wx_hit *>(arena_dev + j.off + px * 32u)
                                                           : nullptr;
        }
        Batch b;
        b.ids = ids;
        Slot &sl = slots[free_slot];
        sl.inflight.push_back(Batch());  // reserve the slot while unlocked
        st.batches++;
        batch_req_sum += (double)ids.size();
        lk.unlock();
        if (!dry) {
            b.status = wx_trace_batch(bg.data(), (uint32_t)bg.size(), reqs.data(), (uint32_t)reqs.size(), nullptr,
                                      sl.stream);
            if (hipEventCreateWithFlags(&b.ev, hipEventDisableTiming) == hipSuccess)
                (void)hipEventRecord(b.ev, sl.stream);
            else
                b.status = set_error(WX_ERR_DEVICE, "server: event create");
        }
        lk.lock();
        sl.inflight.back() = std::move(b);
        cv_slot.notify_all();
    }
}

void wx_server::complete_loop(size_t k) {
    if (!dry) (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    Slot &sl = slots[k];
    for (;;) {
        // a batch is ready once the dispatcher has filled it in (ids non-empty)
        while (!(!sl.inflight.empty() && !sl.inflight.front().ids.empty()) && !(stop && sl.inflight.empty()))
            cv_slot.wait(lk);
        if (sl.inflight.empty()) break;  // stopping and nothing in flight
        if (sl.inflight.front().ids.empty()) continue;
        Batch &b = sl.inflight.front();
        hipEvent_t ev = b.ev;
        int status = b.status;
        lk.unlock();
        if (ev) {
            if (hipEventSynchronize(ev) != hipSuccess && status == WX_OK) status = WX_ERR_DEVICE;
            (void)hipEventDestroy(ev);
        }
        const double t = now_s();
        lk.lock();
        for (uint64_t id : sl.inflight.front().ids) {
            Job &j = jobs[id];
            j.state = Q_DONE;
            j.t_done = t;
            j.status = status;
            st.completed++;
            st.primary_rays += (uint64_t)j.req.width * j.req.height;
            lat.add((t - j.t_submit) * 1e3);
        }
        sl.inflight.pop_front();
        cv_done.notify_all();
        cv_dispatch.notify_all();
        cv_slot.notify_all();
    }
}

extern "C" {

int wx_server_create(const wx_server_config *cfg, wx_server **out) {
    if (!cfg || !out) return set_error(WX_ERR_INVALID, "wx_server_create: NULL argument");
    if (cfg->streams < 1 || cfg->streams > 16) return set_error(WX_ERR_INVALID, "wx_server_create: streams in 1..16");
    if (cfg->max_batch_requests < 1 || cfg->max_batch_requests > 1024)
        return set_error(WX_ERR_INVALID, "wx_server_create: max_batch_requests in 1..1024");
    if (cfg->flags & ~WX_SERVER_DRYRUN) return set_error(WX_ERR_INVALID, "wx_server_create: unknown flags");
    wx_server *s = new (std::nothrow) wx_server();
    if (!s) return set_error(WX_ERR_BUDGET, "wx_server_create: host allocation failed");
    s->cfg = *cfg;
    s->dry = (cfg->flags & WX_SERVER_DRYRUN) != 0u;
    s->arena.init(cfg->hbm_budget);
    s->slots.resize(cfg->streams);
    if (!s->dry) {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
            delete s;
            return set_error(WX_ERR_DEVICE, "wx_server_create: no HIP device");
        }
        (void)hipGetDevice(&s->device);
        if (cfg->hbm_budget) {
            hipError_t e = hipMalloc(reinterpret_cast<void **>(&s->arena_dev), cfg->hbm_budget);
            if (e != hipSuccess) {
                delete s;
                return set_error(WX_ERR_BUDGET, std::string("wx_server_create: arena: ") + hipGetErrorString(e));
            }
        }
        for (Slot &sl : s->slots) {
            if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess) {
                for (Slot &o : s->slots)
                    if (o.stream) (void)hipStreamDestroy(o.stream);
                if (s->arena_dev) (void)hipFree(s->arena_dev);
                delete s;
                return set_error(WX_ERR_DEVICE, "wx_server_create: stream create");
            }
        }
    }
    s->dispatcher = std::thread([s] { s->dispatch_loop(); });
    for (size_t k = 0; k < s->slots.size(); ++k) s->completers.emplace_back([s, k] { s->complete_loop(k); });
    *out = s;
    return WX_OK;
}

int wx_server_destroy(wx_server *s) {
    if (!s) return WX_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_dispatch.notify_all();
    s->cv_slot.notify_all();
    s->dispatcher.join();
    s->cv_slot.notify_all();
    for (auto &t : s->completers) t.join();
    for (Slot &sl : s->slots)
        if (sl.stream) {
            wx::drop_stream_scratch(sl.stream);
            (void)hipStreamDestroy(sl.stream);
        }
    if (s->arena_dev) (void)hipFree(s->arena_dev);
    delete s;
    return WX_OK;
}

int wx_server_add_grid(wx_server *s, wx_grid *g, uint32_t *id) {
    if (!s || (!g && !s->dry) || !id) return set_error(WX_ERR_INVALID, "wx_server_add_grid: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *id = (uint32_t)s->grids.size();
    s->grids.push_back(g);
    return WX_OK;
}

int wx_server_submit(wx_server *s, const wx_serve_request *r, uint64_t *ticket) {
    if (!s || !r || !ticket) return set_error(WX_ERR_INVALID, "wx_server_submit: NULL argument");
    if (r->flags & ~(WX_TRACE_CLIP | WX_TRACE_SHADOW)) return set_error(WX_ERR_INVALID, "wx_server_submit: bad flags");
    if (r->width == 0 || r->height == 0 || r->width > 65536u || r->height > 65536u)
        return set_error(WX_ERR_INVALID, "wx_server_submit: image size");
    std::lock_guard<std::mutex> lk(s->mu);
    if (r->grid >= s->grids.size()) return set_error(WX_ERR_INVALID, "wx_server_submit: unknown grid id");
    if (s->stop) return set_error(WX_ERR_STATE, "wx_server_submit: server is shutting down");
    if (s->cfg.max_queued && s->queue.size() >= s->cfg.max_queued) {
        s->st.rejected++;
        return set_error(WX_ERR_BUSY, "wx_server_submit: request queue full");
    }
    Job j;
    j.req = *r;
    j.bytes = job_bytes(*r);
    s->st.submitted++;
    if (Arena::round(j.bytes) > s->cfg.hbm_budget) {
        s->st.rejected++;
        return set_error(WX_ERR_BUDGET, "wx_server_submit: outputs exceed the whole HBM budget");
    }
    j.t_submit = now_s();
    const uint64_t id = s->next_id++;
    s->jobs.emplace(id, j);
    s->queue.push_back(id);
    *ticket = id;
    s->cv_dispatch.notify_all();
    return WX_OK;
}

int wx_server_wait(wx_server *s, uint64_t ticket, double timeout_s, wx_serve_result *out) {
    if (!s || !out) return set_error(WX_ERR_INVALID, "wx_server_wait: NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] {
        auto it = s->jobs.find(ticket);
        return it == s->jobs.end() || it->second.state == Q_DONE || it->second.state == Q_CANCELLED;
    };
    if (timeout_s < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
        return set_error(WX_ERR_STATE, "wx_server_wait: timed out");
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return set_error(WX_ERR_INVALID, "wx_server_wait: unknown ticket");
    const Job &j = it->second;
    if (j.state == Q_CANCELLED) return set_error(WX_ERR_STATE, "wx_server_wait: request was cancelled");
    std::memset(out, 0, sizeof *out);
    const uint64_t px = (uint64_t)j.req.width * j.req.height;
    out->primary = s->arena_dev ? reinterpret_cast<const wx_hit *>(s->arena_dev + j.off) : nullptr;
    out->shadow = (s->arena_dev && (j.req.flags & WX_TRACE_SHADOW))
                      ? reinterpret_cast<const wx_hit *>(s->arena_dev + j.off + px * 32u)
                      : nullptr;
    out->width = j.req.width;
    out->height = j.req.height;
    out->submit_s = j.t_submit;
    out->launch_s = j.t_launch;
    out->done_s = j.t_done;
    out->batch_requests = j.batch_n;
    out->status = j.status;
    return WX_OK;
}

int wx_server_release(wx_server *s, uint64_t ticket) {
    if (!s) return set_error(WX_ERR_INVALID, "wx_server_release: NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return set_error(WX_ERR_INVALID, "wx_server_release: unknown ticket");
    if (it->second.state == Q_QUEUED || it->second.state == Q_LAUNCHED)
        return set_error(WX_ERR_STATE, "wx_server_release: request not done");
    if (it->second.state == Q_DONE) {
        s->arena.release(it->second.off, it->second.bytes);
        s->in_use -= Arena::round(it->second.bytes);
    }
    s->jobs.erase(it);
    s->cv_dispatch.notify_all();
    return WX_OK;
}

int wx_server_cancel(wx_server *s, uint64_t ticket) {
    if (!s) return set_error(WX_ERR_INVALID, "wx_server_cancel: NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return set_error(WX_ERR_INVALID, "wx_server_cancel: unknown ticket");
    if (it->second.state != Q_QUEUED) return set_error(WX_ERR_STATE, "wx_server_cancel: already launched");
    s->queue.erase(std::find(s->queue.begin(), s->queue.end(), ticket));
    it->second.state = Q_CANCELLED;
    s->st.cancelled++;
    s->cv_done.notify_all();
    s->cv_dispatch.notify_all();
    return WX_OK;
}

int wx_server_get_stats(wx_server *s, wx_server_stats *out) {
    if (!s || !out) return set_error(WX_ERR_INVALID, "wx_server_get_stats: NULL argument");
    // copy the counters and the fixed-size histogram under the lock, read it outside
    std::unique_ptr<LatencyHist> h(new (std::nothrow) LatencyHist());
    if (!h) return set_error(WX_ERR_BUDGET, "wx_server_get_stats: host allocation failed");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        *out = s->st;
        *h = s->lat;
        out->mean_batch_requests = s->st.batches ? s->batch_req_sum / (double)s->st.batches : 0.0;
        out->budget_in_use = s->in_use;
        out->budget_peak = s->peak;
    }
    out->p50_ms = h->percentile(0.50);
    out->p99_ms = h->percentile(0.99);
    out->mean_ms = h->n ? h->sum / (double)h->n : 0.0;
    return WX_OK;
}

}  // extern "C"
