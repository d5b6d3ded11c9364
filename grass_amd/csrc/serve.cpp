// Serving session: continuous batching of random-walk and k-hop requests on one resident graph.
//
// A scheduler thread owns `max_inflight` launch-group slots, each with its own stream, event,
// workspaces and pinned staging.  Whenever a slot is free and requests are queued, it takes
// queued requests in arrival order while their HBM fits the session's arena (memory-based
// admission: inputs, walk matrices and the group's k-hop output), launches them as one group
// (one persistent random-walk launch and one k-hop pipeline) and goes back to watching events.
// Requests that arrive while groups run are therefore admitted as soon as a slot and memory
// free up, and their work overlaps the running groups on the device.  Results stay in HBM until
// the client releases them.  Every random choice is keyed by (seed, instance, ...), so a
// request's output does not depend on its group.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
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

#include "runtime.h"

namespace grass {
namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator over one device region; 256-byte granules, neighbours coalesced on free.
class Arena {
  public:
    hipError_t init(size_t bytes) {
        bytes &= ~(size_t)255;
        hipError_t e = hipMalloc(&base_, bytes);
        if (e != hipSuccess) return e;
        size_ = bytes;
        free_[0] = bytes;
        return hipSuccess;
    }
    ~Arena() {
        if (base_) (void)hipFree(base_);
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            size_t rest = it->second - bytes;
            size_t roff = it->first + bytes;
            free_.erase(it);
            if (rest) free_[roff] = rest;
            in_use += bytes;
            peak = std::max(peak, in_use);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
        in_use -= bytes;
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
    char* at(size_t off) { return (char*)base_ + off; }
    size_t size() const { return size_; }
    size_t in_use = 0, peak = 0;

  private:
    void* base_ = nullptr;
    size_t size_ = 0;
    std::map<size_t, size_t> free_;
};

struct OutBlock {  // a group's k-hop output, shared by its k-hop requests
    size_t off = 0, bytes = 0;
    int refs = 0;
};

enum class St { Pending, Running, Done };

struct Req {
    uint64_t ticket = 0;
    grass_request q{};
    std::vector<uint32_t> data;  // host copy of starts / seeds
    St state = St::Pending;
    int status = GRASS_OK;
    bool cancelled = false, auto_release = false;
    size_t off = 0, bytes = 0;  // own arena region (inputs + walk matrix)
    bool has_mem = false;
    std::shared_ptr<OutBlock> out;
    uint32_t kidx = 0;  // index among its group's k-hop requests
    grass_result res{};
    double arrival_us = 0;
    uint64_t edges = 0;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    Workspace ws_rw, ws_kh;
    PinnedBuf stage;   // request inputs on their way to HBM
    PinnedBuf offs;    // k-hop offsets read back
    bool busy = false;
    double launch_us = 0;
    std::vector<Req*> reqs;
    std::vector<Req*> kreqs;
    std::shared_ptr<OutBlock> out;
    uint32_t H = 0;
    char* out_base = nullptr;  // k-hop block: src[cap], dst[cap], offsets
    uint64_t cap = 0;
    int launch_status = GRASS_OK;
};

uint64_t rw_bytes(const grass_request& q) {
    return (q.starts ? q.walkers * 4 : 0) + q.walkers * ((uint64_t)q.length + 1) * 4;
}

}  // namespace
}  // namespace grass

using namespace grass;

struct grass_session {
    grass_graph* g = nullptr;
    grass_session_config cfg{};
    Arena arena;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req*> pending;
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
    uint64_t next_ticket = 1;
    bool stop = false;
    bool admissible = true;  // false after an admission failed for memory, until a release
    std::thread worker;
    Clock::time_point t0;
    grass_session_counters stats{};

    double now_us() const {
        return std::chrono::duration<double, std::micro>(Clock::now() - t0).count();
    }

    // Bytes a request needs of its own (k-hop outputs are counted in the group block).
    uint64_t own_bytes(const grass_request& q) const {
        if (q.kind == GRASS_REQ_RW) return rw_bytes(q);
        return q.num_seeds * 4;
    }
    uint64_t khop_out_bytes(const grass_request& q) const {
        grass_khop_request k{};
        k.num_seeds = q.num_seeds;
        k.seeds_dev = (const uint32_t*)16;  // only the sizes matter to the bound
        k.hops = q.hops;
        memcpy(k.fanout, q.fanout, sizeof(k.fanout));
        uint64_t e = 0;
        if (khop_bound(g, &k, 1, &e, nullptr, nullptr, nullptr)) return ~0ull;
        return e * 8;
    }

    void free_req_mem(Req* r) {
        if (r->has_mem) {
            arena.release(r->off, r->bytes);
            r->has_mem = false;
        }
        if (r->out) {
            if (--r->out->refs == 0) arena.release(r->out->off, r->out->bytes);
            r->out.reset();
        }
        stats.memory_in_use = arena.in_use;
        admissible = true;  // memory came back: queued requests may fit now
        cv_work.notify_one();
    }

    void finish(Req* r, int status) {
        r->state = St::Done;
        r->status = r->cancelled ? GRASS_ECANCELLED : status;
        r->res.status = r->status;
        r->res.done_us = now_us();
        if (r->status == GRASS_OK) {
            stats.completed++;
            stats.sampled_edges += r->edges;
        } else if (r->cancelled) {
            stats.cancelled++;
        }
        if (r->cancelled || r->auto_release) free_req_mem(r);
        if (r->cancelled) reqs.erase(r->ticket);
    }

    void finalize(Slot& s) {
        const uint64_t* off = (const uint64_t*)s.offs.ptr;
        const uint32_t K = (uint32_t)s.kreqs.size();
        for (Req* r : s.reqs) {
            if (r->q.kind == GRASS_REQ_KHOP && s.launch_status == GRASS_OK) {
                r->res.hops = r->q.hops;
                r->edges = 0;
                for (uint32_t h = 0; h < r->q.hops && h < s.H; ++h) {
                    const uint64_t a = off[(uint64_t)h * (K + 1) + r->kidx], b = off[(uint64_t)h * (K + 1) + r->kidx + 1];
                    r->res.hop_src_dev[h] = (const uint32_t*)s.out_base + a;
                    r->res.hop_dst_dev[h] = (const uint32_t*)s.out_base + s.cap + a;
                    r->res.hop_edges[h] = b - a;
                    r->edges += b - a;
                }
            } else if (r->q.kind == GRASS_REQ_RW) {
                r->edges = r->q.walkers * r->q.length;
            }
            finish(r, s.launch_status);
        }
        s.reqs.clear();
        s.kreqs.clear();
        s.out.reset();
        s.busy = false;
    }

    // k-hop block of a group: src[cap], dst[cap] (cap = sum of the requests' edge bounds), then
    // the hop-major offsets (at most GRASS_KHOP_MAX_HOPS x (K + 1)).
    static uint64_t block_bytes(uint64_t kout, uint32_t K) {
        return kout + 8ull * GRASS_KHOP_MAX_HOPS * (K + 1);
    }

    // Forms and launches one group on slot s; false when nothing could be admitted.
    bool launch(Slot& s) {
        std::vector<Req*> take;
        uint64_t kout = 0;
        uint32_t K = 0;
        while (!pending.empty() && take.size() < cfg.max_batch) {
            Req* r = pending.front();
            const uint64_t own = own_bytes(r->q);
            const bool kreq = r->q.kind == GRASS_REQ_KHOP;
            const uint64_t extra = kreq ? khop_out_bytes(r->q) : 0;
            size_t off = 0;
            if (!arena.alloc(own, &off)) break;
            // the group's k-hop block is allocated after the loop: check there is room for it
            if (kreq) {
                size_t probe = 0;
                const uint64_t bb = block_bytes(kout + extra, K + 1);
                if (!arena.alloc(bb, &probe)) {
                    arena.release(off, own);
                    break;
                }
                arena.release(probe, bb);
            }
            r->off = off;
            r->bytes = own;
            r->has_mem = true;
            kout += extra;
            K += kreq;
            take.push_back(r);
            pending.pop_front();
        }
        // allocate the k-hop block; should the requests' own regions have taken its room,
        // hand k-hop requests back (newest first) until it fits
        std::shared_ptr<OutBlock> blk;
        while (K) {
            blk = std::make_shared<OutBlock>();
            blk->bytes = block_bytes(kout, K);
            if (arena.alloc(blk->bytes, &blk->off)) break;
            blk.reset();
            for (auto it = take.rbegin(); it != take.rend(); ++it) {
                if ((*it)->q.kind != GRASS_REQ_KHOP) continue;
                Req* r = *it;
                kout -= khop_out_bytes(r->q);
                --K;
                free_req_mem(r);
                pending.push_front(r);
                take.erase(std::next(it).base());
                break;
            }
        }
        if (take.empty()) {
            admissible = false;  // nothing fits until memory is released
            return false;
        }
        s.reqs = take;
        s.kreqs.clear();
        size_t n_in = 0;
        for (Req* r : take) {
            n_in += r->data.size();
            if (r->q.kind == GRASS_REQ_KHOP) {
                r->kidx = (uint32_t)s.kreqs.size();
                s.kreqs.push_back(r);
            }
        }
        s.H = 0;
        for (Req* r : s.kreqs) s.H = std::max(s.H, r->q.hops);
        s.out.reset();
        if (K) {
            blk->refs = (int)K;
            s.out = blk;
            for (Req* r : s.kreqs) r->out = blk;
        }
        stats.memory_in_use = arena.in_use;
        stats.memory_peak = arena.peak;
        stats.groups++;
        const double t = now_us();
        s.launch_us = t;
        for (Req* r : take) {
            r->state = St::Running;
            r->res.launch_us = t;
        }
        s.launch_status = enqueue(s, n_in, kout);
        if (hipEventRecord(s.done, s.stream) != hipSuccess && s.launch_status == GRASS_OK)
            s.launch_status = GRASS_EDEVICE;
        s.busy = true;
        return true;
    }

    int enqueue(Slot& s, size_t n_in, uint64_t kout) {
        // inputs: pinned staging -> each request's region
        if (n_in && s.stage.reserve(n_in * 4) != hipSuccess) return GRASS_ENOMEM;
        size_t pos = 0;
        std::vector<grass_rw_request> rw;
        std::vector<grass_khop_request> kh(s.kreqs.size());
        for (Req* r : s.reqs) {
            uint32_t* in_dev = (uint32_t*)arena.at(r->off);
            if (!r->data.empty()) {
                uint32_t* st = (uint32_t*)s.stage.ptr + pos;
                memcpy(st, r->data.data(), r->data.size() * 4);
                if (hipMemcpyAsync(in_dev, st, r->data.size() * 4, hipMemcpyHostToDevice, s.stream) != hipSuccess)
                    return GRASS_EDEVICE;
                pos += r->data.size();
            }
            if (r->q.kind == GRASS_REQ_RW) {
                uint32_t* walks = in_dev + (r->q.starts ? r->q.walkers : 0);
                r->res.kind = GRASS_REQ_RW;
                r->res.walks_dev = walks;
                r->res.walkers = r->q.walkers;
                r->res.length = r->q.length;
                rw.push_back(grass_rw_request{r->q.walkers, r->q.length, r->q.instance,
                                              r->q.starts ? in_dev : nullptr, walks});
            } else {
                r->res.kind = GRASS_REQ_KHOP;
                grass_khop_request& k = kh[r->kidx];
                k.seeds_dev = in_dev;
                k.num_seeds = r->q.num_seeds;
                k.instance = r->q.instance;
                k.hops = r->q.hops;
                memcpy(k.fanout, r->q.fanout, sizeof(k.fanout));
            }
        }
        const uint64_t seed = cfg.seed;
        if (!rw.empty()) {
            int rc = rw_batch_enqueue(g, rw.data(), (uint32_t)rw.size(), seed, s.stream, s.ws_rw, nullptr);
            if (rc) return rc;
        }
        if (!kh.empty()) {
            const uint32_t K = (uint32_t)kh.size();
            char* base = arena.at(s.out->off);
            s.out_base = base;
            const uint64_t cap = kout / 8;
            s.cap = cap;
            grass_khop_output o{};
            o.src_dev = (uint32_t*)base;
            o.dst_dev = (uint32_t*)base + cap;
            o.capacity = cap;
            o.offsets_dev = (uint64_t*)(base + kout);
            if (s.ws_kh.host.reserve(Workspace::kResultBytes) != hipSuccess) return GRASS_ENOMEM;
            int rc = khop_enqueue(g, kh.data(), K, seed, &o, s.stream, s.ws_kh);
            if (rc) return rc;
            const size_t ob = (size_t)std::max<uint32_t>(s.H, 1) * (K + 1) * 8;
            if (s.offs.reserve(ob) != hipSuccess) return GRASS_ENOMEM;
            if (hipMemcpyAsync(s.offs.ptr, o.offsets_dev, ob, hipMemcpyDeviceToHost, s.stream) != hipSuccess)
                return GRASS_EDEVICE;
        }
        return GRASS_OK;
    }

    void run() {
        (void)hipSetDevice(g->device);
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            bool progress = false, busy = false;
            for (Slot& s : slots) {
                if (!s.busy) continue;
                if (hipEventQuery(s.done) == hipErrorNotReady) {
                    busy = true;
                    continue;
                }
                finalize(s);
                progress = true;
            }
            if (progress) cv_done.notify_all();
            for (Slot& s : slots) {
                if (s.busy || pending.empty() || !admissible) continue;
                if (!launch(s)) break;
                progress = busy = true;
            }
            if (stop && pending.empty() && !busy) break;
            if (progress) continue;
            if (busy) {  // groups on the device: poll their events
                lk.unlock();
                std::this_thread::yield();
                lk.lock();
            } else {
                cv_work.wait(lk, [&] { return stop || (!pending.empty() && admissible); });
            }
        }
    }

    int submit(const grass_request* q, uint64_t* ticket, bool auto_release, double arrival_us) {
        if (!q || !ticket) return GRASS_EINVAL;
        std::unique_ptr<Req> r(new (std::nothrow) Req());
        if (!r) return GRASS_ENOMEM;
        r->q = *q;
        if (q->kind == GRASS_REQ_RW) {
            if (q->walkers == 0 || q->walkers > 0xFFFFFFFFull || q->length == 0xFFFFFFFFu) return GRASS_EINVAL;
            if (q->starts) r->data.assign(q->starts, q->starts + q->walkers);
        } else if (q->kind == GRASS_REQ_KHOP) {
            if (q->hops == 0 || q->hops > GRASS_KHOP_MAX_HOPS || (q->num_seeds && !q->seeds)) return GRASS_EINVAL;
            for (uint32_t h = 0; h < q->hops; ++h)
                if (q->fanout[h] == 0 || q->fanout[h] > 64) return GRASS_EINVAL;
            r->data.assign(q->seeds, q->seeds + q->num_seeds);
        } else {
            return GRASS_EINVAL;
        }
        for (uint32_t x : r->data)
            if (x >= g->n) return GRASS_EINVAL;
        // from here on the request refers to its own copy of the ids
        if (q->kind == GRASS_REQ_RW) r->q.starts = r->data.empty() ? nullptr : r->data.data();
        else r->q.seeds = r->data.data();
        const uint64_t need = own_bytes(r->q) + (q->kind == GRASS_REQ_KHOP ? khop_out_bytes(r->q) + 4096 : 0);
        std::lock_guard<std::mutex> lk(mu);
        stats.submitted++;
        if (need + 8 * 8 * (cfg.max_batch + 1) > arena.size()) {
            stats.rejected++;
            return GRASS_ECAPACITY;
        }
        r->ticket = next_ticket++;
        r->auto_release = auto_release;
        r->res.submit_us = now_us();
        r->arrival_us = arrival_us < 0 ? r->res.submit_us : arrival_us;
        *ticket = r->ticket;
        pending.push_back(r.get());
        reqs.emplace(r->ticket, std::move(r));
        cv_work.notify_one();
        return GRASS_OK;
    }
};

extern "C" {

int grass_session_create(grass_graph* g, const grass_session_config* cfg, grass_session** out) {
    if (!g || !out) return GRASS_EINVAL;
    *out = nullptr;
    grass_session* s = new (std::nothrow) grass_session();
    if (!s) return GRASS_ENOMEM;
    s->g = g;
    if (cfg) s->cfg = *cfg;
    if (!s->cfg.memory_budget) s->cfg.memory_budget = 4ull << 30;
    if (!s->cfg.max_batch) s->cfg.max_batch = 512;
    if (!s->cfg.max_inflight) s->cfg.max_inflight = 3;
    s->cfg.max_batch = std::min<uint32_t>(s->cfg.max_batch, 1u << 16);
    int prev = 0;
    (void)hipGetDevice(&prev);
    if (hipSetDevice(g->device) != hipSuccess) {
        delete s;
        return GRASS_EDEVICE;
    }
    int rc = s->arena.init(s->cfg.memory_budget) == hipSuccess ? GRASS_OK : GRASS_ENOMEM;
    s->slots.resize(s->cfg.max_inflight);
    for (Slot& sl : s->slots) {
        if (rc) break;
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess)
            rc = GRASS_EDEVICE;
    }
    (void)hipSetDevice(prev);
    if (rc) {
        grass_session_destroy(s);
        return rc;
    }
    s->t0 = Clock::now();
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return GRASS_OK;
}

int grass_session_submit(grass_session* s, const grass_request* req, uint64_t* ticket) {
    if (!s) return GRASS_EINVAL;
    return s->submit(req, ticket, false, -1.0);
}

int grass_session_poll(grass_session* s, uint64_t ticket) {
    if (!s) return GRASS_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return GRASS_EINVAL;
    Req* r = it->second.get();
    if (r->state != St::Done) return GRASS_ENOTREADY;
    return r->status;
}

int grass_session_wait(grass_session* s, uint64_t ticket, int64_t timeout_us) {
    if (!s) return GRASS_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto done = [&] {
        auto it = s->reqs.find(ticket);
        return it == s->reqs.end() || it->second->state == St::Done;
    };
    if (timeout_us < 0) s->cv_done.wait(lk, done);
    else if (!s->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), done)) return GRASS_ETIMEOUT;
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return GRASS_EINVAL;
    return it->second->status;
}

int grass_session_result(grass_session* s, uint64_t ticket, grass_result* out) {
    if (!s || !out) return GRASS_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return GRASS_EINVAL;
    Req* r = it->second.get();
    if (r->state != St::Done) return GRASS_ENOTREADY;
    *out = r->res;
    return r->status;
}

int grass_session_fetch(grass_session* s, uint64_t ticket, void* host, uint64_t capacity,
                        uint64_t* bytes) {
    if (!s || !bytes) return GRASS_EINVAL;
    grass_result res;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) return GRASS_EINVAL;
        Req* r = it->second.get();
        if (r->state != St::Done) return GRASS_ENOTREADY;
        if (r->status != GRASS_OK) return r->status;
        if (r->auto_release) return GRASS_EINVAL;  // its memory is already gone
        res = r->res;
    }
    // the request's memory stays allocated until the caller releases it
    uint64_t need = 0;
    if (res.kind == GRASS_REQ_RW) {
        need = res.walkers * ((uint64_t)res.length + 1) * 4;
    } else {
        for (uint32_t h = 0; h < res.hops; ++h) need += res.hop_edges[h] * 8;
    }
    *bytes = need;
    if (need > capacity || (need && !host)) return GRASS_ECAPACITY;
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(s->g->device);
    hipError_t e = hipSuccess;
    if (res.kind == GRASS_REQ_RW) {
        e = hipMemcpy(host, res.walks_dev, need, hipMemcpyDeviceToHost);
    } else {
        uint64_t total = need / 8, pos = 0;
        uint32_t* out = (uint32_t*)host;
        for (uint32_t h = 0; h < res.hops && e == hipSuccess; ++h) {
            const uint64_t c = res.hop_edges[h];
            if (!c) continue;
            e = hipMemcpy(out + pos, res.hop_src_dev[h], c * 4, hipMemcpyDeviceToHost);
            if (e == hipSuccess) e = hipMemcpy(out + total + pos, res.hop_dst_dev[h], c * 4, hipMemcpyDeviceToHost);
            pos += c;
        }
    }
    (void)hipSetDevice(prev);
    return status_of(e);
}

int grass_session_release(grass_session* s, uint64_t ticket) {
    if (!s) return GRASS_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return GRASS_EINVAL;
    Req* r = it->second.get();
    if (r->state == St::Running) {  // finish() frees it
        r->cancelled = true;
        return GRASS_OK;
    }
    if (r->state == St::Pending) s->pending.erase(std::find(s->pending.begin(), s->pending.end(), r));
    s->free_req_mem(r);
    s->reqs.erase(it);
    s->cv_work.notify_one();
    return GRASS_OK;
}

int grass_session_cancel(grass_session* s, uint64_t ticket) {
    if (!s) return GRASS_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return GRASS_EINVAL;
    Req* r = it->second.get();
    if (r->state == St::Pending) {
        s->pending.erase(std::find(s->pending.begin(), s->pending.end(), r));
        r->cancelled = true;
        r->state = St::Done;
        r->status = r->res.status = GRASS_ECANCELLED;
        s->stats.cancelled++;
        s->cv_done.notify_all();
        return GRASS_OK;
    }
    if (r->state == St::Running) r->cancelled = true;
    return GRASS_OK;
}

int grass_session_stats(grass_session* s, grass_session_counters* out) {
    if (!s || !out) return GRASS_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->memory_in_use = s->arena.in_use;
    out->memory_peak = s->arena.peak;
    return GRASS_OK;
}

int grass_session_destroy(grass_session* s) {
    if (!s) return GRASS_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (Req* r : s->pending) {
            r->cancelled = true;
            r->state = St::Done;
            r->status = GRASS_ECANCELLED;
        }
        s->pending.clear();
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(s->g->device);
    for (Slot& sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        sl.ws_rw.release();
        sl.ws_kh.release();
        sl.stage.release();
        sl.offs.release();
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    (void)hipSetDevice(prev);
    delete s;
    return GRASS_OK;
}

int grass_session_replay(grass_session* s, const grass_request* reqs, const double* arrival_us,
                         uint32_t count, double* latency_us, double* wall_us) {
    if (!s || (count && (!reqs || !arrival_us || !latency_us))) return GRASS_EINVAL;
    std::vector<uint64_t> tickets(count, 0);
    const double base = s->now_us();
    for (uint32_t i = 0; i < count; ++i) {
        const double due = base + arrival_us[i];
        while (s->now_us() < due) std::this_thread::yield();
        int rc = s->submit(&reqs[i], &tickets[i], true, due);
        if (rc != GRASS_OK) {
            tickets[i] = 0;
            latency_us[i] = -1.0;
        }
    }
    double last = base;
    for (uint32_t i = 0; i < count; ++i) {
        if (!tickets[i]) continue;
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(tickets[i]);
        if (it == s->reqs.end()) {
            latency_us[i] = -1.0;
            continue;
        }
        Req* r = it->second.get();
        s->cv_done.wait(lk, [&] { return r->state == St::Done; });
        latency_us[i] = r->status == GRASS_OK ? r->res.done_us - r->arrival_us : -1.0;
        last = std::max(last, r->res.done_us);
        s->free_req_mem(r);
        s->reqs.erase(it);
    }
    if (wall_us) *wall_us = last - base - (count ? arrival_us[0] : 0.0);
    return GRASS_OK;
}

}  // extern "C"
