// Serving session: continuous batching of random-walk and k-hop requests on one resident graph.
//
// A scheduler thread owns `max_inflight` launch-group slots.  Whenever a slot is free and
// requests are queued, it takes queued requests in arrival order while their HBM fits the
// session's arena (memory-based admission) and the group's edge bound stays under
// max_group_edges, and launches them as one group:
//   * all the group's inputs (walk starts, k-hop seeds) go to HBM in one copy from pinned staging;
//   * the random walks run as one persistent launch on the slot's walk stream while the k-hop
//     requests run as one pipeline on its k-hop stream, so the two kinds overlap;
//   * then the scheduler goes back to watching the slots' events.
// Requests that arrive while groups run are therefore admitted as soon as a slot and memory free
// up, and their work overlaps the running groups on the device.  Walk matrices and k-hop edges
// stay in HBM until the client releases them (the group's input block is freed when the group
// completes).  Every random choice is keyed by (seed, instance, ...), so a request's output does
// not depend on its group.
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

#include "devutil.h"
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
    static size_t round(size_t bytes) { return (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255; }
    bool alloc(size_t bytes, size_t* off) {
        bytes = round(bytes);
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
    // Whether regions of the given sizes could all be allocated now.  The trial allocations are
    // undone and leave the peak as it was.
    bool fits(std::initializer_list<size_t> sizes) {
        std::vector<std::pair<size_t, size_t>> got;
        const size_t peak0 = peak;
        bool ok = true;
        for (size_t b : sizes) {
            if (!b) continue;
            size_t off;
            if (!alloc(b, &off)) {
                ok = false;
                break;
            }
            got.emplace_back(off, b);
        }
        for (auto& p : got) release(p.first, p.second);
        peak = peak0;
        return ok;
    }
    void release(size_t off, size_t bytes) {
        bytes = round(bytes);
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
    size_t off = 0, bytes = 0;  // own arena region: the walk matrix of a random-walk request
    bool has_mem = false;
    std::shared_ptr<OutBlock> out;
    uint32_t kidx = 0;       // index among its group's k-hop requests
    uint64_t kbound = 0;     // k-hop edge bound
    size_t in_pos = 0;       // offset of its ids in the group's input block (words)
    grass_result res{};
    double arrival_us = 0;
    uint64_t edges = 0;
};

struct Slot {
    hipStream_t s_rw = nullptr, s_kh = nullptr;
    hipEvent_t in_ready = nullptr, ev_rw = nullptr, ev_kh = nullptr;
    bool rw_pending = false, kh_pending = false;
    Workspace ws_rw, ws_kh;
    PinnedBuf stage;  // the group's inputs on their way to HBM
    PinnedBuf offs;   // k-hop offsets read back
    PinnedBuf errs;   // the group's error words read back: [0] walks, [1] k-hop
    bool busy = false;
    bool rw_finished = false;  // the group's walks have been completed
    std::vector<Req*> reqs;
    std::vector<Req*> kreqs;
    std::shared_ptr<OutBlock> out;
    size_t in_off = 0, in_bytes = 0;  // input block
    uint32_t H = 0;
    char* out_base = nullptr;  // k-hop block: src[cap], dst[cap], offsets
    uint64_t cap = 0;
    int launch_status = GRASS_OK;
};

// Bytes of a walk matrix; saturates instead of wrapping (such a request is rejected anyway).
uint64_t walk_bytes(const grass_request& q) {
    const uint64_t cols = (uint64_t)q.length + 1;
    if (q.walkers > (UINT64_MAX / 8) / cols) return UINT64_MAX / 2;
    return q.walkers * cols * 4;
}

// k-hop block of a group: src[cap], dst[cap] (cap = sum of the requests' edge bounds), then the
// hop-major offsets (at most GRASS_KHOP_MAX_HOPS x (K + 1)).
uint64_t block_bytes(uint64_t kcap, uint32_t K) { return kcap * 8 + 8ull * GRASS_KHOP_MAX_HOPS * (K + 1); }

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

    uint64_t khop_edge_bound(const grass_request& q) const {
        grass_khop_request k{};
        k.num_seeds = q.num_seeds;
        k.seeds_dev = (const uint32_t*)16;  // only the sizes matter to the bound
        k.hops = q.hops;
        memcpy(k.fanout, q.fanout, sizeof(k.fanout));
        uint64_t e = 0;
        if (khop_bound(g, &k, 1, &e, nullptr, nullptr, nullptr)) return ~0ull >> 8;
        return e;
    }

    void notify_memory() {
        stats.memory_in_use = arena.in_use;
        admissible = true;  // memory came back: queued requests may fit now
        cv_work.notify_one();
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
        notify_memory();
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

    static bool passed(bool pending, hipEvent_t ev) { return !pending || hipEventQuery(ev) != hipErrorNotReady; }

    // Completes the requests of one kind in slot s (the walks finish without waiting for the
    // group's k-hop pipeline).
    void finish_kind(Slot& s, int kind) {
        const uint64_t* off = (const uint64_t*)s.offs.ptr;
        const uint32_t K = (uint32_t)s.kreqs.size();
        // what the kind's kernels recorded in its own error word (a decoded id or a start that is
        // not a vertex) fails the kind's requests of this group, and no other call's
        int status = s.launch_status;
        if (status == GRASS_OK) status = status_of_err(((const uint32_t*)s.errs.ptr)[kind == GRASS_REQ_RW ? 0 : 1]);
        for (Req*& slot_req : s.reqs) {
            Req* r = slot_req;
            if (!r || r->q.kind != kind) continue;
            slot_req = nullptr;  // finish() may delete a cancelled request
            if (kind == GRASS_REQ_KHOP && status == GRASS_OK) {
                r->res.hops = r->q.hops;
                r->edges = 0;
                for (uint32_t h = 0; h < r->q.hops && h < s.H; ++h) {
                    const uint64_t a = off[(uint64_t)h * (K + 1) + r->kidx];
                    const uint64_t b = off[(uint64_t)h * (K + 1) + r->kidx + 1];
                    r->res.hop_src_dev[h] = (const uint32_t*)s.out_base + a;
                    r->res.hop_dst_dev[h] = (const uint32_t*)s.out_base + s.cap + a;
                    r->res.hop_edges[h] = b - a;
                    r->edges += b - a;
                }
            } else if (kind == GRASS_REQ_RW) {
                r->edges = r->q.walkers * r->q.length;
            }
            finish(r, status);
        }
    }

    // Advances slot s; true when some request completed.
    bool progress_slot(Slot& s) {
        bool any = false;
        if (s.launch_status != GRASS_OK && !s.rw_finished) {
            // part of the group may still be queued (the input copy, one kind's launches): let it
            // drain before the group's memory goes back to the arena
            (void)hipStreamSynchronize(s.s_rw);
            (void)hipStreamSynchronize(s.s_kh);
        }
        if (!s.rw_finished && (s.launch_status != GRASS_OK || passed(s.rw_pending, s.ev_rw))) {
            finish_kind(s, GRASS_REQ_RW);
            s.rw_finished = true;
            any = true;
        }
        if (s.rw_finished && (s.launch_status != GRASS_OK || passed(s.kh_pending, s.ev_kh)) &&
            passed(s.rw_pending, s.ev_rw)) {
            finish_kind(s, GRASS_REQ_KHOP);
            if (s.in_bytes) arena.release(s.in_off, s.in_bytes);
            s.in_bytes = 0;
            notify_memory();
            s.reqs.clear();
            s.kreqs.clear();
            s.out.reset();
            s.rw_pending = s.kh_pending = false;
            s.busy = false;
            any = true;
        }
        return any;
    }

    // Forms and launches one group on slot s; false when nothing could be admitted.
    bool launch(Slot& s) {
        std::vector<Req*> take;
        uint64_t kcap = 0, in_words = 0, est = 0;
        uint32_t K = 0;
        while (!pending.empty() && take.size() < cfg.max_batch) {
            Req* r = pending.front();
            const bool kreq = r->q.kind == GRASS_REQ_KHOP;
            const uint64_t bound = kreq ? r->kbound : r->q.walkers * r->q.length;
            if (!take.empty() && est + bound > cfg.max_group_edges) break;  // group is full
            const uint64_t own = kreq ? 0 : walk_bytes(r->q);
            const uint64_t words = in_words + r->data.size();
            const uint64_t kc = kcap + (kreq ? bound : 0);
            const uint32_t kk = K + kreq;
            // the request's own region, the input block and the k-hop block must fit together
            if (!arena.fits({own, words * 4, kk ? block_bytes(kc, kk) : 0})) break;
            if (own) {
                size_t off = 0;
                arena.alloc(own, &off);
                r->off = off;
                r->bytes = own;
                r->has_mem = true;
            }
            r->in_pos = in_words;
            in_words = words;
            kcap = kc;
            K = kk;
            est += bound;
            take.push_back(r);
            pending.pop_front();
        }
        // group blocks; should fragmentation defeat the check above, hand the newest requests back
        std::shared_ptr<OutBlock> blk;
        size_t in_off = 0;
        for (;;) {
            bool ok = true;
            if (in_words && !arena.alloc(in_words * 4, &in_off)) ok = false;
            if (ok && K) {
                blk = std::make_shared<OutBlock>();
                blk->bytes = block_bytes(kcap, K);
                if (!arena.alloc(blk->bytes, &blk->off)) {
                    if (in_words) arena.release(in_off, in_words * 4);
                    blk.reset();
                    ok = false;
                }
            }
            if (ok || take.empty()) break;
            Req* r = take.back();
            take.pop_back();
            in_words -= r->data.size();
            if (r->q.kind == GRASS_REQ_KHOP) {
                kcap -= r->kbound;
                --K;
            }
            free_req_mem(r);
            pending.push_front(r);
        }
        if (take.empty()) {
            admissible = false;  // nothing fits until memory is released
            return false;
        }
        s.reqs = take;
        s.kreqs.clear();
        for (Req* r : take)
            if (r->q.kind == GRASS_REQ_KHOP) {
                r->kidx = (uint32_t)s.kreqs.size();
                s.kreqs.push_back(r);
            }
        s.H = 0;
        for (Req* r : s.kreqs) s.H = std::max(s.H, r->q.hops);
        s.in_off = in_off;
        s.in_bytes = in_words * 4;
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
        for (Req* r : take) {
            r->state = St::Running;
            r->res.launch_us = t;
        }
        s.launch_status = enqueue(s, kcap);
        s.rw_finished = false;
        s.busy = true;
        return true;
    }

    int enqueue(Slot& s, uint64_t kcap) {
        const uint64_t seed = cfg.seed;
        uint32_t* in_dev = (uint32_t*)arena.at(s.in_off);
        uint32_t* errs = (uint32_t*)s.errs.ptr;
        errs[0] = errs[1] = kErrNone;
        // 1. every input of the group in one copy (on the k-hop stream; the walk stream waits)
        if (s.in_bytes) {
            if (s.stage.reserve(s.in_bytes) != hipSuccess) return GRASS_ENOMEM;
            uint32_t* st = (uint32_t*)s.stage.ptr;
            for (Req* r : s.reqs)
                if (!r->data.empty()) memcpy(st + r->in_pos, r->data.data(), r->data.size() * 4);
            if (hipMemcpyAsync(in_dev, st, s.in_bytes, hipMemcpyHostToDevice, s.s_kh) != hipSuccess ||
                hipEventRecord(s.in_ready, s.s_kh) != hipSuccess ||
                hipStreamWaitEvent(s.s_rw, s.in_ready, 0) != hipSuccess)
                return GRASS_EDEVICE;
        }
        std::vector<grass_rw_request> rw;
        std::vector<grass_khop_request> kh(s.kreqs.size());
        for (Req* r : s.reqs) {
            const uint32_t* ids = r->data.empty() ? nullptr : in_dev + r->in_pos;
            if (r->q.kind == GRASS_REQ_RW) {
                uint32_t* walks = (uint32_t*)arena.at(r->off);
                r->res.kind = GRASS_REQ_RW;
                r->res.walks_dev = walks;
                r->res.walkers = r->q.walkers;
                r->res.length = r->q.length;
                rw.push_back(grass_rw_request{r->q.walkers, r->q.length, r->q.instance, ids, walks});
            } else {
                r->res.kind = GRASS_REQ_KHOP;
                grass_khop_request& k = kh[r->kidx];
                k.seeds_dev = ids;
                k.num_seeds = r->q.num_seeds;
                k.instance = r->q.instance;
                k.hops = r->q.hops;
                memcpy(k.fanout, r->q.fanout, sizeof(k.fanout));
            }
        }
        // 2. walks on the walk stream, recording into the slot's own error word
        if (!rw.empty()) {
            hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(1), dim3(64), 0, s.s_rw, s.ws_rw.err, (uint64_t)1, 0u);
            int rc = rw_batch_enqueue(g, rw.data(), (uint32_t)rw.size(), seed, s.s_rw, s.ws_rw, nullptr);
            if (rc == GRASS_OK &&
                hipMemcpyAsync(errs, s.ws_rw.err, 4, hipMemcpyDeviceToHost, s.s_rw) != hipSuccess)
                rc = GRASS_EDEVICE;
            if (hipEventRecord(s.ev_rw, s.s_rw) != hipSuccess && rc == GRASS_OK) rc = GRASS_EDEVICE;
            s.rw_pending = true;
            if (rc) return rc;
        }
        // 3. k-hop pipeline on the k-hop stream
        if (!kh.empty()) {
            const uint32_t K = (uint32_t)kh.size();
            char* base = arena.at(s.out->off);
            s.out_base = base;
            s.cap = kcap;
            grass_khop_output o{};
            o.src_dev = (uint32_t*)base;
            o.dst_dev = (uint32_t*)base + kcap;
            o.capacity = kcap;
            o.offsets_dev = (uint64_t*)(base + kcap * 8);
            hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(1), dim3(64), 0, s.s_kh, s.ws_kh.err, (uint64_t)1, 0u);
            int rc = khop_enqueue(g, kh.data(), K, seed, &o, s.s_kh, s.ws_kh);
            const size_t ob = (size_t)std::max<uint32_t>(s.H, 1) * (K + 1) * 8;
            if (rc == GRASS_OK && s.offs.reserve(ob) != hipSuccess) rc = GRASS_ENOMEM;
            if (rc == GRASS_OK &&
                hipMemcpyAsync(s.offs.ptr, o.offsets_dev, ob, hipMemcpyDeviceToHost, s.s_kh) != hipSuccess)
                rc = GRASS_EDEVICE;
            if (rc == GRASS_OK && hipMemcpyAsync(errs + 1, s.ws_kh.err, 4, hipMemcpyDeviceToHost, s.s_kh) != hipSuccess)
                rc = GRASS_EDEVICE;
            if (hipEventRecord(s.ev_kh, s.s_kh) != hipSuccess && rc == GRASS_OK) rc = GRASS_EDEVICE;
            s.kh_pending = true;
            if (rc) return rc;
        } else if (s.in_bytes) {
            // only walks: the input copy sits on the k-hop stream; wait for it too
            if (hipEventRecord(s.ev_kh, s.s_kh) != hipSuccess) return GRASS_EDEVICE;
            s.kh_pending = true;
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
                if (progress_slot(s)) progress = true;
                if (s.busy) busy = true;
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
        if (q->kind == GRASS_REQ_RW) {
            if (q->walkers == 0 || q->walkers > 0xFFFFFFFFull || q->length == 0xFFFFFFFFu) return GRASS_EINVAL;
        } else if (q->kind == GRASS_REQ_KHOP) {
            if (q->hops == 0 || q->hops > GRASS_KHOP_MAX_HOPS || (q->num_seeds && !q->seeds)) return GRASS_EINVAL;
            for (uint32_t h = 0; h < q->hops; ++h)
                if (q->fanout[h] == 0 || q->fanout[h] > 64) return GRASS_EINVAL;
        } else {
            return GRASS_EINVAL;
        }
        // HBM the request needs in the arena, checked before anything is copied
        uint64_t need = 0, kbound = 0;
        if (q->kind == GRASS_REQ_RW) {
            need = walk_bytes(*q) + (q->starts ? q->walkers * 4 : 0);
        } else if (q->num_seeds > arena.size() / 4) {
            need = UINT64_MAX / 2;  // more seeds than the arena holds
        } else {
            kbound = khop_edge_bound(*q);
            need = q->num_seeds * 4 + block_bytes(kbound, 1);
        }
        if (need + 4096 > arena.size()) {
            std::lock_guard<std::mutex> lk(mu);
            stats.submitted++;
            stats.rejected++;
            return GRASS_ECAPACITY;
        }
        std::unique_ptr<Req> r(new (std::nothrow) Req());
        if (!r) return GRASS_ENOMEM;
        r->q = *q;
        r->kbound = kbound;
        if (q->kind == GRASS_REQ_RW) {
            if (q->starts) r->data.assign(q->starts, q->starts + q->walkers);
        } else {
            r->data.assign(q->seeds, q->seeds + q->num_seeds);
        }
        for (uint32_t x : r->data)
            if (x >= g->n) return GRASS_EINVAL;
        // from here on the request refers to its own copy of the ids
        if (q->kind == GRASS_REQ_RW) r->q.starts = r->data.empty() ? nullptr : r->data.data();
        else r->q.seeds = r->data.data();
        std::lock_guard<std::mutex> lk(mu);
        stats.submitted++;
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

// Sizes a slot's working memory now for the largest group the edge bound admits (seeds, frontier
// and dedup bounds are each at most the group's edge bound when every fanout is at least 1; at
// most kReserveEdges of them), so that serving never grows it: a growth frees and reallocates
// device or pinned memory, and the free waits for the whole device, which stalls every request
// in flight for milliseconds.  Best effort: where memory is short the slot keeps what it got and
// a group beyond it grows its buffers as it would have without this.
static constexpr uint64_t kReserveEdges = 8ull << 20;
static void reserve_slot(const grass_graph* g, const grass_session_config& cfg, Slot& sl) {
    const uint64_t E = std::min<uint64_t>(cfg.max_group_edges, kReserveEdges);
    const uint32_t B = cfg.max_batch;
    if (rw_reserve(sl.ws_rw, B) != GRASS_OK || khop_reserve(g, sl.ws_kh, B, E, E, E) != GRASS_OK ||
        sl.stage.reserve(E * 4) != hipSuccess ||
        sl.offs.reserve((size_t)GRASS_KHOP_MAX_HOPS * (B + 1) * 8) != hipSuccess)
        (void)hipGetLastError();  // a failed allocation must not read as a later kernel's error
}

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
    if (!s->cfg.max_inflight) s->cfg.max_inflight = 2;
    if (!s->cfg.max_group_edges) s->cfg.max_group_edges = 6ull << 20;
    s->cfg.max_batch = std::min<uint32_t>(s->cfg.max_batch, 1u << 16);
    s->cfg.max_inflight = std::min<uint32_t>(s->cfg.max_inflight, 64);
    int prev = 0;
    (void)hipGetDevice(&prev);
    if (!grass::enter_device(g->device)) {
        delete s;
        return GRASS_EDEVICE;
    }
    int rc = s->arena.init(s->cfg.memory_budget) == hipSuccess ? GRASS_OK : GRASS_ENOMEM;
    s->slots.resize(s->cfg.max_inflight);
    for (Slot& sl : s->slots) {
        if (rc) break;
        if (hipStreamCreateWithFlags(&sl.s_rw, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&sl.s_kh, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.in_ready, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ev_rw, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ev_kh, hipEventDisableTiming) != hipSuccess)
            rc = GRASS_EDEVICE;
        // each slot's kernels record errors in words of their own, read back with the group
        else if (hipMalloc((void**)&sl.ws_rw.err, 64) != hipSuccess ||
                 hipMalloc((void**)&sl.ws_kh.err, 64) != hipSuccess || sl.errs.reserve(64) != hipSuccess)
            rc = GRASS_ENOMEM;
        else if (hipMemset(sl.ws_rw.err, 0, 64) != hipSuccess || hipMemset(sl.ws_kh.err, 0, 64) != hipSuccess)
            rc = GRASS_EDEVICE;
        else
            reserve_slot(g, s->cfg, sl);
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
        if (sl.s_rw) (void)hipStreamSynchronize(sl.s_rw);
        if (sl.s_kh) (void)hipStreamSynchronize(sl.s_kh);
        sl.ws_rw.release();
        sl.ws_kh.release();
        sl.stage.release();
        sl.offs.release();
        sl.errs.release();
        for (hipEvent_t e : {sl.in_ready, sl.ev_rw, sl.ev_kh})
            if (e) (void)hipEventDestroy(e);
        if (sl.s_rw) (void)hipStreamDestroy(sl.s_rw);
        if (sl.s_kh) (void)hipStreamDestroy(sl.s_kh);
    }
    (void)hipSetDevice(prev);
    delete s;
    return GRASS_OK;
}

int grass_session_replay(grass_session* s, const grass_request* reqs, const double* arrival_us,
                         uint32_t count, double* latency_us, double* queue_us, double* wall_us) {
    if (!s || (count && (!reqs || !arrival_us || !latency_us))) return GRASS_EINVAL;
    std::vector<uint64_t> tickets(count, 0);
    {
        // every ticket of the replay stays in the map until the submit loop ends: size it now, or
        // its rehashes (under the session lock, about 10 ms at 85 k tickets) stall the worker
        std::lock_guard<std::mutex> lk(s->mu);
        s->reqs.reserve(s->reqs.size() + count);
    }
    const double base = s->now_us();
    for (uint32_t i = 0; i < count; ++i) {
        const double due = base + arrival_us[i];
        while (s->now_us() < due) std::this_thread::yield();
        int rc = s->submit(&reqs[i], &tickets[i], true, due);
        if (rc != GRASS_OK) {
            tickets[i] = 0;
            latency_us[i] = -1.0;
            if (queue_us) queue_us[i] = -1.0;
        }
    }
    double last = base;
    for (uint32_t i = 0; i < count; ++i) {
        if (!tickets[i]) continue;
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(tickets[i]);
        if (it == s->reqs.end()) {
            latency_us[i] = -1.0;
            if (queue_us) queue_us[i] = -1.0;
            continue;
        }
        Req* r = it->second.get();
        s->cv_done.wait(lk, [&] { return r->state == St::Done; });
        latency_us[i] = r->status == GRASS_OK ? r->res.done_us - r->arrival_us : -1.0;
        if (queue_us) queue_us[i] = r->status == GRASS_OK ? r->res.launch_us - r->arrival_us : -1.0;
        last = std::max(last, r->res.done_us);
        s->free_req_mem(r);
        s->reqs.erase(it);
    }
    if (wall_us) *wall_us = last - base - (count ? arrival_us[0] : 0.0);
    return GRASS_OK;
}

}  // extern "C"
