// Serving loop: continuous batching of FFT requests on one GPU.
//
// The reference has no serving layer; its closest mechanism is recording many
// plans into one FftRecorder (src/plan.rs:38-95).  Here a single host event
// loop replays an arrival trace:
//   * arrivals join a FIFO queue;
//   * whenever a stream slot is free, the loop forms a tick from the head of
//     the queue, admitting requests while their memory (input + output +
//     workspace) fits the budget next to the work already on the GPU;
//   * a tick packs every request whose plan the batched kernel can run (1D and
//     2D C2C in f32 and f64, R2C; up to eight dependent passes) and that has
//     fewer points than the batch limit into one persistent batched launch per
//     pass (kern_batch.hip) and runs the others (large transforms, df64)
//     through their resident plans on the slot's second stream;
//   * ticks on different slots overlap on the device; an event per tick marks
//     completion, which frees the slot and the memory.
// Plans, descriptor arenas, pinned staging and workspaces are allocated once
// (mifft_server_prepare); the loop itself allocates nothing.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>

#include "batch_plan.hpp"

using namespace mifft;

namespace {

struct PlanKey {
    int kind, prec, rank, dir, norm;
    uint64_t s0, s1, batch;
    bool operator<(const PlanKey& o) const {
        return std::tie(kind, prec, rank, dir, norm, s0, s1, batch) <
               std::tie(o.kind, o.prec, o.rank, o.dir, o.norm, o.s0, o.s1, o.batch);
    }
};

PlanKey key_of(const mifft_server_request& r) {
    return PlanKey{r.kind, r.precision, r.rank, r.direction, r.normalization, r.shape[0],
                   r.rank > 1 ? r.shape[1] : 1, r.batch};
}

struct Slot {
    hipStream_t s = nullptr;
    hipEvent_t ev = nullptr;
    hipStream_t s2 = nullptr;   // the tick's requests outside the batched launch
    hipEvent_t ev2 = nullptr;
    bool busy = false;
    std::vector<int> reqs;
    uint64_t mem = 0;
    char* ws = nullptr;
    uint64_t ws_cap = 0;
    char* desc = nullptr;       // device arena: ops, tiles
    char* staging = nullptr;    // pinned host copy of the arena
    size_t desc_cap = 0;
    uint32_t* counters = nullptr;
};

double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct mifft_server {
    uint64_t budget = 0;
    int nslots = 0;
    int max_items = 0;
    // Requests of at least this many points skip the batched launch (measured
    // on the serving mix: 2^19 gives 3.9x over sequential dispatch, 2^18 3.8x,
    // 2^20 3.6x, no limit 2.9x).
    uint64_t batch_max_elems = 1ull << 19;
    BatchPlanCache bcache;
    std::map<PlanKey, std::unique_ptr<Plan>> plans;  // standalone plans
    std::vector<Slot> slots;
    char* seq_ws = nullptr;  // sequential mode: one workspace, stream order serialises its users
    uint64_t seq_ws_cap = 0;
    hipStream_t seq_stream = nullptr;
    std::vector<hipEvent_t> seq_events;
    // online submission (mifft_server_submit / _poll)
    struct Item {
        mifft_server_request r;
        uint64_t ticket = 0;
        uint64_t mem = 0, desc = 0;
        bool batched = false;
        double submit_s = 0, admit_s = 0;
    };
    std::deque<Item> queue;                   // submitted, not yet admitted
    std::vector<std::vector<Item>> slot_items;  // per slot, in flight
    std::deque<mifft_server_completion> completions;
    uint64_t next_ticket = 1, online_inflight = 0;
    bool clock_started = false;
    std::chrono::steady_clock::time_point t0;
    // resident plans of the online path: at most max_shapes request shapes,
    // least recently used evicted first among shapes with nothing in flight
    size_t max_shapes = 256;
    uint64_t use_clock = 0;
    std::map<PlanKey, uint64_t> last_use;
    std::map<PlanKey, int> refs;
    ~mifft_server() {
        for (Slot& s : slots) {
            if (s.ev) (void)hipEventDestroy(s.ev);
            if (s.ev2) (void)hipEventDestroy(s.ev2);
            if (s.s2) (void)hipStreamDestroy(s.s2);
            if (s.s) (void)hipStreamDestroy(s.s);
            if (s.ws) (void)hipFree(s.ws);
            if (s.desc) (void)hipFree(s.desc);
            if (s.staging) (void)hipHostFree(s.staging);
            if (s.counters) (void)hipFree(s.counters);
        }
        for (hipEvent_t e : seq_events) (void)hipEventDestroy(e);
        if (seq_stream) (void)hipStreamDestroy(seq_stream);
        if (seq_ws) (void)hipFree(seq_ws);
    }
};

static mifft_config config_of(const mifft_server_request& r) {
    mifft_config c;
    std::memset(&c, 0, sizeof(c));
    if (r.rank < 1 || r.rank > 2) fail(MIFFT_ERR_RANK_TOO_LARGE, "server requests are 1D or 2D");
    c.rank = r.rank;
    c.naxes = r.rank;
    for (int i = 0; i < r.rank; i++) {
        c.shape[i] = r.shape[i];
        c.axes[i] = i;
    }
    c.batch = r.batch;
    c.direction = r.direction;
    c.normalization = r.normalization;
    c.precision = r.precision;
    mifft_tuning_init(&c.tuning);
    return c;
}

// The batched-mode plan of a request, or nullptr if the batched kernel cannot run it.
static const Plan* batched_plan(mifft_server& S, const mifft_server_request& r) {
    return S.bcache.get(config_of(r), (mifft_kind)r.kind);
}

static Plan& standalone(mifft_server& S, const mifft_server_request& r) {
    PlanKey k = key_of(r);
    auto f = S.plans.find(k);
    if (f != S.plans.end()) return *f->second;
    std::unique_ptr<Plan> p = build_plan(config_of(r), (mifft_kind)r.kind);
    Plan& ref = *p;
    S.plans[k] = std::move(p);
    return ref;
}

static uint64_t request_ws(mifft_server& S, const mifft_server_request& r, bool batched) {
    if (batched) return batched_plan(S, r)->ws_bytes;
    return standalone(S, r).ws_bytes;
}

static void prepare(mifft_server& S, const mifft_server_request* reqs, int32_t n) {
    uint64_t max_plan_ws = 0;
    for (int32_t i = 0; i < n; i++) {
        const mifft_server_request& r = reqs[i];
        batched_plan(S, r);
        // the sequential baseline runs every request through a standalone plan
        max_plan_ws = std::max(max_plan_ws, standalone(S, r).ws_bytes);
    }
    if (S.slots.empty()) {
        S.slots.resize(S.nslots);
        const uint64_t ws_cap = S.budget / (2 * (uint64_t)S.nslots);
        const size_t desc_cap = (size_t)S.max_items * 4096 * sizeof(BatchTile) + (size_t)S.max_items * 2 * sizeof(OpDesc);
        for (Slot& sl : S.slots) {
            hip_check(hipStreamCreateWithFlags(&sl.s, hipStreamNonBlocking), "server stream");
            hip_check(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming), "server event");
            hip_check(hipStreamCreateWithFlags(&sl.s2, hipStreamNonBlocking), "server stream");
            hip_check(hipEventCreateWithFlags(&sl.ev2, hipEventDisableTiming), "server event");
            sl.ws_cap = std::max(ws_cap, max_plan_ws);
            hip_check(hipMalloc(&sl.ws, sl.ws_cap), "server workspace");
            sl.desc_cap = desc_cap;
            hip_check(hipMalloc(&sl.desc, desc_cap), "server descriptors");
            hip_check(hipHostMalloc(&sl.staging, desc_cap, hipHostMallocDefault), "server staging");
            hip_check(hipMalloc(&sl.counters, 64), "server counters");
        }
        hip_check(hipStreamCreateWithFlags(&S.seq_stream, hipStreamNonBlocking), "sequential stream");
    }
    if (max_plan_ws > S.seq_ws_cap) {
        if (S.seq_ws) (void)hipFree(S.seq_ws);
        S.seq_ws = nullptr;
        hip_check(hipMalloc(&S.seq_ws, max_plan_ws), "sequential workspace");
        S.seq_ws_cap = max_plan_ws;
    }
    hip_check(hipDeviceSynchronize(), "server prepare");
}

// Launches one tick on a free slot.  Returns the number of requests taken.
static void launch_tick(mifft_server& S, Slot& sl, const mifft_server_request* reqs, const std::vector<int>& take,
                        const std::vector<char>& is_batched) {
    std::vector<BatchEntry> entries;
    std::vector<int> plan_reqs;
    for (int r : take) {
        if (is_batched[r]) entries.push_back(BatchEntry{batched_plan(S, reqs[r]), 1, reqs[r].in, reqs[r].out});
        else plan_reqs.push_back(r);
    }
    BatchBuild b;
    if (!entries.empty()) batch_build(entries.data(), (int32_t)entries.size(), sl.ws, b);
    char* plan_ws = sl.ws + b.ws_bytes;
    const size_t ops_bytes = b.ops.size() * sizeof(OpDesc);
    size_t total = ops_bytes;
    for (auto& v : b.tiles) total += v.size() * sizeof(BatchTile);
    if (!entries.empty()) {
        if (total > sl.desc_cap) fail(MIFFT_ERR_OVER_BUDGET, "tick descriptors exceed the arena");
        std::memcpy(sl.staging, b.ops.data(), ops_bytes);
        const BatchTile* d_tiles[kMaxPhases];
        uint32_t counts[kMaxPhases];
        size_t off = ops_bytes;
        for (int k = 0; k < kMaxPhases; k++) {
            const size_t tb = b.tiles[k].size() * sizeof(BatchTile);
            std::memcpy(sl.staging + off, b.tiles[k].data(), tb);
            d_tiles[k] = (const BatchTile*)(sl.desc + off);
            counts[k] = (uint32_t)b.tiles[k].size();
            off += tb;
        }
        hip_check(hipMemcpyAsync(sl.desc, sl.staging, total, hipMemcpyHostToDevice, sl.s), "descriptor upload");
        batch_enqueue(d_tiles, counts, (const OpDesc*)sl.desc, sl.counters, batch_grid(b.lds_bytes), b.lds_bytes,
                      sl.s);
    }
    // Requests outside the batched launch run concurrently with it on the
    // slot's second stream (own workspace region); the tick completes when
    // both streams have.
    if (!plan_reqs.empty()) {
        for (int r : plan_reqs) run_plan(standalone(S, reqs[r]), sl.s2, reqs[r].in, reqs[r].out, plan_ws);
        hip_check(hipEventRecord(sl.ev2, sl.s2), "tick event");
        hip_check(hipStreamWaitEvent(sl.s, sl.ev2, 0), "tick join");
    }
    hip_check(hipEventRecord(sl.ev, sl.s), "tick event");
}

static void replay(mifft_server& S, const mifft_server_request* reqs, const double* arrival, int32_t n,
                   int32_t mode, mifft_server_result* res) {
    for (int32_t i = 0; i < n; i++) {
        res[i] = mifft_server_result{arrival[i], -1.0, -1.0, 0, 0};
    }
    std::vector<char> is_batched(n, 0);
    std::vector<uint64_t> mem(n, 0), desc(n, 0);
    std::map<const Plan*, uint64_t> desc_of;
    const size_t desc_cap = S.slots.empty() ? 0 : S.slots[0].desc_cap;
    for (int32_t i = 0; i < n; i++) {
        const Plan* bp = mode == 0 ? batched_plan(S, reqs[i]) : nullptr;
        if (bp) {
            auto f = desc_of.find(bp);
            desc[i] = f != desc_of.end() ? f->second : (desc_of[bp] = batch_desc_bytes(*bp, 1));
        }
        // a request whose descriptors alone overflow the tick arena runs through its plan
        // and so does one of batch_max_elems points or more: it fills the GPU
        // by itself, and its standalone kernels (tuned, three-pass splits)
        // outrun the batched classes
        uint64_t elems = reqs[i].batch;
        for (int d = 0; d < reqs[i].rank; d++) elems *= reqs[i].shape[d];
        is_batched[i] = bp != nullptr && desc[i] <= desc_cap && elems < S.batch_max_elems;
        mem[i] = reqs[i].in_bytes + reqs[i].out_bytes + request_ws(S, reqs[i], is_batched[i]);
        res[i].path = is_batched[i] ? 0 : 1;
    }
    hip_check(hipDeviceSynchronize(), "replay start");
    const auto t0 = std::chrono::steady_clock::now();
    int32_t next = 0, done = 0;
    std::deque<int> pending;
    if (mode == 1) {
        // Sequential per-plan dispatch: each arrival is enqueued at once on one
        // stream through its standalone plan; completion in stream order.
        while ((int32_t)S.seq_events.size() < n) {
            hipEvent_t e;
            hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
            S.seq_events.push_back(e);
        }
        int32_t head = 0;
        while (done < n) {
            const double t = now_s(t0);
            while (next < n && arrival[next] <= t) {
                run_plan(standalone(S, reqs[next]), S.seq_stream, reqs[next].in, reqs[next].out, S.seq_ws);
                hip_check(hipEventRecord(S.seq_events[next], S.seq_stream), "event");
                res[next].admit_s = now_s(t0);
                next++;
            }
            while (head < next && hipEventQuery(S.seq_events[head]) == hipSuccess) {
                res[head].done_s = now_s(t0);
                head++;
                done++;
            }
        }
        return;
    }
    uint64_t inflight = 0;
    while (done < n) {
        double t = now_s(t0);
        while (next < n && arrival[next] <= t) pending.push_back(next++);
        for (Slot& sl : S.slots) {
            if (sl.busy && hipEventQuery(sl.ev) == hipSuccess) {
                t = now_s(t0);
                for (int r : sl.reqs) {
                    res[r].done_s = t;
                    done++;
                }
                inflight -= sl.mem;
                sl.busy = false;
                sl.reqs.clear();
                sl.mem = 0;
            }
        }
        if (pending.empty()) continue;
        Slot* free_slot = nullptr;
        for (Slot& sl : S.slots)
            if (!sl.busy) {
                free_slot = &sl;
                break;
            }
        if (!free_slot) continue;
        std::vector<int> take;
        uint64_t tick_mem = 0, batch_ws = 0, plan_ws = 0, tick_desc = 0;
        while (!pending.empty() && (int)take.size() < S.max_items) {
            const int r = pending.front();
            const uint64_t ws = mem[r] - reqs[r].in_bytes - reqs[r].out_bytes;
            if (mem[r] > S.budget || ws > free_slot->ws_cap) {
                pending.pop_front();  // can never be admitted
                res[r].status = MIFFT_ERR_OVER_BUDGET;
                res[r].admit_s = res[r].done_s = now_s(t0);
                res[r].path = 2;
                done++;
                continue;
            }
            const uint64_t nb = is_batched[r] ? batch_ws + ws : batch_ws;
            const uint64_t np = is_batched[r] ? plan_ws : std::max(plan_ws, ws);
            const uint64_t nd = is_batched[r] ? tick_desc + desc[r] : tick_desc;
            if (inflight + tick_mem + mem[r] > S.budget || nb + np > free_slot->ws_cap || nd > desc_cap) break;
            tick_desc = nd;
            batch_ws = nb;
            plan_ws = np;
            tick_mem += mem[r];
            take.push_back(r);
            pending.pop_front();
        }
        if (take.empty()) continue;
        const double ta = now_s(t0);
        for (int r : take) res[r].admit_s = ta;
        launch_tick(S, *free_slot, reqs, take, is_batched);
        free_slot->busy = true;
        free_slot->reqs = take;
        free_slot->mem = tick_mem;
        inflight += tick_mem;
    }
}

// ------------------------------------------------------------- online path
static double online_now(mifft_server& S) {
    if (!S.clock_started) {
        S.t0 = std::chrono::steady_clock::now();
        S.clock_started = true;
    }
    return now_s(S.t0);
}

static size_t online_busy(const mifft_server& S) {
    size_t n = S.queue.size();
    for (const auto& v : S.slot_items) n += v.size();
    return n;
}

// Frees cache room for a new shape: drops least recently used shapes that
// nothing in flight or queued uses.
static void evict_for(mifft_server& S, const PlanKey& k) {
    if (S.plans.count(k) || S.plans.size() < S.max_shapes) return;
    std::vector<std::pair<uint64_t, PlanKey>> cand;
    for (auto& kv : S.plans) {
        auto r = S.refs.find(kv.first);
        if (r == S.refs.end() || r->second == 0) cand.emplace_back(S.last_use[kv.first], kv.first);
    }
    std::sort(cand.begin(), cand.end());
    for (size_t i = 0; i < cand.size() && S.plans.size() >= S.max_shapes; i++) {
        const PlanKey& v = cand[i].second;
        mifft_server_request r{};
        r.kind = v.kind;
        r.precision = v.prec;
        r.rank = v.rank;
        r.direction = v.dir;
        r.normalization = v.norm;
        r.shape[0] = v.s0;
        r.shape[1] = v.s1;
        r.batch = v.batch;
        S.bcache.erase(config_of(r), (mifft_kind)v.kind);
        S.plans.erase(v);
        S.last_use.erase(v);
        S.refs.erase(v);
    }
}

// Completes finished ticks, then admits queued requests into free slots as
// the replay loop does (the same admission rules, one tick per free slot).
static void online_progress(mifft_server& S) {
    for (size_t i = 0; i < S.slots.size(); i++) {
        Slot& sl = S.slots[i];
        if (!sl.busy || hipEventQuery(sl.ev) != hipSuccess) continue;
        const double t = online_now(S);
        for (auto& it : S.slot_items[i]) {
            S.completions.push_back(
                mifft_server_completion{it.ticket, it.submit_s, it.admit_s, t, MIFFT_SUCCESS, it.batched ? 0 : 1});
            S.refs[key_of(it.r)]--;
        }
        S.online_inflight -= sl.mem;
        S.slot_items[i].clear();
        sl.busy = false;
        sl.mem = 0;
    }
    const size_t desc_cap = S.slots.empty() ? 0 : S.slots[0].desc_cap;
    for (size_t i = 0; i < S.slots.size() && !S.queue.empty(); i++) {
        Slot& sl = S.slots[i];
        if (sl.busy) continue;
        std::vector<mifft_server_request> reqs;
        std::vector<char> batched;
        std::vector<mifft_server::Item> items;
        uint64_t tick_mem = 0, batch_ws = 0, plan_ws = 0, tick_desc = 0;
        while (!S.queue.empty() && (int)items.size() < S.max_items) {
            const mifft_server::Item& it = S.queue.front();
            const uint64_t ws = it.mem - it.r.in_bytes - it.r.out_bytes;
            const uint64_t nb = it.batched ? batch_ws + ws : batch_ws;
            const uint64_t np = it.batched ? plan_ws : std::max(plan_ws, ws);
            const uint64_t nd = it.batched ? tick_desc + it.desc : tick_desc;
            if (S.online_inflight + tick_mem + it.mem > S.budget || nb + np > sl.ws_cap || nd > desc_cap) break;
            tick_desc = nd;
            batch_ws = nb;
            plan_ws = np;
            tick_mem += it.mem;
            items.push_back(it);
            S.queue.pop_front();
        }
        if (items.empty()) break;  // the head waits for memory to free up
        const double ta = online_now(S);
        std::vector<int> take;
        for (size_t k = 0; k < items.size(); k++) {
            items[k].admit_s = ta;
            reqs.push_back(items[k].r);
            batched.push_back(items[k].batched ? 1 : 0);
            take.push_back((int)k);
        }
        launch_tick(S, sl, reqs.data(), take, batched);
        sl.busy = true;
        sl.mem = tick_mem;
        S.online_inflight += tick_mem;
        S.slot_items[i] = std::move(items);
    }
}

static uint64_t online_submit(mifft_server& S, const mifft_server_request& r) {
    if (!r.in || !r.out) fail(MIFFT_ERR_NULL_ARGUMENT, "null request buffer");
    if (((uintptr_t)r.in | (uintptr_t)r.out) & 15u) fail(MIFFT_ERR_UNALIGNED, "device buffers must be 16-byte aligned");
    if (S.slots.empty()) prepare(S, nullptr, 0);
    if (S.slot_items.size() != S.slots.size()) S.slot_items.resize(S.slots.size());
    const PlanKey k = key_of(r);
    evict_for(S, k);
    // the plans of a new shape are built here, on first sight (plan tables
    // upload synchronously on the plan's own stream)
    const Plan& sp = standalone(S, r);
    if (r.in_bytes < sp.in_bytes || r.out_bytes < sp.out_bytes)
        fail(MIFFT_ERR_BUFFER_TOO_SMALL, "request buffers smaller than the transform needs");
    if (r.in == r.out && !sp.in_place_kernels)
        fail(MIFFT_ERR_IN_PLACE_UNSUPPORTED, "served real transforms run out of place");
    const Plan* bp = batched_plan(S, r);
    mifft_server::Item it;
    it.r = r;
    if (bp) it.desc = batch_desc_bytes(*bp, 1);
    uint64_t elems = r.batch;
    for (int d = 0; d < r.rank; d++) elems *= r.shape[d];
    const size_t desc_cap = S.slots[0].desc_cap;
    it.batched = bp != nullptr && it.desc <= desc_cap && elems < S.batch_max_elems;
    const uint64_t ws = it.batched ? bp->ws_bytes : sp.ws_bytes;
    it.mem = r.in_bytes + r.out_bytes + ws;
    if (it.mem > S.budget || ws > S.slots[0].ws_cap)
        fail(MIFFT_ERR_OVER_BUDGET, "request needs " + std::to_string(it.mem) + " bytes; the budget is " +
                                        std::to_string(S.budget));
    it.ticket = S.next_ticket++;
    it.submit_s = online_now(S);
    S.last_use[k] = ++S.use_clock;
    S.refs[k]++;
    S.queue.push_back(it);
    online_progress(S);
    return it.ticket;
}

template <class F>
static mifft_status sguard(F&& f) {
    try {
        f();
        return MIFFT_SUCCESS;
    } catch (const Error& e) {
        set_last_detail(e.detail);
        return e.code;
    } catch (...) {
        set_last_detail("unexpected internal error");
        return MIFFT_ERR_HIP;
    }
}

extern "C" {

mifft_status mifft_server_create(uint64_t memory_budget_bytes, int32_t streams, int32_t max_tick_requests,
                                 mifft_server** out) {
    if (!out) return MIFFT_ERR_NULL_ARGUMENT;
    *out = nullptr;
    if (streams < 1 || streams > 4 || max_tick_requests < 1 || memory_budget_bytes == 0) {
        set_last_detail("streams must be 1..4, max_tick_requests >= 1, budget > 0");
        return MIFFT_ERR_INVALID_ENUM;
    }
    return sguard([&] {
        auto s = std::make_unique<mifft_server>();
        s->budget = memory_budget_bytes;
        s->nslots = streams;
        s->max_items = max_tick_requests;
        *out = s.release();
    });
}

mifft_status mifft_server_set_batch_limit(mifft_server* s, uint64_t max_points) {
    if (!s) return MIFFT_ERR_NULL_ARGUMENT;
    s->batch_max_elems = max_points;
    return MIFFT_SUCCESS;
}

mifft_status mifft_server_prepare(mifft_server* s, const mifft_server_request* reqs, int32_t n) {
    if (!s || (!reqs && n)) return MIFFT_ERR_NULL_ARGUMENT;
    return sguard([&] { prepare(*s, reqs, n); });
}

mifft_status mifft_server_replay(mifft_server* s, const mifft_server_request* reqs, const double* arrival_s,
                                 int32_t n, int32_t mode, mifft_server_result* results) {
    if (!s || !reqs || !arrival_s || !results) return MIFFT_ERR_NULL_ARGUMENT;
    if (mode < 0 || mode > 1) return MIFFT_ERR_INVALID_ENUM;
    if (online_busy(*s) > 0) {
        set_last_detail("submitted requests are still queued or in flight: poll them first");
        return MIFFT_ERR_BUSY;
    }
    return sguard([&] {
        prepare(*s, reqs, n);
        replay(*s, reqs, arrival_s, n, mode, results);
        hip_check(hipDeviceSynchronize(), "replay end");
    });
}

mifft_status mifft_server_submit(mifft_server* s, const mifft_server_request* req, uint64_t* ticket) {
    if (!s || !req || !ticket) return MIFFT_ERR_NULL_ARGUMENT;
    *ticket = 0;
    return sguard([&] { *ticket = online_submit(*s, *req); });
}

mifft_status mifft_server_poll(mifft_server* s, int32_t wait, mifft_server_completion* out, int32_t cap,
                               int32_t* count) {
    if (!s || !count || (cap > 0 && !out)) return MIFFT_ERR_NULL_ARGUMENT;
    *count = 0;
    return sguard([&] {
        online_progress(*s);
        while (wait && s->completions.empty() && online_busy(*s) > 0) online_progress(*s);
        int32_t n = 0;
        while (n < cap && !s->completions.empty()) {
            out[n++] = s->completions.front();
            s->completions.pop_front();
        }
        *count = n;
    });
}

mifft_status mifft_server_pending(const mifft_server* s, uint64_t* queued, uint64_t* in_flight, uint64_t* completed) {
    if (!s) return MIFFT_ERR_NULL_ARGUMENT;
    uint64_t f = 0;
    for (const auto& v : s->slot_items) f += v.size();
    if (queued) *queued = s->queue.size();
    if (in_flight) *in_flight = f;
    if (completed) *completed = s->completions.size();
    return MIFFT_SUCCESS;
}

mifft_status mifft_server_set_plan_cache(mifft_server* s, uint64_t max_shapes) {
    if (!s) return MIFFT_ERR_NULL_ARGUMENT;
    if (max_shapes < 1) {
        set_last_detail("the plan cache holds at least one shape");
        return MIFFT_ERR_INVALID_ENUM;
    }
    s->max_shapes = (size_t)max_shapes;
    return MIFFT_SUCCESS;
}

uint64_t mifft_server_cached_shapes(const mifft_server* s) { return s ? s->plans.size() : 0; }

void mifft_server_destroy(mifft_server* s) {
    if (s) {
        for (Slot& sl : s->slots)
            if (sl.busy) (void)hipEventSynchronize(sl.ev);
    }
    delete s;
}

}  // extern "C"
