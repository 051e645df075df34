// Host planner of the persistent batched launch (mifft_batch_*).
//
// Each transform reuses the planner of standalone plans in batched mode
// (256-thread kernels, an 80 KiB LDS budget so that at least two workgroups share a
// CU).  Plans are cached, so the twiddle, chirp and filter tables of equal
// shapes are shared.  Op k of every transform becomes a run of tiles in phase
// k; tiles are sorted by cost (longest first) inside their phase and pulled by
// the persistent kernel from a device counter.
#include <algorithm>
#include <cstring>

#include "batch_plan.hpp"

namespace mifft {

BatchPlanCache::BatchPlanCache() { mifft_tuning_init(&tuning); }

// Ops of a plan in launch order, child plans (the padded transforms of a
// multi-pass Bluestein axis) expanded in place.
static void flat_ops(const Plan& p, std::vector<const Op*>& out) {
    for (const Op& op : p.ops) {
        if (op.kind == OP_SUBPLAN) flat_ops(*op.sub, out);
        else out.push_back(&op);
    }
}

// Batched class of the register Bluestein kernel of convolution length m.
static uint32_t regb_batch_class(int m) {
    switch (m) {
        case 64: return BC_REGB_64;
        case 128: return BC_REGB_128;
        case 256: return BC_REGB_256;
        case 512: return BC_REGB_512;
        case 1024: return BC_REGB_1024;
        case 2048: return BC_REGB_2048;
        case 4096: return BC_REGB_4096;
        case 8192: return BC_REGB_8192;
    }
    fail(MIFFT_ERR_UNSUPPORTED_LENGTH, "no batched class for register Bluestein M=" + std::to_string(m));
    return 0;
}

static std::string batch_reject_reason(const Plan& p) {
    std::vector<const Op*> ops;
    flat_ops(p, ops);
    if (ops.size() > (size_t)kMaxPhases) return "more than " + std::to_string(kMaxPhases) + " dependent passes";
    for (const Op* op : ops) {
        bool ok = false;
        switch (op->kind) {
            case OP_REG: ok = batch_class_for(op->rv) >= 0; break;
            case OP_LDS: ok = batch_lds_class(op->prec, op->ldm, op->stm, op->post) >= 0; break;
            case OP_RADER:
            case OP_BLUESTEIN:
            case OP_CHIRP_PRE:
            case OP_POINTWISE:
            case OP_CHIRP_POST: ok = op->prec == P_F32 || op->prec == P_F64; break;
            default: ok = false;
        }
        if (!ok) return "no batched class for " + op->desc;
    }
    return "";
}

static std::unique_ptr<Plan> build_batched(const mifft_config& c, mifft_kind kind) {
    PlanOptions opt;
    opt.lds_budget = kBatchLdsBudget;
    opt.batch_mode = true;
    return build_plan(c, kind, opt);
}

const Plan* BatchPlanCache::get_1d(uint64_t n, int direction, int normalization, int algorithm,
                                   std::string* why) {
    auto key = std::make_tuple(n, direction, normalization, algorithm);
    auto f = plans_1d.find(key);
    if (f != plans_1d.end()) return f->second.get();
    mifft_config c;
    std::memset(&c, 0, sizeof(c));
    c.rank = 1;
    c.shape[0] = n;
    c.naxes = 1;
    c.axes[0] = 0;
    c.batch = 1;
    c.direction = direction;
    c.normalization = normalization;
    c.precision = MIFFT_F32;
    c.tuning = tuning;
    c.tuning.n_force_rader_axes = c.tuning.n_force_bluestein_axes = 0;
    if (algorithm == 1) {
        c.tuning.force_rader_axes[0] = 0;
        c.tuning.n_force_rader_axes = 1;
    } else if (algorithm == 2) {
        c.tuning.force_bluestein_axes[0] = 0;
        c.tuning.n_force_bluestein_axes = 1;
    }
    std::unique_ptr<Plan> p;
    try {
        p = build_batched(c, MIFFT_C2C);
    } catch (const Error& e) {
        if (e.code != MIFFT_ERR_UNSUPPORTED_LENGTH) throw;
        if (why) *why = e.detail;
        return nullptr;
    }
    const std::string r = batch_reject_reason(*p);
    if (!r.empty()) {
        if (why) *why = "length " + std::to_string(n) + ": " + r;
        return nullptr;
    }
    const Plan* out = p.get();
    plans_1d[key] = std::move(p);
    return out;
}

static std::string config_key(const mifft_config& c, mifft_kind kind) {
    std::string k = std::to_string((int)kind) + "|";
    for (int i = 0; i < c.rank; i++) k += std::to_string(c.shape[i]) + ",";
    k += "|";
    for (int i = 0; i < c.naxes; i++) k += std::to_string(c.axes[i]) + ",";
    k += "|" + std::to_string(c.batch) + "|" + std::to_string(c.direction) + "|" + std::to_string(c.normalization) +
         "|" + std::to_string(c.precision);
    for (int i = 0; i < c.tuning.n_force_rader_axes; i++) k += "|r" + std::to_string(c.tuning.force_rader_axes[i]);
    for (int i = 0; i < c.tuning.n_force_bluestein_axes; i++)
        k += "|b" + std::to_string(c.tuning.force_bluestein_axes[i]);
    return k;
}

void BatchPlanCache::erase(const mifft_config& c, mifft_kind kind) {
    const std::string key = config_key(c, kind);
    plans.erase(key);
    rejected.erase(key);
}

const Plan* BatchPlanCache::get(const mifft_config& c, mifft_kind kind, std::string* why) {
    const std::string key = config_key(c, kind);
    auto f = plans.find(key);
    if (f != plans.end()) return f->second.get();
    auto rj = rejected.find(key);
    if (rj != rejected.end()) {
        if (why) *why = rj->second;
        return nullptr;
    }
    std::unique_ptr<Plan> p;
    try {
        p = build_batched(c, kind);
    } catch (const Error& e) {
        if (e.code != MIFFT_ERR_UNSUPPORTED_LENGTH && e.code != MIFFT_ERR_PRECISION_UNSUPPORTED) throw;
        rejected[key] = e.detail;
        if (why) *why = e.detail;
        return nullptr;
    }
    const std::string r = batch_reject_reason(*p);
    if (!r.empty()) {
        rejected[key] = r;
        if (why) *why = r;
        return nullptr;
    }
    const Plan* out = p.get();
    plans[key] = std::move(p);
    return out;
}

// Device addresses of a plan's buffer references.  A plan built for one line
// and run over `mult` lines keeps its input and output bases and scales its
// workspace offsets (and everything inside the workspace) by `mult`.
struct Binding {
    char* base[3];
    uint64_t scale[3];
    char* at(int ref, uint64_t off) const { return base[ref] + off * scale[ref]; }
};

static void bind_ops(const Plan& p, const Binding& b, uint64_t mult,
                     std::vector<std::pair<const Op*, Binding>>& out) {
    for (const Op& op : p.ops) {
        if (op.kind == OP_SUBPLAN) {
            char* buf = b.at(op.dst, op.dst_off);
            const Binding cb{{buf, buf, b.at(B_WS, op.sub_ws_off)}, {mult, mult, mult}};
            bind_ops(*op.sub, cb, mult, out);
        } else {
            out.emplace_back(&op, b);
        }
    }
}

void batch_build(const BatchEntry* entries, int32_t count, char* ws_base, BatchBuild& out) {
    size_t first[kMaxPhases];
    for (int k = 0; k < kMaxPhases; k++) first[k] = out.tiles[k].size();
    std::vector<std::pair<const Op*, Binding>> flat;
    for (int32_t i = 0; i < count; i++) {
        const BatchEntry& e = entries[i];
        const Plan& P = *e.plan;
        char* ws = ws_base + out.ws_bytes;
        out.ws_bytes += P.ws_bytes * e.mult;
        flat.clear();
        bind_ops(P, Binding{{(char*)e.in, (char*)e.out, ws}, {1, 1, e.mult}}, e.mult, flat);
        if (flat.size() > (size_t)kMaxPhases) fail(MIFFT_ERR_UNSUPPORTED_LENGTH, "too many batched passes");
        for (size_t j = 0; j < flat.size(); j++) {
            const Op& op = *flat[j].first;
            const Binding& bd = flat[j].second;
            OpDesc d;
            std::memset(&d, 0, sizeof(d));
            uint32_t cls = 0, lpw = 1;
            size_t op_lds = 0;
            const size_t es = P.arith == P_F32 ? 8 : 16;
            const bool f64 = op.prec == P_F64;
            bool elt = false;
            uint64_t unit = 0;  // points one line of the op touches (scheduling weight)
            switch (op.kind) {
                case OP_REG:
                    cls = (uint32_t)batch_class_for(op.rv);
                    d.a = op.a;
                    lpw = (uint32_t)op.rv->lines;
                    op_lds = op.rv->lds_bytes;
                    unit = op.rv->n;
                    break;
                case OP_LDS:
                    cls = (uint32_t)batch_lds_class(op.prec, op.ldm, op.stm, op.post);
                    d.a = op.a;
                    d.sc = op.sc;
                    lpw = op.sc.lpw;
                    op_lds = lds_kernel_bytes(op.prec, op.sc.n, op.sc.lpw);
                    unit = op.sc.n;
                    break;
                case OP_RADER:
                case OP_BLUESTEIN:
                    if (f64) cls = op.kind == OP_RADER ? BC_RADER_F64 : BC_BLUESTEIN_F64;
                    else if (op.kind == OP_RADER) cls = sched_has_medium_prime(op.conv.sc) ? BC_RADER_MED : BC_RADER;
                    else cls = BC_BLUESTEIN;
                    d.conv = op.conv;
                    lpw = op.conv.sc.lpw;
                    op_lds = conv_kernel_bytes(op.prec, op.conv.m, op.conv.sc.lpw);
                    if (op.rcv) {
                        // the register Bluestein kernel's own class (same tile body)
                        cls = regb_batch_class(op.rcv->m);
                        lpw = (uint32_t)op.rcv->lines;
                        op_lds = op.rcv->lds_bytes;
                    }
                    unit = 3ull * op.conv.m;  // convolutions count 3x
                    break;
                case OP_CHIRP_PRE:
                case OP_CHIRP_POST:
                    cls = op.kind == OP_CHIRP_PRE ? (f64 ? BC_CHIRP_PRE_F64 : BC_CHIRP_PRE)
                                                  : (f64 ? BC_CHIRP_POST_F64 : BC_CHIRP_POST);
                    d.conv.io = op.a;
                    d.conv.chirp = op.chirp;
                    d.conv.n = op.n;
                    d.conv.m = op.m;
                    elt = true;
                    break;
                case OP_POINTWISE:
                    cls = f64 ? BC_POINTWISE_F64 : BC_POINTWISE;
                    d.conv.io.lines = (uint32_t)op.count;
                    d.conv.filt = op.filt;
                    d.conv.m = op.m;
                    elt = true;
                    break;
                default: fail(MIFFT_ERR_UNSUPPORTED_LENGTH, "no batched class for " + op.desc);
            }
            const bool conv = op.kind != OP_REG && op.kind != OP_LDS;
            LineArgs& la = conv ? d.conv.io : d.a;
            const unsigned __int128 L = (unsigned __int128)la.lines * e.mult;
            if (L > 0x7FFFFFFFu) fail(MIFFT_ERR_LENGTH_TOO_LARGE, "batched transform lines exceed 2^31");
            la.lines = (uint32_t)L;
            if (op.kind == OP_POINTWISE) {
                la.in = la.out = bd.at(op.dst, op.dst_off);
            } else {
                la.in = bd.at(op.src, op.src_off);
                la.out = bd.at(op.dst, op.dst_off);
            }
            out.lds_bytes = std::max(out.lds_bytes, op_lds);
            const uint32_t arg = (uint32_t)out.ops.size();
            out.ops.push_back(d);
            uint32_t ntiles, step;
            uint64_t c64;
            if (elt) {
                // element-wise tiles index lines x m (lines x n for the chirp post) in 32 bits
                const uint64_t total = (uint64_t)la.lines * (op.kind == OP_CHIRP_POST ? op.n : op.m);
                if (total > 0xFFFFFFFFull - kEltTile)
                    fail(MIFFT_ERR_LENGTH_TOO_LARGE, "batched Bluestein workspace exceeds 2^32 points");
                step = kEltTile;
                ntiles = (uint32_t)((total + kEltTile - 1) / kEltTile);
                c64 = (uint64_t)kEltTile * es * 2;
            } else {
                step = lpw;
                ntiles = (la.lines + lpw - 1) / lpw;
                c64 = (uint64_t)lpw * es * unit;
            }
            const uint32_t cost = (uint32_t)std::min<uint64_t>(c64, 0xFFFFFFFFu);
            for (uint32_t t = 0; t < ntiles; t++) out.tiles[j].push_back(BatchTile{cls, arg, t * step, cost});
        }
    }
    for (int k = 0; k < kMaxPhases; k++)
        std::stable_sort(out.tiles[k].begin() + first[k], out.tiles[k].end(),
                         [](const BatchTile& a, const BatchTile& b) { return a.cost > b.cost; });
}

uint64_t batch_desc_bytes(const Plan& p, uint64_t mult) {
    // Bind to placeholder addresses: only the descriptor and tile counts are used.
    static char placeholder[16];
    const BatchEntry e{&p, mult, placeholder, placeholder};
    BatchBuild b;
    batch_build(&e, 1, placeholder, b);
    return b.ops.size() * sizeof(OpDesc) + b.tile_count() * sizeof(BatchTile);
}

uint32_t batch_grid(size_t lds_bytes) {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
    }
    const size_t per_cu = std::max<size_t>(1, std::min<size_t>(4, kLdsBudget / std::max<size_t>(lds_bytes, 16)));
    return (uint32_t)(cus * per_cu);
}

void batch_enqueue(const BatchTile* const* d_tiles, const uint32_t* counts, const OpDesc* d_ops,
                   uint32_t* d_counters, uint32_t grid, size_t lds_bytes, hipStream_t s) {
    hip_check(launch_batch_reset(d_counters, kMaxPhases, s), "batch reset");
    for (int k = 0; k < kMaxPhases; k++)
        if (counts[k])
            hip_check(launch_batch(d_tiles[k], counts[k], d_counters + k, d_ops, std::min(grid, counts[k]),
                                   lds_bytes, s),
                      "batch");
}

}  // namespace mifft

using namespace mifft;

struct mifft_batch {
    BatchPlanCache cache;
    BatchBuild build;
    OpDesc* d_ops = nullptr;
    BatchTile* d_tiles[kMaxPhases] = {};
    uint32_t* d_counters = nullptr;
    void* ws = nullptr;
    uint32_t grid = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~mifft_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (d_ops) (void)hipFree(d_ops);
        for (auto* t : d_tiles)
            if (t) (void)hipFree(t);
        if (d_counters) (void)hipFree(d_counters);
        if (ws) (void)hipFree(ws);
    }
};

static void check_item(const mifft_batch_item& it) {
    if (it.n == 0) fail(MIFFT_ERR_ZERO_LENGTH, "batch item: zero length");
    if (it.batch == 0) fail(MIFFT_ERR_ZERO_BATCH, "batch item: zero batch");
    if (!it.in || !it.out) fail(MIFFT_ERR_NULL_ARGUMENT, "batch item: null buffer");
    if (((uintptr_t)it.in | (uintptr_t)it.out) & 15u)
        fail(MIFFT_ERR_UNALIGNED, "batch item: buffers must be 16-byte aligned");
    if (it.algorithm < 0 || it.algorithm > 2 || it.direction < 0 || it.direction > 1 || it.normalization < 0 ||
        it.normalization > 3)
        fail(MIFFT_ERR_INVALID_ENUM, "batch item: enum out of range");
    if ((unsigned __int128)it.n * it.batch > 0xFFFFFFFFull)
        fail(MIFFT_ERR_LENGTH_TOO_LARGE, "batch item exceeds u32::MAX elements");
}

static void create_batch(mifft_batch& B, const mifft_batch_item* items, int32_t count, const mifft_tuning* t) {
    if (t) B.cache.tuning = *t;
    std::vector<BatchEntry> entries;
    uint64_t ws = 0;
    for (int32_t i = 0; i < count; i++) {
        const mifft_batch_item& it = items[i];
        check_item(it);
        std::string why;
        const Plan* p = B.cache.get_1d(it.n, it.direction, it.normalization, it.algorithm, &why);
        if (!p) fail(MIFFT_ERR_UNSUPPORTED_LENGTH, "batch item " + std::to_string(i) + ": " + why);
        entries.push_back(BatchEntry{p, it.batch, it.in, it.out});
        ws += p->ws_bytes * it.batch;
    }
    if (ws) hip_check(hipMalloc(&B.ws, ws), "batch workspace");
    batch_build(entries.data(), (int32_t)entries.size(), (char*)B.ws, B.build);
    B.grid = batch_grid(B.build.lds_bytes);
    auto upload = [](void** dst, const void* src, size_t bytes, const char* what) {
        hip_check(hipMalloc(dst, std::max<size_t>(bytes, 16)), what);
        if (bytes) hip_check(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice), what);
    };
    upload((void**)&B.d_ops, B.build.ops.data(), B.build.ops.size() * sizeof(OpDesc), "batch ops");
    for (int k = 0; k < kMaxPhases; k++)
        upload((void**)&B.d_tiles[k], B.build.tiles[k].data(), B.build.tiles[k].size() * sizeof(BatchTile),
               "batch tiles");
    hip_check(hipMalloc(&B.d_counters, 64), "batch counters");
}

static void run_batch(const mifft_batch& B, hipStream_t s) {
    uint32_t counts[kMaxPhases];
    for (int k = 0; k < kMaxPhases; k++) counts[k] = (uint32_t)B.build.tiles[k].size();
    batch_enqueue(B.d_tiles, counts, B.d_ops, B.d_counters, B.grid, B.build.lds_bytes, s);
}

template <class F>
static mifft_status bguard(F&& f) {
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

mifft_status mifft_batch_create(const mifft_batch_item* items, int32_t count, const mifft_tuning* tuning,
                                mifft_batch** out) {
    if (!items || !out || count <= 0) return MIFFT_ERR_NULL_ARGUMENT;
    *out = nullptr;
    return bguard([&] {
        auto b = std::make_unique<mifft_batch>();
        create_batch(*b, items, count, tuning);
        *out = b.release();
    });
}

mifft_status mifft_batch_execute(mifft_batch* b, hipStream_t stream) {
    if (!b) return MIFFT_ERR_NULL_ARGUMENT;
    return bguard([&] { run_batch(*b, stream); });
}

mifft_status mifft_batch_capture(mifft_batch* b, hipStream_t stream) {
    if (!b) return MIFFT_ERR_NULL_ARGUMENT;
    return bguard([&] {
        if (b->exec) return;
        hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture");
        try {
            run_batch(*b, stream);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(stream, &b->graph), "end capture");
        hip_check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "graph instantiate");
    });
}

mifft_status mifft_batch_launch(mifft_batch* b, hipStream_t stream) {
    if (!b) return MIFFT_ERR_NULL_ARGUMENT;
    if (!b->exec) return MIFFT_ERR_RECORDER_STATE;
    return bguard([&] { hip_check(hipGraphLaunch(b->exec, stream), "graph launch"); });
}

mifft_status mifft_batch_info(const mifft_batch* b, uint64_t* tiles, uint64_t* phases, uint64_t* lds_bytes,
                              uint64_t* workspace_bytes, uint32_t* grid) {
    if (!b) return MIFFT_ERR_NULL_ARGUMENT;
    if (tiles) *tiles = b->build.tile_count();
    if (phases) *phases = (uint64_t)b->build.phases();
    if (lds_bytes) *lds_bytes = b->build.lds_bytes;
    if (workspace_bytes) *workspace_bytes = b->build.ws_bytes;
    if (grid) *grid = b->grid;
    return MIFFT_SUCCESS;
}

void mifft_batch_destroy(mifft_batch* b) { delete b; }

}  // extern "C"
