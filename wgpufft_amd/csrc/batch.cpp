// Host planner of the persistent batched launch (mifft_batch_*).
//
// Each item (a 1D f32 transform of `n` points over `batch` lines) reuses the
// planner of standalone plans in batched mode (256-thread kernels, a 40 KiB
// LDS budget so that four workgroups share a CU); plans are cached by
// (n, direction, normalization, algorithm), so the twiddle, chirp and filter
// tables of equal lengths are shared.  Every op of every item becomes a run of
// tiles; tiles are sorted by cost (longest first) inside their phase and
// pulled by the persistent kernel from a device counter.
#include <algorithm>
#include <cstring>

#include "batch_plan.hpp"

namespace mifft {

BatchPlanCache::BatchPlanCache() { mifft_tuning_init(&tuning); }

const Plan* BatchPlanCache::get(uint64_t n, int direction, int normalization, int algorithm) {
    auto key = std::make_tuple(n, direction, normalization, algorithm);
    auto f = plans.find(key);
    if (f != plans.end()) return f->second.get();
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
    PlanOptions opt;
    opt.lds_budget = kBatchLdsBudget;
    opt.batch_mode = true;
    std::unique_ptr<Plan> p = build_plan(c, MIFFT_C2C, opt);
    if (p->ops.size() > 2)
        fail(MIFFT_ERR_UNSUPPORTED_LENGTH, "length " + std::to_string(n) + " needs more than two passes");
    for (const Op& op : p->ops) {
        const bool ok = (op.kind == OP_REG && batch_class_for(op.rv) >= 0) ||
                        (op.kind == OP_LDS && op.ldm == LD_COMPLEX && op.stm == ST_COMPLEX) ||
                        op.kind == OP_RADER || op.kind == OP_BLUESTEIN;
        if (!ok)
            fail(MIFFT_ERR_UNSUPPORTED_LENGTH,
                 "length " + std::to_string(n) + " needs a route the batched kernel does not run (" + op.desc + ")");
    }
    const Plan* r = p.get();
    plans[key] = std::move(p);
    return r;
}

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

uint64_t batch_item_workspace(BatchPlanCache& cache, const mifft_batch_item& it) {
    check_item(it);
    const Plan* P = cache.get(it.n, it.direction, it.normalization, it.algorithm);
    return P->ws_bytes * it.batch;
}

bool batch_item_supported(BatchPlanCache& cache, const mifft_batch_item& it) {
    try {
        batch_item_workspace(cache, it);
        return true;
    } catch (const Error&) {
        return false;
    }
}

void batch_build(BatchPlanCache& cache, const mifft_batch_item* items, int32_t count, char* ws_base,
                 BatchBuild& out) {
    const size_t first[2] = {out.tiles[0].size(), out.tiles[1].size()};
    for (int32_t i = 0; i < count; i++) {
        const mifft_batch_item& it = items[i];
        check_item(it);
        const Plan& P = *cache.get(it.n, it.direction, it.normalization, it.algorithm);
        char* ws = ws_base + out.ws_bytes;
        out.ws_bytes += P.ws_bytes * it.batch;
        for (size_t j = 0; j < P.ops.size(); j++) {
            const Op& op = P.ops[j];
            // Plans are built for one line: workspace offsets scale with the line count.
            auto bind = [&](int ref, uint64_t off) -> char* {
                if (ref == B_IN) return (char*)it.in + off;
                if (ref == B_OUT) return (char*)it.out + off;
                return ws + off * it.batch;
            };
            OpDesc d;
            std::memset(&d, 0, sizeof(d));
            uint32_t cls = 0, lpw = 1;
            size_t op_lds = 0;
            switch (op.kind) {
                case OP_REG:
                    cls = (uint32_t)batch_class_for(op.rv);
                    d.a = op.a;
                    lpw = (uint32_t)op.rv->lines;
                    op_lds = op.rv->lds_bytes;
                    break;
                case OP_LDS:
                    cls = op.post ? BC_LDS_POST : BC_LDS;
                    d.a = op.a;
                    d.sc = op.sc;
                    lpw = op.sc.lpw;
                    op_lds = lds_kernel_bytes(P_F32, op.sc.n, op.sc.lpw);
                    break;
                default:  // OP_RADER, OP_BLUESTEIN (BatchPlanCache::get admits nothing else)
                    cls = op.kind == OP_RADER ? BC_RADER : BC_BLUESTEIN;
                    d.conv = op.conv;
                    lpw = op.conv.sc.lpw;
                    op_lds = conv_kernel_bytes(P_F32, op.conv.m, op.conv.sc.lpw);
                    break;
            }
            LineArgs& la = (op.kind == OP_RADER || op.kind == OP_BLUESTEIN) ? d.conv.io : d.a;
            const unsigned __int128 L = (unsigned __int128)la.lines * it.batch;
            if (L > 0x7FFFFFFFu) fail(MIFFT_ERR_LENGTH_TOO_LARGE, "batch item lines exceed 2^31");
            la.lines = (uint32_t)L;
            la.in = bind(op.src, op.src_off);
            la.out = bind(op.dst, op.dst_off);
            out.lds_bytes = std::max(out.lds_bytes, op_lds);
            const uint32_t arg = (uint32_t)out.ops.size();
            out.ops.push_back(d);
            const uint32_t ntiles = (la.lines + lpw - 1) / lpw;
            const uint64_t c64 = (uint64_t)lpw * it.n * (op.kind == OP_BLUESTEIN || op.kind == OP_RADER ? 3 : 1);
            const uint32_t cost = (uint32_t)std::min<uint64_t>(c64, 0xFFFFFFFFu);
            for (uint32_t t = 0; t < ntiles; t++) out.tiles[j].push_back(BatchTile{cls, arg, t * lpw, cost});
        }
    }
    for (int p = 0; p < 2; p++)
        std::stable_sort(out.tiles[p].begin() + first[p], out.tiles[p].end(),
                         [](const BatchTile& a, const BatchTile& b) { return a.cost > b.cost; });
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

void batch_enqueue(const BatchTile* d_tiles0, uint32_t n0, const BatchTile* d_tiles1, uint32_t n1,
                   const OpDesc* d_ops, uint32_t* d_counters, uint32_t grid, size_t lds_bytes,
                   hipStream_t s) {
    hip_check(launch_batch_reset(d_counters, 2, s), "batch reset");
    if (n0) hip_check(launch_batch(d_tiles0, n0, d_counters, d_ops, std::min(grid, n0), lds_bytes, s), "batch");
    if (n1)
        hip_check(launch_batch(d_tiles1, n1, d_counters + 1, d_ops, std::min(grid, n1), lds_bytes, s), "batch");
}

}  // namespace mifft

using namespace mifft;

struct mifft_batch {
    BatchPlanCache cache;
    BatchBuild build;
    OpDesc* d_ops = nullptr;
    BatchTile* d_tiles[2] = {nullptr, nullptr};
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

static void create_batch(mifft_batch& B, const mifft_batch_item* items, int32_t count, const mifft_tuning* t) {
    if (t) B.cache.tuning = *t;
    uint64_t ws = 0;
    for (int32_t i = 0; i < count; i++) ws += batch_item_workspace(B.cache, items[i]);
    if (ws) hip_check(hipMalloc(&B.ws, ws), "batch workspace");
    batch_build(B.cache, items, count, (char*)B.ws, B.build);
    B.grid = batch_grid(B.build.lds_bytes);
    auto upload = [](void** dst, const void* src, size_t bytes, const char* what) {
        hip_check(hipMalloc(dst, std::max<size_t>(bytes, 16)), what);
        if (bytes) hip_check(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice), what);
    };
    upload((void**)&B.d_ops, B.build.ops.data(), B.build.ops.size() * sizeof(OpDesc), "batch ops");
    for (int p = 0; p < 2; p++)
        upload((void**)&B.d_tiles[p], B.build.tiles[p].data(), B.build.tiles[p].size() * sizeof(BatchTile),
               "batch tiles");
    hip_check(hipMalloc(&B.d_counters, 64), "batch counters");
}

static void run_batch(const mifft_batch& B, hipStream_t s) {
    batch_enqueue(B.d_tiles[0], (uint32_t)B.build.tiles[0].size(), B.d_tiles[1], (uint32_t)B.build.tiles[1].size(),
                  B.d_ops, B.d_counters, B.grid, B.build.lds_bytes, s);
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
    if (tiles) *tiles = b->build.tiles[0].size() + b->build.tiles[1].size();
    if (phases) *phases = b->build.tiles[1].empty() ? 1 : 2;
    if (lds_bytes) *lds_bytes = b->build.lds_bytes;
    if (workspace_bytes) *workspace_bytes = b->build.ws_bytes;
    if (grid) *grid = b->grid;
    return MIFFT_SUCCESS;
}

void mifft_batch_destroy(mifft_batch* b) { delete b; }

}  // extern "C"
