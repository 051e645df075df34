// Host planner of the persistent batched launch (mifft_batch_*).
//
// Each item (a 1D transform of `n` points over `batch` lines) reuses the
// planner of standalone plans in batched mode (256-thread kernels, a 40 KiB
// LDS budget so that four workgroups share a CU); plans are cached by
// (n, direction, normalization, algorithm), so the twiddle, chirp and filter
// tables of equal lengths are shared.  Every op of every item becomes a run of
// tiles; tiles are sorted by cost (longest first) inside their phase and
// pulled by the persistent kernel from a device counter.
#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>

#include "batch_kernels.hpp"
#include "plan.hpp"

using namespace mifft;

struct mifft_batch {
    std::vector<std::unique_ptr<Plan>> plans;
    std::vector<OpDesc> ops;
    std::vector<BatchTile> tiles[2];
    OpDesc* d_ops = nullptr;
    BatchTile* d_tiles[2] = {nullptr, nullptr};
    uint32_t* d_counters = nullptr;
    void* ws = nullptr;
    uint64_t ws_bytes = 0;
    size_t lds_bytes = 0;
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


static constexpr size_t kBatchLds = 40 * 1024;

static void build_batch(mifft_batch& B, const mifft_batch_item* items, int32_t count,
                        const mifft_tuning* tuning) {
    std::map<std::tuple<uint64_t, int, int, int>, Plan*> cache;
    uint64_t ws_total = 0;
    // per-item workspace offsets (four-step scratch)
    std::vector<uint64_t> ws_off(count, 0);
    struct Pending {
        int item;
        const Plan* plan;
    };
    std::vector<Pending> pend;
    for (int32_t i = 0; i < count; i++) {
        const mifft_batch_item& it = items[i];
        if (it.n == 0) fail(MIFFT_ERR_ZERO_LENGTH, "batch item " + std::to_string(i) + ": zero length");
        if (it.batch == 0) fail(MIFFT_ERR_ZERO_BATCH, "batch item " + std::to_string(i) + ": zero batch");
        if (!it.in || !it.out) fail(MIFFT_ERR_NULL_ARGUMENT, "batch item " + std::to_string(i) + ": null buffer");
        if (((uintptr_t)it.in | (uintptr_t)it.out) & 15u)
            fail(MIFFT_ERR_UNALIGNED, "batch item " + std::to_string(i) + ": buffers must be 16-byte aligned");
        if (it.algorithm < 0 || it.algorithm > 2) fail(MIFFT_ERR_INVALID_ENUM, "algorithm out of range");
        auto key = std::make_tuple(it.n, (int)it.direction, (int)it.normalization, (int)it.algorithm);
        Plan* P = nullptr;
        auto f = cache.find(key);
        if (f != cache.end()) {
            P = f->second;
        } else {
            mifft_config c;
            std::memset(&c, 0, sizeof(c));
            c.rank = 1;
            c.shape[0] = it.n;
            c.naxes = 1;
            c.axes[0] = 0;
            c.batch = 1;
            c.direction = it.direction;
            c.normalization = it.normalization;
            c.precision = MIFFT_F32;
            if (tuning) c.tuning = *tuning;
            else mifft_tuning_init(&c.tuning);
            c.tuning.n_force_rader_axes = c.tuning.n_force_bluestein_axes = 0;
            if (it.algorithm == 1) {
                c.tuning.force_rader_axes[0] = 0;
                c.tuning.n_force_rader_axes = 1;
            } else if (it.algorithm == 2) {
                c.tuning.force_bluestein_axes[0] = 0;
                c.tuning.n_force_bluestein_axes = 1;
            }
            PlanOptions opt;
            opt.lds_budget = kBatchLds;
            opt.batch_mode = true;
            B.plans.push_back(build_plan(c, MIFFT_C2C, opt));
            P = B.plans.back().get();
            cache[key] = P;
        }
        if (P->ops.size() > 2)
            fail(MIFFT_ERR_UNSUPPORTED_LENGTH, "length " + std::to_string(it.n) + " needs more than two passes");
        const unsigned __int128 ws = (unsigned __int128)P->ws_bytes * it.batch;
        ws_off[i] = ws_total;
        ws_total += (uint64_t)ws;
        pend.push_back({i, P});
    }
    B.ws_bytes = ws_total;
    if (ws_total) hip_check(hipMalloc(&B.ws, ws_total), "batch workspace");
    size_t lds = 0;
    for (const Pending& pd : pend) {
        const mifft_batch_item& it = items[pd.item];
        const Plan& P = *pd.plan;
        for (size_t j = 0; j < P.ops.size(); j++) {
            const Op& op = P.ops[j];
            auto bind = [&](int ref, uint64_t off) -> char* {
                char* base = ref == B_IN ? (char*)it.in : ref == B_OUT ? (char*)it.out
                                                                       : (char*)B.ws + ws_off[pd.item];
                return base + off * (ref == B_WS ? it.batch : 1);
            };
            OpDesc d;
            std::memset(&d, 0, sizeof(d));
            uint32_t cls = 0, lpw = 1, lines = 0;
            size_t op_lds = 0;
            switch (op.kind) {
                case OP_REG: {
                    const int c = batch_class_for(op.rv);
                    if (c < 0) fail(MIFFT_ERR_UNSUPPORTED_LENGTH, "no batched class for " + op.desc);
                    cls = (uint32_t)c;
                    d.a = op.a;
                    lpw = (uint32_t)op.rv->lines;
                    op_lds = op.rv->lds_bytes;
                    break;
                }
                case OP_LDS:
                    if (op.ldm != LD_COMPLEX || op.stm != ST_COMPLEX)
                        fail(MIFFT_ERR_UNSUPPORTED_LENGTH, "batched real modes");
                    cls = op.post ? BC_LDS_POST : BC_LDS;
                    d.a = op.a;
                    d.sc = op.sc;
                    lpw = op.sc.lpw;
                    op_lds = lds_kernel_bytes(P_F32, op.sc.n, op.sc.lpw);
                    break;
                case OP_RADER:
                case OP_BLUESTEIN:
                    cls = op.kind == OP_RADER ? BC_RADER : BC_BLUESTEIN;
                    d.conv = op.conv;
                    lpw = op.conv.sc.lpw;
                    op_lds = conv_kernel_bytes(P_F32, op.conv.m, op.conv.sc.lpw);
                    break;
                default:
                    fail(MIFFT_ERR_UNSUPPORTED_LENGTH,
                         "length " + std::to_string(it.n) + " needs a route the batched kernel does not run (" +
                             op.desc + ")");
            }
            LineArgs& la = (op.kind == OP_RADER || op.kind == OP_BLUESTEIN) ? d.conv.io : d.a;
            const unsigned __int128 L = (unsigned __int128)la.lines * it.batch;
            if (L > 0x7FFFFFFFu) fail(MIFFT_ERR_LENGTH_TOO_LARGE, "batch item lines exceed 2^31");
            la.lines = (uint32_t)L;
            lines = la.lines;
            la.in = bind(op.src, op.src_off);
            la.out = bind(op.dst, op.dst_off);
            lds = std::max(lds, op_lds);
            const uint32_t arg = (uint32_t)B.ops.size();
            B.ops.push_back(d);
            const uint32_t ntiles = (lines + lpw - 1) / lpw;
            const uint64_t cost64 = (uint64_t)lpw * it.n * (op.kind == OP_BLUESTEIN || op.kind == OP_RADER ? 3 : 1);
            const uint32_t cost = (uint32_t)std::min<uint64_t>(cost64, 0xFFFFFFFFu);
            for (uint32_t t = 0; t < ntiles; t++) B.tiles[j].push_back(BatchTile{cls, arg, t * lpw, cost});
        }
    }
    for (auto& v : B.tiles)
        std::stable_sort(v.begin(), v.end(), [](const BatchTile& a, const BatchTile& b) { return a.cost > b.cost; });
    B.lds_bytes = std::max<size_t>(lds, 16);
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, dev), "device properties");
    const uint32_t per_cu = (uint32_t)std::max<size_t>(1, std::min<size_t>(8, kLdsBudget / B.lds_bytes));
    B.grid = (uint32_t)prop.multiProcessorCount * per_cu;
    hip_check(hipMalloc(&B.d_ops, std::max<size_t>(1, B.ops.size()) * sizeof(OpDesc)), "batch ops");
    hip_check(hipMemcpy(B.d_ops, B.ops.data(), B.ops.size() * sizeof(OpDesc), hipMemcpyHostToDevice), "batch ops");
    for (int p = 0; p < 2; p++) {
        hip_check(hipMalloc(&B.d_tiles[p], std::max<size_t>(1, B.tiles[p].size()) * sizeof(BatchTile)), "batch tiles");
        if (!B.tiles[p].empty())
            hip_check(hipMemcpy(B.d_tiles[p], B.tiles[p].data(), B.tiles[p].size() * sizeof(BatchTile),
                                hipMemcpyHostToDevice),
                      "batch tiles");
    }
    hip_check(hipMalloc(&B.d_counters, 64), "batch counters");
}

static void run_batch(const mifft_batch& B, hipStream_t s) {
    hip_check(launch_batch_reset(B.d_counters, 2, s), "batch reset");
    for (int p = 0; p < 2; p++) {
        if (B.tiles[p].empty()) continue;
        const uint32_t nt = (uint32_t)B.tiles[p].size();
        const uint32_t grid = std::min<uint32_t>(B.grid, nt);
        hip_check(launch_batch(B.d_tiles[p], nt, B.d_counters + p, B.d_ops, grid, B.lds_bytes, s), "batch launch");
    }
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
        build_batch(*b, items, count, tuning);
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
    if (tiles) *tiles = b->tiles[0].size() + b->tiles[1].size();
    if (phases) *phases = b->tiles[1].empty() ? 1 : 2;
    if (lds_bytes) *lds_bytes = b->lds_bytes;
    if (workspace_bytes) *workspace_bytes = b->ws_bytes;
    if (grid) *grid = b->grid;
    return MIFFT_SUCCESS;
}

void mifft_batch_destroy(mifft_batch* b) { delete b; }

}  // extern "C"
