// Host planning of persistent batched launches, shared by the batch API
// (batch.cpp) and the serving loop (serve.cpp).
#pragma once
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "batch_kernels.hpp"
#include "plan.hpp"

namespace mifft {

constexpr size_t kBatchLdsBudget = 80 * 1024;

// Batched-mode plans (256-thread kernels, up to 80 KiB LDS per tile: a launch
// runs two workgroups per CU at that size, four when its tiles need at most
// 40 KiB; measured on the serving mix, 80 KiB batches more requests and
// serves faster than 40, 48, 64, 96 or 128 KiB).  `get_1d` builds 1D f32
// plans for one line (shared by every item of that length, whose line count
// scales the ops); `get` builds a plan for a full configuration.  Both return
// nullptr when the batched kernel cannot run the plan (too many passes, or an
// op without a batched class).
struct BatchPlanCache {
    std::map<std::tuple<uint64_t, int, int, int>, std::unique_ptr<Plan>> plans_1d;
    std::map<std::string, std::unique_ptr<Plan>> plans;
    std::map<std::string, std::string> rejected;  // config key -> reason
    mifft_tuning tuning;
    BatchPlanCache();
    const Plan* get_1d(uint64_t n, int direction, int normalization, int algorithm, std::string* why = nullptr);
    const Plan* get(const mifft_config& c, mifft_kind kind, std::string* why = nullptr);
    // Drops the cached plan (or rejection) of a configuration.
    void erase(const mifft_config& c, mifft_kind kind);
};

// One transform of a batch: a batched-mode plan, its line multiplier (the
// plan's ops run over mult times their lines, workspace scaled alike) and its
// bound buffers.
struct BatchEntry {
    const Plan* plan;
    uint64_t mult;
    const void* in;
    void* out;
};

struct BatchBuild {
    std::vector<OpDesc> ops;
    std::vector<BatchTile> tiles[kMaxPhases];  // phase k holds op k of every entry
    size_t lds_bytes = 16;
    uint64_t ws_bytes = 0;
    int phases() const {
        int p = 0;
        for (int k = 0; k < kMaxPhases; k++)
            if (!tiles[k].empty()) p = k + 1;
        return p;
    }
    size_t tile_count() const {
        size_t t = 0;
        for (auto& v : tiles) t += v.size();
        return t;
    }
};

// Appends the entries' ops and tiles to `out`, binding workspace inside
// [ws_base, ws_base + growth of out.ws_bytes).  Tiles are sorted by cost.
void batch_build(const BatchEntry* e, int32_t count, char* ws_base, BatchBuild& out);
// Device descriptor bytes (ops and tiles) one entry adds to a batch.
uint64_t batch_desc_bytes(const Plan& p, uint64_t mult);
// Persistent workgroups for a given LDS footprint: CUs x workgroups per CU.
uint32_t batch_grid(size_t lds_bytes);
// Enqueues the counter reset and one persistent launch per non-empty phase;
// `d_tiles[k]` holds phase k's tiles on the device.
void batch_enqueue(const BatchTile* const* d_tiles, const uint32_t* counts, const OpDesc* d_ops,
                   uint32_t* d_counters, uint32_t grid, size_t lds_bytes, hipStream_t s);

}  // namespace mifft
