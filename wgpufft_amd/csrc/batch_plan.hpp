// Host planning of persistent batched launches, shared by the batch API
// (batch.cpp) and the serving loop (serve.cpp).
#pragma once
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "batch_kernels.hpp"
#include "plan.hpp"

namespace mifft {

// Batched-mode plans (256-thread kernels, 40 KiB LDS), one per
// (n, direction, normalization, algorithm), built for one line and shared by
// every item of that shape.
struct BatchPlanCache {
    std::map<std::tuple<uint64_t, int, int, int>, std::unique_ptr<Plan>> plans;
    mifft_tuning tuning;
    BatchPlanCache();
    const Plan* get(uint64_t n, int direction, int normalization, int algorithm);
};

struct BatchBuild {
    std::vector<OpDesc> ops;
    std::vector<BatchTile> tiles[2];  // phase 0, phase 1 (second pass of four-step splits)
    size_t lds_bytes = 16;
    uint64_t ws_bytes = 0;
};

constexpr size_t kBatchLdsBudget = 40 * 1024;

// Workspace bytes of one item (0 unless its length needs a four-step split);
// throws Error for items the batched kernel cannot run.
uint64_t batch_item_workspace(BatchPlanCache& cache, const mifft_batch_item& it);
// Whether the batched kernel can run the item (no throw).
bool batch_item_supported(BatchPlanCache& cache, const mifft_batch_item& it);
// Appends the items' ops and tiles to `out`, binding four-step scratch inside
// [ws_base, ws_base + out.ws_bytes growth).  Tiles are sorted by cost.
void batch_build(BatchPlanCache& cache, const mifft_batch_item* items, int32_t count, char* ws_base,
                 BatchBuild& out);
// Persistent workgroups for a given LDS footprint: CUs x workgroups per CU.
uint32_t batch_grid(size_t lds_bytes);
// Enqueues counter reset + phase 0 + phase 1 launches.
void batch_enqueue(const BatchTile* d_tiles0, uint32_t n0, const BatchTile* d_tiles1, uint32_t n1,
                   const OpDesc* d_ops, uint32_t* d_counters, uint32_t grid, size_t lds_bytes,
                   hipStream_t s);

}  // namespace mifft
