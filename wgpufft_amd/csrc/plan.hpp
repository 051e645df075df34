// Internal plan representation of the host runtime.
#pragma once
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mifft.h"
#include "launch.h"

namespace mifft {

// Thrown inside the library, converted to mifft_status at the C boundary.
struct Error {
    mifft_status code;
    std::string detail;
};
[[noreturn]] void fail(mifft_status code, const std::string& detail);
void hip_check(hipError_t e, const char* what);
// Sets the thread-local text returned by mifft_last_error_detail().
void set_last_detail(const std::string& detail);

enum BufRef { B_IN = 0, B_OUT = 1, B_WS = 2 };

enum OpKind {
    OP_REG,         // compiled register-resident line kernel
    OP_LDS,         // runtime-schedule LDS line kernel
    OP_RADER,       // fused Rader convolution
    OP_BLUESTEIN,   // fused Bluestein convolution
    OP_CHIRP_PRE,   // multi-pass Bluestein: chirp multiply + zero pad into workspace
    OP_POINTWISE,   // multi-pass Bluestein: filter multiply (+ swap)
    OP_CHIRP_POST,  // multi-pass Bluestein: chirp multiply + crop
    OP_SUBPLAN,     // a child plan run in place on a workspace region
    OP_SCALE_COPY,  // direct route of a one-point transform
    OP_REAL_WIDEN,  // general R2C: real input -> complex lines (im = 0)
    OP_PACK,        // general R2C: complex lines -> packed bins k <= N0/2
    OP_HERM_EXTEND, // general C2R: packed lines -> Hermitian-extended complex lines
    OP_REAL_PART,   // general C2R: complex lines -> real part
    OP_R2C_POST,    // half-length R2C: h-point complex bins -> packed spectrum
    OP_C2R_PRE,     // half-length C2R: packed spectrum -> h-point complex lines
    OP_SMALL_VOLUME, // every axis of small f32 volumes, one workgroup per volume
    OP_DIRECT,       // direct DFT of short prime lines (n = p, m = lines per workgroup, filt = W_p table)
    OP_STOCKHAM,     // one radix stage of a long axis over HBM (st)
    OP_CHUNK         // large-chunk route: a child plan over batches at byte offsets src_off / dst_off
};

struct Plan;

struct Op {
    OpKind kind;
    int src = B_IN, dst = B_OUT;
    uint64_t src_off = 0, dst_off = 0;  // byte offsets into the referenced buffers
    int prec = P_F32;                   // kernel format
    LineArgs a{};
    Sched sc{};
    ConvArgs conv{};
    SmallVolArgs sv{};  // OP_SMALL_VOLUME (volumes in count)
    StockhamStage st{}; // OP_STOCKHAM
    const RegVariant* rv = nullptr;
    const RegConvVariant* rcv = nullptr;  // OP_BLUESTEIN: register-resident form
    int ldm = LD_COMPLEX, stm = ST_COMPLEX, post = 0;
    Plan* sub = nullptr;     // OP_SUBPLAN
    uint64_t sub_ws_off = 0; // child workspace region
    uint32_t n = 0, m = 0;
    const void* chirp = nullptr;
    const void* filt = nullptr;
    uint64_t count = 0;
    double scale = 1.0;
    std::string desc;
    // line ops: lane layout, whether line pairs are allowed, and the
    // tuning-table key of the op's shape
    bool ld_lf = false, st_lf = false, pairable = false;
    bool fold = false;  // the fold pass of a 2D transform (fold register variants only)
    // unit point strides and every line at a 16-byte-aligned offset on both
    // sides: adjacent-butterfly variants (16-byte accesses) may run the op
    bool adj_ok = false;
    std::string tkey;
};

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// A plan-owned device table and what it holds ("stage-twiddles prec=f32
// radices=16,16,16 count=4080", "chirp prec=f32 n=1009", ...).
struct TableInfo {
    const void* p = nullptr;
    uint64_t bytes = 0;
    std::string desc;
};

struct AxisInfo {
    int axis;
    uint64_t n;
    int kind;  // mifft_axis_kind
    std::vector<uint64_t> factors;
};

struct Plan {
    mifft_config cfg{};
    mifft_kind kind = MIFFT_C2C;
    int prec = P_F32;   // storage precision of the plan
    int arith = P_F32;  // P_F32 or P_F64 (df64 computes in f64)
    int route = MIFFT_ROUTE_MIXED_RADIX;
    int device = 0;
    std::vector<AxisInfo> axes;
    std::vector<Op> ops;
    std::vector<DevBuf> owned;
    std::vector<TableInfo> tables;  // described device tables among `owned`
    std::vector<std::unique_ptr<Plan>> children;
    uint64_t in_bytes = 0, out_bytes = 0, ws_bytes = 0;
    bool in_place_kernels = true;
    size_t lds_budget = kLdsBudget;
    bool batch_mode = false;
    bool large_chunk = false;  // ops are OP_CHUNK over child plans (more than u32::MAX elements)
    // Plan-owned workspace used when the caller passes none (one in-flight
    // execution per plan, as the reference's plan-owned temporaries).
    std::mutex ws_mu;
    DevBuf own_ws;
    DevBuf in_copy;  // lazily created input copy for in-place real transforms
    DevBuf stage;    // lazily created staging of strided logical views
    ~Plan();
};

// Planner options: the LDS budget per workgroup, and the batched mode, which
// restricts kernels to those the persistent batched kernel can run
// (256 threads, see batch_kernels.hpp).
struct PlanOptions {
    size_t lds_budget = kLdsBudget;
    bool batch_mode = false;
};
std::unique_ptr<Plan> build_plan(const mifft_config& cfg, mifft_kind kind,
                                 const PlanOptions& opt = PlanOptions());
// Enqueues the plan on `s`.  `ws` must hold ws_bytes when ws_bytes > 0.
void run_plan(const Plan& p, hipStream_t s, const void* in, void* out, void* ws);
// Runs only launch `index` of the flattened launch sequence (child plans
// expanded); returns false if index is out of range.
bool run_plan_launch(const Plan& p, hipStream_t s, bool run, int index, const void* in, void* out, void* ws,
                     std::string* desc);
int plan_launch_count(const Plan& p);
// Size of the static device scratch of mifft_execute_static_scratch
// (scratch.hip) and its device address on `device`, resolved once per device
// with that device current (fails on any other device).
constexpr uint64_t kStaticScratchBytes = 80ull << 20;
void* static_scratch(int device);
// Described device tables of the plan and its child plans, in creation order.
void plan_tables(const Plan& p, std::vector<const TableInfo*>* out);
// The op behind launch `index` of the flattened sequence, or nullptr.
const Op* plan_launch_op(const Plan& p, int index);

// Process-wide tuning table: measured line-kernel choices keyed by op shape
// (precision, length, lane layout, load/store modes, twiddle, pairability,
// line-count bucket).  A choice is a registered variant name ("reg:...") or
// a runtime-schedule LDS shape ("lds:lpw=<lines per workgroup>").
std::string tuning_lookup(const std::string& key);
void tuning_set(const std::string& key, const std::string& choice);
void tuning_clear();
std::vector<std::pair<std::string, std::string>> tuning_entries();
// Parses "key = choice" lines ('#' comments); throws on a malformed line or
// an unknown variant.  Returns the number of entries read.
int tuning_parse(const std::string& text);
// Tunes the kernel choice of every line-op shape of the plans of
// (cfgs[i], kinds[i]), executed back to back as a caller runs them (in place
// on one array when all plans allow it, else each reading the previous
// output; arrays rotated beyond the MALL, one workspace per plan): for each
// shape in launch order, the whole sequence is timed with each candidate
// applied to all ops of that shape, and the fastest is kept and recorded.
// Returns the number of shapes tuned; appends one line per shape to `report`.
int autotune_sequence(const mifft_config* cfgs, const mifft_kind* kinds, int count, int reps, hipStream_t s,
                      std::string* report);
// autotune_sequence of one plan.
int autotune(const mifft_config& cfg, mifft_kind kind, int reps, hipStream_t s, std::string* report);
std::string describe_plan(const Plan& p);
// FftPlan::diagnostics as JSON, with the given blockers (JSON objects).
std::string plan_diagnostics_json(const Plan& p, const std::vector<std::string>& blockers = {});

// Host-side integer helpers (the reference's route-choice arithmetic).
std::vector<uint64_t> factor_supported(uint64_t n, bool* ok);  // runtime/mod.rs:33-54
bool is_prime_u64(uint64_t n);
uint64_t mod_pow_u64(uint64_t b, uint64_t e, uint64_t m);
uint64_t primitive_root(uint64_t p);
uint64_t next_smooth(uint64_t n);
uint64_t rader_conv_len(uint64_t p);
bool medium_rader_schedule(uint64_t p, std::vector<int>* out);
uint64_t bluestein_conv_len(uint64_t n);
void validate_config(const mifft_config& c);
double config_scale(const mifft_config& c);
uint64_t config_total(const mifft_config& c);

}  // namespace mifft
