// C ABI (include/mifft.h) over the host runtime.  Every entry point converts
// internal errors into an mifft_status and a thread-local detail string; no
// exception crosses the boundary.
#include <cstring>
#include <string>

#include <cstdio>
#include <algorithm>
#include "diagnostics.hpp"
#include "plan.hpp"

using namespace mifft;

struct mifft_plan {
    std::unique_ptr<Plan> p;
};

struct mifft_recorder {
    hipStream_t stream = nullptr;
    bool capture = false;
    bool ended = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static thread_local std::string g_detail;
// Diagnostics JSON of the last failing plan creation or execution on this
// thread (FftPlanCreationError / FftExecutionError diagnostics).
static thread_local std::string g_diag;

void mifft::set_last_detail(const std::string& d) {
    g_detail = d;
    g_diag.clear();
}

static mifft_status set_err(mifft_status s, const std::string& d) {
    g_detail = d;
    g_diag.clear();
    return s;
}

static mifft_status write_json(const std::string& d, char* buf, size_t cap, size_t* needed) {
    if (needed) *needed = d.size() + 1;
    if (buf && cap) {
        const size_t n = std::min(cap - 1, d.size());
        std::memcpy(buf, d.data(), n);
        buf[n] = '\0';
    }
    return MIFFT_SUCCESS;
}

template <class F>
static mifft_status guard(F&& f) {
    try {
        f();
        return MIFFT_SUCCESS;
    } catch (const Error& e) {
        return set_err(e.code, e.detail);
    } catch (const std::bad_alloc&) {
        return set_err(MIFFT_ERR_ALLOCATION_FAILED, "host allocation failed");
    } catch (...) {
        return set_err(MIFFT_ERR_HIP, "unexpected internal error");
    }
}

extern "C" {

void mifft_tuning_init(mifft_tuning* t) {
    if (!t) return;
    std::memset(t, 0, sizeof(*t));
    t->workgroup_size = 64;
    t->fused_workgroup_size = 256;
    t->rader_max_prime = 8192;
    t->direct_max_prime = 127;
    t->fuse_long_axes = 1;
    t->fuse_small_volumes = 1;
    t->fused_min_convolution_length = 128;
}

mifft_status mifft_config_init(mifft_config* c, int32_t rank, const uint64_t* shape) {
    if (!c || (rank > 0 && !shape)) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null config or shape");
    if (rank < 0 || rank > MIFFT_MAX_RANK) return set_err(MIFFT_ERR_RANK_TOO_LARGE, "rank out of range");
    std::memset(c, 0, sizeof(*c));
    c->rank = rank;
    for (int i = 0; i < rank; i++) {
        c->shape[i] = shape[i];
        c->axes[i] = i;
    }
    c->naxes = rank;
    c->batch = 1;
    c->direction = MIFFT_FORWARD;
    c->normalization = MIFFT_NORM_INVERSE;
    c->precision = MIFFT_F32;
    mifft_tuning_init(&c->tuning);
    return MIFFT_SUCCESS;
}

mifft_status mifft_config_validate(const mifft_config* c) {
    if (!c) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null config");
    return guard([&] { validate_config(*c); });
}

mifft_status mifft_config_scale(const mifft_config* c, double* scale) {
    if (!c || !scale) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null argument");
    return guard([&] {
        validate_config(*c);
        *scale = config_scale(*c);
    });
}

mifft_status mifft_config_total_len(const mifft_config* c, uint64_t* len) {
    if (!c || !len) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null argument");
    return guard([&] {
        validate_config(*c);
        *len = config_total(*c);
    });
}

mifft_status mifft_plan_create(const mifft_config* c, mifft_kind kind, mifft_plan** out) {
    if (!c || !out) {
        set_err(MIFFT_ERR_NULL_ARGUMENT, "null config or output handle");
        g_diag = creation_failure_json(c, kind, MIFFT_ERR_NULL_ARGUMENT, g_detail);
        return MIFFT_ERR_NULL_ARGUMENT;
    }
    *out = nullptr;
    const mifft_status s = guard([&] {
        auto h = std::make_unique<mifft_plan>();
        h->p = build_plan(*c, kind);
        if (h->p->ws_bytes) {
            hipError_t e = hipMalloc(&h->p->own_ws.p, h->p->ws_bytes);
            if (e != hipSuccess) fail(MIFFT_ERR_ALLOCATION_FAILED, "plan workspace");
            h->p->own_ws.bytes = h->p->ws_bytes;
        }
        *out = h.release();
    });
    if (s != MIFFT_SUCCESS) g_diag = creation_failure_json(c, kind, s, g_detail);
    return s;
}

mifft_status mifft_plan_create_with_diagnostics(const mifft_config* c, mifft_kind kind, mifft_plan** out, char* buf,
                                                size_t cap, size_t* needed) {
    const mifft_status s = mifft_plan_create(c, kind, out);
    if (s == MIFFT_SUCCESS) {
        std::string d;
        const mifft_status g = guard([&] { d = plan_diagnostics_json(*(*out)->p); });
        if (g != MIFFT_SUCCESS) return g;
        return write_json(d, buf, cap, needed);
    }
    const std::string detail = g_detail;
    write_json(g_diag, buf, cap, needed);
    g_detail = detail;
    return s;
}

mifft_status mifft_last_error_diagnostics(char* buf, size_t cap, size_t* needed) {
    return write_json(g_diag.empty() ? std::string("null") : g_diag, buf, cap, needed);
}

void mifft_plan_destroy(mifft_plan* p) { delete p; }

mifft_kind mifft_plan_kind(const mifft_plan* p) { return p ? p->p->kind : MIFFT_C2C; }

mifft_status mifft_plan_config(const mifft_plan* p, mifft_config* out) {
    if (!p || !out) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null argument");
    *out = p->p->cfg;
    return MIFFT_SUCCESS;
}

int32_t mifft_plan_axis_factors(const mifft_plan* p, int32_t axis_index, uint64_t* out, int32_t cap) {
    if (!p || axis_index < 0 || axis_index >= (int32_t)p->p->axes.size()) return -1;
    const auto& f = p->p->axes[axis_index].factors;
    for (int32_t i = 0; i < (int32_t)f.size() && i < cap; i++) out[i] = f[i];
    return (int32_t)f.size();
}

int32_t mifft_plan_factors(const mifft_plan* p, uint64_t* out, int32_t cap) {
    return mifft_plan_axis_factors(p, 0, out, cap);
}

int32_t mifft_plan_axis_kinds(const mifft_plan* p, int32_t* out, int32_t cap) {
    if (!p) return -1;
    const auto& a = p->p->axes;
    for (int32_t i = 0; i < (int32_t)a.size() && i < cap; i++) out[i] = a[i].kind;
    return (int32_t)a.size();
}

mifft_route mifft_plan_route(const mifft_plan* p) {
    return p ? (mifft_route)p->p->route : MIFFT_ROUTE_DIRECT_DFT;
}
uint64_t mifft_plan_workspace_bytes(const mifft_plan* p) { return p ? p->p->ws_bytes : 0; }
uint64_t mifft_plan_required_input_bytes(const mifft_plan* p) { return p ? p->p->in_bytes : 0; }
uint64_t mifft_plan_required_output_bytes(const mifft_plan* p) { return p ? p->p->out_bytes : 0; }
uint64_t mifft_plan_required_bytes(const mifft_plan* p) {
    return p ? std::max(p->p->in_bytes, p->p->out_bytes) : 0;
}

int32_t mifft_plan_packed_shape(const mifft_plan* p, uint64_t* out, int32_t cap) {
    if (!p || p->p->kind == MIFFT_C2C) return 0;
    const mifft_config& c = p->p->cfg;
    for (int32_t i = 0; i < c.rank && i < cap; i++) out[i] = i == 0 ? c.shape[0] / 2 + 1 : c.shape[i];
    return c.rank;
}

int32_t mifft_plan_supports_in_place(const mifft_plan* p) { return p && p->p->in_place_kernels ? 1 : 0; }

int32_t mifft_plan_launch_count(const mifft_plan* p) { return p ? plan_launch_count(*p->p) : 0; }

mifft_status mifft_plan_launch_describe(const mifft_plan* p, int32_t index, char* buf, size_t cap) {
    if (!p || !buf || cap == 0) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null argument");
    std::string d;
    if (!run_plan_launch(*p->p, nullptr, false, index, nullptr, nullptr, nullptr, &d))
        return set_err(MIFFT_ERR_INVALID_ENUM, "launch index out of range");
    std::strncpy(buf, d.c_str(), cap - 1);
    buf[cap - 1] = '\0';
    return MIFFT_SUCCESS;
}

static void copy_out(const std::string& v, char* buf, size_t cap) {
    if (!buf || cap == 0) return;
    const size_t n = std::min(cap - 1, v.size());
    std::memcpy(buf, v.data(), n);
    buf[n] = '\0';
}

mifft_status mifft_plan_launch_tuning(const mifft_plan* p, int32_t index, char* key, size_t key_cap, char* variant,
                                      size_t variant_cap) {
    if (!p) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null plan");
    const Op* op = plan_launch_op(*p->p, index);
    if (!op) return set_err(MIFFT_ERR_INVALID_ENUM, "launch index out of range");
    copy_out(op->tkey, key, key_cap);
    copy_out(op->kind == OP_REG && op->rv ? reg_variant_name(op->rv) : std::string(), variant, variant_cap);
    return MIFFT_SUCCESS;
}

int32_t mifft_plan_table_count(const mifft_plan* p) {
    if (!p) return 0;
    std::vector<const TableInfo*> t;
    plan_tables(*p->p, &t);
    return (int32_t)t.size();
}

mifft_status mifft_plan_table(const mifft_plan* p, int32_t index, char* desc, size_t desc_cap, void* host,
                              uint64_t host_cap, uint64_t* bytes) {
    if (!p) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null plan");
    return guard([&] {
        std::vector<const TableInfo*> t;
        plan_tables(*p->p, &t);
        if (index < 0 || index >= (int32_t)t.size()) fail(MIFFT_ERR_INVALID_ENUM, "table index out of range");
        const TableInfo& ti = *t[index];
        copy_out(ti.desc, desc, desc_cap);
        if (bytes) *bytes = ti.bytes;
        if (host) {
            if (host_cap < ti.bytes)
                fail(MIFFT_ERR_HOST_BUFFER_LENGTH, "table needs " + std::to_string(ti.bytes) + " bytes");
            hip_check(hipMemcpy(host, ti.p, ti.bytes, hipMemcpyDeviceToHost), "table copy");
        }
    });
}

mifft_status mifft_twiddle_table(void* dst, uint64_t n, uint64_t count, int32_t precision, hipStream_t stream) {
    if (!dst) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null buffer");
    if (n == 0) return set_err(MIFFT_ERR_ZERO_LENGTH, "n must be at least 1");
    if (precision != MIFFT_F32 && precision != MIFFT_F64) return set_err(MIFFT_ERR_INVALID_ENUM, "f32 or f64 table");
    if ((uintptr_t)dst & 15u) return set_err(MIFFT_ERR_UNALIGNED, "device buffers must be 16-byte aligned");
    return guard([&] {
        if (count) hip_check(gen_twiddles(precision == MIFFT_F32 ? P_F32 : P_F64, dst, n, count, stream), "twiddles");
    });
}

int32_t mifft_registry_count(void) { return reg_variant_count(); }

mifft_status mifft_registry_name(int32_t index, char* buf, size_t cap) {
    if (!buf || cap == 0) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null argument");
    const RegVariant* v = reg_variant_at(index);
    if (!v) return set_err(MIFFT_ERR_INVALID_ENUM, "registry index out of range");
    copy_out(reg_variant_name(v), buf, cap);
    return MIFFT_SUCCESS;
}

static void check_sizes(const Plan& P, const void* in, uint64_t in_bytes, const void* out, uint64_t out_bytes);

mifft_status mifft_execute_launch(const mifft_plan* p, hipStream_t stream, int32_t index, const void* in,
                                  uint64_t in_bytes, void* out, uint64_t out_bytes, void* workspace,
                                  uint64_t workspace_bytes) {
    if (!p || !in || !out) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null argument");
    return guard([&] {
        const Plan& P = *p->p;
        check_sizes(P, in, in_bytes, out, out_bytes);
        if (in == out && !P.in_place_kernels)
            fail(MIFFT_ERR_IN_PLACE_UNSUPPORTED, "a single launch of this plan cannot run in place");
        if (workspace && ((uintptr_t)workspace & 15u))
            fail(MIFFT_ERR_UNALIGNED, "device buffers must be 16-byte aligned");
        if (P.ws_bytes && workspace && workspace_bytes < P.ws_bytes)
            fail(MIFFT_ERR_WORKSPACE_TOO_SMALL,
                 "workspace needs " + std::to_string(P.ws_bytes) + " bytes, got " + std::to_string(workspace_bytes));
        if (!run_plan_launch(P, stream, true, index, in, out, workspace ? workspace : P.own_ws.p, nullptr))
            fail(MIFFT_ERR_INVALID_ENUM, "launch index out of range");
    });
}

mifft_status mifft_plan_describe(const mifft_plan* p, char* buf, size_t cap) {
    if (!p || !buf || cap == 0) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null argument");
    std::string d = describe_plan(*p->p);
    std::strncpy(buf, d.c_str(), cap - 1);
    buf[cap - 1] = '\0';
    return MIFFT_SUCCESS;
}

mifft_status mifft_plan_diagnostics(const mifft_plan* p, char* buf, size_t cap, size_t* needed) {
    if (!p) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null plan");
    return guard([&] {
        const std::string d = plan_diagnostics_json(*p->p);
        if (needed) *needed = d.size() + 1;
        if (buf && cap) {
            const size_t n = std::min(cap - 1, d.size());
            std::memcpy(buf, d.data(), n);
            buf[n] = '\0';
        }
    });
}

static void check_sizes(const Plan& P, const void* in, uint64_t in_bytes, const void* out,
                        uint64_t out_bytes) {
    if (!in || !out) fail(MIFFT_ERR_NULL_ARGUMENT, "null buffer");
    if (((uintptr_t)in | (uintptr_t)out) & 15u)
        fail(MIFFT_ERR_UNALIGNED, "device buffers must be 16-byte aligned");
    if (in_bytes < P.in_bytes)
        fail(MIFFT_ERR_BUFFER_TOO_SMALL, "input needs " + std::to_string(P.in_bytes) + " bytes, got " +
                                             std::to_string(in_bytes));
    if (out_bytes < P.out_bytes)
        fail(MIFFT_ERR_BUFFER_TOO_SMALL, "output needs " + std::to_string(P.out_bytes) + " bytes, got " +
                                             std::to_string(out_bytes));
}

static void exec_impl(const mifft_plan* p, hipStream_t s, const void* in, uint64_t in_bytes, void* out,
                      uint64_t out_bytes, void* ws, uint64_t ws_bytes) {
    if (!p) fail(MIFFT_ERR_NULL_ARGUMENT, "null plan");
    Plan& P = *p->p;
    check_sizes(P, in, in_bytes, out, out_bytes);
    if (P.ws_bytes && ws && ws_bytes < P.ws_bytes)
        fail(MIFFT_ERR_WORKSPACE_TOO_SMALL,
             "workspace needs " + std::to_string(P.ws_bytes) + " bytes, got " + std::to_string(ws_bytes));
    if (in == out && !P.in_place_kernels) {
        // Real plans change the element size: transform from a plan-owned
        // copy of the input (FftPlan::record_in_place, plan.rs:839-880).
        std::lock_guard<std::mutex> g(P.ws_mu);
        if (!P.in_copy.p) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            (void)hipStreamIsCapturing(s, &cs);
            if (cs != hipStreamCaptureStatusNone)
                fail(MIFFT_ERR_RECORDER_STATE, "first in-place real execution cannot be graph-captured");
            hip_check(hipMalloc(&P.in_copy.p, P.in_bytes), "in-place input copy");
            P.in_copy.bytes = P.in_bytes;
        }
        hip_check(hipMemcpyAsync(P.in_copy.p, in, P.in_bytes, hipMemcpyDeviceToDevice, s), "input copy");
        run_plan(P, s, P.in_copy.p, out, ws ? ws : P.own_ws.p);
        return;
    }
    run_plan(P, s, in, out, ws ? ws : P.own_ws.p);
}

mifft_status mifft_execute_static_scratch(const mifft_plan* p, hipStream_t stream, const void* in, uint64_t in_bytes,
                                          void* out, uint64_t out_bytes) {
    if (!p) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null plan");
    return guard([&] {
        const Plan& P = *p->p;
        check_sizes(P, in, in_bytes, out, out_bytes);
        // in place on a plan whose kernels cannot run in place: the input
        // copy goes into the scratch after the workspace
        const bool copy_in = in == out && !P.in_place_kernels;
        const uint64_t in_off = (P.ws_bytes + 255) / 256 * 256;
        const uint64_t need = copy_in ? in_off + P.in_bytes : P.ws_bytes;
        if (need > kStaticScratchBytes)
            fail(MIFFT_ERR_WORKSPACE_TOO_SMALL, "this execution needs " + std::to_string(need) +
                                                    " bytes of scratch; the static scratch holds " +
                                                    std::to_string(kStaticScratchBytes));
        int dev = -1;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        if (dev != P.device) fail(MIFFT_ERR_INVALID_ENUM, "the plan was created on another device");
        char* sc = (char*)static_scratch(dev);
        if (copy_in) {
            hip_check(hipMemcpyAsync(sc + in_off, in, P.in_bytes, hipMemcpyDeviceToDevice, stream), "input copy");
            run_plan(P, stream, sc + in_off, out, sc);
        } else {
            run_plan(P, stream, in, out, sc);
        }
    });
}

// ------------------------------------------------------------ logical views
// FftLogicalLayout semantics (src/runtime/logical_io.rs:100-178).
static uint64_t checked_mul(uint64_t a, uint64_t b) {
    if (a && b > ~0ull / a) fail(MIFFT_ERR_LAYOUT_TOO_LARGE, "layout value overflows 64 bits");
    return a * b;
}
static uint64_t checked_add(uint64_t a, uint64_t b) {
    if (b > ~0ull - a) fail(MIFFT_ERR_LAYOUT_TOO_LARGE, "layout value overflows 64 bits");
    return a + b;
}

struct LogicalSide {
    uint64_t off = 0, se = 1, sb = 0;  // resolved batch stride
    uint64_t span = 0;                 // elements spanned
    bool direct = true;                // usable as a contiguous buffer at base + off
};

static LogicalSide resolve_layout(const mifft_logical_layout* l, uint64_t per_batch, uint64_t batch) {
    LogicalSide r;
    if (l) {
        r.off = l->element_offset;
        r.se = l->element_stride;
    }
    if (r.se == 0) fail(MIFFT_ERR_LAYOUT_ZERO_STRIDE, "layout element stride must be greater than zero");
    if (per_batch == 0 || batch == 0) return r;
    const uint64_t per_span = checked_add(checked_mul(r.se, per_batch - 1), 1);
    r.sb = (l && l->batch_stride) ? l->batch_stride : per_span;
    if (r.sb < per_span)
        fail(MIFFT_ERR_LAYOUT_BATCH_STRIDE, "layout batch stride is too small: required at least " +
                                                std::to_string(per_span) + " elements, got " + std::to_string(r.sb));
    r.span = checked_add(r.off, checked_add(checked_mul(batch - 1, r.sb), per_span));
    r.direct = r.se == 1 && (batch <= 1 || r.sb == per_batch);
    return r;
}

// Elements per batch and element bytes of a plan's input / output side.
static void endpoint(const Plan& P, bool input, uint64_t* per_batch, uint32_t* es) {
    const bool real = (P.kind == MIFFT_R2C && input) || (P.kind == MIFFT_C2R && !input);
    *es = real ? 4u : (uint32_t)(P.prec == P_F32 ? 8 : 16);
    const uint64_t bytes = input ? P.in_bytes : P.out_bytes;
    *per_batch = bytes / *es / std::max<uint64_t>(1, P.cfg.batch);
}

static uint64_t stage_align(uint64_t b) { return (b + 255) & ~255ull; }

struct LogicalExec {
    LogicalSide in, out;
    uint64_t in_pb = 0, out_pb = 0;
    uint32_t in_es = 0, out_es = 0;
    uint64_t ws_bytes = 0, stage_in = 0, stage_out = 0;  // byte offsets / sizes within the workspace
};

static LogicalExec plan_logical(const Plan& P, const void* in, const void* out, const mifft_logical_layout* li,
                                const mifft_logical_layout* lo) {
    LogicalExec e;
    const uint64_t batch = std::max<uint64_t>(1, P.cfg.batch);
    endpoint(P, true, &e.in_pb, &e.in_es);
    endpoint(P, false, &e.out_pb, &e.out_es);
    e.in = resolve_layout(li, e.in_pb, batch);
    e.out = resolve_layout(lo, e.out_pb, batch);
    // a contiguous side at an offset keeps the kernels' 16-byte alignment
    if (e.in.direct && (((uintptr_t)in + e.in.off * e.in_es) & 15u)) e.in.direct = false;
    if (e.out.direct && (((uintptr_t)out + e.out.off * e.out_es) & 15u)) e.out.direct = false;
    // contiguous sides that overlap without coinciding: read the input from staging
    if (e.in.direct && e.out.direct && in && out) {
        const uintptr_t a0 = (uintptr_t)in + e.in.off * e.in_es, a1 = a0 + P.in_bytes;
        const uintptr_t b0 = (uintptr_t)out + e.out.off * e.out_es, b1 = b0 + P.out_bytes;
        if (a0 != b0 && a0 < b1 && b0 < a1) e.in.direct = false;
    }
    uint64_t b = stage_align(P.ws_bytes);
    if (!e.in.direct) {
        e.stage_in = b;
        b += stage_align(P.in_bytes);
    }
    if (!e.out.direct) {
        e.stage_out = b;
        b += stage_align(P.out_bytes);
    }
    e.ws_bytes = b;
    return e;
}

static void exec_logical(const mifft_plan* p, hipStream_t s, const void* in, uint64_t in_bytes,
                         const mifft_logical_layout* li, void* out, uint64_t out_bytes,
                         const mifft_logical_layout* lo, void* ws, uint64_t ws_bytes, bool caller_ws) {
    if (!p) fail(MIFFT_ERR_NULL_ARGUMENT, "null plan");
    if (!in || !out) fail(MIFFT_ERR_NULL_ARGUMENT, "null buffer");
    if (((uintptr_t)in | (uintptr_t)out) & 15u) fail(MIFFT_ERR_UNALIGNED, "device buffers must be 16-byte aligned");
    Plan& P = *p->p;
    const LogicalExec e = plan_logical(P, in, out, li, lo);
    const uint64_t batch = std::max<uint64_t>(1, P.cfg.batch);
    if (P.large_chunk && !(e.in.direct && e.out.direct))
        fail(MIFFT_ERR_LAYOUT_TOO_LARGE, "strided logical views of large-chunk plans (more than u32::MAX elements) "
                                         "are not supported; use contiguous views");
    const uint64_t need_in = checked_mul(e.in.span, e.in_es), need_out = checked_mul(e.out.span, e.out_es);
    if (in_bytes < need_in)
        fail(MIFFT_ERR_LAYOUT_OUT_OF_BOUNDS, "input layout is out of bounds: required " + std::to_string(need_in) +
                                                 " bytes, got " + std::to_string(in_bytes));
    if (out_bytes < need_out)
        fail(MIFFT_ERR_LAYOUT_OUT_OF_BOUNDS, "output layout is out of bounds: required " +
                                                 std::to_string(need_out) + " bytes, got " + std::to_string(out_bytes));
    if (e.in.direct && e.out.direct) {
        // both sides contiguous: the plain execution at the offsets
        exec_impl(p, s, (const char*)in + e.in.off * e.in_es, P.in_bytes, (char*)out + e.out.off * e.out_es,
                  P.out_bytes, caller_ws ? ws : nullptr, caller_ws ? ws_bytes : 0);
        return;
    }
    char* base = nullptr;
    if (caller_ws) {
        if (!ws) fail(MIFFT_ERR_NULL_ARGUMENT, "null workspace");
        if (ws_bytes < e.ws_bytes)
            fail(MIFFT_ERR_WORKSPACE_TOO_SMALL, "workspace needs " + std::to_string(e.ws_bytes) + " bytes, got " +
                                                    std::to_string(ws_bytes));
        base = (char*)ws;
    } else {
        std::lock_guard<std::mutex> g(P.ws_mu);
        if (P.stage.bytes < e.ws_bytes) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            (void)hipStreamIsCapturing(s, &cs);
            if (cs != hipStreamCaptureStatusNone)
                fail(MIFFT_ERR_RECORDER_STATE, "first strided logical execution cannot be graph-captured");
            if (P.stage.p) hip_check(hipFree(P.stage.p), "staging free");
            P.stage.p = nullptr;
            P.stage.bytes = 0;
            hip_check(hipMalloc(&P.stage.p, e.ws_bytes), "logical view staging");
            P.stage.bytes = e.ws_bytes;
        }
        base = (char*)P.stage.p;
    }
    void* plan_ws = P.ws_bytes ? (void*)base : nullptr;
    const void* src = (const char*)in + e.in.off * e.in_es;
    void* dst = (char*)out + e.out.off * e.out_es;
    if (!e.in.direct) {
        src = base + e.stage_in;
        hip_check(launch_strided_copy(e.in_es, in, (void*)src, e.in_pb, batch, e.in.off, e.in.se, e.in.sb, 0, 1,
                                      e.in_pb, s),
                  "logical input gather");
    }
    if (!e.out.direct) dst = base + e.stage_out;
    // src and dst never alias here (one side is staging)
    if (plan_ws) run_plan(P, s, src, dst, plan_ws);
    else run_plan(P, s, src, dst, nullptr);
    if (!e.out.direct)
        hip_check(launch_strided_copy(e.out_es, dst, out, e.out_pb, batch, 0, 1, e.out_pb, e.out.off, e.out.se,
                                      e.out.sb, s),
                  "logical output scatter");
}

}  // extern "C"

// ------------------------------------------------------- execution blockers
static const char* plan_route_name(const Plan& P) {
    static const char* routes[] = {"direct-dft", "mixed-radix", "rader", "bluestein", "axis-sequence"};
    return P.route >= 0 && P.route <= 4 ? routes[P.route] : "unknown";
}

// What stops executing P over these buffers (plan.rs:1396-1500): unaligned
// pointers, layouts the reference rejects, sides smaller than their logical
// view, a caller workspace smaller than required.  Pointers may be NULL (not
// checked); NULL layouts are contiguous; ws_bytes NULL means the plan's own
// workspace.
static std::vector<Blocker> exec_blockers(const Plan& P, const void* in, uint64_t in_bytes,
                                          const mifft_logical_layout* li, const void* out, uint64_t out_bytes,
                                          const mifft_logical_layout* lo, const void* ws, const uint64_t* ws_bytes) {
    std::vector<Blocker> v;
    const std::string route = plan_route_name(P);
    auto align = [&](const void* ptr, const char* stage, const char* helper) {
        if (ptr && ((uintptr_t)ptr & 15u)) {
            Blocker b;
            b.set("Alignment", "logical-io", stage, "storage binding").req(16).act((uint64_t)((uintptr_t)ptr & 15u));
            b.helper = helper;
            b.reason = "device buffers must be 16-byte aligned";
            v.push_back(b);
        }
    };
    align(in, "storage-window", "");
    align(out, "storage-window", "");
    align(ws, "workspace", "workspace");
    const uint64_t batch = std::max<uint64_t>(1, P.cfg.batch);
    bool layouts_ok = true;
    auto side = [&](bool input, const mifft_logical_layout* l, uint64_t bytes) {
        uint64_t pb = 0;
        uint32_t es = 0;
        endpoint(P, input, &pb, &es);
        const char* stage = input ? "input-logical-view" : "output-logical-view";
        try {
            const LogicalSide r = resolve_layout(l, pb, batch);
            const bool contiguous = !l || (l->element_stride == 1 && (batch <= 1 || r.sb == pb));
            const uint64_t need = checked_mul(r.span, es);
            if (bytes < need) {
                Blocker b;
                b.set("Validation", route, stage, contiguous ? "contiguous" : "strided").req(need).act(bytes);
                b.reason = std::string(input ? "input" : "output") + " view needs " + std::to_string(need) +
                           " bytes, got " + std::to_string(bytes);
                v.push_back(b);
            }
        } catch (const Error& e) {
            layouts_ok = false;
            Blocker b;
            b.reason = e.detail;
            switch (e.code) {
                case MIFFT_ERR_LAYOUT_ZERO_STRIDE:
                    b.set("Layout", "logical-io", "logical-layout", "element-stride").req(1).act(0);
                    break;
                case MIFFT_ERR_LAYOUT_BATCH_STRIDE: {
                    const uint64_t per_span = l ? l->element_stride * (pb - 1) + 1 : pb;
                    b.set("Layout", "logical-io", "logical-layout", "batch-stride").req(per_span).act(l ? l->batch_stride : 0);
                    break;
                }
                default:
                    b.set("DeviceLimit", "logical-io", "logical-layout", "logical layout span").lim(~0ull);
                    break;
            }
            v.push_back(b);
        }
    };
    side(true, li, in_bytes);
    side(false, lo, out_bytes);
    if (ws_bytes && layouts_ok) {
        // plain executions take the plan's workspace; logical views add
        // staging for their strided sides
        uint64_t need = P.ws_bytes;
        if (li || lo) {
            const LogicalExec e = plan_logical(P, in, out, li, lo);
            need = (e.in.direct && e.out.direct) ? P.ws_bytes : e.ws_bytes;
        }
        if (*ws_bytes < need) {
            Blocker b;
            b.set("Workspace", route, "workspace", "").req(need).act(*ws_bytes);
            b.helper = "workspace";
            b.reason = "workspace needs " + std::to_string(need) + " bytes, got " + std::to_string(*ws_bytes);
            v.push_back(b);
        }
    }
    return v;
}

static std::string exec_diagnostics(const Plan& P, const std::vector<Blocker>& bl) {
    std::vector<std::string> js;
    for (const Blocker& b : bl) js.push_back(blocker_json(b));
    return plan_diagnostics_json(P, js);
}

// Runs an execution entry point; on failure the thread's last-error
// diagnostics become the plan's diagnostics with the blockers of these
// buffers, plus one for the error itself when none of them names it
// (FftExecutionError, error.rs:180-208).
template <class F>
static mifft_status guard_exec(const mifft_plan* p, const void* in, uint64_t in_bytes, const mifft_logical_layout* li,
                               const void* out, uint64_t out_bytes, const mifft_logical_layout* lo, const void* ws,
                               const uint64_t* ws_bytes, F&& f) {
    const mifft_status s = guard(std::forward<F>(f));
    if (s == MIFFT_SUCCESS || !p) return s;
    const std::string detail = g_detail;
    try {
        std::vector<Blocker> bl = exec_blockers(*p->p, in, in_bytes, li, out, out_bytes, lo, ws, ws_bytes);
        if (bl.empty()) {
            Blocker b;
            b.reason = detail;
            if (s == MIFFT_ERR_NULL_ARGUMENT) b.set("Validation", plan_route_name(*p->p), "arguments", "null pointer");
            else if (s == MIFFT_ERR_IN_PLACE_UNSUPPORTED) b.set("Validation", plan_route_name(*p->p), "in-place", "");
            else b.set("DeviceLimit", plan_route_name(*p->p), "kernel", "");
            bl.push_back(b);
        }
        g_diag = exec_diagnostics(*p->p, bl);
    } catch (...) {
        g_diag.clear();
    }
    g_detail = detail;
    return s;
}

extern "C" {

mifft_status mifft_execute(const mifft_plan* p, hipStream_t stream, const void* in, uint64_t in_bytes,
                           void* out, uint64_t out_bytes) {
    return guard_exec(p, in, in_bytes, nullptr, out, out_bytes, nullptr, nullptr, nullptr,
                      [&] { exec_impl(p, stream, in, in_bytes, out, out_bytes, nullptr, 0); });
}

mifft_status mifft_execute_with_workspace(const mifft_plan* p, hipStream_t stream, const void* in,
                                          uint64_t in_bytes, void* out, uint64_t out_bytes,
                                          void* workspace, uint64_t workspace_bytes) {
    return guard_exec(p, in, in_bytes, nullptr, out, out_bytes, nullptr, workspace, &workspace_bytes, [&] {
        if (p && p->p->ws_bytes && !workspace) fail(MIFFT_ERR_NULL_ARGUMENT, "null workspace");
        exec_impl(p, stream, in, in_bytes, out, out_bytes, workspace, workspace_bytes);
    });
}

mifft_status mifft_execute_in_place(const mifft_plan* p, hipStream_t stream, void* buf, uint64_t bytes) {
    return guard_exec(p, buf, bytes, nullptr, buf, bytes, nullptr, nullptr, nullptr,
                      [&] { exec_impl(p, stream, buf, bytes, buf, bytes, nullptr, 0); });
}

// FftPlan::diagnostics_for_workspace (plan.rs:1023-1055).
mifft_status mifft_plan_diagnostics_for_workspace(const mifft_plan* p, const void* in, uint64_t in_bytes,
                                                  const void* out, uint64_t out_bytes, const void* workspace,
                                                  uint64_t workspace_bytes, char* buf, size_t cap, size_t* needed) {
    if (!p) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null plan");
    std::string d;
    const mifft_status s = guard([&] {
        d = exec_diagnostics(*p->p, exec_blockers(*p->p, in, in_bytes, nullptr, out, out_bytes, nullptr, workspace,
                                                  &workspace_bytes));
    });
    return s != MIFFT_SUCCESS ? s : write_json(d, buf, cap, needed);
}

// FftPlan::diagnostics_for_views / _for_logical_views(_with_workspace)
// (plan.rs:1057-1300).
mifft_status mifft_plan_diagnostics_for_views(const mifft_plan* p, const void* in, uint64_t in_bytes,
                                              const mifft_logical_layout* in_layout, const void* out,
                                              uint64_t out_bytes, const mifft_logical_layout* out_layout,
                                              const uint64_t* workspace_bytes, char* buf, size_t cap,
                                              size_t* needed) {
    if (!p) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null plan");
    std::string d;
    const mifft_status s = guard([&] {
        d = exec_diagnostics(*p->p, exec_blockers(*p->p, in, in_bytes, in_layout, out, out_bytes, out_layout,
                                                  nullptr, workspace_bytes));
    });
    return s != MIFFT_SUCCESS ? s : write_json(d, buf, cap, needed);
}

mifft_status mifft_recorder_begin(hipStream_t stream, int32_t capture_graph, mifft_recorder** out) {
    if (!out) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null output handle");
    *out = nullptr;
    return guard([&] {
        auto r = std::make_unique<mifft_recorder>();
        r->stream = stream;
        r->capture = capture_graph != 0;
        if (r->capture)
            hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture");
        *out = r.release();
    });
}

hipStream_t mifft_recorder_stream(const mifft_recorder* r) { return r ? r->stream : nullptr; }

mifft_status mifft_record(mifft_recorder* r, const mifft_plan* p, const void* in, uint64_t in_bytes,
                          void* out, uint64_t out_bytes) {
    if (!r) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null recorder");
    if (r->ended) return set_err(MIFFT_ERR_RECORDER_STATE, "recorder already ended");
    return guard([&] { exec_impl(p, r->stream, in, in_bytes, out, out_bytes, nullptr, 0); });
}

mifft_status mifft_record_logical(mifft_recorder* r, const mifft_plan* p, const void* in, uint64_t in_bytes,
                                  const mifft_logical_layout* in_layout, void* out, uint64_t out_bytes,
                                  const mifft_logical_layout* out_layout) {
    if (!r) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null recorder");
    if (r->ended) return set_err(MIFFT_ERR_RECORDER_STATE, "recorder already ended");
    return guard([&] {
        exec_logical(p, r->stream, in, in_bytes, in_layout, out, out_bytes, out_layout, nullptr, 0, false);
    });
}

void mifft_logical_layout_init(mifft_logical_layout* l) {
    if (l) *l = mifft_logical_layout{0, 1, 0};
}

mifft_status mifft_logical_layout_span(const mifft_logical_layout* l, uint64_t per_batch, uint64_t batch,
                                       uint64_t* elements) {
    if (!elements) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null output");
    return guard([&] { *elements = resolve_layout(l, per_batch, batch).span; });
}

mifft_status mifft_plan_logical_workspace_bytes(const mifft_plan* p, const mifft_logical_layout* in_layout,
                                                const mifft_logical_layout* out_layout, uint64_t* bytes) {
    if (!p || !bytes) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null argument");
    return guard([&] {
        // alignment of the caller's buffers is unknown here: count staging for
        // every side that is not contiguous by its layout alone
        const LogicalExec e = plan_logical(*p->p, nullptr, nullptr, in_layout, out_layout);
        *bytes = (e.in.direct && e.out.direct) ? p->p->ws_bytes : e.ws_bytes;
    });
}

mifft_status mifft_execute_logical(const mifft_plan* p, hipStream_t stream, const void* in, uint64_t in_bytes,
                                   const mifft_logical_layout* in_layout, void* out, uint64_t out_bytes,
                                   const mifft_logical_layout* out_layout) {
    return guard_exec(p, in, in_bytes, in_layout, out, out_bytes, out_layout, nullptr, nullptr, [&] {
        exec_logical(p, stream, in, in_bytes, in_layout, out, out_bytes, out_layout, nullptr, 0, false);
    });
}

mifft_status mifft_execute_logical_with_workspace(const mifft_plan* p, hipStream_t stream, const void* in,
                                                  uint64_t in_bytes, const mifft_logical_layout* in_layout,
                                                  void* out, uint64_t out_bytes,
                                                  const mifft_logical_layout* out_layout, void* workspace,
                                                  uint64_t workspace_bytes) {
    return guard_exec(p, in, in_bytes, in_layout, out, out_bytes, out_layout, workspace, &workspace_bytes, [&] {
        exec_logical(p, stream, in, in_bytes, in_layout, out, out_bytes, out_layout, workspace, workspace_bytes,
                     true);
    });
}

mifft_status mifft_record_in_place(mifft_recorder* r, const mifft_plan* p, void* buf, uint64_t bytes) {
    return mifft_record(r, p, buf, bytes, buf, bytes);
}

mifft_status mifft_recorder_end(mifft_recorder* r) {
    if (!r) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null recorder");
    if (r->ended) return set_err(MIFFT_ERR_RECORDER_STATE, "recorder already ended");
    r->ended = true;
    if (!r->capture) return MIFFT_SUCCESS;
    return guard([&] {
        hip_check(hipStreamEndCapture(r->stream, &r->graph), "end capture");
        hip_check(hipGraphInstantiate(&r->exec, r->graph, nullptr, nullptr, 0), "graph instantiate");
    });
}

mifft_status mifft_recorder_launch(mifft_recorder* r, hipStream_t stream) {
    if (!r) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null recorder");
    if (!r->exec) return set_err(MIFFT_ERR_RECORDER_STATE, "no captured graph");
    return guard([&] { hip_check(hipGraphLaunch(r->exec, stream), "graph launch"); });
}

void mifft_recorder_destroy(mifft_recorder* r) {
    if (!r) return;
    if (r->capture && !r->ended) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(r->stream, &g);
        if (g) (void)hipGraphDestroy(g);
    }
    if (r->exec) (void)hipGraphExecDestroy(r->exec);
    if (r->graph) (void)hipGraphDestroy(r->graph);
    delete r;
}

const char* mifft_status_string(mifft_status s) {
    switch (s) {
        case MIFFT_SUCCESS: return "success";
        case MIFFT_ERR_ZERO_LENGTH: return "zero length";
        case MIFFT_ERR_ZERO_BATCH: return "zero batch";
        case MIFFT_ERR_LENGTH_TOO_LARGE: return "length too large";
        case MIFFT_ERR_UNSUPPORTED_LENGTH: return "unsupported length";
        case MIFFT_ERR_PRECISION_UNSUPPORTED: return "precision unsupported";
        case MIFFT_ERR_INVALID_TUNING: return "invalid tuning";
        case MIFFT_ERR_EMPTY_AXES: return "empty axes";
        case MIFFT_ERR_INVALID_AXIS: return "invalid axis";
        case MIFFT_ERR_DUPLICATE_AXIS: return "duplicate axis";
        case MIFFT_ERR_INVALID_REAL_DIRECTION: return "invalid real transform direction";
        case MIFFT_ERR_UNSUPPORTED_REAL_AXES: return "unsupported real axes";
        case MIFFT_ERR_BUFFER_TOO_SMALL: return "buffer too small";
        case MIFFT_ERR_WORKSPACE_TOO_SMALL: return "workspace too small";
        case MIFFT_ERR_IN_PLACE_UNSUPPORTED: return "in-place unsupported";
        case MIFFT_ERR_ALLOCATION_FAILED: return "allocation failed";
        case MIFFT_ERR_HOST_BUFFER_LENGTH: return "host buffer length";
        case MIFFT_ERR_NULL_ARGUMENT: return "null argument";
        case MIFFT_ERR_INVALID_ENUM: return "invalid enum";
        case MIFFT_ERR_RANK_TOO_LARGE: return "rank too large";
        case MIFFT_ERR_HIP: return "HIP runtime error";
        case MIFFT_ERR_RECORDER_STATE: return "recorder state";
        case MIFFT_ERR_OVER_BUDGET: return "over memory budget";
        case MIFFT_ERR_BUSY: return "busy";
        case MIFFT_ERR_UNALIGNED: return "buffer not 16-byte aligned";
        case MIFFT_ERR_LAYOUT_ZERO_STRIDE: return "layout element stride is zero";
        case MIFFT_ERR_LAYOUT_BATCH_STRIDE: return "layout batch stride too small";
        case MIFFT_ERR_LAYOUT_OUT_OF_BOUNDS: return "layout out of bounds";
        case MIFFT_ERR_LAYOUT_TOO_LARGE: return "layout value too large";
    }
    return "unknown status";
}

const char* mifft_last_error_detail(void) { return g_detail.c_str(); }

const char* mifft_version(void) { return "mifft 0.1.0 (gfx950)"; }

mifft_status mifft_autotune(const mifft_config* config, mifft_kind kind, int32_t reps, hipStream_t stream,
                            int32_t* passes_tuned, char* report, uint64_t report_cap) {
    if (!config) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null config");
    return guard([&] {
        std::string text;
        const int n = autotune(*config, kind, reps, stream, &text);
        if (passes_tuned) *passes_tuned = n;
        if (report && report_cap) {
            const size_t k = std::min<size_t>(text.size(), (size_t)report_cap - 1);
            std::memcpy(report, text.data(), k);
            report[k] = 0;
        }
    });
}

mifft_status mifft_autotune_sequence(const mifft_config* configs, const mifft_kind* kinds, int32_t count,
                                     int32_t reps, hipStream_t stream, int32_t* shapes_tuned, char* report,
                                     uint64_t report_cap) {
    if (!configs || !kinds) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null configs or kinds");
    return guard([&] {
        std::string text;
        const int n = autotune_sequence(configs, kinds, count, reps, stream, &text);
        if (shapes_tuned) *shapes_tuned = n;
        if (report && report_cap) {
            const size_t k = std::min<size_t>(text.size(), (size_t)report_cap - 1);
            std::memcpy(report, text.data(), k);
            report[k] = 0;
        }
    });
}

mifft_status mifft_tuning_table_save(const char* path) {
    if (!path) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null path");
    return guard([&] {
        std::string text = "# mifft tuning table: key = choice\n";
        for (auto& kv : tuning_entries()) text += kv.first + " = " + kv.second + "\n";
        FILE* f = std::fopen(path, "w");
        if (!f) fail(MIFFT_ERR_INVALID_TUNING, std::string("cannot write ") + path);
        const size_t w = std::fwrite(text.data(), 1, text.size(), f);
        const int c = std::fclose(f);
        if (w != text.size() || c != 0) fail(MIFFT_ERR_INVALID_TUNING, std::string("cannot write ") + path);
    });
}

mifft_status mifft_tuning_table_load(const char* path) {
    if (!path) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null path");
    return guard([&] {
        FILE* f = std::fopen(path, "r");
        if (!f) fail(MIFFT_ERR_INVALID_TUNING, std::string("cannot read ") + path);
        std::string text;
        char buf[4096];
        size_t r;
        while ((r = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, r);
        std::fclose(f);
        tuning_parse(text);
    });
}

mifft_status mifft_tuning_table_parse(const char* text, int32_t* entries) {
    if (!text) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null text");
    return guard([&] {
        const int n = tuning_parse(text);
        if (entries) *entries = n;
    });
}

void mifft_tuning_table_clear(void) { tuning_clear(); }

int32_t mifft_tuning_table_size(void) { return (int32_t)tuning_entries().size(); }

mifft_status mifft_fill_uniform_f32(float* dst, uint64_t count, uint64_t seed, hipStream_t stream) {
    if (!dst) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null buffer");
    return guard([&] { hip_check(fill_uniform_f32(dst, count, seed, stream), "fill"); });
}

mifft_status mifft_fill_normal_f32(float* dst, uint64_t count, uint64_t seed, hipStream_t stream) {
    if (!dst) return set_err(MIFFT_ERR_NULL_ARGUMENT, "null buffer");
    return guard([&] { hip_check(fill_normal_f32(dst, count, seed, stream), "fill"); });
}

}  // extern "C"
