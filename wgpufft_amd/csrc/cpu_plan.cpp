// C ABI of the host-memory FFT plans (mifft_cpu_*): CpuFftPlan
// (src/cpu.rs:52-361) with the same configuration semantics, buffer layouts
// and error variants as the GPU plans.  The engine (cpu_fft_core.inc) is
// compiled for AVX2 + FMA and for the x86-64 baseline; the first plan picks
// one from what the CPU reports.
#include <cstring>
#include <memory>
#include <string>

#include "cpu_fft.hpp"
#include "plan.hpp"

using namespace mifft;

namespace {
struct Engine {
    void* (*create)(int, int, int, const uint64_t*, const int*, int, uint64_t, int, double);
    void (*destroy)(void*);
    void (*run)(void*, const void*, void*);
    void (*set_max_threads)(unsigned);
    unsigned (*max_threads)();
    const char* isa;
};

const Engine& engine() {
    static const Engine e = [] {
        if (cpu_has_avx2_fma())
            return Engine{cpu_avx2::create, cpu_avx2::destroy, cpu_avx2::run, cpu_avx2::set_max_threads,
                          cpu_avx2::max_threads, "avx2+fma"};
        return Engine{cpu_base::create, cpu_base::destroy, cpu_base::run, cpu_base::set_max_threads,
                      cpu_base::max_threads, "x86-64"};
    }();
    return e;
}

mifft_status cpu_err(mifft_status s, const std::string& d) {
    set_last_detail(d);
    return s;
}
}  // namespace

struct mifft_cpu_plan {
    mifft_config cfg{};
    mifft_kind kind = MIFFT_C2C;
    void* h = nullptr;
    uint64_t total = 0, packed = 0;
    std::vector<uint64_t> packed_shape;
    ~mifft_cpu_plan() {
        if (h) engine().destroy(h);
    }
};

extern "C" {

mifft_status mifft_cpu_plan_create(const mifft_config* c, mifft_kind kind, mifft_cpu_plan** out) {
    if (!c || !out) return cpu_err(MIFFT_ERR_NULL_ARGUMENT, "null config or output handle");
    *out = nullptr;
    try {
        validate_config(*c);
        if (kind != MIFFT_C2C && kind != MIFFT_R2C && kind != MIFFT_C2R)
            return cpu_err(MIFFT_ERR_INVALID_ENUM, "unknown transform kind");
        auto p = std::make_unique<mifft_cpu_plan>();
        p->cfg = *c;
        p->kind = kind;
        uint64_t per = 1;
        for (int i = 0; i < c->rank; i++) per *= c->shape[i];
        p->total = per * c->batch;
        if (kind != MIFFT_C2C) {
            // CpuFftPlan::real (cpu.rs:130-210)
            const int want = kind == MIFFT_R2C ? MIFFT_FORWARD : MIFFT_INVERSE;
            if (c->direction != want)
                return cpu_err(MIFFT_ERR_INVALID_REAL_DIRECTION,
                               std::string(kind == MIFFT_R2C ? "r2c" : "c2r") + " expects the " +
                                   (want == MIFFT_FORWARD ? "forward" : "inverse") + " direction");
            bool all = c->naxes == c->rank;
            for (int i = 0; all && i < c->naxes; i++) all = c->axes[i] == i;
            if (!all) return cpu_err(MIFFT_ERR_UNSUPPORTED_REAL_AXES, "real transforms run over every axis in order");
            if (c->precision == MIFFT_DF64)
                return cpu_err(MIFFT_ERR_PRECISION_UNSUPPORTED, "cpu-real-transforms-support-f32-and-f64");
            uint64_t pp = 1;
            for (int i = 0; i < c->rank; i++) {
                const uint64_t s = i == 0 ? c->shape[0] / 2 + 1 : c->shape[i];
                p->packed_shape.push_back(s);
                pp *= s;
            }
            p->packed = pp * c->batch;
        } else {
            p->packed = p->total;
        }
        int axes[MIFFT_MAX_RANK];
        for (int i = 0; i < c->naxes; i++) axes[i] = c->axes[i];
        const int prec = c->precision == MIFFT_F32 ? 0 : c->precision == MIFFT_F64 ? 1 : 2;
        p->h = engine().create((int)kind, prec, c->rank, c->shape, axes, c->naxes, c->batch,
                               c->direction == MIFFT_INVERSE, config_scale(*c));
        *out = p.release();
        return MIFFT_SUCCESS;
    } catch (const Error& e) {
        return cpu_err(e.code, e.detail);
    } catch (const std::bad_alloc&) {
        return cpu_err(MIFFT_ERR_ALLOCATION_FAILED, "host allocation failed");
    }
}

void mifft_cpu_plan_destroy(mifft_cpu_plan* p) { delete p; }

mifft_kind mifft_cpu_plan_kind(const mifft_cpu_plan* p) { return p ? p->kind : MIFFT_C2C; }

mifft_status mifft_cpu_plan_config(const mifft_cpu_plan* p, mifft_config* out) {
    if (!p || !out) return cpu_err(MIFFT_ERR_NULL_ARGUMENT, "null argument");
    *out = p->cfg;
    return MIFFT_SUCCESS;
}

int32_t mifft_cpu_plan_packed_shape(const mifft_cpu_plan* p, uint64_t* out, int32_t cap) {
    if (!p || p->kind == MIFFT_C2C) return 0;
    for (int32_t i = 0; i < (int32_t)p->packed_shape.size() && i < cap; i++) out[i] = p->packed_shape[i];
    return (int32_t)p->packed_shape.size();
}

static uint64_t words_per_complex(const mifft_cpu_plan* p) { return p->cfg.precision == MIFFT_DF64 ? 4 : 2; }

uint64_t mifft_cpu_plan_required_input_len(const mifft_cpu_plan* p) {
    if (!p) return 0;
    switch (p->kind) {
        case MIFFT_C2C: return p->total * words_per_complex(p);
        case MIFFT_R2C: return p->total;
        default: return p->packed * 2;
    }
}

uint64_t mifft_cpu_plan_required_output_len(const mifft_cpu_plan* p) {
    if (!p) return 0;
    switch (p->kind) {
        case MIFFT_C2C: return p->total * words_per_complex(p);
        case MIFFT_R2C: return p->packed * 2;
        default: return p->total;
    }
}

// CpuFftPlan::check_lengths (cpu.rs:345-361): exact lengths.
static mifft_status check_lengths(const mifft_cpu_plan* p, uint64_t in_len, uint64_t out_len) {
    const uint64_t ei = mifft_cpu_plan_required_input_len(p), eo = mifft_cpu_plan_required_output_len(p);
    if (in_len != ei)
        return cpu_err(MIFFT_ERR_HOST_BUFFER_LENGTH,
                       "input: expected " + std::to_string(ei) + " words, got " + std::to_string(in_len));
    if (out_len != eo)
        return cpu_err(MIFFT_ERR_HOST_BUFFER_LENGTH,
                       "output: expected " + std::to_string(eo) + " words, got " + std::to_string(out_len));
    return MIFFT_SUCCESS;
}

static mifft_status run_checked(const mifft_cpu_plan* p, const void* in, void* out) {
    try {
        engine().run(p->h, in, out);
        return MIFFT_SUCCESS;
    } catch (const std::bad_alloc&) {
        return cpu_err(MIFFT_ERR_ALLOCATION_FAILED, "host allocation failed");
    } catch (...) {
        return cpu_err(MIFFT_ERR_HIP, "unexpected internal error");
    }
}

// CpuFftPlan::execute (cpu.rs:242-265): F32 and Df64 plans on f32 words.
mifft_status mifft_cpu_execute(const mifft_cpu_plan* p, const float* in, uint64_t in_len, float* out,
                               uint64_t out_len) {
    if (!p || !in || !out) return cpu_err(MIFFT_ERR_NULL_ARGUMENT, "null argument");
    if (p->cfg.precision == MIFFT_F64)
        return cpu_err(MIFFT_ERR_PRECISION_UNSUPPORTED, "f64-plans-execute-with-execute_f64");
    const mifft_status s = check_lengths(p, in_len, out_len);
    return s != MIFFT_SUCCESS ? s : run_checked(p, in, out);
}

// CpuFftPlan::execute_f64 (cpu.rs:267-280).
mifft_status mifft_cpu_execute_f64(const mifft_cpu_plan* p, const double* in, uint64_t in_len, double* out,
                                   uint64_t out_len) {
    if (!p || !in || !out) return cpu_err(MIFFT_ERR_NULL_ARGUMENT, "null argument");
    if (p->cfg.precision != MIFFT_F64)
        return cpu_err(MIFFT_ERR_PRECISION_UNSUPPORTED, "execute_f64-requires-an-f64-plan");
    const mifft_status s = check_lengths(p, in_len, out_len);
    return s != MIFFT_SUCCESS ? s : run_checked(p, in, out);
}

// CpuFftPlan::check_in_place (cpu.rs:325-334).
static mifft_status check_in_place(const mifft_cpu_plan* p, uint64_t len) {
    if (p->kind != MIFFT_C2C)
        return cpu_err(MIFFT_ERR_IN_PLACE_UNSUPPORTED, "real-transforms-change-the-element-count");
    return check_lengths(p, len, len);
}

// CpuFftPlan::execute_in_place (cpu.rs:287-309).
mifft_status mifft_cpu_execute_in_place(const mifft_cpu_plan* p, float* data, uint64_t len) {
    if (!p || !data) return cpu_err(MIFFT_ERR_NULL_ARGUMENT, "null argument");
    mifft_status s = check_in_place(p, len);
    if (s != MIFFT_SUCCESS) return s;
    if (p->cfg.precision == MIFFT_F64)
        return cpu_err(MIFFT_ERR_PRECISION_UNSUPPORTED, "f64-plans-execute-with-execute_in_place_f64");
    return run_checked(p, data, data);
}

// CpuFftPlan::execute_in_place_f64 (cpu.rs:311-323).
mifft_status mifft_cpu_execute_in_place_f64(const mifft_cpu_plan* p, double* data, uint64_t len) {
    if (!p || !data) return cpu_err(MIFFT_ERR_NULL_ARGUMENT, "null argument");
    mifft_status s = check_in_place(p, len);
    if (s != MIFFT_SUCCESS) return s;
    if (p->cfg.precision != MIFFT_F64)
        return cpu_err(MIFFT_ERR_PRECISION_UNSUPPORTED, "execute_in_place_f64-requires-an-f64-plan");
    return run_checked(p, data, data);
}

void mifft_cpu_set_max_threads(int32_t n) { engine().set_max_threads(n > 0 ? (unsigned)n : 0u); }

int32_t mifft_cpu_max_threads(void) { return (int32_t)engine().max_threads(); }

const char* mifft_cpu_isa(void) { return engine().isa; }

}  // extern "C"
