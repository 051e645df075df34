// Compiled line-kernel variants and their launchers.
//
// The registry is the gfx950 counterpart of the reference's pipeline cache
// (src/runtime/pipeline_cache.rs:356-616): instead of generating WGSL per
// typed key at plan time, every register-resident variant is instantiated here
// at build time and looked up by (precision, N, layout, load/store mode).
// Lengths without a variant run the runtime-schedule LDS kernel.
#include <algorithm>
#include <mutex>
#include <unordered_map>

#include "launch.h"

namespace mifft {

hipError_t ensure_lds_attr(const void* fn, size_t bytes) {
    if (bytes <= 64 * 1024) return hipSuccess;
    static std::mutex mu;
    static std::unordered_map<const void*, size_t> done;
    std::lock_guard<std::mutex> g(mu);
    auto it = done.find(fn);
    if (it != done.end() && it->second >= bytes) return hipSuccess;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    done[fn] = bytes;
    return hipSuccess;
}

#define MIFFT_RL(...) RadixList<__VA_ARGS__>
#define MIFFT_K(FMT, N, TPL, LINES, LDLF, STLF, LDM, STM, POST, VEC, ...) \
    RegFft<FMT, N, TPL, LINES, LDLF, STLF, LDM, STM, POST, MIFFT_RL(__VA_ARGS__), VEC>
#define REGV(FMT, PREC, N, TPL, LINES, LDLF, STLF, LDM, STM, POST, VEC, RADSTR, ...)              \
    {PREC,                                                                                      \
     N,                                                                                         \
     TPL,                                                                                       \
     LINES,                                                                                     \
     LDLF,                                                                                      \
     STLF,                                                                                      \
     LDM,                                                                                       \
     STM,                                                                                       \
     POST,                                                                                      \
     VEC,                                                                                       \
     MIFFT_K(FMT, N, TPL, LINES, LDLF, STLF, LDM, STM, POST, VEC, __VA_ARGS__)::NT,             \
     MIFFT_K(FMT, N, TPL, LINES, LDLF, STLF, LDM, STM, POST, VEC, __VA_ARGS__)::lds_bytes(),    \
     (const void*)&reg_fft_kernel<MIFFT_K(FMT, N, TPL, LINES, LDLF, STLF, LDM, STM, POST, VEC,  \
                                          __VA_ARGS__)>,                                        \
     RADSTR}

// Persistent pipelined form (twiddles held in registers, see
// RegFft::run_pipelined), at least MINB waves per SIMD (the launch bound's
// register budget).
#define MIFFT_KP(FMT, N, TPL, LINES, LDLF, STLF, LDM, STM, POST, VEC, ...) \
    RegFft<FMT, N, TPL, LINES, LDLF, STLF, LDM, STM, POST, MIFFT_RL(__VA_ARGS__), VEC, true>
#define REGP(FMT, PREC, N, TPL, LINES, LDLF, STLF, LDM, STM, POST, VEC, MINB, RADSTR, ...)         \
    {PREC,                                                                                      \
     N,                                                                                         \
     TPL,                                                                                       \
     LINES,                                                                                     \
     LDLF,                                                                                      \
     STLF,                                                                                      \
     LDM,                                                                                       \
     STM,                                                                                       \
     POST,                                                                                      \
     VEC,                                                                                       \
     MIFFT_KP(FMT, N, TPL, LINES, LDLF, STLF, LDM, STM, POST, VEC, __VA_ARGS__)::NT,            \
     MIFFT_KP(FMT, N, TPL, LINES, LDLF, STLF, LDM, STM, POST, VEC, __VA_ARGS__)::lds_bytes(),   \
     (const void*)&reg_fft_pipe_kernel<MIFFT_KP(FMT, N, TPL, LINES, LDLF, STLF, LDM, STM, POST, \
                                                VEC, __VA_ARGS__),                              \
                                       MINB>,                                                   \
     RADSTR,                                                                                    \
     MINB}

// Fold form (F contiguous lines M/F rows apart per thread, the next axis's
// first radix-F stage on store): REGF one tile of F lines per workgroup, REGFP
// the persistent pipelined form (at least MINB waves per SIMD).
#define MIFFT_KF(FMT, N, TPL, F, LDM, STM, PF, ...) \
    RegFft<FMT, N, TPL, F, false, false, LDM, STM, false, MIFFT_RL(__VA_ARGS__), F, PF, true>
#define REGFM(FMT, PREC, N, TPL, F, LDM, STM, RADSTR, ...)                                            \
    {PREC, N, TPL, F, 0, 0, LDM, STM, 0, F,                                                        \
     MIFFT_KF(FMT, N, TPL, F, LDM, STM, false, __VA_ARGS__)::NT,                                   \
     MIFFT_KF(FMT, N, TPL, F, LDM, STM, false, __VA_ARGS__)::lds_bytes(),                          \
     (const void*)&reg_fft_kernel<MIFFT_KF(FMT, N, TPL, F, LDM, STM, false, __VA_ARGS__)>, RADSTR, 0, F}
#define REGF(FMT, PREC, N, TPL, F, RADSTR, ...) \
    REGFM(FMT, PREC, N, TPL, F, LD_COMPLEX, ST_COMPLEX, RADSTR, __VA_ARGS__)
#define REGFMP(FMT, PREC, N, TPL, F, LDM, STM, MINB, RADSTR, ...)                                  \
    {PREC, N, TPL, F, 0, 0, LDM, STM, 0, F,                                                        \
     MIFFT_KF(FMT, N, TPL, F, LDM, STM, true, __VA_ARGS__)::NT,                                    \
     MIFFT_KF(FMT, N, TPL, F, LDM, STM, true, __VA_ARGS__)::lds_bytes(),                           \
     (const void*)&reg_fft_pipe_kernel<MIFFT_KF(FMT, N, TPL, F, LDM, STM, true, __VA_ARGS__), MINB>, \
     RADSTR, MINB, F}
#define REGFP(FMT, PREC, N, TPL, F, MINB, RADSTR, ...) \
    REGFMP(FMT, PREC, N, TPL, F, LD_COMPLEX, ST_COMPLEX, MINB, RADSTR, __VA_ARGS__)

// Adjacent-butterfly forms (16-byte global accesses) of contiguous complex
// lines: REGVA one line per thread group, REGFA the fold.
#define MIFFT_KA(N, TPL, LINES, ...) \
    RegFft<FmtF32, N, TPL, LINES, false, false, LD_COMPLEX, ST_COMPLEX, false, MIFFT_RL(__VA_ARGS__), 1, false, false, true>
#define REGVA(N, TPL, LINES, RADSTR, ...)                                                            \
    {P_F32, N, TPL, LINES, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, MIFFT_KA(N, TPL, LINES, __VA_ARGS__)::NT, \
     MIFFT_KA(N, TPL, LINES, __VA_ARGS__)::lds_bytes(),                                            \
     (const void*)&reg_fft_kernel<MIFFT_KA(N, TPL, LINES, __VA_ARGS__)>, RADSTR, 0, 0, 1}
#define MIFFT_KFA(N, TPL, F, ...) \
    RegFft<FmtF32, N, TPL, F, false, false, LD_COMPLEX, ST_COMPLEX, false, MIFFT_RL(__VA_ARGS__), F, false, true, true>
#define REGFA(N, TPL, F, RADSTR, ...)                                                                \
    {P_F32, N, TPL, F, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, F, MIFFT_KFA(N, TPL, F, __VA_ARGS__)::NT,       \
     MIFFT_KFA(N, TPL, F, __VA_ARGS__)::lds_bytes(),                                                \
     (const void*)&reg_fft_kernel<MIFFT_KFA(N, TPL, F, __VA_ARGS__)>, RADSTR, 0, F, 1}

// f32 power-of-two lines: 16 points per thread and line, radix-16 stages.
#define F32_POW2(LDLF, STLF, LDM, STM, POST, VEC, L4096, L2048, L1024, L512, L256, L128, L64)          \
    REGV(FmtF32, P_F32, 4096, 256, L4096, LDLF, STLF, LDM, STM, POST, VEC, "16x16x16", 16, 16, 16),    \
        REGV(FmtF32, P_F32, 2048, 128, L2048, LDLF, STLF, LDM, STM, POST, VEC, "16x16x8", 16, 16, 8),  \
        REGV(FmtF32, P_F32, 1024, 64, L1024, LDLF, STLF, LDM, STM, POST, VEC, "16x16x4", 16, 16, 4),   \
        REGV(FmtF32, P_F32, 512, 32, L512, LDLF, STLF, LDM, STM, POST, VEC, "8x8x8", 8, 8, 8),         \
        REGV(FmtF32, P_F32, 256, 16, L256, LDLF, STLF, LDM, STM, POST, VEC, "16x16", 16, 16),          \
        REGV(FmtF32, P_F32, 128, 8, L128, LDLF, STLF, LDM, STM, POST, VEC, "16x8", 16, 8),             \
        REGV(FmtF32, P_F32, 64, 4, L64, LDLF, STLF, LDM, STM, POST, VEC, "16x4", 16, 4)

static const RegVariant kRegistry[] = {
    // contiguous axes (axis 0)
    F32_POW2(0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, 1, 2, 4, 8, 16, 32, 64),
    // strided axes, two adjacent lines per thread: 16-byte accesses and
    // >= 128-byte row segments per workgroup
    F32_POW2(1, 1, LD_COMPLEX, ST_COMPLEX, 0, 2, 8, 16, 32, 32, 32, 64, 128),
    // strided axes whose line count is odd (one line per thread)
    F32_POW2(1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, 4, 8, 16, 16, 16, 32, 64),
    // R2C axis 0: real in, packed half spectrum out
    F32_POW2(0, 0, LD_REAL, ST_PACKED, 0, 1, 1, 2, 4, 8, 16, 32, 64),
    // C2R axis 0: Hermitian-extended packed in, real out
    F32_POW2(0, 0, LD_HERMITIAN, ST_REAL, 0, 1, 1, 2, 4, 8, 16, 32, 64),
    // four-step split passes: columns with twiddle on store; rows with transposed store
    // (line pairs: the columns of a strided long axis, whose balanced splits
    // of 2^14..2^24 points give 128- to 4096-point columns)
    REGV(FmtF32, P_F32, 4096, 256, 8, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 2, "16x16x16", 16, 16, 16),
    REGV(FmtF32, P_F32, 2048, 128, 16, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 2, "16x16x8", 16, 16, 8),
    REGV(FmtF32, P_F32, 1024, 64, 32, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 2, "16x16x4", 16, 16, 4),
    REGV(FmtF32, P_F32, 512, 32, 32, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 2, "8x8x8", 8, 8, 8),
    REGV(FmtF32, P_F32, 256, 16, 32, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 2, "16x16", 16, 16),
    REGV(FmtF32, P_F32, 128, 8, 64, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 2, "16x8", 16, 8),
    F32_POW2(1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, 4, 8, 16, 16, 16, 32, 64),
    F32_POW2(0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, 4, 8, 16, 16, 16, 32, 64),
    // the same passes with a quarter and a half of the lines per workgroup:
    // up to 512 workgroups for the 4096- to 16384-line passes of 2^19..2^21
    // points (64 threads per workgroup at a quarter)
    REGV(FmtF32, P_F32, 256, 16, 4, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "16x16", 16, 16),
    REGV(FmtF32, P_F32, 256, 16, 8, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "16x16", 16, 16),
    REGV(FmtF32, P_F32, 128, 8, 8, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "16x8", 16, 8),
    REGV(FmtF32, P_F32, 128, 8, 16, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "16x8", 16, 8),
    REGV(FmtF32, P_F32, 64, 4, 16, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "16x4", 16, 4),
    REGV(FmtF32, P_F32, 64, 4, 32, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "16x4", 16, 4),
    REGV(FmtF32, P_F32, 256, 16, 4, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16", 16, 16),
    REGV(FmtF32, P_F32, 256, 16, 8, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16", 16, 16),
    REGV(FmtF32, P_F32, 128, 8, 8, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x8", 16, 8),
    REGV(FmtF32, P_F32, 128, 8, 16, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x8", 16, 8),
    REGV(FmtF32, P_F32, 64, 4, 16, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x4", 16, 4),
    REGV(FmtF32, P_F32, 64, 4, 32, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x4", 16, 4),
    // ... and of the plain contiguous and strided axes of small volumes
    // (128 x 128 x 64: 256-workgroup passes at the default widths) and of the
    // 512-point columns of a 1024^2 fold
    REGV(FmtF32, P_F32, 256, 16, 4, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16", 16, 16),
    REGV(FmtF32, P_F32, 256, 16, 8, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16", 16, 16),
    REGV(FmtF32, P_F32, 128, 8, 8, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x8", 16, 8),
    REGV(FmtF32, P_F32, 128, 8, 16, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x8", 16, 8),
    REGV(FmtF32, P_F32, 64, 4, 16, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x4", 16, 4),
    REGV(FmtF32, P_F32, 64, 4, 32, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x4", 16, 4),
    REGV(FmtF32, P_F32, 512, 32, 4, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "8x8x8", 8, 8, 8),
    REGV(FmtF32, P_F32, 512, 32, 8, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "8x8x8", 8, 8, 8),
    REGV(FmtF32, P_F32, 256, 16, 4, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16", 16, 16),
    REGV(FmtF32, P_F32, 256, 16, 8, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16", 16, 16),
    REGV(FmtF32, P_F32, 128, 8, 8, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x8", 16, 8),
    REGV(FmtF32, P_F32, 128, 8, 16, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x8", 16, 8),
    REGV(FmtF32, P_F32, 64, 4, 16, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x4", 16, 4),
    REGV(FmtF32, P_F32, 64, 4, 32, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x4", 16, 4),
    // two-pass splits of 2^19..2^21 points (512- and 1024-point lines, 512 to
    // 2048 of them): column passes with the twiddle on store and row passes
    // with the transposed store at 1 to 8 lines per workgroup
    REGV(FmtF32, P_F32, 1024, 64, 2, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "16x16x4", 16, 16, 4),
    REGV(FmtF32, P_F32, 1024, 64, 4, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "16x16x4", 16, 16, 4),
    REGV(FmtF32, P_F32, 1024, 64, 8, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "16x16x4", 16, 16, 4),
    REGV(FmtF32, P_F32, 512, 32, 2, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "8x8x8", 8, 8, 8),
    REGV(FmtF32, P_F32, 512, 32, 4, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "8x8x8", 8, 8, 8),
    REGV(FmtF32, P_F32, 512, 32, 8, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "8x8x8", 8, 8, 8),
    REGV(FmtF32, P_F32, 1024, 64, 1, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16x4", 16, 16, 4),
    REGV(FmtF32, P_F32, 1024, 64, 2, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16x4", 16, 16, 4),
    REGV(FmtF32, P_F32, 1024, 64, 4, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16x4", 16, 16, 4),
    REGV(FmtF32, P_F32, 512, 32, 2, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "8x8x8", 8, 8, 8),
    REGV(FmtF32, P_F32, 512, 32, 4, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "8x8x8", 8, 8, 8),
    REGV(FmtF32, P_F32, 512, 32, 8, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "8x8x8", 8, 8, 8),
    // smooth lengths of the mixed workload (schedules as wgpuFFT's fused_smooth_factors)
    REGV(FmtF32, P_F32, 120, 8, 32, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "15x8", 15, 8),
    REGV(FmtF32, P_F32, 343, 64, 4, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "7x7x7", 7, 7, 7),
    REGV(FmtF32, P_F32, 1001, 128, 2, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "13x11x7", 13, 11, 7),
    // f64 / df64 256-point lines (256^3 volumes)
    REGV(FmtF64, P_F64, 256, 16, 16, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16", 16, 16),
    REGV(FmtF64, P_F64, 256, 16, 16, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16", 16, 16),
    REGV(FmtDF64, P_DF64, 256, 16, 16, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16", 16, 16),
    REGV(FmtDF64, P_DF64, 256, 16, 16, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16", 16, 16),
    // f64 / df64 256-point passes of the three-pass split of long axes (2^24 =
    // 256^3): columns with twiddle on store; rows with line-first store
    REGV(FmtF64, P_F64, 256, 16, 16, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "16x16", 16, 16),
    REGV(FmtF64, P_F64, 256, 16, 16, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16", 16, 16),
    REGV(FmtDF64, P_DF64, 256, 16, 16, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "16x16", 16, 16),
    REGV(FmtDF64, P_DF64, 256, 16, 16, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16", 16, 16),
    // autotuner alternatives of the f64 three-pass passes: 8 and 32 lines per
    // workgroup, and 32 threads per line (8 points each)
    REGV(FmtF64, P_F64, 256, 16, 8, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "16x16", 16, 16),
    REGV(FmtF64, P_F64, 256, 16, 32, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "16x16", 16, 16),
    REGV(FmtF64, P_F64, 256, 32, 16, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "8x8x4", 8, 8, 4),
    REGV(FmtF64, P_F64, 256, 16, 8, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16", 16, 16),
    REGV(FmtF64, P_F64, 256, 16, 32, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16", 16, 16),
    REGV(FmtF64, P_F64, 256, 32, 16, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "8x8x4", 8, 8, 4),
    REGV(FmtDF64, P_DF64, 256, 16, 32, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "16x16", 16, 16),
    REGV(FmtDF64, P_DF64, 256, 16, 32, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16", 16, 16),
    // 32 lines per workgroup for the 256^3 axis passes
    REGV(FmtF64, P_F64, 256, 16, 32, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16", 16, 16),
    REGV(FmtF64, P_F64, 256, 16, 32, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16", 16, 16),
    REGV(FmtDF64, P_DF64, 256, 16, 32, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16", 16, 16),
    REGV(FmtDF64, P_DF64, 256, 16, 32, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16", 16, 16),
    // f64 / df64 4096-point lines: the two passes of the 2^24 four-step
    // (columns with twiddle on store; rows with transposed store), strided
    // and contiguous axes; component-split LDS exchange at four lines
    REGV(FmtF64, P_F64, 4096, 256, 4, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "16x16x16", 16, 16, 16),
    REGV(FmtF64, P_F64, 4096, 256, 4, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16x16", 16, 16, 16),
    REGV(FmtF64, P_F64, 4096, 256, 4, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16x16", 16, 16, 16),
    REGV(FmtF64, P_F64, 4096, 256, 1, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16x16", 16, 16, 16),
    REGV(FmtDF64, P_DF64, 4096, 256, 4, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "16x16x16", 16, 16, 16),
    REGV(FmtDF64, P_DF64, 4096, 256, 4, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16x16", 16, 16, 16),
    // Alternatives for the autotuner: other workgroup shapes and radix splits
    // of the hot lengths (the first entry of a shape above is the default).
    REGV(FmtF32, P_F32, 4096, 512, 1, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "8x8x8x8", 8, 8, 8, 8),
    REGV(FmtF32, P_F32, 4096, 256, 2, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16x16", 16, 16, 16),
    REGV(FmtF32, P_F32, 4096, 128, 1, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16x16", 16, 16, 16),
    REGV(FmtF32, P_F32, 2048, 256, 1, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "8x8x8x4", 8, 8, 8, 4),
    REGV(FmtF32, P_F32, 2048, 128, 4, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16x8", 16, 16, 8),
    REGV(FmtF32, P_F32, 1024, 128, 2, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "8x8x4x4", 8, 8, 4, 4),
    REGV(FmtF32, P_F32, 1024, 64, 8, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16x4", 16, 16, 4),
    REGV(FmtF32, P_F32, 1024, 128, 1, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "8x8x4x4", 8, 8, 4, 4),
    REGV(FmtF32, P_F32, 1024, 64, 1, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16x4", 16, 16, 4),
    REGV(FmtF32, P_F32, 1024, 64, 2, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16x4", 16, 16, 4),
    REGV(FmtF32, P_F32, 4096, 256, 4, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 2, "16x16x16", 16, 16, 16),
    REGV(FmtF32, P_F32, 2048, 128, 8, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 2, "16x16x8", 16, 16, 8),
    REGV(FmtF32, P_F32, 2048, 256, 8, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 2, "8x8x8x4", 8, 8, 8, 4),
    REGV(FmtF32, P_F32, 1024, 64, 16, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 2, "16x16x4", 16, 16, 4),
    // strided lines one per thread, fewer lines per workgroup: more
    // workgroups for short line counts (1025 packed columns of a 2048^2 R2C)
    REGV(FmtF32, P_F32, 2048, 128, 4, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16x8", 16, 16, 8),
    REGV(FmtF32, P_F32, 2048, 128, 2, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16x8", 16, 16, 8),
    REGV(FmtF32, P_F32, 2048, 256, 4, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "8x8x8x4", 8, 8, 8, 4),
    REGV(FmtF32, P_F32, 2048, 256, 2, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "8x8x8x4", 8, 8, 8, 4),
    REGV(FmtF32, P_F32, 1024, 64, 8, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16x4", 16, 16, 4),
    REGV(FmtF32, P_F32, 1024, 64, 4, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16x4", 16, 16, 4),
    // ... and more threads per line (8 points each): more waves in flight
    // for the 513 to 1025 workgroups of those columns
    REGV(FmtF32, P_F32, 1024, 128, 4, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "8x8x4x4", 8, 8, 4, 4),
    REGV(FmtF32, P_F32, 1024, 128, 2, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "8x8x4x4", 8, 8, 4, 4),
    REGV(FmtF32, P_F32, 1024, 64, 2, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16x4", 16, 16, 4),
    // strided line pairs in 256-thread workgroups (the batched kernel's shape)
    REGV(FmtF32, P_F32, 1024, 64, 8, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 2, "16x16x4", 16, 16, 4),
    REGV(FmtF32, P_F32, 512, 32, 16, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 2, "8x8x8", 8, 8, 8),
    // (256, 128 and 64 points at 32, 64 and 128 line pairs are the strided
    // defaults above)
    // persistent pipelined forms of the 4096x2048 passes
    REGP(FmtF32, P_F32, 4096, 256, 1, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, 2, "16x16x16", 16, 16, 16),
    REGP(FmtF32, P_F32, 4096, 512, 1, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, 1, "8x8x8x8", 8, 8, 8, 8),
    REGP(FmtF32, P_F32, 2048, 256, 4, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 2, 2, "8x8x8x4", 8, 8, 8, 4),
    REGP(FmtF32, P_F32, 1024, 64, 4, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, 2, "16x16x4", 16, 16, 4),
    // persistent pipelined forms of the strided column-pair passes of 2D
    // folds (1024-point columns of a 4096 x 2048 fold)
    REGP(FmtF32, P_F32, 1024, 128, 8, 1, 1, LD_COMPLEX, ST_COMPLEX, 0, 2, 2, "8x8x4x4", 8, 8, 4, 4),
    // persistent pipelined forms of the 2048^2 real passes (axis 0)
    REGP(FmtF32, P_F32, 2048, 128, 1, 0, 0, LD_REAL, ST_PACKED, 0, 1, 2, "16x16x8", 16, 16, 8),
    REGP(FmtF32, P_F32, 2048, 128, 2, 0, 0, LD_REAL, ST_PACKED, 0, 1, 2, "16x16x8", 16, 16, 8),
    REGP(FmtF32, P_F32, 2048, 128, 1, 0, 0, LD_HERMITIAN, ST_REAL, 0, 1, 2, "16x16x8", 16, 16, 8),
    REGP(FmtF32, P_F32, 2048, 128, 2, 0, 0, LD_HERMITIAN, ST_REAL, 0, 1, 2, "16x16x8", 16, 16, 8),
    // fold passes of 2D transforms (rows r + c M/F of the second axis per thread)
    REGF(FmtF32, P_F32, 4096, 256, 2, "16x16x16", 16, 16, 16),
    REGF(FmtF32, P_F32, 4096, 512, 2, "8x8x8x8", 8, 8, 8, 8),
    REGFP(FmtF32, P_F32, 4096, 512, 2, 1, "8x8x8x8", 8, 8, 8, 8),
    REGFP(FmtF32, P_F32, 4096, 512, 2, 2, "8x8x8x8", 8, 8, 8, 8),
    REGFP(FmtF32, P_F32, 4096, 256, 2, 1, "16x16x16", 16, 16, 16),
    REGF(FmtF32, P_F32, 4096, 512, 4, "8x8x8x8", 8, 8, 8, 8),
    REGF(FmtF32, P_F32, 2048, 128, 2, "16x16x8", 16, 16, 8),
    REGF(FmtF32, P_F32, 2048, 256, 2, "8x8x8x4", 8, 8, 8, 4),
    REGF(FmtF32, P_F32, 1024, 64, 2, "16x16x4", 16, 16, 4),
    REGF(FmtF32, P_F32, 1024, 128, 2, "8x8x4x4", 8, 8, 4, 4),
    // R2C fold: real rows r, r + M/2 in, packed rows out
    REGFM(FmtF32, P_F32, 2048, 128, 2, LD_REAL, ST_PACKED, "16x16x8", 16, 16, 8),
    REGFM(FmtF32, P_F32, 2048, 256, 2, LD_REAL, ST_PACKED, "8x8x8x4", 8, 8, 8, 4),
    // C2R unfold: packed rows n, n + M/2 of the halves in, real rows out
    REGFM(FmtF32, P_F32, 2048, 128, 2, LD_HERMITIAN, ST_REAL, "16x16x8", 16, 16, 8),
    REGFM(FmtF32, P_F32, 2048, 256, 2, LD_HERMITIAN, ST_REAL, "8x8x8x4", 8, 8, 8, 4),
    REGFM(FmtF32, P_F32, 1024, 64, 2, LD_REAL, ST_PACKED, "16x16x4", 16, 16, 4),
    REGFM(FmtF32, P_F32, 1024, 64, 2, LD_HERMITIAN, ST_REAL, "16x16x4", 16, 16, 4),
    // smaller tiles: more tiles per persistent workgroup, a shorter pipeline tail
    REGP(FmtF32, P_F32, 1024, 64, 2, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, 2, "16x16x4", 16, 16, 4),
    REGP(FmtF32, P_F32, 1024, 128, 2, 0, 0, LD_COMPLEX, ST_COMPLEX, 0, 1, 2, "8x8x4x4", 8, 8, 4, 4),
    // persistent pipelined form of the first pass of a 2^20-point split
    // (16384 lines of 64 points; with two tiles per workgroup the next tile's
    // loads overlap the current tile's stages and stores: 41.3 vs 43.3 us per
    // pair; the same form of the other short passes measured 1-17 % slower)
    REGP(FmtF32, P_F32, 64, 4, 16, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, 2, "16x4", 16, 4),
    REGV(FmtF64, P_F64, 4096, 256, 2, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "16x16x16", 16, 16, 16),
    REGV(FmtF64, P_F64, 4096, 256, 2, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16x16", 16, 16, 16),
    REGV(FmtF64, P_F64, 4096, 128, 4, 1, 1, LD_COMPLEX, ST_COMPLEX, 1, 1, "16x16x16", 16, 16, 16),
    REGV(FmtF64, P_F64, 4096, 128, 4, 0, 1, LD_COMPLEX, ST_COMPLEX, 0, 1, "16x16x16", 16, 16, 16),
    // adjacent-butterfly forms: 16-byte loads and stores of contiguous lines
    REGFA(4096, 256, 2, "8x8x8x8", 8, 8, 8, 8),
    REGFA(4096, 256, 2, "8x16x8x4", 8, 16, 8, 4),
    REGFA(4096, 128, 2, "8x8x8x8", 8, 8, 8, 8),
    REGFA(4096, 128, 2, "16x16x16", 16, 16, 16),
    REGFA(2048, 128, 2, "8x8x8x4", 8, 8, 8, 4),
    REGFA(1024, 64, 2, "8x8x4x4", 8, 8, 4, 4),
    REGVA(4096, 256, 1, "8x8x8x8", 8, 8, 8, 8),
    REGVA(4096, 128, 1, "16x16x16", 16, 16, 16),
    REGVA(1024, 64, 2, "8x8x4x4", 8, 8, 4, 4),
    REGVA(1024, 64, 4, "8x8x4x4", 8, 8, 4, 4),
    REGVA(1024, 32, 4, "16x16x4", 16, 16, 4),
};

int reg_variant_count() { return (int)(sizeof(kRegistry) / sizeof(kRegistry[0])); }
const RegVariant* reg_variant_at(int i) {
    return (i >= 0 && i < reg_variant_count()) ? &kRegistry[i] : nullptr;
}

std::vector<const RegVariant*> find_reg_variants(int prec, int n, int ld_lf, int st_lf, int ldm, int stm,
                                                 int post, int fold) {
    std::vector<const RegVariant*> out;
    for (const RegVariant& v : kRegistry)
        if (v.prec == prec && v.n == n && v.ld_lf == ld_lf && v.st_lf == st_lf && v.ldm == ldm && v.stm == stm &&
            v.post == post && v.fold == fold)
            out.push_back(&v);
    return out;
}

std::string reg_variant_name(const RegVariant* v) {
    static const char* prec[] = {"f32", "f64", "df64"};
    char b[160];
    snprintf(b, sizeof(b), "reg:%s:n%d:t%d:l%d:lf%d%d:m%d/%d:p%d:v%d:%s", prec[v->prec], v->n, v->tpl, v->lines,
             v->ld_lf, v->st_lf, v->ldm, v->stm, v->post, v->vec, v->radices);
    std::string name = b;
    if (v->persist) name += ":pipe" + std::to_string(v->persist);
    if (v->fold) name += ":fold";
    if (v->adj) name += ":adj";
    return name;
}

const RegVariant* find_reg_variant_named(const std::string& name) {
    for (const RegVariant& v : kRegistry)
        if (reg_variant_name(&v) == name) return &v;
    return nullptr;
}

const RegVariant* find_reg_variant(int prec, int n, int ld_lf, int st_lf, int ldm, int stm,
                                   int post, int vec, int fold) {
    for (const RegVariant& v : kRegistry) {
        if (v.prec == prec && v.n == n && v.ld_lf == ld_lf && v.st_lf == st_lf && v.ldm == ldm &&
            v.stm == stm && v.post == post && v.vec == vec && v.fold == fold && !v.adj)
            return &v;
    }
    return nullptr;
}

uint32_t device_cu_count() {
    static uint32_t cus = 0;
    if (!cus) {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
            cus = (uint32_t)n;
        else
            cus = 256;
    }
    return cus;
}

// Workgroups of a persistent variant resident on one CU at once (registers,
// LDS, waves): its grid holds no more, since a workgroup beyond them would
// start only when one of them has finished all of its tiles.
static uint32_t resident_per_cu(const RegVariant* v) {
    static std::mutex mu;
    static std::unordered_map<const void*, uint32_t> done;
    std::lock_guard<std::mutex> g(mu);
    auto it = done.find(v->fn);
    if (it != done.end()) return it->second;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, v->fn, v->nthreads, v->lds_bytes) != hipSuccess || n < 1)
        n = 1;
    done[v->fn] = (uint32_t)n;
    return (uint32_t)n;
}

hipError_t launch_reg(const RegVariant* v, const LineArgs& a, hipStream_t s) {
    hipError_t e = ensure_lds_attr(v->fn, v->lds_bytes);
    if (e != hipSuccess) return e;
    uint32_t blocks = (a.lines + v->lines - 1) / v->lines;
    if (v->persist) {
        // at most the workgroups resident at once, and at least two tiles
        // per workgroup where that still leaves one workgroup per CU (the
        // pipeline needs a next tile to load)
        const uint32_t tiles = blocks, cus = device_cu_count();
        blocks = std::min<uint32_t>(tiles, cus * resident_per_cu(v));
        if (tiles < 2 * blocks) blocks = std::max<uint32_t>(std::min<uint32_t>(tiles, cus), (tiles + 1) / 2);
    }
    LineArgs copy = a;
    void* args[] = {&copy};
    return hipLaunchKernel(v->fn, dim3(blocks), dim3(v->nthreads), args, v->lds_bytes, s);
}

size_t lds_kernel_bytes(int prec, uint32_t n, uint32_t lpw) {
    const size_t es = prec == P_F32 ? 8 : 16;
    return 2 * (size_t)lpw * lds_line_stride(n) * es;
}

template <class Fmt>
static hipError_t launch_lds_fmt(int ldm, int stm, int post, const LineArgs& a, const Sched& sc,
                                 hipStream_t s) {
    const size_t bytes = lds_kernel_bytes(Fmt::kId == 0 ? P_F32 : P_F64, sc.n, sc.lpw);
    // one workgroup per CU (a tile over half the LDS): twice the threads
    const bool wide = bytes > kLdsBudget / 2;
    const void* fn = nullptr;
    if (ldm == LD_COMPLEX && stm == ST_COMPLEX && !post)
        fn = wide ? (const void*)&lds_fft_kernel<Fmt, LD_COMPLEX, ST_COMPLEX, false, 512>
                  : (const void*)&lds_fft_kernel<Fmt, LD_COMPLEX, ST_COMPLEX, false>;
    else if (ldm == LD_COMPLEX && stm == ST_COMPLEX && post)
        fn = wide ? (const void*)&lds_fft_kernel<Fmt, LD_COMPLEX, ST_COMPLEX, true, 512>
                  : (const void*)&lds_fft_kernel<Fmt, LD_COMPLEX, ST_COMPLEX, true>;
    else if (ldm == LD_REAL && stm == ST_PACKED && !post)
        fn = wide ? (const void*)&lds_fft_kernel<Fmt, LD_REAL, ST_PACKED, false, 512>
                  : (const void*)&lds_fft_kernel<Fmt, LD_REAL, ST_PACKED, false>;
    else if (ldm == LD_HERMITIAN && stm == ST_REAL && !post)
        fn = wide ? (const void*)&lds_fft_kernel<Fmt, LD_HERMITIAN, ST_REAL, false, 512>
                  : (const void*)&lds_fft_kernel<Fmt, LD_HERMITIAN, ST_REAL, false>;
    else
        return hipErrorInvalidValue;
    hipError_t e = ensure_lds_attr(fn, bytes);
    if (e != hipSuccess) return e;
    const uint32_t blocks = (a.lines + sc.lpw - 1) / sc.lpw;
    LineArgs ca = a;
    Sched cs = sc;
    void* args[] = {&ca, &cs};
    return hipLaunchKernel(fn, dim3(blocks), dim3(wide ? 512 : 256), args, bytes, s);
}

hipError_t launch_lds(int prec, int ldm, int stm, int post, const LineArgs& a, const Sched& sc,
                      hipStream_t s) {
    switch (prec) {
        case P_F32: return launch_lds_fmt<FmtF32>(ldm, stm, post, a, sc, s);
        case P_F64: return launch_lds_fmt<FmtF64>(ldm, stm, post, a, sc, s);
        default: return launch_lds_fmt<FmtDF64>(ldm, stm, post, a, sc, s);
    }
}

}  // namespace mifft
