// Host-side view of the kernel library: the compiled kernel-variant registry
// (the counterpart of the reference's typed pipeline-cache keys,
// src/runtime/pipeline_cache.rs:2048-2260) and the launch wrappers.
#pragma once
#include <string>
#include <vector>
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "fft_kernels.hpp"

namespace mifft {

enum Prec { P_F32 = 0, P_F64 = 1, P_DF64 = 2 };

// One compiled register-resident line kernel.
struct RegVariant {
    int prec;
    int n;
    int tpl;
    int lines;
    int ld_lf, st_lf;
    int ldm, stm;
    int post;
    int vec;  // adjacent lines per thread (2: 16-byte pair accesses)
    int nthreads;
    size_t lds_bytes;
    const void* fn;
    const char* radices;  // e.g. "16x16x16"
    // 0: one tile per workgroup.  k > 0: persistent pipelined form, each
    // workgroup loading its next tile before transforming the current one;
    // k is its launch bound's minimum of waves per SIMD (the register
    // budget), and its grid holds as many workgroups per CU as are resident
    // at once (hipOccupancyMaxActiveBlocksPerMultiprocessor).
    int persist = 0;
    // F > 0: fold variant (F contiguous lines M/F rows apart per thread, the
    // next axis's first radix-F stage on store); looked up only by fold passes.
    int fold = 0;
    // 1: stage 0 and the last stage take adjacent butterflies, so global
    // loads and stores move 16 bytes (two points) per lane.
    int adj = 0;
};

// Compute units of the current device (256 when it cannot be queried).
uint32_t device_cu_count();

// Registry lookup; nullptr when no compiled variant matches.
const RegVariant* find_reg_variant(int prec, int n, int ld_lf, int st_lf, int ldm, int stm,
                                   int post, int vec, int fold = 0);
// Every registered variant of a line shape (any lines-per-thread width), the
// default first; stable names for the tuning table.
std::vector<const RegVariant*> find_reg_variants(int prec, int n, int ld_lf, int st_lf, int ldm, int stm,
                                                 int post, int fold = 0);
std::string reg_variant_name(const RegVariant* v);
const RegVariant* find_reg_variant_named(const std::string& name);
int reg_variant_count();
const RegVariant* reg_variant_at(int i);

hipError_t launch_reg(const RegVariant* v, const LineArgs& a, hipStream_t s);

// Runtime-schedule LDS kernel (any {2..16}-smooth schedule whose lines fit LDS).
hipError_t launch_lds(int prec, int ldm, int stm, int post, const LineArgs& a, const Sched& sc,
                      hipStream_t s);
size_t lds_kernel_bytes(int prec, uint32_t n, uint32_t lpw);
constexpr size_t kLdsBudget = 160 * 1024;

// Fused Rader / Bluestein convolution kernels (one line per workgroup pass).
struct ConvArgs {
    LineArgs io;              // the prime / Bluestein line layout (n points)
    Sched sc;                 // schedule of the convolution length M
    const void* filt;         // FFT_M of the filter, scaled by 1/M (arith type)
    const void* chirp;        // Bluestein chirp c[n] (arith type), or nullptr
    const uint32_t* perm_in;  // Rader gather g^q mod p, q < p-1
    const uint32_t* perm_out; // Rader scatter g^-u mod p, u < p-1
    uint32_t n;               // transform length
    uint32_t m;               // convolution length
    DivMagic dn;              // divide by n (Rader)
};
hipError_t launch_rader(int prec, const ConvArgs& c, hipStream_t s);
hipError_t launch_bluestein(int prec, const ConvArgs& c, hipStream_t s);
// Register-resident f32 Bluestein kernels (reg_conv.hpp, kern_regconv.hip),
// one per power-of-two convolution length M: `lines` lines of `tpl` threads
// per workgroup, radix schedule `radices` (stage twiddles in that format).
struct RegConvVariant {
    int m, tpl, lines, nthreads;
    size_t lds_bytes;
    const void* fn;
    const char* radices;
};
const RegConvVariant* find_reg_bluestein(uint32_t m);
hipError_t launch_reg_bluestein(const RegConvVariant* v, const ConvArgs& c, hipStream_t s);
// Direct f32 DFT of prime lines of length p (kern_conv.hip), lpw lines per
// workgroup, lines-fastest slots when lf; wtab = W_p^t, t < p.
size_t direct_prime_bytes(uint32_t p, uint32_t lpw);
hipError_t launch_direct_prime(const LineArgs& a, uint32_t p, uint32_t lpw, bool lf, const void* wtab,
                               hipStream_t s);
size_t conv_kernel_bytes(int prec, uint32_t m, uint32_t lpw);

// Table generation (FP64 on the device, rounded once to the table type).
hipError_t gen_twiddles(int prec, void* out, uint64_t n, uint64_t count, hipStream_t s);
// Stage-twiddle table of a radix schedule (entries: stage_twiddle_count).
uint64_t stage_twiddle_count(const int* rad, int ns);
hipError_t gen_stage_twiddles(int prec, void* out, const int* rad, int ns, hipStream_t s);
hipError_t gen_twiddles_f64_strided(double2* out, uint64_t n, uint64_t count, uint64_t mult,
                                    hipStream_t s);
hipError_t gen_chirp(int prec, void* out, uint64_t n, hipStream_t s);
hipError_t gen_rader_seq(double2* out, uint64_t p, uint64_t m, uint64_t ginv, hipStream_t s);
hipError_t gen_bluestein_seq(double2* out, uint64_t n, uint64_t m, hipStream_t s);
hipError_t round_table(int prec, void* out, const double2* in, uint64_t count, double mul,
                       hipStream_t s);

// Elementwise helpers of the multi-pass Bluestein route and the direct DFT.
hipError_t launch_chirp_pre(int prec, const LineArgs& a, const void* chirp, uint32_t n,
                            uint32_t m, hipStream_t s);
hipError_t launch_pointwise(int prec, void* buf, const void* filt, uint64_t m, uint64_t lines,
                            hipStream_t s);
hipError_t launch_chirp_post(int prec, const LineArgs& a, const void* chirp, uint32_t n,
                             uint32_t m, hipStream_t s);
// General real transforms (f32): real -> complex widening, packing of bins
// k <= n0/2 with a scale, Hermitian extension of packed lines, real part with a
// scale.  Lines are contiguous.
hipError_t launch_real_widen(const void* in, void* out, uint64_t count, hipStream_t s);
// One Stockham radix stage over HBM (kern_util.hip): radix R, L = product of
// the earlier radices, twiddles at twoff in the stage-twiddle table a.tw; lf:
// adjacent threads take adjacent lines (strided axes).
struct StockhamStage {
    uint32_t n, R, L, twoff, lf;
    DivMagic dl;  // divide by L
};
hipError_t launch_stockham_pass(int prec, const LineArgs& a, const StockhamStage& st, hipStream_t s);

// Small-volume kernel (kern_small.hip): all axes of an f32 volume of
// elems <= 4096 points per workgroup, axis a of length sc[a].n at stride S[a].
constexpr int kSmallVolMaxRank = 8;  // MIFFT_MAX_RANK
struct SmallVolArgs {
    const void* in;
    void* out;
    uint32_t elems, buf_elems, rank;
    uint32_t S[kSmallVolMaxRank];
    DivMagic dS[kSmallVolMaxRank];
    DivMagic dL[kSmallVolMaxRank];  // divide by elems / n of the axis (its line count)
    const void* tw;                     // every axis's stage twiddles, back to back
    uint32_t tw_len[kSmallVolMaxRank];  // stage-twiddle entries of each axis
    uint32_t tw_total;                  // their sum (copied to LDS at start)
    const Sched* sc;  // device array of rank schedules
    float scale;
    uint32_t conj_in, conj_out;
    uint32_t direct;  // some axis runs a direct DFT
};
size_t small_volume_lds_bytes(const SmallVolArgs& v);
hipError_t launch_small_volume(const SmallVolArgs& v, uint32_t volumes, hipStream_t s);
// dst[d_off + b d_sb + i d_se] = src[s_off + b s_sb + i s_se] for i < per_batch,
// b < batch, in elements of elem_bytes (4, 8 or 16).
hipError_t launch_strided_copy(uint32_t elem_bytes, const void* src, void* dst, uint64_t per_batch, uint64_t batch,
                               uint64_t s_off, uint64_t s_se, uint64_t s_sb, uint64_t d_off, uint64_t d_se,
                               uint64_t d_sb, hipStream_t s);
hipError_t launch_pack(const void* full, void* packed, uint32_t n0, uint64_t lines, double scale, hipStream_t s);
hipError_t launch_herm_extend(const void* packed, void* full, uint32_t n0, uint64_t lines, hipStream_t s);
hipError_t launch_real_part(const void* full, void* out, uint64_t count, double scale, hipStream_t s);
// Half-length real transforms (N0 = 2h): the bins of h-point complex lines ->
// packed N0-point spectra, and the packed spectra -> h-point complex lines.
hipError_t launch_r2c_post(const void* z, void* x, uint32_t h, uint64_t lines, double scale, const double2* hi,
                           const double2* lo, uint32_t sh, uint32_t mask, hipStream_t s);
hipError_t launch_c2r_pre(const void* x, void* z, uint32_t h, uint64_t lines, const double2* hi, const double2* lo,
                          uint32_t sh, uint32_t mask, hipStream_t s);
hipError_t launch_scale_copy(int prec, const void* in, void* out, uint64_t count, double scale,
                             hipStream_t s);

// Synthetic data: SplitMix64 counter-based uniform[-1,1) f32 and N(0,1) f32.
hipError_t fill_uniform_f32(float* out, uint64_t n, uint64_t seed, hipStream_t s);
hipError_t fill_normal_f32(float* out, uint64_t n, uint64_t seed, hipStream_t s);

DivMagic make_magic(uint32_t d);

// Raises a kernel's dynamic-LDS limit to `bytes` when above the 64 KiB default.
hipError_t ensure_lds_attr(const void* fn, size_t bytes);

}  // namespace mifft
