// Line-FFT kernels for gfx950.
//
// Every axis pass of a plan transforms `lines` independent lines of length N.
// Line g is split as i = g % S1, r = g / S1, j = r % S2, o = r / S2 and its
// element n lives at   in  + i*in_si  + j*in_sj  + o*in_so  + n*in_sn
// and its output k at  out + i*out_si + j*out_sj + o*out_so + k*out_sk.
// A contiguous axis is S1 = 1 (lines N apart); a strided axis of an ND array is
// S1 = stride (adjacent lines adjacent in memory); the two passes of a
// four-step split use S2 for the split index.  This replaces the reference's
// line_base_for_axis (src/runtime/nd_wgsl.rs:1-55, src/math/mod.rs:506-516).
// Element offsets are computed in 32 bits: every buffer a kernel indexes holds
// fewer than 2^32 elements (the per-plan u32 element limit of the reference,
// enforced by the planners).
//
// Inverse transforms are conj(DFT(conj(x))): conj_in / conj_out flip the sign of
// the imaginary part on load / store.
//
// Algorithm: Stockham autosort, radix R_s at stage s with L_s = R_0...R_{s-1}:
// butterfly j (< N/R) reads points j + m*N/R, multiplies point m by
// W_{L R}^{(j mod L) m}, runs a radix-R DFT and writes point m to
// (j - j mod L) R + (j mod L) + m L.  Stage 0 reads HBM, the last stage writes
// HBM, and the stages in between exchange through LDS.
#pragma once
#include "fft_math.hpp"

namespace mifft {

enum LoadMode { LD_COMPLEX = 0, LD_REAL = 1, LD_HERMITIAN = 2 };
enum StoreMode { ST_COMPLEX = 0, ST_PACKED = 1, ST_REAL = 2 };

struct LineArgs {
    const void* in;
    void* out;
    uint32_t lines;
    uint32_t S1, S2;
    DivMagic d1, d2;
    int64_t in_si, in_sj, in_so, in_sn;
    int64_t out_si, out_sj, out_so, out_sk;
    const void* tw;           // W_N^t, t < N, arithmetic type
    const double2* post_hi;   // four-step twiddle W_P^t = hi[t >> post_shift] * lo[t & post_mask]
    const double2* post_lo;
    uint32_t post_shift, post_mask;
    double scale;
    uint32_t conj_in, conj_out;
    // Fold passes (F lines per thread): line c of a fold group sits c times
    // these element offsets after the first, which is at line_base(g / F).
    uint32_t fold_in, fold_out;
};

__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n) {
    // Blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous
    // run of logical tiles so neighbouring tiles share its L2.  XCD x holds
    // blocks x, x + 8, ...: q + (x < r) of them for n = 8q + r.
    const uint32_t x = b & 7u, q = n >> 3, r = n & 7u;
    return x * q + (x < r ? x : r) + (b >> 3);
}

struct LineBase {
    uint32_t in, out;  // element offsets of the line's first point
    uint32_t j;
};
__device__ __forceinline__ LineBase line_base(const LineArgs& a, uint32_t g) {
    uint32_t r = fdiv(g, a.d1);
    uint32_t i = g - r * a.S1;
    uint32_t o = fdiv(r, a.d2);
    uint32_t j = r - o * a.S2;
    LineBase b;
    b.in = i * (uint32_t)a.in_si + j * (uint32_t)a.in_sj + o * (uint32_t)a.in_so;
    b.out = i * (uint32_t)a.out_si + j * (uint32_t)a.out_sj + o * (uint32_t)a.out_so;
    b.j = j;
    return b;
}

// W_P^t as the FP64 product of two FP64 table entries.
__device__ __forceinline__ double2 post_f64(const LineArgs& a, uint64_t t) {
    const double2 h = a.post_hi[t >> a.post_shift];
    const double2 l = a.post_lo[t & a.post_mask];
    return make_double2(h.x * l.x - h.y * l.y, h.x * l.y + h.y * l.x);
}

// W_P^t rounded once from that product.
template <class T>
__device__ __forceinline__ cpx<T> post_twiddle(const LineArgs& a, uint64_t t) {
    const double2 w = post_f64(a, t);
    return mk<T>((T)w.x, (T)w.y);
}

// Four-step twiddles W_P^{jline (p0 + m L)} of one butterfly's outputs m = 0,
// 1, ...: two table products (for p0 and for the step L) and an FP64
// recurrence, so that a radix-R butterfly reads 4 table entries instead of 2R.
// The R - 1 products add about R ulp of FP64 rounding (2e-15 for R = 16).
struct PostSeq {
    double2 w, step;
    __device__ __forceinline__ void init(const LineArgs& a, uint32_t jline, uint32_t p0, uint32_t L) {
        w = post_f64(a, (uint64_t)jline * p0);
        step = post_f64(a, (uint64_t)jline * L);
    }
    template <class T>
    __device__ __forceinline__ cpx<T> next() {
        const cpx<T> r = mk<T>((T)w.x, (T)w.y);
        w = make_double2(w.x * step.x - w.y * step.y, w.x * step.y + w.y * step.x);
        return r;
    }
};

// Complex or real point at element index idx (not the Hermitian mode); NT
// selects the non-temporal load.
template <class Fmt, int LDM, bool NT = true>
__device__ __forceinline__ cpx<typename Fmt::T> load_point_at(const LineArgs& a, uint32_t idx) {
    using T = typename Fmt::T;
    static_assert(LDM != LD_HERMITIAN, "the Hermitian load needs the point index");
    if constexpr (LDM == LD_REAL) return mk<T>(Fmt::ld_real(a.in, idx), T(0));
    else if constexpr (NT) return conj_by(Fmt::ld(a.in, idx), sign_mask(a.conj_in));
    else return conj_by(Fmt::ld_c(a.in, idx), sign_mask(a.conj_in));
}

// Point p of a line starting at element offset `base`.
template <class Fmt, int LDM>
__device__ __forceinline__ cpx<typename Fmt::T> load_point(const LineArgs& a, uint32_t base,
                                                           uint32_t p, uint32_t n) {
    using T = typename Fmt::T;
    cpx<T> v;
    if constexpr (LDM == LD_COMPLEX) {
        v = Fmt::ld(a.in, base + p * (uint32_t)a.in_sn);
    } else if constexpr (LDM == LD_REAL) {
        v = mk<T>(Fmt::ld_real(a.in, base + p * (uint32_t)a.in_sn), T(0));
    } else {
        // Hermitian extension of a packed half line (n/2 + 1 points): the
        // imaginary part of the self-conjugate bins 0 and n/2 is dropped and the
        // upper half mirrors the conjugate (src/math/mod.rs:470-502).
        // Branch-free so that consecutive loads stay in flight together.
        const uint32_t h = n >> 1;
        const bool up = p > h;
        const uint32_t idx = up ? n - p : p;
        v = Fmt::ld(a.in, base + idx * (uint32_t)a.in_sn);
        v.y = (p == 0 || 2 * p == n) ? T(0) : (up ? -v.y : v.y);
    }
    return conj_by(v, sign_mask(a.conj_in));
}

// Scale of the stored real and imaginary parts (the imaginary one carries the
// output conjugation).
template <class T>
struct OutScale {
    T re, im;
};
template <class T>
__device__ __forceinline__ OutScale<T> out_scale(const LineArgs& a) {
    const T s = (T)a.scale;
    return OutScale<T>{s, a.conj_out ? -s : s};
}

// Output point k (of line index jline along the split axis, for the
// four-step twiddle) at element index idx.
template <class Fmt, int STM, bool POST>
__device__ __forceinline__ void store_point_at(const LineArgs& a, uint32_t idx, uint32_t jline, uint32_t k,
                                               uint32_t n, cpx<typename Fmt::T> v) {
    using T = typename Fmt::T;
    if constexpr (POST) v = cmul(v, post_twiddle<T>(a, (uint64_t)jline * k));
    const OutScale<T> s = out_scale<T>(a);
    if constexpr (STM == ST_COMPLEX) {
        Fmt::st(a.out, idx, mk<T>(v.x * s.re, v.y * s.im));
    } else if constexpr (STM == ST_PACKED) {
        if (k <= (n >> 1)) Fmt::st(a.out, idx, mk<T>(v.x * s.re, v.y * s.im));
    } else {
        Fmt::st_real(a.out, idx, v.x * s.re);
    }
}

// Output point k of a line whose output starts at element offset `base`.
template <class Fmt, int STM, bool POST>
__device__ __forceinline__ void store_point(const LineArgs& a, uint32_t base, uint32_t jline,
                                            uint32_t k, uint32_t n, cpx<typename Fmt::T> v) {
    store_point_at<Fmt, STM, POST>(a, base + k * (uint32_t)a.out_sk, jline, k, n, v);
}

// ------------------------------------------------- register-resident kernel
template <int... Rs>
struct RadixList {
    static constexpr int n = sizeof...(Rs);
    static constexpr int r[n] = {Rs...};
    static constexpr int at(int s) { return r[s]; }
    static constexpr int lprod(int s) {
        int p = 1;
        for (int i = 0; i < s; i++) p *= r[i];
        return p;
    }
    static constexpr int prod() { return lprod(n); }
    // Offset of stage s in the stage-twiddle table (see stage_twiddle_index).
    static constexpr int twoff(int s) {
        int o = 0;
        for (int i = 1; i < s; i++) o += (r[i] - 1) * lprod(i);
        return o;
    }
};

// Stage-twiddle table of a radix schedule R_0..R_{S-1}: for s >= 1 a block of
// (R_s - 1) * L_s entries, entry (m-1)*L_s + k = W_{L_s R_s}^(k m).  Adjacent
// butterflies (adjacent k) read adjacent entries, so a wavefront's twiddle loads
// are contiguous.  N - R_0 entries in all.

constexpr int ceil_div(int a, int b) { return (a + b - 1) / b; }
constexpr int lds_pad(int p) { return p + (p >> 4); }

// Two adjacent lines (line g even and g+1, adjacent in memory) in one 16-byte
// access per point; the host guarantees even strides and 16-byte alignment.
// Element index i of lines g and g+1 (adjacent in memory).
template <class Fmt, bool NT = true>
__device__ __forceinline__ void load_pair_at(const LineArgs& a, uint32_t i, cpx<typename Fmt::T>& v0,
                                             cpx<typename Fmt::T>& v1) {
    if constexpr (Fmt::kId == 0) {
        const vf4_t* q4 = reinterpret_cast<const vf4_t*>(reinterpret_cast<const float2*>(a.in) + i);
        vf4_t q;
        if constexpr (NT) q = __builtin_nontemporal_load(q4);
        else q = *q4;
        v0 = mk<float>(q.x, q.y);
        v1 = mk<float>(q.z, q.w);
    } else if constexpr (NT) {
        v0 = Fmt::ld(a.in, i);
        v1 = Fmt::ld(a.in, i + 1);
    } else {
        v0 = Fmt::ld_c(a.in, i);
        v1 = Fmt::ld_c(a.in, i + 1);
    }
    const uint32_t m = sign_mask(a.conj_in);
    v0 = conj_by(v0, m);
    v1 = conj_by(v1, m);
}

// Output point k of lines g and g+1 at element index i.
template <class Fmt, bool POST>
__device__ __forceinline__ void store_pair_at(const LineArgs& a, uint32_t i, uint32_t jline, uint32_t k,
                                              cpx<typename Fmt::T> v0, cpx<typename Fmt::T> v1) {
    using T = typename Fmt::T;
    if constexpr (POST) {
        // the two lines of a pair share j (adjacent i)
        const cpx<T> w = post_twiddle<T>(a, (uint64_t)jline * k);
        v0 = cmul(v0, w);
        v1 = cmul(v1, w);
    }
    const OutScale<T> s = out_scale<T>(a);
    v0 = mk<T>(v0.x * s.re, v0.y * s.im);
    v1 = mk<T>(v1.x * s.re, v1.y * s.im);
    if constexpr (Fmt::kId == 0) {
        *reinterpret_cast<float4*>(reinterpret_cast<float2*>(a.out) + i) = make_float4(v0.x, v0.y, v1.x, v1.y);
    } else {
        Fmt::st(a.out, i, v0);
        Fmt::st(a.out, i + 1, v1);
    }
}

// Template parameters:
//   N, TPL (threads per line group), LINES (lines per workgroup);
//   LD_LF / ST_LF: stage-0 loads / last-stage stores map adjacent lanes to
//   adjacent lines (strided axes) instead of adjacent points (contiguous axes);
//   LDM / STM: load and store modes; POST: four-step twiddle on store;
//   VEC: adjacent lines per thread (2: one 16-byte access covers both lines;
//   the LDS exchange then runs in VEC rounds through one slot per line group).
//   PF: load every stage's twiddles into registers together with the data, so
//   their latency overlaps the data loads instead of following each exchange.
//   FOLD (contiguous lines, F = VEC of 2 or 4): the F lines of a thread are
//   rows r + c Q (c < F) of the next axis (M = F Q points, fold_in /
//   fold_out apart), and the store applies that axis's first radix-F stage,
//   so that its pass runs Q-point lines (1/F of the tile of an M-point line).
//   ADJ (contiguous complex lines): in stage 0 and the last stage a thread
//   takes adjacent butterflies j = 2t + q instead of j = t + q TPL, so its
//   points j and j + 1 are adjacent in memory and each global load and store
//   moves 16 bytes per lane (two points) instead of 8.
template <class Fmt, int N, int TPL, int LINES, bool LD_LF, bool ST_LF, int LDM, int STM, bool POST,
          class RL, int VEC = 1, bool PF = false, bool FOLD = false, bool ADJ = false>
struct RegFft {
    using T = typename Fmt::T;
    static constexpr int NS = RL::n;
    // Twiddle registers of stage s: NB_s butterflies x (R_s - 1).
    static constexpr int tw_count(int s) {
        return s >= NS ? 0 : (s == 0 ? 0 : ceil_div(N / RL::at(s), TPL) * (RL::at(s) - 1)) + tw_count(s + 1);
    }
    static constexpr int tw_base(int s) {
        int b = 0;
        for (int i = 1; i < s; i++) b += ceil_div(N / RL::at(i), TPL) * (RL::at(i) - 1);
        return b;
    }
    static constexpr int TWREG = PF ? (tw_count(0) > 0 ? tw_count(0) : 1) : 1;
    static constexpr int kLines = LINES;
    static constexpr int GROUPS = LINES / VEC;
    static constexpr int NT = TPL * GROUPS;
    // LDS elements between the line groups' exchange slots.  Where a stage
    // maps adjacent lanes to adjacent line groups (LD_LF / ST_LF), a wave's
    // lanes span GROUPS slots at once, and the slot stride mod 32 points
    // decides how their lanes share banks: 8 for two groups, 12 for four, 6
    // for eight spread the slots' lanes evenly over 32 or 64 four-byte banks
    // (counted lane by lane for the stage writes and reads of each variant).
    static constexpr int lds_stride() {
        constexpr int b = lds_pad(N);
        constexpr int g = LINES / VEC;
        constexpr int r = g == 2 ? 8 : g == 4 ? 12 : g == 8 ? 6 : -1;
        if ((LD_LF || ST_LF) && r >= 0 && sizeof(cpx<T>) == 8) return b + ((r - b % 32) + 32) % 32;
        return b + 1;
    }
    static constexpr int LSTR = lds_stride();
    static_assert(VEC == 1 || FOLD || (LDM == LD_COMPLEX && STM == ST_COMPLEX), "pairs are complex only");
    // Unfold (FOLD with the Hermitian load, C2R): the two rows are n and
    // n + H of the halves E_0, E_1 left by the previous axis's half-length
    // inverse transforms; the load forms E_0 + W_M^{-n} E_1 and
    // E_0 - W_M^{-n} E_1 per packed bin (that axis's last radix-2
    // decimation-in-time stage) before the Hermitian extension, and the store
    // writes the two rows as they are.
    static constexpr bool UNFOLD = FOLD && LDM == LD_HERMITIAN;
    static_assert(!FOLD || ((VEC == 2 || VEC == 4) && !LD_LF && !ST_LF && !POST), "fold rows are contiguous lines");
    static_assert(!FOLD || (UNFOLD ? (VEC == 2 && STM == ST_REAL) : STM != ST_REAL),
                  "fold: complex or real in, complex or packed out; unfold: packed in, real out");
    // Lines of 16-byte points whose exchange buffer would exceed 80 KiB go
    // through LDS one component at a time: half the space, twice the rounds
    // (four 4096-point f64 lines per workgroup instead of two).
    static constexpr bool SPLIT = sizeof(cpx<T>) == 16 && (size_t)GROUPS * LSTR * sizeof(cpx<T>) > 80 * 1024;
    static constexpr size_t lds_bytes() {
        return NS > 1 ? (size_t)GROUPS * LSTR * (SPLIT ? sizeof(T) : sizeof(cpx<T>)) : 0;
    }
    // Non-temporal loads only where a load instruction covers whole
    // 128-byte row segments: contiguous lines, or line groups that wide.
    static constexpr bool NTL = !LD_LF || (size_t)GROUPS * VEC * sizeof(cpx<T>) >= 128;
    static constexpr int max_regs(int s) {
        return s >= NS ? 0
                       : (ceil_div(N / RL::at(s), TPL) * RL::at(s) > max_regs(s + 1)
                              ? ceil_div(N / RL::at(s), TPL) * RL::at(s)
                              : max_regs(s + 1));
    }
    static constexpr int EMAX = max_regs(0);
    static_assert(!ADJ || (!LD_LF && !ST_LF && LDM == LD_COMPLEX && STM == ST_COMPLEX && !POST &&
                           (VEC == 1 || FOLD) && sizeof(cpx<T>) == 8),
                  "adjacent butterflies: contiguous f32 complex lines");
    // Stages whose butterflies are taken in adjacent pairs (ADJ): the HBM ones.
    template <int s>
    static constexpr bool adj_stage() {
        return ADJ && (s == 0 || s == NS - 1);
    }
    // Butterfly q of thread t in stage s.
    template <int s>
    static __device__ __forceinline__ uint32_t jof(uint32_t t, int q) {
        constexpr int NB = ceil_div(N / RL::at(s), TPL);
        if constexpr (adj_stage<s>()) {
            static_assert((N / RL::at(s)) % TPL == 0 && NB % 2 == 0, "adjacent butterflies need pairs per thread");
            return t * NB + q;
        } else {
            return t + q * TPL;
        }
    }

    struct Map {
        uint32_t lg, t;  // line group within the workgroup, thread within the group
    };
    // Stage 0 follows the load layout, the last stage the store layout; the
    // stages in between exchange through LDS, so their mapping is free.
    template <int s>
    static __device__ __forceinline__ Map map_of() {
        constexpr bool LF = (s == 0) ? LD_LF : (s == NS - 1 ? ST_LF : false);
        const uint32_t tid = threadIdx.x;
        if constexpr (LF) return Map{tid % GROUPS, tid / GROUPS};
        else return Map{tid / TPL, tid % TPL};
    }

    // Outputs of stage s -> LDS -> inputs of stage s+1, one round per line of
    // the group (VEC rounds through GROUPS slots).
    template <int s>
    static __device__ __forceinline__ void exchange(cpx<T>* lds, cpx<T> (&w)[VEC][EMAX]) {
        constexpr int R = RL::at(s), L = RL::lprod(s), NR = N / R, NB = ceil_div(NR, TPL);
        constexpr int R2 = RL::at(s + 1), NR2 = N / R2, NB2 = ceil_div(NR2, TPL);
        const Map mw = map_of<s>();
        const Map mr = map_of<s + 1>();
#pragma unroll
        for (int v = 0; v < VEC; v++) {
            if constexpr (!SPLIT) {
                cpx<T>* ln = lds + mw.lg * LSTR;
#pragma unroll
                for (int q = 0; q < NB; q++) {
                    const uint32_t j = jof<s>(mw.t, q);
                    if (NR % TPL == 0 || j < (uint32_t)NR) {
                        const uint32_t k = j % L;
#pragma unroll
                        for (int m = 0; m < R; m++) ln[lds_pad((j - k) * R + k + m * L)] = w[v][q * R + m];
                    }
                }
                __syncthreads();
                const cpx<T>* lr = lds + mr.lg * LSTR;
#pragma unroll
                for (int q = 0; q < NB2; q++) {
                    const uint32_t j = jof<s + 1>(mr.t, q);
                    if (NR2 % TPL == 0 || j < (uint32_t)NR2) {
#pragma unroll
                        for (int m = 0; m < R2; m++) w[v][q * R2 + m] = lr[lds_pad(j + m * NR2)];
                    }
                }
                __syncthreads();
            } else {
                // Round c moves component c only; a slot's other component still
                // holds its stage-s value when its own round comes.
                T* lt = reinterpret_cast<T*>(lds);
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    T* ln = lt + mw.lg * LSTR;
#pragma unroll
                    for (int q = 0; q < NB; q++) {
                        const uint32_t j = jof<s>(mw.t, q);
                        if (NR % TPL == 0 || j < (uint32_t)NR) {
                            const uint32_t k = j % L;
#pragma unroll
                            for (int m = 0; m < R; m++)
                                ln[lds_pad((j - k) * R + k + m * L)] = c ? w[v][q * R + m].y : w[v][q * R + m].x;
                        }
                    }
                    __syncthreads();
                    const T* lr = lt + mr.lg * LSTR;
#pragma unroll
                    for (int q = 0; q < NB2; q++) {
                        const uint32_t j = jof<s + 1>(mr.t, q);
                        if (NR2 % TPL == 0 || j < (uint32_t)NR2) {
#pragma unroll
                            for (int m = 0; m < R2; m++) {
                                if (c) w[v][q * R2 + m].y = lr[lds_pad(j + m * NR2)];
                                else w[v][q * R2 + m].x = lr[lds_pad(j + m * NR2)];
                            }
                        }
                    }
                    __syncthreads();
                }
            }
        }
    }

    template <int s>
    static __device__ __forceinline__ void prefetch(const LineArgs& a, cpx<T> (&twr)[TWREG]) {
        if constexpr (s < NS) {
            constexpr int R = RL::at(s), L = RL::lprod(s), NR = N / R, NB = ceil_div(NR, TPL);
            constexpr uint32_t off = RL::twoff(s);
            const Map mp = map_of<s>();
#pragma unroll
            for (int q = 0; q < NB; q++) {
                const uint32_t j = jof<s>(mp.t, q);
                const uint32_t k = (NR % TPL == 0 || j < (uint32_t)NR) ? j % L : 0;
#pragma unroll
                for (int m = 1; m < R; m++)
                    twr[tw_base(s) + q * (R - 1) + m - 1] = ld_tw<T>(a.tw, off + (m - 1) * L + k);
            }
            prefetch<s + 1>(a, twr);
        }
    }

    // Issues the global loads of lines [g0, g0 + LINES) into w, in the layout
    // of stage 0.  Branch-free: a line group past the end reads the last valid
    // group and a butterfly past N/R reads point N/R-1; those values are never
    // stored.  A conditional load per point would make the compiler wait for
    // each load before issuing the next.
    static __device__ __forceinline__ void load_lines(const LineArgs& a, uint32_t g0, cpx<T> (&w)[VEC][EMAX],
                                                      cpx<T> (&tf)[VEC]) {
        constexpr int R = RL::at(0);
        constexpr int NR = N / R;
        constexpr int NB = ceil_div(NR, TPL);
        const Map mp = map_of<0>();
        const uint32_t g = g0 + mp.lg * VEC;
        const uint32_t gl = g < a.lines ? g : (a.lines - VEC);
        LineBase b = line_base(a, FOLD ? gl / VEC : gl);
        // point j + m*NR sits at o0 + (m*NR)*sn: one add per point, the
        // products of constants and the uniform stride stay scalar
        const uint32_t sn = (uint32_t)a.in_sn;
        cpx<T> tin{};
        if constexpr (UNFOLD) {
            const cpx<T> t = post_twiddle<T>(a, b.j);
            tin = mk<T>(t.x, -t.y);  // W_M^{-n}
        }
        // the fold's store twiddles W_M^{r c}, loaded with the tile (a
        // twiddle load at the store would wait for every load issued before
        // it, the next tile's included, in the pipelined form)
        if constexpr (FOLD && !UNFOLD) {
#pragma unroll
            for (int c = 0; c < VEC; c++) tf[c] = post_twiddle<T>(a, (uint64_t)b.j * c);
        }
        if constexpr (ADJ) {
            // adjacent butterflies j, j + 1: points j + m NR and j + 1 + m NR
            // in one 16-byte load (contiguous line, even element index)
#pragma unroll
            for (int q = 0; q < NB; q += 2) {
                const uint32_t o0 = b.in + jof<0>(mp.t, q);
#pragma unroll
                for (int m = 0; m < R; m++) {
#pragma unroll
                    for (int v = 0; v < VEC; v++)
                        load_pair_at<Fmt, NTL>(a, o0 + (uint32_t)(m * NR) + (FOLD ? v * a.fold_in : 0u),
                                               w[v][q * R + m], w[v][(q + 1) * R + m]);
                }
            }
            return;
        }
#pragma unroll
        for (int q = 0; q < NB; q++) {
            uint32_t j = mp.t + q * TPL;
            if constexpr (NR % TPL != 0) j = j < (uint32_t)NR ? j : (uint32_t)(NR - 1);
            const uint32_t o0 = b.in + j * sn;
#pragma unroll
            for (int m = 0; m < R; m++) {
                const uint32_t idx = o0 + (uint32_t)(m * NR) * sn;
                if constexpr (UNFOLD) {
                    // Hermitian extension of the two unfolded packed rows, as load_point
                    const uint32_t p = j + m * NR, h = N >> 1;
                    const bool up = p > h;
                    const uint32_t at = b.in + (up ? N - p : p) * sn;
                    const cpx<T> e0 = Fmt::ld(a.in, at);
                    const cpx<T> e1 = cmul(Fmt::ld(a.in, at + a.fold_in), tin);
                    const bool self = p == 0 || 2 * p == N;
                    const uint32_t cm = sign_mask(a.conj_in);
                    cpx<T> u0 = e0 + e1, u1 = e0 - e1;
                    u0.y = self ? T(0) : (up ? -u0.y : u0.y);
                    u1.y = self ? T(0) : (up ? -u1.y : u1.y);
                    w[0][q * R + m] = conj_by(u0, cm);
                    w[1][q * R + m] = conj_by(u1, cm);
                } else if constexpr (FOLD) {
#pragma unroll
                    for (int v = 0; v < VEC; v++)
                        w[v][q * R + m] = load_point_at<Fmt, LDM, NTL>(a, idx + v * a.fold_in);
                } else if constexpr (VEC == 2)
                    load_pair_at<Fmt, NTL>(a, idx, w[0][q * R + m], w[1][q * R + m]);
                else if constexpr (LDM == LD_HERMITIAN)
                    w[0][q * R + m] = load_point<Fmt, LDM>(a, b.in, j + m * NR, N);
                else
                    w[0][q * R + m] = load_point_at<Fmt, LDM, NTL>(a, idx);
            }
        }
    }

    // Stage s of lines [g0, g0 + LINES) and every later stage; the last one
    // stores.  With LOAD false, stage 0 takes w as already loaded.
    template <int s, bool LOAD = true>
    static __device__ __forceinline__ void stage(const LineArgs& a, cpx<T>* lds, uint32_t g0,
                                                 cpx<T> (&w)[VEC][EMAX], cpx<T> (&twr)[TWREG],
                                                 cpx<T> (&tf)[VEC]) {
        if constexpr (s < NS) {
            constexpr int R = RL::at(s);
            constexpr int L = RL::lprod(s);
            constexpr int NR = N / R;
            constexpr int NB = ceil_div(NR, TPL);
            constexpr bool FIRST = (s == 0), LAST = (s == NS - 1);
            const Map mp = map_of<s>();
            const uint32_t g = g0 + mp.lg * VEC;
            const bool valid = g < a.lines;
            if constexpr (FIRST && LOAD) {
                load_lines(a, g0, w, tf);
                if constexpr (PF) prefetch<1>(a, twr);
            }
            // twiddle + butterfly (one twiddle load serves the VEC lines)
#pragma unroll
            for (int q = 0; q < NB; q++) {
                const uint32_t j = jof<s>(mp.t, q);
                if (NR % TPL == 0 || j < (uint32_t)NR) {
                    if constexpr (L > 1) {
                        const uint32_t k = j % L;
                        constexpr uint32_t off = RL::twoff(s);
#pragma unroll
                        for (int m = 1; m < R; m++) {
                            cpx<T> tw;
                            if constexpr (PF) tw = twr[tw_base(s) + q * (R - 1) + m - 1];
                            else tw = ld_tw<T>(a.tw, off + (m - 1) * L + k);
#pragma unroll
                            for (int v = 0; v < VEC; v++) w[v][q * R + m] = cmul(w[v][q * R + m], tw);
                        }
                    }
#pragma unroll
                    for (int v = 0; v < VEC; v++) Dft<R, T>::run(&w[v][q * R]);
                }
            }
            if constexpr (LAST) {
                LineBase b = valid ? line_base(a, FOLD ? g / VEC : g) : LineBase{0, 0, 0};
                const uint32_t sk = (uint32_t)a.out_sk;
                // fold: rows y[r + b Q] (b < F) of the next axis (M = F Q)
                // leave as z_c = W_M^{r c} sum_b y[r + b Q] W_F^{b c}, its first
                // radix-F decimation-in-frequency stage
                if constexpr (adj_stage<s>()) {
                    // adjacent butterflies j, j + 1 of the last stage write
                    // adjacent outputs j + m L, j + 1 + m L: 16-byte stores
                    if (valid) {
#pragma unroll
                        for (int q = 0; q < NB; q += 2) {
                            const uint32_t j = jof<s>(mp.t, q);
                            const uint32_t o0 = b.out + ((j - j % L) * R + j % L);
#pragma unroll
                            for (int m = 0; m < R; m++) {
                                const uint32_t idx = o0 + (uint32_t)(m * L);
                                if constexpr (FOLD) {
                                    cpx<T> z0[VEC], z1[VEC];
#pragma unroll
                                    for (int c = 0; c < VEC; c++) {
                                        z0[c] = w[c][q * R + m];
                                        z1[c] = w[c][(q + 1) * R + m];
                                    }
                                    Dft<VEC, T>::run(z0);
                                    Dft<VEC, T>::run(z1);
                                    store_pair_at<Fmt, false>(a, idx, b.j, 0, z0[0], z1[0]);
#pragma unroll
                                    for (int c = 1; c < VEC; c++)
                                        store_pair_at<Fmt, false>(a, idx + c * a.fold_out, b.j, 0, cmul(z0[c], tf[c]),
                                                                  cmul(z1[c], tf[c]));
                                } else {
                                    store_pair_at<Fmt, false>(a, idx, b.j, 0, w[0][q * R + m], w[0][(q + 1) * R + m]);
                                }
                            }
                        }
                    }
                    return;
                }
#pragma unroll
                for (int q = 0; q < NB; q++) {
                    const uint32_t j = mp.t + q * TPL;
                    if (valid && (NR % TPL == 0 || j < (uint32_t)NR)) {
                        const uint32_t k = j % L;
                        const uint32_t p0 = (j - k) * R + k;
                        const uint32_t o0 = b.out + p0 * sk;
                        // outputs p0 + m L: four-step twiddles by recurrence (the
                        // lines of a pair share j)
                        PostSeq ps0;
                        if constexpr (POST) ps0.init(a, b.j, p0, L);
#pragma unroll
                        for (int m = 0; m < R; m++) {
                            const uint32_t idx = o0 + (uint32_t)(m * L) * sk;
                            cpx<T> v0 = w[0][q * R + m];
                            cpx<T> t0;
                            if constexpr (POST) {
                                t0 = ps0.template next<T>();
                                v0 = cmul(v0, t0);
                            }
                            if constexpr (UNFOLD) {
#pragma unroll
                                for (int c = 0; c < VEC; c++)
                                    store_point_at<Fmt, STM, false>(a, idx + c * a.fold_out, b.j, p0 + m * L, N,
                                                                    w[c][q * R + m]);
                            } else if constexpr (FOLD) {
                                cpx<T> z[VEC];
#pragma unroll
                                for (int c = 0; c < VEC; c++) z[c] = w[c][q * R + m];
                                Dft<VEC, T>::run(z);
                                store_point_at<Fmt, STM, false>(a, idx, b.j, p0 + m * L, N, z[0]);
#pragma unroll
                                for (int c = 1; c < VEC; c++)
                                    store_point_at<Fmt, STM, false>(a, idx + c * a.fold_out, b.j, p0 + m * L, N,
                                                                    cmul(z[c], tf[c]));
                            } else if constexpr (VEC == 1) {
                                store_point_at<Fmt, STM, false>(a, idx, b.j, p0 + m * L, N, v0);
                            } else {
                                cpx<T> v1 = w[1][q * R + m];
                                if constexpr (POST) v1 = cmul(v1, t0);
                                store_pair_at<Fmt, false>(a, idx, b.j, p0 + m * L, v0, v1);
                            }
                        }
                    }
                }
            } else {
                exchange<s>(lds, w);
                stage<s + 1>(a, lds, g0, w, twr, tf);
            }
        }
    }

    // Transforms lines [g0, g0 + LINES) with the first NT threads of the block.
    static __device__ __forceinline__ void tile(const LineArgs& a, cpx<T>* lds, uint32_t g0) {
        cpx<T> w[VEC][EMAX];
        cpx<T> twr[TWREG];
        cpx<T> tf[VEC];
        stage<0>(a, lds, g0, w, twr, tf);
    }

    static __device__ __forceinline__ void run(const LineArgs& a, cpx<T>* lds) {
        const uint32_t bid = xcd_remap(blockIdx.x, gridDim.x);
        tile(a, lds, bid * LINES);
    }

    // Persistent form: block b transforms tiles b, b + G, b + 2G, ... (G the
    // grid size) and issues the loads of its next tile before transforming the
    // current one, so that every CU keeps reads in flight while it computes
    // and stores (with one tile per block, the whole chip would load, then
    // compute, then store in lockstep).  The exchange barriers wait for LDS
    // only, so the loads stay in flight across them.
    static __device__ __forceinline__ void run_pipelined(const LineArgs& a, cpx<T>* lds) {
        const uint32_t G = gridDim.x;
        const uint32_t ntiles = (a.lines + LINES - 1) / LINES;
        const uint32_t t0 = xcd_remap(blockIdx.x, G);
        if (t0 >= ntiles) return;
        const uint32_t nt = (ntiles - t0 + G - 1) / G;  // tiles of this block
        cpx<T> wa[VEC][EMAX], wb[VEC][EMAX];
        // Stage twiddles depend on the thread's position only: with PF they
        // are loaded once, so that no global load is issued (and waited for)
        // between the next tile's loads and the current tile's stages; the
        // in-order load counter would otherwise drain the next tile's loads.
        cpx<T> twr[TWREG];
        if constexpr (PF) prefetch<1>(a, twr);
        // The loads inside the loop are unconditional: a load under a branch
        // makes the compiler wait for it at the join.
        cpx<T> fa[VEC], fb[VEC];
        load_lines(a, t0 * LINES, wa, fa);
        uint32_t i = 0;
        for (; i + 2 < nt; i += 2) {
            load_lines(a, (t0 + (i + 1) * G) * LINES, wb, fb);
            stage<0, false>(a, lds, (t0 + i * G) * LINES, wa, twr, fa);
            load_lines(a, (t0 + (i + 2) * G) * LINES, wa, fa);
            stage<0, false>(a, lds, (t0 + (i + 1) * G) * LINES, wb, twr, fb);
        }
        if (i + 1 < nt) {
            load_lines(a, (t0 + (i + 1) * G) * LINES, wb, fb);
            stage<0, false>(a, lds, (t0 + i * G) * LINES, wa, twr, fa);
            stage<0, false>(a, lds, (t0 + (i + 1) * G) * LINES, wb, twr, fb);
        } else {
            stage<0, false>(a, lds, (t0 + i * G) * LINES, wa, twr, fa);
        }
    }
};

template <class K>
__global__ void __launch_bounds__(K::NT) reg_fft_kernel(LineArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    K::run(a, reinterpret_cast<cpx<typename K::T>*>(smem_raw));
}

// Persistent pipelined launch of K (grid below the tile count: the
// workgroups resident at once); MINB is the launch bound's minimum of waves
// per SIMD, which sets the register budget.
template <class K, int MINB>
__global__ void __launch_bounds__(K::NT, MINB) reg_fft_pipe_kernel(LineArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    K::run_pipelined(a, reinterpret_cast<cpx<typename K::T>*>(smem_raw));
}

// ------------------------------------------------ runtime-schedule LDS kernel
constexpr int kMaxStages = 24;

struct Sched {
    uint32_t n;    // line length
    uint32_t ns;   // stages
    uint32_t lpw;  // lines per workgroup
    uint32_t ld_lf, st_lf;
    DivMagic dn;    // divide by n
    DivMagic dlpw;  // divide by lpw
    uint8_t rad[kMaxStages];
    uint32_t L[kMaxStages];
    uint32_t twoff[kMaxStages];  // stage offsets in the stage-twiddle table
    DivMagic dnr[kMaxStages];    // divide by n / rad
    DivMagic dl[kMaxStages];     // divide by L
};

// Whether a schedule holds a medium prime radix (17..61), which only the
// kernels instantiated for them run.
__host__ __device__ inline bool sched_has_medium_prime(const Sched& sc) {
    for (uint32_t s = 0; s < sc.ns; s++)
        if (sc.rad[s] > 16) return true;
    return false;
}

// LDS line stride (elements) of the runtime-schedule kernels.
__host__ __device__ constexpr uint32_t lds_line_stride(uint32_t n) { return n + (n >> 4) + 1; }
__device__ __forceinline__ uint32_t lpad(uint32_t p) { return p + (p >> 4); }

// One Stockham stage over `nl` lines resident in LDS: src -> dst.
template <int R, class T>
__device__ __forceinline__ void lds_stage(const cpx<T>* src, cpx<T>* dst, uint32_t nl,
                                          uint32_t n, uint32_t lstr, uint32_t L, DivMagic dnr,
                                          DivMagic dl, uint32_t twoff, const void* tw) {
    const uint32_t nr = n / R;
    const uint32_t total = nl * nr;
    for (uint32_t b = threadIdx.x; b < total; b += blockDim.x) {
        const uint32_t l = fdiv(b, dnr);
        const uint32_t j = b - l * nr;
        const uint32_t kq = fdiv(j, dl);
        const uint32_t k = j - kq * L;
        const cpx<T>* s = src + l * lstr;
        cpx<T> w[R];
#pragma unroll
        for (int m = 0; m < R; m++) w[m] = s[lpad(j + m * nr)];
        if (L > 1) {
#pragma unroll
            for (int m = 1; m < R; m++) w[m] = cmul(w[m], ld_tw<T>(tw, twoff + (m - 1) * L + k));
        }
        Dft<R, T>::run(w);
        cpx<T>* d = dst + l * lstr;
        const uint32_t ob = kq * L * R + k;
#pragma unroll
        for (int m = 0; m < R; m++) d[lpad(ob + m * L)] = w[m];
    }
}

// Output pair q, R - q of a symmetric-pair prime DFT (as DftOdd) with the
// constants W_R^(m q) resolved at compile time, stored to LDS at stride L.
template <int R, int q, int m, class T>
__device__ __forceinline__ void prime_acc(cpx<T>& A, cpx<T>& B, const cpx<T>* ap, const cpx<T>* am) {
    if constexpr (m <= (R - 1) / 2) {
        constexpr int e = (m * q) % R;
        constexpr T c = (T)root_cos<R>(e);
        constexpr T sn = (T)root_sin<R>(e);
        A.x += c * ap[m - 1].x;
        A.y += c * ap[m - 1].y;
        B.x += sn * am[m - 1].x;
        B.y += sn * am[m - 1].y;
        prime_acc<R, q, m + 1, T>(A, B, ap, am);
    }
}
template <int R, int q, class T>
__device__ __forceinline__ void prime_out(cpx<T>* d, uint32_t ob, uint32_t L, cpx<T> v0, const cpx<T>* ap,
                                          const cpx<T>* am) {
    if constexpr (q <= (R - 1) / 2) {
        cpx<T> A = v0, B = mk<T>(T(0), T(0));
        prime_acc<R, q, 1, T>(A, B, ap, am);
        d[lpad(ob + q * L)] = mk<T>(A.x + B.y, A.y - B.x);
        d[lpad(ob + (R - q) * L)] = mk<T>(A.x - B.y, A.y + B.x);
        prime_out<R, q + 1, T>(d, ob, L, v0, ap, am);
    }
}

// Inputs m and R - m of a prime butterfly (twiddled when L > 1) as their
// sum and difference.
template <int R, int m, class T>
__device__ __forceinline__ void prime_in(cpx<T>* ap, cpx<T>* am, cpx<T>& sum, const cpx<T>* s, uint32_t j,
                                         uint32_t nr, uint32_t L, uint32_t k, uint32_t twoff, const void* tw) {
    if constexpr (m <= (R - 1) / 2) {
        cpx<T> x = s[lpad(j + m * nr)], y = s[lpad(j + (R - m) * nr)];
        if (L > 1) {
            x = cmul(x, ld_tw<T>(tw, twoff + (m - 1) * L + k));
            y = cmul(y, ld_tw<T>(tw, twoff + (R - m - 1) * L + k));
        }
        ap[m - 1] = x + y;
        am[m - 1] = x - y;
        sum = sum + ap[m - 1];
        prime_in<R, m + 1, T>(ap, am, sum, s, j, nr, L, k, twoff, tw);
    }
}

// One Stockham stage of a medium prime radix R (17..61) over `nl` lines in
// LDS.  The butterfly is the symmetric-pair DFT of DftOdd, written so that
// only the pair sums and differences stay in registers: inputs j and R - j
// are loaded (and twiddled) together, and each output pair k, R - k goes to
// LDS as soon as it is formed.
template <int R, class T>
__device__ __forceinline__ void lds_stage_prime(const cpx<T>* src, cpx<T>* dst, uint32_t nl, uint32_t n,
                                                uint32_t lstr, uint32_t L, DivMagic dnr, DivMagic dl,
                                                uint32_t twoff, const void* tw) {
    constexpr int H = (R - 1) / 2;
    const uint32_t nr = n / R;
    const uint32_t total = nl * nr;
#pragma nounroll
    for (uint32_t b = threadIdx.x; b < total; b += blockDim.x) {
        const uint32_t l = fdiv(b, dnr);
        const uint32_t j = b - l * nr;
        const uint32_t kq = fdiv(j, dl);
        const uint32_t k = j - kq * L;
        const cpx<T>* s = src + l * lstr;
        const cpx<T> v0 = s[lpad(j)];
        cpx<T> ap[H], am[H];
        cpx<T> sum = v0;
        prime_in<R, 1, T>(ap, am, sum, s, j, nr, L, k, twoff, tw);
        cpx<T>* d = dst + l * lstr;
        const uint32_t ob = kq * L * R + k;
        d[lpad(ob)] = sum;
        prime_out<R, 1, T>(d, ob, L, v0, ap, am);
    }
}

// Transforms nl lines held in LDS buffer `a` (line stride lstr) using `b` as
// ping-pong space; returns the buffer holding the result.  Every thread of the
// block must call it (it synchronises after each stage).  MED adds the
// medium prime radices 17..61 (Rader convolutions over such p - 1) to the
// stage switch; it is a separate instantiation so that the register demand
// of those stages does not reach the other kernels.
template <class T, bool MED>
__device__ __forceinline__ cpx<T>* lds_fft_lines_inl(cpx<T>* a, cpx<T>* b, uint32_t nl, const Sched& sc,
                                                     const void* tw) {
    const uint32_t lstr = lds_line_stride(sc.n);
    cpx<T>* src = a;
    cpx<T>* dst = b;
    for (uint32_t s = 0; s < sc.ns; s++) {
        const uint32_t R = sc.rad[s];
        switch (R) {
#define MIFFT_CASE(RR)                                                                     \
    case RR:                                                                               \
        lds_stage<RR, T>(src, dst, nl, sc.n, lstr, sc.L[s], sc.dnr[s], sc.dl[s], sc.twoff[s], tw); \
        break;
            MIFFT_CASE(2)
            MIFFT_CASE(3)
            MIFFT_CASE(4)
            MIFFT_CASE(5)
            MIFFT_CASE(6)
            MIFFT_CASE(7)
            MIFFT_CASE(8)
            MIFFT_CASE(9)
            MIFFT_CASE(10)
            MIFFT_CASE(11)
            MIFFT_CASE(12)
            MIFFT_CASE(13)
            MIFFT_CASE(14)
            MIFFT_CASE(15)
            MIFFT_CASE(16)
#undef MIFFT_CASE
            // medium primes (MED only)
#define MIFFT_PCASE(RR)                                                                                     \
    case RR:                                                                                                \
        if constexpr (MED)                                                                                  \
            lds_stage_prime<RR, T>(src, dst, nl, sc.n, lstr, sc.L[s], sc.dnr[s], sc.dl[s], sc.twoff[s], tw); \
        break;
            MIFFT_PCASE(17)
            MIFFT_PCASE(19)
            MIFFT_PCASE(23)
            MIFFT_PCASE(29)
            MIFFT_PCASE(31)
            MIFFT_PCASE(37)
            MIFFT_PCASE(41)
            MIFFT_PCASE(43)
            MIFFT_PCASE(47)
            MIFFT_PCASE(53)
            MIFFT_PCASE(59)
            MIFFT_PCASE(61)
#undef MIFFT_PCASE
            default:
                break;
        }
        __syncthreads();
        cpx<T>* t = src;
        src = dst;
        dst = t;
    }
    return src;
}

// The same as a call (the schedule's own kernels), or inlined for MED: as a
// called function the medium prime stages see a divergent radix and a
// caller-independent register budget and measured 256 VGPRs with scratch;
// inlined into the kernel, 175 without.
template <class T, bool MED = false>
__device__ cpx<T>* lds_fft_lines(cpx<T>* a, cpx<T>* b, uint32_t nl, const Sched& sc, const void* tw) {
    if constexpr (MED) return lds_fft_lines_inl<T, true>(a, b, nl, sc, tw);
    else return lds_fft_lines_inl<T, false>(a, b, nl, sc, tw);
}

// Moving `lpw` lines of `len` points between memory and LDS: slot s of
// lpw * tpl slots (tpl = threads per line) owns line l and points
// t, t + tpl, ...; with lines-fastest memory order adjacent slots take adjacent
// lines (coalesced across lines), otherwise adjacent points of one line.  A
// slot computes its line's base once and moves kChunk points per step with
// the loads issued together (clamped indices, no branch per load), so they
// are in flight at once instead of one memory latency per point.
constexpr uint32_t kChunk = 8;
struct LineSlot {
    uint32_t l, t;
};
__device__ __forceinline__ uint32_t slot_tpl(uint32_t lpw) {
    const uint32_t tpl = blockDim.x / lpw;
    return tpl ? tpl : 1u;
}
__device__ __forceinline__ LineSlot line_slot(uint32_t s, uint32_t lpw, uint32_t tpl, bool lf) {
    return lf ? LineSlot{s % lpw, s / lpw} : LineSlot{s / tpl, s % tpl};
}

// Lines [g0, g0 + sc.lpw) through LDS with the runtime schedule `sc`.
template <class Fmt, int LDM, int STM, bool POST>
__device__ void lds_fft_tile(const LineArgs& a, const Sched& sc, unsigned char* smem_raw,
                             uint32_t g0) {
    using T = typename Fmt::T;
    const uint32_t lstr = lds_line_stride(sc.n);
    cpx<T>* bufA = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* bufB = bufA + sc.lpw * lstr;
    const uint32_t n = sc.n, lpw = sc.lpw, tpl = slot_tpl(lpw), slots = tpl * lpw;
    // load (lines past the end read the last line; they are never stored)
    for (uint32_t s = threadIdx.x; s < slots; s += blockDim.x) {
        const LineSlot m = line_slot(s, lpw, tpl, sc.ld_lf);
        const uint32_t g = g0 + m.l;
        const LineBase b = line_base(a, g < a.lines ? g : a.lines - 1);
        for (uint32_t p0 = m.t; p0 < n; p0 += tpl * kChunk) {
            cpx<T> v[kChunk];
#pragma unroll
            for (uint32_t u = 0; u < kChunk; u++) {
                const uint32_t p = p0 + u * tpl;
                v[u] = load_point<Fmt, LDM>(a, b.in, p < n ? p : n - 1, n);
            }
#pragma unroll
            for (uint32_t u = 0; u < kChunk; u++) {
                const uint32_t p = p0 + u * tpl;
                if (p < n) bufA[m.l * lstr + lpad(p)] = v[u];
            }
        }
    }
    __syncthreads();
    cpx<T>* res = lds_fft_lines<T>(bufA, bufB, lpw, sc, a.tw);
    // store
    for (uint32_t s = threadIdx.x; s < slots; s += blockDim.x) {
        const LineSlot m = line_slot(s, lpw, tpl, sc.st_lf);
        const uint32_t g = g0 + m.l;
        if (g >= a.lines) continue;
        const LineBase b = line_base(a, g);
        for (uint32_t p = m.t; p < n; p += tpl)
            store_point<Fmt, STM, POST>(a, b.out, b.j, p, n, res[m.l * lstr + lpad(p)]);
    }
}

// NT threads per workgroup: 256, or 512 where a tile's LDS leaves room for
// one workgroup per CU only (the tile's loops stride by blockDim.x).
template <class Fmt, int LDM, int STM, bool POST, int NT = 256>
__global__ void __launch_bounds__(NT) lds_fft_kernel(LineArgs a, Sched sc) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const uint32_t bid = xcd_remap(blockIdx.x, gridDim.x);
    lds_fft_tile<Fmt, LDM, STM, POST>(a, sc, smem_raw, bid * sc.lpw);
}

}  // namespace mifft
