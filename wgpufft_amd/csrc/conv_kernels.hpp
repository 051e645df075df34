// Fused Rader and Bluestein line transforms as per-tile device functions
// (used by the standalone kernels in kern_conv.hip and by the persistent
// batched kernel in kern_batch.hip).
//
// Each tile keeps `lpw` lines of the
// convolution length M in LDS and runs forward FFT -> filter multiply ->
// inverse FFT without leaving the CU.
//
// Rader (prime p, wgpuFFT src/runtime/rader_axis.rs:202-642): with g a
// primitive root, a[q] = x[g^q], b[t] = W_p^(g^-t) and c = a (*) b (cyclic over
// p-1, or zero-padded to M >= 2p-3), X[0] = x[0] + sum a and
// X[g^-u] = x[0] + c[u].  sum a is bin 0 of FFT_M(a), so no separate
// reduction is needed.
//
// Bluestein (chirp-z, src/runtime/bluestein_axis.rs:179-1041): with
// c[i] = exp(-pi*i*i^2/N), X[k] = c[k] * sum_n (x[n] c[n]) conj(c[k-n]), a linear
// convolution evaluated cyclically over a smooth M >= 2N-1.
//
// The filter spectra (FFT_M of b, scaled by 1/M) are computed once per plan in
// FP64 on the device and rounded once.  The inverse FFT reuses the forward
// engine: IFFT(y) = swap(FFT(swap(y))).
#pragma once
#include "launch.h"

namespace mifft {

// Primes above this stage their lines through LDS in natural order (for
// N = 1009 over 1 GiB 21 % faster; for N = 17, whose line spans three cache
// lines, 18 % slower than gathering straight from memory).
constexpr uint32_t kRaderStagedMin = 64;

// Rader's filter step on the forward spectra in LDS: keeps bin 0 (the sum
// of the permuted inputs) in s0 and leaves swap(A[k] F[k]) for the inverse.
template <class T>
__device__ __forceinline__ void rader_filter(const ConvArgs& c, cpx<T>* r, cpx<T>* s0, uint32_t total,
                                             uint32_t lstr) {
    const uint32_t M = c.m;
    for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
        const uint32_t l = fdiv(e, c.sc.dn);  // sc.n == M
        const uint32_t k = e - l * M;
        cpx<T> v = r[l * lstr + lpad(k)];
        if (k == 0) s0[l] = v;
        r[l * lstr + lpad(k)] = swap_ri(cmul(v, ld_tw<T>(c.filt, k)));
    }
}

// MED: the schedule holds medium prime radices (17..61), f32 only.
template <class Fmt, bool MED = false>
__device__ __forceinline__ void rader_tile(const ConvArgs& c, unsigned char* smem_raw, uint32_t g0) {
    using T = typename Fmt::T;
    const Sched& sc = c.sc;
    const LineArgs& a = c.io;
    const uint32_t M = c.m, p = c.n, lpw = sc.lpw;
    const uint32_t lstr = lds_line_stride(M);
    cpx<T>* bufA = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* bufB = bufA + lpw * lstr;
    cpx<T>* x0 = bufB + lpw * lstr;  // lpw entries
    cpx<T>* s0 = x0 + lpw;           // lpw entries
    // Long lines: the line in natural order into bufB (coalesced loads),
    // then the gather a[q] = x[g^q] (zero-padded to M) and x0 = x[0] from
    // LDS, so that no global load waits on the permutation table's.  Short
    // lines (a few cache lines each) gather straight from memory.  Lines
    // past the end read the last line and are never stored.
    const uint32_t total = lpw * M;
    const uint32_t tpl = slot_tpl(lpw), slots = tpl * lpw;
    const bool staged = p > kRaderStagedMin;
    if (staged) {
        for (uint32_t s = threadIdx.x; s < slots; s += blockDim.x) {
            const LineSlot m = line_slot(s, lpw, tpl, sc.ld_lf);
            const uint32_t g = g0 + m.l;
            const LineBase b = line_base(a, g < a.lines ? g : a.lines - 1);
            for (uint32_t q0 = m.t; q0 < p; q0 += tpl * kChunk) {
                cpx<T> v[kChunk];
#pragma unroll
                for (uint32_t u = 0; u < kChunk; u++) {
                    const uint32_t q = q0 + u * tpl;
                    v[u] = load_point<Fmt, LD_COMPLEX>(a, b.in, q < p ? q : p - 1, p);
                }
#pragma unroll
                for (uint32_t u = 0; u < kChunk; u++) {
                    const uint32_t q = q0 + u * tpl;
                    if (q < p) bufB[m.l * lstr + lpad(q)] = v[u];
                }
            }
        }
        __syncthreads();
        for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
            const uint32_t l = fdiv(e, sc.dn);  // sc.n == M
            const uint32_t q = e - l * M;
            const cpx<T>* xl = bufB + l * lstr;
            if (q == 0) x0[l] = xl[0];
            bufA[l * lstr + lpad(q)] = q < p - 1 ? xl[lpad(c.perm_in[q])] : mk<T>(T(0), T(0));
        }
    } else {
        for (uint32_t s = threadIdx.x; s < slots; s += blockDim.x) {
            const LineSlot m = line_slot(s, lpw, tpl, sc.ld_lf);
            const uint32_t g = g0 + m.l;
            const LineBase b = line_base(a, g < a.lines ? g : a.lines - 1);
            if (m.t == 0) x0[m.l] = load_point<Fmt, LD_COMPLEX>(a, b.in, 0, p);
            for (uint32_t q0 = m.t; q0 < M; q0 += tpl * kChunk) {
                uint32_t src[kChunk];
#pragma unroll
                for (uint32_t u = 0; u < kChunk; u++) {
                    const uint32_t q = q0 + u * tpl;
                    src[u] = c.perm_in[q < p - 1 ? q : p - 2];
                }
                cpx<T> v[kChunk];
#pragma unroll
                for (uint32_t u = 0; u < kChunk; u++) v[u] = load_point<Fmt, LD_COMPLEX>(a, b.in, src[u], p);
#pragma unroll
                for (uint32_t u = 0; u < kChunk; u++) {
                    const uint32_t q = q0 + u * tpl;
                    if (q < M) bufA[m.l * lstr + lpad(q)] = q < p - 1 ? v[u] : mk<T>(T(0), T(0));
                }
            }
        }
    }
    __syncthreads();
    cpx<T>* r2;
    if constexpr (MED) {
        // Both transforms through one inlined copy of the stages (two copies
        // measured 256 VGPRs with scratch, one 176 without).
        cpx<T>* cur = bufA;
        cpx<T>* oth = bufB;
#pragma nounroll
        for (int pass = 0;; pass++) {
            cpx<T>* r = lds_fft_lines_inl<T, true>(cur, oth, lpw, sc, a.tw);
            if (pass == 1) {
                r2 = r;
                break;
            }
            rader_filter<T>(c, r, s0, total, lstr);
            __syncthreads();
            oth = (r == bufA) ? bufB : bufA;
            cur = r;
        }
    } else {
        cpx<T>* r = lds_fft_lines<T>(bufA, bufB, lpw, sc, a.tw);
        cpx<T>* o = (r == bufA) ? bufB : bufA;
        rader_filter<T>(c, r, s0, total, lstr);
        __syncthreads();
        r2 = lds_fft_lines<T>(r, o, lpw, sc, a.tw);
    }
    // X[g^-u] = x0 + c[u], X[0] = x0 + sum a.  Long lines: permuted in LDS
    // (into the buffer the second transform did not end in), then stored in
    // natural order (coalesced); short lines: scattered straight to memory.
    if (staged) {
        cpx<T>* y = (r2 == bufA) ? bufB : bufA;
        for (uint32_t e = threadIdx.x; e < lpw * p; e += blockDim.x) {
            const uint32_t l = fdiv(e, c.dn);
            const uint32_t u = e - l * p;
            if (u < p - 1) y[l * lstr + lpad(c.perm_out[u])] = x0[l] + swap_ri(r2[l * lstr + lpad(u)]);
            else y[l * lstr] = x0[l] + s0[l];
        }
        __syncthreads();
        for (uint32_t s = threadIdx.x; s < slots; s += blockDim.x) {
            const LineSlot m = line_slot(s, lpw, tpl, sc.st_lf);
            const uint32_t g = g0 + m.l;
            if (g >= a.lines) continue;
            const LineBase b = line_base(a, g);
            for (uint32_t k = m.t; k < p; k += tpl)
                store_point<Fmt, ST_COMPLEX, false>(a, b.out, b.j, k, p, y[m.l * lstr + lpad(k)]);
        }
    } else {
        for (uint32_t s = threadIdx.x; s < slots; s += blockDim.x) {
            const LineSlot m = line_slot(s, lpw, tpl, sc.st_lf);
            const uint32_t g = g0 + m.l;
            if (g >= a.lines) continue;
            const LineBase b = line_base(a, g);
            for (uint32_t u = m.t; u < p; u += tpl) {
                cpx<T> v;
                uint32_t k;
                if (u < p - 1) {
                    v = x0[m.l] + swap_ri(r2[m.l * lstr + lpad(u)]);
                    k = c.perm_out[u];
                } else {
                    v = x0[m.l] + s0[m.l];
                    k = 0;
                }
                store_point<Fmt, ST_COMPLEX, false>(a, b.out, b.j, k, p, v);
            }
        }
    }
}

template <class Fmt>
__device__ void bluestein_tile(const ConvArgs& c, unsigned char* smem_raw, uint32_t g0) {
    using T = typename Fmt::T;
    const Sched& sc = c.sc;
    const LineArgs& a = c.io;
    const uint32_t M = c.m, n = c.n, lpw = sc.lpw;
    const uint32_t lstr = lds_line_stride(M);
    cpx<T>* bufA = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* bufB = bufA + lpw * lstr;
    const DivMagic dM = sc.dn;
    const uint32_t total = lpw * M;
    const uint32_t tpl = slot_tpl(lpw), slots = tpl * lpw;
    // gather x[q] c[q] (zero-padded to M); lines past the end read the last
    // line and are never stored
    for (uint32_t s = threadIdx.x; s < slots; s += blockDim.x) {
        const LineSlot m = line_slot(s, lpw, tpl, sc.ld_lf);
        const uint32_t g = g0 + m.l;
        const LineBase b = line_base(a, g < a.lines ? g : a.lines - 1);
        for (uint32_t q0 = m.t; q0 < M; q0 += tpl * kChunk) {
            cpx<T> v[kChunk];
#pragma unroll
            for (uint32_t u = 0; u < kChunk; u++) {
                const uint32_t q = q0 + u * tpl;
                v[u] = load_point<Fmt, LD_COMPLEX>(a, b.in, q < n ? q : n - 1, n);
            }
#pragma unroll
            for (uint32_t u = 0; u < kChunk; u++) {
                const uint32_t q = q0 + u * tpl;
                if (q < M) bufA[m.l * lstr + lpad(q)] = q < n ? cmul(v[u], ld_tw<T>(c.chirp, q)) : mk<T>(T(0), T(0));
            }
        }
    }
    __syncthreads();
    cpx<T>* r = lds_fft_lines<T>(bufA, bufB, lpw, sc, a.tw);
    cpx<T>* o = (r == bufA) ? bufB : bufA;
    for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
        const uint32_t l = fdiv(e, dM);
        const uint32_t k = e - l * M;
        r[l * lstr + lpad(k)] = swap_ri(cmul(r[l * lstr + lpad(k)], ld_tw<T>(c.filt, k)));
    }
    __syncthreads();
    cpx<T>* r2 = lds_fft_lines<T>(r, o, lpw, sc, a.tw);
    for (uint32_t s = threadIdx.x; s < slots; s += blockDim.x) {
        const LineSlot m = line_slot(s, lpw, tpl, sc.st_lf);
        const uint32_t g = g0 + m.l;
        if (g >= a.lines) continue;
        const LineBase b = line_base(a, g);
        for (uint32_t k = m.t; k < n; k += tpl) {
            const cpx<T> v = cmul(swap_ri(r2[m.l * lstr + lpad(k)]), ld_tw<T>(c.chirp, k));
            store_point<Fmt, ST_COMPLEX, false>(a, b.out, b.j, k, n, v);
        }
    }
}

// Element-wise steps of a multi-pass Bluestein convolution (used by the
// standalone kernels in kern_util.hip and by the batched kernel).  The padded
// buffer holds `lines` rows of m points, row-major.
//   pre:  u[l][p] = x_l[p] c[p] for p < n, 0 for n <= p < m
//   filt: u[l][k] = swap(u[l][k] F[k])   (F = FFT_M(conj chirp) / M; the swap
//         turns the next forward FFT into an inverse)
//   post: X_l[k] = swap(u[l][k]) c[k]    for k < n
template <class Fmt>
__device__ __forceinline__ void chirp_pre_elem(const LineArgs& a, const void* chirp, uint32_t n, uint32_t l,
                                               uint32_t p, uint64_t e) {
    using T = typename Fmt::T;
    cpx<T> v = mk<T>(T(0), T(0));
    if (p < n) {
        LineBase b = line_base(a, l);
        v = load_point<Fmt, LD_COMPLEX>(a, b.in, p, n);
        v = cmul(v, ld_tw<T>(chirp, p));
    }
    Fmt::st(a.out, e, v);
}

template <class Fmt>
__device__ __forceinline__ void pointwise_elem(void* buf, const void* filt, uint32_t k, uint64_t e) {
    using T = typename Fmt::T;
    cpx<T> v = Fmt::ld(buf, e);
    Fmt::st(buf, e, swap_ri(cmul(v, ld_tw<T>(filt, k))));
}

template <class Fmt>
__device__ __forceinline__ void chirp_post_elem(const LineArgs& a, const void* chirp, uint32_t n, uint32_t m,
                                                uint32_t l, uint32_t k) {
    using T = typename Fmt::T;
    cpx<T> v = swap_ri(Fmt::ld(a.in, (uint64_t)l * m + k));
    v = cmul(v, ld_tw<T>(chirp, k));
    LineBase b = line_base(a, l);
    store_point<Fmt, ST_COMPLEX, false>(a, b.out, b.j, k, n, v);
}

}  // namespace mifft
