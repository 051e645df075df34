// Register-resident Bluestein convolution (the reference's register Bluestein
// kernel, src/runtime/register_fft.rs:510-664, bluestein_axis.rs:937-969):
// one kernel per axis for f32 lines of N points whose power-of-two
// convolution length M >= 2N - 1 fits a workgroup.
//
//   a[p] = x[p] c[p] (p < N), 0 (N <= p < M)      c[p] = W_2N^(p^2 mod 2N)
//   A = FFT_M(a);  B[k] = swap(A[k] F[k])          F = FFT_M(conj c, padded) / M
//   C = FFT_M(B);  X[k] = swap(C[k]) c[k]          (k < N)
//
// The swap of real and imaginary parts turns the second forward transform
// into an inverse one.  Both transforms run as Stockham stages in registers
// with LDS exchanges between stages, as the line kernels do; the radix
// schedule starts and ends with the same radix, so the first transform's
// outputs (point j + m L_last of butterfly j) sit in exactly the registers
// the second transform's first stage reads (point j + m N/R_0): the filter
// multiply needs no exchange.  Rader primes whose cyclic convolution would
// be zero-padded take this route too (the reference does the same,
// rader_axis.rs:1262-1280).
#pragma once
#include "launch.h"

namespace mifft {

template <int M, int TPL, int LINES, class RL>
struct RegBluestein {
    using T = float;
    static constexpr int kM = M, kTPL = TPL, kLines = LINES;
    static constexpr int NS = RL::n;
    static constexpr int NT = TPL * LINES;
    static constexpr int LSTR = lds_pad(M) + 1;
    static_assert(RL::prod() == M, "radices multiply to M");
    static_assert(RL::at(0) == RL::at(NS - 1), "first and last radix agree");
    static constexpr int max_regs(int s) {
        return s >= NS ? 0
                       : ((M / RL::at(s)) / TPL * RL::at(s) > max_regs(s + 1) ? (M / RL::at(s)) / TPL * RL::at(s)
                                                                              : max_regs(s + 1));
    }
    static constexpr int EMAX = max_regs(0);
    static constexpr size_t lds_bytes() { return (size_t)LINES * LSTR * sizeof(cpx<T>); }

    template <int s>
    static constexpr bool fits() {
        return (M / RL::at(s)) % TPL == 0;
    }

    // Stages s.. of one forward transform of the group's line; input in the
    // layout of stage 0 (register q R_0 + m holds point j + m M/R_0, j = t +
    // q TPL), output in that of the last stage (register q R + m holds output
    // j + m M/R).  Every thread of the workgroup calls it (the exchanges take
    // workgroup barriers).
    template <int s>
    static __device__ __forceinline__ void stages(cpx<T> (&w)[EMAX], cpx<T>* ln, uint32_t t, const void* tw) {
        if constexpr (s < NS) {
            constexpr int R = RL::at(s), L = RL::lprod(s), NR = M / R, NB = NR / TPL;
            static_assert(fits<s>(), "every thread takes whole butterflies");
#pragma unroll
            for (int q = 0; q < NB; q++) {
                const uint32_t j = t + q * TPL;
                if constexpr (L > 1) {
                    const uint32_t k = j % L;
                    constexpr uint32_t off = RL::twoff(s);
#pragma unroll
                    for (int m = 1; m < R; m++) w[q * R + m] = cmul(w[q * R + m], ld_tw<T>(tw, off + (m - 1) * L + k));
                }
                Dft<R, T>::run(&w[q * R]);
            }
            if constexpr (s + 1 < NS) {
                constexpr int R2 = RL::at(s + 1), NR2 = M / R2, NB2 = NR2 / TPL;
#pragma unroll
                for (int q = 0; q < NB; q++) {
                    const uint32_t j = t + q * TPL;
                    const uint32_t k = j % L;
#pragma unroll
                    for (int m = 0; m < R; m++) ln[lds_pad((j - k) * R + k + m * L)] = w[q * R + m];
                }
                __syncthreads();
#pragma unroll
                for (int q = 0; q < NB2; q++) {
                    const uint32_t j = t + q * TPL;
#pragma unroll
                    for (int m = 0; m < R2; m++) w[q * R2 + m] = ln[lds_pad(j + m * NR2)];
                }
                __syncthreads();
                stages<s + 1>(w, ln, t, tw);
            }
        }
    }

    // Lines [g0, g0 + LINES) with the NT threads of the workgroup (the
    // standalone kernel and the batched kernel's class share this body, so a
    // line's result is the same in both).
    static __device__ __forceinline__ void tile(const ConvArgs& c, cpx<T>* lds, uint32_t g0) {
        constexpr int R0 = RL::at(0), NR0 = M / R0, NB0 = NR0 / TPL;
        const LineArgs& a = c.io;
        const uint32_t n = c.n;
        const uint32_t lg = threadIdx.x / TPL, t = threadIdx.x % TPL;
        const uint32_t g = g0 + lg;
        const bool valid = g < a.lines;
        const LineBase b = line_base(a, valid ? g : a.lines - 1);
        cpx<T> w[EMAX];
        // a[p] = x[p] c[p], zero past N; branch-free loads (a point past N
        // reads point N - 1 and is replaced by zero)
#pragma unroll
        for (int q = 0; q < NB0; q++) {
            const uint32_t j = t + q * TPL;
#pragma unroll
            for (int m = 0; m < R0; m++) {
                const uint32_t p = j + m * NR0;
                const cpx<T> v = load_point<FmtF32, LD_COMPLEX>(a, b.in, p < n ? p : n - 1, n);
                const cpx<T> cp = ld_tw<T>(c.chirp, p < n ? p : 0);
                w[q * R0 + m] = p < n ? cmul(v, cp) : mk<T>(T(0), T(0));
            }
        }
        cpx<T>* ln = lds + lg * LSTR;
        stages<0>(w, ln, t, a.tw);
        // B[p] = swap(A[p] F[p]); the last stage's output p = j + m M/R_0 is
        // the first stage's input point p of the same register
#pragma unroll
        for (int q = 0; q < NB0; q++) {
            const uint32_t j = t + q * TPL;
#pragma unroll
            for (int m = 0; m < R0; m++) {
                const uint32_t p = j + m * NR0;
                w[q * R0 + m] = swap_ri(cmul(w[q * R0 + m], ld_tw<T>(c.filt, p)));
            }
        }
        stages<0>(w, ln, t, a.tw);
        if (!valid) return;
#pragma unroll
        for (int q = 0; q < NB0; q++) {
            const uint32_t j = t + q * TPL;
#pragma unroll
            for (int m = 0; m < R0; m++) {
                const uint32_t p = j + m * NR0;
                if (p < n)
                    store_point<FmtF32, ST_COMPLEX, false>(a, b.out, b.j, p, n,
                                                           cmul(swap_ri(w[q * R0 + m]), ld_tw<T>(c.chirp, p)));
            }
        }
    }

    static __device__ __forceinline__ void run(const ConvArgs& c, cpx<T>* lds) {
        tile(c, lds, xcd_remap(blockIdx.x, gridDim.x) * LINES);
    }
};

// The compiled schedules, 256 threads each (the batched kernel's workgroup):
// 16 points per thread up to M = 4096, 32 at M = 8192; first and last radix
// equal.
using RegB64 = RegBluestein<64, 4, 64, RadixList<8, 8>>;
using RegB128 = RegBluestein<128, 8, 32, RadixList<8, 2, 8>>;
using RegB256 = RegBluestein<256, 16, 16, RadixList<16, 16>>;
using RegB512 = RegBluestein<512, 32, 8, RadixList<8, 8, 8>>;
using RegB1024 = RegBluestein<1024, 64, 4, RadixList<16, 4, 16>>;
using RegB2048 = RegBluestein<2048, 128, 2, RadixList<16, 8, 16>>;
using RegB4096 = RegBluestein<4096, 256, 1, RadixList<16, 16, 16>>;
using RegB8192 = RegBluestein<8192, 256, 1, RadixList<8, 16, 8, 8>>;
// 512-thread form of the longest (16 points per thread, 128 registers: two
// workgroups per CU where the 256-thread form's 256 registers allow one) for
// the standalone kernel; every butterfly's arithmetic is that of the
// 256-thread form, which the batched kernel runs (N = 3256 over 1 GiB 7.00
// -> 3.71 ms per pair).  A 512-thread 4096-point form (8x8x8x8) measured
// 5-12 % slower than RegB4096.
using RegB8192W = RegBluestein<8192, 512, 1, RadixList<8, 16, 8, 8>>;

template <class K>
__global__ void __launch_bounds__(K::NT) reg_bluestein_kernel(ConvArgs c) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    K::run(c, reinterpret_cast<cpx<typename K::T>*>(smem_raw));
}

// The same with at least MINB waves per SIMD (the register budget that lets
// two 512-thread workgroups share a CU).
template <class K, int MINB>
__global__ void __launch_bounds__(K::NT, MINB) reg_bluestein_kernel_mb(ConvArgs c) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    K::run(c, reinterpret_cast<cpx<typename K::T>*>(smem_raw));
}

}  // namespace mifft
