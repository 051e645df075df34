// Fused Rader and Bluestein kernels: each workgroup keeps `lpw` lines of the
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
#include "launch.h"

namespace mifft {

template <class Fmt>
__global__ void __launch_bounds__(256) k_rader(ConvArgs c) {
    using T = typename Fmt::T;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const Sched& sc = c.sc;
    const LineArgs& a = c.io;
    const uint32_t M = c.m, p = c.n, lpw = sc.lpw;
    const uint32_t lstr = lds_line_stride(M);
    cpx<T>* bufA = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* bufB = bufA + lpw * lstr;
    cpx<T>* x0 = bufB + lpw * lstr;  // lpw entries
    cpx<T>* s0 = x0 + lpw;           // lpw entries
    const uint32_t bid = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t g0 = bid * lpw;
    const DivMagic dM = sc.dn;  // sc.n == M
    // gather a[q] = x[g^q]
    const uint32_t total = lpw * M;
    for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
        uint32_t l, q;
        if (sc.ld_lf) {
            q = fdiv(e, sc.dlpw);
            l = e - q * lpw;
        } else {
            l = fdiv(e, dM);
            q = e - l * M;
        }
        const uint32_t g = g0 + l;
        cpx<T> v = mk<T>(T(0), T(0));
        if (g < a.lines) {
            LineBase b = line_base(a, g);
            if (q < p - 1) v = load_point<Fmt, LD_COMPLEX>(a, b.in, c.perm_in[q], p);
            if (q == 0) x0[l] = load_point<Fmt, LD_COMPLEX>(a, b.in, 0, p);
        } else if (q == 0) {
            x0[l] = mk<T>(T(0), T(0));
        }
        bufA[l * lstr + lpad(q)] = v;
    }
    __syncthreads();
    cpx<T>* r = lds_fft_lines<T>(bufA, bufB, lpw, sc, a.tw);
    cpx<T>* o = (r == bufA) ? bufB : bufA;
    for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
        const uint32_t l = fdiv(e, dM);
        const uint32_t k = e - l * M;
        cpx<T> v = r[l * lstr + lpad(k)];
        if (k == 0) s0[l] = v;
        r[l * lstr + lpad(k)] = swap_ri(cmul(v, ld_tw<T>(c.filt, k)));
    }
    __syncthreads();
    cpx<T>* r2 = lds_fft_lines<T>(r, o, lpw, sc, a.tw);
    // scatter X[g^-u] = x0 + c[u], X[0] = x0 + sum a
    const uint32_t tot_out = lpw * p;
    for (uint32_t e = threadIdx.x; e < tot_out; e += blockDim.x) {
        uint32_t l, u;
        if (sc.st_lf) {
            u = fdiv(e, sc.dlpw);
            l = e - u * lpw;
        } else {
            l = e / p;
            u = e - l * p;
        }
        const uint32_t g = g0 + l;
        if (g >= a.lines) continue;
        LineBase b = line_base(a, g);
        cpx<T> v;
        uint32_t k;
        if (u < p - 1) {
            v = x0[l] + swap_ri(r2[l * lstr + lpad(u)]);
            k = c.perm_out[u];
        } else {
            v = x0[l] + s0[l];
            k = 0;
        }
        store_point<Fmt, ST_COMPLEX, false>(a, b.out, b.j, k, p, v);
    }
}

template <class Fmt>
__global__ void __launch_bounds__(256) k_bluestein(ConvArgs c) {
    using T = typename Fmt::T;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const Sched& sc = c.sc;
    const LineArgs& a = c.io;
    const uint32_t M = c.m, n = c.n, lpw = sc.lpw;
    const uint32_t lstr = lds_line_stride(M);
    cpx<T>* bufA = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* bufB = bufA + lpw * lstr;
    const uint32_t bid = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t g0 = bid * lpw;
    const DivMagic dM = sc.dn;
    const uint32_t total = lpw * M;
    for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
        uint32_t l, q;
        if (sc.ld_lf) {
            q = fdiv(e, sc.dlpw);
            l = e - q * lpw;
        } else {
            l = fdiv(e, dM);
            q = e - l * M;
        }
        const uint32_t g = g0 + l;
        cpx<T> v = mk<T>(T(0), T(0));
        if (g < a.lines && q < n) {
            LineBase b = line_base(a, g);
            v = cmul(load_point<Fmt, LD_COMPLEX>(a, b.in, q, n), ld_tw<T>(c.chirp, q));
        }
        bufA[l * lstr + lpad(q)] = v;
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
    const uint32_t tot_out = lpw * n;
    for (uint32_t e = threadIdx.x; e < tot_out; e += blockDim.x) {
        uint32_t l, k;
        if (sc.st_lf) {
            k = fdiv(e, sc.dlpw);
            l = e - k * lpw;
        } else {
            l = e / n;
            k = e - l * n;
        }
        const uint32_t g = g0 + l;
        if (g >= a.lines) continue;
        LineBase b = line_base(a, g);
        cpx<T> v = cmul(swap_ri(r2[l * lstr + lpad(k)]), ld_tw<T>(c.chirp, k));
        store_point<Fmt, ST_COMPLEX, false>(a, b.out, b.j, k, n, v);
    }
}

size_t conv_kernel_bytes(int prec, uint32_t m, uint32_t lpw) {
    const size_t es = prec == P_F32 ? 8 : 16;
    return (2 * (size_t)lpw * lds_line_stride(m) + 2 * lpw) * es;
}

template <class Fmt, class K>
static hipError_t launch_conv(K kern, int prec, const ConvArgs& c, hipStream_t s) {
    const size_t bytes = conv_kernel_bytes(prec, c.m, c.sc.lpw);
    ensure_lds_attr((const void*)kern);
    const uint32_t blocks = (c.io.lines + c.sc.lpw - 1) / c.sc.lpw;
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), bytes, s, c);
    return hipGetLastError();
}

hipError_t launch_rader(int prec, const ConvArgs& c, hipStream_t s) {
    switch (prec) {
        case P_F32: return launch_conv<FmtF32>(k_rader<FmtF32>, prec, c, s);
        case P_F64: return launch_conv<FmtF64>(k_rader<FmtF64>, prec, c, s);
        default: return launch_conv<FmtDF64>(k_rader<FmtDF64>, prec, c, s);
    }
}

hipError_t launch_bluestein(int prec, const ConvArgs& c, hipStream_t s) {
    switch (prec) {
        case P_F32: return launch_conv<FmtF32>(k_bluestein<FmtF32>, prec, c, s);
        case P_F64: return launch_conv<FmtF64>(k_bluestein<FmtF64>, prec, c, s);
        default: return launch_conv<FmtDF64>(k_bluestein<FmtDF64>, prec, c, s);
    }
}

}  // namespace mifft
