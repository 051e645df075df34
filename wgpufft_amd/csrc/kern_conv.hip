// Standalone fused Rader / Bluestein kernels (one tile of `lpw` lines per
// workgroup); the algorithms are in conv_kernels.hpp.
#include "conv_kernels.hpp"

namespace mifft {

// NT threads per workgroup: 256, or 512 where a tile's LDS leaves room for
// one workgroup per CU only (the tile's loops stride by blockDim.x, so a
// line's arithmetic does not depend on NT).
template <class Fmt, bool MED = false, int NT = 256>
__global__ void __launch_bounds__(NT) k_rader(ConvArgs c) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    rader_tile<Fmt, MED>(c, smem_raw, xcd_remap(blockIdx.x, gridDim.x) * c.sc.lpw);
}
// With at least MINB waves per SIMD: the medium-prime tiles that share a CU
// with a second one run 512 threads at 128 registers (four waves per SIMD;
// 256 threads at 175 registers left two; the stages of the largest primes
// spill to scratch at this budget and still gain): N = 4093 over 1 GiB 5.27
// -> 4.38 ms per pair, N = 4241 (radix 53) 0.80 -> 0.67 ms.
template <class Fmt, bool MED, int NT, int MINB>
__global__ void __launch_bounds__(NT, MINB) k_rader_mb(ConvArgs c) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    rader_tile<Fmt, MED>(c, smem_raw, xcd_remap(blockIdx.x, gridDim.x) * c.sc.lpw);
}

template <class Fmt, int NT = 256>
__global__ void __launch_bounds__(NT) k_bluestein(ConvArgs c) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    bluestein_tile<Fmt>(c, smem_raw, xcd_remap(blockIdx.x, gridDim.x) * c.sc.lpw);
}

size_t conv_kernel_bytes(int prec, uint32_t m, uint32_t lpw) {
    const size_t es = prec == P_F32 ? 8 : 16;
    return (2 * (size_t)lpw * lds_line_stride(m) + 2 * lpw) * es;
}

template <class Fmt, class K>
static hipError_t launch_conv(K kern, int prec, const ConvArgs& c, hipStream_t s, uint32_t threads = 256) {
    const size_t bytes = conv_kernel_bytes(prec, c.m, c.sc.lpw);
    hipError_t e = ensure_lds_attr((const void*)kern, bytes);
    if (e != hipSuccess) return e;
    const uint32_t blocks = (c.io.lines + c.sc.lpw - 1) / c.sc.lpw;
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), bytes, s, c);
    return hipGetLastError();
}

hipError_t launch_rader(int prec, const ConvArgs& c, hipStream_t s) {
    // one workgroup per CU (a tile over half the LDS): twice the threads
    const bool wide = conv_kernel_bytes(prec, c.m, c.sc.lpw) > kLdsBudget / 2;
    switch (prec) {
        case P_F32:
            if (sched_has_medium_prime(c.sc))
                return wide ? launch_conv<FmtF32>(k_rader<FmtF32, true, 512>, prec, c, s, 512)
                            : launch_conv<FmtF32>(k_rader_mb<FmtF32, true, 512, 4>, prec, c, s, 512);
            return wide ? launch_conv<FmtF32>(k_rader<FmtF32, false, 512>, prec, c, s, 512)
                        : launch_conv<FmtF32>(k_rader<FmtF32>, prec, c, s);
        case P_F64:
            return wide ? launch_conv<FmtF64>(k_rader<FmtF64, false, 512>, prec, c, s, 512)
                        : launch_conv<FmtF64>(k_rader<FmtF64>, prec, c, s);
        default:
            return wide ? launch_conv<FmtDF64>(k_rader<FmtDF64, false, 512>, prec, c, s, 512)
                        : launch_conv<FmtDF64>(k_rader<FmtDF64>, prec, c, s);
    }
}

hipError_t launch_bluestein(int prec, const ConvArgs& c, hipStream_t s) {
    const bool wide = conv_kernel_bytes(prec, c.m, c.sc.lpw) > kLdsBudget / 2;
    switch (prec) {
        case P_F32:
            return wide ? launch_conv<FmtF32>(k_bluestein<FmtF32, 512>, prec, c, s, 512)
                        : launch_conv<FmtF32>(k_bluestein<FmtF32>, prec, c, s);
        case P_F64:
            return wide ? launch_conv<FmtF64>(k_bluestein<FmtF64, 512>, prec, c, s, 512)
                        : launch_conv<FmtF64>(k_bluestein<FmtF64>, prec, c, s);
        default:
            return wide ? launch_conv<FmtDF64>(k_bluestein<FmtDF64, 512>, prec, c, s, 512)
                        : launch_conv<FmtDF64>(k_bluestein<FmtDF64>, prec, c, s);
    }
}

// Direct DFT of prime lines (src/runtime/direct_prime.rs:25-110), for the
// primes up to direct_max_prime whose Rader convolution would need zero
// padding.  A task computes the output pair (k, p - k) from the symmetric and
// antisymmetric sums s_n = x_n + x_{p-n}, d_n = x_n - x_{p-n}, n = 1..(p-1)/2:
//   X_k     = x_0 + sum s_n cos(2 pi n k / p) - i sum d_n sin(2 pi n k / p),
//   X_{p-k} = x_0 + sum s_n cos(2 pi n k / p) + i sum d_n sin(2 pi n k / p);
// task k = 0 sums the line.  The lines and W_p^t (t < p) sit in LDS; slots and
// tasks run lines-fastest on strided axes so that loads and stores coalesce
// across lines.
__global__ void __launch_bounds__(256) k_direct_prime(LineArgs a, uint32_t p, uint32_t lpw, uint32_t lf,
                                                      const void* wtab) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<float>* w = reinterpret_cast<cpx<float>*>(smem_raw);
    cpx<float>* xs = w + p;
    const uint32_t g0 = xcd_remap(blockIdx.x, gridDim.x) * lpw;
    for (uint32_t t = threadIdx.x; t < p; t += blockDim.x) w[t] = ld_tw<float>(wtab, t);
    const uint32_t slots = lpw * p;
    for (uint32_t s = threadIdx.x; s < slots; s += blockDim.x) {
        const uint32_t l = lf ? s % lpw : s / p, n = lf ? s / lpw : s % p;
        const uint32_t g = g0 + l;
        const LineBase b = line_base(a, g < a.lines ? g : a.lines - 1);
        xs[l * p + n] = load_point<FmtF32, LD_COMPLEX>(a, b.in, n, p);
    }
    __syncthreads();
    const uint32_t h = (p - 1) / 2, tasks = lpw * (h + 1);
    for (uint32_t s = threadIdx.x; s < tasks; s += blockDim.x) {
        const uint32_t l = lf ? s % lpw : s / (h + 1), k = lf ? s / lpw : s % (h + 1);
        const uint32_t g = g0 + l;
        if (g >= a.lines) continue;
        const LineBase b = line_base(a, g);
        const cpx<float>* x = xs + l * p;
        const cpx<float> x0 = x[0];
        if (k == 0) {
            cpx<float> sum = x0;
            for (uint32_t n = 1; n < p; n++) sum = mk<float>(sum.x + x[n].x, sum.y + x[n].y);
            store_point<FmtF32, ST_COMPLEX, false>(a, b.out, b.j, 0, p, sum);
            continue;
        }
        float ar = 0.f, ai = 0.f, br = 0.f, bi = 0.f;
        uint32_t t = 0;
        for (uint32_t n = 1; n <= h; n++) {
            t += k;
            if (t >= p) t -= p;
            const cpx<float> u = x[n], v = x[p - n];
            const float c = w[t].x, sn = -w[t].y;  // W_p^t = cos - i sin
            ar += (u.x + v.x) * c;
            ai += (u.y + v.y) * c;
            br += (u.x - v.x) * sn;
            bi += (u.y - v.y) * sn;
        }
        // -i B = (B.y, -B.x)
        store_point<FmtF32, ST_COMPLEX, false>(a, b.out, b.j, k, p, mk<float>(x0.x + ar + bi, x0.y + ai - br));
        store_point<FmtF32, ST_COMPLEX, false>(a, b.out, b.j, p - k, p, mk<float>(x0.x + ar - bi, x0.y + ai + br));
    }
}

size_t direct_prime_bytes(uint32_t p, uint32_t lpw) { return ((size_t)lpw * p + p) * sizeof(cpx<float>); }

hipError_t launch_direct_prime(const LineArgs& a, uint32_t p, uint32_t lpw, bool lf, const void* wtab,
                               hipStream_t s) {
    const size_t bytes = direct_prime_bytes(p, lpw);
    hipError_t e = ensure_lds_attr((const void*)&k_direct_prime, bytes);
    if (e != hipSuccess) return e;
    const uint32_t blocks = (a.lines + lpw - 1) / lpw;
    k_direct_prime<<<blocks, 256, bytes, s>>>(a, p, lpw, lf ? 1u : 0u, wtab);
    return hipGetLastError();
}

}  // namespace mifft
