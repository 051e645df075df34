// Table generation and elementwise kernels.
//
// Twiddles, Bluestein chirps and the Rader/Bluestein filter sequences are
// generated in FP64 on the device with the phase reduced exactly in integers
// (t*8 mod n for the octant, n^2 mod 2N for the chirp, g^-t mod p for Rader),
// evaluated with the ROCm device-library sincospi, and rounded once to the
// table type.  The reference builds the same values on the host in f64
// (src/runtime/twiddle.rs:198-217, src/runtime/bluestein_axis.rs:818-867,
// src/runtime/rader_axis.rs:1318-1336).
#include <cstring>

#include "conv_kernels.hpp"

namespace mifft {

// W_n^t = exp(-2*pi*i*t/n) in FP64.  The angle theta = 2 pi t / n is reduced
// exactly in integers to an octant: 8t = o n + r (0 <= r < n), theta =
// (pi/4)(o + r/n).  sincospi then only sees x = r/(4n) (even octants) or
// (n - r)/(4n) (odd octants, the complement), both in [0, 1/4], where sin and
// cos keep their relative accuracy (no value near zero is formed by
// cancellation); the octant's symmetry gives (cos theta, sin theta).
__device__ double2 root_f64(uint64_t t, uint64_t n) {
    t %= n;
    const unsigned __int128 t8 = (unsigned __int128)t * 8u;
    const uint64_t o = (uint64_t)(t8 / n);
    const uint64_t r = (uint64_t)(t8 - (unsigned __int128)o * n);
    const bool odd = o & 1u;
    const double x = (double)(odd ? n - r : r) / (4.0 * (double)n);
    double s, c;
    sincospi(x, &s, &c);
    // angle within the quarter: alpha = phi (even octant) or pi/2 - phi' (odd)
    const double ca = odd ? s : c, sa = odd ? c : s;
    double ct, st;
    switch ((o >> 1) & 3u) {
        case 0: ct = ca; st = sa; break;
        case 1: ct = -sa; st = ca; break;
        case 2: ct = -ca; st = -sa; break;
        default: ct = sa; st = -ca; break;
    }
    // + 0.0 and 0.0 - st give exact zeros a positive sign, as the exact value
    // rounded once has (cos(pi/2) = 0, sin(0) = 0)
    return make_double2(ct + 0.0, 0.0 - st);
}

__device__ __forceinline__ uint64_t mod_mul(uint64_t a, uint64_t b, uint64_t m) {
    return (uint64_t)(((unsigned __int128)a * b) % m);
}
__device__ uint64_t mod_pow_dev(uint64_t b, uint64_t e, uint64_t m) {
    uint64_t r = 1 % m;
    b %= m;
    while (e) {
        if (e & 1) r = mod_mul(r, b, m);
        b = mod_mul(b, b, m);
        e >>= 1;
    }
    return r;
}

#define GRID_STRIDE(i, n) \
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

template <class T>
__global__ void k_gen_tw(cpx<T>* out, uint64_t n, uint64_t count) {
    GRID_STRIDE(i, count) {
        double2 w = root_f64(i, n);
        out[i] = mk<T>((T)w.x, (T)w.y);
    }
}

// Stage-twiddle table of a radix schedule (layout in fft_kernels.hpp):
// entry off_s + (m-1)*L_s + k = W_{L_s R_s}^(k m).
struct StageTab {
    uint32_t ns;
    uint32_t L[kMaxStages];
    uint32_t R[kMaxStages];
    uint64_t off[kMaxStages + 1];
};

template <class T>
__global__ void k_gen_stage_tw(cpx<T>* out, StageTab st, uint64_t total) {
    GRID_STRIDE(e, total) {
        uint32_t s = 1;
        while (s + 1 < st.ns && e >= st.off[s + 1]) s++;
        const uint64_t r = e - st.off[s];
        const uint64_t L = st.L[s];
        const uint64_t m = r / L + 1, k = r % L;
        double2 w = root_f64(k * m, L * st.R[s]);
        out[e] = mk<T>((T)w.x, (T)w.y);
    }
}

__global__ void k_gen_tw_f64_strided(double2* out, uint64_t n, uint64_t count, uint64_t mult) {
    GRID_STRIDE(i, count) { out[i] = root_f64(mod_mul(i, mult, n), n); }
}

// Forward chirp c[i] = exp(-pi*i*i^2/n) = W_{2n}^(i^2 mod 2n).
template <class T>
__global__ void k_gen_chirp(cpx<T>* out, uint64_t n) {
    GRID_STRIDE(i, n) {
        const uint64_t ph = mod_mul(i, i, 2 * n);
        double2 w = root_f64(ph, 2 * n);
        out[i] = mk<T>((T)w.x, (T)w.y);
    }
}

// Rader filter sequence over the convolution length m: b[t] = W_p^(ginv^t mod p)
// for t < p-1; when m > p-1 the cyclic sequence is embedded with b[m-t] = b[p-1-t].
__global__ void k_rader_seq(double2* out, uint64_t p, uint64_t m, uint64_t ginv) {
    const uint64_t l = p - 1;
    GRID_STRIDE(idx, m) {
        int64_t t = -1;
        if (idx < l) t = (int64_t)idx;
        else if (m != l && idx > m - l) t = (int64_t)(l - (m - idx));
        if (t < 0) {
            out[idx] = make_double2(0.0, 0.0);
        } else {
            out[idx] = root_f64(mod_pow_dev(ginv, (uint64_t)t, p), p);
        }
    }
}

// Bluestein filter sequence: b[i] = conj(c[i]) for |i| < n, wrapped mod m.
__global__ void k_bluestein_seq(double2* out, uint64_t n, uint64_t m) {
    GRID_STRIDE(idx, m) {
        int64_t t = -1;
        if (idx < n) t = (int64_t)idx;
        else if (idx > m - n) t = (int64_t)(m - idx);
        if (t < 0) {
            out[idx] = make_double2(0.0, 0.0);
        } else {
            const uint64_t ph = mod_mul((uint64_t)t, (uint64_t)t, 2 * n);
            double2 w = root_f64(ph, 2 * n);
            out[idx] = make_double2(w.x, -w.y);
        }
    }
}

template <class T>
__global__ void k_round_table(cpx<T>* out, const double2* in, uint64_t count, double mul) {
    GRID_STRIDE(i, count) {
        double2 v = in[i];
        out[i] = mk<T>((T)(v.x * mul), (T)(v.y * mul));
    }
}

// ----------------------------------------------- multi-pass Bluestein steps
// Workspace layout: line l of the convolution at ws[l*m + k], plan format.
template <class Fmt>
__global__ void k_chirp_pre(LineArgs a, const void* chirp, uint32_t n, uint32_t m) {
    const uint64_t total = (uint64_t)a.lines * m;
    GRID_STRIDE(e, total) {
        const uint32_t l = (uint32_t)(e / m);
        chirp_pre_elem<Fmt>(a, chirp, n, l, (uint32_t)(e - (uint64_t)l * m), e);
    }
}

template <class Fmt>
__global__ void k_pointwise(void* buf, const void* filt, uint64_t m, uint64_t total) {
    GRID_STRIDE(e, total) { pointwise_elem<Fmt>(buf, filt, (uint32_t)(e % m), e); }
}

template <class Fmt>
__global__ void k_chirp_post(LineArgs a, const void* chirp, uint32_t n, uint32_t m) {
    const uint64_t total = (uint64_t)a.lines * n;
    GRID_STRIDE(e, total) {
        const uint32_t l = (uint32_t)(e / n);
        chirp_post_elem<Fmt>(a, chirp, n, m, l, (uint32_t)(e - (uint64_t)l * n));
    }
}

template <class Fmt>
__global__ void k_scale_copy(const void* in, void* out, uint64_t count, double s) {
    using T = typename Fmt::T;
    GRID_STRIDE(e, count) { Fmt::st(out, e, scale(Fmt::ld(in, e), (T)s)); }
}

// ------------------------------------------------ general real transforms
// Real transforms whose axis 0 needs Rader, Bluestein or a four-step split run
// that axis as a complex transform of full lines in the workspace: widen the
// real input (im = 0) before it and keep bins k <= N0/2 after it (R2C), or
// Hermitian-extend the packed lines before it and keep the real part after it
// (C2R).  The extension follows the packed convention of the fused real load:
// the imaginary part of bins 0 and N0/2 is dropped and bin N0-k is conj(bin k)
// (src/math/mod.rs:470-502).
__global__ void k_real_widen(const float* __restrict__ in, float2* __restrict__ out, uint64_t count) {
    GRID_STRIDE(i, count) { out[i] = make_float2(in[i], 0.0f); }
}

__global__ void k_pack(const float2* __restrict__ full, float2* __restrict__ packed, uint32_t n0, uint32_t p0,
                       uint64_t lines, float sc) {
    const uint64_t total = lines * p0;
    GRID_STRIDE(e, total) {
        const uint64_t l = e / p0;
        const uint32_t k = (uint32_t)(e - l * p0);
        const float2 v = full[l * n0 + k];
        packed[e] = make_float2(v.x * sc, v.y * sc);
    }
}

__global__ void k_herm_extend(const float2* __restrict__ packed, float2* __restrict__ full, uint32_t n0,
                              uint32_t p0, uint64_t lines) {
    const uint64_t total = lines * n0;
    GRID_STRIDE(e, total) {
        const uint64_t l = e / n0;
        const uint32_t p = (uint32_t)(e - l * n0);
        const bool up = p > (n0 >> 1);
        float2 v = packed[l * p0 + (up ? n0 - p : p)];
        v.y = (p == 0 || 2 * p == n0) ? 0.0f : (up ? -v.y : v.y);
        full[e] = v;
    }
}

__global__ void k_real_part(const float2* __restrict__ full, float* __restrict__ out, uint64_t count, float sc) {
    GRID_STRIDE(i, count) { out[i] = full[i].x * sc; }
}

// Half-length real transforms (N0 = 2h): a real line read as h complex points
// z[n] = x[2n] + i x[2n+1], with Z = DFT_h(z) and Z[h] = Z[0], gives
//   X[k] = A[k] + W_N0^k B[k],  A = (Z[k] + conj Z[h-k]) / 2,
//   B = (Z[k] - conj Z[h-k]) / (2i)            (k <= h: even / odd samples);
// the inverse builds, from the packed bins (imaginary parts of bins 0 and h
// dropped), Z[k] = E + i O with E = X[k] + conj X[h-k] and
// O = (X[k] - conj X[h-k]) W_N0^{-k} (k < h), whose unnormalized half-length
// inverse is x[2n] + i x[2n+1] of the unnormalized full-length one.
// W_N0^k = hi[k >> sh] * lo[k & mask] in FP64, rounded once.
__device__ __forceinline__ float2 half_twiddle(const double2* hi, const double2* lo, uint32_t sh, uint32_t mask,
                                               uint32_t k) {
    const double2 a = hi[k >> sh], b = lo[k & mask];
    return make_float2((float)(a.x * b.x - a.y * b.y), (float)(a.x * b.y + a.y * b.x));
}

__global__ void k_r2c_post(const float2* __restrict__ z, float2* __restrict__ x, uint32_t h, uint64_t lines, float sc,
                           const double2* __restrict__ hi, const double2* __restrict__ lo, uint32_t sh,
                           uint32_t mask) {
    const uint64_t total = lines * (h + 1);
    GRID_STRIDE(e, total) {
        const uint64_t l = e / (h + 1);
        const uint32_t k = (uint32_t)(e - l * (h + 1));
        const float2 a = z[l * h + (k == h ? 0 : k)];
        const float2 b = z[l * h + (k == 0 ? 0 : h - k)];  // conj(b) = Z[h-k]^*
        const float2 A = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
        const float2 d = make_float2(a.x - b.x, a.y + b.y);  // Z[k] - conj Z[h-k]
        const float2 B = make_float2(0.5f * d.y, -0.5f * d.x);  // d / (2i)
        const float2 w = half_twiddle(hi, lo, sh, mask, k);
        const float2 r = make_float2(A.x + w.x * B.x - w.y * B.y, A.y + w.x * B.y + w.y * B.x);
        x[e] = make_float2(r.x * sc, r.y * sc);
    }
}

__global__ void k_c2r_pre(const float2* __restrict__ x, float2* __restrict__ z, uint32_t h, uint64_t lines,
                          const double2* __restrict__ hi, const double2* __restrict__ lo, uint32_t sh,
                          uint32_t mask) {
    const uint64_t total = lines * h;
    GRID_STRIDE(e, total) {
        const uint64_t l = e / h;
        const uint32_t k = (uint32_t)(e - l * h);
        float2 a = x[l * (h + 1) + k];
        float2 b = x[l * (h + 1) + h - k];
        if (k == 0) {  // bins 0 and h are self-conjugate
            a.y = 0.0f;
            b.y = 0.0f;
        }
        const float2 E = make_float2(a.x + b.x, a.y - b.y);   // X[k] + conj X[h-k]
        const float2 d = make_float2(a.x - b.x, a.y + b.y);   // X[k] - conj X[h-k]
        const float2 w = half_twiddle(hi, lo, sh, mask, k);  // W^k; O = d conj(W^k)
        const float2 O = make_float2(d.x * w.x + d.y * w.y, d.y * w.x - d.x * w.y);
        z[e] = make_float2(E.x - O.y, E.y + O.x);
    }
}

// ------------------------------------------------------------ synthetic data
__device__ __forceinline__ uint64_t splitmix_at(uint64_t seed, uint64_t k) {
    // State after k+1 steps of SplitMix64 (Vigna), mixed.
    uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void k_fill_uniform(float* out, uint64_t n, uint64_t seed) {
    GRID_STRIDE(i, n) {
        const uint64_t x = splitmix_at(seed, i);
        const float u = (float)(x >> 40) * (1.0f / 16777216.0f);
        out[i] = 2.0f * u - 1.0f;
    }
}

__global__ void k_fill_normal(float* out, uint64_t n, uint64_t seed) {
    const uint64_t pairs = (n + 1) / 2;
    GRID_STRIDE(k, pairs) {
        const double u1 = (double)((splitmix_at(seed, 2 * k) >> 11) + 1) * (1.0 / 9007199254740992.0);
        const double u2 = (double)(splitmix_at(seed, 2 * k + 1) >> 11) * (1.0 / 9007199254740992.0);
        const double r = sqrt(-2.0 * log(u1));
        double s, c;
        sincospi(2.0 * u2, &s, &c);
        out[2 * k] = (float)(r * c);
        if (2 * k + 1 < n) out[2 * k + 1] = (float)(r * s);
    }
}

// ------------------------------------------------------------- launchers
static inline dim3 grid_for(uint64_t n, int block = 256) {
    uint64_t g = (n + block - 1) / block;
    if (g > 65536u * 16u) g = 65536u * 16u;
    if (g == 0) g = 1;
    return dim3((unsigned)g);
}

hipError_t gen_twiddles(int prec, void* out, uint64_t n, uint64_t count, hipStream_t s) {
    if (prec == P_F32)
        k_gen_tw<float><<<grid_for(count), 256, 0, s>>>((cpx<float>*)out, n, count);
    else
        k_gen_tw<double><<<grid_for(count), 256, 0, s>>>((cpx<double>*)out, n, count);
    return hipGetLastError();
}

uint64_t stage_twiddle_count(const int* rad, int ns) {
    uint64_t L = 1, total = 0;
    for (int s = 0; s < ns; s++) {
        if (s >= 1) total += (uint64_t)(rad[s] - 1) * L;
        L *= (uint64_t)rad[s];
    }
    return total;
}

hipError_t gen_stage_twiddles(int prec, void* out, const int* rad, int ns, hipStream_t s) {
    if (ns > kMaxStages) return hipErrorInvalidValue;
    StageTab st;
    std::memset(&st, 0, sizeof(st));
    st.ns = (uint32_t)ns;
    uint64_t L = 1, off = 0;
    for (int i = 0; i < ns; i++) {
        st.L[i] = (uint32_t)L;
        st.R[i] = (uint32_t)rad[i];
        st.off[i] = off;
        if (i >= 1) off += (uint64_t)(rad[i] - 1) * L;
        L *= (uint64_t)rad[i];
    }
    st.off[ns] = off;
    if (off == 0) return hipSuccess;
    if (prec == P_F32)
        k_gen_stage_tw<float><<<grid_for(off), 256, 0, s>>>((cpx<float>*)out, st, off);
    else
        k_gen_stage_tw<double><<<grid_for(off), 256, 0, s>>>((cpx<double>*)out, st, off);
    return hipGetLastError();
}

hipError_t gen_twiddles_f64_strided(double2* out, uint64_t n, uint64_t count, uint64_t mult,
                                    hipStream_t s) {
    k_gen_tw_f64_strided<<<grid_for(count), 256, 0, s>>>(out, n, count, mult);
    return hipGetLastError();
}

hipError_t gen_chirp(int prec, void* out, uint64_t n, hipStream_t s) {
    if (prec == P_F32)
        k_gen_chirp<float><<<grid_for(n), 256, 0, s>>>((cpx<float>*)out, n);
    else
        k_gen_chirp<double><<<grid_for(n), 256, 0, s>>>((cpx<double>*)out, n);
    return hipGetLastError();
}

hipError_t gen_rader_seq(double2* out, uint64_t p, uint64_t m, uint64_t ginv, hipStream_t s) {
    k_rader_seq<<<grid_for(m), 256, 0, s>>>(out, p, m, ginv);
    return hipGetLastError();
}

hipError_t gen_bluestein_seq(double2* out, uint64_t n, uint64_t m, hipStream_t s) {
    k_bluestein_seq<<<grid_for(m), 256, 0, s>>>(out, n, m);
    return hipGetLastError();
}

hipError_t round_table(int prec, void* out, const double2* in, uint64_t count, double mul,
                       hipStream_t s) {
    if (prec == P_F32)
        k_round_table<float><<<grid_for(count), 256, 0, s>>>((cpx<float>*)out, in, count, mul);
    else
        k_round_table<double><<<grid_for(count), 256, 0, s>>>((cpx<double>*)out, in, count, mul);
    return hipGetLastError();
}

// The multi-pass Bluestein workspace is F32 for f32 plans and F64 otherwise.
hipError_t launch_chirp_pre(int prec, const LineArgs& a, const void* chirp, uint32_t n,
                            uint32_t m, hipStream_t s) {
    const uint64_t total = (uint64_t)a.lines * m;
    if (prec == P_F32) k_chirp_pre<FmtF32><<<grid_for(total), 256, 0, s>>>(a, chirp, n, m);
    else if (prec == P_F64) k_chirp_pre<FmtF64><<<grid_for(total), 256, 0, s>>>(a, chirp, n, m);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_pointwise(int prec, void* buf, const void* filt, uint64_t m, uint64_t lines,
                            hipStream_t s) {
    const uint64_t total = m * lines;
    if (prec == P_F32) k_pointwise<FmtF32><<<grid_for(total), 256, 0, s>>>(buf, filt, m, total);
    else k_pointwise<FmtF64><<<grid_for(total), 256, 0, s>>>(buf, filt, m, total);
    return hipGetLastError();
}

hipError_t launch_chirp_post(int prec, const LineArgs& a, const void* chirp, uint32_t n,
                             uint32_t m, hipStream_t s) {
    const uint64_t total = (uint64_t)a.lines * n;
    if (prec == P_F32) k_chirp_post<FmtF32><<<grid_for(total), 256, 0, s>>>(a, chirp, n, m);
    else if (prec == P_F64) k_chirp_post<FmtF64><<<grid_for(total), 256, 0, s>>>(a, chirp, n, m);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_real_widen(const void* in, void* out, uint64_t count, hipStream_t s) {
    k_real_widen<<<grid_for(count), 256, 0, s>>>((const float*)in, (float2*)out, count);
    return hipGetLastError();
}

hipError_t launch_pack(const void* full, void* packed, uint32_t n0, uint64_t lines, double sc, hipStream_t s) {
    const uint32_t p0 = n0 / 2 + 1;
    k_pack<<<grid_for(lines * p0), 256, 0, s>>>((const float2*)full, (float2*)packed, n0, p0, lines, (float)sc);
    return hipGetLastError();
}

hipError_t launch_herm_extend(const void* packed, void* full, uint32_t n0, uint64_t lines, hipStream_t s) {
    k_herm_extend<<<grid_for(lines * n0), 256, 0, s>>>((const float2*)packed, (float2*)full, n0, n0 / 2 + 1,
                                                       lines);
    return hipGetLastError();
}

hipError_t launch_r2c_post(const void* z, void* x, uint32_t h, uint64_t lines, double sc, const double2* hi,
                           const double2* lo, uint32_t sh, uint32_t mask, hipStream_t s) {
    k_r2c_post<<<grid_for(lines * (h + 1)), 256, 0, s>>>((const float2*)z, (float2*)x, h, lines, (float)sc, hi, lo,
                                                         sh, mask);
    return hipGetLastError();
}

hipError_t launch_c2r_pre(const void* x, void* z, uint32_t h, uint64_t lines, const double2* hi, const double2* lo,
                          uint32_t sh, uint32_t mask, hipStream_t s) {
    k_c2r_pre<<<grid_for(lines * h), 256, 0, s>>>((const float2*)x, (float2*)z, h, lines, hi, lo, sh, mask);
    return hipGetLastError();
}

hipError_t launch_real_part(const void* full, void* out, uint64_t count, double sc, hipStream_t s) {
    k_real_part<<<grid_for(count), 256, 0, s>>>((const float2*)full, (float*)out, count, (float)sc);
    return hipGetLastError();
}

hipError_t launch_scale_copy(int prec, const void* in, void* out, uint64_t count, double sc,
                             hipStream_t s) {
    if (prec == P_F32) k_scale_copy<FmtF32><<<grid_for(count), 256, 0, s>>>(in, out, count, sc);
    else if (prec == P_F64) k_scale_copy<FmtF64><<<grid_for(count), 256, 0, s>>>(in, out, count, sc);
    else k_scale_copy<FmtDF64><<<grid_for(count), 256, 0, s>>>(in, out, count, sc);
    return hipGetLastError();
}

hipError_t fill_uniform_f32(float* out, uint64_t n, uint64_t seed, hipStream_t s) {
    k_fill_uniform<<<grid_for(n), 256, 0, s>>>(out, n, seed);
    return hipGetLastError();
}

hipError_t fill_normal_f32(float* out, uint64_t n, uint64_t seed, hipStream_t s) {
    k_fill_normal<<<grid_for((n + 1) / 2), 256, 0, s>>>(out, n, seed);
    return hipGetLastError();
}

// One Stockham radix-R stage over HBM, out of place (the reference's
// per-radix pass for axes too long for one workgroup when fuse_long_axes is
// off, src/runtime/axis_plan.rs:3540-3665).  Butterfly j < N/R of line
// (i < S, o) reads points j + m N/R at stride S, multiplies point m by
// W_{L R}^{(j mod L) m} from the stage-twiddle table, and writes point m to
// (j - j mod L) R + (j mod L) + m L.  On a contiguous axis adjacent threads take
// adjacent butterflies of one line; on a strided axis, adjacent lines.
template <class Fmt, int R>
__device__ __forceinline__ void stockham_butterfly(const LineArgs& a, const StockhamStage& st, uint32_t line,
                                                   uint32_t j) {
    using T = typename Fmt::T;
    const LineBase b = line_base(a, line);
    const uint32_t nr = st.n / R;
    cpx<T> w[R];
#pragma unroll
    for (int m = 0; m < R; m++) w[m] = load_point<Fmt, LD_COMPLEX>(a, b.in, j + m * nr, st.n);
    const uint32_t k = fdiv(j, st.dl);
    const uint32_t kk = j - k * st.L;
    if (st.L > 1) {
#pragma unroll
        for (int m = 1; m < R; m++) w[m] = cmul(w[m], ld_tw<T>(a.tw, st.twoff + (m - 1) * st.L + kk));
    }
    Dft<R, T>::run(w);
    const uint32_t ob = k * st.L * R + kk;
#pragma unroll
    for (int m = 0; m < R; m++) store_point<Fmt, ST_COMPLEX, false>(a, b.out, b.j, ob + m * st.L, st.n, w[m]);
}

template <class Fmt>
__global__ void __launch_bounds__(256) k_stockham_pass(LineArgs a, StockhamStage st) {
    const uint32_t nr = st.n / st.R;
    const uint64_t total = (uint64_t)a.lines * nr;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t line, j;
        if (st.lf) {
            j = (uint32_t)(t / a.lines);
            line = (uint32_t)(t - (uint64_t)j * a.lines);
        } else {
            line = (uint32_t)(t / nr);
            j = (uint32_t)(t - (uint64_t)line * nr);
        }
        switch (st.R) {
#define MIFFT_SCASE(RR) \
    case RR: stockham_butterfly<Fmt, RR>(a, st, line, j); break;
            MIFFT_SCASE(2)
            MIFFT_SCASE(3)
            MIFFT_SCASE(4)
            MIFFT_SCASE(5)
            MIFFT_SCASE(6)
            MIFFT_SCASE(7)
            MIFFT_SCASE(8)
            MIFFT_SCASE(9)
            MIFFT_SCASE(10)
            MIFFT_SCASE(11)
            MIFFT_SCASE(12)
            MIFFT_SCASE(13)
            MIFFT_SCASE(14)
            MIFFT_SCASE(15)
            MIFFT_SCASE(16)
#undef MIFFT_SCASE
            default: break;
        }
    }
}

hipError_t launch_stockham_pass(int prec, const LineArgs& a, const StockhamStage& st, hipStream_t s) {
    const uint64_t total = (uint64_t)a.lines * (st.n / st.R);
    switch (prec) {
        case P_F32: k_stockham_pass<FmtF32><<<grid_for(total), 256, 0, s>>>(a, st); break;
        case P_F64: k_stockham_pass<FmtF64><<<grid_for(total), 256, 0, s>>>(a, st); break;
        default: k_stockham_pass<FmtDF64><<<grid_for(total), 256, 0, s>>>(a, st); break;
    }
    return hipGetLastError();
}

// Logical-view copies: element (b, i) of a batch of `per_batch` elements moves
// from src[s.off + b s.sb + i s.se] to dst[d.off + b d.sb + i d.se] (elements
// of E).  Gathers a strided caller layout into a contiguous staging array and
// scatters one back (execute_logical_views, src/runtime/c2c.rs:1696-1860).
struct CopyLayout {
    uint64_t off, se, sb;
};
template <class E, bool WIDE>
__global__ void k_strided_copy(const E* __restrict__ src, E* __restrict__ dst, uint32_t per_batch, DivMagic dp,
                               uint64_t total, CopyLayout sl, CopyLayout dl) {
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t b, i;
        if constexpr (WIDE) {
            b = t / per_batch;
            i = t - b * per_batch;
        } else {
            b = fdiv((uint32_t)t, dp);
            i = t - b * per_batch;
        }
        dst[dl.off + b * dl.sb + i * dl.se] = src[sl.off + b * sl.sb + i * sl.se];
    }
}

template <class E>
static hipError_t strided_copy_as(const void* src, void* dst, uint64_t per_batch, uint64_t batch, const CopyLayout& sl,
                                  const CopyLayout& dl, hipStream_t s) {
    const uint64_t total = per_batch * batch;
    const DivMagic dp = make_magic((uint32_t)per_batch);
    if (total < (1ull << 31))
        k_strided_copy<E, false><<<grid_for(total), 256, 0, s>>>((const E*)src, (E*)dst, (uint32_t)per_batch, dp,
                                                                 total, sl, dl);
    else
        k_strided_copy<E, true><<<grid_for(total), 256, 0, s>>>((const E*)src, (E*)dst, (uint32_t)per_batch, dp,
                                                                total, sl, dl);
    return hipGetLastError();
}

hipError_t launch_strided_copy(uint32_t elem_bytes, const void* src, void* dst, uint64_t per_batch, uint64_t batch,
                               uint64_t s_off, uint64_t s_se, uint64_t s_sb, uint64_t d_off, uint64_t d_se,
                               uint64_t d_sb, hipStream_t s) {
    if (per_batch == 0 || batch == 0) return hipSuccess;
    if (per_batch > 0xFFFFFFFFull) return hipErrorInvalidValue;
    const CopyLayout sl{s_off, s_se, s_sb}, dl{d_off, d_se, d_sb};
    switch (elem_bytes) {
        case 4: return strided_copy_as<uint32_t>(src, dst, per_batch, batch, sl, dl, s);
        case 8: return strided_copy_as<uint2>(src, dst, per_batch, batch, sl, dl, s);
        case 16: return strided_copy_as<uint4>(src, dst, per_batch, batch, sl, dl, s);
        default: return hipErrorInvalidValue;
    }
}

DivMagic make_magic(uint32_t d) {
    DivMagic r;
    if (d <= 1) {
        r.m = 1u << 31;
        r.s = 31;
        return r;
    }
    uint32_t l = 0;
    while ((1ull << l) < d) l++;
    const unsigned __int128 num = (unsigned __int128)1 << (31 + l);
    r.m = (uint32_t)((num + d - 1) / d);
    r.s = 31 + l;
    return r;
}

}  // namespace mifft
