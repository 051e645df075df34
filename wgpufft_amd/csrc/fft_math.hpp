// Device-side complex arithmetic, storage formats and in-register DFT
// butterflies for the gfx950 FFT kernels.
//
// Conventions (wgpuFFT src/math/mod.rs:164-167,185): the forward transform
// uses exp(-2*pi*i*k*n/N); the inverse is obtained from the forward kernels by
// exchanging real and imaginary parts on load and on store
// (IDFT(x) = swap(DFT(swap(x)))), so every butterfly below is forward-only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "roots.hpp"

namespace mifft {

template <class T>
struct cpx {
    T x, y;
};

template <class T>
__device__ __forceinline__ cpx<T> mk(T x, T y) {
    cpx<T> r;
    r.x = x;
    r.y = y;
    return r;
}
template <class T>
__device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) {
    return mk<T>(a.x + b.x, a.y + b.y);
}
template <class T>
__device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) {
    return mk<T>(a.x - b.x, a.y - b.y);
}
// a * b
template <class T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <class T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {
    return mk<T>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// -i * a
template <class T>
__device__ __forceinline__ cpx<T> mul_mi(cpx<T> a) {
    return mk<T>(a.y, -a.x);
}
// +i * a
template <class T>
__device__ __forceinline__ cpx<T> mul_pi(cpx<T> a) {
    return mk<T>(-a.y, a.x);
}
template <class T>
__device__ __forceinline__ cpx<T> swap_ri(cpx<T> a) {
    return mk<T>(a.y, a.x);
}
template <class T>
__device__ __forceinline__ cpx<T> scale(cpx<T> a, T s) {
    return mk<T>(a.x * s, a.y * s);
}

// Sign flip by a uniform mask (0 or the sign bit): one XOR, no select.  The
// inverse transform is conj(DFT(conj(x))); the input side conjugates with this
// and the output side folds the sign into its scale factor.
__device__ __forceinline__ float flip_sign(float y, uint32_t m) { return __uint_as_float(__float_as_uint(y) ^ m); }
__device__ __forceinline__ double flip_sign(double y, uint32_t m) {
    return __hiloint2double(__double2hiint(y) ^ (int)m, __double2loint(y));
}
template <class T>
__device__ __forceinline__ cpx<T> conj_by(cpx<T> a, uint32_t m) {
    return mk<T>(a.x, flip_sign(a.y, m));
}
__host__ __device__ __forceinline__ uint32_t sign_mask(uint32_t flag) { return flag ? 0x80000000u : 0u; }

// ------------------------------------------------------------------ formats
// Storage formats of the plan precisions (wgpuFFT README.md:111-123):
//   F32  interleaved [re, im] f32        -> arithmetic in f32
//   F64  interleaved [re, im] f64        -> arithmetic in f64
//   DF64 [re_hi, re_lo, im_hi, im_lo] f32, value = f64(hi) + f64(lo)
//        -> arithmetic in native f64; split on store hi = f32(v), lo = f32(v - hi)
//        (src/math/double_float.rs:20-31).
// Line data is read once per pass: `ld` carries the non-temporal hint,
// `ld_c` is a plain load.  Measured on MI355X: the hint helps streams of
// whole 128-byte rows (contiguous lines) and hurts narrower strided segments
// (a 64-byte segment per row pass ran 36 % slower); non-temporal stores are
// slower everywhere (they give up the MALL reuse of a pass's output by the
// next pass).
typedef float vf2_t __attribute__((ext_vector_type(2)));
typedef float vf4_t __attribute__((ext_vector_type(4)));
typedef double vd2_t __attribute__((ext_vector_type(2)));

struct FmtF32 {
    using T = float;
    static constexpr int kId = 0;
    static __device__ __forceinline__ cpx<float> ld(const void* p, uint64_t i) {
        const vf2_t v = __builtin_nontemporal_load(reinterpret_cast<const vf2_t*>(p) + i);
        return mk<float>(v.x, v.y);
    }
    static __device__ __forceinline__ cpx<float> ld_c(const void* p, uint64_t i) {
        const vf2_t v = reinterpret_cast<const vf2_t*>(p)[i];
        return mk<float>(v.x, v.y);
    }
    static __device__ __forceinline__ void st(void* p, uint64_t i, cpx<float> v) {
        reinterpret_cast<float2*>(p)[i] = make_float2(v.x, v.y);
    }
    static __device__ __forceinline__ float ld_real(const void* p, uint64_t i) {
        return reinterpret_cast<const float*>(p)[i];
    }
    static __device__ __forceinline__ void st_real(void* p, uint64_t i, float v) {
        reinterpret_cast<float*>(p)[i] = v;
    }
};

struct FmtF64 {
    using T = double;
    static constexpr int kId = 1;
    static __device__ __forceinline__ cpx<double> ld(const void* p, uint64_t i) {
        const vd2_t v = __builtin_nontemporal_load(reinterpret_cast<const vd2_t*>(p) + i);
        return mk<double>(v.x, v.y);
    }
    static __device__ __forceinline__ cpx<double> ld_c(const void* p, uint64_t i) {
        const vd2_t v = reinterpret_cast<const vd2_t*>(p)[i];
        return mk<double>(v.x, v.y);
    }
    static __device__ __forceinline__ void st(void* p, uint64_t i, cpx<double> v) {
        reinterpret_cast<double2*>(p)[i] = make_double2(v.x, v.y);
    }
    static __device__ __forceinline__ double ld_real(const void* p, uint64_t i) {
        return reinterpret_cast<const double*>(p)[i];
    }
    static __device__ __forceinline__ void st_real(void* p, uint64_t i, double v) {
        reinterpret_cast<double*>(p)[i] = v;
    }
};

struct FmtDF64 {
    using T = double;
    static constexpr int kId = 2;
    static __device__ __forceinline__ double join(float hi, float lo) {
        return (double)hi + (double)lo;
    }
    static __device__ __forceinline__ float2 split(double v) {
        float hi = (float)v;
        float lo = (float)(v - (double)hi);
        return make_float2(hi, lo);
    }
    static __device__ __forceinline__ cpx<double> ld(const void* p, uint64_t i) {
        const vf4_t v = __builtin_nontemporal_load(reinterpret_cast<const vf4_t*>(p) + i);
        return mk<double>(join(v.x, v.y), join(v.z, v.w));
    }
    static __device__ __forceinline__ cpx<double> ld_c(const void* p, uint64_t i) {
        const vf4_t v = reinterpret_cast<const vf4_t*>(p)[i];
        return mk<double>(join(v.x, v.y), join(v.z, v.w));
    }
    static __device__ __forceinline__ void st(void* p, uint64_t i, cpx<double> v) {
        float2 a = split(v.x), b = split(v.y);
        reinterpret_cast<float4*>(p)[i] = make_float4(a.x, a.y, b.x, b.y);
    }
    static __device__ __forceinline__ double ld_real(const void* p, uint64_t i) {
        float2 v = reinterpret_cast<const float2*>(p)[i];
        return join(v.x, v.y);
    }
    static __device__ __forceinline__ void st_real(void* p, uint64_t i, double v) {
        reinterpret_cast<float2*>(p)[i] = split(v);
    }
};

// Twiddle tables are stored in the arithmetic type (cpx<float> for F32,
// cpx<double> for F64/DF64), each entry rounded once from an FP64 value.
template <class T>
__device__ __forceinline__ cpx<T> ld_tw(const void* tab, uint32_t i) {
    return reinterpret_cast<const cpx<T>*>(tab)[i];
}

// ------------------------------------------------------- constant twiddles
// v * W_R^t with W_R = exp(-2*pi*i/R), t a compile-time constant; the
// multiples of pi/4 are applied without rounding error.
template <int R, int t, class T>
__device__ __forceinline__ cpx<T> twc(cpx<T> v) {
    constexpr int tt = ((t % R) + R) % R;
    if constexpr (tt == 0) {
        return v;
    } else if constexpr ((8 * tt) % R == 0) {
        constexpr int o = 8 * tt / R;  // angle = -o * pi/4
        constexpr T h = (T)0.707106781186547524400844362104849039;
        if constexpr (o == 2) return mul_mi(v);
        else if constexpr (o == 4) return mk<T>(-v.x, -v.y);
        else if constexpr (o == 6) return mul_pi(v);
        else if constexpr (o == 1) return mk<T>((v.x + v.y) * h, (v.y - v.x) * h);
        else if constexpr (o == 3) return mk<T>((v.y - v.x) * h, -(v.x + v.y) * h);
        else if constexpr (o == 5) return mk<T>(-(v.x + v.y) * h, (v.x - v.y) * h);
        else return mk<T>((v.x - v.y) * h, (v.x + v.y) * h);  // o == 7
    } else {
        constexpr T c = (T)root_cos<R>(tt);
        constexpr T s = (T)root_sin<R>(tt);
        // (x + iy)(c - is)
        return mk<T>(v.x * c + v.y * s, v.y * c - v.x * s);
    }
}

// ------------------------------------------------------------ butterflies
template <int R, class T>
struct Dft;

template <class T>
struct Dft<1, T> {
    static __device__ __forceinline__ void run(cpx<T>*) {}
};

template <class T>
struct Dft<2, T> {
    static __device__ __forceinline__ void run(cpx<T>* v) {
        cpx<T> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    }
};

template <class T>
struct Dft<4, T> {
    static __device__ __forceinline__ void run(cpx<T>* v) {
        cpx<T> a0 = v[0] + v[2], a1 = v[0] - v[2];
        cpx<T> a2 = v[1] + v[3], a3 = mul_mi(v[1] - v[3]);
        v[0] = a0 + a2;
        v[2] = a0 - a2;
        v[1] = a1 + a3;
        v[3] = a1 - a3;
    }
};

// Odd prime R by symmetric pairs: with a_j = v_j + v_{R-j}, b_j = v_j - v_{R-j},
// X_k = v_0 + sum_j cos(2pi jk/R) a_j - i sum_j sin(2pi jk/R) b_j and X_{R-k} the
// same with +i; (R-1)^2/2 real multiply pairs instead of (R-1)^2.
template <int R, class T>
struct DftOdd {
    static constexpr int H = (R - 1) / 2;
    template <int k, int j>
    static __device__ __forceinline__ void acc(cpx<T>& A, cpx<T>& B, const cpx<T>* a,
                                               const cpx<T>* b) {
        if constexpr (j <= H) {
            constexpr int m = (j * k) % R;
            constexpr T c = (T)root_cos<R>(m);
            constexpr T s = (T)root_sin<R>(m);
            A.x += c * a[j - 1].x;
            A.y += c * a[j - 1].y;
            B.x += s * b[j - 1].x;
            B.y += s * b[j - 1].y;
            acc<k, j + 1>(A, B, a, b);
        }
    }
    template <int k>
    static __device__ __forceinline__ void outk(cpx<T>* v, cpx<T> v0, const cpx<T>* a,
                                                const cpx<T>* b) {
        if constexpr (k <= H) {
            cpx<T> A = v0, B = mk<T>(T(0), T(0));
            acc<k, 1>(A, B, a, b);
            v[k] = mk<T>(A.x + B.y, A.y - B.x);
            v[R - k] = mk<T>(A.x - B.y, A.y + B.x);
            outk<k + 1>(v, v0, a, b);
        }
    }
    static __device__ __forceinline__ void run(cpx<T>* v) {
        cpx<T> a[H], b[H];
        cpx<T> v0 = v[0];
        cpx<T> sum = v0;
#pragma unroll
        for (int j = 1; j <= H; j++) {
            a[j - 1] = v[j] + v[R - j];
            b[j - 1] = v[j] - v[R - j];
            sum = sum + a[j - 1];
        }
        outk<1>(v, v0, a, b);
        v[0] = sum;
    }
};

template <class T>
struct Dft<3, T> : DftOdd<3, T> {};
template <class T>
struct Dft<5, T> : DftOdd<5, T> {};
template <class T>
struct Dft<7, T> : DftOdd<7, T> {};
template <class T>
struct Dft<11, T> : DftOdd<11, T> {};
template <class T>
struct Dft<13, T> : DftOdd<13, T> {};

// Composite R = R1 * R2 in registers: n = R2*n1 + n2, k = k1 + R1*k2.
// X[k1 + R1 k2] = sum_n2 W_R2^(n2 k2) W_R^(n2 k1) sum_n1 x[R2 n1 + n2] W_R1^(n1 k1).
template <int R1, int R2, class T>
struct DftComp {
    static constexpr int R = R1 * R2;
    template <int n2>
    static __device__ __forceinline__ void first(const cpx<T>* v, cpx<T> (*y)[R1]) {
        if constexpr (n2 < R2) {
#pragma unroll
            for (int n1 = 0; n1 < R1; n1++) y[n2][n1] = v[R2 * n1 + n2];
            Dft<R1, T>::run(y[n2]);
            tw<n2, 1>(y[n2]);
            first<n2 + 1>(v, y);
        }
    }
    template <int n2, int k1>
    static __device__ __forceinline__ void tw(cpx<T>* row) {
        if constexpr (k1 < R1) {
            row[k1] = twc<R, n2 * k1>(row[k1]);
            tw<n2, k1 + 1>(row);
        }
    }
    static __device__ __forceinline__ void run(cpx<T>* v) {
        cpx<T> y[R2][R1];
        first<0>(v, y);
#pragma unroll
        for (int k1 = 0; k1 < R1; k1++) {
            cpx<T> z[R2];
#pragma unroll
            for (int n2 = 0; n2 < R2; n2++) z[n2] = y[n2][k1];
            Dft<R2, T>::run(z);
#pragma unroll
            for (int k2 = 0; k2 < R2; k2++) v[k1 + R1 * k2] = z[k2];
        }
    }
};

template <class T>
struct Dft<8, T> : DftComp<2, 4, T> {};
template <class T>
struct Dft<16, T> : DftComp<4, 4, T> {};
template <class T>
struct Dft<32, T> : DftComp<4, 8, T> {};
template <class T>
struct Dft<64, T> : DftComp<8, 8, T> {};
template <class T>
struct Dft<6, T> : DftComp<2, 3, T> {};
template <class T>
struct Dft<9, T> : DftComp<3, 3, T> {};
template <class T>
struct Dft<10, T> : DftComp<2, 5, T> {};
template <class T>
struct Dft<12, T> : DftComp<4, 3, T> {};
template <class T>
struct Dft<14, T> : DftComp<2, 7, T> {};
template <class T>
struct Dft<15, T> : DftComp<3, 5, T> {};

// ---------------------------------------------------------- fast division
// floor(x / d) for x < 2^31 with a host-computed magic (m, s):
// m = ceil(2^(31+l) / d), s = 31 + l, l = ceil(log2 d); d = 1 gives m = 2^31, s = 31.
struct DivMagic {
    uint32_t m;
    uint32_t s;
};
__device__ __forceinline__ uint32_t fdiv(uint32_t x, DivMagic d) {
    return (uint32_t)(((uint64_t)x * d.m) >> d.s);
}

}  // namespace mifft
