// The host FFT engine (cpu_fft_core.inc, both builds: x86-64 baseline and
// AVX2 + FMA) run under AddressSanitizer and UndefinedBehaviorSanitizer over
// the shapes the CPU plans take (src/cpu.rs:88-361): smooth, prime (Rader-
// and Bluestein-sized), four-step (>= 2^15) and strided axes, axis subsets,
// batches, C2C in f32 / f64 / df64, R2C + C2R in f32 / f64; in-place C2C.
// Checks: forward then inverse returns the input, small 1D lines match a
// direct DFT (f32 bound 2e-4 of the peak, src/cpu.rs:914-1221), and in-place
// equals out-of-place bit for bit.  Exit status 0 when every check passes.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>
#include "cpu_fft.hpp"
using namespace mifft;
typedef void* (*CreateFn)(int, int, int, const uint64_t*, const int*, int, uint64_t, int, double);
typedef void (*RunFn)(void*, const void*, void*);
typedef void (*DestroyFn)(void*);
struct Eng { CreateFn c; RunFn r; DestroyFn d; const char* name; };
static Eng engs[2] = {{cpu_base::create, cpu_base::run, cpu_base::destroy, "base"},
                      {cpu_avx2::create, cpu_avx2::run, cpu_avx2::destroy, "avx2"}};
static int fails = 0;
static double rel(const std::vector<double>& a, const std::vector<double>& b) {
    double n = 0, d = 0;
    for (size_t i = 0; i < a.size(); i++) { n += (a[i] - b[i]) * (a[i] - b[i]); d += b[i] * b[i]; }
    return d > 0 ? std::sqrt(n / d) : std::sqrt(n);
}
static void c2c(const Eng& e, std::vector<uint64_t> shape, std::vector<int> axes, uint64_t batch, int prec) {
    size_t per = 1; for (auto s : shape) per *= s;
    size_t total = per * batch;
    double sc = 1; for (int a : axes) sc *= shape[a];
    std::mt19937 g(per * 7 + batch);
    std::uniform_real_distribution<double> u(-1, 1);
    std::vector<double> x(2 * total); for (auto& v : x) v = u(g);
    void* f = e.c(0, prec, (int)shape.size(), shape.data(), axes.data(), (int)axes.size(), batch, 0, 1.0);
    void* b = e.c(0, prec, (int)shape.size(), shape.data(), axes.data(), (int)axes.size(), batch, 1, 1.0 / sc);
    std::vector<double> y(2 * total);
    double err;
    if (prec == 0) {
        std::vector<float> xf(x.begin(), x.end()), yf(2 * total), zf(2 * total);
        e.r(f, xf.data(), yf.data());
        // direct DFT of 1D lines
        if (shape.size() == 1 && shape[0] <= 4200) {
            size_t n = shape[0];
            double worst = 0, peak = 0;
            for (uint64_t l = 0; l < batch && l < 2; l++)
                for (size_t k = 0; k < n; k++) {
                    std::complex<double> acc = 0;
                    for (size_t t = 0; t < n; t++)
                        acc += std::complex<double>(xf[2 * (l * n + t)], xf[2 * (l * n + t) + 1]) *
                               std::polar(1.0, -2 * M_PI * (double)((k * t) % n) / n);
                    std::complex<double> got(yf[2 * (l * n + k)], yf[2 * (l * n + k) + 1]);
                    worst = std::max(worst, std::abs(got - acc)); peak = std::max(peak, std::abs(acc));
                }
            if (worst > 2e-4 * peak) { printf("DFT FAIL %s n=%zu err %g peak %g\n", e.name, n, worst, peak); fails++; }
        }
        e.r(b, yf.data(), zf.data());
        std::vector<double> z(zf.begin(), zf.end()), x2(xf.begin(), xf.end());
        err = rel(z, x2);
        // in place
        e.r(f, xf.data(), xf.data());
        for (size_t i = 0; i < 2 * total; i++) if (xf[i] != yf[i]) { printf("INPLACE mismatch %s\n", e.name); fails++; break; }
    } else if (prec == 1) {
        std::vector<double> z(2 * total);
        e.r(f, x.data(), y.data());
        e.r(b, y.data(), z.data());
        err = rel(z, x);
    } else {
        std::vector<float> w(4 * total), o(4 * total), r(4 * total);
        for (size_t i = 0; i < 2 * total; i++) { float hi = (float)x[i]; w[2 * i] = hi; w[2 * i + 1] = (float)(x[i] - hi); }
        e.r(f, w.data(), o.data());
        e.r(b, o.data(), r.data());
        std::vector<double> z(2 * total), x2(2 * total);
        for (size_t i = 0; i < 2 * total; i++) { z[i] = (double)r[2 * i] + r[2 * i + 1]; x2[i] = (double)w[2 * i] + w[2 * i + 1]; }
        err = rel(z, x2);
    }
    double tol = prec == 0 ? 1e-5 : 1e-12;
    if (!(err <= tol)) { printf("RT FAIL %s prec %d shape[0]=%llu rank %zu batch %llu err %g\n", e.name, prec, (unsigned long long)shape[0], shape.size(), (unsigned long long)batch, err); fails++; }
    e.d(f); e.d(b);
}
static void real(const Eng& e, std::vector<uint64_t> shape, uint64_t batch, int prec) {
    size_t per = 1; for (auto s : shape) per *= s;
    size_t total = per * batch, pp = (shape[0] / 2 + 1) * (per / shape[0]) * batch;
    double sc = (double)per;
    std::vector<int> axes; for (int i = 0; i < (int)shape.size(); i++) axes.push_back(i);
    void* f = e.c(1, prec, (int)shape.size(), shape.data(), axes.data(), (int)axes.size(), batch, 0, 1.0);
    void* b = e.c(2, prec, (int)shape.size(), shape.data(), axes.data(), (int)axes.size(), batch, 1, 1.0 / sc);
    std::mt19937 g(per + 3);
    std::uniform_real_distribution<double> u(-1, 1);
    double err;
    if (prec == 0) {
        std::vector<float> x(total), X(2 * pp), z(total);
        for (auto& v : x) v = u(g);
        e.r(f, x.data(), X.data()); e.r(b, X.data(), z.data());
        std::vector<double> a(z.begin(), z.end()), c(x.begin(), x.end()); err = rel(a, c);
    } else {
        std::vector<double> x(total), X(2 * pp), z(total);
        for (auto& v : x) v = u(g);
        e.r(f, x.data(), X.data()); e.r(b, X.data(), z.data()); err = rel(z, x);
    }
    double tol = prec == 0 ? 1e-5 : 1e-12;
    if (!(err <= tol)) { printf("REAL FAIL %s prec %d n0=%llu rank %zu err %g\n", e.name, prec, (unsigned long long)shape[0], shape.size(), err); fails++; }
    e.d(f); e.d(b);
}
int main(int argc, char** argv) {
    const uint64_t lens[] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 13, 16, 17, 23, 31, 32, 47, 59, 61, 64, 97, 120, 127, 128,
                             243, 343, 625, 1000, 1001, 1009, 1024, 2003, 2048, 3256, 4093, 4096, 6841};
    for (int ei = 0; ei < 2; ei++) {
        if (ei == 1 && !cpu_has_avx2_fma()) continue;
        const Eng& e = engs[ei];
        for (uint64_t n : lens)
            for (int prec = 0; prec < 3; prec++) c2c(e, {n}, {0}, n < 256 ? 5 : 2, prec);
        // four-step lines (>= 2^15, fewer than 16 lines) and long Bluestein
        c2c(e, {1u << 15}, {0}, 3, 0); c2c(e, {1u << 16}, {0}, 1, 1); c2c(e, {40009}, {0}, 1, 0);
        c2c(e, {3 * (1u << 15)}, {0}, 2, 0);
        // multi-dimensional, strided axes and axis subsets
        c2c(e, {8, 6}, {0, 1}, 3, 0); c2c(e, {17, 5}, {1}, 2, 0); c2c(e, {64, 33}, {0, 1}, 1, 1);
        c2c(e, {5, 7, 9}, {0, 1, 2}, 2, 0); c2c(e, {5, 7, 9}, {2, 0}, 1, 2); c2c(e, {256, 3, 131}, {1, 2}, 1, 0);
        c2c(e, {2, 4099}, {1}, 2, 0); c2c(e, {1000, 70}, {1}, 1, 0); c2c(e, {3, 1 << 15}, {1}, 1, 0);
        for (int prec = 0; prec < 2; prec++) {
            real(e, {8}, 3, prec); real(e, {9}, 2, prec); real(e, {16, 5}, 2, prec); real(e, {33, 4}, 1, prec);
            real(e, {2, 3}, 1, prec); real(e, {1 << 15}, 1, prec); real(e, {1009 * 2}, 2, prec); real(e, {127}, 3, prec);
            real(e, {64, 8, 6}, 1, prec); real(e, {1, 7}, 2, prec);
        }
        printf("%s done\n", e.name);
    }
    printf("fails=%d\n", fails);
    return fails != 0;
}
