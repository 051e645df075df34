/*
 * oracle.c -- FP64 CPU restatement of wgpuFFT's host reference transforms.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product path (the HIP library,
 * the C ABI, the Python mirror of FftConfig/FftPlan) links, loads or calls
 * this file.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py use it, and only as the reference results are compared against.
 *
 * What it restates (file:line in the wgpuFFT sources):
 *   oracle_fft1d            src/math/mod.rs:274-364  fft_f64 / fft_f64_recursive
 *                           (smallest prime factor first, direct DFT for primes
 *                           <= 64, Bluestein above)
 *   bluestein (static)      src/math/mod.rs:366-405  bluestein_f64 (pow2 M >= 2n-1,
 *                           chirp phase from (k*k mod 2n))
 *   oracle_c2c_nd           src/math/mod.rs:214-268  reference_c2c_nd_f64 (axes in
 *                           config order, scale 1/prod(shape) on the last axis only)
 *   oracle_r2c_nd           src/math/mod.rs:81-112   reference_r2c_packed_interleaved
 *   oracle_c2r_nd           src/math/mod.rs:114-153  reference_c2r_from_packed_interleaved
 *                           (conjugate mirror, imaginary part of self-conjugate
 *                           bins dropped, real part of the full inverse)
 *   line_base_for_axis      src/math/mod.rs:506-516
 *   packed coordinates      src/math/mod.rs:470-502
 *   oracle_factor_supported src/runtime/mod.rs:29-54 (greedy [13,11,8,7,5,4,3,2])
 *   oracle_is_prime etc.    src/runtime/axis_policy.rs:83-177
 *   oracle_rader_conv_len   src/runtime/rader_axis.rs:1270-1289
 *   oracle_rader_perm       src/runtime/rader_axis.rs:1301-1304
 *   oracle_bluestein_*      src/runtime/bluestein_axis.rs:858-867,1035-1041
 *   oracle_df64_split/join  src/cpu.rs:489-517, src/math/double_float.rs:20-31
 *   oracle_splitmix64_*     input generator shared with the tests (Vigna's
 *                           SplitMix64; uniform = (x>>40)*2^-24*2-1)
 *
 * Threading: oracle_c2c_nd and friends split the lines of each axis across
 * `nthreads` POSIX threads; results do not depend on the thread count.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    double re, im;
} cplx;

static inline cplx cmul(cplx a, cplx b) {
    cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}

static const double TAU = 6.283185307179586476925286766559;
static const double PI_D = 3.14159265358979323846264338327950;

#define MAX_DIRECT_PRIME 64

/* ---------------------------------------------------------------- 1D FFT */

static void fft_rec(const cplx *in, size_t stride, cplx *out, size_t n, const cplx *roots,
                    size_t root_stride, int sign);

static size_t next_pow2(size_t v) {
    size_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

/* roots[k] = exp(sign * i * tau * k / n), k < n (math/mod.rs:282-287). */
static cplx *make_roots(size_t n, int sign) {
    cplx *roots = (cplx *)malloc(sizeof(cplx) * (n ? n : 1));
    for (size_t k = 0; k < n; k++) {
        double a = (double)sign * TAU * (double)k / (double)n;
        roots[k].re = cos(a);
        roots[k].im = sin(a);
    }
    return roots;
}

static void fft_any(const cplx *in, cplx *out, size_t n, int sign) {
    if (n <= 1) {
        if (n == 1) out[0] = in[0];
        return;
    }
    cplx *roots = make_roots(n, sign);
    fft_rec(in, 1, out, n, roots, 1, sign);
    free(roots);
}

/* bluestein_f64, math/mod.rs:366-405. */
static void bluestein(const cplx *x, cplx *out, size_t n, int sign) {
    cplx *chirp = (cplx *)malloc(sizeof(cplx) * n);
    for (size_t k = 0; k < n; k++) {
        unsigned __int128 sq = ((unsigned __int128)k * k) % (2 * (unsigned __int128)n);
        double a = (double)sign * PI_D * (double)sq / (double)n;
        chirp[k].re = cos(a);
        chirp[k].im = sin(a);
    }
    size_t m = next_pow2(2 * n - 1);
    cplx *a = (cplx *)calloc(m, sizeof(cplx));
    cplx *b = (cplx *)calloc(m, sizeof(cplx));
    cplx *fa = (cplx *)malloc(sizeof(cplx) * m);
    cplx *fb = (cplx *)malloc(sizeof(cplx) * m);
    for (size_t k = 0; k < n; k++) {
        a[k] = cmul(x[k], chirp[k]);
        cplx c = {chirp[k].re, -chirp[k].im};
        b[k] = c;
        if (k > 0) b[m - k] = c;
    }
    fft_any(a, fa, m, -1);
    fft_any(b, fb, m, -1);
    for (size_t k = 0; k < m; k++) fa[k] = cmul(fa[k], fb[k]);
    fft_any(fa, a, m, +1);
    for (size_t k = 0; k < n; k++) {
        cplx v = cmul(a[k], chirp[k]);
        out[k].re = v.re / (double)m;
        out[k].im = v.im / (double)m;
    }
    free(chirp);
    free(a);
    free(b);
    free(fa);
    free(fb);
}

/* fft_f64_recursive, math/mod.rs:300-364: transforms in[0], in[stride], ...
 * (n points) into out; roots[k*root_stride] is the k-th root of this length. */
static void fft_rec(const cplx *in, size_t stride, cplx *out, size_t n, const cplx *roots,
                    size_t root_stride, int sign) {
    size_t p = n;
    for (size_t f = 2; f * f <= n; f++) {
        if (n % f == 0) {
            p = f;
            break;
        }
    }
    if (p == n && n > MAX_DIRECT_PRIME) {
        cplx *vals = (cplx *)malloc(sizeof(cplx) * n);
        for (size_t j = 0; j < n; j++) vals[j] = in[j * stride];
        bluestein(vals, out, n, sign);
        free(vals);
        return;
    }
    if (p == n) {
        for (size_t k = 0; k < n; k++) {
            cplx sum = {0.0, 0.0};
            for (size_t j = 0; j < n; j++) {
                cplx t = cmul(in[j * stride], roots[((j * k) % n) * root_stride]);
                sum.re += t.re;
                sum.im += t.im;
            }
            out[k] = sum;
        }
        return;
    }
    size_t m = n / p;
    for (size_t r = 0; r < p; r++)
        fft_rec(in + r * stride, stride * p, out + r * m, m, roots, root_stride * p, sign);
    cplx column[128];
    cplx *col = p <= 128 ? column : (cplx *)malloc(sizeof(cplx) * p);
    cplx *res = (cplx *)malloc(sizeof(cplx) * p);
    for (size_t k = 0; k < m; k++) {
        for (size_t r = 0; r < p; r++) col[r] = cmul(out[r * m + k], roots[r * k * root_stride]);
        for (size_t s = 0; s < p; s++) {
            cplx sum = {0.0, 0.0};
            for (size_t r = 0; r < p; r++) {
                cplx t = cmul(col[r], roots[((r * s) % p) * m * root_stride]);
                sum.re += t.re;
                sum.im += t.im;
            }
            res[s] = sum;
        }
        for (size_t s = 0; s < p; s++) out[k + m * s] = res[s];
    }
    if (col != column) free(col);
    free(res);
}

/* Unnormalised DFT of n interleaved complex doubles; sign -1 forward, +1 inverse. */
void oracle_fft1d(const double *in, double *out, size_t n, int sign) {
    fft_any((const cplx *)in, (cplx *)out, n, sign < 0 ? -1 : 1);
}

/* Brute-force O(n^2) DFT in long double with exact integer phase reduction.
 * Independent of the recursion above; used to pin it on small n. */
void oracle_dft_naive(const double *in, double *out, size_t n, int sign) {
    const long double tau = 6.283185307179586476925286766559005768L;
    for (size_t k = 0; k < n; k++) {
        long double sr = 0.0L, si = 0.0L;
        for (size_t j = 0; j < n; j++) {
            size_t ph = (size_t)(((unsigned __int128)j * k) % n);
            long double a = (long double)(sign < 0 ? -1 : 1) * tau * (long double)ph / (long double)n;
            long double c = cosl(a), s = sinl(a);
            long double xr = in[2 * j], xi = in[2 * j + 1];
            sr += xr * c - xi * s;
            si += xr * s + xi * c;
        }
        out[2 * k] = (double)sr;
        out[2 * k + 1] = (double)si;
    }
}

/* ------------------------------------------------------------- ND driver */

enum { NORM_NONE = 0, NORM_FORWARD = 1, NORM_INVERSE = 2, NORM_ORTHO = 3 };

/* config.rs:262-286 and math/mod.rs:227-234: scale uses prod(shape), not the
 * product of the transformed axes. */
double oracle_scale(const size_t *shape, int rank, int direction_inverse, int normalization) {
    double len = 1.0;
    for (int i = 0; i < rank; i++) len *= (double)shape[i];
    switch (normalization) {
    case NORM_NONE: return 1.0;
    case NORM_FORWARD: return direction_inverse ? 1.0 : 1.0 / len;
    case NORM_INVERSE: return direction_inverse ? 1.0 / len : 1.0;
    case NORM_ORTHO: return 1.0 / sqrt(len);
    default: return 1.0;
    }
}

static size_t line_base_for_axis(size_t line, int axis, const size_t *shape, const size_t *strides,
                                 int rank) {
    size_t base = 0;
    for (int d = 0; d < rank; d++) {
        if (d == axis) continue;
        size_t coord = line % shape[d];
        line /= shape[d];
        base += coord * strides[d];
    }
    return base;
}

typedef struct {
    const cplx *src;
    cplx *dst;
    const size_t *shape;
    const size_t *strides;
    int rank;
    int axis;
    size_t n;
    size_t stride;
    size_t logical;
    size_t lines_total; /* lines_per_batch * batch */
    size_t lines_per_batch;
    size_t first, last;
    int sign;
    double scale;
} axis_job;

static void *axis_worker(void *arg) {
    axis_job *j = (axis_job *)arg;
    cplx *line = (cplx *)malloc(sizeof(cplx) * j->n);
    cplx *spec = (cplx *)malloc(sizeof(cplx) * j->n);
    cplx *roots = make_roots(j->n, j->sign);
    for (size_t g = j->first; g < j->last; g++) {
        size_t b = g / j->lines_per_batch, l = g % j->lines_per_batch;
        size_t base = b * j->logical + line_base_for_axis(l, j->axis, j->shape, j->strides, j->rank);
        for (size_t t = 0; t < j->n; t++) line[t] = j->src[base + t * j->stride];
        if (j->n == 1)
            spec[0] = line[0];
        else
            fft_rec(line, 1, spec, j->n, roots, 1, j->sign);
        for (size_t t = 0; t < j->n; t++) {
            cplx v = {spec[t].re * j->scale, spec[t].im * j->scale};
            j->dst[base + t * j->stride] = v;
        }
    }
    free(line);
    free(spec);
    free(roots);
    return NULL;
}

static int run_axes(cplx *values, size_t total, const size_t *shape, int rank, const int *axes,
                    int naxes, size_t batch, int sign, double final_scale, int nthreads) {
    size_t strides[16];
    if (rank < 1 || rank > 16) return -1;
    strides[0] = 1;
    for (int i = 1; i < rank; i++) strides[i] = strides[i - 1] * shape[i - 1];
    size_t logical = strides[rank - 1] * shape[rank - 1];
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * (total ? total : 1));
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    for (int ai = 0; ai < naxes; ai++) {
        int axis = axes[ai];
        size_t n = shape[axis];
        size_t lpb = logical / n;
        size_t lines = lpb * batch;
        double scale = (ai + 1 == naxes) ? final_scale : 1.0;
        int nt = nthreads;
        if ((size_t)nt > lines) nt = (int)lines;
        if (nt < 1) nt = 1;
        axis_job jobs[256];
        pthread_t tids[256];
        for (int t = 0; t < nt; t++) {
            axis_job *j = &jobs[t];
            j->src = values;
            j->dst = tmp;
            j->shape = shape;
            j->strides = strides;
            j->rank = rank;
            j->axis = axis;
            j->n = n;
            j->stride = strides[axis];
            j->logical = logical;
            j->lines_total = lines;
            j->lines_per_batch = lpb;
            j->first = lines * (size_t)t / (size_t)nt;
            j->last = lines * (size_t)(t + 1) / (size_t)nt;
            j->sign = sign;
            j->scale = scale;
        }
        if (nt == 1) {
            axis_worker(&jobs[0]);
        } else {
            for (int t = 0; t < nt; t++) pthread_create(&tids[t], NULL, axis_worker, &jobs[t]);
            for (int t = 0; t < nt; t++) pthread_join(tids[t], NULL);
        }
        memcpy(values, tmp, sizeof(cplx) * total);
    }
    free(tmp);
    return 0;
}

/* reference_c2c_nd_f64 (math/mod.rs:214-268).  in/out: interleaved complex
 * doubles, prod(shape)*batch elements each, axis 0 contiguous.  in == out is
 * allowed.  Returns 0, or -1 on a rank outside 1..16. */
int oracle_c2c_nd(const double *in, double *out, const size_t *shape, int rank, const int *axes,
                  int naxes, size_t batch, int direction_inverse, int normalization, int nthreads) {
    size_t logical = 1;
    for (int i = 0; i < rank; i++) logical *= shape[i];
    size_t total = logical * batch;
    if (in != out) memcpy(out, in, sizeof(double) * 2 * total);
    return run_axes((cplx *)out, total, shape, rank, axes, naxes, batch, direction_inverse ? 1 : -1,
                    oracle_scale(shape, rank, direction_inverse, normalization), nthreads);
}

static void packed_shape_for(const size_t *shape, int rank, size_t *packed) {
    for (int i = 0; i < rank; i++) packed[i] = shape[i];
    packed[0] = shape[0] / 2 + 1;
}

/* reference_r2c_packed_interleaved (math/mod.rs:81-112), in f64.  real: prod(shape)*batch
 * doubles; packed: prod(packed_shape)*batch interleaved complex doubles. */
int oracle_r2c_nd(const double *real, double *packed, const size_t *shape, int rank, size_t batch,
                  int normalization, int nthreads) {
    if (rank < 1 || rank > 16) return -1;
    size_t logical = 1;
    for (int i = 0; i < rank; i++) logical *= shape[i];
    size_t total = logical * batch;
    cplx *full = (cplx *)malloc(sizeof(cplx) * total);
    for (size_t i = 0; i < total; i++) {
        full[i].re = real[i];
        full[i].im = 0.0;
    }
    int axes[16];
    for (int i = 0; i < rank; i++) axes[i] = i;
    run_axes(full, total, shape, rank, axes, rank, batch, -1,
             oracle_scale(shape, rank, 0, normalization), nthreads);
    size_t pshape[16], pstr[16], fstr[16];
    packed_shape_for(shape, rank, pshape);
    pstr[0] = fstr[0] = 1;
    for (int i = 1; i < rank; i++) {
        pstr[i] = pstr[i - 1] * pshape[i - 1];
        fstr[i] = fstr[i - 1] * shape[i - 1];
    }
    size_t lp = pstr[rank - 1] * pshape[rank - 1];
    cplx *out = (cplx *)packed;
    for (size_t b = 0; b < batch; b++) {
        for (size_t pi = 0; pi < lp; pi++) {
            size_t rem = pi, fi = 0;
            for (int d = rank - 1; d >= 0; d--) {
                size_t c = rem / pstr[d];
                rem %= pstr[d];
                fi += c * fstr[d];
            }
            out[b * lp + pi] = full[b * logical + fi];
        }
    }
    free(full);
    return 0;
}

/* reference_c2r_from_packed_interleaved (math/mod.rs:114-153), in f64. */
int oracle_c2r_nd(const double *packed, double *real, const size_t *shape, int rank, size_t batch,
                  int normalization, int nthreads) {
    if (rank < 1 || rank > 16) return -1;
    size_t logical = 1;
    for (int i = 0; i < rank; i++) logical *= shape[i];
    size_t total = logical * batch;
    size_t pshape[16], pstr[16], fstr[16];
    packed_shape_for(shape, rank, pshape);
    pstr[0] = fstr[0] = 1;
    for (int i = 1; i < rank; i++) {
        pstr[i] = pstr[i - 1] * pshape[i - 1];
        fstr[i] = fstr[i - 1] * shape[i - 1];
    }
    size_t lp = pstr[rank - 1] * pshape[rank - 1];
    const cplx *pk = (const cplx *)packed;
    cplx *full = (cplx *)malloc(sizeof(cplx) * total);
    size_t nx = pshape[0];
    for (size_t b = 0; b < batch; b++) {
        for (size_t fi = 0; fi < logical; fi++) {
            size_t coords[16], rem = fi;
            for (int d = rank - 1; d >= 0; d--) {
                coords[d] = rem / fstr[d];
                rem %= fstr[d];
            }
            size_t x = coords[0];
            size_t pidx = (x >= nx ? shape[0] - x : x) * pstr[0];
            for (int d = 1; d < rank; d++) {
                size_t c = coords[d];
                if (x >= nx && c != 0) c = shape[d] - c;
                pidx += c * pstr[d];
            }
            cplx v = pk[b * lp + pidx];
            if (x >= nx) v.im = -v.im;
            /* is_self_conjugate_bin, math/mod.rs:481-502 */
            int selfc = (x == 0 || (shape[0] % 2 == 0 && x == shape[0] / 2));
            for (int d = 1; d < rank && selfc; d++) {
                if (shape[d] % 2 == 0) {
                    if (coords[d] != 0 && coords[d] != shape[d] / 2) selfc = 0;
                } else if (coords[d] != 0) {
                    selfc = 0;
                }
            }
            if (selfc) v.im = 0.0;
            full[b * logical + fi] = v;
        }
    }
    int axes[16];
    for (int i = 0; i < rank; i++) axes[i] = i;
    run_axes(full, total, shape, rank, axes, rank, batch, +1,
             oracle_scale(shape, rank, 1, normalization), nthreads);
    for (size_t i = 0; i < total; i++) real[i] = full[i].re;
    free(full);
    (void)lp;
    return 0;
}

/* ------------------------------------------------------- integer helpers */

static const size_t FACTOR_ORDER[8] = {13, 11, 8, 7, 5, 4, 3, 2};

/* factor_supported_length (runtime/mod.rs:33-54).  Writes up to cap radices;
 * returns the count, 0 for len 1, or -1 if len is not {2,3,5,7,11,13}-smooth. */
int oracle_factor_supported(size_t len, size_t *out, int cap) {
    if (len == 0) return -1;
    int count = 0;
    size_t rem = len;
    while (rem > 1) {
        int found = 0;
        for (int i = 0; i < 8; i++) {
            if (rem % FACTOR_ORDER[i] == 0) {
                if (count < cap) out[count] = FACTOR_ORDER[i];
                count++;
                rem /= FACTOR_ORDER[i];
                found = 1;
                break;
            }
        }
        if (!found) return -1;
    }
    return count;
}

int oracle_is_prime(size_t n) {
    if (n < 2) return 0;
    if (n % 2 == 0) return n == 2;
    for (size_t d = 3; d <= n / d; d += 2)
        if (n % d == 0) return 0;
    return 1;
}

size_t oracle_mod_pow(size_t base, size_t exp, size_t mod) {
    unsigned __int128 r = 1, b = base % mod;
    while (exp) {
        if (exp & 1) r = r * b % mod;
        b = b * b % mod;
        exp >>= 1;
    }
    return (size_t)r;
}

/* primitive_root_prime (axis_policy.rs:134-153); 0 if n is not prime. */
size_t oracle_primitive_root(size_t p) {
    if (!oracle_is_prime(p)) return 0;
    if (p == 2) return 1;
    size_t phi = p - 1, fs[64];
    int nf = 0;
    size_t n = phi;
    for (size_t d = 2; d <= n / d; d += (d == 2 ? 1 : 2)) {
        if (n % d == 0) {
            fs[nf++] = d;
            while (n % d == 0) n /= d;
        }
    }
    if (n > 1) fs[nf++] = n;
    for (size_t g = 2; g < p; g++) {
        int ok = 1;
        for (int i = 0; i < nf; i++)
            if (oracle_mod_pow(g, phi / fs[i], p) == 1) {
                ok = 0;
                break;
            }
        if (ok) return g;
    }
    return 0;
}

static int smooth_supported(size_t n) {
    static const size_t r[8] = {2, 3, 4, 5, 7, 8, 11, 13};
    if (n == 0) return 0;
    for (int i = 0; i < 8; i++)
        while (n % r[i] == 0) n /= r[i];
    return n == 1;
}

/* next_smooth_at_least (axis_policy.rs:155-161). */
size_t oracle_next_smooth(size_t min) {
    size_t c = min ? min : 1;
    while (!smooth_supported(c)) c++;
    return c;
}

/* rader_convolution_length (rader_axis.rs:1270-1289). */
size_t oracle_rader_conv_len(size_t n) {
    size_t l = n - 1;
    size_t tmp[64];
    if (l >= 2 && oracle_factor_supported(l, tmp, 64) >= 0) return l;
    size_t mn = 2 * l - 1;
    size_t m = oracle_next_smooth(mn);
    if (oracle_factor_supported(m, tmp, 64) < 0) m = next_pow2(mn);
    return m;
}

/* rader_permutation (rader_axis.rs:1301-1304): perm[k] = g^(k+1) mod n. */
int oracle_rader_perm(size_t n, uint32_t *perm) {
    size_t g = oracle_primitive_root(n);
    if (!g) return -1;
    for (size_t k = 0; k + 1 < n; k++) perm[k] = (uint32_t)oracle_mod_pow(g, k + 1, n);
    return 0;
}

/* bluestein_convolution_length (bluestein_axis.rs:1035-1041). */
size_t oracle_bluestein_conv_len(size_t n) { return oracle_next_smooth(2 * n - 1); }

/* The integer part of bluestein_phase (bluestein_axis.rs:858-867): i*i mod 2n. */
uint64_t oracle_bluestein_phase_index(uint64_t n, uint64_t i) {
    return (uint64_t)(((unsigned __int128)i * i) % (2 * (unsigned __int128)n));
}

/* ------------------------------------------------------------- df64 split */

/* DoubleFloat::from_f64 (double_float.rs:20-27): hi = (f32)v, lo = (f32)(v - hi). */
void oracle_df64_split(const double *v, float *hilo, size_t n) {
    for (size_t i = 0; i < n; i++) {
        float hi = (float)v[i];
        float lo = (float)(v[i] - (double)hi);
        hilo[2 * i] = hi;
        hilo[2 * i + 1] = lo;
    }
}

void oracle_df64_join(const float *hilo, double *v, size_t n) {
    for (size_t i = 0; i < n; i++) v[i] = (double)hilo[2 * i] + (double)hilo[2 * i + 1];
}

/* ------------------------------------------------------ input generation */

static inline uint64_t splitmix64_next(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* n floats uniform in [-1,1): u = (x>>40) * 2^-24, v = 2u - 1 (exact in f32). */
void oracle_splitmix64_uniform_f32(uint64_t seed, float *out, size_t n) {
    uint64_t s = seed;
    for (size_t i = 0; i < n; i++) {
        uint64_t x = splitmix64_next(&s);
        float u = (float)(x >> 40) * (1.0f / 16777216.0f);
        out[i] = 2.0f * u - 1.0f;
    }
}

/* n floats ~ N(0,1): Box-Muller in f64 on two 53-bit uniforms, rounded once. */
void oracle_splitmix64_normal_f32(uint64_t seed, float *out, size_t n) {
    uint64_t s = seed;
    for (size_t i = 0; i < n; i += 2) {
        double u1 = (double)((splitmix64_next(&s) >> 11) + 1) * (1.0 / 9007199254740992.0);
        double u2 = (double)(splitmix64_next(&s) >> 11) * (1.0 / 9007199254740992.0);
        double r = sqrt(-2.0 * log(u1));
        out[i] = (float)(r * cos(TAU * u2));
        if (i + 1 < n) out[i + 1] = (float)(r * sin(TAU * u2));
    }
}
