// Small-volume kernel: every axis of an f32 volume of at most 4096 points in
// one workgroup (SmallVolumePlan, src/runtime/small_volume.rs:65-130,215-275).
// A workgroup loads its volume into LDS, and for each axis gathers the axis
// lines into a contiguous line buffer, runs the runtime-schedule Stockham
// stages there (ping-pong in LDS), or a prime axis's direct DFT, and scatters
// them back; then it stores the volume once.  One launch and one HBM round trip instead of one per axis.
#include "launch.h"

namespace mifft {

// Copies n words with kChunk loads in flight per thread (clamped indices).
__device__ __forceinline__ void copy_words(uint32_t* dst, const uint32_t* src, uint32_t n) {
    for (uint32_t t0 = threadIdx.x; t0 < n; t0 += blockDim.x * kChunk) {
        uint32_t r[kChunk];
#pragma unroll
        for (uint32_t u = 0; u < kChunk; u++) {
            const uint32_t t = t0 + u * blockDim.x;
            r[u] = src[t < n ? t : n - 1];
        }
#pragma unroll
        for (uint32_t u = 0; u < kChunk; u++) {
            const uint32_t t = t0 + u * blockDim.x;
            if (t < n) dst[t] = r[u];
        }
    }
}

__global__ void __launch_bounds__(1024) k_small_volume(SmallVolArgs v) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<float>* vol = reinterpret_cast<cpx<float>*>(smem_raw);
    cpx<float>* bufA = vol + v.elems;
    cpx<float>* bufB = bufA + v.buf_elems;
    cpx<float>* twl = bufB + v.buf_elems;                     // every axis's stage twiddles
    Sched* scl = reinterpret_cast<Sched*>(twl + v.tw_total);  // the axis schedules
    const uint32_t base = blockIdx.x * v.elems;
    // Twiddles and schedules into LDS, so that the stage loops read them at
    // LDS latency; these copies and the volume load keep their global loads
    // in flight together.
    copy_words(reinterpret_cast<uint32_t*>(twl), reinterpret_cast<const uint32_t*>(v.tw), 2 * v.tw_total);
    copy_words(reinterpret_cast<uint32_t*>(scl), reinterpret_cast<const uint32_t*>(v.sc),
               v.rank * (uint32_t)(sizeof(Sched) / 4));
    const uint32_t cm = sign_mask(v.conj_in);
    for (uint32_t t0 = threadIdx.x; t0 < v.elems; t0 += blockDim.x * kChunk) {
        cpx<float> r[kChunk];
#pragma unroll
        for (uint32_t u = 0; u < kChunk; u++) {
            const uint32_t t = t0 + u * blockDim.x;
            r[u] = FmtF32::ld(v.in, base + (t < v.elems ? t : v.elems - 1));
        }
#pragma unroll
        for (uint32_t u = 0; u < kChunk; u++) {
            const uint32_t t = t0 + u * blockDim.x;
            if (t < v.elems) vol[t] = conj_by(r[u], cm);
        }
    }
    __syncthreads();
    uint32_t two = 0;
    for (uint32_t ax = 0; ax < v.rank; ax++) {
        const Sched& sc = scl[ax];
        // per-axis constants in registers: the loops below store to LDS, and
        // through generic pointers the compiler would otherwise reload them
        // from memory in every iteration
        const uint32_t n = sc.n, S = v.S[ax], lstr = lds_line_stride(n), lines = v.elems / n;
        const DivMagic dn = sc.dn, dL = v.dL[ax], dS = v.dS[ax];
        const bool direct = sc.ns == 0 && n > 1;
        const cpx<float>* res;
        if (direct) {
            // a prime axis up to direct_max_prime: the direct DFT of each line,
            // X[k] = sum_p x[p] W_n^{pk}, one output per thread, read straight
            // from the volume (point p of line l = (i, o), i < S, at
            // o S n + i + p S; the threads of one line read the same words)
            // with W_n^t (t < n) from LDS (small_volume.rs:220-245)
            const cpx<float>* wt = twl + two;
            for (uint32_t t = threadIdx.x; t < v.elems; t += blockDim.x) {
                const uint32_t l = fdiv(t, dn), k = t - l * n;
                const uint32_t o = fdiv(l, dS), i = l - o * S;
                const cpx<float>* ln = vol + o * S * n + i;
                cpx<float> acc = mk<float>(0.f, 0.f);
                uint32_t e = 0;
                // unrolled so that the LDS reads of several terms are in
                // flight together (the sum's order is unchanged)
#pragma unroll 4
                for (uint32_t q = 0; q < n; q++) {
                    acc = acc + cmul(ln[q * S], wt[e]);
                    e += k;
                    e = e >= n ? e - n : e;
                }
                bufB[l * lstr + lpad(k)] = acc;
            }
            __syncthreads();
            res = bufB;
        } else {
            // Point p of line l = (i, o), i < S, sits at o S n + i + p S.
            // Adjacent threads take adjacent points of axis 0 but adjacent
            // lines (adjacent i) of the other axes, so that they read
            // adjacent volume words: the points of one line are S words
            // apart, the same LDS bank for S a multiple of 32.
            for (uint32_t t = threadIdx.x; t < v.elems; t += blockDim.x) {
                uint32_t l, p;
                if (S == 1) {
                    l = fdiv(t, dn);
                    p = t - l * n;
                } else {
                    p = fdiv(t, dL);
                    l = t - p * lines;
                }
                const uint32_t o = fdiv(l, dS), i = l - o * S;
                bufA[l * lstr + lpad(p)] = vol[o * S * n + i + p * S];
            }
            __syncthreads();
            res = lds_fft_lines<float>(bufA, bufB, lines, sc, twl + two);
        }
        two += v.tw_len[ax];
        // back into the volume, or after the last axis straight out to
        // memory (adjacent threads write adjacent volume words either way)
        const bool last = ax + 1 == v.rank;
        const float sre = v.scale, sim = v.conj_out ? -v.scale : v.scale;
        for (uint32_t t = threadIdx.x; t < v.elems; t += blockDim.x) {
            uint32_t l, p;
            if (S == 1) {
                l = fdiv(t, dn);
                p = t - l * n;
            } else {
                p = fdiv(t, dL);
                l = t - p * lines;
            }
            const uint32_t o = fdiv(l, dS), i = l - o * S;
            const cpx<float> r = res[l * lstr + lpad(p)];
            if (last)
                FmtF32::st(v.out, base + o * S * n + i + p * S, mk<float>(r.x * sre, r.y * sim));
            else
                vol[o * S * n + i + p * S] = r;
        }
        if (!last) __syncthreads();
    }
}

size_t small_volume_lds_bytes(const SmallVolArgs& v) {
    return ((size_t)v.elems + 2 * (size_t)v.buf_elems + v.tw_total) * sizeof(cpx<float>) + v.rank * sizeof(Sched);
}

hipError_t launch_small_volume(const SmallVolArgs& v, uint32_t volumes, hipStream_t s) {
    const size_t lds = small_volume_lds_bytes(v);
    const hipError_t e = ensure_lds_attr((const void*)&k_small_volume, lds);
    if (e != hipSuccess) return e;
    // about four points per thread (one per thread with a direct-DFT axis,
    // whose outputs each take n terms), whole wavefronts, at most 1024 threads
    uint32_t threads = ((v.elems / (v.direct ? 1 : 4) + 63) / 64) * 64;
    threads = threads < 64 ? 64 : threads > 1024 ? 1024 : threads;
    k_small_volume<<<volumes, threads, lds, s>>>(v);
    return hipGetLastError();
}

}  // namespace mifft
