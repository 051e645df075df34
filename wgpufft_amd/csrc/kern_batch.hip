// Persistent batched kernel: many transforms of uneven lengths and batch
// counts in one launch.
//
// The host splits every transform's line passes into tiles (a line range of
// one pass, sized for one 256-thread workgroup) and sorts them by cost.
// Resident workgroups pull tile indices from a device counter and run the
// tile with the device function of its class (register-resident line kernel,
// runtime-schedule LDS kernel, fused Rader or Bluestein, register Bluestein).  A tile computes
// exactly what the same op computes in a standalone plan, so every transform
// is bit-identical alone and in the mix.  Passes that depend on a previous
// pass of the same transform (the next axis, the second half of a four-step
// split, the steps of a multi-pass Bluestein convolution) go into later
// phases, each launched after the previous one on the same stream.
//
// This is the gfx950 counterpart of recording many plans into one shared
// compute pass with FftRecorder (src/plan.rs:38-95, src/runtime/recorder.rs).
#include "batch_kernels.hpp"
#include "reg_conv.hpp"

namespace mifft {

// The tile claimed by the workgroup, broadcast through LDS.
__shared__ uint32_t s_tile;

__device__ __forceinline__ uint32_t claim_tile(uint32_t* counter) {
    if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t t = __builtin_amdgcn_readfirstlane(s_tile);
    __syncthreads();  // the next claim overwrites s_tile
    return t;
}

// Every class is a separate (not inlined) function: inlined into one switch,
// the register allocation of the kernel would be the worst path's, and the
// whole persistent kernel would run at one wave per SIMD.  A call boundary
// waits for every outstanding memory operation, so a class function keeps
// claiming tiles while they are of its own class (tiles are sorted by cost,
// so a class comes in runs): the stores of one tile stay in flight while the
// next tile loads.  It returns the first claimed tile of another class (or
// past the end).
template <class Body>
__device__ __forceinline__ uint32_t tile_run(const BatchTile* __restrict__ tiles, uint32_t ntiles,
                                             uint32_t* counter, const OpDesc* __restrict__ ops, uint32_t t,
                                             Body&& body) {
    t = __builtin_amdgcn_readfirstlane(t);
    const uint32_t cls = tiles[t].cls;
    for (;;) {
        const BatchTile tl = tiles[t];
        body(ops[tl.arg], tl.g0);
        __syncthreads();  // the next tile reuses LDS
        t = claim_tile(counter);
        if (t >= ntiles || tiles[t].cls != cls) return t;
    }
}

#define MIFFT_TILE_ARGS const BatchTile *__restrict__ tiles, uint32_t ntiles, uint32_t *counter, \
                        const OpDesc *__restrict__ ops, unsigned char *smem, uint32_t t
#define MIFFT_TILE_PASS tiles, ntiles, counter, ops, smem, t

template <class K>
__device__ __noinline__ uint32_t reg_tiles(MIFFT_TILE_ARGS) {
    return tile_run(tiles, ntiles, counter, ops, t, [&](const OpDesc& d, uint32_t g0) {
        K::tile(d.a, reinterpret_cast<cpx<typename K::T>*>(smem), g0);
    });
}
template <class Fmt, int LDM, int STM, bool POST>
__device__ __noinline__ uint32_t lds_tiles(MIFFT_TILE_ARGS) {
    return tile_run(tiles, ntiles, counter, ops, t, [&](const OpDesc& d, uint32_t g0) {
        lds_fft_tile<Fmt, LDM, STM, POST>(d.a, d.sc, smem, g0);
    });
}
template <class Fmt, bool MED = false>
__device__ __noinline__ uint32_t rader_tiles(MIFFT_TILE_ARGS) {
    return tile_run(tiles, ntiles, counter, ops, t,
                    [&](const OpDesc& d, uint32_t g0) { rader_tile<Fmt, MED>(d.conv, smem, g0); });
}
template <class K>
__device__ __noinline__ uint32_t regb_tiles(MIFFT_TILE_ARGS) {
    return tile_run(tiles, ntiles, counter, ops, t, [&](const OpDesc& d, uint32_t g0) {
        K::tile(d.conv, reinterpret_cast<cpx<typename K::T>*>(smem), g0);
    });
}
template <class Fmt>
__device__ __noinline__ uint32_t bluestein_tiles(MIFFT_TILE_ARGS) {
    return tile_run(tiles, ntiles, counter, ops, t,
                    [&](const OpDesc& d, uint32_t g0) { bluestein_tile<Fmt>(d.conv, smem, g0); });
}

// Element-wise step KIND (0 chirp pre, 1 filter, 2 chirp post) over elements
// [g0, g0 + kEltTile) of the op's index space (lines x m, or lines x n for the
// chirp post); the host keeps that space below 2^32.
template <class Fmt, int KIND>
__device__ __forceinline__ void elt_tile(const ConvArgs& c, uint32_t g0) {
    const uint32_t per = KIND == 2 ? c.n : c.m;
    const uint32_t total = c.io.lines * per;
    const uint32_t end = total - g0 < kEltTile ? total : g0 + kEltTile;
    for (uint32_t e = g0 + threadIdx.x; e < end; e += 256) {
        const uint32_t l = e / per;
        const uint32_t p = e - l * per;
        if (KIND == 0) chirp_pre_elem<Fmt>(c.io, c.chirp, c.n, l, p, e);
        else if (KIND == 1) pointwise_elem<Fmt>(c.io.out, c.filt, p, e);
        else chirp_post_elem<Fmt>(c.io, c.chirp, c.n, c.m, l, p);
    }
}
template <class Fmt, int KIND>
__device__ __noinline__ uint32_t elt_tiles(MIFFT_TILE_ARGS) {
    return tile_run(tiles, ntiles, counter, ops, t,
                    [&](const OpDesc& d, uint32_t g0) { elt_tile<Fmt, KIND>(d.conv, g0); });
}

#define MIFFT_BK(N, TPL, LINES, LDLF, STLF, POST, VEC, ...) \
    RegFft<FmtF32, N, TPL, LINES, LDLF, STLF, LD_COMPLEX, ST_COMPLEX, POST, RadixList<__VA_ARGS__>, VEC>

// Up to four workgroups per CU (a launch's LDS is the largest of its tiles,
// at most 80 KiB: two per CU then): at most 128 VGPRs per lane.
__global__ void __launch_bounds__(256, 4) k_batch(const BatchTile* __restrict__ tiles, uint32_t ntiles,
                                               uint32_t* counter, const OpDesc* __restrict__ ops) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t t = claim_tile(counter);
    while (t < ntiles) {  // uniform: every wave of the block leaves together
        switch (tiles[t].cls) {
            case BC_LDS: t = lds_tiles<FmtF32, LD_COMPLEX, ST_COMPLEX, false>(MIFFT_TILE_PASS); break;
            case BC_LDS_POST: t = lds_tiles<FmtF32, LD_COMPLEX, ST_COMPLEX, true>(MIFFT_TILE_PASS); break;
            case BC_LDS_R2C: t = lds_tiles<FmtF32, LD_REAL, ST_PACKED, false>(MIFFT_TILE_PASS); break;
            case BC_LDS_C2R: t = lds_tiles<FmtF32, LD_HERMITIAN, ST_REAL, false>(MIFFT_TILE_PASS); break;
            case BC_LDS_F64: t = lds_tiles<FmtF64, LD_COMPLEX, ST_COMPLEX, false>(MIFFT_TILE_PASS); break;
            case BC_LDS_POST_F64: t = lds_tiles<FmtF64, LD_COMPLEX, ST_COMPLEX, true>(MIFFT_TILE_PASS); break;
            case BC_RADER: t = rader_tiles<FmtF32>(MIFFT_TILE_PASS); break;
            case BC_BLUESTEIN: t = bluestein_tiles<FmtF32>(MIFFT_TILE_PASS); break;
            case BC_RADER_F64: t = rader_tiles<FmtF64>(MIFFT_TILE_PASS); break;
            case BC_RADER_MED: t = rader_tiles<FmtF32, true>(MIFFT_TILE_PASS); break;
            case BC_BLUESTEIN_F64: t = bluestein_tiles<FmtF64>(MIFFT_TILE_PASS); break;
            case BC_REG_C256_F64:
                t = reg_tiles<RegFft<FmtF64, 256, 16, 16, false, false, LD_COMPLEX, ST_COMPLEX, false,
                                     RadixList<16, 16>, 1>>(MIFFT_TILE_PASS);
                break;
            case BC_REG_S256_F64:
                t = reg_tiles<RegFft<FmtF64, 256, 16, 16, true, true, LD_COMPLEX, ST_COMPLEX, false,
                                     RadixList<16, 16>, 1>>(MIFFT_TILE_PASS);
                break;
            case BC_CHIRP_PRE: t = elt_tiles<FmtF32, 0>(MIFFT_TILE_PASS); break;
            case BC_CHIRP_PRE_F64: t = elt_tiles<FmtF64, 0>(MIFFT_TILE_PASS); break;
            case BC_POINTWISE: t = elt_tiles<FmtF32, 1>(MIFFT_TILE_PASS); break;
            case BC_POINTWISE_F64: t = elt_tiles<FmtF64, 1>(MIFFT_TILE_PASS); break;
            case BC_CHIRP_POST: t = elt_tiles<FmtF32, 2>(MIFFT_TILE_PASS); break;
            case BC_CHIRP_POST_F64: t = elt_tiles<FmtF64, 2>(MIFFT_TILE_PASS); break;
            // register-resident variants with 256 threads (the registry entries of the same shape)
            case BC_REG_C4096: t = reg_tiles<MIFFT_BK(4096, 256, 1, 0, 0, 0, 1, 16, 16, 16)>(MIFFT_TILE_PASS); break;
            case BC_REG_C2048: t = reg_tiles<MIFFT_BK(2048, 128, 2, 0, 0, 0, 1, 16, 16, 8)>(MIFFT_TILE_PASS); break;
            case BC_REG_C1024: t = reg_tiles<MIFFT_BK(1024, 64, 4, 0, 0, 0, 1, 16, 16, 4)>(MIFFT_TILE_PASS); break;
            case BC_REG_C512: t = reg_tiles<MIFFT_BK(512, 32, 8, 0, 0, 0, 1, 8, 8, 8)>(MIFFT_TILE_PASS); break;
            case BC_REG_C256: t = reg_tiles<MIFFT_BK(256, 16, 16, 0, 0, 0, 1, 16, 16)>(MIFFT_TILE_PASS); break;
            case BC_REG_C128: t = reg_tiles<MIFFT_BK(128, 8, 32, 0, 0, 0, 1, 16, 8)>(MIFFT_TILE_PASS); break;
            case BC_REG_C64: t = reg_tiles<MIFFT_BK(64, 4, 64, 0, 0, 0, 1, 16, 4)>(MIFFT_TILE_PASS); break;
            case BC_REG_C120: t = reg_tiles<MIFFT_BK(120, 8, 32, 0, 0, 0, 1, 15, 8)>(MIFFT_TILE_PASS); break;
            case BC_REG_C343: t = reg_tiles<MIFFT_BK(343, 64, 4, 0, 0, 0, 1, 7, 7, 7)>(MIFFT_TILE_PASS); break;
            case BC_REG_C1001: t = reg_tiles<MIFFT_BK(1001, 128, 2, 0, 0, 0, 1, 13, 11, 7)>(MIFFT_TILE_PASS); break;
            case BC_REG_P256: t = reg_tiles<MIFFT_BK(256, 16, 32, 1, 1, 1, 2, 16, 16)>(MIFFT_TILE_PASS); break;
            case BC_REG_P256V1: t = reg_tiles<MIFFT_BK(256, 16, 16, 1, 1, 1, 1, 16, 16)>(MIFFT_TILE_PASS); break;
            case BC_REG_T256: t = reg_tiles<MIFFT_BK(256, 16, 16, 0, 1, 0, 1, 16, 16)>(MIFFT_TILE_PASS); break;
            case BC_REG_P128: t = reg_tiles<MIFFT_BK(128, 8, 64, 1, 1, 1, 2, 16, 8)>(MIFFT_TILE_PASS); break;
            case BC_REG_T128: t = reg_tiles<MIFFT_BK(128, 8, 32, 0, 1, 0, 1, 16, 8)>(MIFFT_TILE_PASS); break;
            case BC_REG_S1024V2: t = reg_tiles<MIFFT_BK(1024, 64, 8, 1, 1, 0, 2, 16, 16, 4)>(MIFFT_TILE_PASS); break;
            case BC_REG_S512V2: t = reg_tiles<MIFFT_BK(512, 32, 16, 1, 1, 0, 2, 8, 8, 8)>(MIFFT_TILE_PASS); break;
            case BC_REG_S256V2: t = reg_tiles<MIFFT_BK(256, 16, 32, 1, 1, 0, 2, 16, 16)>(MIFFT_TILE_PASS); break;
            case BC_REG_S128V2: t = reg_tiles<MIFFT_BK(128, 8, 64, 1, 1, 0, 2, 16, 8)>(MIFFT_TILE_PASS); break;
            case BC_REG_S64V2: t = reg_tiles<MIFFT_BK(64, 4, 128, 1, 1, 0, 2, 16, 4)>(MIFFT_TILE_PASS); break;
            case BC_REGB_64: t = regb_tiles<RegB64>(MIFFT_TILE_PASS); break;
            case BC_REGB_128: t = regb_tiles<RegB128>(MIFFT_TILE_PASS); break;
            case BC_REGB_256: t = regb_tiles<RegB256>(MIFFT_TILE_PASS); break;
            case BC_REGB_512: t = regb_tiles<RegB512>(MIFFT_TILE_PASS); break;
            case BC_REGB_1024: t = regb_tiles<RegB1024>(MIFFT_TILE_PASS); break;
            case BC_REGB_2048: t = regb_tiles<RegB2048>(MIFFT_TILE_PASS); break;
            case BC_REGB_4096: t = regb_tiles<RegB4096>(MIFFT_TILE_PASS); break;
            case BC_REGB_8192: t = regb_tiles<RegB8192>(MIFFT_TILE_PASS); break;
            default: t = claim_tile(counter); break;  // unknown class: skip the tile
        }
    }
}

__global__ void k_batch_reset(uint32_t* counters, uint32_t n) {
    if (threadIdx.x < n) counters[threadIdx.x] = 0u;
}

hipError_t launch_batch(const BatchTile* tiles, uint32_t ntiles, uint32_t* counter, const OpDesc* ops,
                        uint32_t grid, size_t lds_bytes, hipStream_t s) {
    hipError_t e = ensure_lds_attr((const void*)&k_batch, lds_bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_batch, dim3(grid), dim3(256), lds_bytes, s, tiles, ntiles, counter, ops);
    return hipGetLastError();
}

hipError_t launch_batch_reset(uint32_t* counters, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, counters, n);
    return hipGetLastError();
}

int batch_lds_class(int prec, int ldm, int stm, int post) {
    if (prec == P_F32) {
        if (ldm == LD_COMPLEX && stm == ST_COMPLEX) return post ? BC_LDS_POST : BC_LDS;
        if (!post && ldm == LD_REAL && stm == ST_PACKED) return BC_LDS_R2C;
        if (!post && ldm == LD_HERMITIAN && stm == ST_REAL) return BC_LDS_C2R;
        return -1;
    }
    if (prec == P_F64 && ldm == LD_COMPLEX && stm == ST_COMPLEX) return post ? BC_LDS_POST_F64 : BC_LDS_F64;
    return -1;
}

int batch_class_for(const RegVariant* v) {
    if (!v || v->persist || v->adj || v->nthreads != 256 || v->ldm != LD_COMPLEX || v->stm != ST_COMPLEX) return -1;
    if (v->prec == P_F64 && v->n == 256 && v->vec == 1 && !v->post)
        return v->ld_lf == 0 && v->st_lf == 0 ? BC_REG_C256_F64 : (v->ld_lf && v->st_lf ? BC_REG_S256_F64 : -1);
    if (v->prec != P_F32) return -1;
    const int key = v->n * 1000 + v->ld_lf * 100 + v->st_lf * 10 + v->post;
    if (v->vec == 1) {
        switch (key) {
            case 4096000: return BC_REG_C4096;
            case 2048000: return BC_REG_C2048;
            case 1024000: return BC_REG_C1024;
            case 512000: return BC_REG_C512;
            case 256000: return BC_REG_C256;
            case 128000: return BC_REG_C128;
            case 64000: return BC_REG_C64;
            case 120000: return BC_REG_C120;
            case 343000: return BC_REG_C343;
            case 1001000: return BC_REG_C1001;
            case 256111: return BC_REG_P256V1;
            case 256010: return BC_REG_T256;
            case 128010: return BC_REG_T128;
            default: return -1;
        }
    }
    switch (key) {
        case 256111: return BC_REG_P256;
        case 128111: return BC_REG_P128;
        case 1024110: return BC_REG_S1024V2;
        case 512110: return BC_REG_S512V2;
        case 256110: return BC_REG_S256V2;
        case 128110: return BC_REG_S128V2;
        case 64110: return BC_REG_S64V2;
        default: return -1;
    }
}

}  // namespace mifft
