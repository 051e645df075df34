// Compiled register-resident Bluestein kernels (reg_conv.hpp) and their
// launcher: one schedule per power-of-two convolution length.
#include "reg_conv.hpp"

namespace mifft {

hipError_t ensure_lds_attr(const void* fn, size_t bytes);

#define REGB(K, RADSTR) {K::kM, K::kTPL, K::kLines, K::NT, K::lds_bytes(), (const void*)&reg_bluestein_kernel<K>, RADSTR}

static const RegConvVariant kRegBluestein[] = {
    {RegB8192W::kM, RegB8192W::kTPL, RegB8192W::kLines, RegB8192W::NT, RegB8192W::lds_bytes(),
     (const void*)&reg_bluestein_kernel_mb<RegB8192W, 4>, "8x16x8x8"},
    REGB(RegB64, "8x8"),
    REGB(RegB128, "8x2x8"),
    REGB(RegB256, "16x16"),
    REGB(RegB512, "8x8x8"),
    REGB(RegB1024, "16x4x16"),
    REGB(RegB2048, "16x8x16"),
    REGB(RegB4096, "16x16x16"),
    REGB(RegB8192, "8x16x8x8"),
};

const RegConvVariant* find_reg_bluestein(uint32_t m) {
    for (const RegConvVariant& v : kRegBluestein)
        if ((uint32_t)v.m == m) return &v;
    return nullptr;
}

hipError_t launch_reg_bluestein(const RegConvVariant* v, const ConvArgs& c, hipStream_t s) {
    hipError_t e = ensure_lds_attr(v->fn, v->lds_bytes);
    if (e != hipSuccess) return e;
    const uint32_t blocks = (c.io.lines + v->lines - 1) / v->lines;
    ConvArgs copy = c;
    void* args[] = {&copy};
    return hipLaunchKernel(v->fn, dim3(blocks), dim3(v->nthreads), args, v->lds_bytes, s);
}

}  // namespace mifft
