// Standalone fused Rader / Bluestein kernels (one tile of `lpw` lines per
// workgroup); the algorithms are in conv_kernels.hpp.
#include "conv_kernels.hpp"

namespace mifft {

template <class Fmt>
__global__ void __launch_bounds__(256) k_rader(ConvArgs c) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    rader_tile<Fmt>(c, smem_raw, xcd_remap(blockIdx.x, gridDim.x) * c.sc.lpw);
}

template <class Fmt>
__global__ void __launch_bounds__(256) k_bluestein(ConvArgs c) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    bluestein_tile<Fmt>(c, smem_raw, xcd_remap(blockIdx.x, gridDim.x) * c.sc.lpw);
}

size_t conv_kernel_bytes(int prec, uint32_t m, uint32_t lpw) {
    const size_t es = prec == P_F32 ? 8 : 16;
    return (2 * (size_t)lpw * lds_line_stride(m) + 2 * lpw) * es;
}

template <class Fmt, class K>
static hipError_t launch_conv(K kern, int prec, const ConvArgs& c, hipStream_t s) {
    const size_t bytes = conv_kernel_bytes(prec, c.m, c.sc.lpw);
    hipError_t e = ensure_lds_attr((const void*)kern, bytes);
    if (e != hipSuccess) return e;
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
