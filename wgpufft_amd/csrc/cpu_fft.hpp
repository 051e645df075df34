// Entry points of the host-memory FFT engine (cpu_fft_core.inc), compiled
// twice: cpu_avx2 (AVX2 + FMA) and cpu_base (x86-64 baseline).
#pragma once
#include <cstdint>

namespace mifft {
#define MIFFT_CPU_DECLS                                                                                          \
    void* create(int kind, int prec, int rank, const uint64_t* shape, const int* axes, int naxes, uint64_t batch, \
                 int inverse, double scale);                                                                     \
    void destroy(void* h);                                                                                       \
    void run(void* h, const void* in, void* out);                                                                \
    void set_max_threads(unsigned n);                                                                            \
    unsigned max_threads();
namespace cpu_avx2 {
MIFFT_CPU_DECLS
}
namespace cpu_base {
MIFFT_CPU_DECLS
}
#undef MIFFT_CPU_DECLS
// Whether the CPU runs AVX2 and FMA (defined in the host-only baseline unit).
bool cpu_has_avx2_fma();
}  // namespace mifft
