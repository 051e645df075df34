// The host FFT engine built for the x86-64 baseline.
#define CPU_NS cpu_base
#include "cpu_fft.hpp"
#include "cpu_fft_core.inc"

bool mifft::cpu_has_avx2_fma() {
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
}
