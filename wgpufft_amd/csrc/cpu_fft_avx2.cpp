// The host FFT engine built with AVX2 + FMA (selected at run time when the CPU has them).
#define CPU_NS cpu_avx2
#include "cpu_fft.hpp"
#include "cpu_fft_core.inc"
