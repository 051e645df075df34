// Shared declarations of the persistent batched kernel (kern_batch.hip) and its
// host planner (batch.cpp).
#pragma once
#include "conv_kernels.hpp"
#include "launch.h"

namespace mifft {

// Tile classes: the device function a tile runs.
enum BatchClass : uint32_t {
    BC_LDS = 0,
    BC_LDS_POST,
    BC_RADER,
    BC_BLUESTEIN,
    BC_REG_C4096,
    BC_REG_C2048,
    BC_REG_C1024,
    BC_REG_C512,
    BC_REG_C256,
    BC_REG_C128,
    BC_REG_C64,
    BC_REG_C120,
    BC_REG_C343,
    BC_REG_C1001,
    BC_REG_P256,    // four-step column pass with twiddle, 256 points, line pairs
    BC_REG_P256V1,  // the same, one line per thread
    BC_REG_T256,    // four-step row pass with transposed store, 256 points
    BC_REG_P128,
    BC_REG_T128,
    BC_LDS_R2C,       // f32 real load, packed store (R2C axis 0)
    BC_LDS_C2R,       // f32 Hermitian load, real store (C2R axis 0)
    BC_LDS_F64,
    BC_LDS_POST_F64,
    BC_RADER_F64,
    BC_BLUESTEIN_F64,
    BC_REG_C256_F64,  // f64 256-point lines, contiguous
    BC_REG_S256_F64,  // f64 256-point lines, strided
    // element-wise steps of a multi-pass Bluestein convolution (tile = a run
    // of kEltTile elements)
    BC_CHIRP_PRE,
    BC_CHIRP_PRE_F64,
    BC_POINTWISE,
    BC_POINTWISE_F64,
    BC_CHIRP_POST,
    BC_CHIRP_POST_F64,
    // strided line pairs (axes of ND arrays), 256 threads
    BC_REG_S1024V2,
    BC_REG_S512V2,
    BC_REG_S256V2,
    BC_REG_S128V2,
    BC_REG_S64V2,
    // register-resident Bluestein, one class per convolution length
    BC_REGB_64,
    BC_REGB_128,
    BC_REGB_256,
    BC_REGB_512,
    BC_REGB_1024,
    BC_REGB_2048,
    BC_REGB_4096,
    BC_REGB_8192,
    // fused Rader convolution with medium prime radices (17..61), f32
    BC_RADER_MED,
    BC_COUNT
};

// Dependent passes of one transform run in successive phases (one persistent
// launch each); ND transforms, four-step splits and multi-pass Bluestein
// (chirp, padded FFT, filter, padded FFT, chirp) need more than one.
constexpr int kMaxPhases = 8;
// Elements per tile of the element-wise classes (16 per thread).
constexpr uint32_t kEltTile = 4096;

struct BatchTile {
    uint32_t cls;  // BatchClass
    uint32_t arg;  // index into the OpDesc array
    uint32_t g0;   // first line of the tile within its op
    uint32_t cost; // scheduling weight (host only)
};

// Arguments of one op of one transform, device pointers already bound.
struct OpDesc {
    LineArgs a;   // line ops (register and LDS classes)
    Sched sc;     // LDS classes
    ConvArgs conv;  // Rader / Bluestein classes; element-wise classes (io, chirp, filt, n, m)
};

// Class of a register variant usable inside the batched kernel (256 threads,
// complex in and out), or -1.
int batch_class_for(const RegVariant* v);
// Class of a runtime-schedule LDS op (precision, load/store modes, twiddle on
// store), or -1.
int batch_lds_class(int prec, int ldm, int stm, int post);

hipError_t launch_batch(const BatchTile* tiles, uint32_t ntiles, uint32_t* counter, const OpDesc* ops,
                        uint32_t grid, size_t lds_bytes, hipStream_t s);
hipError_t launch_batch_reset(uint32_t* counters, uint32_t n, hipStream_t s);

}  // namespace mifft
