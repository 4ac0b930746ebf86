// dfft_blue_fixed.hpp -- the compile-time Bluestein kernel and its catalog
// entry macros (tables in dfft_blue.hip and dfft_blue_fixed2.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "dfft_axis.hpp"
#include "dfft_fixed.hpp"

namespace dfft {

// Bluestein tile on the compile-time engine (blue_fixed_tile)
template <class T, int OP, bool CONTIG, bool SPLIT, class RL>
__global__ __launch_bounds__(1024) void k_blue_fixed(AxisPass a) {
    extern __shared__ __align__(16) unsigned char dfft_smem[];
    blue_fixed_tile<RL, T, OP, CONTIG, SPLIT>(a, a.in, a.out, pass_tile(a), (int)threadIdx.x,
                                              reinterpret_cast<cx<T> *>(dfft_smem));
}

// Rader tile on the compile-time engine (rader_fixed_tile): C2C of a prime
// length n = RL::L + 1
template <class T, bool CONTIG, bool SPLIT, class RL>
__global__ __launch_bounds__(1024) void k_rader_fixed(AxisPass a) {
    extern __shared__ __align__(16) unsigned char dfft_smem[];
    rader_fixed_tile<RL, T, OP_C2C, CONTIG, SPLIT>(a, a.in, a.out, pass_tile(a), (int)threadIdx.x,
                                                   reinterpret_cast<cx<T> *>(dfft_smem));
}

// Rader catalog entries: L is the convolution length M = n - 1, split exchange
#define DFFT_RADER_FX(T, P, CONTIG, RL)                                                      \
    {P, OP_C2C, CONTIG, 2, RL::L, fixed_tpl<RL, 16>(), RL::S, RL::R, (axis_kernel_t)k_rader_fixed<T, CONTIG, true, RL>}
#define DFFT_RADER_M(M)                                                                      \
    DFFT_RADER_FX(double, 1, 1, RadicesOf16<M>), DFFT_RADER_FX(double, 1, 0, RadicesOf16<M>),   \
        DFFT_RADER_FX(float, 0, 1, RadicesOf16<M>), DFFT_RADER_FX(float, 0, 0, RadicesOf16<M>)
// FP64 with the fewest stages of radices up to 10 (the two-FFT FP64 tile
// spills with 12-16-point butterflies at the 1024-thread register bound)
#define DFFT_RADER_M10(M)                                                                    \
    DFFT_RADER_FX(double, 1, 1, RadicesCap10<M>), DFFT_RADER_FX(double, 1, 0, RadicesCap10<M>), \
        DFFT_RADER_FX(float, 0, 1, RadicesOf16<M>), DFFT_RADER_FX(float, 0, 0, RadicesOf16<M>)
// FP64 with radix-8 lists where those spill too
#define DFFT_RADER_M8(M)                                                                     \
    DFFT_RADER_FX(double, 1, 1, RadicesOf8<M>), DFFT_RADER_FX(double, 1, 0, RadicesOf8<M>),     \
        DFFT_RADER_FX(float, 0, 1, RadicesOf16<M>), DFFT_RADER_FX(float, 0, 0, RadicesOf16<M>)

// variant 2: split exchange; 3: full exchange
#define DFFT_BLUE_FX(T, P, OPV, CONTIG, VAR, RL)                                            \
    {P, OPV, CONTIG, VAR, RL::L, fixed_tpl<RL, 16>(), RL::S, RL::R,                          \
     (axis_kernel_t)k_blue_fixed<T, OPV, CONTIG, VAR == 2, RL>}
// C2C of convolution length M, both precisions and layouts
#define DFFT_BLUE_M(M)                                                                      \
    DFFT_BLUE_FX(double, 1, OP_C2C, 1, 2, RadicesOf16<M>), DFFT_BLUE_FX(double, 1, OP_C2C, 0, 2, RadicesOf16<M>), \
        DFFT_BLUE_FX(float, 0, OP_C2C, 1, 2, RadicesOf16<M>), DFFT_BLUE_FX(float, 0, OP_C2C, 0, 2, RadicesOf16<M>)

// the same with radix-8 lists in FP64: the 16-radix FP64 schedules of the
// larger mixed lengths (16*16*6 for 1536, ...) need more than 128 VGPRs and
// spill at the 1024-thread bound
#define DFFT_BLUE_M8(M)                                                                     \
    DFFT_BLUE_FX(double, 1, OP_C2C, 1, 2, RadicesOf8<M>), DFFT_BLUE_FX(double, 1, OP_C2C, 0, 2, RadicesOf8<M>), \
        DFFT_BLUE_FX(float, 0, OP_C2C, 1, 2, RadicesOf16<M>), DFFT_BLUE_FX(float, 0, OP_C2C, 0, 2, RadicesOf16<M>)

// the real-input / real-output ops of one precision along the contiguous
// axis: packed R2C (full exchange), R2C_FULL, C2R and C2R_FULL (split)
#define DFFT_BLUE_REAL(T, P, RL)                                                            \
    DFFT_BLUE_FX(T, P, OP_R2C, 1, 3, RL), DFFT_BLUE_FX(T, P, OP_R2C_FULL, 1, 2, RL),          \
        DFFT_BLUE_FX(T, P, OP_C2R, 1, 2, RL), DFFT_BLUE_FX(T, P, OP_C2R_FULL, 1, 2, RL)

}  // namespace dfft
