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
