// dfft_blue_fixed.hpp -- the compile-time Bluestein kernel and its catalog
// entry macros (tables in dfft_blue.hip and dfft_blue_fixed2.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "dfft_axis.hpp"
#include "dfft_fixed.hpp"

namespace dfft {

// C2C Bluestein tile on the compile-time engine (blue_fixed_tile)
template <class T, bool CONTIG, class RL>
__global__ __launch_bounds__(1024) void k_blue_fixed(AxisPass a) {
    extern __shared__ __align__(16) unsigned char dfft_smem[];
    blue_fixed_tile<RL, T, CONTIG>(a, a.in, a.out, pass_tile(a), (int)threadIdx.x,
                                   reinterpret_cast<cx<T> *>(dfft_smem));
}

#define DFFT_BLUE_FX(T, P, CONTIG, RL)                                                      \
    {P, OP_C2C, CONTIG, 2, RL::L, fixed_tpl<RL, 16>(), RL::S, RL::R,                        \
     (axis_kernel_t)k_blue_fixed<T, CONTIG, RL>}
// convolution length M, both precisions and layouts
#define DFFT_BLUE_M(M)                                                                      \
    DFFT_BLUE_FX(double, 1, 1, RadicesOf16<M>), DFFT_BLUE_FX(double, 1, 0, RadicesOf16<M>), \
        DFFT_BLUE_FX(float, 0, 1, RadicesOf16<M>), DFFT_BLUE_FX(float, 0, 0, RadicesOf16<M>)

// the same with radix-8 lists in FP64: the 16-radix FP64 schedules of the
// larger mixed lengths (16*16*6 for 1536, ...) need more than 128 VGPRs and
// spill at the 1024-thread bound
#define DFFT_BLUE_M8(M)                                                                     \
    DFFT_BLUE_FX(double, 1, 1, RadicesOf8<M>), DFFT_BLUE_FX(double, 1, 0, RadicesOf8<M>),   \
        DFFT_BLUE_FX(float, 0, 1, RadicesOf16<M>), DFFT_BLUE_FX(float, 0, 0, RadicesOf16<M>)

}  // namespace dfft
