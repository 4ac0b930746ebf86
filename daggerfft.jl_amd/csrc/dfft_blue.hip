// dfft_blue.hip -- Bluestein (chirp-z) axis-pass kernels for gfx950.
//
// Line lengths with a prime factor above 7 (FFTW and cuFFT take any length)
// run as two M-point Stockham FFTs of a cyclic convolution in LDS, M >= 2n-1
// 7-smooth (blue_tile in dfft_axis.hpp), one launch per pass like every other
// axis pass, around the same fused load/store maps of each op.
#include <hip/hip_runtime.h>

#include "dfft_axis.hpp"
#include "dfft_fixed.hpp"

namespace dfft {

template <class T, int OP, bool CONTIG>
__global__ __launch_bounds__(1024) void k_axis_blue(AxisPass a) {
    extern __shared__ __align__(16) unsigned char dfft_smem[];
    blue_tile<8, T, OP, CONTIG>(a, a.in, a.out, pass_tile(a), (int)threadIdx.x,
                                reinterpret_cast<cx<T> *>(dfft_smem));
}

#define DFFT_BLUE_OPS(T, CONTIG)                                                                          \
    {(axis_kernel_t)k_axis_blue<T, OP_C2C, CONTIG>,      (axis_kernel_t)k_axis_blue<T, OP_R2C, CONTIG>,  \
     (axis_kernel_t)k_axis_blue<T, OP_R2C_FULL, CONTIG>, (axis_kernel_t)k_axis_blue<T, OP_C2R, CONTIG>,  \
     (axis_kernel_t)k_axis_blue<T, OP_C2R_FULL, CONTIG>, (axis_kernel_t)k_axis_blue<T, OP_DCT2, CONTIG>, \
     (axis_kernel_t)k_axis_blue<T, OP_DCT3, CONTIG>,     (axis_kernel_t)k_axis_blue<T, OP_DHT, CONTIG>,  \
     (axis_kernel_t)k_axis_blue<T, OP_DST2, CONTIG>,     (axis_kernel_t)k_axis_blue<T, OP_DST3, CONTIG>, \
     (axis_kernel_t)k_axis_blue<T, OP_DCT1, CONTIG>,     (axis_kernel_t)k_axis_blue<T, OP_DST1, CONTIG>, \
     (axis_kernel_t)k_axis_blue<T, OP_DCT4, CONTIG>,     (axis_kernel_t)k_axis_blue<T, OP_DST4, CONTIG>}

// [precision][contig][op]
static axis_kernel_t g_blue_kernels[2][2][OP_COUNT] = {
    {DFFT_BLUE_OPS(float, false), DFFT_BLUE_OPS(float, true)},
    {DFFT_BLUE_OPS(double, false), DFFT_BLUE_OPS(double, true)},
};

axis_kernel_t blue_kernel(int prec, int contig, int op) {
    if (op < 0 || op >= OP_COUNT) return nullptr;
    return g_blue_kernels[prec ? 1 : 0][contig ? 1 : 0][op];
}

}  // namespace dfft
