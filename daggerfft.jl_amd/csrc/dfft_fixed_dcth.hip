// dfft_fixed_dcth.hip -- REDFT10 / REDFT01 of n = 2H reals (H a power of two,
// 8..512) through an H-point complex FFT: DCT-II with whole complex values in
// LDS (its post-twiddle pairs bins k and H-k there), DCT-III with the split
// exchange (its pre-twiddle reads global memory only); both layouts.
#include "dfft_fixed.hpp"

namespace dfft {

#define DFFT_DCTH(T, P, H)                                             \
    DFFT_FX(T, P, OP_DCT2H, 1, 3, 16, false, RadicesOf16<H>),         \
        DFFT_FX(T, P, OP_DCT2H, 0, 3, 16, false, RadicesOf16<H>),     \
        DFFT_FX(T, P, OP_DCT3H, 1, 2, 16, true, RadicesOf16<H>),      \
        DFFT_FX(T, P, OP_DCT3H, 0, 2, 16, true, RadicesOf16<H>)

DFFT_TABLE(g_fixed_dcth,
    DFFT_DCTH(double, 1, 8), DFFT_DCTH(double, 1, 16), DFFT_DCTH(double, 1, 32), DFFT_DCTH(double, 1, 64),
    DFFT_DCTH(double, 1, 128), DFFT_DCTH(double, 1, 256), DFFT_DCTH(double, 1, 512),
    DFFT_DCTH(float, 0, 8), DFFT_DCTH(float, 0, 16), DFFT_DCTH(float, 0, 32), DFFT_DCTH(float, 0, 64),
    DFFT_DCTH(float, 0, 128), DFFT_DCTH(float, 0, 256), DFFT_DCTH(float, 0, 512)
)

}  // namespace dfft
