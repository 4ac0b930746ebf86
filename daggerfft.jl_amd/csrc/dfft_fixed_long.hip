// dfft_fixed_long.hip -- C2C catalog kernels of the power-of-two lengths
// 2048..16384: 16 values per thread with the split-LDS exchange (variant 2),
// fewest-stage radix lists (16*16*8 for 2048 ... 16*16*16*4 for 16384),
// both layouts and precisions.  A line of 16384 takes 1024 threads; its
// tile holds the real or imaginary parts of one line in LDS (131-147 KB in
// FP64), so these are also the longest lines the library transforms.
#include "dfft_fixed.hpp"

namespace dfft {

#define DFFT_LONG(T, P, L)                                              \
    DFFT_FX(T, P, OP_C2C, 1, 2, 16, true, RadicesOf16<L>),             \
        DFFT_FX(T, P, OP_C2C, 0, 2, 16, true, RadicesOf16<L>)

// FP64 4096 as 16*16*16 needs 129+ VGPRs and spills at the 1024-thread
// bound; 8*8*8*8 (one more stage) fits and is listed first, the default
#define DFFT_LONG8(T, P, L)                                             \
    DFFT_FX(T, P, OP_C2C, 1, 2, 16, true, RadicesOf8<L>),              \
        DFFT_FX(T, P, OP_C2C, 0, 2, 16, true, RadicesOf8<L>)

DFFT_TABLE(g_fixed_long,
    DFFT_LONG(double, 1, 2048), DFFT_LONG8(double, 1, 4096), DFFT_LONG(double, 1, 4096), DFFT_LONG(double, 1, 8192),
    DFFT_LONG(double, 1, 16384),
    DFFT_LONG(float, 0, 2048), DFFT_LONG(float, 0, 4096), DFFT_LONG(float, 0, 8192),
    DFFT_LONG(float, 0, 16384)
)

}  // namespace dfft
