// dfft_fixed_pow2.hip -- C2C catalog kernels of the power-of-two lengths
// 64..1024: 8 values per thread (variant 1) and 16 values per thread with
// the split-LDS exchange (variant 2), both layouts.
#include "dfft_fixed.hpp"

namespace dfft {

// the radix-8 lists of both variants (8*8*8 for 512, ...) are the measured ones
#define DFFT_POW2(T, P, L)                                              \
    DFFT_FX(T, P, OP_C2C, 1, 1, 8, false, RadicesOf8<L>),            \
        DFFT_FX(T, P, OP_C2C, 1, 2, 16, true, RadicesOf8<L>),        \
        DFFT_FX(T, P, OP_C2C, 0, 1, 8, false, RadicesOf8<L>),        \
        DFFT_FX(T, P, OP_C2C, 0, 2, 16, true, RadicesOf8<L>)

// the split schedule again with the fewest-stage lists (16*16 for 256,
// 16*16*2 for 512, ...): the autotuner picks between the two
#define DFFT_POW2_16(T, P, L)                                           \
    DFFT_FX(T, P, OP_C2C, 1, 2, 16, true, RadicesOf16<L>),             \
        DFFT_FX(T, P, OP_C2C, 0, 2, 16, true, RadicesOf16<L>)

DFFT_TABLE(g_fixed_pow2,
    DFFT_POW2(double, 1, 64), DFFT_POW2(double, 1, 128), DFFT_POW2(double, 1, 256),
    DFFT_POW2(double, 1, 512), DFFT_POW2(double, 1, 1024),
    DFFT_POW2(float, 0, 64), DFFT_POW2(float, 0, 128), DFFT_POW2(float, 0, 256),
    DFFT_POW2(float, 0, 512), DFFT_POW2(float, 0, 1024),
    DFFT_POW2_16(double, 1, 64), DFFT_POW2_16(double, 1, 128), DFFT_POW2_16(double, 1, 256),
    DFFT_POW2_16(double, 1, 512), DFFT_POW2_16(double, 1, 1024),
    DFFT_POW2_16(float, 0, 64), DFFT_POW2_16(float, 0, 128), DFFT_POW2_16(float, 0, 256),
    DFFT_POW2_16(float, 0, 512), DFFT_POW2_16(float, 0, 1024)
)

}  // namespace dfft
