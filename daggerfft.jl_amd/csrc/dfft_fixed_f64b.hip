// dfft_fixed_f64b.hip -- FP64 C2C catalog kernels, lengths 96..245
// (every {2,3,5,7}-smooth length in the range except the powers of two
// 64..1024, which have their own kernels in dfft_fixed_pow2.hip)
#include "dfft_fixed.hpp"

namespace dfft {

DFFT_TABLE(g_fixed_f64b,
    DFFT_C2C_F64(96), DFFT_C2C_F64(98), DFFT_C2C_F64(100), DFFT_C2C_F64(105), DFFT_C2C_F64(108), DFFT_C2C_F64(112),
    DFFT_C2C_F64(120), DFFT_C2C_F64(125), DFFT_C2C_F64(126), DFFT_C2C_F64(135), DFFT_C2C_F64(140), DFFT_C2C_F64(144),
    DFFT_C2C_F64(147), DFFT_C2C_F64(150), DFFT_C2C_F64(160), DFFT_C2C_F64(162), DFFT_C2C_F64(168), DFFT_C2C_F64(175),
    DFFT_C2C_F64(180), DFFT_C2C_F64(189), DFFT_C2C_F64(192), DFFT_C2C_F64(196), DFFT_C2C_F64(200), DFFT_C2C_F64(210),
    DFFT_C2C_F64(216), DFFT_C2C_F64(224), DFFT_C2C_F64(225), DFFT_C2C_F64(240), DFFT_C2C_F64(243), DFFT_C2C_F64(245)
)

}  // namespace dfft
