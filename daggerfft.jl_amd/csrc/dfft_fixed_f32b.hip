// dfft_fixed_f32b.hip -- FP32 C2C catalog kernels, lengths 96..245
// (every {2,3,5,7}-smooth length in the range except the powers of two
// 64..1024, which have their own kernels in dfft_fixed_pow2.hip)
#include "dfft_fixed.hpp"

namespace dfft {

DFFT_TABLE(g_fixed_f32b,
    DFFT_C2C_F32(96), DFFT_C2C_F32(98), DFFT_C2C_F32(100), DFFT_C2C_F32(105), DFFT_C2C_F32(108), DFFT_C2C_F32(112),
    DFFT_C2C_F32(120), DFFT_C2C_F32(125), DFFT_C2C_F32(126), DFFT_C2C_F32(135), DFFT_C2C_F32(140), DFFT_C2C_F32(144),
    DFFT_C2C_F32(147), DFFT_C2C_F32(150), DFFT_C2C_F32(160), DFFT_C2C_F32(162), DFFT_C2C_F32(168), DFFT_C2C_F32(175),
    DFFT_C2C_F32(180), DFFT_C2C_F32(189), DFFT_C2C_F32(192), DFFT_C2C_F32(196), DFFT_C2C_F32(200), DFFT_C2C_F32(210),
    DFFT_C2C_F32(216), DFFT_C2C_F32(224), DFFT_C2C_F32(225), DFFT_C2C_F32(240), DFFT_C2C_F32(243), DFFT_C2C_F32(245)
)

}  // namespace dfft
