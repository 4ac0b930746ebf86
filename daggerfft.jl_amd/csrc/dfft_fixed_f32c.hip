// dfft_fixed_f32c.hip -- FP32 C2C catalog kernels, lengths 250..504
// (every {2,3,5,7}-smooth length in the range except the powers of two
// 64..1024, which have their own kernels in dfft_fixed_pow2.hip)
#include "dfft_fixed.hpp"

namespace dfft {

DFFT_TABLE(g_fixed_f32c,
    DFFT_C2C_F32(250), DFFT_C2C_F32(252), DFFT_C2C_F32(270), DFFT_C2C_F32(280), DFFT_C2C_F32(288), DFFT_C2C_F32(294),
    DFFT_C2C_F32(300), DFFT_C2C_F32(315), DFFT_C2C_F32(320), DFFT_C2C_F32(324), DFFT_C2C_F32(336), DFFT_C2C_F32(343),
    DFFT_C2C_F32(350), DFFT_C2C_F32(360), DFFT_C2C_F32(375), DFFT_C2C_F32(378), DFFT_C2C_F32(384), DFFT_C2C_F32(392),
    DFFT_C2C_F32(400), DFFT_C2C_F32(405), DFFT_C2C_F32(420), DFFT_C2C_F32(432), DFFT_C2C_F32(441), DFFT_C2C_F32(448),
    DFFT_C2C_F32(450), DFFT_C2C_F32(480), DFFT_C2C_F32(486), DFFT_C2C_F32(490), DFFT_C2C_F32(500), DFFT_C2C_F32(504)
)

}  // namespace dfft
