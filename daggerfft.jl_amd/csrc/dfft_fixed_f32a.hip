// dfft_fixed_f32a.hip -- FP32 C2C catalog kernels, lengths 16..90
// (every {2,3,5,7}-smooth length in the range except the powers of two
// 64..1024, which have their own kernels in dfft_fixed_pow2.hip)
#include "dfft_fixed.hpp"

namespace dfft {

DFFT_TABLE(g_fixed_f32a,
    DFFT_C2C_F32(16), DFFT_C2C_F32(18), DFFT_C2C_F32(20), DFFT_C2C_F32(21), DFFT_C2C_F32(24), DFFT_C2C_F32(25),
    DFFT_C2C_F32(27), DFFT_C2C_F32(28), DFFT_C2C_F32(30), DFFT_C2C_F32(32), DFFT_C2C_F32(35), DFFT_C2C_F32(36),
    DFFT_C2C_F32(40), DFFT_C2C_F32(42), DFFT_C2C_F32(45), DFFT_C2C_F32(48), DFFT_C2C_F32(49), DFFT_C2C_F32(50),
    DFFT_C2C_F32(54), DFFT_C2C_F32(56), DFFT_C2C_F32(60), DFFT_C2C_F32(63), DFFT_C2C_F32(70), DFFT_C2C_F32(72),
    DFFT_C2C_F32(75), DFFT_C2C_F32(80), DFFT_C2C_F32(81), DFFT_C2C_F32(84), DFFT_C2C_F32(90)
)

}  // namespace dfft
