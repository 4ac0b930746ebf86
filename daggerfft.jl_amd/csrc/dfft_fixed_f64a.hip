// dfft_fixed_f64a.hip -- FP64 C2C catalog kernels, lengths 16..90
// (every {2,3,5,7}-smooth length in the range except the powers of two
// 64..1024, which have their own kernels in dfft_fixed_pow2.hip)
#include "dfft_fixed.hpp"

namespace dfft {

DFFT_TABLE(g_fixed_f64a,
    DFFT_C2C_F64(16), DFFT_C2C_F64(18), DFFT_C2C_F64(20), DFFT_C2C_F64(21), DFFT_C2C_F64(24), DFFT_C2C_F64(25),
    DFFT_C2C_F64(27), DFFT_C2C_F64(28), DFFT_C2C_F64(30), DFFT_C2C_F64(32), DFFT_C2C_F64(35), DFFT_C2C_F64(36),
    DFFT_C2C_F64(40), DFFT_C2C_F64(42), DFFT_C2C_F64(45), DFFT_C2C_F64(48), DFFT_C2C_F64(49), DFFT_C2C_F64(50),
    DFFT_C2C_F64(54), DFFT_C2C_F64(56), DFFT_C2C_F64(60), DFFT_C2C_F64(63), DFFT_C2C_F64(70), DFFT_C2C_F64(72),
    DFFT_C2C_F64(75), DFFT_C2C_F64(80), DFFT_C2C_F64(81), DFFT_C2C_F64(84), DFFT_C2C_F64(90)
)

}  // namespace dfft
