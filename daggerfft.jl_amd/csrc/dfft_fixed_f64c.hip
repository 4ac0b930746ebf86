// dfft_fixed_f64c.hip -- FP64 C2C catalog kernels, lengths 250..504
// (every {2,3,5,7}-smooth length in the range except the powers of two
// 64..1024, which have their own kernels in dfft_fixed_pow2.hip)
#include "dfft_fixed.hpp"

namespace dfft {

DFFT_TABLE(g_fixed_f64c,
    DFFT_C2C_F64(250), DFFT_C2C_F64(252), DFFT_C2C_F64(270), DFFT_C2C_F64(280), DFFT_C2C_F64(288), DFFT_C2C_F64(294),
    DFFT_C2C_F64(300), DFFT_C2C_F64(315), DFFT_C2C_F64(320), DFFT_C2C_F64(324), DFFT_C2C_F64(336), DFFT_C2C_F64(343),
    DFFT_C2C_F64(350), DFFT_C2C_F64(360), DFFT_C2C_F64(375), DFFT_C2C_F64(378), DFFT_C2C_F64(384), DFFT_C2C_F64(392),
    DFFT_C2C_F64(400), DFFT_C2C_F64(405), DFFT_C2C_F64(420), DFFT_C2C_F64(432), DFFT_C2C_F64(441), DFFT_C2C_F64(448),
    DFFT_C2C_F64(450), DFFT_C2C_F64(480), DFFT_C2C_F64(486), DFFT_C2C_F64(490), DFFT_C2C_F64(500), DFFT_C2C_F64(504)
)

}  // namespace dfft
