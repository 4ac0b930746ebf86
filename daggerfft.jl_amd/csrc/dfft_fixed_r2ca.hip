// dfft_fixed_r2ca.hip -- R2C catalog kernels, half-lengths 16..98
// (every {2,3,5,7}-smooth half-length in the range: real lines of 2H points)
#include "dfft_fixed.hpp"

namespace dfft {

DFFT_TABLE(g_fixed_r2ca,
    DFFT_R2C(16), DFFT_R2C(18), DFFT_R2C(20), DFFT_R2C(21), DFFT_R2C(24), DFFT_R2C(25),
    DFFT_R2C(27), DFFT_R2C(28), DFFT_R2C(30), DFFT_R2C(32), DFFT_R2C(35), DFFT_R2C(36),
    DFFT_R2C(40), DFFT_R2C(42), DFFT_R2C(45), DFFT_R2C(48), DFFT_R2C(49), DFFT_R2C(50),
    DFFT_R2C(54), DFFT_R2C(56), DFFT_R2C(60), DFFT_R2C(63), DFFT_R2C(64), DFFT_R2C(70),
    DFFT_R2C(72), DFFT_R2C(75), DFFT_R2C(80), DFFT_R2C(81), DFFT_R2C(84), DFFT_R2C(90),
    DFFT_R2C(96), DFFT_R2C(98)
)

}  // namespace dfft
