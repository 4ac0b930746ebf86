// dfft_fixed_r2ca.hip -- R2C catalog kernels, half-lengths 16..98
// (every {2,3,5,7}-smooth half-length in the range: real lines of 2H points;
// full LDS exchange only)
#include "dfft_fixed.hpp"

namespace dfft {

DFFT_TABLE(g_fixed_r2ca,
    DFFT_R2C_FULL(16), DFFT_R2C_FULL(18), DFFT_R2C_FULL(20), DFFT_R2C_FULL(21), DFFT_R2C_FULL(24), DFFT_R2C_FULL(25),
    DFFT_R2C_FULL(27), DFFT_R2C_FULL(28), DFFT_R2C_FULL(30), DFFT_R2C_FULL(32), DFFT_R2C_FULL(35), DFFT_R2C_FULL(36),
    DFFT_R2C_FULL(40), DFFT_R2C_FULL(42), DFFT_R2C_FULL(45), DFFT_R2C_FULL(48), DFFT_R2C_FULL(49), DFFT_R2C_FULL(50),
    DFFT_R2C_FULL(54), DFFT_R2C_FULL(56), DFFT_R2C_FULL(60), DFFT_R2C_FULL(63), DFFT_R2C_FULL(64), DFFT_R2C_FULL(70),
    DFFT_R2C_FULL(72), DFFT_R2C_FULL(75), DFFT_R2C_FULL(80), DFFT_R2C_FULL(81), DFFT_R2C_FULL(84), DFFT_R2C_FULL(90),
    DFFT_R2C_FULL(96), DFFT_R2C_FULL(98)
)

}  // namespace dfft
