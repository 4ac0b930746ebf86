// dfft_fixed_r2cb.hip -- R2C catalog kernels, half-lengths 100..256
// (every {2,3,5,7}-smooth half-length in the range: real lines of 2H points)
#include "dfft_fixed.hpp"

namespace dfft {

DFFT_TABLE(g_fixed_r2cb,
    DFFT_R2C(100), DFFT_R2C(105), DFFT_R2C(108), DFFT_R2C(112), DFFT_R2C(120), DFFT_R2C(125),
    DFFT_R2C(126), DFFT_R2C(128), DFFT_R2C(135), DFFT_R2C(140), DFFT_R2C(144), DFFT_R2C(147),
    DFFT_R2C(150), DFFT_R2C(160), DFFT_R2C(162), DFFT_R2C(168), DFFT_R2C(175), DFFT_R2C(180),
    DFFT_R2C(189), DFFT_R2C(192), DFFT_R2C(196), DFFT_R2C(200), DFFT_R2C(210), DFFT_R2C(216),
    DFFT_R2C(224), DFFT_R2C(225), DFFT_R2C(240), DFFT_R2C(243), DFFT_R2C(245), DFFT_R2C(250),
    DFFT_R2C(252), DFFT_R2C(256)
)

}  // namespace dfft
