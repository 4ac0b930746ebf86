// dfft_fixed_c2rb.hip -- C2R catalog kernels, half-lengths 100..256
// (every {2,3,5,7}-smooth half-length in the range: real lines of 2H points)
#include "dfft_fixed.hpp"

namespace dfft {

DFFT_TABLE(g_fixed_c2rb,
    DFFT_C2R(100), DFFT_C2R(105), DFFT_C2R(108), DFFT_C2R(112), DFFT_C2R(120), DFFT_C2R(125),
    DFFT_C2R(126), DFFT_C2R(128), DFFT_C2R(135), DFFT_C2R(140), DFFT_C2R(144), DFFT_C2R(147),
    DFFT_C2R(150), DFFT_C2R(160), DFFT_C2R(162), DFFT_C2R(168), DFFT_C2R(175), DFFT_C2R(180),
    DFFT_C2R(189), DFFT_C2R(192), DFFT_C2R(196), DFFT_C2R(200), DFFT_C2R(210), DFFT_C2R(216),
    DFFT_C2R(224), DFFT_C2R(225), DFFT_C2R(240), DFFT_C2R(243), DFFT_C2R(245), DFFT_C2R(250),
    DFFT_C2R(252), DFFT_C2R(256)
)

}  // namespace dfft
