// dfft_fixed_c2ra.hip -- C2R catalog kernels, half-lengths 16..98
// (every {2,3,5,7}-smooth half-length in the range: real lines of 2H points)
#include "dfft_fixed.hpp"

namespace dfft {

DFFT_TABLE(g_fixed_c2ra,
    DFFT_C2R(16), DFFT_C2R(18), DFFT_C2R(20), DFFT_C2R(21), DFFT_C2R(24), DFFT_C2R(25),
    DFFT_C2R(27), DFFT_C2R(28), DFFT_C2R(30), DFFT_C2R(32), DFFT_C2R(35), DFFT_C2R(36),
    DFFT_C2R(40), DFFT_C2R(42), DFFT_C2R(45), DFFT_C2R(48), DFFT_C2R(49), DFFT_C2R(50),
    DFFT_C2R(54), DFFT_C2R(56), DFFT_C2R(60), DFFT_C2R(63), DFFT_C2R(64), DFFT_C2R(70),
    DFFT_C2R(72), DFFT_C2R(75), DFFT_C2R(80), DFFT_C2R(81), DFFT_C2R(84), DFFT_C2R(90),
    DFFT_C2R(96), DFFT_C2R(98)
)

}  // namespace dfft
