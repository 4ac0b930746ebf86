// dfft_blue_fixed2.hip -- compile-time C2C Bluestein kernels of the
// convolution lengths 3*2^k (48..6144) and 5*2^k (40..5120): with the powers
// of two (dfft_blue.hip) the planner finds one within 4/3 of 2n-1.
#include "dfft_blue_fixed.hpp"

namespace dfft {

DFFT_TABLE(g_blue_mixed, DFFT_BLUE_M(48), DFFT_BLUE_M(96), DFFT_BLUE_M(192), DFFT_BLUE_M(384), DFFT_BLUE_M(768),
           DFFT_BLUE_M8(1536), DFFT_BLUE_M8(3072), DFFT_BLUE_M8(6144), DFFT_BLUE_M(40), DFFT_BLUE_M(80),
           DFFT_BLUE_M(160), DFFT_BLUE_M(320), DFFT_BLUE_M(640), DFFT_BLUE_M(1280), DFFT_BLUE_M8(2560),
           DFFT_BLUE_M8(5120))

}  // namespace dfft
