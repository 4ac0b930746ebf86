// dfft_rader_c.hip -- compile-time Rader kernels (C2C of prime lengths
// n = M + 1 with 7-smooth M), part c: M = 1296..3360.  Both precisions
// and layouts; the planner takes them over Bluestein for these lengths.
#include "dfft_blue_fixed.hpp"

namespace dfft {

DFFT_TABLE(g_rader_c, DFFT_RADER_M10(1296), DFFT_RADER_M10(1372), DFFT_RADER_M10(1458), DFFT_RADER_M10(1470), DFFT_RADER_M10(1600), DFFT_RADER_M10(1620),
           DFFT_RADER_M10(1800), DFFT_RADER_M10(2016), DFFT_RADER_M10(2160), DFFT_RADER_M10(2250), DFFT_RADER_M10(2268), DFFT_RADER_M10(2520),
           DFFT_RADER_M10(2592), DFFT_RADER_M10(2646), DFFT_RADER_M10(2688), DFFT_RADER_M10(2800), DFFT_RADER_M10(2916), DFFT_RADER_M10(3000),
           DFFT_RADER_M10(3136), DFFT_RADER_M10(3360))

}  // namespace dfft
