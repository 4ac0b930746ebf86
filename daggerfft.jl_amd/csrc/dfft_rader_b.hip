// dfft_rader_b.hip -- compile-time Rader kernels (C2C of prime lengths
// n = M + 1 with 7-smooth M), part b: M = 400..1200.  Both precisions
// and layouts; the planner takes them over Bluestein for these lengths.
#include "dfft_blue_fixed.hpp"

namespace dfft {

DFFT_TABLE(g_rader_b, DFFT_RADER_M10(400), DFFT_RADER_M10(420), DFFT_RADER_M10(432), DFFT_RADER_M10(448), DFFT_RADER_M10(486), DFFT_RADER_M10(490),
           DFFT_RADER_M10(540), DFFT_RADER_M10(576), DFFT_RADER_M10(600), DFFT_RADER_M10(630), DFFT_RADER_M10(640), DFFT_RADER_M10(672),
           DFFT_RADER_M10(700), DFFT_RADER_M10(750), DFFT_RADER_M10(756), DFFT_RADER_M10(768), DFFT_RADER_M10(810), DFFT_RADER_M10(882),
           DFFT_RADER_M10(1008), DFFT_RADER_M10(1050), DFFT_RADER_M10(1152), DFFT_RADER_M10(1200))

}  // namespace dfft
