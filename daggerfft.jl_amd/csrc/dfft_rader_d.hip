// dfft_rader_d.hip -- compile-time Rader kernels (C2C of prime lengths
// n = M + 1 with 7-smooth M), part d: M = 3456..8100.  Both precisions
// and layouts; the planner takes them over Bluestein for these lengths.
#include "dfft_blue_fixed.hpp"

namespace dfft {

DFFT_TABLE(g_rader_d, DFFT_RADER_M10(3456), DFFT_RADER_M10(3528), DFFT_RADER_M10(3888), DFFT_RADER_M10(4000), DFFT_RADER_M10(4050), DFFT_RADER_M10(4200),
           DFFT_RADER_M10(4480), DFFT_RADER_M10(4800), DFFT_RADER_M10(4860), DFFT_RADER_M10(5880), DFFT_RADER_M10(6300), DFFT_RADER_M10(6480),
           DFFT_RADER_M10(7000), DFFT_RADER_M10(7056), DFFT_RADER_M10(7350), DFFT_RADER_M10(7560), DFFT_RADER_M10(7680), DFFT_RADER_M10(7840),
           DFFT_RADER_M10(8100))

}  // namespace dfft
