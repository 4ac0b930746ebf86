// dfft_rader_a.hip -- compile-time Rader kernels (C2C of prime lengths
// n = M + 1 with 7-smooth M), part a: M = 10..378.  Both precisions
// and layouts; the planner takes them over Bluestein for these lengths.
#include "dfft_blue_fixed.hpp"

namespace dfft {

DFFT_TABLE(g_rader_a, DFFT_RADER_M10(10), DFFT_RADER_M10(12), DFFT_RADER_M10(16), DFFT_RADER_M10(18), DFFT_RADER_M10(28), DFFT_RADER_M10(30),
           DFFT_RADER_M10(36), DFFT_RADER_M10(40), DFFT_RADER_M10(42), DFFT_RADER_M10(60), DFFT_RADER_M10(70), DFFT_RADER_M10(72),
           DFFT_RADER_M10(96), DFFT_RADER_M10(100), DFFT_RADER_M10(108), DFFT_RADER_M10(112), DFFT_RADER_M10(126), DFFT_RADER_M10(150),
           DFFT_RADER_M10(162), DFFT_RADER_M10(180), DFFT_RADER_M10(192), DFFT_RADER_M10(196), DFFT_RADER_M10(210), DFFT_RADER_M10(240),
           DFFT_RADER_M10(250), DFFT_RADER_M10(256), DFFT_RADER_M10(270), DFFT_RADER_M10(280), DFFT_RADER_M10(336), DFFT_RADER_M10(378))

}  // namespace dfft
