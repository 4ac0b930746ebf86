// dfft_fixed_realb.hip -- fused C2R (split exchange) catalog kernels of the
// {2,3,5,7}-smooth half-lengths 576..1024, both
// precisions, contiguous axis: real lines of 2H points (1-9% faster than the
// runtime engine).  R2C of these half-lengths stays on the runtime engine:
// its split post-twiddle would hold (H/2)/TPL partial-sum pairs per thread,
// too many registers, and the full-exchange kernel measured 2-6% slower.
#include "dfft_fixed.hpp"

namespace dfft {

DFFT_TABLE(g_fixed_realb,
    DFFT_C2R_SPLIT(576),
    DFFT_C2R_SPLIT(588),
    DFFT_C2R_SPLIT(600),
    DFFT_C2R_SPLIT(625),
    DFFT_C2R_SPLIT(630),
    DFFT_C2R_SPLIT(640),
    DFFT_C2R_SPLIT(648),
    DFFT_C2R_SPLIT(672),
    DFFT_C2R_SPLIT(675),
    DFFT_C2R_SPLIT(686),
    DFFT_C2R_SPLIT(700),
    DFFT_C2R_SPLIT(720),
    DFFT_C2R_SPLIT(729),
    DFFT_C2R_SPLIT(735),
    DFFT_C2R_SPLIT(750),
    DFFT_C2R_SPLIT(756),
    DFFT_C2R_SPLIT(768),
    DFFT_C2R_SPLIT(784),
    DFFT_C2R_SPLIT(800),
    DFFT_C2R_SPLIT(810),
    DFFT_C2R_SPLIT(840),
    DFFT_C2R_SPLIT(864),
    DFFT_C2R_SPLIT(875),
    DFFT_C2R_SPLIT(882),
    DFFT_C2R_SPLIT(896),
    DFFT_C2R_SPLIT(900),
    DFFT_C2R_SPLIT(945),
    DFFT_C2R_SPLIT(960),
    DFFT_C2R_SPLIT(972),
    DFFT_C2R_SPLIT(980),
    DFFT_C2R_SPLIT(1000),
    DFFT_C2R_SPLIT(1008),
    DFFT_C2R_SPLIT(1024))

}  // namespace dfft
