// dfft_fixed_reala.hip -- fused C2R (split exchange) catalog kernels of the
// {2,3,5,7}-smooth half-lengths 270..567, both
// precisions, contiguous axis: real lines of 2H points (1-9% faster than the
// runtime engine).  R2C of these half-lengths stays on the runtime engine:
// its split post-twiddle would hold (H/2)/TPL partial-sum pairs per thread,
// too many registers, and the full-exchange kernel measured 2-6% slower.
#include "dfft_fixed.hpp"

namespace dfft {

DFFT_TABLE(g_fixed_reala,
    DFFT_C2R_SPLIT(270),
    DFFT_C2R_SPLIT(280),
    DFFT_C2R_SPLIT(288),
    DFFT_C2R_SPLIT(294),
    DFFT_C2R_SPLIT(300),
    DFFT_C2R_SPLIT(315),
    DFFT_C2R_SPLIT(320),
    DFFT_C2R_SPLIT(324),
    DFFT_C2R_SPLIT(336),
    DFFT_C2R_SPLIT(343),
    DFFT_C2R_SPLIT(350),
    DFFT_C2R_SPLIT(360),
    DFFT_C2R_SPLIT(375),
    DFFT_C2R_SPLIT(378),
    DFFT_C2R_SPLIT(384),
    DFFT_C2R_SPLIT(392),
    DFFT_C2R_SPLIT(400),
    DFFT_C2R_SPLIT(405),
    DFFT_C2R_SPLIT(420),
    DFFT_C2R_SPLIT(432),
    DFFT_C2R_SPLIT(441),
    DFFT_C2R_SPLIT(448),
    DFFT_C2R_SPLIT(450),
    DFFT_C2R_SPLIT(480),
    DFFT_C2R_SPLIT(486),
    DFFT_C2R_SPLIT(490),
    DFFT_C2R_SPLIT(500),
    DFFT_C2R_SPLIT(504),
    DFFT_C2R_SPLIT(512),
    DFFT_C2R_SPLIT(525),
    DFFT_C2R_SPLIT(540),
    DFFT_C2R_SPLIT(560),
    DFFT_C2R_SPLIT(567))

}  // namespace dfft
