// dfft_blue_real32.hip -- compile-time Bluestein kernels of the real ops (packed R2C,
// R2C_FULL, C2R, C2R_FULL) along the contiguous axis, ComplexF32/Float32, every
// convolution length of the C2C tables.
#include "dfft_blue_fixed.hpp"

namespace dfft {

DFFT_TABLE(g_blue_real32,
    DFFT_BLUE_REAL(float, 0, RadicesOf16<32>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<64>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<128>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<256>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<512>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<1024>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<2048>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<4096>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<8192>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<48>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<96>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<192>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<384>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<768>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<40>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<80>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<160>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<320>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<640>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<1280>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<1536>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<3072>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<6144>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<2560>),
    DFFT_BLUE_REAL(float, 0, RadicesOf16<5120>))

}  // namespace dfft
