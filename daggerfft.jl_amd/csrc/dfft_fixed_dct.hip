// dfft_fixed_dct.hip -- DCT-II / DCT-III catalog kernels of the power-of-two
// lengths 16..1024 (the DCT configurations run 128 and 256).
#include "dfft_fixed.hpp"

namespace dfft {

DFFT_TABLE(g_fixed_dct,
    DFFT_DCT(double, 1, 16), DFFT_DCT(double, 1, 32), DFFT_DCT(double, 1, 64), DFFT_DCT(double, 1, 128),
    DFFT_DCT(double, 1, 256), DFFT_DCT(double, 1, 512), DFFT_DCT(double, 1, 1024),
    DFFT_DCT(float, 0, 16), DFFT_DCT(float, 0, 32), DFFT_DCT(float, 0, 64), DFFT_DCT(float, 0, 128),
    DFFT_DCT(float, 0, 256), DFFT_DCT(float, 0, 512), DFFT_DCT(float, 0, 1024)
)

}  // namespace dfft
