// dfft_blue.hip -- Bluestein (chirp-z) axis-pass kernels for gfx950.
//
// Line lengths with a prime factor above 7 (FFTW and cuFFT take any length)
// run as two M-point Stockham FFTs of a cyclic convolution in LDS, M >= 2n-1
// 7-smooth (blue_tile in dfft_axis.hpp), one launch per pass like every other
// axis pass, around the same fused load/store maps of each op.
#include <hip/hip_runtime.h>

#include <vector>

#include "dfft_axis.hpp"
#include "dfft_blue_fixed.hpp"
#include "dfft_fixed.hpp"
#include "dfft_kernels.hpp"

namespace dfft {

template <class T, int OP, bool CONTIG>
__global__ __launch_bounds__(1024) void k_axis_blue(AxisPass a) {
    extern __shared__ __align__(16) unsigned char dfft_smem[];
    blue_tile<8, T, OP, CONTIG>(a, a.in, a.out, pass_tile(a), (int)threadIdx.x,
                                reinterpret_cast<cx<T> *>(dfft_smem));
}

#define DFFT_BLUE_OPS(T, CONTIG)                                                                          \
    {(axis_kernel_t)k_axis_blue<T, OP_C2C, CONTIG>,      (axis_kernel_t)k_axis_blue<T, OP_R2C, CONTIG>,  \
     (axis_kernel_t)k_axis_blue<T, OP_R2C_FULL, CONTIG>, (axis_kernel_t)k_axis_blue<T, OP_C2R, CONTIG>,  \
     (axis_kernel_t)k_axis_blue<T, OP_C2R_FULL, CONTIG>, (axis_kernel_t)k_axis_blue<T, OP_DCT2, CONTIG>, \
     (axis_kernel_t)k_axis_blue<T, OP_DCT3, CONTIG>,     (axis_kernel_t)k_axis_blue<T, OP_DHT, CONTIG>,  \
     (axis_kernel_t)k_axis_blue<T, OP_DST2, CONTIG>,     (axis_kernel_t)k_axis_blue<T, OP_DST3, CONTIG>, \
     (axis_kernel_t)k_axis_blue<T, OP_DCT1, CONTIG>,     (axis_kernel_t)k_axis_blue<T, OP_DST1, CONTIG>, \
     (axis_kernel_t)k_axis_blue<T, OP_DCT4, CONTIG>,     (axis_kernel_t)k_axis_blue<T, OP_DST4, CONTIG>}

// [precision][contig][op]
static axis_kernel_t g_blue_kernels[2][2][OP_COUNT] = {
    {DFFT_BLUE_OPS(float, false), DFFT_BLUE_OPS(float, true)},
    {DFFT_BLUE_OPS(double, false), DFFT_BLUE_OPS(double, true)},
};

axis_kernel_t blue_kernel(int prec, int contig, int op) {
    if (op < 0 || op >= OP_COUNT) return nullptr;
    return g_blue_kernels[prec ? 1 : 0][contig ? 1 : 0][op];
}

// Bluestein on the compile-time engine (k_blue_fixed, dfft_blue_fixed.hpp):
// C2C of convolution lengths 2^k (32..8192) here, 3*2^k and 5*2^k in
// dfft_blue_fixed2.hip, the real ops in dfft_blue_real64/32.hip
DFFT_TABLE(g_blue_pow2, DFFT_BLUE_M(32), DFFT_BLUE_M(64), DFFT_BLUE_M(128), DFFT_BLUE_M(256), DFFT_BLUE_M(512),
           DFFT_BLUE_M(1024), DFFT_BLUE_M(2048), DFFT_BLUE_M(4096), DFFT_BLUE_M(8192), DFFT_BLUE_M(16384))
extern const FixedTable g_blue_mixed, g_blue_real64, g_blue_real32;

static const FixedEntry *blue_fixed_entry(int prec, int contig, int op, int M) {
    for (const FixedTable *t : {&g_blue_pow2, &g_blue_mixed, &g_blue_real64, &g_blue_real32})
        for (int i = 0; i < t->n; ++i) {
            const FixedEntry &e = t->e[i];
            if (e.prec == (prec ? 1 : 0) && e.contig == (contig ? 1 : 0) && e.op == op && e.L == M) return &e;
        }
    return nullptr;
}

std::vector<int> blue_fixed_lengths(int prec, int contig, int op) {
    std::vector<int> out;
    for (const FixedTable *t : {&g_blue_pow2, &g_blue_mixed, &g_blue_real64, &g_blue_real32})
        for (int i = 0; i < t->n; ++i)
            if (t->e[i].prec == (prec ? 1 : 0) && t->e[i].contig == (contig ? 1 : 0) && t->e[i].op == op)
                out.push_back(t->e[i].L);
    return out;
}

bool blue_fixed_schedule(int prec, int contig, int op, int M, std::vector<int> &radices, int &tpl, int &variant) {
    const FixedEntry *e = blue_fixed_entry(prec, contig, op, M);
    if (!e) return false;
    radices.assign(e->R, e->R + e->S);
    tpl = e->tpl;
    variant = e->variant;
    return true;
}

axis_kernel_t blue_fixed_kernel(int prec, int contig, int op, int M) {
    const FixedEntry *e = blue_fixed_entry(prec, contig, op, M);
    return e ? e->fn : nullptr;
}

// Rader kernels (dfft_rader_*.hip): C2C of prime lengths M + 1
extern const FixedTable g_rader_a, g_rader_b, g_rader_c, g_rader_d;

static const FixedEntry *rader_fixed_entry(int prec, int contig, int M) {
    for (const FixedTable *t : {&g_rader_a, &g_rader_b, &g_rader_c, &g_rader_d})
        for (int i = 0; i < t->n; ++i) {
            const FixedEntry &e = t->e[i];
            if (e.prec == (prec ? 1 : 0) && e.contig == (contig ? 1 : 0) && e.L == M) return &e;
        }
    return nullptr;
}

bool rader_fixed_schedule(int prec, int contig, int M, std::vector<int> &radices, int &tpl) {
    const FixedEntry *e = rader_fixed_entry(prec, contig, M);
    if (!e) return false;
    radices.assign(e->R, e->R + e->S);
    tpl = e->tpl;
    return true;
}

axis_kernel_t rader_fixed_kernel(int prec, int contig, int M) {
    const FixedEntry *e = rader_fixed_entry(prec, contig, M);
    return e ? e->fn : nullptr;
}

}  // namespace dfft
