// dfft_fixed.hpp -- the catalog of compile-time axis-pass kernels.
//
// A catalog kernel has its length, radices, values per thread and threads
// per line fixed at compile time: every stage is unrolled with constant
// radix and constant index arithmetic (j mod Ns is a mask or a multiply),
// which keeps the register count low enough for several tiles per CU.
//
// Variants (AxisPass::variant):
//   1  8 values per thread, whole complex values exchanged through LDS
//   2  16 values per thread, real and imaginary parts exchanged in two
//      rounds through half the LDS (C2C, R2C, C2R and DCT-II/III)
//   3  16 values per thread, whole complex values exchanged through LDS
//      (the mixed-radix lengths: radix-3/5/7 stages leave at most 1-2 of
//      the 16 registers idle, against 3 of 8)
// The catalog is spread over several translation units so it compiles in
// parallel; fixed_entry() in dfft_kernels.hip searches all of them.
#pragma once
#include <hip/hip_runtime.h>

#include "dfft_axis.hpp"

namespace dfft {

using axis_kernel_t = void (*)(AxisPass);

template <class T, int OP, bool CONTIG, int EMAX, bool SPLIT, class RL, int TPLC>
__global__ __launch_bounds__(1024) void k_axis_fixed(AxisPass a) {
    extern __shared__ __align__(16) unsigned char dfft_smem[];
    axis_tile<EMAX, T, OP, CONTIG, SPLIT, RL, TPLC>(a, a.in, a.out, pass_tile(a), (int)threadIdx.x,
                                                     reinterpret_cast<cx<T> *>(dfft_smem));
}

struct FixedEntry {
    int prec, op, contig, variant, L, tpl, S;
    const int *R;
    axis_kernel_t fn;
};

struct FixedTable {
    const FixedEntry *e;
    int n;
};

// the Bluestein kernels (dfft_blue.hip): runtime engine of (precision, layout,
// op); compile-time C2C engine of (precision, layout, convolution length)
axis_kernel_t blue_kernel(int prec, int contig, int op);
axis_kernel_t blue_fixed_kernel(int prec, int contig, int op, int M);
// the Rader kernels (dfft_rader_*.hip): C2C of prime length M + 1
axis_kernel_t rader_fixed_kernel(int prec, int contig, int M);

#define DFFT_FX(T, P, OPV, CONTIG, VAR, EMAXV, SPLIT, RL)                                             \
    {P, OPV, CONTIG, VAR, RL::L, fixed_tpl<RL, EMAXV>(), RL::S, RL::R,                                 \
     (axis_kernel_t)k_axis_fixed<T, OPV, CONTIG, EMAXV, SPLIT, RL, fixed_tpl<RL, EMAXV>()>}

// C2C of length L, 16 values per thread: full (3) and split (2) LDS exchange,
// contiguous and strided
#define DFFT_C2C_16(T, P, L)                                           \
    DFFT_FX(T, P, OP_C2C, 1, 3, 16, false, RadicesOf16<L>),         \
        DFFT_FX(T, P, OP_C2C, 1, 2, 16, true, RadicesOf16<L>),      \
        DFFT_FX(T, P, OP_C2C, 0, 3, 16, false, RadicesOf16<L>),     \
        DFFT_FX(T, P, OP_C2C, 0, 2, 16, true, RadicesOf16<L>)
#define DFFT_C2C_F64(L) DFFT_C2C_16(double, 1, L)
#define DFFT_C2C_F32(L) DFFT_C2C_16(float, 0, L)
// R2C of a real line of 2H points through an H-point FFT (contiguous axis):
// full (3) and split (2) LDS exchange.  The split schedule runs the
// post-twiddle in two rounds (real parts, then imaginary parts) and keeps two
// partial sums per output bin of the thread between them, so it is built
// only where a line has enough threads (H >= 100: 7 or more) to hold them in
// registers; below that the FP64 kernels spill.
#define DFFT_R2C_FULL(H)                                               \
    DFFT_FX(double, 1, OP_R2C, 1, 3, 16, false, RadicesOf16<H>),    \
        DFFT_FX(float, 0, OP_R2C, 1, 3, 16, false, RadicesOf16<H>)
#define DFFT_R2C(H)                                                    \
    DFFT_R2C_FULL(H), DFFT_FX(double, 1, OP_R2C, 1, 2, 16, true, RadicesOf16<H>), \
        DFFT_FX(float, 0, OP_R2C, 1, 2, 16, true, RadicesOf16<H>)
// C2R of H+1 bins to a real line of 2H points (contiguous axis): the
// pre-twiddle reads bins p and H-p from global memory, so both exchanges apply
#define DFFT_C2R(H)                                                    \
    DFFT_FX(double, 1, OP_C2R, 1, 3, 16, false, RadicesOf16<H>),    \
        DFFT_FX(float, 0, OP_C2R, 1, 3, 16, false, RadicesOf16<H>), \
        DFFT_FX(double, 1, OP_C2R, 1, 2, 16, true, RadicesOf16<H>), \
        DFFT_FX(float, 0, OP_C2R, 1, 2, 16, true, RadicesOf16<H>)

// C2R with the split exchange only (the long half-lengths)
#define DFFT_C2R_SPLIT(H)                                              \
    DFFT_FX(double, 1, OP_C2R, 1, 2, 16, true, RadicesOf16<H>),     \
        DFFT_FX(float, 0, OP_C2R, 1, 2, 16, true, RadicesOf16<H>)

// DCT-II / DCT-III (REDFT10 / REDFT01, Makhoul's n-point algorithm) of
// length L, split LDS exchange, both layouts: their fused load and store maps
// touch only global memory, so the split exchange applies
#define DFFT_DCT(T, P, L)                                              \
    DFFT_FX(T, P, OP_DCT2, 1, 2, 16, true, RadicesOf16<L>),           \
        DFFT_FX(T, P, OP_DCT2, 0, 2, 16, true, RadicesOf16<L>),       \
        DFFT_FX(T, P, OP_DCT3, 1, 2, 16, true, RadicesOf16<L>),       \
        DFFT_FX(T, P, OP_DCT3, 0, 2, 16, true, RadicesOf16<L>)

// C2C of a length 1025..4096 that is not a power of two: split exchange,
// both layouts (RL: the radix list)
#define DFFT_MID_RL(T, P, RL)                                          \
    DFFT_FX(T, P, OP_C2C, 1, 2, 16, true, RL), DFFT_FX(T, P, OP_C2C, 0, 2, 16, true, RL)
#define DFFT_MID(T, P, L) DFFT_MID_RL(T, P, RadicesOf16<L>)

#define DFFT_TABLE(name, ...)                                           \
    static const FixedEntry name##_e[] = {__VA_ARGS__};                 \
    extern const FixedTable name;                                       \
    const FixedTable name = {name##_e, (int)(sizeof(name##_e) / sizeof(name##_e[0]))};

}  // namespace dfft
