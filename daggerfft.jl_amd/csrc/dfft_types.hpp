// dfft_types.hpp -- plain-old-data descriptors shared by the host planner and
// the HIP kernels.  One AxisPass describes one kernel launch: a batch of 1D
// transforms of length L along one axis of a column-major array.
#pragma once
#include <stdint.h>

namespace dfft {

constexpr int MAX_STAGES = 14;
// AxisPass::swz: contiguous tile ranges per XCD (gfx950: 8 XCDs, workgroups dealt round robin)
constexpr int SWZ_XCD = -8;

// Operation fused into an axis pass.
enum PassOp : int {
    OP_C2C = 0,       // complex -> complex (sign via `inverse`)
    OP_R2C = 1,       // real line of length 2H -> H+1 bins via an H-point complex FFT + post-twiddle
    OP_R2C_FULL = 2,  // real line of length L (odd) -> L/2+1 bins via an L-point complex FFT
    OP_C2R = 3,       // H+1 bins -> real line of length 2H via pre-twiddle + H-point inverse FFT
    OP_C2R_FULL = 4,  // L/2+1 bins -> real line of length L (odd) via an L-point inverse FFT
    OP_DCT2 = 5,      // REDFT10 via an L-point FFT of the even/odd reordered line (Makhoul 1980)
    OP_DCT3 = 6,      // REDFT01 via a pre-twiddled L-point inverse FFT (Makhoul 1980)
    // the other FFTW r2r kinds; n = logical line length (AxisPass::nlog)
    OP_DHT = 7,       // DHT: L = n; H = Re F - Im F of the real line's DFT
    OP_DST2 = 8,      // RODFT10: the DCT-II path on (-1)^j x_j, written in reverse order
    OP_DST3 = 9,      // RODFT01: the DCT-III path on the reversed line, times (-1)^k
    OP_DCT1 = 10,     // REDFT00: L = 2(n-1) FFT of the even extension, real part of bins 0..n-1
    OP_DST1 = 11,     // RODFT00: L = 2(n+1) FFT of the odd extension, -Im of bins 1..n
    OP_DCT4 = 12,     // REDFT11: L = 2n FFT of x_j e^{-i pi j/2n}; 2 Re(e^{-i pi (2k+1)/4n} F_k)
    OP_DST4 = 13,     // RODFT11: same FFT; -2 Im(e^{-i pi (2k+1)/4n} F_k)
    OP_COUNT = 14,    // ops of the runtime engine (k_axis); the two below are catalog-only
    // REDFT10 / REDFT01 of n = 2H reals (H even) through an H-point complex FFT
    OP_DCT2H = 14,    // FFT of z_m = v_2m + i v_2m+1 (Makhoul's v), then R2C and DCT-II post-twiddles
    OP_DCT3H = 15     // DCT-III and C2R pre-twiddles, H-point inverse FFT, z unpacked into the line
};

// Stockham schedule of one line length.
struct StagePlan {
    int L;                    // complex FFT length
    int S;                    // number of stages
    int R[MAX_STAGES];        // radix of each stage (2,3,4,5,7,8)
    int tw[MAX_STAGES];       // offset (complex elements) of each stage's twiddle block
    int post;                 // offset of the pre/post-twiddle block (R2C/C2R/DCT), -1 if none
};

// Geometry + schedule of one launch.  Element strides are counted in
// elements of the array's own type (real or complex).
struct AxisPass {
    const void *in;
    void *out;
    const void *tw;           // twiddle table of the pass precision (complex)
    StagePlan sp;
    int op;
    int inverse;              // 1: backward transform (conjugate on load/store)
    double scale;             // applied on store
    int contig;               // 1: each line is contiguous (axis 0); 0: columns contiguous
    int TPL;                  // threads per line
    int TC;                   // lines (contig) / columns (strided) per workgroup
    int lds_pitch;            // contig: elements per line slot; strided: elements per row
    int pad_shift;            // contig: slot index p -> p + (p >> pad_shift)
    int nlog;                 // logical line length (real length for R2C/C2R/R2R)
    int nt;                   // 1: streaming (nontemporal) global loads/stores
    int paired_loads;         // ComplexF64 C2C on the compile-time engine: 1 issues the first
                              // stage's loads in the load map's order (a couple in flight per
                              // wave at a time), 0 issues them all before any is used
                              // (paired_loads_for)
    int variant;              // 0: runtime schedule (any length); 1: compile-time schedule of this
                              // length, 8 values/thread; 2: compile-time schedule, 16 values/thread,
                              // real and imaginary parts exchanged in two rounds (half the LDS);
                              // 3: compile-time schedule, 16 values/thread
    int swz;                  // workgroup b runs tile (b * swz) mod (number of tiles); 1 = in order.
                              // Coprime to the tile count, so every tile runs once; spreads the
                              // tiles running at one time over the array (HBM channel balance
                              // for lines with a large power-of-two stride).  SWZ_XCD: each of
                              // the 8 XCDs (which take workgroups round robin) runs one
                              // contiguous range of tiles, so tiles that share cache lines
                              // (rows not aligned to 128 B) share that XCD's L2
    int64_t nlines;           // contig: number of lines; strided: number of columns
    int64_t nouter;           // strided: number of outer slices (1 for contig)
    int64_t in_sm, in_so;     // input: stride along the line, stride between lines/outer
    int64_t out_sm, out_so;   // output: same
    // Bluestein (chirp-z) pass: a length sp.L with a prime factor > 7 runs its
    // DFT as a cyclic convolution of length bsp.L >= 2*sp.L - 1, through two
    // bsp.L-point Stockham FFTs in LDS (sp then holds no stages, only sp.L
    // and the op's pre/post block).  0: an ordinary pass.
    int blue;
    int chirp;                // offset of c_p = exp(-i pi p^2 / sp.L), p < sp.L
    int bhat;                 // offset of FFT_M(b)_k / M, b_m = conj(c_m) = b_{M-m}, M = bsp.L
    StagePlan bsp;            // schedule of the convolution length M
    StagePlan bsp2;           // compile-time Bluestein (variant 2): the second FFT's schedule,
                              // bsp's radices reversed (its first stage takes the first FFT's
                              // last-stage outputs straight from registers)
};

// AxisPass::paired_loads for a pass whose lines have stride `in_sm`
// elements of `elem_bytes`: the first stage's loads are all issued at once
// (faster on every measured shape: 3-9% on 384x240x210, the 64-transform
// batch and the served mix) except along ComplexF64 lines whose stride is a
// power of two of 8 KiB or more, where 512^3's strided passes measured 2-3%
// slower that way.  ComplexF32 kernels hold only the all-at-once order:
// with both orders the 512-point strided kernel took 65 VGPRs, one past the
// 64 that let two 1024-thread workgroups share a CU, and ran 24% slower,
// while all at once alone matched the paired order on 512^3.
inline int paired_loads_for(int64_t in_sm, int elem_bytes) {
    const int64_t b = in_sm * (int64_t)elem_bytes;
    return (elem_bytes == 16 && b >= 8192 && (b & (b - 1)) == 0) ? 1 : 0;
}

}  // namespace dfft
