// dfft_kernels.hpp -- host-side launchers of the HIP kernels.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include <vector>

#include "dfft_types.hpp"

namespace dfft {

int axis_lds_limit();
hipError_t launch_axis(const AxisPass &a, int prec, int64_t nblocks, int threads, size_t lds, hipStream_t stream);
void *axis_kernel_fn(const AxisPass &a, int prec);
// compile-time schedule of (precision, layout, variant, op, length), if one is
// built: `radices` in = the wanted radix list (empty: the default one), out =
// the kernel's list
bool fixed_schedule(int prec, int contig, int variant, int op, int L, std::vector<int> &radices, int &tpl);
// every radix list the catalog holds for (precision, layout, variant, op, length)
std::vector<std::vector<int>> fixed_lists(int prec, int contig, int variant, int op, int L);
// compile-time Bluestein schedule of (precision, layout, op, convolution
// length M) (dfft_blue*.hip), if one is built; variant 2 split / 3 full exchange
bool blue_fixed_schedule(int prec, int contig, int op, int M, std::vector<int> &radices, int &tpl, int &variant);
// the convolution lengths that have such a schedule
std::vector<int> blue_fixed_lengths(int prec, int contig, int op);
// compile-time Rader schedule (C2C, split exchange) of the prime length M + 1, if built
bool rader_fixed_schedule(int prec, int contig, int M, std::vector<int> &radices, int &tpl);
hipError_t launch_fill(void *out, int64_t nscalar, int prec, uint64_t seed, hipStream_t stream);
hipError_t launch_box(const void *src, int64_t s_off, int64_t s_s1, int64_t s_s2, void *dst, int64_t d_off,
                      int64_t d_s1, int64_t d_s2, const int64_t *cnt, int esize, hipStream_t stream);
hipError_t launch_copy16(const void *src, void *dst, int64_t n16, hipStream_t stream);
hipError_t launch_permute(const void *in, void *out, const int64_t *n, const int *perm, int esize,
                          hipStream_t stream);

}  // namespace dfft
