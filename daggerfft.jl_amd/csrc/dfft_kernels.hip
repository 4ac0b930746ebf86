// dfft_kernels.hip -- HIP kernels of libdaggerfft_hip for gfx950.
//
//   k_axis      one Stockham axis pass per launch (C2C / R2C / C2R / every r2r kind);
//               catalog kernels in dfft_fixed_*.hip, Bluestein kernels in dfft_blue.hip
//   k_fill      SplitMix64 counter-based uniform[0,1) input generator (k_fill16:
//               the same values, 16 bytes per lane per store)
//   k_box       box gather / scatter / copy (pack_kernel!, unpack_kernel!,
//               local_copy_kernel!; src/fftgpu.jl:452-495), coalesced along axis 0
//   k_permute   LDS-tiled 3D axis permutation (the pencil transpose)
//   k_copy16    plain streaming copy (the measured HBM copy ceiling)
#include <hip/hip_runtime.h>

#include "dfft_axis.hpp"
#include "dfft_fixed.hpp"
#include "dfft_kernels.hpp"
#include "dfft_plan.hpp"

#include <algorithm>
#include <vector>

namespace dfft {

// ---------------------------------------------------------------------------
// axis pass
// ---------------------------------------------------------------------------
template <class T, int OP, bool CONTIG, int EMAX>
__global__ __launch_bounds__(1024) void k_axis(AxisPass a) {
    extern __shared__ __align__(16) unsigned char dfft_smem[];
    axis_tile<EMAX, T, OP, CONTIG>(a, a.in, a.out, pass_tile(a), (int)threadIdx.x,
                                   reinterpret_cast<cx<T> *>(dfft_smem));
}

#define DFFT_OPS(T, CONTIG)                                                                   \
    {(axis_kernel_t)k_axis<T, OP_C2C, CONTIG, 8>, (axis_kernel_t)k_axis<T, OP_R2C, CONTIG, 8>, \
     (axis_kernel_t)k_axis<T, OP_R2C_FULL, CONTIG, 8>, (axis_kernel_t)k_axis<T, OP_C2R, CONTIG, 8>, \
     (axis_kernel_t)k_axis<T, OP_C2R_FULL, CONTIG, 8>, (axis_kernel_t)k_axis<T, OP_DCT2, CONTIG, 8>, \
     (axis_kernel_t)k_axis<T, OP_DCT3, CONTIG, 8>, (axis_kernel_t)k_axis<T, OP_DHT, CONTIG, 8>,   \
     (axis_kernel_t)k_axis<T, OP_DST2, CONTIG, 8>, (axis_kernel_t)k_axis<T, OP_DST3, CONTIG, 8>,  \
     (axis_kernel_t)k_axis<T, OP_DCT1, CONTIG, 8>, (axis_kernel_t)k_axis<T, OP_DST1, CONTIG, 8>,  \
     (axis_kernel_t)k_axis<T, OP_DCT4, CONTIG, 8>, (axis_kernel_t)k_axis<T, OP_DST4, CONTIG, 8>}

// runtime-schedule kernels: [precision][contig][op]
static axis_kernel_t g_axis_kernels[2][2][OP_COUNT] = {
    {DFFT_OPS(float, false), DFFT_OPS(float, true)},
    {DFFT_OPS(double, false), DFFT_OPS(double, true)},
};

// ---------------------------------------------------------------------------
// compile-time schedules (the catalog tables live in dfft_fixed_*.hip)
// ---------------------------------------------------------------------------
extern const FixedTable g_fixed_pow2, g_fixed_f64a, g_fixed_f64b, g_fixed_f64c, g_fixed_f32a, g_fixed_f32b,
    g_fixed_f32c, g_fixed_r2ca, g_fixed_r2cb, g_fixed_c2ra, g_fixed_c2rb, g_fixed_dct, g_fixed_long, g_fixed_mid64a,
    g_fixed_mid64b, g_fixed_mid32a, g_fixed_mid32b, g_fixed_reala, g_fixed_realb, g_fixed_r2c32, g_fixed_dcth;
static const FixedTable *const g_fixed_tables[] = {
    &g_fixed_pow2,  &g_fixed_f64a,  &g_fixed_f64b,   &g_fixed_f64c,   &g_fixed_f32a,   &g_fixed_f32b,
    &g_fixed_f32c,  &g_fixed_r2ca,  &g_fixed_r2cb,   &g_fixed_c2ra,   &g_fixed_c2rb,   &g_fixed_dct,
    &g_fixed_long,  &g_fixed_mid64a, &g_fixed_mid64b, &g_fixed_mid32a, &g_fixed_mid32b, &g_fixed_reala,
    &g_fixed_realb, &g_fixed_r2c32, &g_fixed_dcth};

// The catalog kernel of (precision, layout, variant, op, length) with radix
// list R[0..S) -- or, with R == nullptr, the first one listed (the default).
static const FixedEntry *fixed_entry(int prec, int contig, int variant, int op, int L, const int *R, int S) {
    if (variant < 1) return nullptr;
    for (const FixedTable *t : g_fixed_tables)
        for (int i = 0; i < t->n; ++i) {
            const FixedEntry &e = t->e[i];
            if (e.prec != (prec ? 1 : 0) || e.op != op || e.contig != (contig ? 1 : 0) || e.variant != variant ||
                e.L != L)
                continue;
            if (R && (e.S != S || !std::equal(R, R + S, e.R))) continue;
            return &e;
        }
    return nullptr;
}

bool fixed_schedule(int prec, int contig, int variant, int op, int L, std::vector<int> &radices, int &tpl) {
    const FixedEntry *e = radices.empty() ? fixed_entry(prec, contig, variant, op, L, nullptr, 0)
                                          : fixed_entry(prec, contig, variant, op, L, radices.data(),
                                                        (int)radices.size());
    if (!e) return false;
    radices.assign(e->R, e->R + e->S);
    tpl = e->tpl;
    return true;
}

std::vector<std::vector<int>> fixed_lists(int prec, int contig, int variant, int op, int L) {
    std::vector<std::vector<int>> out;
    for (const FixedTable *t : g_fixed_tables)
        for (int i = 0; i < t->n; ++i) {
            const FixedEntry &e = t->e[i];
            if (e.prec == (prec ? 1 : 0) && e.op == op && e.contig == (contig ? 1 : 0) && e.variant == variant &&
                e.L == L)
                out.emplace_back(e.R, e.R + e.S);
        }
    return out;
}

static axis_kernel_t axis_kernel(const AxisPass &a, int prec) {
    if (a.blue == 2) return rader_fixed_kernel(prec, a.contig, a.bsp.L);
    if (a.blue) return a.variant ? blue_fixed_kernel(prec, a.contig, a.op, a.bsp.L) : blue_kernel(prec, a.contig, a.op);
    if (a.variant >= 1) {
        const FixedEntry *e = fixed_entry(prec, a.contig, a.variant, a.op, a.sp.L, a.sp.R, a.sp.S);
        return e ? e->fn : nullptr;
    }
    if (a.op < 0 || a.op >= OP_COUNT) return nullptr;  // catalog-only ops
    return g_axis_kernels[prec ? 1 : 0][a.contig ? 1 : 0][a.op];
}

// LDS per workgroup of the device: read once (a function-local static is
// initialised once even when several threads create plans at the same time)
int axis_lds_limit() {
    static const int limit = [] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
            return v;
        return 65536;
    }();
    return limit;
}

hipError_t launch_axis(const AxisPass &a, int prec, int64_t nblocks, int threads, size_t lds,
                       hipStream_t stream) {
    axis_kernel_t k = axis_kernel(a, prec);
    if (!k) return hipErrorInvalidDeviceFunction;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k, dim3((unsigned)nblocks), dim3(threads), lds, stream, a);
    return hipGetLastError();
}

void *axis_kernel_fn(const AxisPass &a, int prec) {
    return (void *)axis_kernel(a, prec);
}

// ---------------------------------------------------------------------------
// input generator: u(seed, i) = (mix64(seed + (i+1) * golden) >> 11) * 2^-53
// ---------------------------------------------------------------------------
__device__ __forceinline__ double splitmix_u(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * 0x1.0p-53;
}

template <class T>
__global__ __launch_bounds__(256) void k_fill(T *out, int64_t nscalar, uint64_t seed) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nscalar; i += stride)
        out[i] = (T)splitmix_u(seed, (uint64_t)i);
}

// The same values, 16 bytes per lane per store: thread i writes scalars
// V i .. V i + V - 1 (V = 4 floats or 2 doubles) of a 16-byte aligned `out`;
// the scalars past the last whole vector are written one per lane.  With
// one 4-byte store per lane, the ComplexF32 inputs of the served stream
// filled at well under the write bandwidth.
template <class T>
struct alignas(16) FillVec {
    T v[16 / sizeof(T)];
};
template <class T>
__global__ __launch_bounds__(256) void k_fill16(T *out, int64_t nscalar, uint64_t seed) {
    constexpr int V = 16 / sizeof(T);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x, nvec = nscalar / V;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    FillVec<T> *o = reinterpret_cast<FillVec<T> *>(out);
    for (int64_t i = t; i < nvec; i += stride) {
        FillVec<T> f;
#pragma unroll
        for (int k = 0; k < V; ++k) f.v[k] = (T)splitmix_u(seed, (uint64_t)(i * V + k));
        o[i] = f;
    }
    for (int64_t i = nvec * V + t; i < nscalar; i += stride) out[i] = (T)splitmix_u(seed, (uint64_t)i);
}

hipError_t launch_fill(void *out, int64_t nscalar, int prec, uint64_t seed, hipStream_t stream) {
    if (nscalar <= 0) return hipSuccess;
    if (((uintptr_t)out & 15) == 0) {
        const int V = prec ? 2 : 4;
        int64_t blocks = (nscalar / V + 255) / 256;
        if (blocks < 1) blocks = 1;
        if (blocks > 65536) blocks = 65536;
        if (prec) hipLaunchKernelGGL(k_fill16<double>, dim3((unsigned)blocks), dim3(256), 0, stream, (double *)out, nscalar, seed);
        else hipLaunchKernelGGL(k_fill16<float>, dim3((unsigned)blocks), dim3(256), 0, stream, (float *)out, nscalar, seed);
        return hipGetLastError();
    }
    int64_t blocks = (nscalar + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (prec) hipLaunchKernelGGL(k_fill<double>, dim3((unsigned)blocks), dim3(256), 0, stream, (double *)out, nscalar, seed);
    else hipLaunchKernelGGL(k_fill<float>, dim3((unsigned)blocks), dim3(256), 0, stream, (float *)out, nscalar, seed);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// plain copy, 16 B per lane, streaming: the HBM copy ceiling the axis passes
// are measured against.  Each thread issues its four loads before its four
// stores, so a workgroup has 16 KB in flight.
// ---------------------------------------------------------------------------
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_copy16(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst, int64_t n16) {
    const int64_t base = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    u32x4 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (base + 256 * k < n16) v[k] = __builtin_nontemporal_load(src + base + 256 * k);
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (base + 256 * k < n16) __builtin_nontemporal_store(v[k], dst + base + 256 * k);
}

hipError_t launch_copy16(const void *src, void *dst, int64_t n16, hipStream_t stream) {
    if (n16 <= 0) return hipSuccess;
    const int64_t blocks = (n16 + 1023) / 1024;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, stream, (const u32x4 *)src, (u32x4 *)dst, n16);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// box gather / scatter / copy.  Threads run along axis 0 of the box, so
// both the strided array side and the packed buffer side are contiguous
// bursts; blocks enumerate (i-chunk, j, k) of the box.
//   dst[d_off + i*1 + j*d_s1 + k*d_s2] = src[s_off + i + j*s_s1 + k*s_s2]
// ---------------------------------------------------------------------------
struct BoxArgs {
    const void *src;
    void *dst;
    int64_t s_off, s_s1, s_s2;
    int64_t d_off, d_s1, d_s2;
    int64_t c0, c1, c2;
    int64_t nchunk0;
};

template <class E>
__global__ __launch_bounds__(256) void k_box(BoxArgs b) {
    const E *src = reinterpret_cast<const E *>(b.src);
    E *dst = reinterpret_cast<E *>(b.dst);
    const int64_t blk = blockIdx.x;
    const int64_t ic = blk % b.nchunk0;
    const int64_t jk = blk / b.nchunk0;
    const int64_t j = jk % b.c1, k = jk / b.c1;
    const int64_t i = ic * 256 + threadIdx.x;
    if (i < b.c0) dst[b.d_off + i + j * b.d_s1 + k * b.d_s2] = src[b.s_off + i + j * b.s_s1 + k * b.s_s2];
}

hipError_t launch_box(const void *src, int64_t s_off, int64_t s_s1, int64_t s_s2, void *dst, int64_t d_off,
                      int64_t d_s1, int64_t d_s2, const int64_t *cnt, int esize, hipStream_t stream) {
    BoxArgs b{src, dst, s_off, s_s1, s_s2, d_off, d_s1, d_s2, cnt[0], cnt[1], cnt[2], (cnt[0] + 255) / 256};
    const int64_t nblocks = b.nchunk0 * cnt[1] * cnt[2];
    if (nblocks == 0) return hipSuccess;
    if (nblocks > 0x7fffffffLL) return hipErrorInvalidValue;
    switch (esize) {
    case 4: hipLaunchKernelGGL(k_box<uint32_t>, dim3((unsigned)nblocks), dim3(256), 0, stream, b); break;
    case 8: hipLaunchKernelGGL(k_box<uint2>, dim3((unsigned)nblocks), dim3(256), 0, stream, b); break;
    case 16: hipLaunchKernelGGL(k_box<uint4>, dim3((unsigned)nblocks), dim3(256), 0, stream, b); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// axis permutation through an LDS tile.  Input axis 0 (contiguous in the
// input) and input axis `a` (contiguous in the output) span a TILE x TILE
// tile; the remaining axis indexes the tile's plane.
//   out[sum_d idx[d] * ost[d]] = in[idx[0] + idx[1]*n0 + idx[2]*n0*n1]
// ---------------------------------------------------------------------------
struct PermArgs {
    const void *in;
    void *out;
    int64_t n[3];
    int64_t ist[3];
    int64_t ost[3];
    int a;           // input axis that becomes output axis 0 (a != 0)
    int b;           // remaining input axis
    int64_t t0, ta;  // tiles along axis 0 and axis a
};

template <class E, int TILE>
__global__ __launch_bounds__(256) void k_permute(PermArgs p) {
    __shared__ E tile[TILE][TILE + 1];
    const E *in = reinterpret_cast<const E *>(p.in);
    E *out = reinterpret_cast<E *>(p.out);
    int64_t blk = blockIdx.x;
    const int64_t tx = blk % p.t0;
    blk /= p.t0;
    const int64_t ty = blk % p.ta;
    const int64_t zb = blk / p.ta;
    const int lx = threadIdx.x % TILE, ly = threadIdx.x / TILE;
    constexpr int ROWS = 256 / TILE;
    const int64_t i0 = tx * TILE, q0 = ty * TILE;
    for (int r = ly; r < TILE; r += ROWS) {
        const int64_t i = i0 + lx, q = q0 + r;
        if (i < p.n[0] && q < p.n[p.a])
            tile[r][lx] = in[i * p.ist[0] + q * p.ist[p.a] + zb * p.ist[p.b]];
    }
    __syncthreads();
    for (int r = ly; r < TILE; r += ROWS) {
        const int64_t q = q0 + lx, i = i0 + r;
        if (i < p.n[0] && q < p.n[p.a])
            out[i * p.ost[0] + q * p.ost[p.a] + zb * p.ost[p.b]] = tile[lx][r];
    }
}

hipError_t launch_permute(const void *in, void *out, const int64_t *n, const int *perm, int esize,
                          hipStream_t stream) {
    // output strides per input axis: output axis d is input axis perm[d]
    int64_t on[3] = {n[perm[0]], n[perm[1]], n[perm[2]]};
    int64_t ost[3];
    ost[perm[0]] = 1;
    ost[perm[1]] = on[0];
    ost[perm[2]] = on[0] * on[1];
    int64_t ist[3] = {1, n[0], n[0] * n[1]};
    if (perm[0] == 0) {
        // axis 0 stays contiguous: a strided row copy, one launch per plane of the last input axis
        int64_t cnt[3] = {n[0], n[1], n[2]};
        return launch_box(in, 0, ist[1], ist[2], out, 0, ost[1], ost[2], cnt, esize, stream);
    }
    PermArgs p{};
    p.in = in;
    p.out = out;
    for (int d = 0; d < 3; ++d) { p.n[d] = n[d]; p.ist[d] = ist[d]; p.ost[d] = ost[d]; }
    p.a = perm[0];
    p.b = 3 - p.a;  // axes {0, a, b} = {0, 1, 2}
    const int TILE = esize == 16 ? 32 : 64;
    p.t0 = (n[0] + TILE - 1) / TILE;
    p.ta = (n[p.a] + TILE - 1) / TILE;
    const int64_t nblocks = p.t0 * p.ta * n[p.b];
    if (nblocks == 0) return hipSuccess;
    if (nblocks > 0x7fffffffLL) return hipErrorInvalidValue;
    switch (esize) {
    case 4: hipLaunchKernelGGL((k_permute<uint32_t, 64>), dim3((unsigned)nblocks), dim3(256), 0, stream, p); break;
    case 8: hipLaunchKernelGGL((k_permute<uint2, 64>), dim3((unsigned)nblocks), dim3(256), 0, stream, p); break;
    case 16: hipLaunchKernelGGL((k_permute<uint4, 32>), dim3((unsigned)nblocks), dim3(256), 0, stream, p); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace dfft
