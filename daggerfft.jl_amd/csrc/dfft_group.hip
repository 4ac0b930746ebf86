// dfft_group.hip -- grouped launches: tiles of many independent axis passes
// in one kernel launch.
//
// A group is a sequence of launches on one stream.  Each launch holds a list
// of items (an item = one axis pass over a slice of two pointer slots); its
// workgroups are dealt out to the items' tiles through a block -> item map,
// exactly as the single-pass kernel deals them to one pass.  Items inside a
// launch are independent; every dependency between items is a launch
// boundary, so no workgroup ever waits for another.
//
// Users:
//   * dfft_batch_*: launch l runs pass l of every transform of the batch;
//   * Slab plans: the 2D stage is software-pipelined over chunks of z-planes:
//     launch i runs the dims[0] pass of chunk i together with the dims[1]
//     pass of chunk i-1, which the previous launch just wrote, so that chunk
//     is read back while it is still in the 256 MB memory-side Infinity
//     Cache.  The dims[2] pass follows as one ordinary launch.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/daggerfft_hip.h"
#include "dfft_axis.hpp"
#include "dfft_group.hpp"
#include "dfft_kernels.hpp"
#include "dfft_plan.hpp"

namespace dfft {

struct GroupArgs {
    const GroupItem *items;     // items of this launch
    const int *block_item;      // block -> item (index into `items`)
    void *slots[GROUP_MAX_SLOTS];
};

template <class T>
__device__ __forceinline__ void group_tile(const AxisPass &a, const void *in, void *out, int64_t tile, int tid,
                                           void *smem) {
    cx<T> *lds = reinterpret_cast<cx<T> *>(smem);
    switch (a.op) {
    case OP_C2C:
        if (a.contig) axis_tile<EMAX, T, OP_C2C, true>(a, in, out, tile, tid, lds);
        else axis_tile<EMAX, T, OP_C2C, false>(a, in, out, tile, tid, lds);
        break;
    case OP_R2C:
        if (a.contig) axis_tile<EMAX, T, OP_R2C, true>(a, in, out, tile, tid, lds);
        else axis_tile<EMAX, T, OP_R2C, false>(a, in, out, tile, tid, lds);
        break;
    case OP_C2R:
        if (a.contig) axis_tile<EMAX, T, OP_C2R, true>(a, in, out, tile, tid, lds);
        else axis_tile<EMAX, T, OP_C2R, false>(a, in, out, tile, tid, lds);
        break;
    default:
        break;
    }
}

// Compile-time C2C schedules inside grouped launches: for each length the
// catalog schedule single-pass plans choose by default -- contiguous powers
// of two from 64 take 8 values per thread, every other length and layout 16
// values per thread (radices up to 16) with the split exchange (real and
// imaginary parts in two rounds through half the LDS), which keeps each
// tile's LDS at 37 KB or less: a launch reserves its largest item's LDS for
// every workgroup.  The lengths are the 7-smooth multiples of 16 up to 256.
template <class T, int L, bool CONTIG>
struct GFix {
    static constexpr bool P8 = CONTIG && (L & (L - 1)) == 0 && L >= 64;
    static constexpr int EMAX = P8 ? 8 : 16;
    static constexpr bool SPLIT = !P8;
    using RL = typename std::conditional<P8, RadicesOf8<L>, RadicesOf16<L>>::type;
    static constexpr int TPL = fixed_tpl<RL, EMAX>();
};

template <class T, int L, bool CONTIG>
__device__ __forceinline__ void gfix_run(const AxisPass &a, const void *in, void *out, int64_t tile, int tid,
                                         void *smem) {
    using G = GFix<T, L, CONTIG>;
    axis_tile<G::EMAX, T, OP_C2C, CONTIG, G::SPLIT, typename G::RL, G::TPL>(a, in, out, tile, tid,
                                                                             reinterpret_cast<cx<T> *>(smem));
}

// X(index, L); entry id = 4 * index + 2 * prec + contig
#define DFFT_GFIX_LENGTHS(X)                                                                          \
    X(0, 16) X(1, 32) X(2, 48) X(3, 64) X(4, 80) X(5, 96) X(6, 112) X(7, 128) X(8, 144) X(9, 160)     \
        X(10, 192) X(11, 224) X(12, 240) X(13, 256)

__device__ __forceinline__ void gfix_tile(int id, const AxisPass &a, const void *in, void *out, int64_t tile,
                                          int tid, void *smem) {
    switch (id) {
#define DFFT_GFIX_CASE(i, L)                                            \
    case 4 * i + 0: gfix_run<float, L, false>(a, in, out, tile, tid, smem); break;  \
    case 4 * i + 1: gfix_run<float, L, true>(a, in, out, tile, tid, smem); break;   \
    case 4 * i + 2: gfix_run<double, L, false>(a, in, out, tile, tid, smem); break; \
    case 4 * i + 3: gfix_run<double, L, true>(a, in, out, tile, tid, smem); break;
        DFFT_GFIX_LENGTHS(DFFT_GFIX_CASE)
#undef DFFT_GFIX_CASE
    default: break;
    }
}

// The same dispatch restricted to one precision (PM 1: ComplexF32, 2:
// ComplexF64): a kernel holding one precision's schedules only has that
// precision's register count (ComplexF32 89 VGPRs, 20 waves per CU, against
// 156 and 12 for both).
template <int PM>
__device__ __forceinline__ void gfix_tile_prec(int id, const AxisPass &a, const void *in, void *out, int64_t tile,
                                               int tid, void *smem) {
    switch (id) {
#define DFFT_GFIX_CASE_P(i, L)                                                                    \
    case 4 * i + 0:                                                                               \
        if constexpr (PM == 1) gfix_run<float, L, false>(a, in, out, tile, tid, smem);            \
        break;                                                                                    \
    case 4 * i + 1:                                                                               \
        if constexpr (PM == 1) gfix_run<float, L, true>(a, in, out, tile, tid, smem);             \
        break;                                                                                    \
    case 4 * i + 2:                                                                               \
        if constexpr (PM == 2) gfix_run<double, L, false>(a, in, out, tile, tid, smem);           \
        break;                                                                                    \
    case 4 * i + 3:                                                                               \
        if constexpr (PM == 2) gfix_run<double, L, true>(a, in, out, tile, tid, smem);            \
        break;
        DFFT_GFIX_LENGTHS(DFFT_GFIX_CASE_P)
#undef DFFT_GFIX_CASE_P
    default: break;
    }
}

#define DFFT_GFIX_E(T, P, L, C) \
    {L, P, C ? 1 : 0, GFix<T, L, C>::EMAX, GFix<T, L, C>::SPLIT ? 1 : 0, GFix<T, L, C>::TPL, GFix<T, L, C>::RL::S, \
     GFix<T, L, C>::RL::R}
#define DFFT_GFIX_ENTRIES(i, L) \
    DFFT_GFIX_E(float, 0, L, false), DFFT_GFIX_E(float, 0, L, true), DFFT_GFIX_E(double, 1, L, false), \
        DFFT_GFIX_E(double, 1, L, true),
static const GFixEntry g_gfix[] = {DFFT_GFIX_LENGTHS(DFFT_GFIX_ENTRIES)};
#undef DFFT_GFIX_ENTRIES
#undef DFFT_GFIX_E

int gfix_lookup(int prec, int contig, int L) {
    for (int i = 0; i < (int)(sizeof(g_gfix) / sizeof(g_gfix[0])); ++i)
        if (g_gfix[i].L == L && g_gfix[i].prec == (prec ? 1 : 0) && g_gfix[i].contig == (contig ? 1 : 0)) return i;
    return -1;
}

const GFixEntry &gfix_entry(int id) { return g_gfix[id]; }

// A pointer the whole wavefront shares, held in scalar registers (the item
// table is read through vector loads, so the compiler cannot tell)
template <class P>
__device__ __forceinline__ P *uniform_ptr(P *p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return reinterpret_cast<P *>(lo | (hi << 32));
}

// One workgroup's tile of a grouped launch: block `blk` of the level whose
// items and block -> item map are given; the slots come from the kernel
// arguments.  FIX: the launch holds compile-time schedules.
template <bool FIX, class Args>
__device__ __forceinline__ void group_block(const Args &g, const GroupItem *items, const int *block_item, int64_t blk,
                                            int tid, void *smem) {
    const int item = __builtin_amdgcn_readfirstlane(block_item[blk]);
    const GroupItem &it = items[item];
    const void *in = uniform_ptr(reinterpret_cast<const unsigned char *>(g.slots[it.in_slot]) + it.in_off);
    void *out = uniform_ptr(reinterpret_cast<unsigned char *>(g.slots[it.out_slot]) + it.out_off);
    const int64_t tile = blk - it.first_block;
    if constexpr (FIX) {
        const int gfix = __builtin_amdgcn_readfirstlane(it.gfix);
        if (gfix >= 0) {
            // a private copy of the pass (held in registers: the compile-time
            // schedules index it with constants only), so its fields are not
            // reloaded after the tile's global stores, which the compiler
            // cannot tell apart from the item table
            const AxisPass a = it.ap;
            gfix_tile(gfix, a, in, out, tile, tid, smem);
            return;
        }
    }
    if (it.prec) group_tile<double>(it.ap, in, out, tile, tid, smem);
    else group_tile<float>(it.ap, in, out, tile, tid, smem);
}

// FIX: the launch holds compile-time schedules (GFIX_THREADS threads only);
// its register budget is 12 waves per CU (168 VGPRs): the 16-value schedules
// spill at 128.  Launches of runtime-engine items only take the kernel
// without them, whose register count follows from the runtime engine alone.
template <int THREADS, bool FIX>
__global__ __launch_bounds__(THREADS, FIX ? 3 : 1) void k_group(GroupArgs g) {
    extern __shared__ __align__(16) unsigned char dfft_smem[];
    group_block<FIX>(g, g.items, g.block_item, (int64_t)blockIdx.x, (int)threadIdx.x, dfft_smem);
}

// A launch whose items all run compile-time schedules of one precision
// (PM 1: ComplexF32, 89 VGPRs, five workgroups per CU; 2: ComplexF64, 156
// VGPRs, three per CU).  (ComplexF32 tiles of 512 threads, for row segments
// of 32 columns, two workgroups per CU: the 64-transform mix 1.93 -> 2.06 ms.)
template <int PM>
__global__ __launch_bounds__(GFIX_THREADS, PM == 1 ? 5 : 3) void k_group_prec(GroupArgs g) {
    extern __shared__ __align__(16) unsigned char dfft_smem[];
    const int item = __builtin_amdgcn_readfirstlane(g.block_item[blockIdx.x]);
    const GroupItem &it = g.items[item];
    const void *in = uniform_ptr(reinterpret_cast<const unsigned char *>(g.slots[it.in_slot]) + it.in_off);
    void *out = uniform_ptr(reinterpret_cast<unsigned char *>(g.slots[it.out_slot]) + it.out_off);
    const int64_t tile = (int64_t)blockIdx.x - it.first_block;
    const AxisPass a = it.ap;  // a private copy, as in k_group
    gfix_tile_prec<PM>(__builtin_amdgcn_readfirstlane(it.gfix), a, in, out, tile, (int)threadIdx.x, dfft_smem);
}

// The whole group as one cooperative launch: every workgroup of the grid is
// resident at once (hipLaunchCooperativeKernel refuses a grid larger than the
// device holds), walks the tiles of each of the group's launches with a
// stride of the grid size, and meets the others at a grid barrier wherever
// the next launch belongs to the next level, before it reads what this level
// wrote (a level split by precision into two launches needs none between
// them).  Every workgroup runs every launch's loop, so each reaches each
// barrier; no workgroup waits on a particular other one.  (Tiles dealt from
// per-level ticket counters instead measured within 2%, and counters would
// make two launches of one group on different streams race.)
constexpr int PERSIST_MAX_LAUNCHES = 8;
struct PersistArgs {
    const GroupItem *items;  // every item, launch after launch
    const int *block_item;   // every launch's block -> item map, concatenated
    int nlaunch;
    int sync_after[PERSIST_MAX_LAUNCHES];  // 1: the next launch is the next level
    int first_item[PERSIST_MAX_LAUNCHES];
    int64_t map_off[PERSIST_MAX_LAUNCHES];
    int64_t nblocks[PERSIST_MAX_LAUNCHES];
    void *slots[GROUP_MAX_SLOTS];
};

__global__ __launch_bounds__(GFIX_THREADS, 3) void k_group_persist(PersistArgs g) {
    extern __shared__ __align__(16) unsigned char dfft_smem[];
    for (int l = 0; l < g.nlaunch; ++l) {
        const GroupItem *items = g.items + g.first_item[l];
        const int *map = g.block_item + g.map_off[l];
        const int nb = (int)g.nblocks[l];  // <= 0x7fffffff - grid (group_persistent_prepare)
        for (int b = (int)blockIdx.x; b < nb; b += (int)gridDim.x) {
            // the thread index passes an empty asm statement every tile, so
            // the compiler cannot hoist the tiles' per-thread index
            // arithmetic out of the loop (held across it, the 56 schedules'
            // values spill)
            int tid = (int)threadIdx.x;
            asm volatile("" : "+v"(tid));
            group_block<true>(g, items, map, b, tid, dfft_smem);
            __syncthreads();  // the next tile reuses the LDS
        }
        // the grid barrier orders memory too: after a workgroup barrier its
        // arriving workitem writes the L2 back, and invalidates it once all
        // workgroups have arrived
        if (g.sync_after[l]) cooperative_groups::this_grid().sync();
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct GroupLaunch {
    int first_item = 0, nitems = 0;
    int64_t nblocks = 0;
    int threads = 256;
    bool fixed = false;  // holds compile-time schedules (k_group<GFIX_THREADS, true>)
    int prec = 0;        // 1 / 2: compile-time schedules of ComplexF32 / ComplexF64 only (k_group_prec)
    int level = 0;       // the level (dependency step) the launch belongs to
    size_t lds = 0;
    int64_t map_off = 0;  // offset of this launch's block -> item map
};

struct GroupExec {
    std::vector<GroupItem> items;
    std::vector<GroupLaunch> launches;
    GroupItem *d_items = nullptr;
    int *d_map = nullptr;
    int persist_grid = 0;  // group_persistent_prepare: workgroups of the single launch
    size_t persist_lds = 0;
};

void group_destroy(GroupExec *g) {
    if (!g) return;
    if (g->d_items) (void)hipFree(g->d_items);
    if (g->d_map) (void)hipFree(g->d_map);
    delete g;
}

static const void *group_kernel(int threads, bool fixed, int prec) {
    if (prec == 1) return (const void *)k_group_prec<1>;
    if (prec == 2) return (const void *)k_group_prec<2>;
    if (fixed) return (const void *)k_group<GFIX_THREADS, true>;
    switch (threads) {
    case 1024: return (const void *)k_group<1024, false>;
    case 512: return (const void *)k_group<512, false>;
    default: return (const void *)k_group<256, false>;
    }
}

int group_create(GroupExec **out, const std::vector<std::vector<GroupItem>> &levels) {
    *out = nullptr;
    GroupExec *g = new (std::nothrow) GroupExec();
    if (!g) return DFFT_ERR_OOM;
    std::vector<int> map;
    const int lim = axis_lds_limit();
    // one launch of the items `lv` (already checked) at level `level`;
    // prec 1 / 2: all of them run compile-time ComplexF32 / ComplexF64 schedules
    auto add_launch = [&](const std::vector<GroupItem> &lv, int level, int prec) -> int {
        GroupLaunch L;
        L.first_item = (int)g->items.size();
        L.nitems = (int)lv.size();
        L.map_off = (int64_t)map.size();
        L.level = level;
        L.prec = prec;
        int need = 1;
        bool fixed = false;
        for (size_t i = 0; i < lv.size(); ++i) {
            GroupItem it = lv[i];
            fixed = fixed || it.gfix >= 0;
            it.first_block = L.nblocks;
            for (int64_t b = 0; b < it.ntiles; ++b) map.push_back((int)i);
            L.nblocks += it.ntiles;
            need = std::max(need, it.threads);
            L.lds = std::max(L.lds, it.lds);
            g->items.push_back(it);
        }
        L.threads = need <= 256 ? 256 : need <= 512 ? 512 : 1024;
        // the compile-time schedules are compiled into the 256-thread kernels only
        L.fixed = fixed;
        if (L.lds > (size_t)lim || L.nblocks > 0x7fffffffLL || (fixed && L.threads != GFIX_THREADS))
            return DFFT_ERR_UNSUPPORTED_SIZE;
        g->launches.push_back(L);
        return DFFT_OK;
    };
    int level = 0;
    for (const auto &lv : levels) {
        if (lv.empty()) continue;
        bool all_fixed = true;
        for (const GroupItem &it : lv) {
            if (it.in_slot < 0 || it.in_slot >= GROUP_MAX_SLOTS || it.out_slot < 0 || it.out_slot >= GROUP_MAX_SLOTS ||
                it.ntiles < 1 || it.threads < 1 || it.threads > 1024 ||
                it.gfix >= (int)(sizeof(g_gfix) / sizeof(g_gfix[0]))) {
                group_destroy(g);
                return DFFT_ERR_ARG;
            }
            all_fixed = all_fixed && it.gfix >= 0 && it.threads <= GFIX_THREADS;
        }
        int rc = DFFT_OK;
        if (all_fixed) {
            // a level of compile-time schedules only: one launch per precision,
            // ComplexF64 first, each in the kernel holding that precision's
            // schedules alone (fewer registers, more waves per CU)
            for (int prec = 1; prec >= 0 && rc == DFFT_OK; --prec) {
                std::vector<GroupItem> part;
                for (const GroupItem &it : lv)
                    if (it.prec == prec) part.push_back(it);
                if (!part.empty()) rc = add_launch(part, level, prec ? 2 : 1);
            }
        } else {
            rc = add_launch(lv, level, 0);
        }
        if (rc != DFFT_OK) {
            group_destroy(g);
            return rc;
        }
        ++level;
    }
    if (g->items.empty()) {
        group_destroy(g);
        return DFFT_ERR_ARG;
    }
    if (hipMalloc(&g->d_items, sizeof(GroupItem) * g->items.size()) != hipSuccess ||
        hipMalloc(&g->d_map, sizeof(int) * map.size()) != hipSuccess) {
        group_destroy(g);
        return DFFT_ERR_OOM;
    }
    if (hipMemcpy(g->d_items, g->items.data(), sizeof(GroupItem) * g->items.size(), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(g->d_map, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice) != hipSuccess) {
        group_destroy(g);
        return DFFT_ERR_HIP;
    }
    for (const auto &L : g->launches)
        if (L.lds > 65536 && hipFuncSetAttribute(group_kernel(L.threads, L.fixed, L.prec), hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 lim) != hipSuccess) {
            group_destroy(g);
            return DFFT_ERR_HIP;
        }
    *out = g;
    return DFFT_OK;
}

int group_launch(const GroupExec *g, void *const *slots, int nslots, hipStream_t st) {
    if (nslots > GROUP_MAX_SLOTS) return DFFT_ERR_ARG;
    GroupArgs a{};
    for (int i = 0; i < nslots; ++i) a.slots[i] = slots[i];
    for (const auto &L : g->launches) {
        a.items = g->d_items + L.first_item;
        a.block_item = g->d_map + L.map_off;
        if (L.prec == 1)
            hipLaunchKernelGGL(k_group_prec<1>, dim3((unsigned)L.nblocks), dim3(GFIX_THREADS), L.lds, st, a);
        else if (L.prec == 2)
            hipLaunchKernelGGL(k_group_prec<2>, dim3((unsigned)L.nblocks), dim3(GFIX_THREADS), L.lds, st, a);
        else if (L.fixed)
            hipLaunchKernelGGL((k_group<GFIX_THREADS, true>), dim3((unsigned)L.nblocks), dim3(GFIX_THREADS), L.lds, st,
                               a);
        else if (L.threads == 1024)
            hipLaunchKernelGGL((k_group<1024, false>), dim3((unsigned)L.nblocks), dim3(1024), L.lds, st, a);
        else if (L.threads == 512)
            hipLaunchKernelGGL((k_group<512, false>), dim3((unsigned)L.nblocks), dim3(512), L.lds, st, a);
        else
            hipLaunchKernelGGL((k_group<256, false>), dim3((unsigned)L.nblocks), dim3(256), L.lds, st, a);
        if (hipError_t e = hipGetLastError()) return hip_fail(e, "grouped launch");
    }
    return DFFT_OK;
}

int group_persistent_prepare(GroupExec *g) {
    if (g->launches.empty() || g->launches.size() > (size_t)PERSIST_MAX_LAUNCHES) return DFFT_ERR_UNSUPPORTED;
    size_t lds = 0;
    int64_t most = 0;
    for (const auto &L : g->launches) {
        // one kernel for every level: the 256-thread tiles only
        if (L.threads != GFIX_THREADS) return DFFT_ERR_UNSUPPORTED_SIZE;
        lds = std::max(lds, L.lds);
        most = std::max(most, L.nblocks);
    }
    int dev = 0, coop = 0, cus = 0, per = 0;
    if (hipError_t e = hipGetDevice(&dev)) return hip_fail(e, "hipGetDevice");
    if (hipError_t e = hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev))
        return hip_fail(e, "hipDeviceGetAttribute");
    if (!coop) return DFFT_ERR_UNSUPPORTED;
    if (hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev))
        return hip_fail(e, "hipDeviceGetAttribute");
    if (lds > 65536) {
        if (hipError_t e = hipFuncSetAttribute((const void *)k_group_persist,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, axis_lds_limit()))
            return hip_fail(e, "hipFuncSetAttribute");
    }
    if (hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_group_persist, GFIX_THREADS, lds))
        return hip_fail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    if (per < 1 || cus < 1) return DFFT_ERR_UNSUPPORTED_SIZE;
    g->persist_grid = (int)std::min<int64_t>((int64_t)per * cus, most);
    // the kernel's tile counter (int) steps by the grid size past the level's count
    if (most > (int64_t)0x7fffffff - g->persist_grid) return DFFT_ERR_UNSUPPORTED_SIZE;
    if (getenv("DFFT_TUNE_VERBOSE"))
        fprintf(stderr, "persistent group: %d workgroups (%d per CU x %d CUs), %zu B LDS, levels of", g->persist_grid,
                per, cus, lds);
    if (getenv("DFFT_TUNE_VERBOSE")) {
        for (const auto &L : g->launches) fprintf(stderr, " %lld", (long long)L.nblocks);
        fprintf(stderr, " tiles\n");
    }
    g->persist_lds = lds;
    return DFFT_OK;
}

int group_persistent_grid(const GroupExec *g) { return g->persist_grid; }

int group_launch_persistent(const GroupExec *g, void *const *slots, int nslots, hipStream_t st) {
    if (nslots > GROUP_MAX_SLOTS || g->persist_grid < 1) return DFFT_ERR_ARG;
    PersistArgs a{};
    a.items = g->d_items;
    a.block_item = g->d_map;
    a.nlaunch = (int)g->launches.size();
    for (int l = 0; l < a.nlaunch; ++l) {
        a.sync_after[l] = l + 1 < a.nlaunch && g->launches[l + 1].level != g->launches[l].level;
        a.first_item[l] = g->launches[l].first_item;
        a.map_off[l] = g->launches[l].map_off;
        a.nblocks[l] = g->launches[l].nblocks;
    }
    for (int i = 0; i < nslots; ++i) a.slots[i] = slots[i];
    void *kargs[] = {&a};
    if (hipError_t e = hipLaunchCooperativeKernel((const void *)k_group_persist, dim3((unsigned)g->persist_grid),
                                                  dim3(GFIX_THREADS), kargs, (unsigned)g->persist_lds, st))
        return hip_fail(e, "cooperative launch");
    return DFFT_OK;
}

// The launches of a group recorded once as a hipGraph (kernel nodes in a
// chain) and replayed; re-recorded when the slot pointers change.
struct GroupGraph {
    std::mutex mu;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<void *> slots;
};

static void graph_reset(GroupGraph *gg) {
    if (gg->exec) (void)hipGraphExecDestroy(gg->exec);
    if (gg->graph) (void)hipGraphDestroy(gg->graph);
    gg->exec = nullptr;
    gg->graph = nullptr;
    gg->slots.clear();
}

GroupGraph *group_graph_new() { return new (std::nothrow) GroupGraph(); }

void group_graph_destroy(GroupGraph *gg) {
    if (!gg) return;
    graph_reset(gg);
    delete gg;
}

int group_launch_graph(const GroupExec *g, GroupGraph *gg, void *const *slots, int nslots, hipStream_t st) {
    if (nslots > GROUP_MAX_SLOTS) return DFFT_ERR_ARG;
    std::lock_guard<std::mutex> lk(gg->mu);
    std::vector<void *> cur(slots, slots + nslots);
    if (!gg->exec || gg->slots != cur) {
        graph_reset(gg);
        hipGraph_t graph;
        if (hipError_t e = hipGraphCreate(&graph, 0)) return hip_fail(e, "hipGraphCreate");
        hipGraphNode_t prev = nullptr;
        for (const auto &L : g->launches) {
            GroupArgs a{};
            for (int i = 0; i < nslots; ++i) a.slots[i] = slots[i];
            a.items = g->d_items + L.first_item;
            a.block_item = g->d_map + L.map_off;
            void *kargs[] = {&a};
            hipKernelNodeParams kp{};
            kp.func = const_cast<void *>(group_kernel(L.threads, L.fixed, L.prec));
            kp.gridDim = dim3((unsigned)L.nblocks);
            kp.blockDim = dim3(L.threads);
            kp.sharedMemBytes = (unsigned)L.lds;
            kp.kernelParams = kargs;
            hipGraphNode_t node;
            if (hipError_t e = hipGraphAddKernelNode(&node, graph, prev ? &prev : nullptr, prev ? 1 : 0, &kp)) {
                (void)hipGraphDestroy(graph);
                return hip_fail(e, "hipGraphAddKernelNode");
            }
            prev = node;
        }
        hipGraphExec_t exec;
        if (hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0)) {
            (void)hipGraphDestroy(graph);
            return hip_fail(e, "hipGraphInstantiate");
        }
        gg->graph = graph;
        gg->exec = exec;
        gg->slots = cur;
    }
    if (hipError_t e = hipGraphLaunch(gg->exec, st)) return hip_fail(e, "hipGraphLaunch");
    return DFFT_OK;
}

int group_num_launches(const GroupExec *g) { return (int)g->launches.size(); }

int64_t group_blocks(const GroupExec *g) {
    int64_t s = 0;
    for (const auto &L : g->launches) s += L.nblocks;
    return s;
}

GroupItem make_group_item(const PassPlan &pp, int in_slot, int64_t in_off, int out_slot, int64_t out_off) {
    GroupItem it{};
    it.ap = pp.ap;
    it.ap.in = nullptr;
    it.ap.out = nullptr;
    it.prec = pp.prec;
    it.in_slot = in_slot;
    it.out_slot = out_slot;
    it.in_off = in_off;
    it.out_off = out_off;
    it.ntiles = pp.nblocks;
    it.threads = pp.threads;
    it.lds = pp.lds;
    it.gfix = -1;
    return it;
}

}  // namespace dfft

using namespace dfft;

// ---------------------------------------------------------------------------
// batch API: launch l runs pass l of every transform
// ---------------------------------------------------------------------------
struct dfft_batch {
    int count = 0;
    std::vector<dfft_batch_desc> descs;
    GroupExec *g = nullptr;
    GroupGraph *graph = nullptr;  // DFFT_FLAG_GRAPH: the launches replayed as one hipGraph
    bool persistent = false;      // DFFT_FLAG_BATCH_PERSISTENT: the launches as one cooperative launch
    void *d_tw64 = nullptr, *d_tw32 = nullptr;
    size_t tw64_bytes = 0, tw32_bytes = 0;
    double alg_bytes = 0;
    // lanes (default): every transform runs its own plan (the catalog
    // kernels) on one of a few streams forked from the caller's stream
    bool lanes = false;
    std::vector<dfft_plan *> plans;
    std::vector<int> order;     // transforms, largest first
    std::vector<int> lane_of;   // transform -> stream (balanced by bytes)
    std::vector<hipStream_t> streams;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joins;
    bool use_graph = false;
    std::mutex gmu;             // the lanes recorded as one hipGraph (stream capture)
    hipStream_t origin = nullptr;
    hipGraph_t lgraph = nullptr;
    hipGraphExec_t lexec = nullptr;
    std::vector<void *> lslots;
};

static void batch_free(dfft_batch *b) {
    if (!b) return;
    group_graph_destroy(b->graph);
    group_destroy(b->g);
    if (b->d_tw64) (void)hipFree(b->d_tw64);
    if (b->d_tw32) (void)hipFree(b->d_tw32);
    if (b->lexec) (void)hipGraphExecDestroy(b->lexec);
    if (b->lgraph) (void)hipGraphDestroy(b->lgraph);
    for (dfft_plan *p : b->plans) dfft_plan_destroy(p);
    for (hipStream_t st : b->streams) (void)hipStreamDestroy(st);
    if (b->origin) (void)hipStreamDestroy(b->origin);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (hipEvent_t e : b->joins) (void)hipEventDestroy(e);
    delete b;
}

namespace dfft {
int plan_pass_count(const dfft_plan *p);
PassPlan plan_pass(const dfft_plan *p, int i);
}

constexpr int BATCH_LANES = 4;

// Geometry of a pass that runs compile-time schedule e in a grouped launch:
// as many lines (contiguous) or columns (strided) per tile as fill
// GFIX_THREADS threads within 64 KB of LDS.  The same rules as the planner's
// for single-pass launches.
static bool gfix_schedule(PassPlan &pp, const GFixEntry &e) {
    AxisPass &a = pp.ap;
    const int es = pp.prec ? 16 : 8;
    const int lds_es = e.split ? es / 2 : es;
    const int rows = a.sp.L + 1;
    const size_t lim = 65536;
    if (e.S > MAX_STAGES) return false;
    a.variant = 0;  // informational: the grouped kernel dispatches on the entry
    a.sp.S = e.S;
    for (int s = 0; s < e.S; ++s) a.sp.R[s] = e.R[s];
    a.TPL = e.tpl;
    int tc = std::max(1, GFIX_THREADS / e.tpl);
    if (a.contig) {
        a.pad_shift = pp.prec ? 3 : 4;
        a.lds_pitch = rows + (rows >> a.pad_shift) + 1;
        while (tc > 1 && (size_t)tc * a.lds_pitch * lds_es > lim) --tc;
        pp.lds = (size_t)tc * a.lds_pitch * lds_es;
        pp.nblocks = (a.nlines + tc - 1) / tc;
    } else {
        a.pad_shift = 30;
        // TC < 2 * columns, which the strided tile mapping assumes
        while (tc > 1 && tc / 2 >= a.nlines) tc /= 2;
        while (tc > 1 && (size_t)rows * tc * lds_es > lim) tc /= 2;
        a.lds_pitch = tc;
        pp.lds = (size_t)rows * tc * lds_es;
        pp.nblocks = (a.nlines * a.nouter + tc - 1) / tc;
    }
    a.TC = tc;
    a.swz = 1;
    pp.threads = e.tpl * tc;
    return pp.threads <= GFIX_THREADS && pp.lds <= lim && pp.nblocks <= 0x7fffffffLL;
}

// Schedule of one batch pass inside a grouped launch of workgroups of up to
// max_threads threads.  With max_threads == GFIX_THREADS a C2C pass whose
// length the grouped kernel has a compile-time schedule for takes it (gfix =
// its entry), with as many lines or columns per tile as fill the workgroup;
// every other pass takes the runtime engine (gfix = -1).
static int schedule_grouped(PassPlan &pp, uint32_t flags, int max_threads, int &gfix) {
    gfix = -1;
    if (max_threads == GFIX_THREADS && pp.ap.op == OP_C2C) {
        const int id = gfix_lookup(pp.prec, pp.ap.contig, pp.ap.sp.L);
        if (id >= 0) {
            PassPlan q = pp;
            if (gfix_schedule(q, gfix_entry(id))) {
                pp = q;
                gfix = id;
                return DFFT_OK;
            }
        }
    }
    Tuning tn;
    const Tuning *tune = nullptr;
    if (!(flags & DFFT_FLAG_NO_TUNE) && tuning_lookup(pp, tn) && tn.tpl <= max_threads) tune = &tn;
    return schedule_pass(pp, tune, max_threads, 65536, 256, true);
}

// lanes mode: builds one plan per transform and the streams/events
static int lanes_create(dfft_batch *b, uint32_t flags) {
    const int dims[3] = {0, 1, 2};
    std::vector<double> bytes(b->count);
    for (int t = 0; t < b->count; ++t) {
        const dfft_batch_desc &d = b->descs[t];
        dfft_plan *p = nullptr;
        // (DFFT_FLAG_MEASURE: each transform's untuned passes are timed first)
        const uint32_t pf = (d.flags & DFFT_FLAG_UNNORMALIZED) | (flags & (DFFT_FLAG_NO_TUNE | DFFT_FLAG_MEASURE));
        int rc = dfft_plan_create(&p, 3, d.n, d.kind, d.sign, nullptr, d.precision, DFFT_PENCIL, dims, 3, pf);
        if (rc != DFFT_OK) return rc;
        b->plans.push_back(p);
        dfft_plan_algorithmic_bytes(p, &bytes[t]);
        b->alg_bytes += bytes[t];
    }
    b->order.resize(b->count);
    std::iota(b->order.begin(), b->order.end(), 0);
    std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) { return bytes[x] > bytes[y]; });
    const int nl = std::min(BATCH_LANES, b->count);
    std::vector<double> load(nl, 0.0);
    b->lane_of.assign(b->count, 0);
    for (int t : b->order) {
        const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        b->lane_of[t] = l;
        load[l] += bytes[t];
    }
    for (int l = 0; l < nl; ++l) {
        hipStream_t st;
        hipEvent_t ev;
        if (hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) return hip_fail(e, "hipStreamCreateWithFlags");
        b->streams.push_back(st);
        if (hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) return hip_fail(e, "hipEventCreate");
        b->joins.push_back(ev);
    }
    if (hipError_t e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming)) return hip_fail(e, "hipEventCreate");
    if (b->use_graph) {
        if (hipError_t e = hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking))
            return hip_fail(e, "hipStreamCreateWithFlags");
    }
    return DFFT_OK;
}

// lanes mode: fork from `st`, each transform's passes on its stream, join back into `st`
static int lanes_enqueue(dfft_batch *b, void *const *d_in, void *const *d_out, hipStream_t st) {
    if (hipError_t e = hipEventRecord(b->fork, st)) return hip_fail(e, "hipEventRecord");
    for (hipStream_t s : b->streams)
        if (hipError_t e = hipStreamWaitEvent(s, b->fork, 0)) return hip_fail(e, "hipStreamWaitEvent");
    for (int t : b->order) {
        int rc = dfft_execute(b->plans[t], d_in[t], d_out[t], nullptr, b->streams[b->lane_of[t]]);
        if (rc != DFFT_OK) return rc;
    }
    for (size_t l = 0; l < b->streams.size(); ++l) {
        if (hipError_t e = hipEventRecord(b->joins[l], b->streams[l])) return hip_fail(e, "hipEventRecord");
        if (hipError_t e = hipStreamWaitEvent(st, b->joins[l], 0)) return hip_fail(e, "hipStreamWaitEvent");
    }
    return DFFT_OK;
}

static int lanes_execute(dfft_batch *b, void *const *d_in, void *const *d_out, hipStream_t st) {
    if (!b->use_graph) return lanes_enqueue(b, d_in, d_out, st);
    std::lock_guard<std::mutex> lk(b->gmu);
    std::vector<void *> cur;
    for (int t = 0; t < b->count; ++t) {
        cur.push_back(d_in[t]);
        cur.push_back(d_out[t]);
    }
    if (!b->lexec || cur != b->lslots) {
        if (b->lexec) (void)hipGraphExecDestroy(b->lexec);
        if (b->lgraph) (void)hipGraphDestroy(b->lgraph);
        b->lexec = nullptr;
        b->lgraph = nullptr;
        // record the fork/join pattern from a private stream: the lanes join the
        // capture through the fork event, and the graph keeps their parallel branches
        if (hipError_t e = hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal))
            return hip_fail(e, "hipStreamBeginCapture");
        int rc = lanes_enqueue(b, d_in, d_out, b->origin);
        hipGraph_t graph = nullptr;
        hipError_t ec = hipStreamEndCapture(b->origin, &graph);
        if (rc != DFFT_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (ec != hipSuccess) return hip_fail(ec, "hipStreamEndCapture");
        hipGraphExec_t exec;
        if (hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0)) {
            (void)hipGraphDestroy(graph);
            return hip_fail(e, "hipGraphInstantiate");
        }
        b->lgraph = graph;
        b->lexec = exec;
        b->lslots = cur;
    }
    if (hipError_t e = hipGraphLaunch(b->lexec, st)) return hip_fail(e, "hipGraphLaunch");
    return DFFT_OK;
}

// The schedule cache key of a batch's launch schedule: its transforms'
// descriptors (sizes, kind, sign, precision, flags), FNV-1a hashed.
static std::string batch_key(int count, const dfft_batch_desc *descs) {
    uint64_t h = 1469598103934665603ULL;
    auto mix = [&](int64_t v) {
        for (int k = 0; k < 8; ++k) {
            h ^= (uint64_t)((v >> (8 * k)) & 0xff);
            h *= 1099511628211ULL;
        }
    };
    for (int t = 0; t < count; ++t) {
        const dfft_batch_desc &d = descs[t];
        for (int a = 0; a < 3; ++a) mix(d.n[a]);
        mix(d.kind);
        mix(d.sign);
        mix(d.precision);
        mix(d.flags);
    }
    char buf[64];
    snprintf(buf, sizeof buf, "batch/%d/%016llx", count, (unsigned long long)h);
    return buf;
}

// Launch-schedule values of a batch's cache entry (Tuning::variant)
constexpr int BATCH_PICK_GROUPED = 1, BATCH_PICK_LANES = 2;

extern "C" int dfft_batch_create(dfft_batch **batch, int count, const dfft_batch_desc *descs, uint32_t flags);

// DFFT_FLAG_MEASURE: times the grouped schedule (when every pass fits it) and
// the lanes schedule (each transform's own plan, its passes autotuned) on
// scratch arrays of the batch's sizes, best of three, and returns the faster
// (BATCH_PICK_*), or a negative DFFT_ERR_* code.  The single cooperative
// launch is not a candidate: after a process's first cooperative launch,
// work spread over several streams runs slower for the rest of the process
// (the lanes schedule of the 64-transform mix 2.17 -> 4.03 ms), so timing
// it would slow whatever the caller runs next.
static int measure_batch(int count, const dfft_batch_desc *descs, uint32_t flags) {
    const uint32_t base = flags & ~(uint32_t)(DFFT_FLAG_GRAPH | DFFT_FLAG_BATCH_GROUPED | DFFT_FLAG_BATCH_LANES |
                                              DFFT_FLAG_BATCH_PERSISTENT);
    std::vector<void *> in(count, nullptr), out(count, nullptr);
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = DFFT_OK;
    for (int t = 0; t < count && rc == DFFT_OK; ++t) {
        const dfft_batch_desc &d = descs[t];
        const int64_t N = d.n[0] * d.n[1] * d.n[2];
        const int64_t NC = (d.kind == DFFT_R2C ? d.n[0] / 2 + 1 : d.n[0]) * d.n[1] * d.n[2];
        const size_t es = d.precision ? 8 : 4;
        const size_t ib = (size_t)N * es * (d.kind == DFFT_R2C ? 1 : 2), ob = (size_t)NC * es * 2;
        if (hipMalloc(&in[t], ib) != hipSuccess || hipMalloc(&out[t], ob) != hipSuccess) rc = DFFT_ERR_OOM;
        else if (hipError_t e = launch_fill(in[t], (int64_t)(ib / es), d.precision, 17 + t, nullptr))
            rc = hip_fail(e, "input fill");
    }
    if (rc == DFFT_OK) {
        if (hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) rc = hip_fail(e, "hipStreamCreate");
        else if (hipError_t e2 = hipEventCreate(&e0)) rc = hip_fail(e2, "hipEventCreate");
        else if (hipError_t e3 = hipEventCreate(&e1)) rc = hip_fail(e3, "hipEventCreate");
    }
    if (rc == DFFT_OK && hipDeviceSynchronize() != hipSuccess) rc = hip_fail(hipGetLastError(), "fill");
    float best[2] = {-1.0f, -1.0f};
    for (int k = 0; k < 2 && rc == DFFT_OK; ++k) {
        dfft_batch *b = nullptr;
        const uint32_t sf = base | (k == 0 ? DFFT_FLAG_BATCH_GROUPED : DFFT_FLAG_BATCH_LANES);
        const int crc = dfft_batch_create(&b, count, descs, sf);
        if (crc != DFFT_OK) {
            if (k == 1) rc = crc;  // lanes always builds; a grouped batch may not
            continue;
        }
        bool ok = dfft_batch_execute(b, in.data(), out.data(), st) == DFFT_OK;  // warm-up
        float m = 1e30f;
        for (int trial = 0; trial < 3 && ok; ++trial) {
            ok = hipEventRecord(e0, st) == hipSuccess;
            for (int r = 0; r < 2 && ok; ++r) ok = dfft_batch_execute(b, in.data(), out.data(), st) == DFFT_OK;
            ok = ok && hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
            float ms = 0;
            if (ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) m = std::min(m, ms / 2);
        }
        if (ok) best[k] = m;
        else rc = hip_fail(hipGetLastError(), "batch timing");
        dfft_batch_destroy(b);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    for (int t = 0; t < count; ++t) {
        if (in[t]) (void)hipFree(in[t]);
        if (out[t]) (void)hipFree(out[t]);
    }
    if (rc != DFFT_OK) return -rc;
    if (getenv("DFFT_TUNE_VERBOSE"))
        fprintf(stderr, "tune batch of %d: grouped %.4f ms, lanes %.4f ms\n", count, best[0], best[1]);
    return (best[0] >= 0 && best[0] <= best[1]) ? BATCH_PICK_GROUPED : BATCH_PICK_LANES;
}

extern "C" {

int dfft_batch_create(dfft_batch **batch, int count, const dfft_batch_desc *descs, uint32_t flags) {
    if (!batch || count < 1 || !descs) return DFFT_ERR_ARG;
    *batch = nullptr;
    if (2 * count > GROUP_MAX_SLOTS) return DFFT_ERR_ARG;
    for (int t = 0; t < count; ++t) {
        const dfft_batch_desc &d = descs[t];
        if (d.kind != DFFT_C2C && d.kind != DFFT_R2C) return DFFT_ERR_UNSUPPORTED;
        if (d.sign != -1 && d.sign != 1) return DFFT_ERR_ARG;
        if (d.kind == DFFT_R2C && (d.sign != -1 || d.n[0] % 2 != 0)) return DFFT_ERR_ARG;
        if (d.precision != DFFT_F32 && d.precision != DFFT_F64) return DFFT_ERR_ARG;
        for (int a = 0; a < 3; ++a)
            if (d.n[a] < 1 || d.n[a] > (1 << 20)) return DFFT_ERR_ARG;
    }
    if (!(flags & DFFT_FLAG_NO_TUNE)) wisdom_autoload();
    dfft_batch *b = new (std::nothrow) dfft_batch();
    if (!b) return DFFT_ERR_OOM;
    b->count = count;
    b->descs.assign(descs, descs + count);
    if ((flags & DFFT_FLAG_BATCH_GROUPED) && (flags & DFFT_FLAG_BATCH_LANES)) {
        batch_free(b);
        return DFFT_ERR_ARG;
    }
    // the persistent schedule is the grouped one in a single launch: nothing
    // left for a graph to chain, and no lanes
    if ((flags & DFFT_FLAG_BATCH_PERSISTENT) && (flags & (DFFT_FLAG_BATCH_LANES | DFFT_FLAG_GRAPH))) {
        batch_free(b);
        return DFFT_ERR_ARG;
    }
    b->persistent = (flags & DFFT_FLAG_BATCH_PERSISTENT) != 0;
    bool grouped = (flags & (DFFT_FLAG_BATCH_GROUPED | DFFT_FLAG_BATCH_PERSISTENT)) != 0;
    if (!grouped && !(flags & DFFT_FLAG_BATCH_LANES)) {
        // default: the schedule cache's pick for this mix (DFFT_FLAG_MEASURE
        // times both schedules when there is none); without one, grouped
        // when every pass has a compile-time grouped schedule
        const std::string key = batch_key(count, descs);
        Tuning pick;
        bool have = !(flags & DFFT_FLAG_NO_TUNE) && tuning_lookup_key(key, pick);
        if (!have && (flags & DFFT_FLAG_MEASURE) && !(flags & DFFT_FLAG_NO_TUNE)) {
            const int m = measure_batch(count, descs, flags);
            if (m < 0) {
                batch_free(b);
                return -m;
            }
            pick = Tuning{};
            pick.variant = m;
            tuning_store_key(key, pick);
            have = true;
        }
        if (have) {
            grouped = pick.variant == BATCH_PICK_GROUPED;
        } else {
            grouped = true;
            for (int t = 0; t < count && grouped; ++t) {
                grouped = descs[t].kind == DFFT_C2C;
                for (int i = 0; i < 3 && grouped; ++i)
                    grouped = gfix_lookup(descs[t].precision, 1, (int)descs[t].n[i]) >= 0;
            }
        }
    }
    if (!grouped) {
        b->lanes = true;
        b->use_graph = (flags & DFFT_FLAG_GRAPH) != 0;
        int rc = lanes_create(b, flags);
        if (rc != DFFT_OK) {
            batch_free(b);
            return rc;
        }
        *batch = b;
        return DFFT_OK;
    }
    TwiddleTable tt64, tt32;
    tt64.prec = 1;
    tt32.prec = 0;
    std::vector<std::vector<PassPlan>> lv(3);
    std::vector<std::vector<int>> lv_t(3);
    int rc = DFFT_OK;
    for (int t = 0; t < count && rc == DFFT_OK; ++t) {
        const dfft_batch_desc &d = descs[t];
        const int prec = d.precision;
        const int es_c = prec ? 16 : 8;
        const bool inv = d.sign > 0;
        const bool norm = !(d.flags & DFFT_FLAG_UNNORMALIZED);
        const int64_t n[3] = {d.n[0], d.n[1], d.n[2]};
        int64_t nc[3] = {n[0], n[1], n[2]};
        std::vector<int> order = {0, 1, 2};
        if (d.kind == DFFT_R2C) nc[0] = n[0] / 2 + 1;
        else if (inv) std::reverse(order.begin(), order.end());
        const double N = (double)(n[0] * n[1] * n[2]), NC = (double)(nc[0] * nc[1] * nc[2]);
        for (int i = 0; i < 3; ++i) {
            const int ax = order[i];
            PassPlan pp{};
            pp.prec = prec;
            pp.axis = ax;
            pp.src = i == 0 ? BUF_IN : BUF_OUT;
            pp.dst = BUF_OUT;
            AxisPass &a = pp.ap;
            if (d.kind == DFFT_R2C && i == 0) {
                a.op = OP_R2C;
                a.sp.L = (int)(n[0] / 2);
                a.nlog = (int)n[0];
                a.scale = 1.0;
                set_geometry(a, 0, n, nc);
                a.nt = 1;
                pp.alg_bytes = N * es_c / 2 + NC * es_c;
            } else {
                const int64_t *gsz = d.kind == DFFT_R2C ? nc : n;
                const int L = (int)gsz[ax];
                a.op = OP_C2C;
                a.sp.L = L;
                a.nlog = L;
                a.inverse = inv ? 1 : 0;
                a.scale = (inv && norm) ? 1.0 / L : 1.0;
                set_geometry(a, ax, gsz, gsz);
                a.nt = 1;
                a.paired_loads = paired_loads_for(a.in_sm, prec ? 16 : 8);
                pp.alg_bytes = 2.0 * (d.kind == DFFT_R2C ? NC : N) * es_c;
            }
            b->alg_bytes += pp.alg_bytes;
            lv[i].push_back(pp);
            lv_t[i].push_back(t);
        }
    }
    // schedules, level by level: the grouped kernel's compile-time schedules
    // where it has one (256-thread tiles); if a pass of the level does not
    // fit 256 threads, the whole level runs the runtime engine in wider
    // workgroups.  (A level whose passes all take compile-time schedules runs
    // one launch per precision: group_create.)
    std::vector<std::vector<int>> lv_g(3);
    for (int l = 0; l < 3 && rc == DFFT_OK; ++l) {
        std::vector<PassPlan> sched = lv[l];
        std::vector<int> gf(sched.size(), -1);
        bool wide = false;
        for (size_t k = 0; k < sched.size() && !wide; ++k)
            wide = schedule_grouped(sched[k], flags, GFIX_THREADS, gf[k]) != DFFT_OK;
        if (wide) {
            sched = lv[l];
            gf.assign(sched.size(), -1);
            for (size_t k = 0; k < sched.size() && rc == DFFT_OK; ++k)
                rc = schedule_grouped(sched[k], flags, 512, gf[k]);
        }
        if (rc != DFFT_OK) break;
        for (PassPlan &pp : sched) (pp.prec ? tt64 : tt32).add(pp.ap);
        lv[l] = sched;
        lv_g[l] = gf;
    }
    if (rc == DFFT_OK && !tt64.tab.empty()) rc = tt64.upload(&b->d_tw64, &b->tw64_bytes);
    if (rc == DFFT_OK && !tt32.tab.empty()) rc = tt32.upload(&b->d_tw32, &b->tw32_bytes);
    if (rc != DFFT_OK) {
        batch_free(b);
        return rc;
    }
    std::vector<std::vector<GroupItem>> levels(3);
    for (int l = 0; l < 3; ++l) {
        // larger passes first: their tiles start earliest
        std::vector<int> idx(lv[l].size());
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(),
                         [&](int x, int y) { return lv[l][x].alg_bytes > lv[l][y].alg_bytes; });
        for (int k : idx) {
            PassPlan pp = lv[l][k];
            const int t = lv_t[l][k];
            pp.ap.tw = pp.prec ? b->d_tw64 : b->d_tw32;
            const int in_slot = pp.src == BUF_IN ? 2 * t : 2 * t + 1;
            levels[l].push_back(make_group_item(pp, in_slot, 0, 2 * t + 1, 0));
            levels[l].back().gfix = lv_g[l][k];
        }
    }
    rc = group_create(&b->g, levels);
    if (rc == DFFT_OK && b->persistent) rc = group_persistent_prepare(b->g);
    if (rc == DFFT_OK && (flags & DFFT_FLAG_GRAPH) && !(b->graph = group_graph_new())) rc = DFFT_ERR_OOM;
    if (rc != DFFT_OK) {
        batch_free(b);
        return rc;
    }
    *batch = b;
    return DFFT_OK;
}

int dfft_batch_execute(dfft_batch *b, void *const *d_in, void *const *d_out, void *stream) {
    if (!b || !d_in || !d_out) return DFFT_ERR_ARG;
    std::vector<void *> slots(2 * b->count);
    for (int t = 0; t < b->count; ++t) {
        if (!d_in[t] || !d_out[t]) return DFFT_ERR_ARG;
        if (d_in[t] == d_out[t] && b->descs[t].kind != DFFT_C2C) return DFFT_ERR_ARG;
        slots[2 * t] = d_in[t];
        slots[2 * t + 1] = d_out[t];
    }
    if (b->lanes) return lanes_execute(b, d_in, d_out, (hipStream_t)stream);
    if (b->persistent) return group_launch_persistent(b->g, slots.data(), (int)slots.size(), (hipStream_t)stream);
    if (b->graph) return group_launch_graph(b->g, b->graph, slots.data(), (int)slots.size(), (hipStream_t)stream);
    return group_launch(b->g, slots.data(), (int)slots.size(), (hipStream_t)stream);
}

int dfft_batch_num_items(const dfft_batch *b, int64_t *items) {
    if (!b || !items) return DFFT_ERR_ARG;
    if (b->lanes) {
        int64_t n = 0;
        for (const dfft_plan *p : b->plans)
            for (int i = 0; i < plan_pass_count(p); ++i) n += plan_pass(p, i).nblocks;
        *items = n;
    } else {
        *items = group_blocks(b->g);
    }
    return DFFT_OK;
}

int dfft_batch_num_launches(const dfft_batch *b, int *launches) {
    if (!b || !launches) return DFFT_ERR_ARG;
    if (b->lanes) {
        int n = 0;
        for (const dfft_plan *p : b->plans) n += plan_pass_count(p);
        *launches = n;
    } else {
        *launches = b->persistent ? 1 : group_num_launches(b->g);
    }
    return DFFT_OK;
}

int dfft_batch_workspace_size(const dfft_batch *b, size_t *bytes) {
    if (!b || !bytes) return DFFT_ERR_ARG;
    *bytes = 0;
    return DFFT_OK;
}

int dfft_batch_algorithmic_bytes(const dfft_batch *b, double *bytes) {
    if (!b || !bytes) return DFFT_ERR_ARG;
    *bytes = b->alg_bytes;
    return DFFT_OK;
}

int dfft_batch_destroy(dfft_batch *b) {
    if (!b) return DFFT_ERR_ARG;
    batch_free(b);
    return DFFT_OK;
}

}  // extern "C"
