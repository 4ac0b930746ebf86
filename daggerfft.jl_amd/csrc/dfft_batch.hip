// dfft_batch.hip -- many independent 3D transforms in one persistent launch.
//
// Every (transform, axis pass) becomes an item with a run of tickets, one per
// tile.  Items are ordered by pass level (all first passes, then all second
// passes, ...), largest transforms first inside a level.  A resident
// workgroup takes the next ticket from one global counter, waits until the
// item its ticket belongs to may start (the same transform's previous pass
// has finished all its tiles), runs the tile with the same device code as the
// single-transform kernels, and counts the tile done.
//
// Deadlock freedom: an item only waits on items of an earlier pass level,
// whose tickets are all smaller.  A ticket is only ever held by a running
// workgroup, and workgroups never wait on larger tickets, so by induction on
// the ticket order every wait ends.  The wait is still bounded by a spin
// limit: on expiry the workgroup raises an error flag and every later item
// skips its work, so the grid always drains.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <numeric>
#include <vector>

#include "../../include/daggerfft_hip.h"
#include "dfft_axis.hpp"
#include "dfft_kernels.hpp"
#include "dfft_plan.hpp"

namespace dfft {

constexpr int BATCH_THREADS = 512;
constexpr size_t BATCH_LDS = 65536;

struct BatchItem {
    AxisPass ap;
    int prec;
    int dep;            // item that must finish first (-1: none)
    int64_t ntiles;
    int64_t first;      // first ticket
    int64_t dep_need;   // tiles of `dep`
};

struct BatchCtl {
    unsigned long long ticket;
    unsigned int err;
    unsigned int pad;
};

__device__ __forceinline__ unsigned ld_acquire(const unsigned *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T>
__device__ __forceinline__ void batch_tile(const AxisPass &a, int64_t tile, int tid, void *smem) {
    cx<T> *lds = reinterpret_cast<cx<T> *>(smem);
    if (a.op == OP_C2C) {
        if (a.contig) axis_tile<EMAX, T, OP_C2C, true>(a, tile, tid, lds);
        else axis_tile<EMAX, T, OP_C2C, false>(a, tile, tid, lds);
    } else {  // OP_R2C: always the first pass of an R2C transform
        if (a.contig) axis_tile<EMAX, T, OP_R2C, true>(a, tile, tid, lds);
        else axis_tile<EMAX, T, OP_R2C, false>(a, tile, tid, lds);
    }
}

__global__ __launch_bounds__(BATCH_THREADS) void k_batch(const BatchItem *__restrict__ items, int nitems,
                                                         int64_t total, BatchCtl *ctl, unsigned *done,
                                                         unsigned long long spin_limit) {
    extern __shared__ __align__(16) unsigned char dfft_smem[];
    __shared__ int64_t s_ticket;
    __shared__ int s_skip;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) s_ticket = (int64_t)atomicAdd(&ctl->ticket, 1ULL);
        __syncthreads();
        // uniform across the workgroup: keep it in scalar registers
        const int64_t tk = (int64_t)__builtin_amdgcn_readfirstlane((int)(s_ticket >> 32)) << 32 |
                           (uint32_t)__builtin_amdgcn_readfirstlane((int)(s_ticket & 0xffffffff));
        if (tk >= total) break;
        // item holding ticket tk (items sorted by `first`)
        int lo = 0, hi = nitems - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (items[mid].first <= tk) lo = mid;
            else hi = mid - 1;
        }
        lo = __builtin_amdgcn_readfirstlane(lo);
        const BatchItem &it = items[lo];
        if (tid == 0) {
            int skip = __hip_atomic_load(&ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
            if (!skip && it.dep >= 0) {
                unsigned long long spins = 0;
                while (ld_acquire(&done[it.dep]) < (unsigned)it.dep_need) {
                    if (++spins > spin_limit) {
                        atomicOr(&ctl->err, 1u);
                        skip = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            s_skip = skip;
        }
        __syncthreads();
        if (!s_skip) {
            // the producer's writes are visible to every lane after the acquire above
            __atomic_thread_fence(__ATOMIC_ACQUIRE);
            if (it.prec) batch_tile<double>(it.ap, tk - it.first, tid, dfft_smem);
            else batch_tile<float>(it.ap, tk - it.first, tid, dfft_smem);
        }
        __threadfence();
        __syncthreads();
        if (tid == 0) __hip_atomic_fetch_add(&done[lo], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ void k_batch_reset(BatchCtl *ctl, unsigned *done, int nitems) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        ctl->ticket = 0;
        ctl->err = 0;
    }
    if (i < nitems) done[i] = 0;
}

}  // namespace dfft

using namespace dfft;

struct dfft_batch {
    int count = 0;
    std::vector<dfft_batch_desc> descs;
    std::vector<BatchItem> items;      // ticket order
    std::vector<int> item_transform;   // transform of each item
    std::vector<int> item_src, item_dst;
    int64_t total = 0;
    size_t lds = 0;
    int grid = 0;
    void *d_tw64 = nullptr, *d_tw32 = nullptr;
    size_t tw64_bytes = 0, tw32_bytes = 0;
    BatchItem *d_items = nullptr;
    BatchItem *h_items = nullptr;       // pinned staging of the item table
    BatchCtl *d_ctl = nullptr;
    unsigned *d_done = nullptr;
    unsigned long long spin_limit = 1ULL << 32;
    std::mutex mu;
};

static void batch_free(dfft_batch *b) {
    if (!b) return;
    if (b->d_tw64) (void)hipFree(b->d_tw64);
    if (b->d_tw32) (void)hipFree(b->d_tw32);
    if (b->d_items) (void)hipFree(b->d_items);
    if (b->h_items) (void)hipHostFree(b->h_items);
    if (b->d_ctl) (void)hipFree(b->d_ctl);
    if (b->d_done) (void)hipFree(b->d_done);
    delete b;
}

extern "C" {

int dfft_batch_create(dfft_batch **batch, int count, const dfft_batch_desc *descs, uint32_t flags) {
    if (!batch || count < 1 || !descs) return DFFT_ERR_ARG;
    *batch = nullptr;
    for (int t = 0; t < count; ++t) {
        const dfft_batch_desc &d = descs[t];
        if (d.kind != DFFT_C2C && d.kind != DFFT_R2C) return DFFT_ERR_UNSUPPORTED;
        if (d.sign != -1 && d.sign != 1) return DFFT_ERR_ARG;
        if (d.kind == DFFT_R2C && d.sign != -1) return DFFT_ERR_ARG;
        if (d.precision != DFFT_F32 && d.precision != DFFT_F64) return DFFT_ERR_ARG;
        for (int a = 0; a < 3; ++a)
            if (d.n[a] < 1 || d.n[a] > (1 << 20)) return DFFT_ERR_ARG;
    }
    dfft_batch *b = nullptr;
    try {
        b = new dfft_batch();
    } catch (...) {
        return DFFT_ERR_OOM;
    }
    b->count = count;
    b->descs.assign(descs, descs + count);
    TwiddleTable tt64, tt32;
    tt64.prec = 1;
    tt32.prec = 0;
    // passes per transform, in stage order
    struct Tmp {
        PassPlan pp;
        int t, level;
        double work;
    };
    std::vector<Tmp> tmp;
    int rc = DFFT_OK;
    for (int t = 0; t < count && rc == DFFT_OK; ++t) {
        const dfft_batch_desc &d = descs[t];
        const int prec = d.precision;
        const int es_c = prec ? 16 : 8;
        const bool inv = d.sign > 0;
        const bool norm = !(d.flags & DFFT_FLAG_UNNORMALIZED);
        int64_t n[3] = {d.n[0], d.n[1], d.n[2]};
        int64_t nc[3] = {n[0], n[1], n[2]};
        const int64_t N = n[0] * n[1] * n[2];
        std::vector<int> order = {0, 1, 2};
        if (d.kind == DFFT_R2C) nc[0] = n[0] / 2 + 1;
        else if (inv) std::reverse(order.begin(), order.end());
        for (int i = 0; i < 3; ++i) {
            const int ax = order[i];
            PassPlan pp{};
            pp.prec = prec;
            pp.axis = ax;
            pp.src = i == 0 ? BUF_IN : BUF_OUT;
            pp.dst = BUF_OUT;
            AxisPass &a = pp.ap;
            if (d.kind == DFFT_R2C && i == 0) {
                const int L = (int)n[0];
                if (L % 2 != 0) { rc = DFFT_ERR_UNSUPPORTED_SIZE; break; }
                a.op = OP_R2C;
                a.sp.L = L / 2;
                a.nlog = L;
                a.scale = 1.0;
                a.inverse = 0;
                // axis 0 of a rank-3 array: contiguous lines
                a.contig = 1;
                a.nlines = n[1] * n[2];
                a.nouter = 1;
                a.in_sm = 1;
                a.in_so = n[0];
                a.out_sm = 1;
                a.out_so = nc[0];
                pp.alg_bytes = (double)N * es_c / 2 + (double)nc[0] * n[1] * n[2] * es_c;
            } else {
                const int64_t *g = d.kind == DFFT_R2C ? nc : n;
                const int L = (int)g[ax];
                a.op = OP_C2C;
                a.sp.L = L;
                a.nlog = L;
                a.inverse = inv ? 1 : 0;
                a.scale = (inv && norm) ? 1.0 / L : 1.0;
                const int64_t st[3] = {1, g[0], g[0] * g[1]};
                if (st[ax] == 1) {
                    a.contig = 1;
                    a.nlines = g[0] * g[1] * g[2] / g[ax];
                    a.nouter = 1;
                    a.in_sm = a.out_sm = 1;
                    a.in_so = a.out_so = g[ax];
                } else if (ax == 1) {
                    a.contig = 0;
                    a.nlines = g[0];
                    a.nouter = g[2];
                    a.in_sm = a.out_sm = st[1];
                    a.in_so = a.out_so = st[2];
                } else {
                    a.contig = 0;
                    a.nlines = g[0] * g[1];
                    a.nouter = 1;
                    a.in_sm = a.out_sm = st[2];
                    a.in_so = a.out_so = 0;
                }
                pp.alg_bytes = 2.0 * (double)(g[0] * g[1] * g[2]) * es_c;
            }
            Tuning tn;
            const Tuning *tune = nullptr;
            if (!(flags & DFFT_FLAG_NO_TUNE) && tuning_lookup(pp, tn) && tn.tpl <= BATCH_THREADS) tune = &tn;
            rc = schedule_pass(pp, tune, BATCH_THREADS, BATCH_LDS);
            if (rc != DFFT_OK) break;
            (prec ? tt64 : tt32).add(pp.ap);
            tmp.push_back(Tmp{pp, t, i, pp.alg_bytes});
        }
    }
    if (rc != DFFT_OK) {
        batch_free(b);
        return rc;
    }
    // ticket order: by pass level, larger work first within a level
    std::vector<int> idx(tmp.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) {
        if (tmp[x].level != tmp[y].level) return tmp[x].level < tmp[y].level;
        return tmp[x].work > tmp[y].work;
    });
    std::vector<int> pos(tmp.size());
    for (size_t k = 0; k < idx.size(); ++k) pos[idx[k]] = (int)k;
    // index of (transform, level) -> tmp index
    std::vector<std::vector<int>> by_t(count, std::vector<int>(3, -1));
    for (size_t k = 0; k < tmp.size(); ++k) by_t[tmp[k].t][tmp[k].level] = (int)k;
    int64_t first = 0;
    for (size_t k = 0; k < idx.size(); ++k) {
        const Tmp &m = tmp[idx[k]];
        BatchItem it{};
        it.ap = m.pp.ap;
        it.prec = m.pp.prec;
        it.ntiles = m.pp.nblocks;
        it.first = first;
        first += it.ntiles;
        it.dep = -1;
        it.dep_need = 0;
        if (m.level > 0) {
            const int prev = by_t[m.t][m.level - 1];
            it.dep = pos[prev];
            it.dep_need = tmp[prev].pp.nblocks;
        }
        b->items.push_back(it);
        b->item_transform.push_back(m.t);
        b->item_src.push_back(m.pp.src);
        b->item_dst.push_back(m.pp.dst);
        b->lds = std::max(b->lds, m.pp.lds);
    }
    b->total = first;
    // device resources
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0) {
        batch_free(b);
        return DFFT_ERR_HIP;
    }
    b->grid = (int)std::min<int64_t>((int64_t)ncu * 2, b->total);
    if ((tt64.tab.size() && (rc = tt64.upload(&b->d_tw64, &b->tw64_bytes)) != DFFT_OK) ||
        (tt32.tab.size() && (rc = tt32.upload(&b->d_tw32, &b->tw32_bytes)) != DFFT_OK)) {
        batch_free(b);
        return rc;
    }
    for (auto &it : b->items) it.ap.tw = it.prec ? b->d_tw64 : b->d_tw32;
    const size_t ib = sizeof(BatchItem) * b->items.size();
    if (hipMalloc(&b->d_items, ib) != hipSuccess || hipHostMalloc(&b->h_items, ib, hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&b->d_ctl, sizeof(BatchCtl)) != hipSuccess ||
        hipMalloc(&b->d_done, sizeof(unsigned) * b->items.size()) != hipSuccess) {
        batch_free(b);
        return DFFT_ERR_OOM;
    }
    if (b->lds > 65536 &&
        hipFuncSetAttribute((const void *)k_batch, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b->lds) !=
            hipSuccess) {
        batch_free(b);
        return DFFT_ERR_HIP;
    }
    *batch = b;
    return DFFT_OK;
}

int dfft_batch_execute(dfft_batch *b, void *const *d_in, void *const *d_out, void *stream) {
    if (!b || !d_in || !d_out) return DFFT_ERR_ARG;
    for (int t = 0; t < b->count; ++t) {
        if (!d_in[t] || !d_out[t]) return DFFT_ERR_ARG;
        if (d_in[t] == d_out[t] && b->descs[t].kind != DFFT_C2C) return DFFT_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> g(b->mu);
    // the pinned staging buffer may still feed an earlier copy on this stream
    if (hipStreamSynchronize(st) != hipSuccess) return DFFT_ERR_HIP;
    for (size_t k = 0; k < b->items.size(); ++k) {
        BatchItem it = b->items[k];
        const int t = b->item_transform[k];
        it.ap.in = b->item_src[k] == BUF_IN ? d_in[t] : d_out[t];
        it.ap.out = d_out[t];
        b->h_items[k] = it;
    }
    if (hipMemcpyAsync(b->d_items, b->h_items, sizeof(BatchItem) * b->items.size(), hipMemcpyHostToDevice, st) !=
        hipSuccess)
        return DFFT_ERR_HIP;
    const int nit = (int)b->items.size();
    hipLaunchKernelGGL(k_batch_reset, dim3((nit + 255) / 256), dim3(256), 0, st, b->d_ctl, b->d_done, nit);
    hipLaunchKernelGGL(k_batch, dim3(b->grid), dim3(BATCH_THREADS), b->lds, st, b->d_items, nit, b->total, b->d_ctl,
                       b->d_done, b->spin_limit);
    return hipGetLastError() == hipSuccess ? DFFT_OK : DFFT_ERR_HIP;
}

int dfft_batch_num_items(const dfft_batch *b, int64_t *items) {
    if (!b || !items) return DFFT_ERR_ARG;
    *items = b->total;
    return DFFT_OK;
}

int dfft_batch_workspace_size(const dfft_batch *b, size_t *bytes) {
    if (!b || !bytes) return DFFT_ERR_ARG;
    *bytes = 0;
    return DFFT_OK;
}

int dfft_batch_algorithmic_bytes(const dfft_batch *b, double *bytes) {
    if (!b || !bytes) return DFFT_ERR_ARG;
    double s = 0;
    (void)s;
    *bytes = 0;
    for (int t = 0; t < b->count; ++t) {
        const dfft_batch_desc &d = b->descs[t];
        const double es_c = d.precision ? 16.0 : 8.0;
        const double N = (double)d.n[0] * d.n[1] * d.n[2];
        if (d.kind == DFFT_C2C) *bytes += 3 * 2 * N * es_c;
        else {
            const double NC = (double)(d.n[0] / 2 + 1) * d.n[1] * d.n[2];
            *bytes += N * es_c / 2 + NC * es_c + 2 * 2 * NC * es_c;
        }
    }
    return DFFT_OK;
}

int dfft_batch_error(dfft_batch *b, void *stream, int *flag) {
    if (!b || !flag) return DFFT_ERR_ARG;
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return DFFT_ERR_HIP;
    BatchCtl c;
    if (hipMemcpy(&c, b->d_ctl, sizeof(c), hipMemcpyDeviceToHost) != hipSuccess) return DFFT_ERR_HIP;
    *flag = (int)c.err;
    return DFFT_OK;
}

int dfft_batch_set_spin_limit(dfft_batch *b, uint64_t limit) {
    if (!b || limit == 0) return DFFT_ERR_ARG;
    b->spin_limit = limit;
    return DFFT_OK;
}

int dfft_batch_destroy(dfft_batch *b) {
    if (!b) return DFFT_ERR_ARG;
    batch_free(b);
    return DFFT_OK;
}

}  // extern "C"
