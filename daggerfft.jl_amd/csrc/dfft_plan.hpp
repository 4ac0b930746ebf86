// dfft_plan.hpp -- host-side pass descriptors shared by the planner, the
// batched launcher and the autotuner.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "dfft_types.hpp"

struct dfft_plan;  // the C ABI's plan object (include/daggerfft_hip.h)

namespace dfft {

constexpr int EMAX = 8;  // complex values a thread holds per stage (variant 0; variant 1 holds 16)

enum BufId : int { BUF_IN = 0, BUF_OUT = 1, BUF_WORK = 2 };

struct PassPlan {
    AxisPass ap{};
    int prec = 1;
    int axis = 0;
    int src = BUF_IN, dst = BUF_OUT;
    int64_t nblocks = 0;
    int threads = 0;
    size_t lds = 0;
    double alg_bytes = 0;
};

// A schedule override from the autotuner cache.
struct Tuning {
    std::vector<int> radices;
    int tpl = 0;
    int tc = 0;
    int pad_shift = 0;
    int variant = 0;  // see AxisPass::variant
    int swz = 1;      // see AxisPass::swz
};

// Records `what` and HIP's message as this thread's last error (dfft_last_error)
// and returns DFFT_ERR_HIP.
int hip_fail(int hip_status, const char *what);
// Records `msg` as this thread's last error (dfft_last_error).
void set_last_error(const char *msg);
bool factorize(int L, std::vector<int> &R);
void set_geometry(AxisPass &a, int axis, const int64_t nin[3], const int64_t nout[3]);
// grouped: the pass runs in a grouped launch (dfft_group.hip), whose kernel
// follows the runtime schedule with 8 values per thread -- only schedules
// with that shape (variants 0 and 1) qualify.
int schedule_pass(PassPlan &pp, const Tuning *tune, int max_threads = 1024, size_t max_lds = 0,
                  int contig_threads = 256, bool grouped = false);

// Offsets of one pass's blocks in a TwiddleTable.
struct TwiddleBlocks {
    StagePlan sp, bsp, bsp2;
    int chirp = 0, bhat = 0;
};

// Stage twiddle + pre/post blocks (+ Bluestein blocks) of every distinct
// (op, L, radices) of one precision, in one device table.
struct TwiddleTable {
    int prec = 1;
    std::vector<long double> tab;           // (re, im) pairs
    std::map<std::string, TwiddleBlocks> seen;
    void add(AxisPass &ap);                 // sets the offsets in ap.sp (and ap.bsp, chirp, bhat)
    int upload(void **d_ptr, size_t *bytes) const;
};
bool tuning_lookup(const PassPlan &pp, Tuning &t);
// schedule cache entries under a key of their own (a batch's launch schedule);
// storing saves the cache when a save path is set
bool tuning_lookup_key(const std::string &key, Tuning &t);
void tuning_store_key(const std::string &key, const Tuning &t);
// Writes `entries` as a stamped wisdom file.
int wisdom_write(const char *path, const std::vector<std::pair<std::string, Tuning>> &entries);
// DFFT_FLAG_MEASURE: tune the passes of this plan that have no cache entry.
int measure_plan(int rank, const int64_t *n, int kind, int sign, const int *r2r_kinds, int precision, int decomp,
                 const int *dims, int ndims, uint32_t flags);
void wisdom_autoload();
std::string tuning_key(const PassPlan &pp);
void cis_ld(int64_t k, int64_t n, long double *c, long double *s);

// Plan execution shared by dfft_execute and dfft_execute_static.
// Argument checks of an execution (DFFT_OK or DFFT_ERR_ARG).
int plan_check_io(const ::dfft_plan *p, const void *d_in, const void *d_out);
// Device workspace bytes the plan's passes use (0: none).
size_t plan_work_bytes(const ::dfft_plan *p);
// The device the plan was created on.
int plan_device(const ::dfft_plan *p);
// Enqueues the plan's launches on `stream` with `work` as its workspace;
// graph: replay the plan's hipGraph if it was created with DFFT_FLAG_GRAPH.
int plan_enqueue(const ::dfft_plan *p, const void *d_in, void *d_out, void *work, void *stream, bool graph);

}  // namespace dfft
