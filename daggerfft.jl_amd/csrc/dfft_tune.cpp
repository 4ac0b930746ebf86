// dfft_tune.cpp -- the autotuner and its schedule cache (the analogue of the
// reference's FFTW wisdom: PLAN_CACHE + import/export_wisdom,
// src/fft.jl:113-124,162).
//
// For every distinct axis pass of a plan the tuner times candidate schedules
// on the device -- radix order, threads per line, lines/columns per
// workgroup, LDS padding -- and records the fastest under a key of (precision,
// op, length, layout, size class).  dfft_plan_create and dfft_batch_create
// consult the cache unless DFFT_FLAG_NO_TUNE is set.  The cache can be saved
// to and loaded from a text file (one entry per line).
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <set>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/daggerfft_hip.h"
#include "dfft_kernels.hpp"
#include "dfft_plan.hpp"

namespace dfft {

static std::mutex g_mu;
static std::unordered_map<std::string, Tuning> g_cache;
static std::once_flag g_wisdom_once;
static bool g_wisdom_done = false;  // guarded by g_mu
static std::string g_autosave;      // guarded by g_mu: where tuning results are saved ("" none)

// The stamp of a wisdom file: the HIP version the library's kernels were
// compiled with and the GPU architecture its schedules were measured on.
// The compiler decides the kernels' code and so which schedule is fastest;
// the runtime in a process may be another build (PyTorch ships its own), and
// does not change the kernels.  A stamped file of another build or
// architecture is refused (FFTW likewise ignores wisdom of another FFTW
// build); hand-written files without a stamp are schedule overrides and
// always load.
static int build_version() { return HIP_VERSION; }

static std::string device_arch() {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        (void)hipGetLastError();
        return "unknown";
    }
    std::string arch = prop.gcnArchName;
    return arch.substr(0, arch.find(':'));
}

static std::string stamp_line() {
    return "# daggerfft_hip schedule cache hip=" + std::to_string(build_version()) + " arch=" + device_arch() +
           ": key radices tpl tc pad_shift variant swizzle\n";
}

static void write_entry(std::ostream &f, const std::string &key, const Tuning &t) {
    f << key << " ";
    for (size_t i = 0; i < t.radices.size(); ++i) f << (i ? "," : "") << t.radices[i];
    if (t.radices.empty()) f << "-";
    f << " " << t.tpl << " " << t.tc << " " << t.pad_shift << " " << t.variant << " " << t.swz << "\n";
}

// Writes `entries` under the stamp (the caller holds no lock).
int wisdom_write(const char *path, const std::vector<std::pair<std::string, Tuning>> &entries) {
    if (!path || !*path) return DFFT_ERR_ARG;
    std::ofstream f(path);
    if (!f) return DFFT_ERR_ARG;
    f << stamp_line();
    for (const auto &kv : entries) write_entry(f, kv.first, kv.second);
    return f.good() ? DFFT_OK : DFFT_ERR_ARG;
}

static std::vector<std::pair<std::string, Tuning>> snapshot() {
    std::lock_guard<std::mutex> g(g_mu);
    std::vector<std::pair<std::string, Tuning>> v(g_cache.begin(), g_cache.end());
    std::sort(v.begin(), v.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    return v;
}

// save_wisdom() after the cache changed (src/fft.jl:122-124,162), when a
// save path is set (dfft_tune_autosave or $DFFT_WISDOM_SAVE)
static void autosave() {
    std::string path;
    {
        std::lock_guard<std::mutex> g(g_mu);
        path = g_autosave;
    }
    if (!path.empty()) (void)wisdom_write(path.c_str(), snapshot());
}

static int size_class(const AxisPass &a) {
    const double lines = (double)a.nlines * (double)a.nouter;
    return (int)std::floor(std::log2(std::max(1.0, lines * a.sp.L)));
}

std::string tuning_key(const PassPlan &pp) {
    const AxisPass &a = pp.ap;
    // strided passes over outer slices (axis 1) and over whole planes (axis 2) differ
    const int layout = a.contig ? 1 : (a.nouter > 1 ? 2 : 0);
    return std::to_string(pp.prec) + "/" + std::to_string(a.op) + "/" + std::to_string(a.sp.L) + "/" +
           std::to_string(layout) + "/" + std::to_string(size_class(a));
}

static int load_file(const char *path, bool quiet = false);

// The schedule cache shipped beside the library (the reference imports its
// FFTW wisdom file at module init, src/fft.jl:116-120): $DFFT_WISDOM if set
// ("off" disables), else <library dir>/wisdom/<gfx arch>.txt.  A missing file
// is not an error.
void wisdom_autoload() {
    std::call_once(g_wisdom_once, [] {
        {
            std::lock_guard<std::mutex> g(g_mu);
            if (g_wisdom_done) return;
            g_wisdom_done = true;
        }
        if (const char *e = getenv("DFFT_WISDOM_SAVE")) {
            std::lock_guard<std::mutex> g(g_mu);
            g_autosave = e;
        }
        std::string path;
        if (const char *e = getenv("DFFT_WISDOM")) {
            if (!strcmp(e, "off") || !*e) return;
            path = e;
        } else {
            Dl_info info;
            if (!dladdr((void *)&wisdom_autoload, &info) || !info.dli_fname) return;
            std::string lib = info.dli_fname;
            const size_t slash = lib.rfind('/');
            const std::string dir = slash == std::string::npos ? "." : lib.substr(0, slash);
            const std::string arch = device_arch();
            if (arch == "unknown") return;
            path = dir + "/wisdom/" + arch + ".txt";
        }
        (void)load_file(path.c_str(), true);
    });
}

bool tuning_lookup(const PassPlan &pp, Tuning &t) {
    std::lock_guard<std::mutex> g(g_mu);
    auto it = g_cache.find(tuning_key(pp));
    if (it == g_cache.end()) return false;
    t = it->second;
    return true;
}

static void tuning_store(const std::string &key, const Tuning &t) {
    std::lock_guard<std::mutex> g(g_mu);
    g_cache[key] = t;
}

bool tuning_lookup_key(const std::string &key, Tuning &t) {
    std::lock_guard<std::mutex> g(g_mu);
    auto it = g_cache.find(key);
    if (it == g_cache.end()) return false;
    t = it->second;
    return true;
}

void tuning_store_key(const std::string &key, const Tuning &t) {
    tuning_store(key, t);
    autosave();
}

// Radix orders to try for length L: all splits of the power-of-two part
// into 8/4/2 factors (as multisets), each in descending and ascending order,
// followed/preceded by the odd factors.
static std::vector<std::vector<int>> radix_candidates(int L) {
    std::vector<std::vector<int>> out;
    int p2 = 0, m = L;
    while (m % 2 == 0) { m /= 2; ++p2; }
    std::vector<int> odd;
    for (int r : {3, 5, 7})
        while (m % r == 0) { odd.push_back(r); m /= r; }
    if (m != 1) return out;
    std::set<std::vector<int>> seen;
    for (int n8 = 0; 3 * n8 <= p2; ++n8)
        for (int n4 = 0; 3 * n8 + 2 * n4 <= p2; ++n4) {
            const int n2 = p2 - 3 * n8 - 2 * n4;
            if (n2 > 2) continue;  // long radix-2 chains are never competitive
            std::vector<int> pw;
            pw.insert(pw.end(), n8, 8);
            pw.insert(pw.end(), n4, 4);
            pw.insert(pw.end(), n2, 2);
            for (int dir = 0; dir < 2; ++dir) {
                std::vector<int> a = pw;
                if (dir) std::reverse(a.begin(), a.end());
                std::vector<int> c1 = a, c2 = odd;
                c1.insert(c1.end(), odd.begin(), odd.end());
                c2.insert(c2.end(), a.begin(), a.end());
                for (auto &c : {c1, c2})
                    if ((int)c.size() <= MAX_STAGES && seen.insert(c).second) out.push_back(c);
            }
        }
    if (L == 1) out.push_back({});
    return out;
}

static std::vector<Tuning> candidates(const PassPlan &pp, int level, int max_threads) {
    const AxisPass &a = pp.ap;
    std::vector<Tuning> out;
    if (a.blue) {
        // a convolution pass (Rader, Bluestein) keeps its engine and takes
        // only its tile width and swizzle from the cache
        const int var = a.blue == 2 ? 2 : a.variant ? 3 : 0;
        const std::vector<int> tcs = a.contig ? std::vector<int>{2, 4, 6, 8, 12, 16, 24, 32}
                                              : std::vector<int>{8, 16, 32, 64};
        const std::vector<int> swizzles =
            a.contig ? std::vector<int>{1, SWZ_XCD} : std::vector<int>{1, 97, 1021, SWZ_XCD};
        for (int tc : tcs) {
            if (tc * a.TPL > max_threads) continue;
            for (int sw : swizzles) {
                Tuning t;
                t.variant = var;
                t.tc = tc;
                t.swz = sw;
                out.push_back(t);
            }
        }
        return out;
    }
    std::vector<int> tcs = a.contig ? std::vector<int>{1, 2, 4, 8, 16} : std::vector<int>{4, 8, 16, 32, 64};
    std::vector<int> pads = a.contig ? (level >= 2 ? std::vector<int>{3, 4, 5} : std::vector<int>{3, 4})
                                     : std::vector<int>{0};
    // strided passes also try spreading the tiles that run at one time over
    // the array (AxisPass::swz; odd multipliers, coprime to power-of-two tile
    // counts); both layouts try contiguous tile ranges per XCD
    std::vector<int> swizzles = a.contig ? std::vector<int>{1, SWZ_XCD} : std::vector<int>{1, 97, 1021, SWZ_XCD};
    // compile-time schedules of this length: every radix list the catalog
    // holds, times the tile shape
    for (int var : {1, 2, 3}) {
        for (const auto &list : fixed_lists(pp.prec, a.contig, var, a.op, a.sp.L)) {
            std::vector<int> R = list;
            int tpl = 0;
            if (!fixed_schedule(pp.prec, a.contig, var, a.op, a.sp.L, R, tpl)) continue;
            for (int tc : tcs) {
                if (tpl * tc > max_threads) continue;
                for (int ps : pads)
                    for (int sw : swizzles) {
                        Tuning t;
                        t.variant = var;
                        t.radices = R;
                        t.tc = tc;
                        t.pad_shift = ps;
                        t.swz = sw;
                        out.push_back(t);
                    }
            }
        }
    }
    // runtime schedules: radix orders x threads per line x tile shape
    auto rads = radix_candidates(a.sp.L);
    if (level <= 1 && rads.size() > 4) rads.resize(4);
    for (const auto &R : rads) {
        std::set<int> tpls;
        for (int e : {8, 4}) {
            int tpl = std::max(1, (a.sp.L + e - 1) / e);
            for (;;) {  // smallest thread count >= L/e that holds every stage in EMAX registers
                bool ok = true;
                for (int r : R)
                    if ((a.sp.L / r + tpl - 1) / tpl > EMAX / r) { ok = false; break; }
                if (ok || tpl >= max_threads) break;
                ++tpl;
            }
            if (tpl <= max_threads) tpls.insert(tpl);
        }
        for (int tpl : tpls)
            for (int tc : tcs) {
                if (tpl * tc > max_threads) continue;
                for (int ps : pads) {
                    Tuning t;
                    t.radices = R;
                    t.tpl = tpl;
                    t.tc = tc;
                    t.pad_shift = ps;
                    t.variant = 0;
                    out.push_back(t);
                }
            }
    }
    return out;
}

}  // namespace dfft

using namespace dfft;

struct dfft_plan;  // defined in dfft_plan.cpp
namespace dfft {
int plan_pass_count(const dfft_plan *p);
PassPlan plan_pass(const dfft_plan *p, int i);
size_t plan_buffer_bytes(const dfft_plan *p, int buf);
}  // namespace dfft

namespace dfft {
static int autotune_impl(int rank, const int64_t *n, int kind, int sign, const int *r2r_kinds, int precision,
                         int decomp, const int *dims, int ndims, uint32_t flags, int level, void *stream,
                         bool only_missing);

// DFFT_FLAG_MEASURE at plan creation: the passes that have no schedule yet
// are timed first (FFTW.MEASURE, src/fft.jl:127), and the cache is saved if a
// save path is set.
int measure_plan(int rank, const int64_t *n, int kind, int sign, const int *r2r_kinds, int precision, int decomp,
                 const int *dims, int ndims, uint32_t flags) {
    return autotune_impl(rank, n, kind, sign, r2r_kinds, precision, decomp, dims, ndims,
                         flags & ~(uint32_t)DFFT_FLAG_MEASURE, 1, nullptr, true);
}
}  // namespace dfft

extern "C" {

int dfft_autotune(int rank, const int64_t *n, int kind, int sign, const int *r2r_kinds, int precision, int decomp,
                  const int *dims, int ndims, uint32_t flags, int level, void *stream) {
    return autotune_impl(rank, n, kind, sign, r2r_kinds, precision, decomp, dims, ndims,
                         flags & ~(uint32_t)DFFT_FLAG_MEASURE, level, stream, false);
}

}  // extern "C"

namespace dfft {
static int autotune_impl(int rank, const int64_t *n, int kind, int sign, const int *r2r_kinds, int precision,
                         int decomp, const int *dims, int ndims, uint32_t flags, int level, void *stream,
                         bool only_missing) {
    if (level < 0) return DFFT_ERR_ARG;
    if (level == 0) return DFFT_OK;
    dfft_plan *p = nullptr;
    int rc = dfft_plan_create(&p, rank, n, kind, sign, r2r_kinds, precision, decomp, dims, ndims,
                              flags | DFFT_FLAG_NO_TUNE);
    if (rc != DFFT_OK) return rc;
    hipStream_t st = (hipStream_t)stream;
    void *buf[3] = {nullptr, nullptr, nullptr};
    for (int b = 0; b < 3; ++b) {
        size_t nb = plan_buffer_bytes(p, b);
        if (nb && hipMalloc(&buf[b], nb) != hipSuccess) {
            rc = DFFT_ERR_OOM;
            break;
        }
        if (nb) (void)launch_fill(buf[b], (int64_t)(nb / (precision ? 8 : 4)), precision, 7 + b, st);
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == DFFT_OK) {
        if (hipError_t e = hipEventCreate(&e0)) rc = hip_fail(e, "hipEventCreate");
        else if (hipError_t e2 = hipEventCreate(&e1)) rc = hip_fail(e2, "hipEventCreate");
    }
    std::set<std::string> done;
    const int reps = level >= 2 ? 5 : 3;
    const bool verbose = getenv("DFFT_TUNE_VERBOSE") != nullptr;
    // times `n` launches of a scheduled candidate; < 0 if it cannot run
    auto time_pass = [&](PassPlan pp, int n) -> float {
        TwiddleTable tt;
        tt.prec = pp.prec;
        tt.add(pp.ap);
        void *d_tw = nullptr;
        size_t twb = 0;
        if (tt.upload(&d_tw, &twb) != DFFT_OK) return -1.0f;
        pp.ap.tw = d_tw;
        pp.ap.in = buf[pp.src];
        pp.ap.out = buf[pp.dst];
        bool ok = launch_axis(pp.ap, pp.prec, pp.nblocks, pp.threads, pp.lds, st) == hipSuccess;
        float ms = -1.0f;
        if (ok) {
            (void)hipEventRecord(e0, st);
            for (int r = 0; r < n && ok; ++r)
                ok = launch_axis(pp.ap, pp.prec, pp.nblocks, pp.threads, pp.lds, st) == hipSuccess;
            (void)hipEventRecord(e1, st);
            ok = ok && hipEventSynchronize(e1) == hipSuccess;
            if (ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) ms /= n;
            else ms = -1.0f;
        }
        (void)hipFree(d_tw);
        if (!ok) (void)hipGetLastError();
        return ms;
    };
    for (int i = 0; rc == DFFT_OK && i < plan_pass_count(p); ++i) {
        PassPlan base = plan_pass(p, i);
        const std::string key = tuning_key(base);
        if (!done.insert(key).second) continue;
        Tuning have;
        if (only_missing && tuning_lookup(base, have)) continue;
        // first a quick timing of every candidate, then the best few again
        // with more launches, keeping the best of three trials: candidates
        // within a few per cent of each other are otherwise ranked by noise
        std::vector<std::pair<float, size_t>> quick;
        std::vector<std::pair<Tuning, PassPlan>> cands;
        for (const Tuning &cand : candidates(base, level, 1024)) {
            PassPlan pp = base;
            if (schedule_pass(pp, &cand) != DFFT_OK) continue;
            const float ms = time_pass(pp, reps);
            if (ms < 0) continue;
            quick.emplace_back(ms, cands.size());
            cands.emplace_back(cand, pp);
            if (verbose) {
                std::string rad;
                for (int r : cand.radices) rad += (rad.empty() ? "" : "x") + std::to_string(r);
                fprintf(stderr,
                        "tune %s radices=%s var=%d tpl=%d tc=%d pad=%d swz=%d threads=%d lds=%zu: %.4f ms %.1f GB/s\n",
                        key.c_str(), rad.c_str(), pp.ap.variant, pp.ap.TPL, pp.ap.TC, pp.ap.pad_shift, pp.ap.swz,
                        pp.threads, pp.lds, ms, pp.alg_bytes / (ms * 1e-3) / 1e9);
            }
        }
        if (quick.empty()) continue;
        std::sort(quick.begin(), quick.end());
        const size_t top = std::min<size_t>(quick.size(), 4);
        float best = 1e30f;
        size_t best_i = quick[0].second;
        for (size_t k = 0; k < top; ++k) {
            const size_t ci = quick[k].second;
            float m = 1e30f;
            for (int trial = 0; trial < 3; ++trial) {
                const float ms = time_pass(cands[ci].second, 3 * reps);
                if (ms >= 0 && ms < m) m = ms;
            }
            if (m < best) {
                best = m;
                best_i = ci;
            }
        }
        tuning_store(key, cands[best_i].first);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (void *b : buf)
        if (b) (void)hipFree(b);
    dfft_plan_destroy(p);
    if (rc == DFFT_OK) autosave();
    return rc;
}
}  // namespace dfft

extern "C" {

int dfft_tune_save(const char *path) { return wisdom_write(path, snapshot()); }

int dfft_tune_autosave(const char *path) {
    std::lock_guard<std::mutex> g(g_mu);
    g_autosave = path ? path : "";
    return DFFT_OK;
}

int dfft_tune_load(const char *path) {
    if (!path) return DFFT_ERR_ARG;
    return load_file(path);
}

}  // extern "C"

namespace dfft {
static int load_file(const char *path, bool quiet) {
    std::ifstream f(path);
    if (!f) return DFFT_ERR_ARG;
    std::string line;
    std::vector<std::pair<std::string, Tuning>> entries;
    while (std::getline(f, line)) {
        if (line.empty()) continue;
        if (line[0] == '#') {
            // the stamp: hip=<HIP version of the build> arch=<gfx arch>
            std::istringstream hs(line);
            std::string tok;
            while (hs >> tok) {
                if (!tok.empty() && tok.back() == ':') tok.pop_back();
                const bool hip = tok.rfind("hip=", 0) == 0, arch = tok.rfind("arch=", 0) == 0;
                if (!hip && !arch) continue;
                const std::string val = tok.substr(hip ? 4 : 5);
                const std::string cur = hip ? std::to_string(build_version()) : device_arch();
                if (val != cur && val != "unknown" && cur != "unknown" && cur != "0") {
                    if (!quiet) {
                        const std::string msg = std::string("wisdom file ") + path + " has " + tok +
                                                ", this process runs " + (hip ? "hip=" : "arch=") + cur;
                        set_last_error(msg.c_str());
                    }
                    return DFFT_ERR_WISDOM;
                }
            }
            continue;
        }
        std::istringstream ss(line);
        std::string key, rad;
        Tuning t;
        if (!(ss >> key >> rad >> t.tpl >> t.tc >> t.pad_shift)) return DFFT_ERR_ARG;
        if (!(ss >> t.variant)) t.variant = 0;
        if (!(ss >> t.swz)) t.swz = 1;
        if (t.variant < 0 || t.variant > 3 || (t.swz < 1 && t.swz != SWZ_XCD)) return DFFT_ERR_ARG;
        if (rad != "-") {
            std::stringstream rs(rad);
            std::string tok;
            while (std::getline(rs, tok, ',')) {
                const int r = std::atoi(tok.c_str());
                // the runtime engine runs radices up to 8; catalog kernels also the composite ones
                const bool basic = r == 2 || r == 3 || r == 4 || r == 5 || r == 7 || r == 8;
                const bool composite = r == 6 || r == 9 || r == 10 || r == 12 || r == 14 || r == 15 || r == 16;
                if (!basic && !(composite && t.variant >= 1)) return DFFT_ERR_ARG;
                t.radices.push_back(r);
            }
        }
        if (t.tpl < 0 || t.tpl > 1024 || t.tc < 0 || t.tc > 1024 || t.pad_shift < 0 || t.pad_shift > 30)
            return DFFT_ERR_ARG;
        entries.emplace_back(key, t);
    }
    // all or nothing: a file that fails to parse leaves the cache as it was
    for (const auto &kv : entries) tuning_store(kv.first, kv.second);
    return DFFT_OK;
}

}  // namespace dfft

extern "C" {

int dfft_tune_clear(void) {
    std::lock_guard<std::mutex> g(g_mu);
    g_cache.clear();
    g_wisdom_done = true;  // a cleared cache stays empty
    return DFFT_OK;
}

int dfft_tune_entries(int *count) {
    if (!count) return DFFT_ERR_ARG;
    std::lock_guard<std::mutex> g(g_mu);
    *count = (int)g_cache.size();
    return DFFT_OK;
}

}  // extern "C"
