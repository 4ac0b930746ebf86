// dfft_plan.cpp -- planner and C ABI of libdaggerfft_hip.
//
// A plan is an immutable list of axis passes.  Each pass is one kernel
// launch over the whole array (one GPU holds the whole block, so the
// reference's chunk stages collapse to one launch per transformed axis and
// its redistribution collapses to strided addressing inside the pass).
// Stage order follows the reference drivers: forward dims[0], dims[1],
// dims[2]; inverse reversed (fft!/ifft!, src/fft.jl:651-724; Slab
// src/fft.jl:727-776).
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/daggerfft_hip.h"
#include "dfft_group.hpp"
#include "dfft_kernels.hpp"
#include "dfft_plan.hpp"

namespace dfft {

static thread_local char g_last_error[256] = "";

int hip_fail(int hip_status, const char *what) {
    snprintf(g_last_error, sizeof g_last_error, "%s: %s (hip status %d)", what,
             hipGetErrorString((hipError_t)hip_status), hip_status);
    return DFFT_ERR_HIP;
}

void set_last_error(const char *msg) { snprintf(g_last_error, sizeof g_last_error, "%s", msg); }

// ---------------------------------------------------------------------------
// twiddles: exp(-2 pi i k / n) evaluated in long double with octant folding,
// rounded once to the pass precision
// ---------------------------------------------------------------------------
void cis_ld(int64_t k, int64_t n, long double *c, long double *s) {
    k %= n;
    if (k < 0) k += n;
    long double x = (long double)k / (long double)n;
    bool neg_s = false, neg_c = false;
    if (x > 0.5L) { x = 1.0L - x; neg_s = true; }
    if (x > 0.25L) { x = 0.5L - x; neg_c = true; }
    const long double TWO_PI = 6.283185307179586476925286766559005768L;
    long double cc, ss;
    if (x > 0.125L) {
        long double y = (0.25L - x) * TWO_PI;
        cc = sinl(y);
        ss = cosl(y);
    } else {
        long double y = x * TWO_PI;
        cc = cosl(y);
        ss = sinl(y);
    }
    *c = neg_c ? -cc : cc;
    *s = neg_s ? -ss : ss;
}

static void push_w(std::vector<long double> &tab, int64_t k, int64_t n) {
    long double c, s;
    cis_ld(k, n, &c, &s);
    tab.push_back(c);
    tab.push_back(-s);
}

bool factorize(int L, std::vector<int> &R) {
    R.clear();
    int n = L;
    while (n % 8 == 0 && n > 0) { R.push_back(8); n /= 8; }
    while (n % 4 == 0 && n > 0) { R.push_back(4); n /= 4; }
    while (n % 2 == 0 && n > 0) { R.push_back(2); n /= 2; }
    while (n % 3 == 0 && n > 0) { R.push_back(3); n /= 3; }
    while (n % 5 == 0 && n > 0) { R.push_back(5); n /= 5; }
    while (n % 7 == 0 && n > 0) { R.push_back(7); n /= 7; }
    return n == 1 && (int)R.size() <= MAX_STAGES;
}

// Build the stage twiddle blocks (+ the pre/post block of fused ops) into
// `tab` (pairs of long double); offsets are in complex elements.
static void build_twiddles(StagePlan &sp, int op, int nlog, std::vector<long double> &tab) {
    int Ns = 1;
    for (int s = 0; s < sp.S; ++s) {
        const int R = sp.R[s];
        sp.tw[s] = (int)(tab.size() / 2);
        if (s > 0)
            for (int k = 0; k < Ns; ++k)
                for (int r = 1; r < R; ++r) push_w(tab, (int64_t)k * r, (int64_t)Ns * R);
        Ns *= R;
    }
    sp.post = -1;
    if (op == OP_R2C || op == OP_C2R) {
        // w^k = exp(-2 pi i k / (2H)), k = 0..H
        sp.post = (int)(tab.size() / 2);
        for (int k = 0; k <= sp.L; ++k) push_w(tab, k, 2 * (int64_t)sp.L);
    } else if (op == OP_DCT2 || op == OP_DCT3 || op == OP_DST2 || op == OP_DST3) {
        // exp(-i pi k / 2L), k = 0..L-1
        sp.post = (int)(tab.size() / 2);
        for (int k = 0; k < sp.L; ++k) push_w(tab, k, 4 * (int64_t)sp.L);
    } else if (op == OP_DCT2H || op == OP_DCT3H) {
        // e^{-2 pi i k/n} then e^{-i pi k/2n}, k = 0..H (n = 2H = 2 sp.L)
        sp.post = (int)(tab.size() / 2);
        for (int k = 0; k <= sp.L; ++k) push_w(tab, k, 2 * (int64_t)sp.L);
        for (int k = 0; k <= sp.L; ++k) push_w(tab, k, 8 * (int64_t)sp.L);
    } else if (op == OP_DCT4 || op == OP_DST4) {
        // exp(-2 pi i m / 4L), m = 0..L-1 (L = 2n): pre-twiddle e^{-i pi p/2n} at
        // m = 2p, post-twiddle e^{-i pi (2k+1)/4n} at m = 2k+1
        sp.post = (int)(tab.size() / 2);
        for (int m = 0; m < sp.L; ++m) push_w(tab, m, 4 * (int64_t)sp.L);
    }
    (void)nlog;
}

// Bluestein blocks of a length-n pass convolving over M = ap.bsp.L: the
// stage twiddles of M, the chirp c_p = exp(-i pi p^2 / n) = w_{2n}^{p^2 mod 2n}
// (p < n; the exponent reduced exactly in integers), and bhat = FFT_M(b) / M
// for b_m = conj(c_m) at m and M-m (|m| < n), zero elsewhere.  b is even, so
// bhat_k = (b_0 + 2 sum_{m=1}^{n-1} b_m cos(2 pi m k / M)) / M, summed in long
// double.
static void build_bluestein(AxisPass &ap, std::vector<long double> &tab) {
    const int n = ap.sp.L, M = ap.bsp.L;
    build_twiddles(ap.bsp, OP_C2C, M, tab);
    if (ap.variant) build_twiddles(ap.bsp2, OP_C2C, M, tab);
    ap.chirp = (int)(tab.size() / 2);
    std::vector<long double> br(n), bi(n);
    for (int p = 0; p < n; ++p) {
        const int64_t e = ((int64_t)p * p) % (2 * (int64_t)n);
        long double c, sn;
        cis_ld(e, 2 * (int64_t)n, &c, &sn);  // c_p = c - i sn
        tab.push_back(c);
        tab.push_back(-sn);
        br[p] = c;  // b_p = conj(c_p) = c + i sn
        bi[p] = sn;
    }
    std::vector<long double> cosM(M);
    for (int k = 0; k < M; ++k) {
        long double c, sn;
        cis_ld(k, M, &c, &sn);
        cosM[k] = c;
    }
    ap.bhat = (int)(tab.size() / 2);
    for (int k = 0; k < M; ++k) {
        long double re = br[0], im = bi[0];
        int64_t mk = 0;  // m*k mod M
        for (int m = 1; m < n; ++m) {
            mk += k;
            if (mk >= M) mk -= M;
            re += 2 * br[m] * cosM[mk];
            im += 2 * bi[m] * cosM[mk];
        }
        tab.push_back(re / M);
        tab.push_back(im / M);
    }
}

static int64_t powmod(int64_t b, int64_t e, int64_t m) {
    int64_t r = 1;
    b %= m;
    for (; e > 0; e >>= 1, b = b * b % m)
        if (e & 1) r = r * b % m;
    return r;
}

// The smallest primitive root of the prime n: g^((n-1)/q) != 1 for every
// prime factor q of n - 1.
static int primitive_root(int n) {
    std::vector<int> qs;
    int m = n - 1;
    for (int q = 2; q * q <= m; ++q)
        if (m % q == 0) {
            qs.push_back(q);
            while (m % q == 0) m /= q;
        }
    if (m > 1) qs.push_back(m);
    for (int g = 2; g < n; ++g) {
        bool ok = true;
        for (int q : qs)
            if (powmod(g, (n - 1) / q, n) == 1) {
                ok = false;
                break;
            }
        if (ok) return g;
    }
    return 1;  // n = 2
}

static bool is_prime(int n) {
    if (n < 2) return false;
    for (int q = 2; q * q <= n; ++q)
        if (n % q == 0) return false;
    return true;
}

// Rader blocks of a prime length-n pass convolving over M = n - 1 (Rader
// 1968): the stage twiddles of M (both radix orders), the index pairs
// (g^p mod n, g^-p mod n), p < M, stored as the two parts of one table entry
// (exact in either precision for n < 2^24), and bhat = FFT_M(b) / M for
// b_m = w_n^(g^-m), summed directly in long double.
static void build_rader(AxisPass &ap, std::vector<long double> &tab) {
    const int n = ap.sp.L, M = ap.bsp.L;
    build_twiddles(ap.bsp, OP_C2C, M, tab);
    build_twiddles(ap.bsp2, OP_C2C, M, tab);
    const int g = primitive_root(n);
    std::vector<int> gp(M), gi(M);
    int64_t x = 1;
    for (int p = 0; p < M; ++p) {
        gp[p] = (int)x;
        x = x * g % n;
    }
    for (int p = 0; p < M; ++p) gi[p] = gp[(M - p) % M];
    ap.chirp = (int)(tab.size() / 2);
    for (int p = 0; p < M; ++p) {
        tab.push_back((long double)gp[p]);
        tab.push_back((long double)gi[p]);
    }
    std::vector<long double> br(M), bi(M), cM(M), sM(M);
    for (int m = 0; m < M; ++m) {
        long double c, sn;
        cis_ld(gi[m], n, &c, &sn);  // b_m = c - i sn
        br[m] = c;
        bi[m] = -sn;
    }
    for (int k = 0; k < M; ++k) cis_ld(k, M, &cM[k], &sM[k]);
    ap.bhat = (int)(tab.size() / 2);
    for (int k = 0; k < M; ++k) {
        long double re = 0, im = 0;
        int64_t mk = 0;  // m*k mod M
        for (int m = 0; m < M; ++m) {
            // b_m exp(-2 pi i m k / M)
            re += br[m] * cM[mk] + bi[m] * sM[mk];
            im += bi[m] * cM[mk] - br[m] * sM[mk];
            mk += k;
            if (mk >= M) mk -= M;
        }
        tab.push_back(re / M);
        tab.push_back(im / M);
    }
}

void TwiddleTable::add(AxisPass &ap) {
    std::string key = std::to_string(ap.op) + ":" + std::to_string(ap.sp.L);
    for (int s = 0; s < ap.sp.S; ++s) key += "," + std::to_string(ap.sp.R[s]);
    if (ap.blue) {
        key += "/B" + std::to_string(ap.blue) + "." + std::to_string(ap.variant) + ":" + std::to_string(ap.bsp.L);
        for (int s = 0; s < ap.bsp.S; ++s) key += "," + std::to_string(ap.bsp.R[s]);
    }
    auto it = seen.find(key);
    if (it != seen.end()) {
        ap.sp = it->second.sp;
        ap.bsp = it->second.bsp;
        ap.bsp2 = it->second.bsp2;
        ap.chirp = it->second.chirp;
        ap.bhat = it->second.bhat;
    } else {
        build_twiddles(ap.sp, ap.op, ap.nlog, tab);
        if (ap.blue == 2) build_rader(ap, tab);
        else if (ap.blue) build_bluestein(ap, tab);
        seen[key] = TwiddleBlocks{ap.sp, ap.bsp, ap.bsp2, ap.chirp, ap.bhat};
    }
}

int TwiddleTable::upload(void **d_ptr, size_t *bytes) const {
    const size_t es = prec ? sizeof(double) : sizeof(float);
    const size_t nb = std::max<size_t>(tab.size() * es, 16);
    std::vector<unsigned char> host(nb, 0);
    if (prec) {
        double *d = reinterpret_cast<double *>(host.data());
        for (size_t i = 0; i < tab.size(); ++i) d[i] = (double)tab[i];
    } else {
        float *f = reinterpret_cast<float *>(host.data());
        for (size_t i = 0; i < tab.size(); ++i) f[i] = (float)tab[i];
    }
    if (hipMalloc(d_ptr, nb) != hipSuccess) {
        *d_ptr = nullptr;
        return DFFT_ERR_OOM;
    }
    if (hipError_t e = hipMemcpy(*d_ptr, host.data(), nb, hipMemcpyHostToDevice)) return hip_fail(e, "twiddle upload");
    *bytes = nb;
    return DFFT_OK;
}

static int default_tpl(int L, const std::vector<int> &R, int emax, int target_e) {
    int tpl = std::max(1, (L + target_e - 1) / target_e);
    for (;;) {
        bool ok = true;
        for (int r : R) {
            const int nb = L / r;
            if ((nb + tpl - 1) / tpl > emax / r) { ok = false; break; }
        }
        if (ok || tpl >= 1024) break;
        ++tpl;
    }
    return tpl;
}

static bool tpl_ok(int L, const std::vector<int> &R, int emax, int tpl) {
    for (int r : R)
        if ((L / r + tpl - 1) / tpl > emax / r) return false;
    return tpl >= 1 && tpl <= 1024;
}

// ---------------------------------------------------------------------------
// schedule of one pass: radices + launch shape
// ---------------------------------------------------------------------------
int schedule_pass(PassPlan &pp, const Tuning *tune, int max_threads, size_t max_lds, int contig_threads,
                  bool grouped) {
    AxisPass &a = pp.ap;
    // A length with a prime factor > 7 runs Bluestein's convolution over a
    // 7-smooth M >= 2n-1 (runtime engine; the grouped kernel has no Bluestein
    // tile).  Each of its two FFTs costs about (stages x threads per line)
    // barrier-separated LDS rounds, so M is the candidate up to twice the
    // minimum with the least of that product (the smallest M on ties).
    // C2C passes, and the real ops along the contiguous axis, take the
    // compile-time engine instead (both FFTs unrolled, the exchange between
    // them in registers) over the catalog length (2^k, 3*2^k, 5*2^k) >= 2n-1
    // with the least stages x threads per line, unless the schedule cache
    // names the runtime engine (variant 0).
    a.blue = 0;
    a.bsp = StagePlan{};
    a.bsp2 = StagePlan{};
    std::vector<int> blue_fixed_R;
    int blue_fixed_tpl = 0, blue_fixed_variant = 0;
    const Tuning *conv_tune = nullptr;
    {
        std::vector<int> f;
        std::vector<int> rader_R;
        int rader_tpl = 0;
        if (!factorize(a.sp.L, f)) {
            if (grouped) return DFFT_ERR_UNSUPPORTED_SIZE;
            const int m0 = 2 * a.sp.L - 1;
            const bool runtime = tune && tune->variant == 0;
            a.blue = 1;
            int Mf = 0;
            // A prime C2C length whose n-1 has a compile-time Rader kernel may
            // run Rader's convolution of length n-1 instead (decided below)
            const bool rader_ok = !runtime && !(tune && tune->variant == 3) && a.op == OP_C2C &&
                                  is_prime(a.sp.L) &&
                                  rader_fixed_schedule(pp.prec, a.contig, a.sp.L - 1, rader_R, rader_tpl);
            if (!runtime) {
                int64_t best = 0;
                for (int m : blue_fixed_lengths(pp.prec, a.contig, a.op)) {
                    std::vector<int> r;
                    int t = 0, var = 0;
                    if (m < m0 || !blue_fixed_schedule(pp.prec, a.contig, a.op, m, r, t, var)) continue;
                    const int64_t cost = (int64_t)r.size() * t;
                    if (!Mf || cost < best || (cost == best && m < Mf)) {
                        Mf = m;
                        best = cost;
                    }
                }
            }
            // Rader convolves over n-1 instead of M >= 2n-1, but every stage of
            // its two FFTs is a barrier-separated LDS round trip like
            // Bluestein's: it wins only when its FFT has no more stages
            // (measured, FP64: 257 Rader over 8*8*4 1.7x faster than
            // Bluestein over 16*8*5; 251 over 2*5*5*5 1.15x slower than over
            // 16*16*2)
            // (the schedule cache can name Rader: variant 2)
            if (rader_ok) {
                std::vector<int> br;
                int bt = 0, bv = 0;
                if ((tune && tune->variant == 2) || !Mf ||
                    !blue_fixed_schedule(pp.prec, a.contig, a.op, Mf, br, bt, bv) || rader_R.size() <= br.size())
                    a.blue = 2;
            }
            if (a.blue == 2) {
                a.bsp.L = a.sp.L - 1;
                blue_fixed_R = rader_R;
                blue_fixed_tpl = rader_tpl;
                blue_fixed_variant = 2;
            } else if (Mf) {
                (void)blue_fixed_schedule(pp.prec, a.contig, a.op, Mf, blue_fixed_R, blue_fixed_tpl,
                                          blue_fixed_variant);
                a.bsp.L = Mf;
            } else {
                blue_fixed_R.clear();
                int M = 0;
                int64_t best = 0;
                for (int m = m0; m <= 2 * m0; ++m) {
                    if (!factorize(m, f)) continue;
                    const int64_t cost = (int64_t)f.size() * default_tpl(m, f, EMAX, EMAX);
                    if (!M || cost < best) {
                        M = m;
                        best = cost;
                    }
                }
                a.bsp.L = M;
            }
            a.sp.S = 0;
            conv_tune = tune;  // a convolution pass takes only its tile width and swizzle from the cache
            tune = nullptr;
        }
    }
    StagePlan &fp = a.blue ? a.bsp : a.sp;  // the plan of the FFT the tile runs
    const int L = fp.L;
    const int es = pp.prec ? 16 : 8;
    // variants 1-3: a compile-time schedule of this length (see dfft_fixed.hpp).
    // Default (measured on gfx950): the split-LDS schedule wherever the
    // catalog has it -- half the LDS per tile doubles the tiles in flight per
    // CU, worth 20-40% on every mixed-radix length -- except contiguous
    // powers of two, where 8 values per thread is as fast; then 16 values.
    std::vector<int> prefs;
    const bool pow2 = (L & (L - 1)) == 0;
    if (a.blue) prefs = {};
    else if (tune) prefs = {tune->variant};
    else if (a.contig && pow2) prefs = {1, 2, 3};
    else prefs = {2, 1, 3};
    if (tune && tune->variant >= 2) prefs.push_back(1);
    std::vector<int> R;
    int fixed_tpl = 0;
    a.variant = 0;
    for (int want : prefs) {
        if (want < 1 || (grouped && want != 1)) continue;
        // a tuned compile-time schedule names its radix list (several per length)
        R.clear();
        if (tune && want == tune->variant) R = tune->radices;
        if (fixed_schedule(pp.prec, a.contig, want, a.op, L, R, fixed_tpl)) {
            a.variant = want;
            break;
        }
    }
    if (!blue_fixed_R.empty()) {
        a.variant = blue_fixed_variant;
        R = blue_fixed_R;
        fixed_tpl = blue_fixed_tpl;
    }
    if (!a.variant) {
        if (a.op >= OP_COUNT) return DFFT_ERR_UNSUPPORTED_SIZE;  // catalog-only ops
        // a tuned radix list for the runtime engine (radices up to 8); a
        // catalog kernel's list (variant >= 1) may hold composite radices
        if (tune && tune->variant == 0 && !tune->radices.empty()) R = tune->radices;
        else if (!factorize(L, R)) return DFFT_ERR_UNSUPPORTED_SIZE;
    }
    int prod = 1;
    for (int r : R) prod *= r;
    if (prod != L || (int)R.size() > MAX_STAGES) return DFFT_ERR_UNSUPPORTED_SIZE;
    fp.S = (int)R.size();
    for (int s = 0; s < fp.S; ++s) fp.R[s] = R[s];
    if (a.blue && a.variant) {
        a.bsp2 = a.bsp;
        std::reverse(a.bsp2.R, a.bsp2.R + a.bsp2.S);
    }
    const int emax = a.variant >= 2 ? 16 : EMAX;
    // split exchange: half the LDS; a contiguous Rader line stages whole
    // complex values for its permuted gather and scatter (measured, 257 FP64:
    // 0.331 ms scattered, 0.297 ms staged in a slot per line; staging real and
    // imaginary parts in two rounds through half the slot, 0.314 ms; the
    // columnar staging is sized below)
    const int lds_es = (a.variant == 2 && !(a.blue == 2 && a.contig)) ? es / 2 : es;
    int tpl = a.variant ? fixed_tpl
              : (tune && tune->variant == 0 && tune->tpl > 0 && tpl_ok(L, R, emax, tune->tpl))
                  ? tune->tpl
                  : default_tpl(L, R, emax, emax);
    if (L == 1) tpl = 1;
    if (tpl > max_threads) return DFFT_ERR_UNSUPPORTED_SIZE;
    a.TPL = tpl;
    const int rows = L + 1;  // C2R stages H+1 bins
    const int lds_limit = max_lds ? (int)std::min<size_t>(max_lds, (size_t)axis_lds_limit()) : axis_lds_limit();
    if (a.contig && tpl <= 64 && a.blue == 2) {
        // Rader along contiguous lines of at most 64 threads: a columnar tile
        // of TC lines whose LDS holds element i of line ls at i * P + ls, P =
        // TC rounded up to odd, for the n staged complex values of each line
        // (rader_fixed_tile).  (The same staging for C2C Bluestein tiles, whose
        // accesses are not permuted, gained 0-2% and was not kept.)
        const int es_c = pp.prec ? 16 : 8;
        // (default: about 192 threads per tile; tuned, 257 took 12 lines of 16
        // threads and 251 took 4 of 25)
        int tc = (tune && tune->tc > 0) ? tune->tc : (conv_tune && conv_tune->tc > 0) ? conv_tune->tc
                                                                                     : std::max(1, 192 / tpl);
        const size_t per_row = std::max((size_t)a.sp.L * es_c, (size_t)a.bsp.L * lds_es);
        auto lds_of = [&](int c) { return per_row * (size_t)(c | 1); };
        while (tc > 1 && (int64_t)tc * tpl > max_threads) --tc;
        while (tc > 1 && lds_of(tc) > (size_t)lds_limit) --tc;
        if (lds_of(tc) > (size_t)lds_limit || (int64_t)tc * tpl > max_threads) return DFFT_ERR_UNSUPPORTED_SIZE;
        a.pad_shift = 30;
        a.TC = tc;
        a.lds_pitch = tc | 1;
        pp.threads = tpl * tc;
        pp.lds = lds_of(tc);
        pp.nblocks = (a.nlines + tc - 1) / tc;
    } else if (a.contig) {
        a.pad_shift = (tune && tune->pad_shift > 0) ? tune->pad_shift : (pp.prec ? 3 : 4);
        a.lds_pitch = rows + (rows >> a.pad_shift) + 1;
        // 128-thread workgroups by default (64 for the split schedule): many
        // small tiles in flight per CU
        const int wg = contig_threads > 256 ? contig_threads : a.variant == 2 ? 64 : 128;
        int tc = (tune && tune->tc > 0) ? tune->tc : (conv_tune && conv_tune->tc > 0) ? conv_tune->tc
                                                                                     : std::max(1, wg / tpl);
        if (a.blue && tpl < 512 && !(conv_tune && conv_tune->tc > 0)) {
            // Bluestein lines have odd thread counts (105 for M = 525): take
            // the line count in 128..512 threads that leaves the fewest idle
            // lanes in the last wavefront
            double util = 0;
            for (int c = std::max(1, (128 + tpl - 1) / tpl); c * tpl <= 512; ++c) {
                const int th = c * tpl;
                const double u = (double)th / (double)(((th + 63) / 64) * 64);
                if (u > util + 1e-9) {
                    util = u;
                    tc = c;
                }
            }
        }
        while (tc > 1 && (int64_t)tc * tpl > max_threads) --tc;
        const int contig_lds = contig_threads > 256 ? lds_limit : std::min(lds_limit, 65536);
        while (tc > 1 && (size_t)tc * a.lds_pitch * lds_es > (size_t)contig_lds) --tc;
        if ((size_t)a.lds_pitch * lds_es > (size_t)lds_limit) return DFFT_ERR_UNSUPPORTED_SIZE;
        a.TC = tc;
        pp.threads = tpl * tc;
        pp.lds = (size_t)tc * a.lds_pitch * lds_es;
        pp.nblocks = (a.nlines + tc - 1) / tc;
    } else {
        a.pad_shift = 30;
        // 256-byte (FP64) / 512-byte rows; the split schedule doubles the row over outer slices
        int tc = (tune && tune->tc > 0) ? tune->tc : (conv_tune && conv_tune->tc > 0) ? conv_tune->tc
                 : (a.variant >= 2 || pp.prec) ? 16 : 32;
        // convolution tiles: the widest row segments in up to 512 threads
        // (measured, 257-point Rader FP64: 16 -> 32 columns 2.98 -> 3.52
        // TB/s; 800-thread tiles are slower again)
        if (a.blue && !(conv_tune && conv_tune->tc > 0)) {
            tc = 32;
            while (tc > 8 && tc * tpl > 512) tc /= 2;
        }
        int64_t cols = a.nlines;
        while (tc > 1 && tc / 2 >= cols) tc /= 2;
        while (tc > 1 && (int64_t)tc * tpl > max_threads) tc /= 2;
        while (tc > 1 && (size_t)rows * tc * lds_es > (size_t)lds_limit) tc /= 2;
        // Bluestein tiles are issue-bound: two per CU (66 KB of LDS each)
        // beat one wider tile, down to 8 columns (measured: the 1024-point
        // convolution of a 509-point line 2.18 -> 2.06 ms)
        while (a.blue && tc > 8 && (size_t)rows * tc * lds_es > 67584) tc /= 2;
        if ((size_t)rows * tc * lds_es > (size_t)lds_limit || (int64_t)tc * tpl > max_threads)
            return DFFT_ERR_UNSUPPORTED_SIZE;
        a.TC = tc;
        a.lds_pitch = tc;
        pp.threads = tpl * tc;
        pp.lds = (size_t)rows * tc * lds_es;
        // tiles run over the columns of all outer slices back to back
        pp.nblocks = (a.nlines * a.nouter + tc - 1) / tc;
    }
    if (pp.nblocks > 0x7fffffffLL) return DFFT_ERR_UNSUPPORTED_SIZE;
    a.swz = 1;
    if (!tune) tune = conv_tune;
    if (tune && tune->swz > 1 && std::gcd<int64_t, int64_t>(tune->swz, pp.nblocks) == 1) a.swz = tune->swz;
    if (tune && tune->swz == SWZ_XCD) a.swz = SWZ_XCD;
    // untuned strided passes whose rows are not 128-byte aligned: neighbouring
    // tiles share the cache lines at their row ends, so each XCD takes one
    // contiguous range of tiles and the shared lines stay in its L2
    // (measured, tuned: the strided passes of 127^3, 131^3, 251^3 and 257^3
    // 4-17% faster, and 128-point columns at a 1009-element stride 7.5%; on
    // aligned rows and contiguous lines the order is neutral or slower)
    if (!tune && !a.contig) {
        const int64_t eb = (int64_t)(pp.prec ? 8 : 4) * (a.op == OP_C2C ? 2 : 1);
        if ((a.in_sm * eb) % 128 != 0 || (a.out_sm * eb) % 128 != 0) a.swz = SWZ_XCD;
    }
    return DFFT_OK;
}

static void strides_of(const int64_t n[3], int64_t st[3]) {
    st[0] = 1;
    st[1] = n[0];
    st[2] = n[0] * n[1];
}

// geometry of a pass along `axis` from array nin (input type) to nout (output type)
void set_geometry(AxisPass &a, int axis, const int64_t nin[3], const int64_t nout[3]) {
    int64_t si[3], so[3];
    strides_of(nin, si);
    strides_of(nout, so);
    if (si[axis] == 1) {
        // lines contiguous and back to back
        a.contig = 1;
        a.nlines = (nin[0] * nin[1] * nin[2]) / nin[axis];
        a.nouter = 1;
        a.in_sm = 1;
        a.in_so = nin[axis];
        a.out_sm = 1;
        a.out_so = nout[axis];
    } else if (axis == 1) {
        a.contig = 0;
        a.nlines = nin[0];
        a.nouter = nin[2];
        a.in_sm = si[1];
        a.in_so = si[2];
        a.out_sm = so[1];
        a.out_so = so[2];
    } else {  // axis 2: the whole (i, j) plane is contiguous columns
        a.contig = 0;
        a.nlines = nin[0] * nin[1];
        a.nouter = 1;
        a.in_sm = si[2];
        a.in_so = 0;
        a.out_sm = so[2];
        a.out_so = 0;
    }
}

}  // namespace dfft

using namespace dfft;

// ---------------------------------------------------------------------------
// plan object
// ---------------------------------------------------------------------------
struct dfft_plan {
    int rank = 3;
    int64_t n[3] = {1, 1, 1};    // logical real-space size
    int kind = DFFT_C2C, sign = -1, prec = DFFT_F64, decomp = DFFT_PENCIL;
    int ndims = 0;
    int dims[3] = {0, 0, 0};
    int r2r[3] = {0, 0, 0};
    uint32_t flags = 0;
    size_t in_bytes = 0, out_bytes = 0, work_bytes = 0;
    std::vector<PassPlan> passes;
    void *d_tw = nullptr;
    size_t tw_bytes = 0;
    void *own_work = nullptr;
    int device = 0;
    // graph replay cache
    mutable std::mutex gmu;
    mutable hipGraphExec_t gexec = nullptr;
    mutable hipGraph_t graph = nullptr;
    mutable const void *g_in = nullptr;
    mutable void *g_out = nullptr, *g_work = nullptr;
    // Slab: plane-pipelined 2D stage + 1D stage in one persistent launch
    GroupExec *slab = nullptr;  // Slab software pipeline (null: one launch per pass)
};

static const char *op_name(int op) {
    switch (op) {
    case OP_C2C: return "C2C";
    case OP_R2C: return "R2C";
    case OP_R2C_FULL: return "R2C_FULL";
    case OP_C2R: return "C2R";
    case OP_C2R_FULL: return "C2R_FULL";
    case OP_DCT2: return "DCT2";
    case OP_DCT3: return "DCT3";
    case OP_DHT: return "DHT";
    case OP_DST2: return "DST2";
    case OP_DST3: return "DST3";
    case OP_DCT1: return "DCT1";
    case OP_DST1: return "DST1";
    case OP_DCT4: return "DCT4";
    case OP_DST4: return "DST4";
    case OP_DCT2H: return "DCT2H";
    case OP_DCT3H: return "DCT3H";
    }
    return "?";
}

static double r2r_logical(int kind, int64_t n) {
    switch (kind) {
    case DFFT_REDFT00: return 2.0 * (double)(n - 1);
    case DFFT_RODFT00: return 2.0 * (double)(n + 1);
    case DFFT_DHT: return (double)n;
    default: return 2.0 * (double)n;
    }
}

namespace dfft {

// Builds passes; twiddle tables are appended to `tab` (long double pairs),
// one block per distinct (L, op, radices) in this plan.
int build_plan(dfft_plan *p) {
    const int es_r = p->prec ? 8 : 4;
    const int es_c = 2 * es_r;
    const int64_t *n = p->n;
    int64_t nc[3] = {n[0], n[1], n[2]};
    const int a0 = p->dims[0];
    if (p->kind == DFFT_R2C || p->kind == DFFT_C2R) nc[a0] = n[a0] / 2 + 1;
    const int64_t N = n[0] * n[1] * n[2];
    const int64_t NC = nc[0] * nc[1] * nc[2];
    const bool inv = p->sign > 0;
    const bool norm = !(p->flags & DFFT_FLAG_UNNORMALIZED);

    std::vector<int> order;
    for (int i = 0; i < p->ndims; ++i) order.push_back(p->dims[i]);

    auto add = [&](int axis, int op, int L, int nlog, const int64_t *nin, const int64_t *nout, int src, int dst,
                   double scale, bool inverse, double bytes) {
        PassPlan pp{};
        pp.prec = p->prec;
        pp.axis = axis;
        pp.src = src;
        pp.dst = dst;
        pp.alg_bytes = bytes;
        AxisPass &a = pp.ap;
        a.op = op;
        a.inverse = inverse ? 1 : 0;
        a.scale = scale;
        a.nlog = nlog;
        a.sp.L = L;
        a.nt = 1;  // every element is read once and written once per pass
        set_geometry(a, axis, nin, nout);
        a.paired_loads = paired_loads_for(a.in_sm, es_c);
        p->passes.push_back(pp);
    };

    if (p->kind == DFFT_C2C) {
        p->in_bytes = p->out_bytes = (size_t)N * es_c;
        if (inv) std::reverse(order.begin(), order.end());
        // The 1D transforms along distinct axes commute.  The first pass is the
        // out-of-place one, and out-of-place is cheapest on the contiguous axis,
        // so it goes first unless the reference's stage order is requested.
        // The Slab plane pipeline keeps the stage order: it chunks the two
        // in-plane passes over z-planes and needs the dims[2] pass at one end.
        const bool plane_pipeline = p->decomp == DFFT_SLAB && (p->flags & DFFT_FLAG_PLANE_PIPELINE);
        if (!(p->flags & DFFT_FLAG_STAGE_ORDER) && !plane_pipeline) {
            auto it = std::find(order.begin(), order.end(), 0);
            if (it != order.end()) std::rotate(order.begin(), it, it + 1);
        }
        for (size_t i = 0; i < order.size(); ++i) {
            const int ax = order[i];
            const int L = (int)n[ax];
            add(ax, OP_C2C, L, L, n, n, i == 0 ? BUF_IN : BUF_OUT, BUF_OUT, (inv && norm) ? 1.0 / L : 1.0, inv,
                2.0 * N * es_c);
        }
    } else if (p->kind == DFFT_R2C) {
        p->in_bytes = (size_t)N * es_r;
        p->out_bytes = (size_t)NC * es_c;
        const int L = (int)n[a0];
        if (L % 2 == 0) add(a0, OP_R2C, L / 2, L, n, nc, BUF_IN, BUF_OUT, 1.0, false, (double)N * es_r + (double)NC * es_c);
        else add(a0, OP_R2C_FULL, L, L, n, nc, BUF_IN, BUF_OUT, 1.0, false, (double)N * es_r + (double)NC * es_c);
        for (size_t i = 1; i < order.size(); ++i)
            add(order[i], OP_C2C, (int)nc[order[i]], (int)nc[order[i]], nc, nc, BUF_OUT, BUF_OUT, 1.0, false,
                2.0 * NC * es_c);
    } else if (p->kind == DFFT_C2R) {
        p->in_bytes = (size_t)NC * es_c;
        p->out_bytes = (size_t)N * es_r;
        const bool destroy = (p->flags & DFFT_FLAG_DESTROY_INPUT) != 0;
        const int mid = destroy ? BUF_IN : BUF_WORK;
        int cur = BUF_IN;
        for (size_t i = order.size(); i-- > 1;) {
            const int ax = order[i];
            const int L = (int)nc[ax];
            add(ax, OP_C2C, L, L, nc, nc, cur, mid, norm ? 1.0 / L : 1.0, true, 2.0 * NC * es_c);
            cur = mid;
        }
        if (cur == BUF_WORK) p->work_bytes = (size_t)NC * es_c;
        const int L = (int)n[a0];
        const double sc = norm ? 1.0 / L : 1.0;
        if (L % 2 == 0) add(a0, OP_C2R, L / 2, L, nc, n, cur, BUF_OUT, sc, true, (double)NC * es_c + (double)N * es_r);
        else add(a0, OP_C2R_FULL, L, L, nc, n, cur, BUF_OUT, sc, true, (double)NC * es_c + (double)N * es_r);
    } else {  // R2R
        p->in_bytes = p->out_bytes = (size_t)N * es_r;
        std::vector<int> kinds(p->r2r, p->r2r + p->ndims);
        if (inv) {
            std::reverse(order.begin(), order.end());
            std::reverse(kinds.begin(), kinds.end());
        }
        for (size_t i = 0; i < order.size(); ++i) {
            const int ax = order[i];
            const int nl = (int)n[ax];
            int op, L;  // fused op and the length of its complex FFT
            switch (kinds[i]) {
            case DFFT_REDFT10: op = OP_DCT2; L = nl; break;
            case DFFT_REDFT01: op = OP_DCT3; L = nl; break;
            case DFFT_RODFT10: op = OP_DST2; L = nl; break;
            case DFFT_RODFT01: op = OP_DST3; L = nl; break;
            case DFFT_DHT: op = OP_DHT; L = nl; break;
            case DFFT_REDFT11: op = OP_DCT4; L = 2 * nl; break;
            case DFFT_RODFT11: op = OP_DST4; L = 2 * nl; break;
            case DFFT_REDFT00:
                if (nl < 2) return DFFT_ERR_ARG;  // FFTW's REDFT00 needs n > 1
                op = OP_DCT1;
                L = 2 * (nl - 1);
                break;
            case DFFT_RODFT00: op = OP_DST1; L = 2 * (nl + 1); break;
            default: return DFFT_ERR_UNSUPPORTED;  // R2HC / HC2R: not in the reference's R2R kinds
            }
            const double sc = (p->flags & DFFT_FLAG_NORMALIZE_R2R) ? 1.0 / r2r_logical(kinds[i], nl) : 1.0;
            add(ax, op, L, nl, n, n, i == 0 ? BUF_IN : BUF_OUT, BUF_OUT, sc, false, 2.0 * N * es_r);
            // REDFT10 / REDFT01 of n = 4k reals through an n/2-point complex FFT
            // where the catalog has it: half the FFT work per line, half the
            // threads per line.  Measured on gfx950 (Float64 passes of
            // 256^3, 512x512x64, 1024x256x128, 64x64x1024): faster for REDFT10
            // from n = 256 (strided 256-point pass 0.148 -> 0.102 ms) and
            // for REDFT01 from n = 512; slower below (128^3 REDFT10 passes
            // 10.4-10.7 -> 11.6-11.9 us: the lines' threads no longer hide
            // the latency), so shorter lines keep the n-point FFT, and so do
            // arrays under 4M elements (256^2 REDFT10: 16 workgroups per
            // launch, 7.0 -> 8.3 us).
            const bool half_ok = (p->flags & DFFT_FLAG_DCT_HALF_LENGTH) ||
                                 (!(p->flags & DFFT_FLAG_DCT_FULL_LENGTH) && nl >= (op == OP_DCT2 ? 256 : 512) &&
                                  N >= (1LL << 22));
            if ((op == OP_DCT2 || op == OP_DCT3) && nl % 4 == 0 && half_ok) {
                AxisPass &a = p->passes.back().ap;
                const int hop = op == OP_DCT2 ? OP_DCT2H : OP_DCT3H;
                std::vector<int> R;
                int tpl = 0;
                if (fixed_schedule(p->prec, a.contig, hop == OP_DCT2H ? 3 : 2, hop, nl / 2, R, tpl)) {
                    a.op = hop;
                    a.sp.L = nl / 2;
                }
            }
        }
    }

    // schedules + twiddle tables (deduplicated per (op, L, radices))
    TwiddleTable tt;
    tt.prec = p->prec;
    for (auto &pp : p->passes) {
        const Tuning *tune = nullptr;
        Tuning t;
        if (!(p->flags & DFFT_FLAG_NO_TUNE) && tuning_lookup(pp, t)) tune = &t;
        int rc = schedule_pass(pp, tune);
        if (rc != DFFT_OK) return rc;
        tt.add(pp.ap);
    }
    // Slab: the 2D stage (dims[0], dims[1]) is software-pipelined over chunks
    // of z-planes; the dims[2] pass follows as one launch
    const bool slab_ok = p->decomp == DFFT_SLAB && (p->flags & DFFT_FLAG_PLANE_PIPELINE) && p->rank == 3 &&
                         p->ndims == 3 && n[2] > 1 &&
                         p->dims[2] == 2 && p->kind != DFFT_R2R && !(p->flags & DFFT_FLAG_GRAPH) &&
                         std::all_of(p->passes.begin(), p->passes.end(), [](const PassPlan &q) {
                             // ops the grouped kernel dispatches (it has no Bluestein tile)
                             return (q.ap.op == OP_C2C || q.ap.op == OP_R2C || q.ap.op == OP_C2R) && !q.ap.blue;
                         });
    // chunk: planes whose dims[0]-pass output is about `chunk_bytes`
    int64_t chunk_bytes = 32LL << 20;
    if (const char *e = getenv("DFFT_SLAB_CHUNK_BYTES")) chunk_bytes = std::max(1LL, atoll(e));
    struct ChunkPass {
        PassPlan pp;
        int64_t k0, planes;
    };
    std::vector<ChunkPass> cpa, cpb;  // chunk passes of the first / second 2D-stage pass
    PassPlan line;                    // the dims[2] pass, scheduled for the grouped kernel
    if (slab_ok) {
        const bool inv = p->sign > 0;
        const int ia = inv ? 1 : 0, ib = inv ? 2 : 1;
        // The line pass must be the z (axis 2) pass at the pipeline's end that
        // reads (inverse) or writes (forward) the whole array, and the two
        // chunked passes must stay inside z-planes; anything else would make a
        // chunk address planes outside its own.
        if (p->passes.size() != 3 || p->passes[inv ? 0 : 2].axis != 2 || p->passes[ia].axis == 2 ||
            p->passes[ib].axis == 2)
            return DFFT_ERR_UNSUPPORTED;
        line = p->passes[inv ? 0 : 2];
        {
            const Tuning *tune = nullptr;
            Tuning t;
            if (!(p->flags & DFFT_FLAG_NO_TUNE) && tuning_lookup(line, t)) tune = &t;
            int rc2 = schedule_pass(line, tune, 1024, 0, 256, true);
            if (rc2 != DFFT_OK) return rc2;
            tt.add(line.ap);
        }
        auto sizes = [&](const PassPlan &q, int64_t *nin, int64_t *nout) {
            const bool in_real = q.ap.op == OP_R2C || q.ap.op == OP_R2C_FULL;
            const bool out_real = q.ap.op == OP_C2R || q.ap.op == OP_C2R_FULL;
            for (int d = 0; d < 3; ++d) {
                nin[d] = in_real ? n[d] : nc[d];
                nout[d] = out_real ? n[d] : nc[d];
            }
        };
        int64_t ain[3], aout[3];
        sizes(p->passes[ia], ain, aout);
        const int64_t plane_out = aout[0] * aout[1] * ((p->passes[ia].ap.op == OP_C2R) ? es_r : es_c);
        const int64_t K = std::max<int64_t>(1, std::min<int64_t>(n[2], chunk_bytes / std::max<int64_t>(1, plane_out)));
        for (int which = 0; which < 2; ++which) {
            const PassPlan &full = p->passes[which == 0 ? ia : ib];
            int64_t nin[3], nout[3];
            sizes(full, nin, nout);
            for (int64_t k0 = 0; k0 < n[2]; k0 += K) {
                const int64_t planes = std::min(K, n[2] - k0);
                PassPlan q = full;
                const int64_t pin[3] = {nin[0], nin[1], planes}, pout[3] = {nout[0], nout[1], planes};
                set_geometry(q.ap, full.axis, pin, pout);
                const Tuning *tune = nullptr;
                Tuning t;
                if (!(p->flags & DFFT_FLAG_NO_TUNE) && tuning_lookup(q, t)) tune = &t;
                int rc2 = schedule_pass(q, tune, 1024, 0, 1024, true);
                if (rc2 != DFFT_OK) return rc2;
                q.alg_bytes = full.alg_bytes * (double)planes / (double)n[2];
                q.ap.nt = 0;  // the next launch re-reads this chunk from the Infinity Cache
                tt.add(q.ap);
                (which == 0 ? cpa : cpb).push_back(ChunkPass{q, k0, planes});
            }
        }
    }
    int rc = tt.upload(&p->d_tw, &p->tw_bytes);
    if (rc != DFFT_OK) return rc;
    for (auto &pp : p->passes) pp.ap.tw = p->d_tw;
    if (slab_ok) {
        const bool inv = p->sign > 0;
        auto plane_bytes = [&](const PassPlan &q, bool input) -> int64_t {
            const bool real = input ? (q.ap.op == OP_R2C || q.ap.op == OP_R2C_FULL)
                                    : (q.ap.op == OP_C2R || q.ap.op == OP_C2R_FULL);
            const int64_t *sz = real ? n : nc;
            return sz[0] * sz[1] * (real ? es_r : es_c);
        };
        auto item = [&](ChunkPass &c) {
            c.pp.ap.tw = p->d_tw;
            return make_group_item(c.pp, c.pp.src, c.k0 * plane_bytes(c.pp, true), c.pp.dst,
                                   c.k0 * plane_bytes(c.pp, false));
        };
        std::vector<std::vector<GroupItem>> levels;
        line.ap.tw = p->d_tw;
        if (inv) levels.push_back({make_group_item(line, line.src, 0, line.dst, 0)});
        const size_t C = cpa.size();
        for (size_t i = 0; i <= C; ++i) {
            std::vector<GroupItem> lv;
            if (i < C) lv.push_back(item(cpa[i]));
            if (i >= 1) lv.push_back(item(cpb[i - 1]));
            levels.push_back(lv);
        }
        if (!inv) levels.push_back({make_group_item(line, line.src, 0, line.dst, 0)});
        rc = group_create(&p->slab, levels);
        if (rc != DFFT_OK) return rc;
    }
    return DFFT_OK;
}

}  // namespace dfft

namespace dfft {
int plan_pass_count(const dfft_plan *p) { return (int)p->passes.size(); }
PassPlan plan_pass(const dfft_plan *p, int i) { return p->passes[i]; }
size_t plan_buffer_bytes(const dfft_plan *p, int buf) {
    return buf == BUF_IN ? p->in_bytes : buf == BUF_OUT ? p->out_bytes : p->work_bytes;
}
// device memory the plan itself holds (twiddle tables, its own workspace if allocated)
size_t plan_resident_bytes(const dfft_plan *p) { return p->tw_bytes + (p->own_work ? p->work_bytes : 0); }
}  // namespace dfft

static void free_plan(dfft_plan *p) {
    if (!p) return;
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    group_destroy(p->slab);
    if (p->d_tw) (void)hipFree(p->d_tw);
    if (p->own_work) (void)hipFree(p->own_work);
    delete p;
}

static void *resolve(int buf, const void *in, void *out, void *work) {
    switch (buf) {
    case BUF_IN: return const_cast<void *>(in);
    case BUF_OUT: return out;
    default: return work;
    }
}

// Trace ranges around each pass's launch, named as the reference's NVTX
// ranges around its stages ("DIM 1", "DIM 2", "DIM 3", src/fft.jl:654-686):
// the Julia dim (axis + 1) the pass transforms.  They bracket the enqueue on
// the host; rocprofv3 --marker-trace shows them beside the kernels.
static const char *const kDimRange[3] = {"DIM 1", "DIM 2", "DIM 3"};

static int run_passes(const dfft_plan *p, const void *d_in, void *d_out, void *work, hipStream_t st) {
    for (const auto &pp : p->passes) {
        AxisPass a = pp.ap;
        a.in = resolve(pp.src, d_in, d_out, work);
        a.out = resolve(pp.dst, d_in, d_out, work);
        roctxRangePush(kDimRange[pp.axis]);
        hipError_t e = launch_axis(a, pp.prec, pp.nblocks, pp.threads, pp.lds, st);
        roctxRangePop();
        if (e) return hip_fail(e, "axis pass launch");
    }
    return DFFT_OK;
}

static int run_graph(const dfft_plan *p, const void *d_in, void *d_out, void *work, hipStream_t st) {
    std::lock_guard<std::mutex> g(p->gmu);
    if (!p->gexec || p->g_in != d_in || p->g_out != d_out || p->g_work != work) {
        if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
        if (p->graph) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
        hipGraph_t graph;
        if (hipGraphCreate(&graph, 0) != hipSuccess) return DFFT_ERR_HIP;
        hipGraphNode_t prev = nullptr;
        std::vector<AxisPass> args(p->passes.size());
        for (size_t i = 0; i < p->passes.size(); ++i) {
            const auto &pp = p->passes[i];
            args[i] = pp.ap;
            args[i].in = resolve(pp.src, d_in, d_out, work);
            args[i].out = resolve(pp.dst, d_in, d_out, work);
            if (pp.lds > 65536)
                (void)hipFuncSetAttribute(axis_kernel_fn(args[i], pp.prec), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)pp.lds);
            void *kargs[] = {&args[i]};
            hipKernelNodeParams kp{};
            kp.func = axis_kernel_fn(args[i], pp.prec);
            kp.gridDim = dim3((unsigned)pp.nblocks);
            kp.blockDim = dim3(pp.threads);
            kp.sharedMemBytes = (unsigned)pp.lds;
            kp.kernelParams = kargs;
            hipGraphNode_t node;
            if (hipGraphAddKernelNode(&node, graph, prev ? &prev : nullptr, prev ? 1 : 0, &kp) != hipSuccess) {
                (void)hipGraphDestroy(graph);
                return DFFT_ERR_HIP;
            }
            prev = node;
        }
        hipGraphExec_t exec;
        if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return DFFT_ERR_HIP;
        }
        p->graph = graph;
        p->gexec = exec;
        p->g_in = d_in;
        p->g_out = d_out;
        p->g_work = work;
    }
    return hipGraphLaunch(p->gexec, st) == hipSuccess ? DFFT_OK : DFFT_ERR_HIP;
}

namespace dfft {
int plan_check_io(const dfft_plan *p, const void *d_in, const void *d_out) {
    if (!p || !d_in || !d_out) return DFFT_ERR_ARG;
    if (d_in == d_out && (p->kind == DFFT_R2C || p->kind == DFFT_C2R)) return DFFT_ERR_ARG;
    return DFFT_OK;
}

size_t plan_work_bytes(const dfft_plan *p) { return p->work_bytes; }

int plan_device(const dfft_plan *p) { return p->device; }

int plan_enqueue(const dfft_plan *p, const void *d_in, void *d_out, void *work, void *stream, bool graph) {
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (p->slab) {
        void *slots[3] = {const_cast<void *>(d_in), d_out, work};
        roctxRangePush("DIM 1+2 (plane pipeline), DIM 3");
        const int rc = group_launch(p->slab, slots, 3, st);
        roctxRangePop();
        return rc;
    }
    if (graph && (p->flags & DFFT_FLAG_GRAPH)) {
        roctxRangePush("dfft plan graph");
        const int rc = run_graph(p, d_in, d_out, work, st);
        roctxRangePop();
        return rc;
    }
    return run_passes(p, d_in, d_out, work, st);
}
}  // namespace dfft

extern "C" {

int dfft_plan_create(dfft_plan **plan, int rank, const int64_t *n, int kind, int sign, const int *r2r_kinds,
                     int precision, int decomp, const int *dims, int ndims, uint32_t flags) {
    if (!plan || !n || !dims) return DFFT_ERR_ARG;
    *plan = nullptr;
    if (rank < 1 || rank > 3 || ndims < 1 || ndims > rank) return DFFT_ERR_ARG;
    if (kind < DFFT_C2C || kind > DFFT_R2R) return DFFT_ERR_ARG;
    if (sign != -1 && sign != 1) return DFFT_ERR_ARG;
    if (precision != DFFT_F32 && precision != DFFT_F64) return DFFT_ERR_ARG;
    if (decomp != DFFT_PENCIL && decomp != DFFT_SLAB) return DFFT_ERR_ARG;
    if (kind == DFFT_R2C && sign != -1) return DFFT_ERR_ARG;
    if (kind == DFFT_C2R && sign != 1) return DFFT_ERR_ARG;
    if (kind == DFFT_R2R && !r2r_kinds) return DFFT_ERR_ARG;
    dfft_plan *p = nullptr;
    try {
        p = new dfft_plan();
    } catch (...) {
        return DFFT_ERR_OOM;
    }
    p->rank = rank;
    for (int d = 0; d < rank; ++d) {
        // every length must fit an int (line lengths are int on the device)
        if (n[d] < 1 || n[d] >= (1LL << 31)) { delete p; return DFFT_ERR_ARG; }
        p->n[d] = n[d];
    }
    p->kind = kind;
    p->sign = sign;
    p->prec = precision;
    p->decomp = decomp;
    p->flags = flags;
    p->ndims = ndims;
    bool used[3] = {false, false, false};
    for (int i = 0; i < ndims; ++i) {
        if (dims[i] < 0 || dims[i] >= rank || used[dims[i]]) { delete p; return DFFT_ERR_ARG; }
        used[dims[i]] = true;
        p->dims[i] = dims[i];
        if (kind == DFFT_R2R) {
            if (r2r_kinds[i] < DFFT_R2HC || r2r_kinds[i] > DFFT_RODFT11) { delete p; return DFFT_ERR_ARG; }
            p->r2r[i] = r2r_kinds[i];
        }
    }
    const int64_t total = p->n[0] * p->n[1] * p->n[2];
    if (total > (1LL << 40)) { delete p; return DFFT_ERR_ARG; }
    if (hipError_t e = hipGetDevice(&p->device)) {
        delete p;
        return hip_fail(e, "hipGetDevice");
    }
    if (!(flags & DFFT_FLAG_NO_TUNE)) wisdom_autoload();
    if ((flags & DFFT_FLAG_MEASURE) && !(flags & DFFT_FLAG_NO_TUNE)) {
        const int mrc = measure_plan(rank, n, kind, sign, r2r_kinds, precision, decomp, dims, ndims, flags);
        if (mrc != DFFT_OK) {
            delete p;
            return mrc;
        }
    }
    int rc;
    try {
        rc = build_plan(p);
    } catch (const std::bad_alloc &) {
        rc = DFFT_ERR_OOM;
    } catch (...) {
        rc = DFFT_ERR_ARG;
    }
    if (rc != DFFT_OK) {
        free_plan(p);
        return rc;
    }
    *plan = p;
    return DFFT_OK;
}

int dfft_plan_workspace_size(const dfft_plan *p, size_t *bytes) {
    if (!p || !bytes) return DFFT_ERR_ARG;
    *bytes = p->work_bytes;
    return DFFT_OK;
}

int dfft_plan_io_size(const dfft_plan *p, size_t *in_bytes, size_t *out_bytes) {
    if (!p) return DFFT_ERR_ARG;
    if (in_bytes) *in_bytes = p->in_bytes;
    if (out_bytes) *out_bytes = p->out_bytes;
    return DFFT_OK;
}

int dfft_execute(const dfft_plan *p, const void *d_in, void *d_out, void *d_work, void *stream) {
    if (int rc = plan_check_io(p, d_in, d_out)) return rc;
    void *work = d_work;
    if (p->work_bytes && !work) {
        dfft_plan *mp = const_cast<dfft_plan *>(p);
        std::lock_guard<std::mutex> g(mp->gmu);
        if (!mp->own_work && hipMalloc(&mp->own_work, p->work_bytes) != hipSuccess) {
            mp->own_work = nullptr;
            return DFFT_ERR_OOM;
        }
        work = mp->own_work;
    }
    return plan_enqueue(p, d_in, d_out, work, stream, true);
}

int dfft_execute_pass(const dfft_plan *p, int pass, const void *d_in, void *d_out, void *d_work, void *stream) {
    if (!p || !d_in || !d_out || pass < 0 || pass >= (int)p->passes.size()) return DFFT_ERR_ARG;
    if (d_in == d_out && (p->kind == DFFT_R2C || p->kind == DFFT_C2R)) return DFFT_ERR_ARG;
    void *work = d_work;
    if (p->work_bytes && !work) {
        dfft_plan *mp = const_cast<dfft_plan *>(p);
        std::lock_guard<std::mutex> g(mp->gmu);
        if (!mp->own_work && hipMalloc(&mp->own_work, p->work_bytes) != hipSuccess) {
            mp->own_work = nullptr;
            return DFFT_ERR_OOM;
        }
        work = mp->own_work;
    }
    const PassPlan &pp = p->passes[pass];
    AxisPass a = pp.ap;
    a.in = resolve(pp.src, d_in, d_out, work);
    a.out = resolve(pp.dst, d_in, d_out, work);
    roctxRangePush(kDimRange[pp.axis]);
    hipError_t e = launch_axis(a, pp.prec, pp.nblocks, pp.threads, pp.lds, (hipStream_t)stream);
    roctxRangePop();
    return e == hipSuccess ? DFFT_OK : hip_fail(e, "axis pass launch");
}

int dfft_plan_export_wisdom(const dfft_plan *p, const char *path) {
    if (!p || !path) return DFFT_ERR_ARG;
    std::vector<std::pair<std::string, Tuning>> entries;
    for (const auto &pp : p->passes) {
        const AxisPass &a = pp.ap;
        Tuning t;
        // a Bluestein pass keeps only its engine (variant 0: runtime, else the
        // compile-time default); every other pass its radix list and tile shape
        if (!a.blue)
            for (int s = 0; s < a.sp.S; ++s) t.radices.push_back(a.sp.R[s]);
        t.tpl = a.variant ? 0 : a.TPL;
        t.tc = a.TC;
        t.pad_shift = a.contig ? a.pad_shift : 0;
        // variant of a Bluestein pass: 0 runtime engine, 3 compile-time
        // Bluestein, 2 Rader (or the compile-time default where no Rader kernel exists)
        t.variant = a.blue == 1 ? (a.variant ? 3 : 0) : a.variant;
        t.swz = a.swz;
        const std::string key = tuning_key(pp);
        if (std::none_of(entries.begin(), entries.end(), [&](const auto &e) { return e.first == key; }))
            entries.emplace_back(key, t);
    }
    return wisdom_write(path, entries);
}

int dfft_plan_destroy(dfft_plan *p) {
    if (!p) return DFFT_ERR_ARG;
    free_plan(p);
    return DFFT_OK;
}

int dfft_plan_num_passes(const dfft_plan *p, int *npasses) {
    if (!p || !npasses) return DFFT_ERR_ARG;
    *npasses = (int)p->passes.size();
    return DFFT_OK;
}

int dfft_plan_describe(const dfft_plan *p, int pass, char *buf, size_t buflen) {
    if (!p || !buf || pass < 0 || pass >= (int)p->passes.size()) return DFFT_ERR_ARG;
    const PassPlan &pp = p->passes[pass];
    const AxisPass &a = pp.ap;
    // radices of the FFT the tile runs (a Bluestein pass: of its convolution length)
    const StagePlan &fp = a.blue ? a.bsp : a.sp;
    std::string rad, rl;
    for (int s = 0; s < fp.S; ++s) {
        rad += (s ? "x" : "") + std::to_string(fp.R[s]);
        rl += (s ? "," : "") + std::to_string(fp.R[s]);
    }
    if (rad.empty()) rad = "1";
    // the kernel's name as profilers print it (spaces removed)
    const char *T = pp.prec ? "double" : "float";
    std::string kname;
    if (a.blue == 2)
        kname = std::string("k_rader_fixed<") + T + "," + (a.contig ? "true" : "false") + ",true,dfft::Radices<" + rl +
                ">>";
    else if (a.blue && a.variant)
        kname = std::string("k_blue_fixed<") + T + "," + std::to_string(a.op) + "," + (a.contig ? "true" : "false") +
                "," + (a.variant == 2 ? "true" : "false") + ",dfft::Radices<" + rl + ">>";
    else if (a.blue)
        kname = std::string("k_axis_blue<") + T + "," + std::to_string(a.op) + "," + (a.contig ? "true" : "false") + ">";
    else if (a.variant >= 1)
        kname = std::string("k_axis_fixed<") + T + "," + std::to_string(a.op) + "," + (a.contig ? "true" : "false") +
                "," + (a.variant == 2 ? "16,true" : a.variant == 3 ? "16,false" : "8,false") + ",dfft::Radices<" + rl +
                ">," + std::to_string(a.TPL) + ">";
    else
        kname = std::string("k_axis<") + T + "," + std::to_string(a.op) + "," + (a.contig ? "true" : "false") + ",8>";
    snprintf(buf, buflen,
             "{\"pass\":%d,\"axis\":%d,\"op\":\"%s\",\"L\":%d,\"nlog\":%d,\"radices\":\"%s\",\"TPL\":%d,\"TC\":%d,"
             "\"contig\":%d,\"variant\":%d,\"threads\":%d,\"blocks\":%lld,\"lds\":%zu,\"bytes\":%.0f,\"prec\":\"%s\","
             "\"swz\":%d,\"bluestein\":%d,\"rader\":%d,\"paired_loads\":%d,\"kernel\":\"%s\"}",
             pass, pp.axis, op_name(a.op), a.sp.L, a.nlog, rad.c_str(), a.TPL, a.TC, a.contig, a.variant, pp.threads,
             (long long)pp.nblocks, pp.lds, pp.alg_bytes, pp.prec ? "f64" : "f32", a.swz, a.blue == 1 ? a.bsp.L : 0,
             a.blue == 2 ? a.bsp.L : 0, a.paired_loads, kname.c_str());
    return DFFT_OK;
}

int dfft_plan_schedule(const dfft_plan *p, int *schedule, int *launches) {
    if (!p || !schedule) return DFFT_ERR_ARG;
    *schedule = p->slab ? DFFT_SCHEDULE_PLANE_PIPELINE : DFFT_SCHEDULE_PASSES;
    if (launches) *launches = p->slab ? group_num_launches(p->slab) : (int)p->passes.size();
    return DFFT_OK;
}

int dfft_plan_algorithmic_bytes(const dfft_plan *p, double *bytes) {
    if (!p || !bytes) return DFFT_ERR_ARG;
    double b = 0;
    for (const auto &pp : p->passes) b += pp.alg_bytes;
    *bytes = b;
    return DFFT_OK;
}

const char *dfft_last_error(void) { return g_last_error; }

const char *dfft_strerror(int s) {
    switch (s) {
    case DFFT_OK: return "success";
    case DFFT_ERR_ARG: return "invalid argument";
    case DFFT_ERR_UNSUPPORTED_SIZE: return "unsupported transform length (prime factor > 7 or too long)";
    case DFFT_ERR_HIP: return "HIP runtime error";
    case DFFT_ERR_OOM: return "device out of memory";
    case DFFT_ERR_BUDGET: return "request exceeds the HBM budget";
    case DFFT_ERR_STATE: return "object used in the wrong state";
    case DFFT_ERR_NOT_READY: return "request not finished";
    case DFFT_ERR_UNSUPPORTED: return "transform kind not implemented on the GPU path";
    case DFFT_ERR_SCRATCH: return "the plan's workspace exceeds the static scratch buffer (DFFT_STATIC_SCRATCH_BYTES)";
    case DFFT_ERR_WISDOM: return "wisdom file stamped for another HIP build or GPU architecture";
    }
    return "unknown status";
}

int dfft_version(int *major, int *minor) {
    if (major) *major = 0;
    if (minor) *minor = 1;
    return DFFT_OK;
}

static bool esize_ok(int e) { return e == 4 || e == 8 || e == 16; }

int dfft_box_pack(void *d_buf, const void *d_src, const int64_t *src_shape, const int64_t *lo, const int64_t *cnt,
                  int64_t offset, int esize, void *stream) {
    if (!d_buf || !d_src || !src_shape || !lo || !cnt || !esize_ok(esize) || offset < 0) return DFFT_ERR_ARG;
    for (int d = 0; d < 3; ++d)
        if (lo[d] < 0 || cnt[d] < 0 || lo[d] + cnt[d] > src_shape[d]) return DFFT_ERR_ARG;
    const int64_t s1 = src_shape[0], s2 = src_shape[0] * src_shape[1];
    hipError_t e = launch_box(d_src, lo[0] + lo[1] * s1 + lo[2] * s2, s1, s2, d_buf, offset, cnt[0], cnt[0] * cnt[1],
                              cnt, esize, (hipStream_t)stream);
    return e == hipSuccess ? DFFT_OK : DFFT_ERR_HIP;
}

int dfft_box_unpack(void *d_dst, const void *d_buf, const int64_t *dst_shape, const int64_t *lo, const int64_t *cnt,
                    int64_t offset, int esize, void *stream) {
    if (!d_buf || !d_dst || !dst_shape || !lo || !cnt || !esize_ok(esize) || offset < 0) return DFFT_ERR_ARG;
    for (int d = 0; d < 3; ++d)
        if (lo[d] < 0 || cnt[d] < 0 || lo[d] + cnt[d] > dst_shape[d]) return DFFT_ERR_ARG;
    const int64_t s1 = dst_shape[0], s2 = dst_shape[0] * dst_shape[1];
    hipError_t e = launch_box(d_buf, offset, cnt[0], cnt[0] * cnt[1], d_dst, lo[0] + lo[1] * s1 + lo[2] * s2, s1, s2,
                              cnt, esize, (hipStream_t)stream);
    return e == hipSuccess ? DFFT_OK : DFFT_ERR_HIP;
}

int dfft_box_copy(void *d_dst, const int64_t *dst_shape, const int64_t *dlo, const void *d_src,
                  const int64_t *src_shape, const int64_t *slo, const int64_t *cnt, int esize, void *stream) {
    if (!d_dst || !d_src || !dst_shape || !src_shape || !dlo || !slo || !cnt || !esize_ok(esize)) return DFFT_ERR_ARG;
    for (int d = 0; d < 3; ++d) {
        if (cnt[d] < 0 || dlo[d] < 0 || slo[d] < 0) return DFFT_ERR_ARG;
        if (dlo[d] + cnt[d] > dst_shape[d] || slo[d] + cnt[d] > src_shape[d]) return DFFT_ERR_ARG;
    }
    const int64_t s1 = src_shape[0], s2 = src_shape[0] * src_shape[1];
    const int64_t d1 = dst_shape[0], d2 = dst_shape[0] * dst_shape[1];
    hipError_t e = launch_box(d_src, slo[0] + slo[1] * s1 + slo[2] * s2, s1, s2, d_dst,
                              dlo[0] + dlo[1] * d1 + dlo[2] * d2, d1, d2, cnt, esize, (hipStream_t)stream);
    return e == hipSuccess ? DFFT_OK : DFFT_ERR_HIP;
}

int dfft_permute(void *d_out, const void *d_in, const int64_t *in_shape, const int *perm, int esize, void *stream) {
    if (!d_out || !d_in || !in_shape || !perm || !esize_ok(esize)) return DFFT_ERR_ARG;
    bool seen[3] = {false, false, false};
    for (int d = 0; d < 3; ++d) {
        if (perm[d] < 0 || perm[d] > 2 || seen[perm[d]] || in_shape[d] < 1) return DFFT_ERR_ARG;
        seen[perm[d]] = true;
    }
    if (d_out == d_in) return DFFT_ERR_ARG;
    hipError_t e = launch_permute(d_in, d_out, in_shape, perm, esize, (hipStream_t)stream);
    return e == hipSuccess ? DFFT_OK : DFFT_ERR_HIP;
}

int dfft_device_copy(void *d_dst, const void *d_src, size_t bytes, void *stream) {
    if (!d_dst || !d_src || bytes % 16 || ((uintptr_t)d_dst | (uintptr_t)d_src) % 16) return DFFT_ERR_ARG;
    const uintptr_t a = (uintptr_t)d_src, b = (uintptr_t)d_dst;
    if (a < b + bytes && b < a + bytes) return DFFT_ERR_ARG;  // overlapping ranges
    hipError_t e = launch_copy16(d_src, d_dst, (int64_t)(bytes / 16), (hipStream_t)stream);
    return e == hipSuccess ? DFFT_OK : hip_fail(e, "copy launch");
}

int dfft_fill_uniform(void *d_out, int64_t nelem, int is_complex, int precision, uint64_t seed, void *stream) {
    if (!d_out || nelem < 0 || (precision != DFFT_F32 && precision != DFFT_F64)) return DFFT_ERR_ARG;
    hipError_t e = launch_fill(d_out, is_complex ? 2 * nelem : nelem, precision, seed, (hipStream_t)stream);
    return e == hipSuccess ? DFFT_OK : DFFT_ERR_HIP;
}

}  // extern "C"
