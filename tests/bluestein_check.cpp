// Host check of the planner's Bluestein and Rader choices (no GPU: scheduling
// reads only the catalogs).  For every length 2..2000 and the ops C2C / R2C /
// C2R along the contiguous axis (and C2C along a strided one), both
// precisions:
//   * a pass is a convolution pass (Bluestein or Rader) exactly when its FFT
//     length has a prime factor above 7;
//   * only a C2C pass of a prime length n with 7-smooth n-1 may be a Rader
//     pass, with M = n-1 (where its FFT has no more stages than Bluestein's);
//     every other convolution pass is a Bluestein pass with M >= 2n-1; the
//     radices multiply to M;
//   * the compile-time engine (variant >= 2) runs the reversed list in its
//     second FFT, and C2C always gets it;
//   * the 257-point C2C pass is a Rader pass over 256, the 263-point one a
//     Bluestein pass over 640 (3 stages of 16*8*5).
// Prints one line per failure and a summary; exit status 1 on any failure.
#include <algorithm>
#include <cstdio>
#include <vector>

#include "daggerfft_hip.h"
#include "dfft_plan.hpp"

using namespace dfft;

static bool smooth(int n) {
    for (int p : {2, 3, 5, 7})
        while (n % p == 0) n /= p;
    return n == 1;
}

static bool prime(int n) {
    if (n < 2) return false;
    for (int q = 2; q * q <= n; ++q)
        if (n % q == 0) return false;
    return true;
}

static int bad = 0, checked = 0, raders = 0;

static void fail(const char *what, int op, int n, int prec, int contig) {
    printf("FAIL %s: op %d n %d prec %d contig %d\n", what, op, n, prec, contig);
    ++bad;
}

static void check(int op, int n, int prec, int contig) {
    PassPlan pp{};
    pp.prec = prec;
    AxisPass &a = pp.ap;
    a.op = op;
    a.sp.L = (op == OP_R2C || op == OP_C2R) ? n / 2 : n;
    a.nlog = n;
    a.scale = 1.0;
    const int64_t shape[3] = {contig ? n : 3, contig ? 3 : n, 2};
    const int64_t half[3] = {contig ? n / 2 + 1 : 3, contig ? 3 : n / 2 + 1, 2};
    const bool real = op != OP_C2C;
    set_geometry(a, contig ? 0 : 1, op == OP_C2R ? half : shape, op == OP_R2C ? half : shape);
    a.nt = 1;
    const int rc = schedule_pass(pp, nullptr);
    ++checked;
    const int L = a.sp.L;
    if (rc != DFFT_OK) {
        // the host has no device: 64 KB of LDS, which long convolutions exceed
        if (smooth(L) || 2 * L - 1 <= 1024) fail("not scheduled", op, n, prec, contig);
        return;
    }
    if ((a.blue != 0) == smooth(L)) fail("convolution pass iff not 7-smooth", op, n, prec, contig);
    if (!a.blue) return;
    const bool rader_ok = op == OP_C2C && prime(L) && smooth(L - 1);
    if (a.blue == 2 && !rader_ok) fail("Rader only for C2C primes with 7-smooth n-1", op, n, prec, contig);
    const int M = a.bsp.L;
    int prod = 1;
    for (int s = 0; s < a.bsp.S; ++s) prod *= a.bsp.R[s];
    if (a.blue == 2) {
        ++raders;
        if (M != L - 1 || prod != M) fail("Rader convolution length", op, n, prec, contig);
    } else if (M < 2 * L - 1 || prod != M) {
        fail("convolution length", op, n, prec, contig);
    }
    if (a.variant >= 2) {
        std::vector<int> r(a.bsp.R, a.bsp.R + a.bsp.S), r2(a.bsp2.R, a.bsp2.R + a.bsp2.S);
        std::reverse(r.begin(), r.end());
        if (r != r2 || a.bsp2.L != M) fail("second FFT runs the reversed list", op, n, prec, contig);
    } else if (!real) {
        fail("C2C Bluestein on the compile-time engine", op, n, prec, contig);
    }
    if (op == OP_C2C && n == 257 && (a.blue != 2 || M != 256)) fail("257 -> Rader over 256", op, n, prec, contig);
    if (op == OP_C2C && n == 263 && (M != 640 || a.bsp.S != 3)) fail("263 -> 640", op, n, prec, contig);
}

int main() {
    for (int prec = 0; prec < 2; ++prec)
        for (int n = 2; n <= 2000; ++n) {
            check(OP_C2C, n, prec, 1);
            check(OP_C2C, n, prec, 0);
            if (n % 2 == 0) {
                check(OP_R2C, n, prec, 1);
                check(OP_C2R, n, prec, 1);
            }
        }
    printf("%d passes checked (%d Rader), %d failures\n", checked, raders, bad);
    return bad ? 1 : 0;
}
