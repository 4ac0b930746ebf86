// Host check of the planner against a schedule cache: every axis pass of the
// transforms read from stdin ("kind n0 n1 n2 precision" per line, kind 0 =
// C2C, 1 = R2C; dims (0,1,2)) is scheduled as a plan schedules it and as a
// grouped launch schedules it.  Prints one line per failure and a summary.
#include <cstdio>
#include <vector>

#include "daggerfft_hip.h"
#include "dfft_plan.hpp"

using namespace dfft;

static int check(PassPlan pp, bool grouped) {
    Tuning t;
    const Tuning *tune = tuning_lookup(pp, t) ? &t : nullptr;
    const int rc = grouped ? schedule_pass(pp, tune, 512, 65536, 256, true) : schedule_pass(pp, tune);
    if (rc != DFFT_OK) return rc;
    if (grouped && pp.ap.variant > 1) return -1;  // the grouped kernel runs 8-value schedules only
    int prod = 1;
    for (int s = 0; s < pp.ap.sp.S; ++s) prod *= pp.ap.sp.R[s];
    if (prod != pp.ap.sp.L) return -2;
    return DFFT_OK;
}

int main(int argc, char **argv) {
    if (argc > 1 && dfft_tune_load(argv[1]) != DFFT_OK) {
        printf("cannot load %s\n", argv[1]);
        return 1;
    }
    int kind, prec, checked = 0, bad = 0;
    long long a0, a1, a2;
    while (scanf("%d %lld %lld %lld %d", &kind, &a0, &a1, &a2, &prec) == 5) {
        const int64_t n[3] = {a0, a1, a2};
        int64_t nc[3] = {a0, a1, a2};
        if (kind == 1) nc[0] = n[0] / 2 + 1;
        for (int ax = 0; ax < 3; ++ax) {
            PassPlan pp{};
            pp.prec = prec;
            pp.axis = ax;
            AxisPass &a = pp.ap;
            a.scale = 1.0;
            a.nt = 1;
            if (kind == 1 && ax == 0) {
                a.op = OP_R2C;
                a.sp.L = (int)(n[0] / 2);
                a.nlog = (int)n[0];
                set_geometry(a, 0, n, nc);
            } else {
                a.op = OP_C2C;
                a.sp.L = (int)nc[ax];
                a.nlog = a.sp.L;
                set_geometry(a, ax, nc, nc);
            }
            for (int g = 0; g < 2; ++g) {
                if (g && kind == 1 && ax == 0 && n[0] % 2) continue;
                ++checked;
                const int rc = check(pp, g != 0);
                if (rc != DFFT_OK) {
                    ++bad;
                    printf("FAIL kind %d %lldx%lldx%lld prec %d axis %d grouped %d: %d\n", kind, a0, a1, a2, prec, ax,
                           g, rc);
                }
            }
        }
    }
    printf("checked %d bad %d\n", checked, bad);
    return bad ? 2 : 0;
}
