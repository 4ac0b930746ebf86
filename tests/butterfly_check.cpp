// Host check of the DFT butterflies in dfft_math.hpp (the same templates the
// kernels run): each radix on random complex inputs against a direct
// long-double DFT.  Prints "R precision max_rel_err" lines.
#include <cmath>
#include <cstdio>
#include <random>

#include "dfft_math.hpp"

using namespace dfft;

template <int R, class T>
static double check(std::mt19937_64 &g) {
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    double worst = 0;
    for (int trial = 0; trial < 200; ++trial) {
        long double xr[R], xi[R];
        cx<T> v[R];
        for (int r = 0; r < R; ++r) {
            v[r] = cx<T>{(T)u(g), (T)u(g)};
            xr[r] = v[r].x;
            xi[r] = v[r].y;
        }
        bfly<R>(v);
        long double num = 0, den = 0;
        for (int k = 0; k < R; ++k) {
            long double yr = 0, yi = 0;
            for (int r = 0; r < R; ++r) {
                const long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)((r * k) % R) / R;
                yr += xr[r] * cosl(a) - xi[r] * sinl(a);
                yi += xr[r] * sinl(a) + xi[r] * cosl(a);
            }
            num += (v[k].x - yr) * (v[k].x - yr) + (v[k].y - yi) * (v[k].y - yi);
            den += yr * yr + yi * yi;
        }
        const double e = (double)sqrtl(num / den);
        if (e > worst) worst = e;
    }
    return worst;
}

template <int R>
static void both(std::mt19937_64 &g) {
    printf("%d f64 %.3e\n", R, check<R, double>(g));
    printf("%d f32 %.3e\n", R, check<R, float>(g));
}

int main() {
    std::mt19937_64 g(12345);
    both<2>(g); both<3>(g); both<4>(g); both<5>(g); both<6>(g); both<7>(g); both<8>(g);
    both<9>(g); both<10>(g); both<12>(g); both<14>(g); both<15>(g); both<16>(g);
    return 0;
}
