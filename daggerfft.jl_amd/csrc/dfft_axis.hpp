// dfft_axis.hpp -- device body of one axis-pass tile (HIP, gfx950).
//
// A tile is TC lines of one axis, transformed by a Stockham autosort FFT
// (Stockham 1966; the formulation of Govindaraju et al., SC'08): stage s with
// radix R and Ns = R_0*...*R_{s-1} reads x[j + r*L/R], twiddles by
// w_{Ns*R}^{r*(j mod Ns)}, runs an R-point DFT and writes
// y[(j - j mod Ns)*R + j mod Ns + r*Ns].  Stage 0 reads straight from HBM and
// the last stage writes straight to HBM in natural order; both patterns are
// unit-stride across threads, so every global access is coalesced.  The
// stages in between exchange through LDS.
//
// Two layouts:
//   CONTIG   lines are contiguous (axis 0): a line's TPL threads are adjacent
//            lanes and walk the line; LDS holds one padded slot per line.
//   STRIDED  lines run across rows (axes 1, 2): lanes are adjacent columns,
//            so each row segment of TC elements is one contiguous burst; LDS
//            is row-major [p][column].
//
// The fused ops (R2C/C2R half-length packing, DCT-II/III reordering and
// twiddles) live in the stage-0 loader and the last-stage storer.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include "dfft_math.hpp"
#include "dfft_types.hpp"

namespace dfft {

// Global element access.  Streaming (nontemporal) accesses skip cache
// allocation: each element of a pass is read once and written once.
template <class T> struct vec2;
template <> struct vec2<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct vec2<float> { typedef float type __attribute__((ext_vector_type(2))); };

template <class T>
__device__ __forceinline__ cx<T> gload(const cx<T> *p, bool nt) {
    typedef typename vec2<T>::type V;
    V v = nt ? __builtin_nontemporal_load(reinterpret_cast<const V *>(p)) : *reinterpret_cast<const V *>(p);
    return cx<T>{v.x, v.y};
}
template <class T>
__device__ __forceinline__ void gstore(cx<T> *p, cx<T> x, bool nt) {
    typedef typename vec2<T>::type V;
    V v = {x.x, x.y};
    if (nt) __builtin_nontemporal_store(v, reinterpret_cast<V *>(p));
    else *reinterpret_cast<V *>(p) = v;
}
template <class T>
__device__ __forceinline__ T gload(const T *p, bool nt) {
    return nt ? __builtin_nontemporal_load(p) : *p;
}
template <class T>
__device__ __forceinline__ void gstore(T *p, T x, bool nt) {
    if (nt) __builtin_nontemporal_store(x, p);
    else *p = x;
}

template <class T, int OP, bool CONTIG>
struct TileCtx {
    const AxisPass &a;
    const void *in;
    void *out;
    cx<T> *lds;
    int ls, t;
    bool valid;
    int64_t ib, ob;

    __device__ __forceinline__ int li(int p) const {
        if constexpr (CONTIG) return ls * a.lds_pitch + p + (p >> a.pad_shift);
        else return p * a.lds_pitch + ls;
    }

    __device__ __forceinline__ const cx<T> *post() const {
        return reinterpret_cast<const cx<T> *>(a.tw) + a.sp.post;
    }

    // what the runtime stages read: FFT length, threads per line, twiddle table
    typedef T value_type;
    __device__ __forceinline__ int len() const { return a.sp.L; }
    __device__ __forceinline__ int tpl() const { return a.TPL; }
    __device__ __forceinline__ const cx<T> *twb() const { return reinterpret_cast<const cx<T> *>(a.tw); }
    __device__ __forceinline__ int stage_tw(int s) const { return a.sp.tw[s]; }

    // value entering the FFT at position p (stage 0)
    __device__ __forceinline__ cx<T> load0(int p) const {
        const T *inr = reinterpret_cast<const T *>(in);
        const cx<T> *inc = reinterpret_cast<const cx<T> *>(in);
        if constexpr (OP == OP_C2C) {
            if (!valid) return cx<T>{0, 0};
            cx<T> v = gload(inc + ib + (int64_t)p * a.in_sm, a.nt);
            return a.inverse ? conj(v) : v;
        } else if constexpr (OP == OP_R2C) {
            if (!valid) return cx<T>{0, 0};
            if constexpr (CONTIG) {
                // adjacent reals (2p, 2p+1) of a unit-stride line: one vector load
                return gload(inc + (ib >> 1) + p, a.nt);
            } else {
                return cx<T>{gload(inr + ib + (int64_t)(2 * p) * a.in_sm, a.nt),
                             gload(inr + ib + (int64_t)(2 * p + 1) * a.in_sm, a.nt)};
            }
        } else if constexpr (OP == OP_R2C_FULL) {
            if (!valid) return cx<T>{0, 0};
            return cx<T>{gload(inr + ib + (int64_t)p * a.in_sm, a.nt), (T)0};
        } else if constexpr (OP == OP_C2R) {
            // Z_p = A + i w^-p B from bins p and H-p (conjugated: the inverse
            // H-point FFT runs as a forward FFT on conjugates).  Both bins are
            // read from global memory with cached loads: the tile reads each
            // line forwards and backwards at the same time, so the second read
            // of a bin hits the cache.  The DC and Nyquist imaginary parts are
            // ignored (FFTW's c2r input has no slot for them).
            if (!valid) return cx<T>{0, 0};
            const int H = a.sp.L;
            cx<T> xk = gload(inc + ib + (int64_t)p * a.in_sm, false);
            cx<T> xh = gload(inc + ib + (int64_t)(H - p) * a.in_sm, false);
            if (p == 0) {
                xk.y = 0;
                xh.y = 0;
            }
            cx<T> xc = conj(xh);
            cx<T> A = xk + xc, B = xk - xc;
            cx<T> wB = cmul(conj(post()[p]), B);
            cx<T> z = A + mul_pos_i(wB);
            return conj(z);
        } else if constexpr (OP == OP_C2R_FULL) {
            if (!valid) return cx<T>{0, 0};
            const int L = a.sp.L;
            cx<T> x;
            if (2 * p <= L) x = gload(inc + ib + (int64_t)p * a.in_sm, a.nt);
            else x = conj(gload(inc + ib + (int64_t)(L - p) * a.in_sm, a.nt));
            if (p == 0 || 2 * p == L) x.y = 0;
            return conj(x);
        } else if constexpr (OP == OP_DCT2 || OP == OP_DST2) {
            if (!valid) return cx<T>{0, 0};
            const int L = a.sp.L;
            int q = (p < (L + 1) / 2) ? 2 * p : 2 * (L - 1 - p) + 1;
            T x = gload(inr + ib + (int64_t)q * a.in_sm, a.nt);
            if constexpr (OP == OP_DST2) x = (q & 1) ? -x : x;
            return cx<T>{x, (T)0};
        } else if constexpr (OP == OP_DCT3 || OP == OP_DST3) {
            if (!valid) return cx<T>{0, 0};
            const int L = a.sp.L;
            // DST-III reads the line reversed: x'_p = x_{L-1-p}
            const int ip = OP == OP_DCT3 ? p : L - 1 - p;
            const int ir = OP == OP_DCT3 ? L - p : p - 1;
            T xp = gload(inr + ib + (int64_t)ip * a.in_sm, a.nt);
            T xr = p == 0 ? (T)0 : gload(inr + ib + (int64_t)ir * a.in_sm, a.nt);
            // conj( exp(+i pi p / 2L) (x_p - i x_{L-p}) ) = w_p (x_p + i x_{L-p})
            return cmul(post()[p], cx<T>{xp, xr});
        } else if constexpr (OP == OP_DHT) {
            if (!valid) return cx<T>{0, 0};
            return cx<T>{gload(inr + ib + (int64_t)p * a.in_sm, a.nt), (T)0};
        } else if constexpr (OP == OP_DCT1) {
            // even extension x_0 .. x_{n-1}, x_{n-2} .. x_1
            if (!valid) return cx<T>{0, 0};
            const int n = a.nlog, q = p < n ? p : a.sp.L - p;
            return cx<T>{gload(inr + ib + (int64_t)q * a.in_sm, a.nt), (T)0};
        } else if constexpr (OP == OP_DST1) {
            // odd extension 0, x_0 .. x_{n-1}, 0, -x_{n-1} .. -x_0
            if (!valid) return cx<T>{0, 0};
            const int n = a.nlog;
            if (p == 0 || p == n + 1) return cx<T>{0, 0};
            if (p <= n) return cx<T>{gload(inr + ib + (int64_t)(p - 1) * a.in_sm, a.nt), (T)0};
            return cx<T>{-gload(inr + ib + (int64_t)(a.sp.L - p - 1) * a.in_sm, a.nt), (T)0};
        } else if constexpr (OP == OP_DCT2H) {
            // z_p = v_2p + i v_2p+1 of Makhoul's v_i = x[2i] (i < H), x[2n-2i-1] (i >= H)
            if (!valid) return cx<T>{0, 0};
            const int H = a.sp.L, n = 2 * H;
            auto v = [&](int i) {
                const int q = i < H ? 2 * i : 2 * n - 2 * i - 1;
                return gload(inr + ib + (int64_t)q * a.in_sm, a.nt);
            };
            return cx<T>{v(2 * p), v(2 * p + 1)};
        } else if constexpr (OP == OP_DCT3H) {
            // U_k = e^{+i pi k/2n} (X_k - i X_{n-k}) (X_n = 0) is the spectrum of
            // Makhoul's v; Z_p = (U_p + conj U_{H-p}) + i e^{2 pi i p/n} (U_p -
            // conj U_{H-p}) that of z_j = v_2j + i v_2j+1 (the inverse H-point FFT
            // runs as a forward FFT on conjugates).  Each X_k is read by two
            // positions: cached loads.
            if (!valid) return cx<T>{0, 0};
            const int H = a.sp.L, n = 2 * H;
            const cx<T> *pw = post(), *ww = post() + H + 1;
            auto X = [&](int k) { return k >= n ? (T)0 : gload(inr + ib + (int64_t)k * a.in_sm, false); };
            auto U = [&](int k) { return cmul(conj(ww[k]), cx<T>{X(k), -X(n - k)}); };
            const cx<T> Up = U(p), Uh = U(H - p);
            const cx<T> A = Up + conj(Uh), B = Up - conj(Uh);
            return conj(A + mul_pos_i(cmul(conj(pw[p]), B)));
        } else {  // OP_DCT4, OP_DST4: x_p e^{-i pi p / 2n}, zero-padded to 2n
            if (!valid || p >= a.nlog) return cx<T>{0, 0};
            T x = gload(inr + ib + (int64_t)p * a.in_sm, a.nt);
            return scl(post()[2 * p], x);
        }
    }

    // last-stage output V at natural position p
    __device__ __forceinline__ void store_last(int p, cx<T> v) const {
        T *outr = reinterpret_cast<T *>(out);
        cx<T> *outc = reinterpret_cast<cx<T> *>(out);
        const T s = (T)a.scale;
        if constexpr (OP == OP_R2C || OP == OP_DCT2H) {
            lds[li(p)] = v;  // post-processed after a barrier
        } else if constexpr (OP == OP_DCT3H) {
            // z_p = conj(v) = v_2p + i v_2p+1, v_i at x[2i] (i < H), x[2n-2i-1] (i >= H)
            if (!valid) return;
            const int H = a.sp.L, n = 2 * H;
            auto put = [&](int i, T y) {
                const int q = i < H ? 2 * i : 2 * n - 2 * i - 1;
                gstore(outr + ob + (int64_t)q * a.out_sm, y * s, a.nt);
            };
            put(2 * p, v.x);
            put(2 * p + 1, -v.y);
        } else {
            if (!valid) return;
            if constexpr (OP == OP_C2C) {
                cx<T> o = a.inverse ? conj(v) : v;
                gstore(outc + ob + (int64_t)p * a.out_sm, scl(o, s), a.nt);
            } else if constexpr (OP == OP_R2C_FULL) {
                if (2 * p <= a.sp.L) gstore(outc + ob + (int64_t)p * a.out_sm, scl(v, s), a.nt);
            } else if constexpr (OP == OP_C2R) {
                if constexpr (CONTIG) {
                    gstore(reinterpret_cast<cx<T> *>(out) + (ob >> 1) + p, cx<T>{v.x * s, -v.y * s}, a.nt);
                } else {
                    gstore(outr + ob + (int64_t)(2 * p) * a.out_sm, v.x * s, a.nt);
                    gstore(outr + ob + (int64_t)(2 * p + 1) * a.out_sm, -v.y * s, a.nt);
                }
            } else if constexpr (OP == OP_C2R_FULL) {
                gstore(outr + ob + (int64_t)p * a.out_sm, v.x * s, a.nt);
            } else if constexpr (OP == OP_DCT2 || OP == OP_DST2) {
                cx<T> w = post()[p];
                const int q = OP == OP_DCT2 ? p : a.sp.L - 1 - p;
                gstore(outr + ob + (int64_t)q * a.out_sm, (T)2 * (w.x * v.x - w.y * v.y) * s, a.nt);
            } else if constexpr (OP == OP_DCT3 || OP == OP_DST3) {
                const int L = a.sp.L;
                int q = (p < (L + 1) / 2) ? 2 * p : 2 * (L - 1 - p) + 1;
                const T y = (OP == OP_DST3 && (q & 1)) ? -v.x : v.x;
                gstore(outr + ob + (int64_t)q * a.out_sm, y * s, a.nt);
            } else if constexpr (OP == OP_DHT) {
                gstore(outr + ob + (int64_t)p * a.out_sm, (v.x - v.y) * s, a.nt);
            } else if constexpr (OP == OP_DCT1) {
                if (p < a.nlog) gstore(outr + ob + (int64_t)p * a.out_sm, v.x * s, a.nt);
            } else if constexpr (OP == OP_DST1) {
                if (p >= 1 && p <= a.nlog) gstore(outr + ob + (int64_t)(p - 1) * a.out_sm, -v.y * s, a.nt);
            } else {  // OP_DCT4, OP_DST4
                if (p < a.nlog) {
                    const cx<T> z = cmul(post()[2 * p + 1], v);
                    gstore(outr + ob + (int64_t)p * a.out_sm, (T)2 * (OP == OP_DCT4 ? z.x : -z.y) * s, a.nt);
                }
            }
        }
    }
};

// Stage position in the schedule: where its input comes from and where its
// output goes.  SK_FIRST_LDS is a first stage (no twiddles) whose input is
// already in LDS in natural order (the second FFT of a Bluestein pass).
enum StageKind : int { SK_FIRST = 0, SK_MID = 1, SK_LAST = 2, SK_ONLY = 3, SK_FIRST_LDS = 4 };

// One Stockham stage with compile-time radix R.  Each thread owns up to
// EMAX/R butterflies j = t, t+TPL, ...  C is the tile context: its load0 and
// store_last are the fused global maps of the first and last stage.
template <int R, int EMAX, int SK, class C>
__device__ __forceinline__ void run_stage(const C &c, int Ns, int twoff) {
    typedef typename C::value_type T;
    constexpr bool global_in = SK == SK_FIRST || SK == SK_ONLY;
    constexpr bool first = global_in || SK == SK_FIRST_LDS;
    constexpr bool last = SK == SK_LAST || SK == SK_ONLY;
    constexpr int NBM = EMAX / R;
    const int L = c.len(), NB = L / R, TPL = c.tpl();
    cx<T> v[NBM][R];
#pragma unroll
    for (int b = 0; b < NBM; ++b) {
        const int j = c.t + b * TPL;
        if (j < NB) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if constexpr (global_in) v[b][r] = c.load0(j + r * NB);
                else v[b][r] = c.lds[c.li(j + r * NB)];
            }
        }
    }
    // every read of the shared slots must finish before anyone overwrites them
    if constexpr (!global_in) __syncthreads();
#pragma unroll
    for (int b = 0; b < NBM; ++b) {
        const int j = c.t + b * TPL;
        if (j < NB) {
            if constexpr (!first) {
                const int k = j % Ns;
                const cx<T> *w = c.twb() + twoff + k * (R - 1);
#pragma unroll
                for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], w[r - 1]);
            }
            bfly<R>(v[b]);
            if constexpr (last) {
#pragma unroll
                for (int r = 0; r < R; ++r) c.store_last(j + r * NB, v[b][r]);
            } else {
                const int k = first ? 0 : j % Ns;
                const int base = (j - k) * R + k;
#pragma unroll
                for (int r = 0; r < R; ++r) c.lds[c.li(base + r * Ns)] = v[b][r];
            }
        }
    }
    if constexpr (!last) __syncthreads();
}

template <int SK, int EMAX, class C>
__device__ __forceinline__ void dispatch_stage(const C &c, int R, int Ns, int twoff) {
    switch (R) {
    case 8: run_stage<8, EMAX, SK>(c, Ns, twoff); break;
    case 4: run_stage<4, EMAX, SK>(c, Ns, twoff); break;
    case 2: run_stage<2, EMAX, SK>(c, Ns, twoff); break;
    case 3: run_stage<3, EMAX, SK>(c, Ns, twoff); break;
    case 5: run_stage<5, EMAX, SK>(c, Ns, twoff); break;
    case 7: run_stage<7, EMAX, SK>(c, Ns, twoff); break;
    default: break;
    }
}

// Fixed-schedule engine (compile-time length, radices and threads per line).
// Per-thread registers: stage s with radix R keeps butterfly b (j = t + b*TPL,
// j < L/R) in v[b*R .. b*R+R-1].  Between stages the values move through
// LDS: either whole complex values (one round), or -- split mode, half the
// LDS -- the real parts in one round and the imaginary parts in a second
// round through the same buffer.
enum LdsPart : int { PART_FULL = 0, PART_RE = 1, PART_IM = 2 };

template <int R, int EMAX, int LC, int TPLC, class C>
__device__ __forceinline__ void st_load_global(const C &c, cx<typename C::value_type> *v) {
    constexpr int NBM = EMAX / R;
    const int NB = LC / R, TPL = TPLC;
#pragma unroll
    for (int b = 0; b < NBM; ++b) {
        const int j = c.t + b * TPL;
        if (j < NB) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b * R + r] = c.load0(j + r * NB);
        }
    }
}

// twiddle by w_{Ns*R}^{r (j mod Ns)} (not in the first stage), then the R-point DFT
template <int R, int EMAX, int LC, int TPLC, int Ns, class C>
__device__ __forceinline__ void st_compute(const C &c, cx<typename C::value_type> *v, int twoff) {
    typedef typename C::value_type T;
    constexpr bool twiddle = Ns > 1;
    constexpr int NBM = EMAX / R;
    const int NB = LC / R, TPL = TPLC;
#pragma unroll
    for (int b = 0; b < NBM; ++b) {
        const int j = c.t + b * TPL;
        if (j < NB) {
            if constexpr (twiddle) {
                const int k = j % Ns;
                const cx<T> *w = c.twb() + twoff + k * (R - 1);
#pragma unroll
                for (int r = 1; r < R; ++r) v[b * R + r] = cmul(v[b * R + r], w[r - 1]);
            }
            bfly<R>(v + b * R);
        }
    }
}

template <int R, int EMAX, int LC, int TPLC, class C>
__device__ __forceinline__ void st_store_global(const C &c, const cx<typename C::value_type> *v) {
    constexpr int NBM = EMAX / R;
    const int NB = LC / R, TPL = TPLC;
#pragma unroll
    for (int b = 0; b < NBM; ++b) {
        const int j = c.t + b * TPL;
        if (j < NB) {
#pragma unroll
            for (int r = 0; r < R; ++r) c.store_last(j + r * NB, v[b * R + r]);
        }
    }
}

// DCT-II (REDFT10) and DCT-III (REDFT01) along contiguous lines move the
// reals in pairs.  FFT position p holds x[2p] for p < L/2 and x[2(L-1-p)+1]
// for p >= L/2, so the pair (x[2m], x[2m+1]), m < L/2, feeds positions m and
// L-1-m: one vector access per pair, unit-stride across lanes (single reals
// at stride 2 fetch each cache line twice).  With R even and NB = L/R =
// NBM * TPL, position m = j + r*NB (r < R/2) belongs to butterfly j = t + b*TPL
// and position L-1-m to butterfly NB-1-j = (TPL-1-t) + (NBM-1-b)*TPL, slot
// R-1-r: the line's mirror lane TPL-1-t, reached by a lane shuffle (a line's
// TPL lanes are adjacent and aligned within a wavefront when TPL divides 64).
template <int R, int EMAX, int LC, int TPLC>
constexpr bool dct_pairs_ok() {
    return R % 2 == 0 && LC % 2 == 0 && (LC / R) == (EMAX / R) * TPLC && TPLC <= 64 && (64 % TPLC) == 0;
}

template <int R, int EMAX, int LC, int TPLC, class T>
__device__ __forceinline__ void st_load_dct2_pairs(const TileCtx<T, OP_DCT2, true> &c, cx<T> *v) {
    constexpr int NBM = EMAX / R, NB = LC / R, H = R / 2;
    const cx<T> *pairs = reinterpret_cast<const cx<T> *>(c.in) + (c.ib >> 1);
    const bool ld = c.valid && c.t < TPLC;
    T odd[NBM][H];
#pragma unroll
    for (int b = 0; b < NBM; ++b) {
        const int j = c.t + b * TPLC;
#pragma unroll
        for (int r = 0; r < H; ++r) {
            const cx<T> z = ld ? gload(pairs + j + r * NB, c.a.nt) : cx<T>{0, 0};
            v[b * R + r] = cx<T>{z.x, (T)0};
            odd[b][r] = z.y;
        }
    }
    const int mirror = TPLC - 1 - (c.t & (TPLC - 1));
#pragma unroll
    for (int b = 0; b < NBM; ++b)
#pragma unroll
        for (int r = 0; r < H; ++r) v[b * R + R - 1 - r] = cx<T>{__shfl(odd[NBM - 1 - b][r], mirror, TPLC), (T)0};
}

template <int R, int EMAX, int LC, int TPLC, class T>
__device__ __forceinline__ void st_store_dct3_pairs(const TileCtx<T, OP_DCT3, true> &c, const cx<T> *v) {
    constexpr int NBM = EMAX / R, NB = LC / R, H = R / 2;
    cx<T> *pairs = reinterpret_cast<cx<T> *>(c.out) + (c.ob >> 1);
    const T s = (T)c.a.scale;
    const int mirror = TPLC - 1 - (c.t & (TPLC - 1));
#pragma unroll
    for (int b = 0; b < NBM; ++b) {
        const int j = c.t + b * TPLC;
#pragma unroll
        for (int r = 0; r < H; ++r) {
            // position L-1-m from the mirror lane (every lane takes part in the shuffle)
            const T w = __shfl(v[(NBM - 1 - b) * R + R - 1 - r].x, mirror, TPLC);
            if (c.valid && c.t < TPLC) gstore(pairs + j + r * NB, cx<T>{v[b * R + r].x * s, w * s}, c.a.nt);
        }
    }
}

// The half-length DCTs along contiguous lines move the reals in quads: with
// n = 2H and H even, x[4m..4m+3] hold v_2m, v_n-1-2m, v_2m+1, v_n-2-2m, that
// is z_m = (x[4m], x[4m+2]) and z_{H-1-m} = (x[4m+3], x[4m+1]) (m < H/2):
// positions m and H-1-m of the H-point FFT, on mirror lanes as above.
template <int R, int EMAX, int LC, int TPLC, class T>
__device__ __forceinline__ void st_load_dct2h_quads(const TileCtx<T, OP_DCT2H, true> &c, cx<T> *v) {
    constexpr int NBM = EMAX / R, NB = LC / R, HR = R / 2;
    const cx<T> *pairs = reinterpret_cast<const cx<T> *>(c.in) + (c.ib >> 1);
    const bool ld = c.valid && c.t < TPLC;
    cx<T> far[NBM][HR];
#pragma unroll
    for (int b = 0; b < NBM; ++b) {
        const int j = c.t + b * TPLC;
#pragma unroll
        for (int r = 0; r < HR; ++r) {
            const int m = j + r * NB;
            const cx<T> lo = ld ? gload(pairs + 2 * m, c.a.nt) : cx<T>{0, 0};
            const cx<T> hi = ld ? gload(pairs + 2 * m + 1, c.a.nt) : cx<T>{0, 0};
            v[b * R + r] = cx<T>{lo.x, hi.x};
            far[b][r] = cx<T>{hi.y, lo.y};
        }
    }
    const int mirror = TPLC - 1 - (c.t & (TPLC - 1));
#pragma unroll
    for (int b = 0; b < NBM; ++b)
#pragma unroll
        for (int r = 0; r < HR; ++r)
            v[b * R + R - 1 - r] = cx<T>{__shfl(far[NBM - 1 - b][r].x, mirror, TPLC),
                                         __shfl(far[NBM - 1 - b][r].y, mirror, TPLC)};
}

template <int R, int EMAX, int LC, int TPLC, class T>
__device__ __forceinline__ void st_store_dct3h_quads(const TileCtx<T, OP_DCT3H, true> &c, const cx<T> *v) {
    constexpr int NBM = EMAX / R, NB = LC / R, HR = R / 2;
    cx<T> *pairs = reinterpret_cast<cx<T> *>(c.out) + (c.ob >> 1);
    const T s = (T)c.a.scale;
    const int mirror = TPLC - 1 - (c.t & (TPLC - 1));
#pragma unroll
    for (int b = 0; b < NBM; ++b) {
        const int j = c.t + b * TPLC;
#pragma unroll
        for (int r = 0; r < HR; ++r) {
            // position m = j + r*NB here, H-1-m on the mirror lane; z = conj(v)
            const cx<T> w = v[(NBM - 1 - b) * R + R - 1 - r];
            const cx<T> f{__shfl(w.x, mirror, TPLC), __shfl(w.y, mirror, TPLC)};
            const cx<T> o = v[b * R + r];
            if (c.valid && c.t < TPLC) {
                const int m = j + r * NB;
                gstore(pairs + 2 * m, cx<T>{o.x * s, -f.y * s}, c.a.nt);
                gstore(pairs + 2 * m + 1, cx<T>{-o.y * s, f.x * s}, c.a.nt);
            }
        }
    }
}

// DCT-II outputs from the H-point FFT Z of z_m = v_2m + i v_2m+1 in LDS:
// V_k = 1/2 [(Z_k + conj Z_{H-k}) - i e^{-2 pi i k/n} (Z_k - conj Z_{H-k})]
// (k = 0..H) is the n-point DFT of Makhoul's v, and with u = e^{-i pi k/2n} V_k
// the DCT-II is X_k = 2 Re u and X_{n-k} = -2 Im u.
template <class T, bool CONTIG>
__device__ __forceinline__ void dct2h_post_lds(const TileCtx<T, OP_DCT2H, CONTIG> &c, int TPL) {
    __syncthreads();
    const int H = c.a.sp.L, n = 2 * H;
    const T s = (T)c.a.scale;
    T *outr = reinterpret_cast<T *>(c.out);
    const cx<T> *pw = c.post(), *ww = c.post() + H + 1;
    for (int k = c.t; k <= H; k += TPL) {
        const cx<T> zk = c.lds[c.li(k == H ? 0 : k)];
        const cx<T> zc = conj(c.lds[c.li(k == 0 ? 0 : H - k)]);
        const cx<T> A = zk + zc, B = zk - zc;
        const cx<T> V = scl(A + mul_neg_i(cmul(pw[k], B)), (T)0.5);
        const cx<T> u = cmul(ww[k], V);
        if (c.valid) {
            gstore(outr + c.ob + (int64_t)k * c.a.out_sm, (T)2 * u.x * s, c.a.nt);
            if (k > 0 && k < H) gstore(outr + c.ob + (int64_t)(n - k) * c.a.out_sm, (T)-2 * u.y * s, c.a.nt);
        }
    }
}

// Stockham write of stage outputs: y[(j - j mod Ns) R + j mod Ns + r Ns]
template <int R, int EMAX, int PART, int LC, int TPLC, int Ns, class C>
__device__ __forceinline__ void st_write_lds(const C &c, const cx<typename C::value_type> *v) {
    typedef typename C::value_type T;
    constexpr int NBM = EMAX / R;
    const int NB = LC / R, TPL = TPLC;
    T *ldsr = reinterpret_cast<T *>(c.lds);
#pragma unroll
    for (int b = 0; b < NBM; ++b) {
        const int j = c.t + b * TPL;
        if (j < NB) {
            const int k = j % Ns;
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = c.li(base + r * Ns);
                if constexpr (PART == PART_FULL) c.lds[i] = v[b * R + r];
                else if constexpr (PART == PART_RE) ldsr[i] = v[b * R + r].x;
                else ldsr[i] = v[b * R + r].y;
            }
        }
    }
}

// Stockham read of stage inputs: x[j + r L/R]
template <int R, int EMAX, int PART, int LC, int TPLC, class C>
__device__ __forceinline__ void st_read_lds(const C &c, cx<typename C::value_type> *v) {
    typedef typename C::value_type T;
    constexpr int NBM = EMAX / R;
    const int NB = LC / R, TPL = TPLC;
    const T *ldsr = reinterpret_cast<const T *>(c.lds);
#pragma unroll
    for (int b = 0; b < NBM; ++b) {
        const int j = c.t + b * TPL;
        if (j < NB) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = c.li(j + r * NB);
                if constexpr (PART == PART_FULL) v[b * R + r] = c.lds[i];
                else if constexpr (PART == PART_RE) v[b * R + r].x = ldsr[i];
                else v[b * R + r].y = ldsr[i];
            }
        }
    }
}

template <int... Rs>
struct Radices {
    static constexpr int S = sizeof...(Rs);
    static constexpr int R[S] = {Rs...};
    static constexpr int L = (Rs * ...);
    // product of the radices before stage s
    static constexpr int ns(int s) {
        int p = 1;
        for (int i = 0; i < s; ++i) p *= R[i];
        return p;
    }
};

// Compile-time factorizations.
//   EMAX 8:  the runtime planner's greedy order -- radix-8 stages first, then
//            4, 2, 3, 5, 7.
//   EMAX 16: the fewest stages with radices up to 16 (the composite
//            butterflies 6, 9, 10, 12, 14, 15, 16 included), ties broken by
//            the fewest threads per line: every stage is one round trip
//            through LDS, so 486 = 9*9*6 (3 stages) beats 2*3*3*3*3*3 (6).
struct Factors {
    int S;
    int R[MAX_STAGES];
};
constexpr int stage_tpl(const int *R, int S, int L, int emax) {
    int t = 1;
    for (int s = 0; s < S; ++s) {
        const int per = emax / R[s], need = (L / R[s] + per - 1) / per;
        if (need > t) t = need;
    }
    return t;
}
constexpr void min_stage_search(int rem, int start, Factors &cur, int L, int emax, Factors &best, int &best_tpl,
                                int rmax = 16) {
    constexpr int set[13] = {16, 15, 14, 12, 10, 9, 8, 7, 6, 5, 4, 3, 2};
    if (rem == 1) {
        const int t = stage_tpl(cur.R, cur.S, L, emax);
        if (best.S < 0 || cur.S < best.S || (cur.S == best.S && t < best_tpl)) {
            best = cur;
            best_tpl = t;
        }
        return;
    }
    if (cur.S >= MAX_STAGES || (best.S >= 0 && cur.S + 1 > best.S)) return;
    for (int i = start; i < 13; ++i) {
        const int r = set[i];
        if (r > emax || r > rmax || rem % r) continue;
        cur.R[cur.S++] = r;
        min_stage_search(rem / r, i, cur, L, emax, best, best_tpl, rmax);
        --cur.S;
    }
}
// the fewest stages with radices up to rmax, 16 values per thread
constexpr Factors factors_capped(int L, int rmax) {
    Factors cur{0, {0}}, best{-1, {0}};
    int best_tpl = 0;
    min_stage_search(L, 0, cur, L, 16, best, best_tpl, rmax);
    return best;
}
constexpr Factors factors_of(int L, int emax) {
    Factors f{0, {0}};
    if (emax < 16) {
        constexpr int order[6] = {8, 4, 2, 3, 5, 7};
        int n = L;
        for (int i = 0; i < 6; ++i)
            while (n > 1 && n % order[i] == 0 && f.S < MAX_STAGES) {
                f.R[f.S++] = order[i];
                n /= order[i];
            }
        if (n != 1) f.S = -1;
        return f;
    }
    Factors cur{0, {0}}, best{-1, {0}};
    int best_tpl = 0;
    min_stage_search(L, 0, cur, L, emax, best, best_tpl);
    return best;
}
template <int L, int EMAX, class Seq>
struct RadicesOfImpl;
template <int L, int EMAX, int... I>
struct RadicesOfImpl<L, EMAX, std::integer_sequence<int, I...>> {
    using type = Radices<factors_of(L, EMAX).R[I]...>;
};
template <int L, int EMAX>
using RadicesOf = typename RadicesOfImpl<L, EMAX, std::make_integer_sequence<int, factors_of(L, EMAX).S>>::type;
template <int L, int RMAX, class Seq>
struct RadicesCapImpl;
template <int L, int RMAX, int... I>
struct RadicesCapImpl<L, RMAX, std::integer_sequence<int, I...>> {
    using type = Radices<factors_capped(L, RMAX).R[I]...>;
};
// fewest stages with radices up to 10 (16 values per thread): the FP64
// two-FFT tiles spill with the larger composite butterflies
template <int L>
using RadicesCap10 = typename RadicesCapImpl<L, 10, std::make_integer_sequence<int, factors_capped(L, 10).S>>::type;
template <int L>
using RadicesOf8 = RadicesOf<L, 8>;  // (single template argument: usable inside macros)
template <int L>
using RadicesOf16 = RadicesOf<L, 16>;

// threads per line such that each stage's L/R butterflies fit EMAX/R per thread
template <class RL, int EMAX>
constexpr int fixed_tpl() {
    int t = 1;
    for (int s = 0; s < RL::S; ++s) {
        const int nb = RL::L / RL::R[s], per = EMAX / RL::R[s];
        const int need = (nb + per - 1) / per;
        if (need > t) t = need;
    }
    return t;
}

template <class RL, int EMAX, int TPLC, bool SPLIT, int s, class C>
__device__ __forceinline__ void fixed_stages_from(const C &c, cx<typename C::value_type> *v) {
    if constexpr (s + 1 < RL::S) {
        constexpr int Rs = RL::R[s], Rn = RL::R[s + 1], Ns = RL::ns(s), L = RL::L;
        if constexpr (!SPLIT) {
            st_write_lds<Rs, EMAX, PART_FULL, L, TPLC, Ns>(c, v);
            __syncthreads();
            st_read_lds<Rn, EMAX, PART_FULL, L, TPLC>(c, v);
            __syncthreads();
        } else {
            st_write_lds<Rs, EMAX, PART_RE, L, TPLC, Ns>(c, v);
            __syncthreads();
            st_read_lds<Rn, EMAX, PART_RE, L, TPLC>(c, v);
            __syncthreads();
            st_write_lds<Rs, EMAX, PART_IM, L, TPLC, Ns>(c, v);
            __syncthreads();
            st_read_lds<Rn, EMAX, PART_IM, L, TPLC>(c, v);
            __syncthreads();
        }
        st_compute<Rn, EMAX, L, TPLC, Ns * Rs>(c, v, c.stage_tw(s + 1));
        fixed_stages_from<RL, EMAX, TPLC, SPLIT, s + 1>(c, v);
    }
}

// R2C post-twiddle with the split exchange.  The last stage leaves the
// H-point FFT Z of the packed line in registers (Z_p at p = j + r*NB); bin k
// needs Z_k and Z_{H-k} (indices mod H).  With Z_k = a + ib and
// Z_{H-k} = c + id and w = w_k = e^{-i pi k/H}:
//   X_k = 1/2 [(Z_k + conj Z_{H-k}) - i w (Z_k - conj Z_{H-k})]
//   Re X_k = 1/2 [(a + c) + Im w (a - c) + Re w (b + d)]
//   Im X_k = 1/2 [(b - d) - Re w (a - c) + Im w (b + d)]
// The real parts go through LDS first and leave partial sums, then the
// imaginary parts finish them; each thread takes the pairs (k, H-k),
// k = t, t+TPL, ... <= H/2, which read the same four reals.
template <int R, int EMAX, int LC, int TPLC, class T, int OP, bool CONTIG>
__device__ __forceinline__ void r2c_post_split(const TileCtx<T, OP, CONTIG> &c, const cx<T> *v) {
    constexpr int NBM = EMAX / R, NB = LC / R, TPL = TPLC, H = LC;
    constexpr int NK = (H / 2 + TPL) / TPL;  // pairs k = t + i*TPL <= H/2
    T *ldsr = reinterpret_cast<T *>(c.lds);
    const cx<T> *pw = c.post();
    T pr[NK][2], pim[NK][2];
#pragma unroll
    for (int part = 0; part < 2; ++part) {
        // the last read of the shared slots (if any) ended with a barrier
#pragma unroll
        for (int b = 0; b < NBM; ++b) {
            const int j = c.t + b * TPL;
            if (j < NB) {
#pragma unroll
                for (int r = 0; r < R; ++r) ldsr[c.li(j + r * NB)] = part ? v[b * R + r].y : v[b * R + r].x;
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NK; ++i) {
            const int k = c.t + i * TPL;
            if (k <= H / 2) {
                const T zk = ldsr[c.li(k)], zh = ldsr[c.li(k == 0 ? 0 : H - k)];
                const cx<T> wk = pw[k], wh = pw[H - k];
                if (part == 0) {  // zk = a, zh = c
                    pr[i][0] = (zk + zh) + wk.y * (zk - zh);
                    pim[i][0] = -wk.x * (zk - zh);
                    pr[i][1] = (zh + zk) + wh.y * (zh - zk);
                    pim[i][1] = -wh.x * (zh - zk);
                } else {  // zk = b, zh = d
                    const T s = (T)0.5 * (T)c.a.scale;
                    const cx<T> xk{s * (pr[i][0] + wk.x * (zk + zh)), s * (pim[i][0] + (zk - zh) + wk.y * (zk + zh))};
                    const cx<T> xh{s * (pr[i][1] + wh.x * (zh + zk)), s * (pim[i][1] + (zh - zk) + wh.y * (zh + zk))};
                    if (c.valid) {
                        cx<T> *outc = reinterpret_cast<cx<T> *>(c.out);
                        gstore(outc + c.ob + (int64_t)k * c.a.out_sm, xk, c.a.nt);
                        if (H - k != k) gstore(outc + c.ob + (int64_t)(H - k) * c.a.out_sm, xh, c.a.nt);
                    }
                }
            }
        }
        if (part == 0) __syncthreads();
    }
}

// Stage-0 loads of a C2C tile: the positions st_load_global reads, with the
// C2C load map of TileCtx::load0 reordered so that every load is issued
// before any loaded value is used.  Lines past the array's end read line 0
// of the pass (in bounds; those lines compute on it and are never stored)
// instead of taking zeros, and the inverse transform's conjugate follows the
// loads.  With the map's validity select and conjugate after each load, the
// compiler waited for every pair of loads before issuing the next.
template <int R, int EMAX, int LC, int TPLC, class C>
__device__ __forceinline__ void c2c_load_global(const C &c, cx<typename C::value_type> *v) {
    typedef typename C::value_type T;
    constexpr int NBM = EMAX / R;
    const int NB = LC / R;
    const cx<T> *inc = reinterpret_cast<const cx<T> *>(c.in) + (c.valid ? c.ib : 0);
#pragma unroll
    for (int b = 0; b < NBM; ++b) {
        const int j = c.t + b * TPLC;
        if (j < NB) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b * R + r] = gload(inc + (int64_t)(j + r * NB) * c.a.in_sm, c.a.nt);
        }
    }
    if (c.a.inverse) {
#pragma unroll
        for (int b = 0; b < NBM; ++b) {
            if (c.t + b * TPLC < NB) {
#pragma unroll
                for (int r = 0; r < R; ++r) v[b * R + r] = conj(v[b * R + r]);
            }
        }
    }
}

// All stages of a fixed schedule: global -> registers, LDS exchanges, registers -> global.
template <class RL, int EMAX, int TPLC, bool SPLIT, class T, int OP, bool CONTIG>
__device__ __forceinline__ void run_fixed_stages(const TileCtx<T, OP, CONTIG> &c) {
    constexpr int L = RL::L, R0 = RL::R[0], RN = RL::R[RL::S - 1];
    cx<T> v[EMAX];
    if constexpr (OP == OP_C2C && sizeof(T) == 4) c2c_load_global<R0, EMAX, L, TPLC>(c, v);
    else if constexpr (OP == OP_C2C) {
        if (c.a.paired_loads) st_load_global<R0, EMAX, L, TPLC>(c, v);
        else c2c_load_global<R0, EMAX, L, TPLC>(c, v);
    }
    else if constexpr (OP == OP_DCT2 && CONTIG && dct_pairs_ok<R0, EMAX, L, TPLC>())
        st_load_dct2_pairs<R0, EMAX, L, TPLC>(c, v);
    else if constexpr (OP == OP_DCT2H && CONTIG && dct_pairs_ok<R0, EMAX, L, TPLC>())
        st_load_dct2h_quads<R0, EMAX, L, TPLC>(c, v);
    else st_load_global<R0, EMAX, L, TPLC>(c, v);
    st_compute<R0, EMAX, L, TPLC, 1>(c, v, 0);
    fixed_stages_from<RL, EMAX, TPLC, SPLIT, 0>(c, v);
    if constexpr (OP == OP_R2C && SPLIT) r2c_post_split<RN, EMAX, L, TPLC>(c, v);
    else if constexpr (OP == OP_DCT3 && CONTIG && dct_pairs_ok<RN, EMAX, L, TPLC>())
        st_store_dct3_pairs<RN, EMAX, L, TPLC>(c, v);
    else if constexpr (OP == OP_DCT3H && CONTIG && dct_pairs_ok<RN, EMAX, L, TPLC>())
        st_store_dct3h_quads<RN, EMAX, L, TPLC>(c, v);
    else st_store_global<RN, EMAX, L, TPLC>(c, v);
}

// Runs tile `tile` of pass `a` with the calling workgroup.  Threads with
// tid >= TPL*TC (a persistent workgroup larger than the tile needs) take part
// in every barrier but touch no memory.
// With RL (a Radices<...> list) the stages are the compile-time schedule RL
// with TPLC threads per line; otherwise they follow the runtime schedule a.sp.
struct NoRadices {};

// the tile a single-pass launch's workgroup runs (AxisPass::swz)
__device__ __forceinline__ int64_t pass_tile(const AxisPass &a) {
    const uint64_t b = blockIdx.x;
    if (a.swz > 1) return (int64_t)((b * (uint64_t)a.swz) % (uint64_t)gridDim.x);
    if (a.swz == SWZ_XCD) {
        // workgroup b runs on XCD b mod 8 as that XCD's (b / 8)-th; XCD k
        // takes tiles [k q + min(k, r), ...) with G = 8 q + r: a bijection on [0, G)
        const uint32_t G = gridDim.x, q = G >> 3, r = G & 7, k = (uint32_t)b & 7, j = (uint32_t)b >> 3;
        return (int64_t)k * q + (k < r ? k : r) + j;
    }
    return (int64_t)b;
}

// Thread -> (line slot ls, thread-in-line t) and the tile's global bases.
template <bool CONTIG>
__device__ __forceinline__ void tile_map(const AxisPass &a, int64_t tile, int tid, int TPL, int &ls, int &t,
                                         bool &valid, int64_t &ib, int64_t &ob) {
    const int TC = a.TC;
    const bool active = tid < TPL * TC;
    if constexpr (CONTIG) {
        ls = tid / TPL;
        t = tid - ls * TPL;
    } else {
        ls = tid % TC;
        t = tid / TC;
    }
    if (!active) {
        ls = 0;
        t = 1 << 28;  // beyond every butterfly / bin index: skips all work
    }
    if constexpr (CONTIG) {
        const int64_t line = tile * TC + ls;
        valid = active && line < a.nlines;
        ib = line * a.in_so;
        ob = line * a.out_so;
    } else {
        // The columns of all outer slices form one flat index f = tile*TC + ls
        // (column f mod nlines of slice f / nlines), so a row whose length is
        // not a multiple of TC (the n/2+1 columns of an R2C array) leaves no
        // ragged tile: a tile may end one slice and begin the next.
        const int64_t f0 = tile * TC;
        int64_t o = f0 / a.nlines;
        int64_t col = f0 - o * a.nlines + ls;
        // TC < 2*nlines (the planner halves TC until it is), so col < 3*nlines
        while (col >= a.nlines) {
            col -= a.nlines;
            ++o;
        }
        valid = active && o < a.nouter;
        ib = col + o * a.in_so;
        ob = col + o * a.out_so;
    }
}

// R2C post-processing of the H-point FFT Z that the last stage left in LDS
// (natural order): X_k = 1/2 [(Z_k + conj Z_{H-k}) - i w^k (Z_k - conj Z_{H-k})],
// k = 0..H.
template <class T, bool CONTIG>
__device__ __forceinline__ void r2c_post_lds(const TileCtx<T, OP_R2C, CONTIG> &c, int TPL) {
    __syncthreads();
    const int H = c.a.sp.L;
    const T s = (T)c.a.scale;
    cx<T> *outc = reinterpret_cast<cx<T> *>(c.out);
    const cx<T> *pw = c.post();
    for (int k = c.t; k <= H; k += TPL) {
        cx<T> zk = c.lds[c.li(k == H ? 0 : k)];
        cx<T> zc = conj(c.lds[c.li(k == 0 ? 0 : H - k)]);
        cx<T> A = zk + zc, B = zk - zc;
        cx<T> X = scl(A + mul_neg_i(cmul(pw[k], B)), (T)0.5 * s);
        if (c.valid) gstore(outc + c.ob + (int64_t)k * c.a.out_sm, X, c.a.nt);
    }
}

template <int EMAX, class T, int OP, bool CONTIG, bool SPLIT = false, class RL = NoRadices, int TPLC = 0>
__device__ __forceinline__ void axis_tile(const AxisPass &a, const void *in, void *out, int64_t tile, int tid,
                                          cx<T> *lds) {
    constexpr bool FIXED = !std::is_same<RL, NoRadices>::value;
    const int TPL = FIXED ? TPLC : a.TPL;
    int ls, t;
    bool valid;
    int64_t ib, ob;
    tile_map<CONTIG>(a, tile, tid, TPL, ls, t, valid, ib, ob);
    TileCtx<T, OP, CONTIG> c{a, in, out, lds, ls, t, valid, ib, ob};
    const StagePlan &sp = a.sp;

    if (sp.S == 0) {  // length-1 transform: only the fused load/store maps apply
        if (t == 0) c.store_last(0, c.load0(0));
    }
    if constexpr (FIXED) {
        run_fixed_stages<RL, EMAX, TPLC, SPLIT>(c);
    } else if (sp.S == 1) {
        dispatch_stage<SK_ONLY, EMAX>(c, sp.R[0], 1, sp.tw[0]);
    } else if (sp.S > 1) {
        dispatch_stage<SK_FIRST, EMAX>(c, sp.R[0], 1, sp.tw[0]);
        int Ns = sp.R[0];
        for (int s = 1; s < sp.S - 1; ++s) {
            dispatch_stage<SK_MID, EMAX>(c, sp.R[s], Ns, sp.tw[s]);
            Ns *= sp.R[s];
        }
        dispatch_stage<SK_LAST, EMAX>(c, sp.R[sp.S - 1], Ns, sp.tw[sp.S - 1]);
    }

    if constexpr (OP == OP_R2C && !(FIXED && SPLIT)) r2c_post_lds(c, TPL);
    if constexpr (OP == OP_DCT2H) dct2h_post_lds(c, TPL);
}

// ---------------------------------------------------------------------------
// Bluestein (chirp-z) tiles, for line lengths with a prime factor > 7
// (Bluestein 1970; Rabiner, Schafer & Rader 1969).  With c_m = exp(-i pi m^2/n)
// and jk = (j^2 + k^2 - (k-j)^2)/2, the length-n DFT the op's maps ask for is
//   X_k = c_k sum_j (x_j c_j) conj(c_{k-j}),
// a cyclic convolution of length M >= 2n-1 with b_m = conj(c_m) = b_{M-m}.
// Phase 0 is an M-point FFT of a_j = x_j c_j (zero for j >= n) whose last
// stage leaves conj(A_p bhat_p) in LDS, bhat = FFT_M(b) / M.  Phase 1 is an
// M-point FFT of that; the conjugate of its output is the convolution
// (IFFT_M(y) = conj(FFT_M(conj y)) / M), and c_k times it is X_k for k < n.
// The op's own load0/store_last give x_j and take X_k, so every fused op
// (R2C/C2R packing, the r2r kinds) runs unchanged around the chirp-z core.
// ---------------------------------------------------------------------------
template <class T, int OP, bool CONTIG, int PHASE>
struct BlueCtx : TileCtx<T, OP, CONTIG> {
    typedef TileCtx<T, OP, CONTIG> Op;  // the op's maps; a.sp.L = n
    // Loaded once: reading them through `a` inside the stage code lets the
    // compiler merge the radix cases' tails over pointers into the kernel
    // argument, which then gets copied to scratch.
    const cx<T> *chirp, *bhat;
    int n, M;

    __device__ __forceinline__ BlueCtx(const Op &c)
        : Op(c), chirp(c.twb() + c.a.chirp), bhat(c.twb() + c.a.bhat), n(c.a.sp.L), M(c.a.bsp.L) {}
    __device__ __forceinline__ int len() const { return M; }
    // stage twiddle blocks of the compile-time engine: the first FFT's
    // schedule, then its reverse for the second
    __device__ __forceinline__ int stage_tw(int s) const {
        return PHASE == 0 ? this->a.bsp.tw[s] : this->a.bsp2.tw[s];
    }

    __device__ __forceinline__ cx<T> load0(int p) const {  // phase 0 only
        if (p >= n) return cx<T>{0, 0};
        return cmul(Op::load0(p), chirp[p]);
    }
    __device__ __forceinline__ void store_last(int p, cx<T> v) const {
        if constexpr (PHASE == 0) {
            this->lds[this->li(p)] = conj(cmul(v, bhat[p]));
        } else {
            if (p < n) Op::store_last(p, cmul(conj(v), chirp[p]));
        }
    }
};

template <int EMAX, bool LDS_IN, class C>
__device__ __forceinline__ void blue_fft(const C &c, const StagePlan &bp) {
    // the planner gives M >= 2n-1 >= 21, so at least two stages
    dispatch_stage<LDS_IN ? SK_FIRST_LDS : SK_FIRST, EMAX>(c, bp.R[0], 1, bp.tw[0]);
    int Ns = bp.R[0];
    for (int s = 1; s < bp.S - 1; ++s) {
        dispatch_stage<SK_MID, EMAX>(c, bp.R[s], Ns, bp.tw[s]);
        Ns *= bp.R[s];
    }
    dispatch_stage<SK_LAST, EMAX>(c, bp.R[bp.S - 1], Ns, bp.tw[bp.S - 1]);
}

template <int EMAX, class T, int OP, bool CONTIG>
__device__ __forceinline__ void blue_tile(const AxisPass &a, const void *in, void *out, int64_t tile, int tid,
                                          cx<T> *lds) {
    int ls, t;
    bool valid;
    int64_t ib, ob;
    tile_map<CONTIG>(a, tile, tid, a.TPL, ls, t, valid, ib, ob);
    TileCtx<T, OP, CONTIG> c{a, in, out, lds, ls, t, valid, ib, ob};
    blue_fft<EMAX, false>(BlueCtx<T, OP, CONTIG, 0>(c), a.bsp);
    // phase 0's last stage wrote LDS after its own barrier; phase 1 reads it
    __syncthreads();
    blue_fft<EMAX, true>(BlueCtx<T, OP, CONTIG, 1>(c), a.bsp);
    if constexpr (OP == OP_R2C) r2c_post_lds(c, a.TPL);
}

// A radix list reversed (a valid Stockham schedule of the same length).
template <class RL, class Seq>
struct ReversedImpl;
template <class RL, int... I>
struct ReversedImpl<RL, std::integer_sequence<int, I...>> {
    using type = Radices<RL::R[RL::S - 1 - I]...>;
};
template <class RL>
using Reversed = typename ReversedImpl<RL, std::make_integer_sequence<int, RL::S>>::type;

// Between the two FFTs of a compile-time Bluestein tile.  The first FFT's
// last stage (radix R) leaves output p = j + r*M/R of butterfly j in
// v[b*R + r]; the second FFT runs the reversed radix list, so its first
// stage has the same radix and reads exactly those positions into the same
// registers: conj(A_p * bhat_p) is formed in place, with no LDS round trip.
template <int R, int EMAX, int LC, int TPLC, class C>
__device__ __forceinline__ void blue_turn(const C &c, cx<typename C::value_type> *v) {
    constexpr int NBM = EMAX / R;
    const int NB = LC / R;
#pragma unroll
    for (int b = 0; b < NBM; ++b) {
        const int j = c.t + b * TPLC;
        if (j < NB) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b * R + r] = conj(cmul(v[b * R + r], c.bhat[j + r * NB]));
        }
    }
}

// Bluestein tile on the compile-time engine: convolution length RL::L,
// 16 values per thread, around the op's load/store maps.  SPLIT: the
// exchanges move real and imaginary parts in two rounds through half the
// LDS; the packed R2C op needs whole complex values in LDS for its
// post-twiddle, so it runs the full exchange.
template <class RL, class T, int OP, bool CONTIG, bool SPLIT>
__device__ __forceinline__ void blue_fixed_tile(const AxisPass &a, const void *in, void *out, int64_t tile, int tid,
                                                cx<T> *lds) {
    static_assert(!(OP == OP_R2C && SPLIT), "the packed R2C post-twiddle reads complex LDS slots");
    constexpr int EMAX = 16, TPLC = fixed_tpl<RL, EMAX>(), M = RL::L;
    using RR = Reversed<RL>;
    int ls, t;
    bool valid;
    int64_t ib, ob;
    tile_map<CONTIG>(a, tile, tid, TPLC, ls, t, valid, ib, ob);
    TileCtx<T, OP, CONTIG> c{a, in, out, lds, ls, t, valid, ib, ob};
    const BlueCtx<T, OP, CONTIG, 0> c0(c);
    const BlueCtx<T, OP, CONTIG, 1> c1(c);
    cx<T> v[EMAX];
    st_load_global<RL::R[0], EMAX, M, TPLC>(c0, v);
    st_compute<RL::R[0], EMAX, M, TPLC, 1>(c0, v, 0);
    fixed_stages_from<RL, EMAX, TPLC, SPLIT, 0>(c0, v);
    blue_turn<RL::R[RL::S - 1], EMAX, M, TPLC>(c0, v);
    st_compute<RR::R[0], EMAX, M, TPLC, 1>(c1, v, 0);
    fixed_stages_from<RR, EMAX, TPLC, SPLIT, 0>(c1, v);
    st_store_global<RR::R[RR::S - 1], EMAX, M, TPLC>(c1, v);
    if constexpr (OP == OP_R2C) r2c_post_lds(c, TPLC);
}

// ---------------------------------------------------------------------------
// Rader tiles, for prime line lengths n whose n-1 is 7-smooth (Rader 1968).
// With a primitive root g mod n, j = g^q and k = g^-p turn the DFT into
//   X_{g^-p} = x_0 + sum_{q<n-1} x_{g^q} w_n^{g^(q-p)} = x_0 + (a (*) b)_p,
// a cyclic convolution of length M = n-1 (no zero padding) of
// a_q = x_{g^q} and b_m = w_n^{g^-m}, and X_0 = x_0 + sum_q a_q = x_0 + A_0.
// The two M-point FFTs run as in the compile-time Bluestein tile (the
// spectrum product in registers between them); the chirp block holds the
// index pairs (g^p mod n, g^-p mod n) and bhat holds FFT_M(b) / M.  M is
// about n instead of 2n-1 or more, so a Rader pass does 2-2.5x less FFT work
// than the Bluestein pass of the same n.  C2C only.
// ---------------------------------------------------------------------------
template <class T, int OP, bool CONTIG, int PHASE>
struct RaderCtx : TileCtx<T, OP, CONTIG> {
    typedef TileCtx<T, OP, CONTIG> Op;  // the op's maps; a.sp.L = n
    const cx<T> *perm, *bhat;           // perm[p] = (g^p mod n, g^-p mod n) as values of T
    cx<T> x0;                           // the line's x_0 as the op loads it

    __device__ __forceinline__ RaderCtx(const Op &c, cx<T> x0_)
        : Op(c), perm(c.twb() + c.a.chirp), bhat(c.twb() + c.a.bhat), x0(x0_) {}
    __device__ __forceinline__ int len() const { return this->a.bsp.L; }
    __device__ __forceinline__ int stage_tw(int s) const {
        return PHASE == 0 ? this->a.bsp.tw[s] : this->a.bsp2.tw[s];
    }
    __device__ __forceinline__ cx<T> load0(int p) const {  // phase 0: a_p = x_{g^p}
        return Op::load0((int)perm[p].x);
    }
    __device__ __forceinline__ void store_last(int p, cx<T> v) const {  // phase 1: X_{g^-p}
        Op::store_last((int)perm[p].y, x0 + conj(v));
    }
};

// A contiguous line's permuted reads and writes, staged through LDS (complex
// values; the tile's LDS is sized for them): the lines move between global
// memory and LDS in natural order, coalesced, and the permutations x_{g^q} /
// X_{g^-p} become LDS addressing.  CL false: columnar staging (the strided
// layout, element p of line ls at p * P + ls): lanes run across the tile's
// lines, so each gather, scatter and FFT exchange touches consecutive LDS
// words.  CL true: a padded slot per line (lines of more than 64 threads,
// whose tiles hold too few lines for the columnar layout; the permuted
// accesses take bank conflicts there).
template <class T, int OP, bool CL, int PHASE>
struct RaderStagedCtx : RaderCtx<T, OP, CL, PHASE> {
    typedef RaderCtx<T, OP, CL, PHASE> Base;
    __device__ __forceinline__ RaderStagedCtx(const typename Base::Op &c, cx<T> x0_) : Base(c, x0_) {}
    __device__ __forceinline__ cx<T> load0(int p) const { return this->lds[this->li((int)this->perm[p].x)]; }
    __device__ __forceinline__ void store_last(int p, cx<T> v) const {
        this->lds[this->li((int)this->perm[p].y)] = this->x0 + conj(v);
    }
};

// Stage-0 loads of a Rader tile read straight from global memory: the
// positions st_load_global reads, a_p = x_{g^p}, in three steps so that no
// load waits for another -- every permuted offset g^p * stride first (the
// index table), then every data load (positions past the line and lines past
// the array's end read element 0 of the array; st_compute skips those
// positions and such lines are never stored), then the inverse transform's
// conjugate.  The C2C load map of TileCtx::load0, reordered.
template <int R, int EMAX, int LC, int TPLC, class C>
__device__ __forceinline__ void rader_load_global(const C &c, cx<typename C::value_type> *v) {
    typedef typename C::value_type T;
    constexpr int NBM = EMAX / R;
    const int NB = LC / R;
    const cx<T> *inc = reinterpret_cast<const cx<T> *>(c.in);
    int64_t off[NBM * R];
#pragma unroll
    for (int b = 0; b < NBM; ++b) {
        const int j = c.t + b * TPLC;
#pragma unroll
        for (int r = 0; r < R; ++r)
            off[b * R + r] = (c.valid && j < NB) ? c.ib + (int64_t)(int)c.perm[j + r * NB].x * c.a.in_sm : 0;
    }
#pragma unroll
    for (int e = 0; e < NBM * R; ++e) v[e] = gload(inc + off[e], c.a.nt);
    if (c.a.inverse) {
#pragma unroll
        for (int e = 0; e < NBM * R; ++e) v[e] = conj(v[e]);
    }
}

template <class RL, class T, int OP, bool CONTIG, bool SPLIT>
__device__ __forceinline__ void rader_fixed_tile(const AxisPass &a, const void *in, void *out, int64_t tile, int tid,
                                                 cx<T> *lds) {
    static_assert(OP == OP_C2C, "Rader passes are C2C");
    constexpr int EMAX = 16, TPLC = fixed_tpl<RL, EMAX>(), M = RL::L, N = M + 1;
    using RR = Reversed<RL>;
    cx<T> v[EMAX];
    if constexpr (CONTIG && TPLC <= 64) {
        // Columnar tile of TC contiguous lines: thread tid works on line
        // ls = tid % TC as its thread t = tid / TC, and element i of line ls
        // sits in LDS at i * P + ls (P = a.lds_pitch, odd and >= TC; the
        // planner sizes the LDS for N * P complex values).
        const int TC = a.TC, P = a.lds_pitch, NT = TPLC * TC;
        const bool active = tid < NT;
        const int ls = active ? tid % TC : 0;
        const int t = active ? tid / TC : (1 << 28);  // inactive: beyond every butterfly
        const int64_t line0 = tile * (int64_t)TC;
        const bool valid = active && line0 + ls < a.nlines;
        // the op's load and store maps of line `line` of the tile
        auto line_ctx = [&](int line) {
            const int64_t g = line0 + line;
            return TileCtx<T, OP, true>{a, in, out, lds, line, 0, g < a.nlines, g * a.in_so, g * a.out_so};
        };
        // Element f of the tile's flattened lines (f < TC * N) is copied by
        // thread f % NT as its copy f / NT: NPT = ceil(N / TPLC) copies per
        // thread, a compile-time count.  The loads are the C2C load map
        // (TileCtx::load0) split in two: every raw load is issued first, at
        // an in-bounds address (copies past the tile and lines past the
        // array's end read element 0; those lines are never stored), and the
        // inverse transform's conjugate is applied as the values go to LDS,
        // so the loads of a group of CH copies are in flight together (groups
        // of 8 keep the staging registers below the kernel's 128-VGPR bound).
        constexpr int NPT = (N + TPLC - 1) / TPLC, CH = 8;
        // the tile's lines into LDS in natural order, flattened over the lines
        if (active) {
            const cx<T> *inc = reinterpret_cast<const cx<T> *>(in);
#pragma unroll
            for (int k0 = 0; k0 < NPT; k0 += CH) {
                cx<T> s[CH];
#pragma unroll
                for (int kk = 0; kk < CH && k0 + kk < NPT; ++kk) {
                    const int f = tid + (k0 + kk) * NT, line = f / N, i = f - line * N;
                    const int64_t g = line0 + line;
                    const int64_t off = (f < TC * N && g < a.nlines) ? g * a.in_so + (int64_t)i * a.in_sm : 0;
                    s[kk] = gload(inc + off, a.nt);
                }
#pragma unroll
                for (int kk = 0; kk < CH && k0 + kk < NPT; ++kk) {
                    const int f = tid + (k0 + kk) * NT, line = f / N, i = f - line * N;
                    if (f < TC * N) lds[i * P + line] = a.inverse ? conj(s[kk]) : s[kk];
                }
            }
        }
        __syncthreads();
        const TileCtx<T, OP, false> c{a, in, out, lds, ls, t, valid, 0, 0};  // LDS: p * P + ls
        const cx<T> x0 = active ? lds[ls] : cx<T>{0, 0};
        const RaderStagedCtx<T, OP, false, 0> c0(c, x0);
        const RaderStagedCtx<T, OP, false, 1> c1(c, x0);
        st_load_global<RL::R[0], EMAX, M, TPLC>(c0, v);
        __syncthreads();  // the first exchange reuses the staging area
        st_compute<RL::R[0], EMAX, M, TPLC, 1>(c0, v, 0);
        fixed_stages_from<RL, EMAX, TPLC, SPLIT, 0>(c0, v);
        const cx<T> a0 = v[0];
        blue_turn<RL::R[RL::S - 1], EMAX, M, TPLC>(c0, v);
        st_compute<RR::R[0], EMAX, M, TPLC, 1>(c1, v, 0);
        fixed_stages_from<RR, EMAX, TPLC, SPLIT, 0>(c1, v);
        // the last exchange ended with a barrier (or, with a one-stage FFT, the
        // barrier after the gather): X into LDS at g^-p, then out in natural order
        st_store_global<RR::R[RR::S - 1], EMAX, M, TPLC>(c1, v);
        if (t == 0) lds[ls] = x0 + a0;
        __syncthreads();
        if (active) {
#pragma unroll
            for (int k = 0; k < NPT; ++k) {
                const int f = tid + k * NT, line = f / N, i = f - line * N;
                if (f < TC * N) {
                    const auto lc = line_ctx(line);
                    if (lc.valid) lc.store_last(i, lds[i * P + line]);
                }
            }
        }
    } else if constexpr (CONTIG) {
        int ls, t;
        bool valid;
        int64_t ib, ob;
        tile_map<CONTIG>(a, tile, tid, TPLC, ls, t, valid, ib, ob);
        TileCtx<T, OP, CONTIG> c{a, in, out, lds, ls, t, valid, ib, ob};
        // the whole line into its LDS slot in natural order
        constexpr int NPT = (N + TPLC - 1) / TPLC;
#pragma unroll
        for (int k = 0; k < NPT; ++k) {
            const int i = t + k * TPLC;
            if (i < N) lds[c.li(i)] = c.load0(i);
        }
        __syncthreads();
        const cx<T> x0 = t < (1 << 28) ? lds[c.li(0)] : cx<T>{0, 0};
        const RaderStagedCtx<T, OP, true, 0> c0(c, x0);
        const RaderStagedCtx<T, OP, true, 1> c1(c, x0);
        st_load_global<RL::R[0], EMAX, M, TPLC>(c0, v);
        __syncthreads();  // the first exchange reuses the slot
        st_compute<RL::R[0], EMAX, M, TPLC, 1>(c0, v, 0);
        fixed_stages_from<RL, EMAX, TPLC, SPLIT, 0>(c0, v);
        const cx<T> a0 = v[0];
        blue_turn<RL::R[RL::S - 1], EMAX, M, TPLC>(c0, v);
        st_compute<RR::R[0], EMAX, M, TPLC, 1>(c1, v, 0);
        fixed_stages_from<RR, EMAX, TPLC, SPLIT, 0>(c1, v);
        // the last exchange ended with a barrier: X into the slot at g^-p, then out in natural order
        st_store_global<RR::R[RR::S - 1], EMAX, M, TPLC>(c1, v);
        if (t == 0) lds[c.li(0)] = x0 + a0;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NPT; ++k) {
            const int i = t + k * TPLC;
            if (i < N) c.store_last(i, lds[c.li(i)]);
        }
    } else {
        int ls, t;
        bool valid;
        int64_t ib, ob;
        tile_map<CONTIG>(a, tile, tid, TPLC, ls, t, valid, ib, ob);
        TileCtx<T, OP, CONTIG> c{a, in, out, lds, ls, t, valid, ib, ob};
        const cx<T> x0 = t < (1 << 28) ? c.load0(0) : cx<T>{0, 0};
        const RaderCtx<T, OP, CONTIG, 0> c0(c, x0);
        const RaderCtx<T, OP, CONTIG, 1> c1(c, x0);
        // FP64: the reordered loads (3-4% faster); FP32: its offsets cost a
        // wave of occupancy on the large radices, which measured 4% slower
        if constexpr (sizeof(T) == 8) rader_load_global<RL::R[0], EMAX, M, TPLC>(c0, v);
        else st_load_global<RL::R[0], EMAX, M, TPLC>(c0, v);
        st_compute<RL::R[0], EMAX, M, TPLC, 1>(c0, v, 0);
        fixed_stages_from<RL, EMAX, TPLC, SPLIT, 0>(c0, v);
        // position 0 of the first FFT's output (A_0 = sum of a) sits in v[0]
        // of the line's thread 0
        const cx<T> a0 = v[0];
        blue_turn<RL::R[RL::S - 1], EMAX, M, TPLC>(c0, v);
        st_compute<RR::R[0], EMAX, M, TPLC, 1>(c1, v, 0);
        fixed_stages_from<RR, EMAX, TPLC, SPLIT, 0>(c1, v);
        st_store_global<RR::R[RR::S - 1], EMAX, M, TPLC>(c1, v);
        if (t == 0) c.store_last(0, x0 + a0);
    }
}

}  // namespace dfft
