/*
 * dfft_oracle.c -- CPU oracle for the DaggerFFT 3D transform path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, calls or
 * loads this file; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker.
 *
 * What it restates (reference files are Julia; FFTW/cuFFT do the arithmetic):
 *   - per-axis batched 1D transforms, the work of apply_fft!/mul! on one chunk
 *     (src/fft.jl:126-174, src/fftgpu.jl:83-124).  FFTW.jl and CUDA.jl are
 *     unpinned in Project.toml:4-18; their published definitions are restated
 *     here: forward X_k = sum_j x_j exp(-2 pi i jk/n), unnormalised; backward
 *     with exp(+...); Julia's ifft scales by 1/n per transformed dim.
 *   - r2c / c2r along the first transformed dim (FFTW rfft/irfft semantics,
 *     README.md:38-44): output length n/2+1, c2r ignores the imaginary parts
 *     of the DC (and, for even n, the Nyquist) bin.
 *   - the r2r kinds listed in R2R_SUPPORTED_KINDS (src/fft.jl:1-11) by their
 *     FFTW-manual definitions, summed directly in O(n^2) (independent of the
 *     FFT-based method the GPU path uses).
 *   - the box gather / scatter / copy of pack_data!, unpack_data!, local_copy!
 *     (src/fft.jl:180-231) and of the GPU kernels pack_kernel!,
 *     unpack_kernel!, local_copy_kernel! (src/fftgpu.jl:452-495), 0-based.
 *   - the project's counter-based SplitMix64 input generator.
 *   - spot bins of a 3D DFT by direct separable summation in long double,
 *     the full-size check that needs no FFT (dfo_spot_bins).
 *
 * Arrays are column-major (Julia layout): element (i,j,k) of an n0 x n1 x n2
 * array sits at i + n0*(j + n1*k).  Complex data is interleaved (re, im).
 * The 1D FFT is a textbook recursive mixed-radix decimation-in-time FFT
 * (Cooley-Tukey 1965) in double arithmetic with twiddles evaluated in long
 * double and rounded once.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cd;

/* ------------------------------------------------------------------ */
/* Input generator: SplitMix64 (Steele, Lea, Flood 2014) in counter form. */
/* ------------------------------------------------------------------ */
static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t dfo_splitmix64(uint64_t seed, uint64_t i) {
    return mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ULL);
}

double dfo_uniform(uint64_t seed, uint64_t i) {
    return (double)(dfo_splitmix64(seed, i) >> 11) * 0x1.0p-53;
}

/* real element e -> u(e); complex element e -> (u(2e), u(2e+1)) */
void dfo_fill_f64(uint64_t seed, int64_t n, double *out) {
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < n; ++e) out[e] = dfo_uniform(seed, (uint64_t)e);
}
void dfo_fill_c128(uint64_t seed, int64_t n, double *out) {
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < 2 * n; ++e) out[e] = dfo_uniform(seed, (uint64_t)e);
}
void dfo_fill_f32(uint64_t seed, int64_t n, float *out) {
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < n; ++e) out[e] = (float)dfo_uniform(seed, (uint64_t)e);
}
void dfo_fill_c64(uint64_t seed, int64_t n, float *out) {
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < 2 * n; ++e) out[e] = (float)dfo_uniform(seed, (uint64_t)e);
}

/* ------------------------------------------------------------------ */
/* exp(-2 pi i k / n) with octant folding in long double               */
/* ------------------------------------------------------------------ */
static void cis_ld(int64_t k, int64_t n, long double *c, long double *s) {
    /* returns cos(2 pi k/n), sin(2 pi k/n) */
    k %= n;
    if (k < 0) k += n;
    long double x = (long double)k / (long double)n; /* in [0,1) */
    int neg_s = 0;
    if (x > 0.5L) { x = 1.0L - x; neg_s = 1; }     /* sin(2pi(1-x)) = -sin */
    int neg_c = 0;
    if (x > 0.25L) { x = 0.5L - x; neg_c = 1; }    /* cos(pi - a) = -cos a */
    long double cc, ss;
    if (x > 0.125L) {
        long double y = (0.25L - x) * 6.283185307179586476925286766559005768L;
        cc = sinl(y); ss = cosl(y);
    } else {
        long double y = x * 6.283185307179586476925286766559005768L;
        cc = cosl(y); ss = sinl(y);
    }
    *c = neg_c ? -cc : cc;
    *s = neg_s ? -ss : ss;
}

/* table of exp(sign * 2 pi i k / n), k in [0,n) */
static cd *make_twiddles(int64_t n, int sign) {
    cd *t = (cd *)malloc(sizeof(cd) * (size_t)(n > 0 ? n : 1));
    for (int64_t k = 0; k < n; ++k) {
        long double c, s;
        cis_ld(k, n, &c, &s);
        t[k].re = (double)c;
        t[k].im = (double)(sign < 0 ? -s : s);
    }
    return t;
}

static int smallest_factor(int64_t n) {
    if (n % 4 == 0 && n > 4) return 4;
    if (n % 2 == 0) return 2;
    for (int64_t p = 3; p * p <= n; p += 2)
        if (n % p == 0) return (int)p;
    return (int)n;
}

/* out[0..n) = DFT of in[0], in[is], ..., in[(n-1) is]
 * tw: table of w^k, w = exp(sign 2 pi i / ntop); this level uses w^(tws*k). */
static void fft_rec(const cd *in, cd *out, int64_t n, int64_t is,
                    const cd *tw, int64_t ntop, int64_t tws) {
    if (n == 1) { out[0] = in[0]; return; }
    int p = smallest_factor(n);
    int64_t m = n / p;
    for (int r = 0; r < p; ++r)
        fft_rec(in + r * is, out + r * m, m, is * p, tw, ntop, tws * p);
    cd t[64];
    cd *tt = p <= 64 ? t : (cd *)malloc(sizeof(cd) * (size_t)p);
    for (int64_t k = 0; k < m; ++k) {
        for (int r = 0; r < p; ++r) {
            cd y = out[r * m + k];
            if (r == 0 || k == 0) { tt[r] = y; continue; }
            cd w = tw[((int64_t)r * k * tws) % ntop];
            tt[r].re = y.re * w.re - y.im * w.im;
            tt[r].im = y.re * w.im + y.im * w.re;
        }
        for (int q = 0; q < p; ++q) {
            double ar = 0.0, ai = 0.0;
            for (int r = 0; r < p; ++r) {
                cd w = tw[(((int64_t)r * q) % p) * m * tws % ntop];
                ar += tt[r].re * w.re - tt[r].im * w.im;
                ai += tt[r].re * w.im + tt[r].im * w.re;
            }
            out[k + m * q].re = ar;
            out[k + m * q].im = ai;
        }
    }
    if (tt != t) free(tt);
}

/* 1D DFT of a contiguous complex line (n elements), sign -1 forward / +1 backward */
void dfo_fft1d(const double *in, double *out, int64_t n, int sign) {
    cd *tw = make_twiddles(n, sign);
    fft_rec((const cd *)in, (cd *)out, n, 1, tw, n, 1);
    free(tw);
}

/* brute-force O(n^2) DFT accumulated in long double (for pinning) */
void dfo_dft_naive(const double *in, double *out, int64_t n, int sign) {
    for (int64_t k = 0; k < n; ++k) {
        long double ar = 0, ai = 0;
        for (int64_t j = 0; j < n; ++j) {
            long double c, s;
            cis_ld(j * k, n, &c, &s);
            if (sign < 0) s = -s;
            ar += (long double)in[2 * j] * c - (long double)in[2 * j + 1] * s;
            ai += (long double)in[2 * j] * s + (long double)in[2 * j + 1] * c;
        }
        out[2 * k] = (double)ar;
        out[2 * k + 1] = (double)ai;
    }
}

/* ------------------------------------------------------------------ */
/* line enumeration over a column-major rank<=3 array                  */
/* ------------------------------------------------------------------ */
typedef struct { int64_t n[3]; int64_t st[3]; } geom;

static geom make_geom(const int64_t *shape) {
    geom g;
    for (int d = 0; d < 3; ++d) g.n[d] = shape[d] > 0 ? shape[d] : 1;
    g.st[0] = 1; g.st[1] = g.n[0]; g.st[2] = g.n[0] * g.n[1];
    return g;
}

/* base offset (in elements) of line l along axis a */
static inline int64_t line_base(const geom *g, int a, int64_t l) {
    int o1 = a == 0 ? 1 : 0, o2 = a == 2 ? 1 : 2;
    int64_t i1 = l % g->n[o1], i2 = l / g->n[o1];
    return i1 * g->st[o1] + i2 * g->st[o2];
}
static inline int64_t n_lines(const geom *g, int a) {
    return (g->n[0] * g->n[1] * g->n[2]) / g->n[a];
}

/* In-place batched C2C along axis (0-based), scaled by `scale`. */
int dfo_c2c_axis(double *a, const int64_t *shape, int axis, int sign, double scale) {
    if (axis < 0 || axis > 2) return 1;
    geom g = make_geom(shape);
    int64_t n = g.n[axis], st = g.st[axis], nl = n_lines(&g, axis);
    cd *tw = make_twiddles(n, sign);
    cd *A = (cd *)a;
#pragma omp parallel
    {
        cd *bi = (cd *)malloc(sizeof(cd) * (size_t)n), *bo = (cd *)malloc(sizeof(cd) * (size_t)n);
#pragma omp for schedule(static)
        for (int64_t l = 0; l < nl; ++l) {
            int64_t b = line_base(&g, axis, l);
            for (int64_t m = 0; m < n; ++m) bi[m] = A[b + m * st];
            fft_rec(bi, bo, n, 1, tw, n, 1);
            for (int64_t m = 0; m < n; ++m) {
                A[b + m * st].re = bo[m].re * scale;
                A[b + m * st].im = bo[m].im * scale;
            }
        }
        free(bi); free(bo);
    }
    free(tw);
    return 0;
}

/* R2C along axis: real array of `shape` -> complex array with shape[axis] -> n/2+1 */
int dfo_r2c_axis(const double *in, double *out, const int64_t *shape, int axis) {
    if (axis < 0 || axis > 2) return 1;
    geom gi = make_geom(shape);
    int64_t n = gi.n[axis], h = n / 2 + 1;
    int64_t so[3] = {gi.n[0], gi.n[1], gi.n[2]};
    so[axis] = h;
    geom go = make_geom(so);
    int64_t nl = n_lines(&gi, axis);
    cd *tw = make_twiddles(n, -1);
    cd *O = (cd *)out;
#pragma omp parallel
    {
        cd *bi = (cd *)malloc(sizeof(cd) * (size_t)n), *bo = (cd *)malloc(sizeof(cd) * (size_t)n);
#pragma omp for schedule(static)
        for (int64_t l = 0; l < nl; ++l) {
            int64_t b = line_base(&gi, axis, l), bo_ = line_base(&go, axis, l);
            for (int64_t m = 0; m < n; ++m) { bi[m].re = in[b + m * gi.st[axis]]; bi[m].im = 0.0; }
            fft_rec(bi, bo, n, 1, tw, n, 1);
            for (int64_t m = 0; m < h; ++m) O[bo_ + m * go.st[axis]] = bo[m];
        }
        free(bi); free(bo);
    }
    free(tw);
    return 0;
}

/* C2R along axis: complex (shape with axis -> n/2+1) -> real `shape`, times scale.
 * Imaginary parts of bin 0 (and bin n/2 for even n) are ignored. */
int dfo_c2r_axis(const double *in, double *out, const int64_t *shape, int axis, double scale) {
    if (axis < 0 || axis > 2) return 1;
    geom go = make_geom(shape);
    int64_t n = go.n[axis], h = n / 2 + 1;
    int64_t si[3] = {go.n[0], go.n[1], go.n[2]};
    si[axis] = h;
    geom gi = make_geom(si);
    int64_t nl = n_lines(&go, axis);
    cd *tw = make_twiddles(n, +1);
    const cd *I = (const cd *)in;
#pragma omp parallel
    {
        cd *bi = (cd *)malloc(sizeof(cd) * (size_t)n), *bo = (cd *)malloc(sizeof(cd) * (size_t)n);
#pragma omp for schedule(static)
        for (int64_t l = 0; l < nl; ++l) {
            int64_t b = line_base(&gi, axis, l), bo_ = line_base(&go, axis, l);
            for (int64_t m = 0; m < h; ++m) bi[m] = I[b + m * gi.st[axis]];
            bi[0].im = 0.0;
            if (n % 2 == 0) bi[n / 2].im = 0.0;
            for (int64_t m = h; m < n; ++m) { bi[m].re = bi[n - m].re; bi[m].im = -bi[n - m].im; }
            fft_rec(bi, bo, n, 1, tw, n, 1);
            for (int64_t m = 0; m < n; ++m) out[bo_ + m * go.st[axis]] = bo[m].re * scale;
        }
        free(bi); free(bo);
    }
    free(tw);
    return 0;
}

/* ------------------------------------------------------------------ */
/* R2R kinds, numbered as FFTW's fftw_r2r_kind:                         */
/* R2HC=0 HC2R=1 DHT=2 REDFT00=3 REDFT01=4 REDFT10=5 REDFT11=6          */
/* RODFT00=7 RODFT01=8 RODFT10=9 RODFT11=10                             */
/* ------------------------------------------------------------------ */
enum { K_DHT = 2, K_REDFT00 = 3, K_REDFT01 = 4, K_REDFT10 = 5, K_REDFT11 = 6,
       K_RODFT00 = 7, K_RODFT01 = 8, K_RODFT10 = 9, K_RODFT11 = 10 };

/* one line: y = R2R_kind(x), direct summation in long double */
static int r2r_line(const double *x, double *y, int64_t n, int kind,
                    const double *cosT, const double *sinT, int64_t P) {
    for (int64_t k = 0; k < n; ++k) {
        long double acc = 0.0L;
        switch (kind) {
        case K_DHT:
            for (int64_t j = 0; j < n; ++j) {
                int64_t id = (j * k) % P;
                acc += (long double)x[j] * ((long double)cosT[id] + (long double)sinT[id]);
            }
            break;
        case K_REDFT00:
            if (n == 1) { acc = x[0]; break; }
            acc = (long double)x[0] + ((k & 1) ? -(long double)x[n - 1] : (long double)x[n - 1]);
            for (int64_t j = 1; j < n - 1; ++j) acc += 2.0L * x[j] * cosT[(j * k) % P];
            break;
        case K_REDFT10:
            for (int64_t j = 0; j < n; ++j) acc += 2.0L * x[j] * cosT[((2 * j + 1) * k) % P];
            break;
        case K_REDFT01:
            acc = x[0];
            for (int64_t j = 1; j < n; ++j) acc += 2.0L * x[j] * cosT[(j * (2 * k + 1)) % P];
            break;
        case K_REDFT11:
            for (int64_t j = 0; j < n; ++j) acc += 2.0L * x[j] * cosT[((2 * j + 1) * (2 * k + 1)) % P];
            break;
        case K_RODFT00:
            for (int64_t j = 0; j < n; ++j) acc += 2.0L * x[j] * sinT[((j + 1) * (k + 1)) % P];
            break;
        case K_RODFT10:
            for (int64_t j = 0; j < n; ++j) acc += 2.0L * x[j] * sinT[((2 * j + 1) * (k + 1)) % P];
            break;
        case K_RODFT01:
            acc = (k & 1) ? -(long double)x[n - 1] : (long double)x[n - 1];
            for (int64_t j = 0; j < n - 1; ++j) acc += 2.0L * x[j] * sinT[((j + 1) * (2 * k + 1)) % P];
            break;
        case K_RODFT11:
            for (int64_t j = 0; j < n; ++j) acc += 2.0L * x[j] * sinT[((2 * j + 1) * (2 * k + 1)) % P];
            break;
        default:
            return 1;
        }
        y[k] = (double)acc;
    }
    return 0;
}

/* period P such that the kernel's angle is 2 pi * idx / P */
static int64_t r2r_period(int kind, int64_t n) {
    switch (kind) {
    case K_DHT: return n;
    case K_REDFT00: return n > 1 ? 2 * (n - 1) : 1;
    case K_RODFT00: return 2 * (n + 1);
    case K_REDFT10: case K_REDFT01: case K_RODFT10: case K_RODFT01: return 4 * n;
    case K_REDFT11: case K_RODFT11: return 8 * n;
    default: return 0;
    }
}

/* In-place R2R of `kind` along axis, times scale. */
int dfo_r2r_axis(double *a, const int64_t *shape, int axis, int kind, double scale) {
    if (axis < 0 || axis > 2) return 1;
    geom g = make_geom(shape);
    int64_t n = g.n[axis], st = g.st[axis], nl = n_lines(&g, axis);
    int64_t P = r2r_period(kind, n);
    if (P == 0) return 1;
    double *cosT = (double *)malloc(sizeof(double) * (size_t)P);
    double *sinT = (double *)malloc(sizeof(double) * (size_t)P);
    for (int64_t i = 0; i < P; ++i) {
        long double c, s;
        cis_ld(i, P, &c, &s);
        cosT[i] = (double)c; sinT[i] = (double)s;
    }
    int bad = 0;
#pragma omp parallel
    {
        double *x = (double *)malloc(sizeof(double) * (size_t)n), *y = (double *)malloc(sizeof(double) * (size_t)n);
#pragma omp for schedule(static) reduction(|:bad)
        for (int64_t l = 0; l < nl; ++l) {
            int64_t b = line_base(&g, axis, l);
            for (int64_t m = 0; m < n; ++m) x[m] = a[b + m * st];
            bad |= r2r_line(x, y, n, kind, cosT, sinT, P);
            for (int64_t m = 0; m < n; ++m) a[b + m * st] = y[m] * scale;
        }
        free(x); free(y);
    }
    free(cosT); free(sinT);
    return bad;
}

/* ------------------------------------------------------------------ */
/* Box maps (src/fft.jl:180-231; src/fftgpu.jl:452-495), 0-based.       */
/* Elements are opaque `esize`-byte records: bit-exact copies.         */
/* ------------------------------------------------------------------ */

/* buffer[offset + i + nx*(j + ny*k)] = src[lo0+i, lo1+j, lo2+k]  (pack_data!) */
void dfo_pack(void *buffer, const void *src, const int64_t *src_shape,
              const int64_t *lo, const int64_t *cnt, int64_t offset, int64_t esize) {
    geom g = make_geom(src_shape);
    unsigned char *B = (unsigned char *)buffer;
    const unsigned char *S = (const unsigned char *)src;
    for (int64_t k = 0; k < cnt[2]; ++k)
        for (int64_t j = 0; j < cnt[1]; ++j)
            for (int64_t i = 0; i < cnt[0]; ++i) {
                int64_t bi = offset + i + cnt[0] * (j + cnt[1] * k);
                int64_t si = (lo[0] + i) + g.st[1] * (lo[1] + j) + g.st[2] * (lo[2] + k);
                memcpy(B + bi * esize, S + si * esize, (size_t)esize);
            }
}

/* dst[lo0+i, lo1+j, lo2+k] = buffer[offset + i + nx*(j + ny*k)]  (unpack_data!) */
void dfo_unpack(void *dst, const void *buffer, const int64_t *dst_shape,
                const int64_t *lo, const int64_t *cnt, int64_t offset, int64_t esize) {
    geom g = make_geom(dst_shape);
    unsigned char *D = (unsigned char *)dst;
    const unsigned char *B = (const unsigned char *)buffer;
    for (int64_t k = 0; k < cnt[2]; ++k)
        for (int64_t j = 0; j < cnt[1]; ++j)
            for (int64_t i = 0; i < cnt[0]; ++i) {
                int64_t bi = offset + i + cnt[0] * (j + cnt[1] * k);
                int64_t di = (lo[0] + i) + g.st[1] * (lo[1] + j) + g.st[2] * (lo[2] + k);
                memcpy(D + di * esize, B + bi * esize, (size_t)esize);
            }
}

/* dst[dlo+(i,j,k)] = src[slo+(i,j,k)] over cnt  (local_copy!) */
void dfo_local_copy(void *dst, const int64_t *dst_shape, const int64_t *dlo,
                    const void *src, const int64_t *src_shape, const int64_t *slo,
                    const int64_t *cnt, int64_t esize) {
    geom gd = make_geom(dst_shape), gs = make_geom(src_shape);
    unsigned char *D = (unsigned char *)dst;
    const unsigned char *S = (const unsigned char *)src;
    for (int64_t k = 0; k < cnt[2]; ++k)
        for (int64_t j = 0; j < cnt[1]; ++j)
            for (int64_t i = 0; i < cnt[0]; ++i) {
                int64_t di = (dlo[0] + i) + gd.st[1] * (dlo[1] + j) + gd.st[2] * (dlo[2] + k);
                int64_t si = (slo[0] + i) + gs.st[1] * (slo[1] + j) + gs.st[2] * (slo[2] + k);
                memcpy(D + di * esize, S + si * esize, (size_t)esize);
            }
}

/* axis permutation (out[perm-ordered] = in): out index (p0,p1,p2) where p = in index
 * permuted so that out axis d is in axis perm[d]. */
void dfo_permute(void *out, const void *in, const int64_t *in_shape, const int *perm, int64_t esize) {
    geom gi = make_geom(in_shape);
    int64_t on[3] = {gi.n[perm[0]], gi.n[perm[1]], gi.n[perm[2]]};
    unsigned char *O = (unsigned char *)out;
    const unsigned char *I = (const unsigned char *)in;
    int64_t idx[3];
    for (idx[2] = 0; idx[2] < gi.n[2]; ++idx[2])
        for (idx[1] = 0; idx[1] < gi.n[1]; ++idx[1])
            for (idx[0] = 0; idx[0] < gi.n[0]; ++idx[0]) {
                int64_t si = idx[0] + gi.st[1] * idx[1] + gi.st[2] * idx[2];
                int64_t di = idx[perm[0]] + on[0] * (idx[perm[1]] + on[1] * idx[perm[2]]);
                memcpy(O + di * esize, I + si * esize, (size_t)esize);
            }
}

/* ------------------------------------------------------------------ */
/* Spot bins of a 3D DFT by direct separable summation (long double)   */
/* ------------------------------------------------------------------ */
/* X[k0,k1,k2] = sum_{i,j,l} x[i,j,l] w0^(k0 i) w1^(k1 j) w2^(k2 l),
 * w_d = exp(sign 2 pi i / n_d), for each of the nb bins (0-based index
 * triples in `bins`).  The sum over i runs first for every (j, l), then
 * over j, then over l: n0 n1 n2 + n1 n2 + n2 complex multiply-adds per bin,
 * all in long double with long-double twiddles, so the reference value is
 * independent of any FFT factorization and exact to about 1e-18 relative
 * to sum |x|.  x: interleaved complex, column-major n0 x n1 x n2.
 * out: nb interleaved complex results, rounded to double. */
void dfo_spot_bins(const double *x, const int64_t *shape, const int64_t *bins, int64_t nb, int sign,
                   double *out) {
    const int64_t n0 = shape[0], n1 = shape[1], n2 = shape[2];
    long double *c0 = (long double *)malloc(sizeof(long double) * 2 * (size_t)n0);
    long double *c1 = (long double *)malloc(sizeof(long double) * 2 * (size_t)n1);
    long double *c2 = (long double *)malloc(sizeof(long double) * 2 * (size_t)n2);
    long double *a = (long double *)malloc(sizeof(long double) * 2 * (size_t)(n1 * n2));
    for (int64_t b = 0; b < nb; ++b) {
        const int64_t k0 = bins[3 * b], k1 = bins[3 * b + 1], k2 = bins[3 * b + 2];
        /* w_d^(k m) = exp(sign 2 pi i (k m mod n) / n), reduced exactly in integers */
        for (int64_t m = 0; m < n0; ++m) {
            long double c, sn;
            cis_ld((k0 * m) % n0, n0, &c, &sn);
            c0[2 * m] = c; c0[2 * m + 1] = sign < 0 ? -sn : sn;
        }
        for (int64_t m = 0; m < n1; ++m) {
            long double c, sn;
            cis_ld((k1 * m) % n1, n1, &c, &sn);
            c1[2 * m] = c; c1[2 * m + 1] = sign < 0 ? -sn : sn;
        }
        for (int64_t m = 0; m < n2; ++m) {
            long double c, sn;
            cis_ld((k2 * m) % n2, n2, &c, &sn);
            c2[2 * m] = c; c2[2 * m + 1] = sign < 0 ? -sn : sn;
        }
        /* a[j,l] = sum_i x[i,j,l] w0^(k0 i) */
#pragma omp parallel for schedule(static)
        for (int64_t jl = 0; jl < n1 * n2; ++jl) {
            const double *row = x + 2 * n0 * jl;
            long double sr = 0, si = 0;
            for (int64_t i = 0; i < n0; ++i) {
                const long double xr = row[2 * i], xi = row[2 * i + 1];
                sr += xr * c0[2 * i] - xi * c0[2 * i + 1];
                si += xr * c0[2 * i + 1] + xi * c0[2 * i];
            }
            a[2 * jl] = sr; a[2 * jl + 1] = si;
        }
        long double tr = 0, ti = 0;
        for (int64_t l = 0; l < n2; ++l) {
            long double sr = 0, si = 0;
            for (int64_t j = 0; j < n1; ++j) {
                const long double ar = a[2 * (j + n1 * l)], ai = a[2 * (j + n1 * l) + 1];
                sr += ar * c1[2 * j] - ai * c1[2 * j + 1];
                si += ar * c1[2 * j + 1] + ai * c1[2 * j];
            }
            tr += sr * c2[2 * l] - si * c2[2 * l + 1];
            ti += sr * c2[2 * l + 1] + si * c2[2 * l];
        }
        out[2 * b] = (double)tr;
        out[2 * b + 1] = (double)ti;
    }
    free(c0); free(c1); free(c2); free(a);
}

int dfo_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void dfo_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
