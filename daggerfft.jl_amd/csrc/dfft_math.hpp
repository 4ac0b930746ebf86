// dfft_math.hpp -- complex arithmetic and radix-{2,3,4,5,7,8} DFT butterflies.
//
// Compiles for host and device (DFFT_HD), so the host unit tests check the
// same butterflies the kernels run.  All butterflies compute the forward DFT
// y_k = sum_r x_r exp(-2 pi i r k / R); inverse transforms conjugate on load
// and store.  The constants are cos/sin of 2 pi k / R, written to 40
// significant digits (evaluated with mpmath), so each rounds correctly to
// double and float.
#pragma once

#if defined(__HIPCC__) || defined(__HIP__)
#define DFFT_HD __host__ __device__ __forceinline__
#else
#define DFFT_HD inline
#endif

namespace dfft {

template <class T>
struct alignas(2 * sizeof(T)) cx {
    T x, y;
};

template <class T> DFFT_HD cx<T> operator+(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T> DFFT_HD cx<T> operator-(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T> DFFT_HD cx<T> cmul(cx<T> a, cx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class T> DFFT_HD cx<T> scl(cx<T> a, T s) { return {a.x * s, a.y * s}; }
template <class T> DFFT_HD cx<T> conj(cx<T> a) { return {a.x, -a.y}; }
template <class T> DFFT_HD cx<T> mul_neg_i(cx<T> a) { return {a.y, -a.x}; }  // -i * a
template <class T> DFFT_HD cx<T> mul_pos_i(cx<T> a) { return {-a.y, a.x}; }  // +i * a

struct K {
    static constexpr double SQRT1_2 = 0.7071067811865475244008443621048490392848;
    static constexpr double SQRT3_2 = 0.8660254037844386467637231707529361834714;
    static constexpr double C5_1 = 0.3090169943749474241022934171828190588602;
    static constexpr double C5_2 = -0.8090169943749474241022934171828190588602;
    static constexpr double S5_1 = 0.9510565162951535721164393333793821434057;
    static constexpr double S5_2 = 0.5877852522924731291687059546390727685977;
    static constexpr double C7_1 = 0.6234898018587335305250048840042398106323;
    static constexpr double C7_2 = -0.2225209339563144042889025644967947594664;
    static constexpr double C7_3 = -0.9009688679024191262361023195074450511659;
    static constexpr double S7_1 = 0.7818314824680298087084445266740577502323;
    static constexpr double S7_2 = 0.9749279121818236070181316829939312172328;
    static constexpr double S7_3 = 0.4338837391175581204757683328483587546099;
};

template <class T> DFFT_HD void bfly2(cx<T> *v) {
    cx<T> t = v[0];
    v[0] = t + v[1];
    v[1] = t - v[1];
}

template <class T> DFFT_HD void bfly4(cx<T> &a0, cx<T> &a1, cx<T> &a2, cx<T> &a3) {
    cx<T> t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = mul_neg_i(a1 - a3);
    a0 = t0 + t2;
    a2 = t0 - t2;
    a1 = t1 + t3;
    a3 = t1 - t3;
}
template <class T> DFFT_HD void bfly4(cx<T> *v) { bfly4(v[0], v[1], v[2], v[3]); }

template <class T> DFFT_HD void bfly8(cx<T> *v) {
    const T c = (T)K::SQRT1_2;
    cx<T> u0 = v[0] + v[4], u1 = v[1] + v[5], u2 = v[2] + v[6], u3 = v[3] + v[7];
    cx<T> w0 = v[0] - v[4], w1 = v[1] - v[5], w2 = v[2] - v[6], w3 = v[3] - v[7];
    w1 = cx<T>{(w1.x + w1.y) * c, (w1.y - w1.x) * c};   // * exp(-i pi/4)
    w2 = mul_neg_i(w2);                                 // * exp(-i pi/2)
    w3 = cx<T>{(w3.y - w3.x) * c, -(w3.x + w3.y) * c};  // * exp(-3 i pi/4)
    bfly4(u0, u1, u2, u3);
    bfly4(w0, w1, w2, w3);
    v[0] = u0; v[1] = w0; v[2] = u1; v[3] = w1;
    v[4] = u2; v[5] = w2; v[6] = u3; v[7] = w3;
}

template <class T> DFFT_HD void bfly3(cx<T> *v) {
    const T s = (T)K::SQRT3_2;
    cx<T> t1 = v[1] + v[2];
    cx<T> t2 = v[0] - scl(t1, (T)0.5);
    cx<T> t3 = scl(v[1] - v[2], s);
    v[0] = v[0] + t1;
    v[1] = cx<T>{t2.x + t3.y, t2.y - t3.x};  // t2 - i t3
    v[2] = cx<T>{t2.x - t3.y, t2.y + t3.x};  // t2 + i t3
}

template <class T> DFFT_HD void bfly5(cx<T> *v) {
    const T c1 = (T)K::C5_1, c2 = (T)K::C5_2, s1 = (T)K::S5_1, s2 = (T)K::S5_2;
    cx<T> b1 = v[1] + v[4], b2 = v[2] + v[3], d1 = v[1] - v[4], d2 = v[2] - v[3];
    cx<T> a0 = v[0];
    cx<T> p1 = cx<T>{a0.x + c1 * b1.x + c2 * b2.x, a0.y + c1 * b1.y + c2 * b2.y};
    cx<T> p2 = cx<T>{a0.x + c2 * b1.x + c1 * b2.x, a0.y + c2 * b1.y + c1 * b2.y};
    cx<T> q1 = cx<T>{s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y};
    cx<T> q2 = cx<T>{s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y};
    v[0] = a0 + b1 + b2;
    v[1] = p1 + mul_neg_i(q1);
    v[4] = p1 + mul_pos_i(q1);
    v[2] = p2 + mul_neg_i(q2);
    v[3] = p2 + mul_pos_i(q2);
}

template <class T> DFFT_HD void bfly7(cx<T> *v) {
    const T c1 = (T)K::C7_1, c2 = (T)K::C7_2, c3 = (T)K::C7_3;
    const T s1 = (T)K::S7_1, s2 = (T)K::S7_2, s3 = (T)K::S7_3;
    cx<T> b1 = v[1] + v[6], b2 = v[2] + v[5], b3 = v[3] + v[4];
    cx<T> d1 = v[1] - v[6], d2 = v[2] - v[5], d3 = v[3] - v[4];
    cx<T> a0 = v[0];
    cx<T> A1 = cx<T>{a0.x + c1 * b1.x + c2 * b2.x + c3 * b3.x, a0.y + c1 * b1.y + c2 * b2.y + c3 * b3.y};
    cx<T> A2 = cx<T>{a0.x + c2 * b1.x + c3 * b2.x + c1 * b3.x, a0.y + c2 * b1.y + c3 * b2.y + c1 * b3.y};
    cx<T> A3 = cx<T>{a0.x + c3 * b1.x + c1 * b2.x + c2 * b3.x, a0.y + c3 * b1.y + c1 * b2.y + c2 * b3.y};
    cx<T> B1 = cx<T>{s1 * d1.x + s2 * d2.x + s3 * d3.x, s1 * d1.y + s2 * d2.y + s3 * d3.y};
    cx<T> B2 = cx<T>{s2 * d1.x - s3 * d2.x - s1 * d3.x, s2 * d1.y - s3 * d2.y - s1 * d3.y};
    cx<T> B3 = cx<T>{s3 * d1.x - s1 * d2.x + s2 * d3.x, s3 * d1.y - s1 * d2.y + s2 * d3.y};
    v[0] = a0 + b1 + b2 + b3;
    v[1] = A1 + mul_neg_i(B1);
    v[6] = A1 + mul_pos_i(B1);
    v[2] = A2 + mul_neg_i(B2);
    v[5] = A2 + mul_pos_i(B2);
    v[3] = A3 + mul_neg_i(B3);
    v[4] = A3 + mul_pos_i(B3);
}

template <int R, class T> DFFT_HD void bfly(cx<T> *v) {
    if constexpr (R == 2) bfly2(v);
    else if constexpr (R == 3) bfly3(v);
    else if constexpr (R == 4) bfly4(v);
    else if constexpr (R == 5) bfly5(v);
    else if constexpr (R == 7) bfly7(v);
    else if constexpr (R == 8) bfly8(v);
    else static_assert(R == 2, "unsupported radix");
}

}  // namespace dfft
