// regfft.hpp -- register-resident DFTs of small power-of-two sizes with
// compile-time twiddle constants, for the four-step (R1 x R2) line FFTs.
//
// DftReg<T, R, INV>::run(v) transforms v[0..R) in place, natural order in and
// out; forward uses exp(-2 pi i jk/R), inverse exp(+2 pi i jk/R), both
// unnormalised.  R = A*B is split recursively (A-point DFTs over stride-A
// sub-sequences, twiddles W_R^{a kb}, B-point DFTs across), so every index is
// a compile-time constant and the data stays in VGPRs.
#pragma once

namespace fm {

// constexpr sine/cosine (Taylor series after reduction to [0, pi/4]); used
// only to generate twiddle constants at compile time.
constexpr double kPiCe = 3.14159265358979323846264338327950288;

constexpr double ce_sin_series(double x) {
    double term = x, sum = x;
    for (int n = 1; n < 14; ++n) {
        term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double ce_cos_series(double x) {
    double term = 1.0, sum = 1.0;
    for (int n = 1; n < 14; ++n) {
        term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
        sum += term;
    }
    return sum;
}
// cos and sin of 2 pi j / R for integer 0 <= j < R (exact octant symmetry)
constexpr double ce_cos_frac(long j, long R) {
    j %= R;
    if (j < 0) j += R;
    // reduce to the first octant with exact rational arithmetic on j/R
    if (8 * j <= R) return ce_cos_series(2.0 * kPiCe * (double)j / (double)R);
    if (4 * j <= R) return ce_sin_series(2.0 * kPiCe * (double)(R - 4 * j) / (4.0 * (double)R));
    if (2 * j <= R) return -ce_cos_frac(R - 2 * j, 2 * R);
    return ce_cos_frac(R - j, R);
}
constexpr double ce_sin_frac(long j, long R) {
    j %= R;
    if (j < 0) j += R;
    if (8 * j <= R) return ce_sin_series(2.0 * kPiCe * (double)j / (double)R);
    if (4 * j <= R) return ce_cos_series(2.0 * kPiCe * (double)(R - 4 * j) / (4.0 * (double)R));
    if (2 * j <= R) return ce_sin_frac(R - 2 * j, 2 * R);
    return -ce_sin_frac(R - j, R);
}

template <int R>
struct TwConst {
    double c[R], s[R];
    constexpr TwConst() : c(), s() {
        for (int j = 0; j < R; ++j) {
            c[j] = ce_cos_frac(j, R);
            s[j] = ce_sin_frac(j, R);
        }
    }
};

template <int R>
struct TwHolder {
    static constexpr TwConst<R> tw = TwConst<R>();
};
template <int R>
constexpr TwConst<R> TwHolder<R>::tw;

template <class T, int R, bool INV>
struct DftReg;

template <class T, bool INV>
struct DftReg<T, 1, INV> {
    __device__ __forceinline__ static void run(C2<T> *) {}
};
template <class T, bool INV>
struct DftReg<T, 2, INV> {
    __device__ __forceinline__ static void run(C2<T> *v) { Bfly<T, 2, INV>::run(v); }
};
template <class T, bool INV>
struct DftReg<T, 4, INV> {
    __device__ __forceinline__ static void run(C2<T> *v) { Bfly<T, 4, INV>::run(v); }
};
template <class T, bool INV>
struct DftReg<T, 8, INV> {
    __device__ __forceinline__ static void run(C2<T> *v) { Bfly<T, 8, INV>::run(v); }
};

// R = A * B: A-point... see file header.
template <class T, int R, int A, bool INV>
struct DftSplit {
    static constexpr int B = R / A;
    __device__ __forceinline__ static void run(C2<T> *v) {
        C2<T> t[A][B];
#pragma unroll
        for (int a = 0; a < A; ++a) {
#pragma unroll
            for (int b = 0; b < B; ++b) t[a][b] = v[a + A * b];
            DftReg<T, B, INV>::run(t[a]);
        }
#pragma unroll
        for (int a = 1; a < A; ++a)
#pragma unroll
            for (int kb = 1; kb < B; ++kb) {
                const T c = T(TwHolder<R>::tw.c[(a * kb) % R]);
                const T s = INV ? T(TwHolder<R>::tw.s[(a * kb) % R]) : T(-TwHolder<R>::tw.s[(a * kb) % R]);
                t[a][kb] = cmul(t[a][kb], C2<T>{c, s});
            }
#pragma unroll
        for (int kb = 0; kb < B; ++kb) {
            C2<T> col[A];
#pragma unroll
            for (int a = 0; a < A; ++a) col[a] = t[a][kb];
            DftReg<T, A, INV>::run(col);
#pragma unroll
            for (int ka = 0; ka < A; ++ka) v[kb + B * ka] = col[ka];
        }
    }
};

template <class T, bool INV>
struct DftReg<T, 16, INV> {
    __device__ __forceinline__ static void run(C2<T> *v) { DftSplit<T, 16, 4, INV>::run(v); }
};
template <class T, bool INV>
struct DftReg<T, 32, INV> {
    __device__ __forceinline__ static void run(C2<T> *v) { DftSplit<T, 32, 4, INV>::run(v); }
};
template <class T, bool INV>
struct DftReg<T, 64, INV> {
    __device__ __forceinline__ static void run(C2<T> *v) { DftSplit<T, 64, 8, INV>::run(v); }
};

}  // namespace fm
