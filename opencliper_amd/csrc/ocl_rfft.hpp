// Register-resident DFTs for gfx950 and the two four-step halves used by every fast FFT pass of this library.
//
// A length L = 16 * P transform (P in {2,4,...,20}) is split n = a + 16 b, k = k1 + P k2:
//   FA ("a-first"):  the thread with role a holds x[a + 16 b] (b < P)  -> DFT_P in registers -> twiddle W_L^{a k1}
//                    -> one LDS transpose -> the thread with role k1 holds X[k1 + P k2] (k2 < 16) after a DFT_16.
//   FB ("k1-first"): the thread with role k1 holds x[k1 + P k2] (k2 < 16) -> DFT_16 -> twiddle W_L^{k1 n2}
//                    -> one LDS transpose -> the thread with role n2 holds X[n2 + 16 n1] (n1 < P) after a DFT_P.
// FA consumes what FB produces and vice versa, so forward/inverse pairs need no reordering.  Inverse transforms
// conjugate on the way in and out.  Small DFTs are fully unrolled with compile-time twiddles (cos/sin evaluated
// by constexpr Taylor series in double, rounded once to float).
#pragma once
#include "ocl_device.hpp"

namespace rf {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double csin(double x) {
    double t = x, s = x;
    for (int i = 1; i < 40; ++i) { t *= -x * x / ((2.0 * i) * (2.0 * i + 1)); s += t; }
    return s;
}
constexpr double ccos(double x) {
    double t = 1, s = 1;
    for (int i = 1; i < 40; ++i) { t *= -x * x / ((2.0 * i - 1) * (2.0 * i)); s += t; }
    return s;
}
template <int N> struct Tw { float re[N]; float im[N]; };
template <int N> constexpr Tw<N> make_tw() {
    Tw<N> t{};
    for (int m = 0; m < N; ++m) {
        double x = 2 * kPi * m / N;
        if (x > kPi) x -= 2 * kPi;
        t.re[m] = (float)ccos(x);
        t.im[m] = (float)-csin(x);   // exp(-2 pi i m / N)
    }
    return t;
}
template <int N> inline constexpr Tw<N> kTw = make_tw<N>();

template <int N> struct Split {              // N = A * B with A the first radix
    static constexpr int A = (N % 4 == 0 && N != 4) ? 4 : (N % 2 == 0 && N != 2) ? 2 : (N % 3 == 0 && N != 3) ? 3
                             : (N % 5 == 0 && N != 5) ? 5 : N;
    static constexpr int B = N / A;
};

template <int N> __device__ __forceinline__ void rdft(float2* v);

template <int N> __device__ __forceinline__ void rdft_split(float2* v) {
    constexpr int A = Split<N>::A, B = Split<N>::B;
    float2 t[N];
#pragma unroll
    for (int a = 0; a < B; ++a) {
        float2 u[A];
#pragma unroll
        for (int b = 0; b < A; ++b) u[b] = v[a + B * b];
        rdft<A>(u);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) {
            const int m = (a * k1) % N;
            // exact twiddles -i, -1, +i as swaps and sign changes (a multiply by a literal 0 does not fold away)
            if (m == 0) t[k1 * B + a] = u[k1];
            else if (4 * m == N) t[k1 * B + a] = make_float2(u[k1].y, -u[k1].x);
            else if (2 * m == N) t[k1 * B + a] = make_float2(-u[k1].x, -u[k1].y);
            else if (4 * m == 3 * N) t[k1 * B + a] = make_float2(-u[k1].y, u[k1].x);
            else t[k1 * B + a] = c_mul(u[k1], make_float2(kTw<N>.re[m], kTw<N>.im[m]));
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
        float2 u[B];
#pragma unroll
        for (int a = 0; a < B; ++a) u[a] = t[k1 * B + a];
        rdft<B>(u);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = u[k2];
    }
}

// forward DFT of N values held in registers, natural order in and out
template <int N> __device__ __forceinline__ void rdft(float2* v) {
    if constexpr (N == 1) {
    } else if constexpr (N == 2) {
        dft2(v);
    } else if constexpr (N == 4) {
        dft4(v[0], v[1], v[2], v[3]);
    } else if constexpr (N == 8) {
        dft8(v);
    } else if constexpr (N == 3 || N == 5 || N == 7) {
        dft_odd<N>(v);
    } else {
        rdft_split<N>(v);
    }
}

}  // namespace rf
