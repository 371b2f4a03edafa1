// ORACLE — test infrastructure only.
//
// CPU restatement of OpenCLIPER's NESTA-TV path (FP32 "reference-faithful" form and FP64 oracle form, one
// template).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library, and
// only as the checker; the product path (opencliper_amd) never links or calls it.
//
// What each function follows (paths relative to the reference checkout):
//   fft1d / fft2d ........ clFFT conventions used by src/processes/FFT.cpp:127-146,336-449,510-525
//                          (2-D C2C, forward e^{-2*pi*i}, unscaled; backward e^{+2*pi*i}, scaled by 1/(H*W);
//                          no fftshift).  The clFFT kernels themselves are third-party and not in the checkout.
//   ttv_fwd / ttv_adj .... src/kernels/temporalTV.cl:49-76 (operator_tTV) and :78-107 (operator_tTVadj = -D^T)
//   vecnorm .............. src/kernels/vectorNormalization.cl:42-56
//   cep .................. src/kernels/complexElementProd.cl:42-76 + src/processes/ComplexElementProd.cpp:117
//   apply_mask ........... src/kernels/applyMask.cl:5-40
//   ximage_sum ........... src/kernels/xImageSum.cl:42-61
//   abs / abspow2 ........ src/kernels/complexAbs.cl:43-70, src/kernels/complexAbsPow2.cl:43-71
//   nrm2/dotc/csum/asum .. CLBlast Scnrm2 / Cdotc (conjugates x) / Scsum / Scasum as called from
//                          src/processes/nesta/NestaUp.cpp (summation order here: sequential, in double)
//   op_A / op_At ......... src/processes/nesta/NestaUp.cpp:162-190 / :117-151
//   normest .............. src/processes/nesta/NestaUp.cpp:291-439 (power method, drand48 restart)
//   nesta ................ src/processes/nesta/NestaUp.cpp:442-910 (continuation NESTA, stop test 1)
//   StopRule ............. NestaUp.cpp:645,649,799-824 (ring of objective values, armed / ending hits)
//   simple_recon ......... src/processes/examples/SimpleMRIRecon.cpp:59-101
//   rsos ................. src/kernels/rss.cl (rss_kernel): sqrt(sum_c |in_c|^2) in the real part, float sum in coil order
//   simple_recon_sos ..... src/processes/examples/SimpleMRIReconSOS.cpp (in-place inverse FFT, then RSoS)
//   mc_fwd / mc_adj ...... motion compensation: src/processes/MotionCompensation.cpp (copy, then the interpolator
//                          kernel, src/kernels/optimizer.cl:832-867) and src/processes/AdjointMotionCompensation.cpp
//                          (copy, initZeroRect, initialize.cl:55-69, then adjointInterpolator, optimizer.cl:880-930);
//                          box rows / columns as Interpolator.cpp:31-47 builds r1margin / r2margin
//   nesta with MC ........ the argsMC != nullptr branch: U = tTV(MC(x)), U^T = MC^T(tTV^T(u)) (NestaUp.cpp:201-278),
//                          x_ref taken from the output instead of A^H y (:483-486), normest over the same U
//
// Pinning (each checked in tests/test_oracle.py): fft vs brute-force DFT and numpy.fft; ttv vs closed form on
// integers and the adjoint identity; normest's LCG vs libc drand48 from the POSIX default state and its result vs
// the closed-form sigma_max of the periodic difference; vecnorm vs the Huber identity; op_A/op_At adjointness and
// E^H E = I when fully sampled; nesta vs an independent numpy FP64 implementation (tests/golden/make_golden.py);
// mc_fwd vs scipy.ndimage.map_coordinates (order 1), identity and integer-shift meshes, and <MC x, y> = <x, MC^T y>;
// normest with MC vs the largest singular value of the explicit matrix of U; nesta with an identity MC vs nesta;
// the stop rule on scripted objective sequences whose outcomes follow by hand from NestaUp.cpp:799-824.

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <climits>
#include <vector>
#include <algorithm>
#include <omp.h>

namespace {

template <class T> struct Cx { T re, im; };

// ------------------------------------------------------------------------------------------------------------------
// Mixed-radix Stockham FFT (natural order in and out).  Twiddles computed in double, rounded to T.
// ------------------------------------------------------------------------------------------------------------------
template <class T> struct Plan {
    int n = 0;
    std::vector<int> radix;
    std::vector<Cx<T>> w;   // w[m] = exp(-2*pi*i*m/n)
};

static std::vector<int> factorize(int n) {
    std::vector<int> r;
    while (n % 4 == 0) { r.push_back(4); n /= 4; }
    for (int p : {2, 3, 5, 7}) while (n % p == 0) { r.push_back(p); n /= p; }
    for (int p = 11; n > 1; p += 2) while (n % p == 0) { r.push_back(p); n /= p; }
    return r;
}

template <class T> Plan<T> make_plan(int n) {
    Plan<T> p;
    p.n = n;
    p.radix = factorize(n);
    p.w.resize(n);
    for (int m = 0; m < n; ++m) {
        double a = -2.0 * M_PI * (double)m / (double)n;
        p.w[m].re = (T)std::cos(a);
        p.w[m].im = (T)std::sin(a);
    }
    return p;
}

// Forward transform of one sequence (in -> out, both length n, unit stride).  tmp has length n.
template <class T> void fft_forward(const Plan<T>& P, const Cx<T>* in, Cx<T>* out, Cx<T>* tmp) {
    const int n = P.n;
    if (n == 1) { out[0] = in[0]; return; }
    const Cx<T>* src = in;
    Cx<T>* bufs[2] = {out, tmp};
    // Choose the first destination so that the last stage lands in `out`.
    int ns = (int)P.radix.size();
    int cur = (ns % 2 == 1) ? 0 : 1;
    int Ls = 1;
    Cx<T> v[64], u[64];
    for (int s = 0; s < ns; ++s) {
        const int R = P.radix[s];
        const int m = n / R;
        Cx<T>* dst = bufs[cur];
        for (int j = 0; j < m; ++j) {
            const int k = j % Ls;
            for (int r = 0; r < R; ++r) {
                Cx<T> a = src[j + r * m];
                long tix = ((long)r * k * (n / (Ls * R))) % n;
                Cx<T> t = P.w[tix];
                v[r].re = a.re * t.re - a.im * t.im;
                v[r].im = a.re * t.im + a.im * t.re;
            }
            if (R == 2) {
                u[0].re = v[0].re + v[1].re; u[0].im = v[0].im + v[1].im;
                u[1].re = v[0].re - v[1].re; u[1].im = v[0].im - v[1].im;
            } else if (R == 4) {   // X_q = sum_r v_r (-i)^{qr}
                const T ar = v[0].re + v[2].re, ai = v[0].im + v[2].im, br = v[0].re - v[2].re, bi = v[0].im - v[2].im;
                const T cr = v[1].re + v[3].re, ci = v[1].im + v[3].im, dr = v[1].re - v[3].re, di = v[1].im - v[3].im;
                u[0].re = ar + cr; u[0].im = ai + ci;
                u[2].re = ar - cr; u[2].im = ai - ci;
                u[1].re = br + di; u[1].im = bi - dr;      // b - i d
                u[3].re = br - di; u[3].im = bi + dr;      // b + i d
            } else {
                for (int q = 0; q < R; ++q) {
                    T sr = 0, si = 0;
                    for (int r = 0; r < R; ++r) {
                        Cx<T> t = P.w[((long)(q * r) % R) * m];
                        sr += v[r].re * t.re - v[r].im * t.im;
                        si += v[r].re * t.im + v[r].im * t.re;
                    }
                    u[q].re = sr; u[q].im = si;
                }
            }
            const int base = (j / Ls) * Ls * R + k;
            for (int q = 0; q < R; ++q) dst[base + q * Ls] = u[q];
        }
        src = dst;
        cur ^= 1;
        Ls *= R;
    }
}

// 2-D transform of `batch` images [H][W] (W fastest).  dir < 0: forward, unscaled.  dir > 0: backward, 1/(H*W).
template <class T> void fft2d(const Cx<T>* in, Cx<T>* out, long batch, int H, int W, int dir) {
    Plan<T> pw = make_plan<T>(W), ph = make_plan<T>(H);
    const long N = (long)H * W;
    const bool inv = dir > 0;
    const T scale = inv ? (T)(1.0 / (double)N) : (T)1;
    const int CB = 8;   // columns gathered together (one 64-byte line of complex64)
#pragma omp parallel
    {
        std::vector<Cx<T>> a(std::max(H, W)), b(std::max(H, W)), t(std::max(H, W)), blk((size_t)CB * H);
#pragma omp for schedule(static)
        for (long img = 0; img < batch; ++img) {
            const Cx<T>* I = in + img * N;
            Cx<T>* O = out + img * N;
            for (int h = 0; h < H; ++h) {
                for (int x = 0; x < W; ++x) { a[x] = I[(long)h * W + x]; if (inv) a[x].im = -a[x].im; }
                fft_forward(pw, a.data(), b.data(), t.data());
                for (int x = 0; x < W; ++x) O[(long)h * W + x] = b[x];
            }
            for (int x0 = 0; x0 < W; x0 += CB) {
                const int nb = std::min(CB, W - x0);
                for (int h = 0; h < H; ++h)
                    for (int q = 0; q < nb; ++q) blk[(size_t)q * H + h] = O[(long)h * W + x0 + q];
                for (int q = 0; q < nb; ++q) {
                    fft_forward(ph, &blk[(size_t)q * H], b.data(), t.data());
                    for (int h = 0; h < H; ++h) blk[(size_t)q * H + h] = b[h];
                }
                for (int h = 0; h < H; ++h)
                    for (int q = 0; q < nb; ++q) {
                        Cx<T> z = blk[(size_t)q * H + h];
                        if (inv) { z.im = -z.im; z.re *= scale; z.im *= scale; }
                        O[(long)h * W + x0 + q] = z;
                    }
            }
        }
    }
}

// ------------------------------------------------------------------------------------------------------------------
// Element-wise operators (layouts: XData [F][N]; KData [F][C][N]; sensitivity maps [C][N]; masks [F][N] uint8)
// ------------------------------------------------------------------------------------------------------------------

// temporalTV.cl:49-76: out_0 = in_0 - in_{F-1}; out_f = in_f - in_{f-1}
template <class T> void ttv_fwd(const Cx<T>* in, Cx<T>* out, int F, long N) {
#pragma omp parallel for schedule(static)
    for (long p = 0; p < N; ++p) {
        for (int f = 0; f < F; ++f) {
            int g = (f == 0) ? F - 1 : f - 1;
            out[f * N + p].re = in[f * N + p].re - in[g * N + p].re;
            out[f * N + p].im = in[f * N + p].im - in[g * N + p].im;
        }
    }
}

// temporalTV.cl:78-107: out_f = in_{f+1} - in_f (f < F-1); out_{F-1} = in_0 - in_{F-1}
template <class T> void ttv_adj(const Cx<T>* in, Cx<T>* out, int F, long N) {
#pragma omp parallel for schedule(static)
    for (long p = 0; p < N; ++p) {
        for (int f = 0; f < F; ++f) {
            int g = (f == F - 1) ? 0 : f + 1;
            out[f * N + p].re = in[g * N + p].re - in[f * N + p].re;
            out[f * N + p].im = in[g * N + p].im - in[f * N + p].im;
        }
    }
}

// vectorNormalization.cl:42-56: out = in / max(mu, |in|)
template <class T> void vecnorm(const Cx<T>* in, Cx<T>* out, long n, T mu) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        T len = std::sqrt(in[i].re * in[i].re + in[i].im * in[i].im);
        T fac = std::max(mu, len);
        out[i].re = in[i].re / fac;
        out[i].im = in[i].im / fac;
    }
}

// complexElementProd.cl:42-76.  in_coils == false: input is an XData [F][N] broadcast over coils.
template <class T> void cep(const Cx<T>* in, const Cx<T>* S, Cx<T>* out, int F, int C, long N, bool in_coils,
                            bool conj) {
#pragma omp parallel for schedule(static)
    for (long p = 0; p < N; ++p) {
        for (int f = 0; f < F; ++f)
            for (int c = 0; c < C; ++c) {
                Cx<T> a = in_coils ? in[((long)f * C + c) * N + p] : in[(long)f * N + p];
                Cx<T> s = S[(long)c * N + p];
                if (conj) s.im = -s.im;
                Cx<T> r;
                r.re = a.re * s.re - a.im * s.im;
                r.im = a.re * s.im + a.im * s.re;
                out[((long)f * C + c) * N + p] = r;
            }
    }
}

// applyMask.cl:5-40: zero where mask == 0
template <class T> void apply_mask(Cx<T>* d, const uint8_t* M, int F, int C, long N) {
#pragma omp parallel for schedule(static)
    for (long p = 0; p < N; ++p)
        for (int f = 0; f < F; ++f)
            if (M[(long)f * N + p] == 0)
                for (int c = 0; c < C; ++c) { d[((long)f * C + c) * N + p].re = 0; d[((long)f * C + c) * N + p].im = 0; }
}

// xImageSum.cl:42-61: x_f = sum_c in_{f,c}, sequential over c
template <class T> void ximage_sum(const Cx<T>* in, Cx<T>* out, int F, int C, long N) {
#pragma omp parallel for schedule(static)
    for (long p = 0; p < N; ++p)
        for (int f = 0; f < F; ++f) {
            T sr = 0, si = 0;
            for (int c = 0; c < C; ++c) { sr += in[((long)f * C + c) * N + p].re; si += in[((long)f * C + c) * N + p].im; }
            out[(long)f * N + p].re = sr;
            out[(long)f * N + p].im = si;
        }
}

// BLAS-1 reductions: sequential summation in double, result rounded to T.
template <class T> T nrm2(const Cx<T>* x, long n) {
    double s = 0;
    for (long i = 0; i < n; ++i) s += (double)x[i].re * x[i].re + (double)x[i].im * x[i].im;
    return (T)std::sqrt(s);
}
template <class T> T dotc_re(const Cx<T>* x, const Cx<T>* y, long n) {   // Re sum conj(x) y
    double s = 0;
    for (long i = 0; i < n; ++i) s += (double)x[i].re * y[i].re + (double)x[i].im * y[i].im;
    return (T)s;
}
template <class T> T asum(const Cx<T>* x, long n) {
    double s = 0;
    for (long i = 0; i < n; ++i) s += std::fabs((double)x[i].re) + std::fabs((double)x[i].im);
    return (T)s;
}
template <class T> T sum_abspow2(const Cx<T>* x, long n) {  // ComplexAbsPow2 (rounded to T per element) + Scsum
    double s = 0;
    for (long i = 0; i < n; ++i) { T v = x[i].re * x[i].re + x[i].im * x[i].im; s += (double)v; }
    return (T)s;
}
template <class T> T max_abs(const Cx<T>* x, long n) {   // ComplexAbs + host max loop (NestaUp.cpp:521-567)
    T m = (T)LONG_MIN;
    for (long i = 0; i < n; ++i) { T v = std::sqrt(x[i].re * x[i].re + x[i].im * x[i].im); if (v > m) m = v; }
    return m;
}

// y = a*x + y, x = a*x, copy
template <class T> void axpy(long n, T a, const Cx<T>* x, Cx<T>* y) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) { y[i].re += a * x[i].re; y[i].im += a * x[i].im; }
}
template <class T> void scal(long n, T a, Cx<T>* x) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) { x[i].re *= a; x[i].im *= a; }
}
template <class T> void copy(long n, const Cx<T>* x, Cx<T>* y) { std::memcpy(y, x, sizeof(Cx<T>) * n); }

// ------------------------------------------------------------------------------------------------------------------
// Encoding operators (NestaUp.cpp:117-190)
// ------------------------------------------------------------------------------------------------------------------
template <class T> struct Prob {
    int F, C, H, W;
    long N;
    const Cx<T>* y;       // [F][C][N]
    const Cx<T>* S;       // [C][N]
    const uint8_t* M;     // [F][N]
};

// operatorA: out = mask(FFT_fwd(S * x)); aux is KData-sized scratch
template <class T> void op_A(const Prob<T>& P, const Cx<T>* x, Cx<T>* out, Cx<T>* aux) {
    cep(x, P.S, aux, P.F, P.C, P.N, false, false);
    fft2d(aux, out, (long)P.F * P.C, P.H, P.W, -1);
    apply_mask(out, P.M, P.F, P.C, P.N);
}

// operatorAt: x = sum_c conj(S_c) * FFT_bwd(r)  (single coil: plain copy)
template <class T> void op_At(const Prob<T>& P, const Cx<T>* r, Cx<T>* x, Cx<T>* aux) {
    fft2d(r, aux, (long)P.F * P.C, P.H, P.W, +1);
    cep(aux, P.S, aux, P.F, P.C, P.N, true, true);
    if (P.C != 1) ximage_sum(aux, x, P.F, P.C, P.N);
    else copy((long)P.F * P.N, aux, x);
}

// ------------------------------------------------------------------------------------------------------------------
// Motion compensation.  Box: rows r0 .. r0+nr-1 and columns c0 .. c0+nc-1 (0-based) from the 1-based bound box
// [b0, b1, b2, b3] and margins m0 = H / 40, m1 = W / 40 (integer division): r0 = b1 - m0 - 2, nr = b3 - b1 + 2 m0 + 3,
// c0 = b0 - m1 - 2, nc = b2 - b0 + 2 m1 + 3 (Interpolator.cpp:31-47).  For box pixel (r, c) and frame k the output
// position is (hor, ver) = mesh x at [r + c H] minus 1, and the sample position (nhor, nver) = xn at
// [c + r W + k H W] (and + H W F) minus 1.  The reference reads and writes outside the frame when a position leaves
// it; here such neighbours read as zero and such writes are dropped.
// ------------------------------------------------------------------------------------------------------------------
struct McBox { long r0, nr, c0, nc; };

static bool mc_box(int H, int W, const unsigned bb[4], McBox* b) {
    const long m0 = H / 40, m1 = W / 40;
    b->r0 = (long)bb[1] - m0 - 2;
    b->nr = (long)bb[3] - (long)bb[1] + 2 * m0 + 3;
    b->c0 = (long)bb[0] - m1 - 2;
    b->nc = (long)bb[2] - (long)bb[0] + 2 * m1 + 3;
    return b->nr > 0 && b->nc > 0 && b->r0 >= 0 && b->c0 >= 0 && b->r0 + b->nr <= H && b->c0 + b->nc <= W;
}

struct McOp {
    const float* xn;   // [2][F][H][W]
    const int* x;      // [2][W][H]
    McBox box;
};

template <class T> struct McSample {
    bool ok;           // output position inside the frame
    long out;          // output index within the frame
    long q[4];         // neighbour indices (-1 outside): left-up, right-up, left-down, right-down
    T nhor, nver;
    long prehor, prever;
};

template <class T> McSample<T> mc_sample(const McOp& m, int F, int H, int W, long r, long c, int k) {
    const long N = (long)H * W;
    McSample<T> s;
    const long hor = (long)m.x[r + c * H] - 1, ver = (long)m.x[r + c * H + N] - 1;
    s.ok = hor >= 0 && hor < W && ver >= 0 && ver < H;
    s.out = hor + ver * W;
    s.nhor = (T)m.xn[c + r * W + (long)k * N] - (T)1;
    s.nver = (T)m.xn[c + r * W + (long)k * N + N * F] - (T)1;
    s.prehor = s.prever = 0;
    for (int i = 0; i < 4; ++i) s.q[i] = -1;
    if (!(s.nhor > (T)-2 && s.nhor < (T)(W + 1) && s.nver > (T)-2 && s.nver < (T)(H + 1))) return s;
    s.prehor = (long)std::floor(s.nhor);
    s.prever = (long)std::floor(s.nver);
    const long hs[2] = {s.prehor, s.prehor + 1}, vs[2] = {s.prever, s.prever + 1};
    for (int v = 0; v < 2; ++v)
        for (int h = 0; h < 2; ++h)
            if (hs[h] >= 0 && hs[h] < W && vs[v] >= 0 && vs[v] < H) s.q[2 * v + h] = hs[h] + vs[v] * W;
    return s;
}

template <class T> void mc_fwd(const McOp& m, const Cx<T>* in, Cx<T>* out, int F, int H, int W) {
    const long N = (long)H * W;
    std::memcpy(out, in, sizeof(Cx<T>) * N * F);
    for (int k = 0; k < F; ++k)
        for (long i = 0; i < m.box.nr; ++i)
            for (long j = 0; j < m.box.nc; ++j) {
                const McSample<T> s = mc_sample<T>(m, F, H, W, m.box.r0 + i, m.box.c0 + j, k);
                if (!s.ok) continue;
                const Cx<T>* fr = in + (long)k * N;
                Cx<T> p[4];
                for (int t = 0; t < 4; ++t) p[t] = s.q[t] >= 0 ? fr[s.q[t]] : Cx<T>{0, 0};
                const T wl = (T)(s.prehor + 1) - s.nhor, wr = s.nhor - (T)s.prehor;    // (posthor - nhor), (nhor - prehor)
                const T wu = (T)(s.prever + 1) - s.nver, wd = s.nver - (T)s.prever;
                const Cx<T> a1{wl * p[0].re + wr * p[1].re, wl * p[0].im + wr * p[1].im};
                const Cx<T> a2{wl * p[2].re + wr * p[3].re, wl * p[2].im + wr * p[3].im};
                out[(long)k * N + s.out] = Cx<T>{wu * a1.re + wd * a2.re, wu * a1.im + wd * a2.im};
            }
}

template <class T> void mc_adj(const McOp& m, const Cx<T>* in, Cx<T>* out, int F, int H, int W) {
    const long N = (long)H * W;
    std::memcpy(out, in, sizeof(Cx<T>) * N * F);
    for (int k = 0; k < F; ++k)                                   // initZeroRect over the box
        for (long i = 0; i < m.box.nr; ++i)
            for (long j = 0; j < m.box.nc; ++j) out[(long)k * N + (m.box.c0 + j) + (m.box.r0 + i) * W] = Cx<T>{0, 0};
    for (int k = 0; k < F; ++k)
        for (long i = 0; i < m.box.nr; ++i)
            for (long j = 0; j < m.box.nc; ++j) {
                const McSample<T> s = mc_sample<T>(m, F, H, W, m.box.r0 + i, m.box.c0 + j, k);
                if (!s.ok) continue;
                const Cx<T> i0 = in[(long)k * N + s.out];
                const T Whh = s.nhor - (T)s.prehor, Wvv = s.nver - (T)s.prever;
                for (int t = 0; t < 4; ++t) {                     // weights as written at optimizer.cl:913-928
                    if (s.q[t] < 0) continue;
                    const T hw = (t & 1) ? Whh : (1 - Whh), vw = (t < 2) ? (1 - Wvv) : Wvv;
                    Cx<T>& o = out[(long)k * N + s.q[t]];
                    o.re += i0.re * hw * vw;
                    o.im += i0.im * hw * vw;
                }
            }
}

// U = tTV o MC (or tTV) and U^T = MC^T o tTV^T (or tTV^T, then a copy): NestaUp.cpp:201-278
template <class T> void op_U(const McOp* m, const Cx<T>* x, Cx<T>* out, Cx<T>* aux, int F, int H, int W) {
    if (m) {
        mc_fwd(*m, x, aux, F, H, W);
        ttv_fwd(aux, out, F, (long)H * W);
    } else {
        ttv_fwd(x, out, F, (long)H * W);
    }
}
template <class T> void op_Ut(const McOp* m, const Cx<T>* u, Cx<T>* out, Cx<T>* aux, int F, int H, int W) {
    ttv_adj(u, aux, F, (long)H * W);
    if (m) mc_adj(*m, aux, out, F, H, W);
    else copy((long)F * H * W, aux, out);
}

// ------------------------------------------------------------------------------------------------------------------
// drand48 from the POSIX default state (X0 = 0x1234ABCD330E, a = 0x5DEECE66D, c = 0xB, modulus 2^48).
// ------------------------------------------------------------------------------------------------------------------
struct Drand48 {
    uint64_t x = 0x1234ABCD330EULL;
    long long pos = 0;        // draws taken since the default state
    double next() {
        x = (0x5DEECE66DULL * x + 0xBULL) & ((1ULL << 48) - 1);
        ++pos;
        return (double)x / 281474976710656.0;
    }
    // the state after k draws from the default state: x -> a x + c applied k times, by squaring the affine map
    void seek(long long k) {
        const uint64_t m = (1ULL << 48) - 1;
        uint64_t A = 1, Cc = 0;                 // accumulated map x -> A x + Cc
        uint64_t a = 0x5DEECE66DULL, c = 0xBULL; // current power of the step map
        for (long long e = k; e > 0; e >>= 1) {
            if (e & 1) { A = (a * A) & m; Cc = (a * Cc + c) & m; }
            c = ((a + 1) * c) & m;
            a = (a * a) & m;
        }
        x = (A * 0x1234ABCD330EULL + Cc) & m;
        pos = k;
    }
};
// The reference never seeds drand48 (NestaUp.cpp:375), so every myNormest of a process continues one stream from the
// POSIX default state.  The restatement keeps that stream here; the ctypes binding seeks it before a call (to 0 for a
// reconstruction that runs first in its process, or to where a previous one left it).
Drand48 g_drand48;

// myNormest (NestaUp.cpp:291-439): power method for the spectral norm of U = tTV (or tTV o MC with mc, frames of
// H x W), FP-typed like the reference.
template <class T> T normest(int F, long N, int* iters_out, const McOp* mc = nullptr, int H = 0, int W = 0) {
    const long n = (long)F * N;
    const T tol = (T)1e-3;
    const unsigned maxiter = 100;
    unsigned cnt = 0;
    std::vector<Cx<T>> x(n, Cx<T>{(T)1, (T)0}), Sx(n), aux(n), auxMC(mc ? n : 0);
    T e = nrm2(x.data(), n);
    scal(n, (T)1 / e, x.data());
    T e0 = 0;
    Drand48& rng = g_drand48;
    while ((std::abs(e - e0) > (tol * e)) && (cnt < maxiter)) {
        e0 = e;
        if (mc) op_U(mc, x.data(), Sx.data(), auxMC.data(), F, H, W);
        else ttv_fwd(x.data(), Sx.data(), F, N);
        T nnz = asum(Sx.data(), n);
        if (nnz == 0) {
            for (long i = 0; i < n; ++i) { Sx[i].re = (T)rng.next(); Sx[i].im = 0; }
        }
        e = nrm2(Sx.data(), n);
        if (mc) {
            op_Ut(mc, Sx.data(), x.data(), aux.data(), F, H, W);
        } else {
            ttv_adj(Sx.data(), aux.data(), F, N);
            copy(n, aux.data(), x.data());
        }
        T nx = nrm2(x.data(), n);
        scal(n, (T)1 / nx, x.data());
        cnt = cnt + 1;
    }
    if (iters_out) *iters_out = (int)cnt;
    return e;
}

}  // namespace

// ======================================================================================================================
// C ABI (ctypes).  Complex arrays are interleaved (re, im).
// ======================================================================================================================
extern "C" {

struct orc_params {          // NestaUp::LaunchParameters field order (include/.../nesta/NestaUp.hpp:67-83)
    double lambda_i, mu_f, La;
    int maxIntIter;
    double tolVar;
    int verbose, maxIter, stoptest, miniter;
};

#define ORC_MAX_STAGES 16
#define ORC_TRACE 4096
struct orc_report {
    double normU, maxXref, maxUXref, lambda, muf, mu0, gamma, gammat, final_tolVar, final_mu;
    int normest_iters, stages;
    int iters[ORC_MAX_STAGES];       // inner iterations executed per stage
    int ntrace;                      // number of f values recorded
    double trace_f[ORC_TRACE];       // f per inner iteration (all stages, in order)
    double trace_l2[ORC_TRACE];      // ||r||^2 per inner iteration
    double trace_l1[ORC_TRACE];      // fx - mu/2 ||u||^2 per inner iteration
};

int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }

void orc_fft2d_f32(const float* in, float* out, long batch, int H, int W, int dir) {
    fft2d((const Cx<float>*)in, (Cx<float>*)out, batch, H, W, dir);
}
void orc_fft2d_f64(const double* in, double* out, long batch, int H, int W, int dir) {
    fft2d((const Cx<double>*)in, (Cx<double>*)out, batch, H, W, dir);
}
void orc_ttv_f32(const float* in, float* out, int F, long N, int dir) {
    if (dir) ttv_fwd((const Cx<float>*)in, (Cx<float>*)out, F, N); else ttv_adj((const Cx<float>*)in, (Cx<float>*)out, F, N);
}
void orc_ttv_f64(const double* in, double* out, int F, long N, int dir) {
    if (dir) ttv_fwd((const Cx<double>*)in, (Cx<double>*)out, F, N); else ttv_adj((const Cx<double>*)in, (Cx<double>*)out, F, N);
}
void orc_vecnorm_f32(const float* in, float* out, long n, float mu) { vecnorm((const Cx<float>*)in, (Cx<float>*)out, n, mu); }
void orc_cep_f32(const float* in, const float* S, float* out, int F, int C, long N, int in_coils, int conj) {
    cep((const Cx<float>*)in, (const Cx<float>*)S, (Cx<float>*)out, F, C, N, in_coils != 0, conj != 0);
}
void orc_apply_mask_f32(float* d, const uint8_t* M, int F, int C, long N) { apply_mask((Cx<float>*)d, M, F, C, N); }
void orc_ximage_sum_f32(const float* in, float* out, int F, int C, long N) {
    ximage_sum((const Cx<float>*)in, (Cx<float>*)out, F, C, N);
}
double orc_nrm2_f32(const float* x, long n) { return nrm2((const Cx<float>*)x, n); }
double orc_dotc_re_f32(const float* x, const float* y, long n) { return dotc_re((const Cx<float>*)x, (const Cx<float>*)y, n); }
double orc_asum_f32(const float* x, long n) { return asum((const Cx<float>*)x, n); }

void orc_op_A_f64(const double* x, const double* S, const uint8_t* M, double* out, int F, int C, int H, int W) {
    Prob<double> P{F, C, H, W, (long)H * W, nullptr, (const Cx<double>*)S, M};
    std::vector<Cx<double>> aux((long)F * C * H * W);
    op_A(P, (const Cx<double>*)x, (Cx<double>*)out, aux.data());
}
void orc_op_At_f64(const double* r, const double* S, double* x, int F, int C, int H, int W) {
    Prob<double> P{F, C, H, W, (long)H * W, nullptr, (const Cx<double>*)S, nullptr};
    std::vector<Cx<double>> aux((long)F * C * H * W);
    op_At(P, (const Cx<double>*)r, (Cx<double>*)x, aux.data());
}

double orc_normest_f32(int F, long N, int* iters) { return normest<float>(F, N, iters); }
double orc_normest_f64(int F, long N, int* iters) { return normest<double>(F, N, iters); }
double orc_drand48_nth(long k) {  // value of the k-th (1-based) drand48() call from the default state
    Drand48 r; double v = 0; for (long i = 0; i < k; ++i) v = r.next(); return v;
}
double orc_drand48_nth_seek(long k) {  // the same value through the jump-ahead of the stream (k >= 1)
    Drand48 r; r.seek(k - 1); return r.next();
}
void orc_drand48_seek(long long k) { g_drand48.seek(k); }
long long orc_drand48_position() { return g_drand48.pos;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------------------------
// NESTA (NestaUp::launch, NestaUp.cpp:442-910)
// ------------------------------------------------------------------------------------------------------------------
namespace {

// Stop test 1 (NestaUp.cpp:799-824): a ring of the last miniter objective values, its slot miniter-1 reset to
// LONG_MAX at each stage start while the other slots keep the previous stage's values (:645); from k = miniter on,
// qp = (mean - f) / mean; the first qp <= tolVar of a stage arms the rule and skips that iteration's ring write, the
// next one (armed) ends the stage after the current update.
template <class T> struct StopRule {
    std::vector<T> fmean;
    int miniter;
    bool qpDone = false;
    explicit StopRule(int mi) : fmean(std::max(mi, 1), (T)0), miniter(mi) {}
    void start_stage() {                                                                   // (:645,649)
        fmean[miniter - 1] = (T)LONG_MAX;
        qpDone = false;
    }
    // iteration k with objective fx; true when the stage ends after this iteration
    bool step(int k, T fx, T tolVar) {
        bool end = false, skip_write = false;
        if (k >= miniter) {
            T mean = 0;
            for (int ne = 0; ne < miniter; ++ne) mean += fmean[miniter - ne - 1];
            mean = mean / (T)miniter;
            const T qp = (mean - fx) / mean;
            if (qp <= tolVar) {
                if (qpDone) end = true;
                qpDone = true;
                skip_write = true;
            }
        }
        if (!skip_write) fmean[(miniter - 1) - ((k + 1) % miniter)] = fx;
        return end;
    }
};

// mc == nullptr: the reference's argsMC == nullptr path (x_ref = A^H y_s).  With mc, x_ref is the caller's xref
// (the previous reconstruction) and U, U^T, normest include the motion compensation.
template <class T> int nesta(const Prob<T>& P, Cx<T>* xref, const orc_params& prm, orc_report* rep,
                             const McOp* mc = nullptr) {
    const int F = P.F, C = P.C;
    const long N = P.N, nX = (long)F * N, nK = (long)F * C * N;
    std::vector<Cx<T>> yS(nK), aux(nK), Res(nK), AuxRes(nK);
    std::vector<Cx<T>> Ux(nX), Wk(nX), Xk(nX), Yk(nX), Zk(nX), Uk(nX), AuxFx(nX), Df(nX), ARes(nX), tmp(nX);
    std::vector<Cx<T>> auxMC(nX);

    // y_s = sqrt(N) * y ; x_ref = A^H(y_s) without motion compensation                    (:468-486)
    copy(nK, P.y, yS.data());
    const T sqrtN = (T)std::sqrt((T)(P.H * P.W));
    scal(nK, sqrtN, yS.data());
    if (!mc) op_At(P, yS.data(), xref, aux.data());
    op_U(mc, xref, Ux.data(), auxMC.data(), F, P.H, P.W);                                  // (:491)

    const T maxXref = max_abs(xref, nX);                                                   // (:521-545)
    const T maxUXref = max_abs(Ux.data(), nX);                                             // (:547-568)
    const int miniter = prm.miniter;
    const T lambda = (T)prm.lambda_i * maxXref;                                            // (:570-583)
    const T muf = (T)prm.mu_f * maxXref;
    T mu0 = (T)(0.9 * (double)maxUXref);                                                 // double product, as in C++
    int ne_it = 0;
    const T normU = normest<T>(F, N, &ne_it, mc, P.H, P.W);
    const T muL = lambda / (T)prm.La;
    if (muL > mu0) mu0 = muL;
    const T gamma = (T)std::pow((double)(muf / mu0), 1.0 / prm.maxIntIter);
    T mu = mu0;
    const T gammat = (T)std::pow((double)(T)prm.tolVar / 0.1, 1.0 / prm.maxIntIter);  // float / double literal
    T tolVar = (T)0.1;

    if (rep) {
        std::memset(rep, 0, sizeof(*rep));
        rep->normU = normU; rep->maxXref = maxXref; rep->maxUXref = maxUXref; rep->lambda = lambda;
        rep->muf = muf; rep->mu0 = mu0; rep->gamma = gamma; rep->gammat = gammat; rep->normest_iters = ne_it;
    }
    StopRule<T> stop(miniter);
    const T invSqrtN = (T)(1.0 / std::sqrt((double)(P.H * P.W)));

    for (int nl = 1; nl <= prm.maxIntIter; ++nl) {
        T Ak, Lmu, Lmu1;
        mu = mu * gamma;                                                                   // (:608-609)
        tolVar = tolVar * gammat;
        std::fill(Wk.begin(), Wk.end(), Cx<T>{0, 0});                                      // (:615-641)
        copy(nX, xref, Xk.data());
        std::fill(Yk.begin(), Yk.end(), Cx<T>{0, 0});
        std::fill(Zk.begin(), Zk.end(), Cx<T>{0, 0});
        Ak = 0;                                                                            // (:643-649)
        Lmu = normU / mu;
        stop.start_stage();
        Lmu = lambda * Lmu + (T)prm.La;
        Lmu1 = (T)1 / Lmu;
        int done_k = 0;
        for (int k = 0; k < prm.maxIter; ++k) {
            ++done_k;
            op_U(mc, Xk.data(), Uk.data(), auxMC.data(), F, P.H, P.W);                     // (:669)
            copy(nX, Uk.data(), AuxFx.data());                                             // (:675)
            vecnorm(Uk.data(), Uk.data(), nX, mu);                                         // (:680-683)
            const T normUk = nrm2(Uk.data(), nX);                                          // (:686)
            T fx = dotc_re(Uk.data(), AuxFx.data(), nX);                                   // (:691)
            op_Ut(mc, Uk.data(), Df.data(), tmp.data(), F, P.H, P.W);                      // (:696)
            scal(nX, (T)-1, Df.data());                                                    // (:698-704)
            op_A(P, Xk.data(), Res.data(), aux.data());                                    // (:710)
            scal(nK, invSqrtN, Res.data());                                                // (:712-718)
            axpy(nK, (T)-1, P.y, Res.data());                                              // (:720-727)
            copy(nK, Res.data(), AuxRes.data());                                           // (:732)
            const T l2 = sum_abspow2(Res.data(), nK);                                      // (:738-746)
            scal(nK, sqrtN, AuxRes.data());                                                // (:749-753)
            op_At(P, AuxRes.data(), ARes.data(), aux.data());                              // (:756)
            scal(nX, lambda, Df.data());                                                   // (:759-764)
            axpy(nX, (T)1, ARes.data(), Df.data());                                        // (:766-772)
            copy(nX, Xk.data(), Yk.data());                                                // (:776-779)
            axpy(nX, -Lmu1, Df.data(), Yk.data());                                         // (:781-786)

            const T l1 = (T)(fx - (mu / (T)2) * std::pow((double)normUk, 2));              // (:797-798)
            fx = (T)0.5 * l2 + lambda * l1;
            if (rep && rep->ntrace < ORC_TRACE) {
                rep->trace_f[rep->ntrace] = fx; rep->trace_l2[rep->ntrace] = l2; rep->trace_l1[rep->ntrace] = l1;
                rep->ntrace++;
            }
            if (prm.stoptest == 1 && stop.step(k, fx, tolVar)) k = prm.maxIter;            // (:799-824)
            const T apk = (T)0.5 * ((T)k + (T)1);                                          // (:829-831)
            Ak = Ak + apk;
            const T tauk = (T)2 / ((T)k + (T)3);
            axpy(nX, apk, Df.data(), Wk.data());                                           // (:837)
            copy(nX, xref, Zk.data());                                                     // (:843)
            axpy(nX, -Lmu1, Wk.data(), Zk.data());                                         // (:850)
            copy(nX, Zk.data(), Xk.data());                                                // (:858)
            scal(nX, tauk, Xk.data());                                                     // (:865)
            axpy(nX, (T)1 - tauk, Yk.data(), Xk.data());                                   // (:873)
        }
        copy(nX, Xk.data(), xref);                                                         // (:889)
        if (rep && nl <= ORC_MAX_STAGES) { rep->iters[nl - 1] = done_k; rep->stages = nl; }
    }
    if (rep) { rep->final_tolVar = tolVar; rep->final_mu = mu; }
    return 0;
}

}  // namespace

// rss.cl: out_f = (sqrt(sum_c |in_{f,c}|^2), 0), the sum taken in coil order in T
template <class T> void rsos(const Cx<T>* in, Cx<T>* out, int F, int C, long N) {
#pragma omp parallel for schedule(static)
    for (long p = 0; p < N; ++p)
        for (int f = 0; f < F; ++f) {
            T s = 0;
            for (int c = 0; c < C; ++c) {
                const Cx<T> v = in[((long)f * C + c) * N + p];
                s += v.re * v.re + v.im * v.im;
            }
            out[(long)f * N + p].re = std::sqrt(s);
            out[(long)f * N + p].im = 0;
        }
}

extern "C" {

// Full NESTA-TV reconstruction.  y: [F][C][H][W], S: [C][H][W], M: [F][H][W] u8, x_out: [F][H][W].
int orc_nesta_f32(const float* y, const float* S, const uint8_t* M, float* x_out, int F, int C, int H, int W,
                  const orc_params* prm, orc_report* rep) {
    if (!prm || prm->miniter < 1 || F < 1 || C < 1) return -1;
    Prob<float> P{F, C, H, W, (long)H * W, (const Cx<float>*)y, (const Cx<float>*)S, M};
    return nesta(P, (Cx<float>*)x_out, *prm, rep);
}
int orc_nesta_f64(const double* y, const double* S, const uint8_t* M, double* x_out, int F, int C, int H, int W,
                  const orc_params* prm, orc_report* rep) {
    if (!prm || prm->miniter < 1 || F < 1 || C < 1) return -1;
    Prob<double> P{F, C, H, W, (long)H * W, (const Cx<double>*)y, (const Cx<double>*)S, M};
    return nesta(P, (Cx<double>*)x_out, *prm, rep);
}

// SimpleMRIRecon (SimpleMRIRecon.cpp:59-101): x_f = sum_c conj(S_c) * IFFT2_{1/N}(y_{f,c})
int orc_simple_recon_f32(const float* y, const float* S, float* x, int F, int C, int H, int W) {
    long nK = (long)F * C * H * W;
    std::vector<Cx<float>> aux(nK);
    Prob<float> P{F, C, H, W, (long)H * W, nullptr, (const Cx<float>*)S, nullptr};
    op_At(P, (const Cx<float>*)y, (Cx<float>*)x, aux.data());
    return 0;
}
int orc_simple_recon_f64(const double* y, const double* S, double* x, int F, int C, int H, int W) {
    long nK = (long)F * C * H * W;
    std::vector<Cx<double>> aux(nK);
    Prob<double> P{F, C, H, W, (long)H * W, nullptr, (const Cx<double>*)S, nullptr};
    op_At(P, (const Cx<double>*)y, (Cx<double>*)x, aux.data());
    return 0;
}

// The stop test on a scripted objective sequence: f[0..n) feeds nstages stages of at most maxIter iterations each
// (tolVar fixed); iters[s] receives the iterations stage s ran.  Returns the number of values consumed, or -1 when f
// runs out first.
int orc_stop_rule_f64(const double* f, int n, int miniter, double tolVar, int maxIter, int nstages, int* iters) {
    if (miniter < 1 || maxIter < 1) return -1;
    StopRule<double> stop(miniter);
    int used = 0;
    for (int st = 0; st < nstages; ++st) {
        stop.start_stage();
        int done = 0;
        for (int k = 0; k < maxIter; ++k) {
            if (used >= n) return -1;
            ++done;
            if (stop.step(k, f[used++], tolVar)) k = maxIter;
        }
        iters[st] = done;
    }
    return used;
}

// Motion compensation of F frames of H x W (adjoint = 0: MotionCompensation, 1: AdjointMotionCompensation).
// xn: [2][F][H][W] float, x: [2][W][H] int, bb: bound box (1-based).  Returns -1 when the box leaves the frame.
int orc_mc_f64(const double* in, double* out, const float* xn, const int* x, const unsigned* bb, int F, int H, int W,
               int adjoint) {
    McOp m{xn, x, {}};
    if (!mc_box(H, W, bb, &m.box)) return -1;
    if (adjoint) mc_adj(m, (const Cx<double>*)in, (Cx<double>*)out, F, H, W);
    else mc_fwd(m, (const Cx<double>*)in, (Cx<double>*)out, F, H, W);
    return 0;
}
int orc_mc_f32(const float* in, float* out, const float* xn, const int* x, const unsigned* bb, int F, int H, int W,
               int adjoint) {
    McOp m{xn, x, {}};
    if (!mc_box(H, W, bb, &m.box)) return -1;
    if (adjoint) mc_adj(m, (const Cx<float>*)in, (Cx<float>*)out, F, H, W);
    else mc_fwd(m, (const Cx<float>*)in, (Cx<float>*)out, F, H, W);
    return 0;
}
double orc_normest_mc_f64(const float* xn, const int* x, const unsigned* bb, int F, int H, int W, int* iters) {
    McOp m{xn, x, {}};
    if (!mc_box(H, W, bb, &m.box)) return -1.0;
    return normest<double>(F, (long)H * W, iters, &m, H, W);
}
// NESTA with motion compensation: x_io holds x_ref (the previous reconstruction) on entry and the result on exit
int orc_nesta_mc_f64(const double* y, const double* S, const uint8_t* M, double* x_io, int F, int C, int H, int W,
                     const float* xn, const int* x, const unsigned* bb, const orc_params* prm, orc_report* rep) {
    if (!prm || prm->miniter < 1 || F < 1 || C < 1) return -1;
    McOp m{xn, x, {}};
    if (!mc_box(H, W, bb, &m.box)) return -1;
    Prob<double> P{F, C, H, W, (long)H * W, (const Cx<double>*)y, (const Cx<double>*)S, M};
    return nesta(P, (Cx<double>*)x_io, *prm, rep, &m);
}
int orc_nesta_mc_f32(const float* y, const float* S, const uint8_t* M, float* x_io, int F, int C, int H, int W,
                     const float* xn, const int* x, const unsigned* bb, const orc_params* prm, orc_report* rep) {
    if (!prm || prm->miniter < 1 || F < 1 || C < 1) return -1;
    McOp m{xn, x, {}};
    if (!mc_box(H, W, bb, &m.box)) return -1;
    Prob<float> P{F, C, H, W, (long)H * W, (const Cx<float>*)y, (const Cx<float>*)S, M};
    return nesta(P, (Cx<float>*)x_io, *prm, rep, &m);
}

void orc_rsos_f32(const float* in, float* out, int F, int C, long N) {
    rsos((const Cx<float>*)in, (Cx<float>*)out, F, C, N);
}

// SimpleMRIReconSOS.cpp: inverse 2-D FFT (scaled 1/N) of every coil image in place, then rss.cl
int orc_simple_recon_sos_f64(double* y, double* x, int F, int C, int H, int W) {
    fft2d((const Cx<double>*)y, (Cx<double>*)y, (long)F * C, H, W, +1);
    rsos((const Cx<double>*)y, (Cx<double>*)x, F, C, (long)H * W);
    return 0;
}

}  // extern "C"
