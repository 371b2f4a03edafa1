// Groupwise B-spline registration (the reference's GroupwiseRegistration: src/processes/GroupwiseRegistration.cpp,
// groupwiseRegistration/{creabspline,optimizer,bspline}.cpp and about 30 kernels in src/kernels/optimizer.cl,
// creabspline.cl, reductions.cl, initialize.cl): the cardiac motion estimator whose result (the deformed mesh xn, the
// coefficients T, the mesh x and the bound box) feeds motion-compensated NESTA.
//
// What the reference computes, for a square N x N image of F frames with a circular ROI M:
//   I0      = 500 |x| / max|x| (the normalised frames) and its central-difference gradients d1 (along columns) and
//             d2 (along rows), one-sided at the edges;
//   u_d     = sum_{p,q} T[d][f][lo(r)+p][lo(c)+q] B(r)[p] B(c)[q]: a cubic B-spline displacement with control spacing
//             Dp, for the box pixels (r, c) (the ROI's bounding box); xn = (c + 1 + u_0, r + 1 + u_1);
//   It      = I0 resampled bilinearly at xn; V = variance of It over frames; H = sum_M V;
//   dH      = sum over the support of control point (P, Q) of M dy dx_d beta(i) beta(j), with dy = 2 (It - mean)/F and
//             dx_d the gradient d_d resampled at xn;
//   T      -= Wn (dH - mean over frames of dH); Wn x1.2 when H fell, /2 when it rose; stop after nmax iterations or
//             when |Wn (dH - mean)| < et F 2 and the fall of H < eh H0.
// The reference also builds a regularisation gradient (optimizer.cl:317-406, reductions.cl:56-111, optimizer.cl:408-
// 486).  It is the same for every frame, so the frame-mean projection of updateTransformation (optimizer.cl:519-535)
// removes it exactly; it changes T only by rounding and is not computed here.  The regularisation enters the cost
// (sumLambdaMetric, optimizer.cl:138-187) only when the weights summed into an int accumulator
// (GroupwiseRegistration.cpp:441-443) are non-zero; then the eval pass adds it (gwr_reg_terms below).
//
// Layout in HBM: T [2][F][Cp][Cp] (the reference's TData, {Cp1, Cp0, F, nt}); xn [2][F][N][N]; I4 [F][N][N] float4
// {I0, d1, d2, 0} so one 16-byte load serves the three bilinear samples of a neighbour; w [2][F][L][L] (M dy dx_d over
// the L x L box).  Per iteration three launches: gwr_grad_update (one workgroup per control point: the support sums,
// the projection and the update of T, and the block's sum of squared steps), gwr_eval (16 box pixels x 16 frame lanes
// per workgroup: xn, It, V, w for the next gradient, the block's share of H) and gwr_step (one workgroup: the step
// logic, evolution / stopCondition, GroupwiseRegistration.cpp:456-548, on the device).  Every launch after the run
// has stopped returns at once, so the host enqueues all nmax iterations without synchronising.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ocl_internal.hpp"

namespace {

constexpr int GWR_PIX = 16;          // box pixels per eval workgroup
constexpr int GWR_FL = OCL_NT / GWR_PIX;   // frame lanes per eval workgroup (frames f = lane, lane + 16, ...)
constexpr int GWR_FMAX = 64;         // frames (eval keeps It, d1, d2 of its 64 pixels for every frame in LDS)
constexpr int GWR_SUP_MAX = 32;      // support diameter (the grad pass keeps a lane's support weights in registers)

struct GwrState {
    float Wn, p2, p4, eh;
    int it, stopped, nmax, flagW;
    unsigned maxbits;
    float H0, Hprev;
};

struct GwrDev {
    int N, F, L, box0, Cp, dp, use_lambda;
    float Dp, lam0, lam1, lam2, lam3;
    const float4* Btab;     // [N] value taps of a pixel coordinate
    const float4* B1tab;    // [N] first-derivative taps
    const float4* B2tab;    // [N] second-derivative taps
    const int* loT;         // [N] 0-based control index of a pixel's first tap
    const float* mask;      // [N][N]
    const int* gs;          // [Cp] 0-based first pixel of a control point's support
    const float* beta;      // [dp] support weights (the same for every control point)
    float4* I4;             // [F][N][N]
    float* T;               // [2][F][Cp][Cp]
    float* xn;              // [2][F][N][N]
    float* w;               // [2][F][L][L]
    float* reg;             // optional registered frames [F][N][N]
    double* partH;          // [eval blocks]
    double* partD;          // [Cp * Cp]
    float* hist;            // [4][nmax + 1]: H, Tdif, Hdif, Wn
    GwrState* st;
};

// ---------------------------------------------------------------------------------------------------- set-up passes
__global__ __launch_bounds__(OCL_NT) void k_gwr_init(GwrDev g, float Wn0, float p2, float eh, int nmax, int flagW,
                                                     int zeroT) {
    const long N2 = (long)g.N * g.N, tot = 2L * g.F * N2;
    for (long i = (long)blockIdx.x * OCL_NT + threadIdx.x; i < tot; i += (long)gridDim.x * OCL_NT) {
        const int d = (int)(i / ((long)g.F * N2));
        const long p = i % N2;
        const int r = (int)(p / g.N), c = (int)(p - (long)r * g.N);
        g.xn[i] = d == 0 ? (float)(c + 1) : (float)(r + 1);            // the identity mesh, 1-based
    }
    if (zeroT) {
        const long nT = 2L * g.F * g.Cp * g.Cp;
        for (long i = (long)blockIdx.x * OCL_NT + threadIdx.x; i < nT; i += (long)gridDim.x * OCL_NT) g.T[i] = 0.f;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        GwrState s{};
        s.Wn = Wn0;
        s.p2 = p2;
        s.eh = eh;
        s.it = 1;
        s.nmax = nmax;
        s.flagW = flagW;
        *g.st = s;
    }
}

// max |x| over every frame (GroupwiseRegistration.cpp:566-590); |x| >= 0, so the float bits order like the values
__global__ __launch_bounds__(OCL_NT) void k_gwr_absmax(const float2* __restrict__ x, long n, GwrState* st) {
    __shared__ double red[16];
    double m = 0.0;
    for (long i = (long)blockIdx.x * OCL_NT + threadIdx.x; i < n; i += (long)gridDim.x * OCL_NT) {
        const float2 v = x[i];
        m = fmax(m, (double)hypotf(v.x, v.y));
    }
    m = block_max(m, red);
    if (threadIdx.x == 0) atomicMax(&st->maxbits, __float_as_uint((float)m));
}

// dataNormalization (initialize.cl:87-102) and gradient (optimizer.cl:189-234): I4 = {I0, d1, d2, 0}
__global__ __launch_bounds__(OCL_NT) void k_gwr_prep(const float2* __restrict__ x, GwrDev g) {
    const int N = g.N;
    const long N2 = (long)N * N, tot = (long)g.F * N2;
    const float factor = 500.0f / __uint_as_float(g.st->maxbits);
    for (long i = (long)blockIdx.x * OCL_NT + threadIdx.x; i < tot; i += (long)gridDim.x * OCL_NT) {
        const long p = i % N2, fb = i - p;
        const int r = (int)(p / N), c = (int)(p - (long)r * N);
        auto I = [&](int rr, int cc) {
            const float2 v = x[fb + (long)rr * N + cc];
            return factor * hypotf(v.x, v.y);
        };
        const float v = I(r, c);
        const float d1 = c == 0 ? I(r, 1) - I(r, 0) : c == N - 1 ? I(r, N - 1) - I(r, N - 2)
                                                                 : (I(r, c + 1) - I(r, c - 1)) / 2;
        const float d2 = r == 0 ? I(1, c) - I(0, c) : r == N - 1 ? I(N - 1, c) - I(N - 2, c)
                                                                 : (I(r + 1, c) - I(r - 1, c)) / 2;
        g.I4[i] = make_float4(v, d1, d2, 0.f);
        if (g.reg) g.reg[i] = v;
    }
}

// ---------------------------------------------------------------------------------------------------- the eval pass
// Bilinear resampling with the interpolator's arithmetic (optimizer.cl:832-867); neighbours outside the frame read 0
__device__ __forceinline__ float4 gwr_bilinear(const float4* __restrict__ fr, int N, float nhor, float nver) {
    float4 q[4];
    const bool inside = nhor > -2.0f && nhor < (float)(N + 1) && nver > -2.0f && nver < (float)(N + 1);
    const int ph = inside ? (int)floorf(nhor) : 0, pv = inside ? (int)floorf(nver) : 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int h = ph + (k & 1), v = pv + (k >> 1);
        q[k] = (inside && h >= 0 && h < N && v >= 0 && v < N) ? fr[h + (long)v * N] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    const float wl = (float)(ph + 1) - nhor, wr = nhor - (float)ph;
    const float wu = (float)(pv + 1) - nver, wd = nver - (float)pv;
    float4 o;
    o.x = wu * (wl * q[0].x + wr * q[1].x) + wd * (wl * q[2].x + wr * q[3].x);
    o.y = wu * (wl * q[0].y + wr * q[1].y) + wd * (wl * q[2].y + wr * q[3].y);
    o.z = wu * (wl * q[0].z + wr * q[1].z) + wd * (wl * q[2].z + wr * q[3].z);
    o.w = 0.f;
    return o;
}

__device__ __forceinline__ float tap(const float4& b, int k) {
    return k == 0 ? b.x : k == 1 ? b.y : k == 2 ? b.z : b.w;
}

// sum_{p,q} T[lr+p][lc+q] a[p] b[q] over one Cp x Cp grid
__device__ __forceinline__ float gwr_spline(const float* __restrict__ Tg, int Cp, int lr, int lc, const float4& a,
                                            const float4& b) {
    float s = 0.f;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const float* row = Tg + (long)(lr + p) * Cp + lc;
        const float rq = fmaf(row[0], b.x, fmaf(row[1], b.y, fmaf(row[2], b.z, row[3] * b.w)));
        s = fmaf(tap(a, p), rq, s);
    }
    return s;
}

// The regularisation terms of one box pixel (row R, column C) and frame f, weighted as sumLambdaMetric weights them.
// Regularization (optimizer.cl:746-830) evaluates the spline with the control row taken from the pixel's column
// coordinate and the control column from its row coordinate, the row weights from B(R) and the column weights from
// B(C); this keeps that arrangement.
__device__ float gwr_reg_terms(const GwrDev& g, int R, int C, int f) {
    const int Cp = g.Cp;
    const int lR = g.loT[R], lC = g.loT[C];
    const float4 bR = g.Btab[R], bC = g.Btab[C], b1R = g.B1tab[R], b1C = g.B1tab[C], b2R = g.B2tab[R],
                 b2C = g.B2tab[C];
    const int f1 = f == 0 ? g.F - 1 : f - 1, f2 = f >= 2 ? f - 2 : f + g.F - 2;
    const float Dp = g.Dp, Dp2 = g.Dp * g.Dp;
    float acc = 0.f;
    for (int d = 0; d < 2; ++d) {
        const float* T0 = g.T + ((long)d * g.F + f) * Cp * Cp;
        const float* T1 = g.T + ((long)d * g.F + f1) * Cp * Cp;
        const float* T2 = g.T + ((long)d * g.F + f2) * Cp * Cp;
        float x0 = 0.f, x1 = 0.f, xx0 = 0.f, xx1 = 0.f, xy = 0.f, t1 = 0.f, t2 = 0.f;
        for (int n = 0; n < 4; ++n)          // control row lC + n, weights from the pixel row R
            for (int m = 0; m < 4; ++m) {    // control column lR + m, weights from the pixel column C
                const long o = (long)(lC + n) * Cp + lR + m;
                const float v = T0[o], s1 = T1[o], s2 = T2[o];
                x0 += v * (Dp * tap(bR, n) * tap(b1C, m));
                x1 += v * (Dp * tap(b1R, n) * tap(bC, m));
                xx0 += v * (Dp2 * tap(bR, n) * tap(b2C, m));
                xx1 += v * (Dp2 * tap(b2R, n) * tap(bC, m));
                xy += v * (Dp2 * tap(b1R, n) * tap(b1C, m));
                const float bb = tap(bR, n) * tap(bC, m);
                t1 += (v - s1) * bb;
                t2 += (v - 2.0f * s1 + s2) * bb;
            }
        if (d == 0) x0 += 1.f; else x1 += 1.f;
        acc += g.lam1 * 2 * (xy * xy) + g.lam3 * (t2 * t2) + g.lam2 * (t1 * t1);
        acc += g.lam0 * (x0 * x0) + g.lam1 * (xx0 * xx0) + g.lam0 * (x1 * x1) + g.lam1 * (xx1 * xx1);
    }
    return acc;
}

// mode 0: the evaluation before the loop (optimizer.cpp:236-259): the metric of the unwarped frames, dy from them, dx
// resampled at the starting coefficients.  mode 1: the evaluation that closes iteration it (optimizer.cpp:318-356).
template <int REG>
__global__ __launch_bounds__(OCL_NT) void k_gwr_eval(GwrDev g, int mode) {
    __shared__ float sI[GWR_FMAX][GWR_PIX], s1[GWR_FMAX][GWR_PIX], s2[GWR_FMAX][GWR_PIX];
    __shared__ float smean[GWR_PIX], sreg[GWR_FL][GWR_PIX];
    __shared__ double red[16];
    if (mode == 1 && g.st->stopped) return;
    const int N = g.N, F = g.F, L = g.L, Cp = g.Cp;
    const long N2 = (long)N * N, L2 = (long)L * L;
    const int px = threadIdx.x & (GWR_PIX - 1), fl = threadIdx.x / GWR_PIX;
    const long p = (long)blockIdx.x * GWR_PIX + px;
    const bool valid = p < L2;
    const int r = g.box0 + (valid ? (int)(p / L) : 0), c = g.box0 + (valid ? (int)(p % L) : 0);
    if (valid) {
        const float4 br = g.Btab[r], bc = g.Btab[c];
        const int lr = g.loT[r], lc = g.loT[c];
        const long pix = (long)r * N + c;
        for (int f = fl; f < F; f += GWR_FL) {
            const float u0 = gwr_spline(g.T + (long)f * Cp * Cp, Cp, lr, lc, br, bc);
            const float u1 = gwr_spline(g.T + ((long)F + f) * Cp * Cp, Cp, lr, lc, br, bc);
            const float xh = u0 + (float)(c + 1), xv = u1 + (float)(r + 1);
            g.xn[(long)f * N2 + pix] = xh;
            g.xn[((long)F + f) * N2 + pix] = xv;
            const float4 s = gwr_bilinear(g.I4 + (long)f * N2, N, xh - 1.0f, xv - 1.0f);
            const float it = mode == 0 ? g.I4[(long)f * N2 + pix].x : s.x;
            sI[f][px] = it;
            s1[f][px] = s.y;
            s2[f][px] = s.z;
            if (g.reg && mode == 1) g.reg[(long)f * N2 + pix] = it;
        }
        if (REG && mode == 1) {
            float a = 0.f;
            for (int f = fl; f < F; f += GWR_FL) a += gwr_reg_terms(g, r, c, f);
            sreg[fl][px] = a;
        }
    }
    __syncthreads();
    double hpix = 0.0;
    if (valid && fl == 0) {
        float s = 0.f;
        for (int f = 0; f < F; ++f) s += sI[f][px];
        const float mean = s / (float)F;
        float v = 0.f;
        for (int f = 0; f < F; ++f) {
            const float e = sI[f][px] - mean;
            v += e * e;
        }
        v = v / (float)F;
        smean[px] = mean;
        if (REG && mode == 1)
            for (int q = 0; q < GWR_FL; ++q) v += sreg[q][px];
        hpix = (double)(g.mask[(long)r * N + c] * v);
    }
    __syncthreads();
    if (valid) {
        const float m = g.mask[(long)r * N + c], mean = smean[px], factor = (float)(2.0 / F);
        for (int f = fl; f < F; f += GWR_FL) {
            const float dy = (sI[f][px] - mean) * factor;
            g.w[(long)f * L2 + p] = m * (dy * s1[f][px]);
            g.w[((long)F + f) * L2 + p] = m * (dy * s2[f][px]);
        }
    }
    double v1[1] = {hpix};
    block_sum<1>(v1, red);
    if (threadIdx.x == 0) g.partH[blockIdx.x] = v1[0];
}

// The step logic after an eval pass (evolution / stopCondition, GroupwiseRegistration.cpp:456-548), one workgroup:
// H and the squared step norm from every block's partial sums in a fixed order.  A launch of its own: taking it in
// the eval pass's last workgroup would need an agent-scope release fence in every workgroup, which writes back the
// L2 of its XCD.
__global__ __launch_bounds__(OCL_NT) void k_gwr_step(GwrDev g, int mode, int nblocks) {
    __shared__ double red[16];
    GwrState* st = g.st;
    if (mode == 1 && st->stopped) return;
    double sums[2] = {0.0, 0.0};
    for (int b = threadIdx.x; b < nblocks; b += OCL_NT) sums[0] += g.partH[b];
    if (mode == 1)
        for (int b = threadIdx.x; b < g.Cp * g.Cp; b += OCL_NT) sums[1] += g.partD[b];
    block_sum<2>(sums, red);
    if (threadIdx.x != 0) return;
    const float Hf = (float)sums[0];
    const int nmax = st->nmax;
    if (mode == 0) {
        st->H0 = Hf;
        st->Hprev = Hf;
        st->p4 = Hf * st->eh;
        g.hist[0] = Hf;
        return;
    }
    const int it = st->it;
    const float Tdif = sqrtf((float)sums[1]);
    const float Hdif = st->Hprev - Hf;
    float Wn = st->Wn;
    if (st->flagW) Wn = Hdif > 0 ? (float)((double)Wn * 1.2) : Wn / 2;
    g.hist[it] = Hf;
    g.hist[(nmax + 1) + it] = Tdif;
    g.hist[2 * (nmax + 1) + it] = Hdif;
    g.hist[3 * (nmax + 1) + it] = Wn;
    st->Wn = Wn;
    st->Hprev = Hf;
    st->stopped = (it == nmax) || (Tdif < st->p2 && Hdif < st->p4);
    st->it = it + 1;
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ---------------------------------------------------------------------------------------------------- the update pass
// One workgroup per control point (P, Q): dH[d][f] = sum over the support of w beta(i) beta(j) (gradientJoint,
// auxiliarMask, auxiliarCost, cost6DReduction, cost6D: optimizer.cl:442-486, 938-1025, reductions.cl:113-142), then
// updateTransformation (optimizer.cl:519-535): T -= Wn (dH - mean over frames).
// KT: support steps per lane, K = ceil(dp^2 / 64) rounded up to 4, 8 or 16 (steps k >= K carry weight 0)
template <int KT>
__global__ __launch_bounds__(OCL_NT) void k_gwr_grad_update(GwrDev g) {
    __shared__ float sdH[2 * GWR_FMAX];
    __shared__ double red[16];
    if (g.st->stopped) return;
    const int Cp = g.Cp, dp = g.dp, F = g.F, L = g.L;
    const long L2 = (long)L * L;
    const int P = blockIdx.x / Cp, Q = blockIdx.x - P * Cp;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = g.gs[P] - g.box0, c0 = g.gs[Q] - g.box0;
    const int K = (dp * dp + 63) / 64;            // <= KT (the launch picks KT)
    int off[KT];
    float wt[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) {
        off[k] = -1;
        wt[k] = 0.f;
        if (k >= K) continue;
        const int s = lane + 64 * k;
        if (s >= dp * dp) continue;
        const int i = s / dp, j = s - (s / dp) * dp;
        const int rr = r0 + i, cc = c0 + j;
        if (rr < 0 || rr >= L || cc < 0 || cc >= L) continue;      // outside the box the mask is zero
        off[k] = rr * L + cc;
        wt[k] = g.beta[i] * g.beta[j];
    }
    // (d, f) planes handled by a wave ten at a time (all of them for up to 20 frames), so their support loads are in
    // flight together
    constexpr int GB = 10, NW = OCL_NT / 64;
    for (int pr0 = wave; pr0 < 2 * F; pr0 += NW * GB) {
        float a[GB];
#pragma unroll
        for (int b = 0; b < GB; ++b) a[b] = 0.f;
        // branch-free loads (a branch per load puts each in a block of its own, waited on there): a support point
        // outside the box reads position 0 with weight 0, which adds exactly 0; planes past 2F read the last plane
        // and their sums are never stored.  Per plane the support terms are summed in k order, as before.
        int o[KT];
#pragma unroll
        for (int k = 0; k < KT; ++k) o[k] = off[k] < 0 ? 0 : off[k];
#pragma unroll
        for (int b = 0; b < GB; ++b) {
            const float* wp = g.w + (long)min(pr0 + NW * b, 2 * F - 1) * L2;
            float s = 0.f;
#pragma unroll
            for (int k = 0; k < KT; ++k) s = fmaf(wp[o[k]], wt[k], s);
            a[b] = s;
        }
#pragma unroll
        for (int b = 0; b < GB; ++b) {
            const float sum = wave_sum_f(a[b]);
            const int pr = pr0 + NW * b;
            if (lane == 0 && pr < 2 * F) sdH[pr] = sum;
        }
    }
    __syncthreads();
    double dif2 = 0.0;
    if (threadIdx.x < 2 * F) {
        const int d = threadIdx.x / F, f = threadIdx.x - d * F;
        float s = 0.f;
        for (int q = 0; q < F; ++q) s += sdH[d * F + q];
        const float mean = s / (float)F;
        const float Dif = g.st->Wn * (sdH[threadIdx.x] - mean);
        float* t = g.T + ((long)d * F + f) * Cp * Cp + (long)P * Cp + Q;
        *t = *t - Dif;
        dif2 = (double)(Dif * Dif);
    }
    double v1[1] = {dif2};
    block_sum<1>(v1, red);
    if (threadIdx.x == 0) g.partD[blockIdx.x] = v1[0];
}

// ---------------------------------------------------------------------------------------------------- host tables
// The reference's B-spline functions (groupwiseRegistration/bspline.cpp:113-131, 198-219, 267-285), E = 3: the first
// branch of the cubic and of its first derivative is evaluated in double (pow), the rest in float.
float h_bspline3(float x) {
    const float val = std::fabs(x);
    if (val < 1) return (float)(0.5 * std::pow((double)val, 3) - std::pow((double)val, 2) + 0.6666);
    if (val < 2) return (2 - val) * (2 - val) * (2 - val) / 6;
    return 0.f;
}
float h_b1spline3(float x) {
    if (x <= 1 && x >= 0) return (float)(1.5 * std::pow((double)x, 2) - 2 * (double)x);
    if (x >= -1 && x < 0) return (float)(-1.5 * std::pow((double)x, 2) - 2 * (double)x);
    if (x > 1 && x < 2) return -(x - 2) * (x - 2) / 2;
    if (x < -1 && x > -2) return (x + 2) * (x + 2) / 2;
    return 0.f;
}
float h_b2spline3(float x) {
    const float val = std::fabs(x);
    if (val < 1) return 3 * val - 2;
    if (val < 2) return -val + 2;
    return 0.f;
}

}  // namespace

struct ocl_gwr {
    ocl_ctx* ctx = nullptr;
    int N = 0, F = 0, L = 0, box0 = 0, Cp = 0, dp = 0, areaX = 0, use_lambda = 0, nmax = 0, flagW = 1;
    unsigned bb[4] = {0, 0, 0, 0};
    float Wn0 = 0, p2 = 0, eh = 0, Dpf = 0, lam[4] = {0, 0, 0, 0};
    int eval_blocks = 0;
    // device buffers
    float4 *Btab = nullptr, *B1tab = nullptr, *B2tab = nullptr, *I4 = nullptr;
    int *loT = nullptr, *gs = nullptr, *mesh = nullptr;
    float *mask = nullptr, *beta = nullptr, *w = nullptr, *hist = nullptr;
    double *partH = nullptr, *partD = nullptr;
    GwrState* st = nullptr;
    std::vector<float> hist_host;
};

namespace {

void gwr_free(ocl_gwr* g) {
    void* bufs[] = {g->Btab, g->B1tab, g->B2tab, g->I4, g->loT, g->gs, g->mesh, g->mask, g->beta, g->w, g->hist,
                    g->partH, g->partD, g->st};
    for (void* b : bufs)
        if (b) (void)dev_free(b);
}

template <class Tp>
hipError_t upload(ocl_ctx* c, Tp** dst, const std::vector<Tp>& v) {
    hipError_t e = dev_malloc((void**)dst, sizeof(Tp) * v.size());
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(*dst, v.data(), sizeof(Tp) * v.size(), hipMemcpyHostToDevice, c->stream);
}

// GroupwiseRegistration::init (GroupwiseRegistration.cpp:93-419) and creabspline (creabspline.cpp:57-258) for a square
// image with a circular ROI: geometry checks and the tables the kernels read.
int gwr_plan(ocl_gwr* g, const ocl_gwr_params* prm) {
    const int N = g->N, F = g->F;
    // setMask (GroupwiseRegistration.cpp:485-535): X[i + j N] for i over columns, j over rows
    const int radius2 = (int)std::pow((double)prm->radius, 2);
    const float center = (float)(N + 1) / 2;
    std::vector<float> mask((size_t)N * N, 0.f);
    int mn_i = N, mx_i = 0, mn_j = N, mx_j = 0, area = 0;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            const float Cv = (float)std::pow((double)(1 - center + (float)j), 2);
            const float Dv = (float)std::pow((double)(1 - center + (float)i), 2);
            if (Cv + Dv <= radius2) {
                mask[(size_t)i + (size_t)j * N] = 1.f;
                mn_i = std::min(mn_i, i), mx_i = std::max(mx_i, i), mn_j = std::min(mn_j, j), mx_j = std::max(mx_j, j);
                ++area;
            }
        }
    if (area == 0) return fail(OCL_E_INVALID, "ocl_gwr_create: the ROI radius selects no pixel");
    const unsigned bb[4] = {(unsigned)mn_i + 1, (unsigned)mn_j + 1, (unsigned)mx_i + 1, (unsigned)mx_j + 1};
    std::memcpy(g->bb, bb, sizeof bb);
    g->areaX = area;
    const int L = (int)(bb[2] - bb[0] + 1);
    if ((int)(bb[3] - bb[1] + 1) != L || bb[0] != bb[1])
        return fail(OCL_E_UNSUPPORTED, "ocl_gwr_create: asymmetric ROI bounding box");
    g->L = L;
    g->box0 = (int)bb[0] - 1;
    const int margin = N / 40;
    if ((int)bb[1] - margin - 2 < 0 || (int)bb[3] + margin > N - 1)
        return fail(OCL_E_UNSUPPORTED, "ocl_gwr_create: the ROI and its interpolation margin leave the frame");
    // Bspline ts (GroupwiseRegistration.cpp:177-196)
    const int E = 3, Dp = prm->Dp[0];
    const float c0 = (float)(bb[0] + bb[2]) / 2, c1 = (float)(bb[1] + bb[3]) / 2;
    const float rp = (float)((E + 1) * Dp / 2);
    int C[2][2];
    const float cc[2] = {c0, c1};
    for (int l = 0; l < 2; ++l)
        for (int b = 0; b < 2; ++b) C[l][b] = (int)std::floor((std::fabs(cc[l] - (float)bb[l + b * 2]) + rp) / Dp);
    const int Cp0 = C[0][0] + C[0][1] + 1, Cp1 = C[1][0] + C[1][1] + 1;
    const int dp = (int)(2 * std::ceil(rp) + 1);
    if (Cp0 != Cp1 || c0 != c1 || C[0][0] != C[1][0])
        return fail(OCL_E_UNSUPPORTED, "ocl_gwr_create: asymmetric control grid");
    if (dp > GWR_SUP_MAX) return fail(OCL_E_UNSUPPORTED, "ocl_gwr_create: B-spline support above 32 pixels");
    const int Cp = Cp0;
    g->Cp = Cp;
    g->dp = dp;
    g->Dpf = (float)Dp;
    // per-pixel taps: first control index (coef, creabspline.cpp:96-109, rebased at :140-151) and the value and
    // derivative taps (BBAux, :113-131)
    std::vector<int> loT(N);
    std::vector<float4> Bt(N), B1t(N), B2t(N);
    std::vector<float> coef(N);
    for (int i = 0; i < N; ++i) coef[i] = std::floor((float)(i + 1 - c0) / Dp) + ((E + (-1.0)) / 2) * (-1.0);
    const float off = coef[bb[0]];
    for (int n = 0; n < N; ++n) {
        loT[n] = (int)(coef[n] + 1 - off) - 1;
        float v[3][4];
        for (int t = 0; t < 4; ++t) {
            const int k = (int)coef[n] + t;
            const float arg = ((float)(n + 1) - c0) / Dp - (float)k;
            v[0][t] = h_bspline3(arg), v[1][t] = h_b1spline3(arg), v[2][t] = h_b2spline3(arg);
        }
        Bt[n] = make_float4(v[0][0], v[0][1], v[0][2], v[0][3]);
        B1t[n] = make_float4(v[1][0], v[1][1], v[1][2], v[1][3]);
        B2t[n] = make_float4(v[2][0], v[2][1], v[2][2], v[2][3]);
    }
    for (int n = g->box0; n < g->box0 + L; ++n)
        if (loT[n] < 0 || loT[n] + 3 >= Cp)
            return fail(OCL_E_UNSUPPORTED, "ocl_gwr_create: a box pixel's taps leave the control grid");
    // control points and supports (controlPointsLocation, creabspline.cl:50-60; coefg, creabspline.cpp:161-167;
    // BBAux1, :176-212; clamp, :226-236)
    const float crp = std::ceil(rp);
    std::vector<int> gs(Cp);
    std::vector<float> beta(dp);
    for (int n = 0; n < Cp; ++n) {
        const float pu = c0 + (float)Dp * (float)(-C[0][0] + n);
        const float lo = std::ceil(pu - crp);
        if (lo < 1 || lo + (dp - 1) > N)
            return fail(OCL_E_UNSUPPORTED, "ocl_gwr_create: a control point's support leaves the frame");
        gs[n] = (int)lo - 1;
        for (int s = 0; s < dp; ++s) {
            const float b = h_bspline3(((float)((int)lo + s) - pu) / Dp);
            if (n == 0) beta[s] = b;
            else if (b != beta[s])
                return fail(OCL_E_UNSUPPORTED, "ocl_gwr_create: support weights differ between control points");
        }
    }
    g->Wn0 = prm->W / (float)area;
    g->p2 = prm->et * (float)F * 2.0f;
    g->eh = prm->eh;
    g->nmax = (int)prm->nmax;
    g->flagW = prm->flagW ? 1 : 0;
    int acc = 0;                       // sumLambda: std::accumulate with an int accumulator
    for (int k = 0; k < 4; ++k) {
        g->lam[k] = prm->lambda[k];
        acc = (int)((float)acc + prm->lambda[k]);
    }
    g->use_lambda = acc != 0;
    // the mesh x (GroupwiseRegistration.cpp:150-164): x[i + j N] = j + 1, x[i + j N + N^2] = i + 1
    std::vector<int> mesh((size_t)2 * N * N);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            mesh[(size_t)i + (size_t)j * N] = j + 1;
            mesh[(size_t)i + (size_t)j * N + (size_t)N * N] = i + 1;
        }
    ocl_ctx* c = g->ctx;
    g->eval_blocks = (int)(((long)L * L + GWR_PIX - 1) / GWR_PIX);
    HIPCHK(upload(c, &g->Btab, Bt));
    HIPCHK(upload(c, &g->B1tab, B1t));
    HIPCHK(upload(c, &g->B2tab, B2t));
    HIPCHK(upload(c, &g->loT, loT));
    HIPCHK(upload(c, &g->gs, gs));
    HIPCHK(upload(c, &g->beta, beta));
    HIPCHK(upload(c, &g->mask, mask));
    HIPCHK(upload(c, &g->mesh, mesh));
    HIPCHK(dev_malloc((void**)&g->I4, sizeof(float4) * (size_t)F * N * N));
    HIPCHK(dev_malloc((void**)&g->w, sizeof(float) * 2 * (size_t)F * L * L));
    HIPCHK(dev_malloc((void**)&g->hist, sizeof(float) * 4 * (size_t)(g->nmax + 1)));
    HIPCHK(dev_malloc((void**)&g->partH, sizeof(double) * g->eval_blocks));
    HIPCHK(dev_malloc((void**)&g->partD, sizeof(double) * Cp * Cp));
    HIPCHK(dev_malloc((void**)&g->st, sizeof(GwrState)));
    HIPCHK(hipStreamSynchronize(c->stream));
    return OCL_OK;
}

GwrDev gwr_dev(const ocl_gwr* g, float* T, float* xn, float* reg) {
    GwrDev d{};
    d.N = g->N, d.F = g->F, d.L = g->L, d.box0 = g->box0, d.Cp = g->Cp, d.dp = g->dp, d.use_lambda = g->use_lambda;
    d.Dp = g->Dpf, d.lam0 = g->lam[0], d.lam1 = g->lam[1], d.lam2 = g->lam[2], d.lam3 = g->lam[3];
    d.Btab = g->Btab, d.B1tab = g->B1tab, d.B2tab = g->B2tab, d.loT = g->loT, d.mask = g->mask, d.gs = g->gs;
    d.beta = g->beta, d.I4 = g->I4, d.T = T, d.xn = xn, d.w = g->w, d.reg = reg, d.partH = g->partH;
    d.partD = g->partD, d.hist = g->hist, d.st = g->st;
    return d;
}

}  // namespace

extern "C" int ocl_gwr_create(ocl_ctx* c, int F, int H, int W, const ocl_gwr_params* prm, ocl_gwr** out) {
    if (!c || !prm || !out) return fail(OCL_E_INVALID, "ocl_gwr_create: null argument");
    *out = nullptr;
    if (H != W) return fail(OCL_E_UNSUPPORTED, "ocl_gwr_create: the registration needs a square image");
    if (H < 8 || F < 2 || F > GWR_FMAX)
        return fail(OCL_E_UNSUPPORTED, "ocl_gwr_create: needs H >= 8 and 2 <= F <= 64 frames");
    if (prm->radius == 0)
        return fail(OCL_E_UNSUPPORTED, "ocl_gwr_create: the wholebox ROI (radius 0) is not supported: the reference "
                                       "fills its bound box in an order its consumers do not read");
    if (prm->E != 3) return fail(OCL_E_UNSUPPORTED, "ocl_gwr_create: only cubic B-splines (E = 3)");
    if (prm->Dp[0] != prm->Dp[1] || prm->Dp[0] <= 0)
        return fail(OCL_E_UNSUPPORTED, "ocl_gwr_create: needs Dp[0] == Dp[1] > 0");
    if (prm->nmax < 1 || prm->nmax > 100000) return fail(OCL_E_INVALID, "ocl_gwr_create: nmax out of range");
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    ocl_gwr* g = new ocl_gwr();
    g->ctx = c;
    g->N = H;
    g->F = F;
    const int rc = gwr_plan(g, prm);
    if (rc != OCL_OK) {
        const std::string m = ocl_last_error();
        gwr_free(g);
        delete g;
        return fail(rc, m);
    }
    *out = g;
    return OCL_OK;
}

extern "C" int ocl_gwr_destroy(ocl_gwr* g) {
    if (!g) return OCL_OK;
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    gwr_free(g);
    delete g;
    return OCL_OK;
}

extern "C" int ocl_gwr_info(const ocl_gwr* g, ocl_gwr_info_t* info) {
    if (!g || !info) return fail(OCL_E_INVALID, "ocl_gwr_info: null argument");
    info->Cp = g->Cp;
    info->support = g->dp;
    info->box = g->L;
    info->areaX = g->areaX;
    info->use_lambda = g->use_lambda;
    for (int k = 0; k < 4; ++k) info->bound_box[k] = g->bb[k];
    info->T_elems = 2L * g->F * g->Cp * g->Cp;
    info->xn_elems = 2L * g->F * g->N * g->N;
    info->mesh = g->mesh;
    return OCL_OK;
}

extern "C" int ocl_gwr_run(ocl_gwr* g, const void* x, float* T, int T_in, float* xn, float* registered,
                           ocl_gwr_report* report) {
    if (!g || !x || !T || !xn) return fail(OCL_E_INVALID, "ocl_gwr_run: null argument");
    ocl_ctx* c = g->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    const GwrDev d = gwr_dev(g, T, xn, registered);
    const long tot = (long)g->F * g->N * g->N;
    const int grid = grid_for(2 * tot);
    hipLaunchKernelGGL(k_gwr_init, dim3(grid), dim3(OCL_NT), 0, c->stream, d, g->Wn0, g->p2, g->eh, g->nmax,
                       g->flagW, T_in ? 0 : 1);
    HIPCHK(hipGetLastError());
    // one atomic per workgroup into a single word: few workgroups, each striding over the frames
    hipLaunchKernelGGL(k_gwr_absmax, dim3(std::min(grid_for(tot), 2 * c->num_cus)), dim3(OCL_NT), 0, c->stream,
                       (const float2*)x, tot, g->st);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_gwr_prep, dim3(grid_for(tot)), dim3(OCL_NT), 0, c->stream, (const float2*)x, d);
    HIPCHK(hipGetLastError());
    void (*eval)(GwrDev, int) = g->use_lambda ? k_gwr_eval<1> : k_gwr_eval<0>;
    hipLaunchKernelGGL(eval, dim3(g->eval_blocks), dim3(OCL_NT), 0, c->stream, d, 0);
    hipLaunchKernelGGL(k_gwr_step, dim3(1), dim3(OCL_NT), 0, c->stream, d, 0, g->eval_blocks);
    HIPCHK(hipGetLastError());
    for (int it = 1; it <= g->nmax; ++it) {
        const int K = (g->dp * g->dp + 63) / 64;     // dp <= GWR_SUP_MAX = 32: K <= 16
        const auto grad = K <= 4 ? k_gwr_grad_update<4> : K <= 8 ? k_gwr_grad_update<8> : k_gwr_grad_update<16>;
        hipLaunchKernelGGL(grad, dim3(g->Cp * g->Cp), dim3(OCL_NT), 0, c->stream, d);
        hipLaunchKernelGGL(eval, dim3(g->eval_blocks), dim3(OCL_NT), 0, c->stream, d, 1);
        hipLaunchKernelGGL(k_gwr_step, dim3(1), dim3(OCL_NT), 0, c->stream, d, 1, g->eval_blocks);
    }
    HIPCHK(hipGetLastError());
    GwrState st;
    g->hist_host.resize(4 * (size_t)(g->nmax + 1));
    HIPCHK(hipMemcpyAsync(&st, g->st, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(g->hist_host.data(), g->hist, sizeof(float) * g->hist_host.size(), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int iters = st.it - 1;
    if (report) {
        std::memset(report, 0, sizeof *report);
        report->iters = iters;
        report->H0 = st.H0;
        report->Wn = st.Wn;
        std::memcpy(&report->oldmax, &st.maxbits, sizeof(float));
        const int n = std::min(iters, OCL_GWR_TRACE - 1);
        report->ntrace = n;
        const size_t S = (size_t)g->nmax + 1;
        report->H[0] = st.H0;
        for (int k = 1; k <= n; ++k) {
            report->H[k] = g->hist_host[k];
            report->Tdif[k - 1] = g->hist_host[S + k];
            report->Hdif[k - 1] = g->hist_host[2 * S + k];
            report->Wn_trace[k - 1] = g->hist_host[3 * S + k];
        }
    }
    if (!(st.maxbits > 0) || !std::isfinite(st.H0))
        return fail(OCL_E_INVALID, "ocl_gwr_run: the input frames are zero or not finite");
    return OCL_OK;
}
