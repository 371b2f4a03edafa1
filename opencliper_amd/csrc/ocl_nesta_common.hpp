// Pieces shared by the FP32 (ocl_nesta.hip) and split-precision (ocl_nesta_mfma.hip) fused NESTA passes: slice lookup
// in a launch table, the temporal-TV reduction workgroups of pass B, and the per-pixel NESTA update of pass CA.
#pragma once
#include "ocl_kernels.hpp"

__device__ __forceinline__ int find_slice(const int* __restrict__ off, int ns, int b) {
    int lo = 0, hi = ns;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= b) lo = mid; else hi = mid;
    }
    return lo;
}

// Temporal-TV terms of x_k for one chunk of one frame (NestaUp.cpp:669-691): sum |u|^2 and sum Re(conj(u) Ux) with
// Ux = x_f - x_{f-1} (periodic) and u = Ux / max(mu, |Ux|).  Written to partU[2 uid .. 2 uid + 1].
__device__ __forceinline__ void nesta_U_task(const SliceDesc& d, int uid, double* red) {
    const NState* st = d.st;
    const int per_f = (int)((d.N + d.u_chunk - 1) / d.u_chunk);
    const int f = uid / per_f;
    const long p0 = (long)(uid - f * per_f) * d.u_chunk;
    const long p1 = min(d.N, p0 + d.u_chunk);
    const int fm = f == 0 ? d.F - 1 : f - 1;
    const float2* xk = st->cur ? d.x1 : d.x0;
    const float mu = (float)st->mu;
    double v[2] = {0.0, 0.0};
    for (long p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
        const float2 ux = c_sub(xk[(long)f * d.N + p], xk[(long)fm * d.N + p]);
        const float nf = fmaxf(mu, sqrtf(c_abs2(ux)));
        const float2 u = make_float2(ux.x / nf, ux.y / nf);
        v[0] += (double)u.x * u.x + (double)u.y * u.y;
        v[1] += (double)u.x * ux.x + (double)u.y * ux.y;
    }
    block_sum<2>(v, red);
    if (threadIdx.x == 0) { d.partU[2 * uid] = v[0]; d.partU[2 * uid + 1] = v[1]; }
}

// Coefficients of the iteration prepared by the scalar step, read once per workgroup.
struct UpdCoef {
    const float2* xk;
    float2* xo;
    float mu, lam, L1, apk, tau;
    int sstart;
};

__device__ __forceinline__ UpdCoef upd_coef(const SliceDesc& d) {
    const NState* st = d.st;
    UpdCoef u;
    u.xk = st->it_cur ? d.x1 : d.x0;
    u.xo = st->it_cur ? d.x0 : d.x1;
    u.mu = (float)st->it_mu;
    u.lam = (float)st->lambda;
    u.L1 = (float)st->it_L1;
    u.apk = (float)st->it_apk;
    u.tau = (float)st->it_tau;
    u.sstart = st->it_sstart;
    return u;
}

// NESTA update of one pixel p of frame f (NestaUp.cpp:696-875): df = lambda (u_f - u_{f+1}) + A^H r (the temporal-TV
// gradient D^T u with the periodic wrap), y_k = x_k - df / L_mu, W_k += apk df, z_k = x_ref - W_k / L_mu,
// x_{k+1} = tau z_k + (1 - tau) y_k.  The first iteration of a stage sets x_ref = x_k and W_k = 0.
// Writes W_k, x_ref (stage start) and x_{k+1}; returns x_{k+1}.
__device__ __forceinline__ float2 nesta_update_px(const SliceDesc& d, const UpdCoef& u, int f, int fm, int fp, long p,
                                                  float2 ares) {
    const long N = d.N;
    const float2 xf = u.xk[(long)f * N + p], xm = u.xk[(long)fm * N + p], xp = u.xk[(long)fp * N + p];
    const float2 uxf = c_sub(xf, xm), uxp = c_sub(xp, xf);
    const float nf = fmaxf(u.mu, sqrtf(c_abs2(uxf))), np = fmaxf(u.mu, sqrtf(c_abs2(uxp)));
    const float2 uf = make_float2(uxf.x / nf, uxf.y / nf), up = make_float2(uxp.x / np, uxp.y / np);
    const float2 df = make_float2(u.lam * (uf.x - up.x) + ares.x, u.lam * (uf.y - up.y) + ares.y);
    float2 xr, wo;
    if (u.sstart) {
        xr = xf;
        d.xref[(long)f * N + p] = xf;
        wo = make_float2(0.f, 0.f);
    } else {
        xr = d.xref[(long)f * N + p];
        wo = d.Wk[(long)f * N + p];
    }
    const float2 yk = make_float2(xf.x - u.L1 * df.x, xf.y - u.L1 * df.y);
    const float2 wn = make_float2(wo.x + u.apk * df.x, wo.y + u.apk * df.y);
    d.Wk[(long)f * N + p] = wn;
    const float2 z = make_float2(xr.x - u.L1 * wn.x, xr.y - u.L1 * wn.y);
    const float2 xn = make_float2(u.tau * z.x + (1.0f - u.tau) * yk.x, u.tau * z.y + (1.0f - u.tau) * yk.y);
    u.xo[(long)f * N + p] = xn;
    return xn;
}
