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

// Temporal-TV terms of x_k (of MC(x_k) with motion compensation) over one chunk of pixels and one group of
// OCL_U_FRAMES frames (NestaUp.cpp:669-691):
// sum |u|^2 and sum Re(conj(u) Ux) with Ux = x_f - x_{f-1} (periodic in f) and u = Ux / max(mu, |Ux|).  Each thread
// walks the group's frames with the previous frame in registers, so x_k is read (1 + 1 / OCL_U_FRAMES) times.
// Written to partU[2 uid + {0, 1}].
__device__ __forceinline__ void nesta_U_task(const SliceDesc& d, int uid, double* red) {
    const NState* st = d.st;
    const int nchunks = (int)((d.N + d.u_chunk - 1) / d.u_chunk);
    const int fg = uid / nchunks, pc = uid - fg * nchunks;
    const int f0 = fg * OCL_U_FRAMES, f1 = min(d.F, f0 + OCL_U_FRAMES);
    const long p0 = (long)pc * d.u_chunk;
    const long p1 = min(d.N, p0 + d.u_chunk);
    const float2* xk = d.mcz ? d.mcz : (st->cur ? d.x1 : d.x0);
    const float mu = (float)st->mu;
    double v[2] = {0.0, 0.0};
    for (long p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
        float2 prev = ldg2(xk + (long)(f0 == 0 ? d.F - 1 : f0 - 1) * d.N, (unsigned)p);
#pragma unroll 4
        for (int f = f0; f < f1; ++f) {
            const float2 cur = ldg2(xk + (long)f * d.N, (unsigned)p);
            const float2 ux = c_sub(cur, prev);
            const float nf = fmaxf(mu, sqrtf(c_abs2(ux)));
            const float2 u = make_float2(ux.x / nf, ux.y / nf);
            v[0] += (double)u.x * u.x + (double)u.y * u.y;
            v[1] += (double)u.x * ux.x + (double)u.y * ux.y;
            prev = cur;
        }
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

// The same coefficients with the frame-f rows of every image buffer resolved once per workgroup (wave-uniform bases for
// 32-bit indexed global accesses).
struct UpdRow {
    const float2 *xf, *xm, *xp, *xref_in;
    const float2* g;      // motion compensation: MC^T(D^T u) of frame f, replacing the local temporal-TV gradient
    float2 *xref_out, *Wk, *xo;
    float mu, imu, lam, L1, apk, tau;
    int sstart;
};

__device__ __forceinline__ UpdRow upd_row(const SliceDesc& d, int f) {
    const UpdCoef u = upd_coef(d);
    const int fm = f == 0 ? d.F - 1 : f - 1, fp = f == d.F - 1 ? 0 : f + 1;
    UpdRow r;
    r.xf = u.xk + (long)f * d.N;
    r.xm = u.xk + (long)fm * d.N;
    r.xp = u.xk + (long)fp * d.N;
    r.xref_in = d.xref + (long)f * d.N;
    r.xref_out = d.xref + (long)f * d.N;
    r.Wk = d.Wk + (long)f * d.N;
    r.xo = u.xo + (long)f * d.N;
    r.g = d.mcg ? d.mcg + (long)f * d.N : nullptr;
    r.mu = u.mu; r.imu = 1.0f / u.mu; r.lam = u.lam; r.L1 = u.L1; r.apk = u.apk; r.tau = u.tau;
    r.sstart = u.sstart;
    return r;
}

// The arithmetic of the NESTA update of one pixel (NestaUp.cpp:696-875) from its loaded operands:
// df = lambda (u_f - u_{f+1}) + A^H r (the temporal-TV gradient D^T u with the periodic wrap, u = Ux / max(mu, |Ux|);
// with motion compensation lambda MC^T(D^T u), given in a), y_k = x_k - df / L_mu, W_k += apk df,
// z_k = x_ref - W_k / L_mu, x_{k+1} = tau z_k + (1 - tau) y_k.
// a, b: x_{f-1}, x_{f+1} (MC: a = MC^T(D^T u), b unused); xr, wo: x_ref and W_k (x_k and 0 at a stage start).
// Returns x_{k+1}; W_{k+1} in wn.
template <bool MC>
__device__ __forceinline__ float2 nesta_update_math(const UpdRow& r, float2 xf, float2 a, float2 b, float2 xr, float2 wo,
                                                    float2 ares, float2& wn) {
    float2 df;
    if constexpr (MC) {
        df = make_float2(r.lam * a.x + ares.x, r.lam * a.y + ares.y);
    } else {
        const float2 uxf = c_sub(xf, a), uxp = c_sub(b, xf);
        // u = Ux / max(mu, |Ux|) as Ux * min(1/mu, rsqrt(|Ux|^2)) (rsqrt(0) = inf)
        const float inf_ = fminf(r.imu, __builtin_amdgcn_rsqf(c_abs2(uxf)));
        const float inp_ = fminf(r.imu, __builtin_amdgcn_rsqf(c_abs2(uxp)));
        const float2 uf = c_scale(uxf, inf_), up = c_scale(uxp, inp_);
        df = make_float2(r.lam * (uf.x - up.x) + ares.x, r.lam * (uf.y - up.y) + ares.y);
    }
    const float2 yk = make_float2(xf.x - r.L1 * df.x, xf.y - r.L1 * df.y);
    wn = make_float2(wo.x + r.apk * df.x, wo.y + r.apk * df.y);
    const float2 z = make_float2(xr.x - r.L1 * wn.x, xr.y - r.L1 * wn.y);
    return make_float2(r.tau * z.x + (1.0f - r.tau) * yk.x, r.tau * z.y + (1.0f - r.tau) * yk.y);
}

// NESTA update of pixel ub + p of the frame resolved in r0 (ub wave-uniform, p per lane).  The first iteration of a
// stage sets x_ref = x_k and W_k = 0.  Writes W_k, x_ref (stage start) and x_{k+1}; returns x_{k+1}.
// MC (compile time): the table uses motion compensation, so the gradient comes from r.g.
template <bool MC = false>
__device__ __forceinline__ float2 nesta_update_at(const UpdRow& r0, long ub, unsigned p, float2 ares) {
    UpdRow r = r0;
    r.xf += ub; r.xm += ub; r.xp += ub; r.xref_in += ub; r.xref_out += ub; r.Wk += ub; r.xo += ub;
    const float2 xf = ldg2(r.xf, p);
    float2 a, b = make_float2(0.f, 0.f);
    if constexpr (MC) {
        a = ldg2(r.g + ub, p);
    } else {
        a = ldg2(r.xm, p);
        b = ldg2(r.xp, p);
    }
    float2 xr, wo;
    if (r.sstart) {
        xr = xf;
        stg2(r.xref_out, p, xf);
        wo = make_float2(0.f, 0.f);
    } else {
        xr = ldg2(r.xref_in, p);
        wo = ldg2(r.Wk, p);
    }
    float2 wn;
    const float2 xn = nesta_update_math<MC>(r, xf, a, b, xr, wo, ares, wn);
    stg2(r.Wk, p, wn);
    stg2(r.xo, p, xn);
    return xn;
}

template <bool MC = false>
__device__ __forceinline__ float2 nesta_update_at(const UpdRow& r, unsigned p, float2 ares) {
    return nesta_update_at<MC>(r, 0, p, ares);
}
