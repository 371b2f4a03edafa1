// Gram form of the fused NESTA iteration for ky-line masks (a mask constant along every k-space row), H = 16 P.
// Included by ocl_nesta.hip only.
//
// With a line mask the row transform commutes with the mask: M F_2 = F_x M_y F_y, where F_x transforms along the rows
// (length W) and M_y F_y acts on columns only.  Write V_c = W F_x^{-1} y_c (the row-inverse-transformed k-space,
// computed once at setup by pass B, mode 1).  Per (frame, coil) and column, with K = F_y(S_c x) the column transform:
//   r_c = M F_2(S_c x) / sqrt(N) - y_c = F_x (rho_c) / W,      rho_c = m K W / sqrt(N) - V_c   (m: the row's mask)
//   ||r_c||^2 = (1/W) ||rho_c||^2,      A^H(sqrt(N) r) = sum_c conj(S_c) conjDFT_y(rho_c) / sqrt(N)
// so one pass over column tiles evaluates the residual, its norm and A^H r without any row transform and without
// writing k-space: per coil, forward column DFT of S_c x_k (four-step FA), the residual at the carried rows in
// registers, inverse column DFT (four-step FB) straight from those registers, conj(S_c) accumulate.  The NESTA update
// (NestaUp.cpp:696-875) follows in the same pass, together with the partial sums of ||u||^2 and Re<u, Ux> of x_k.
//
// The update needs the objective of x_k for the stop test (NestaUp.cpp:799-831), which only the scalar step after
// the pass knows.  The pass therefore applies the coefficients of an iteration that does not stop; x and W are both
// ping-ponged, so when the stop test ends the stage at this iteration (the reference then takes apk and tau with
// k = maxIter) the next pass recomputes the same iteration from the same x_k and W_k with those coefficients.
#pragma once
#include "ocl_nesta_fast.hpp"

#ifndef GRAM_UPD_ROWS
#define GRAM_UPD_ROWS 8      // rows of the update whose operands are loaded together
#endif

// NESTA update of the P pixels h = t + 16 n1 of this thread's column from x_k (in registers), its neighbour frames,
// x_ref and W_k; accumulates |u_f|^2 and Re(conj(u_f) Ux_f) of x_k for the objective.  SST: first iteration of a
// stage (x_ref := x_k, W_k := 0).
template <int P, bool SST>
__device__ __forceinline__ void gram_update(const SliceDesc& d, const NState* st, int f, unsigned p, bool wv,
                                            const float2* xk, const float2* acc, double* uterm) {
    // loads of CH rows are issued together (the coil loop's registers are free here): each group waits on memory once
    constexpr int CH = GRAM_UPD_ROWS < P ? GRAM_UPD_ROWS : P;
    const long N = d.N, W = d.W;
    const int fm = f == 0 ? d.F - 1 : f - 1, fp = f == d.F - 1 ? 0 : f + 1;
    // buffer resources over the slice's image arrays: one 32-bit lane offset, scalar row offsets (no 64-bit
    // per-lane address arithmetic on the vector unit)
    // (the ping-pong flags are read as wavefront-uniform values, so the resources stay in scalar registers)
    const long xbytes = 8 * (long)d.F * N;
    const int cur = __builtin_amdgcn_readfirstlane(st->it_cur), wcur = __builtin_amdgcn_readfirstlane(st->it_wcur);
    const BufRes rxb = buf_res(cur ? d.x1 : d.x0, xbytes), rxo = buf_res(cur ? d.x0 : d.x1, xbytes);
    const BufRes rwi = buf_res(wcur ? d.Wk2 : d.Wk, xbytes), rwo = buf_res(wcur ? d.Wk : d.Wk2, xbytes);
    const BufRes rxr = buf_res(d.xref, xbytes);
    const unsigned vo = p * 8u;
    const unsigned so_f = (unsigned)(8 * (long)f * N), so_m = (unsigned)(8 * (long)fm * N),
                   so_p = (unsigned)(8 * (long)fp * N);
    const float mu = (float)st->it_mu, imu = 1.0f / mu, lam = (float)st->lambda, L1 = (float)st->it_L1;
    const float apk = (float)st->it_apk, tau = (float)st->it_tau, sc = d.inv_sqrtN;
    float u2 = 0.f, fx = 0.f;
#pragma unroll
    for (int n0 = 0; n0 < P; n0 += CH) {
        float2 a[CH], b[CH], xr[CH], wo[CH];
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            if (n0 + j >= P) break;
            const unsigned ub = (unsigned)(8 * 16 * (n0 + j) * W);
            a[j] = bld2(rxb, vo, so_m + ub);
            b[j] = bld2(rxb, vo, so_p + ub);
            if constexpr (!SST) {
                xr[j] = bld2(rxr, vo, so_f + ub);
                wo[j] = bld2(rwi, vo, so_f + ub);
            }
        }
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            const int n1 = n0 + j;
            if (n1 >= P) break;
            const unsigned ub = (unsigned)(8 * 16 * n1 * W);
            const float2 xf = xk[n1];
            if constexpr (SST) {
                xr[j] = xf;
                wo[j] = make_float2(0.f, 0.f);
            }
            // u = Ux / max(mu, |Ux|) as Ux * min(1/mu, rsqrt(|Ux|^2)), for frames f and f + 1
            const float2 uxf = c_sub(xf, a[j]), uxp = c_sub(b[j], xf);
            const float2 uf = c_scale(uxf, fminf(imu, __builtin_amdgcn_rsqf(c_abs2(uxf))));
            const float2 up = c_scale(uxp, fminf(imu, __builtin_amdgcn_rsqf(c_abs2(uxp))));
            const float2 ar = c_scale(acc[n1], sc);
            const float2 df = make_float2(lam * (uf.x - up.x) + ar.x, lam * (uf.y - up.y) + ar.y);
            const float2 yk = make_float2(xf.x - L1 * df.x, xf.y - L1 * df.y);
            const float2 wn = make_float2(wo[j].x + apk * df.x, wo[j].y + apk * df.y);
            const float2 z = make_float2(xr[j].x - L1 * wn.x, xr[j].y - L1 * wn.y);
            const float2 x1 = make_float2(tau * z.x + (1.0f - tau) * yk.x, tau * z.y + (1.0f - tau) * yk.y);
            if (wv) {
                u2 += c_abs2(uf);
                fx += uf.x * uxf.x + uf.y * uxf.y;
                if constexpr (SST) bst2(rxr, vo, so_f + ub, xf);
                bst2(rwo, vo, so_f + ub, wn);
                bst2(rxo, vo, so_f + ub, x1);
            }
        }
    }
    uterm[0] += (double)u2;
    uterm[1] += (double)fx;
}

// One workgroup = 16 columns of one frame, 16 roles per column (thread (s, t): column s of the tile, role t).
// mode 0: iteration (residual, A^H r, update, partial sums in partG); mode 1: setup, x0 = x_ref = A^H(sqrt(N) y) from
// the staged V (no forward transform, no update).
template <int P>
__device__ __forceinline__ void nesta_G_fast(const SliceDesc& d, int local, int mode, float2* lds, double* red) {
    constexpr int H = 16 * P;
    constexpr int NQ = P > 16 ? 2 : 1;
    const int s = threadIdx.x & 15, t = threadIdx.x >> 4;
    float2* X = lds + s;
    const int tile = local % d.ncol_tiles;
    const int f = local / d.ncol_tiles;
    const int C = d.C, W = d.W, nmax = d.nmax;
    const long N = d.N;
    const int w = tile * 16 + s;
    const bool wv = w < W;
    const int wc = wv ? w : 0;
    const unsigned lane = (unsigned)(t * W + w), lanec = (unsigned)(t * W + wc);
    const NState* st = d.st;
    // per-role tables (as in the column pass): twk[k1][n2] = W_H^{k1 n2}, twa[a][k1] = W_H^{a k1}, rpk[k1][k2] = slot of
    // row k1 + P k2 (nmax when the row carries nothing), mkf[k1][k2] = W / sqrt(N) when that row is a sampled (mask)
    // row, else 0 (the factor of K in the residual)
    float2* twk = lds + P * FAST_CA_KS;
    float2* twa = P == 16 ? twk : twk + fast_twk_len(P, 16);
    int* rpk = (int*)(twk + fast_twk_len(P, 16) + fast_twa_len(P, P));
    float* mkf = (float*)(rpk + H);
    float2* kst = (float2*)(rpk + H) + GRAM_SPARE;
    const float sW = (float)W * d.inv_sqrtN;
    const int kss = (nmax + 1) * 16;                   // slab stride; slot nmax of each slab stays zero
    const int nst = d.nrows[f] * 16;                   // staged values per coil
    for (int i = threadIdx.x; i < H; i += blockDim.x) {
        const int k1 = i >> 4, n2 = i & 15;
        twk[i] = d.ph.tw[k1 * n2];
        if (P != 16) {
            const int a = i / P, k = i - a * P;
            twa[i] = d.ph.tw[a * k];
        }
        const int slot = d.rowpos[(long)f * H + k1 + P * n2];
        rpk[i] = slot < 0 ? nmax : slot;
        mkf[i] = (slot >= 0 && d.rowmask[(long)f * nmax + slot]) ? sW : 0.f;
    }
    if (threadIdx.x < 32) kst[(threadIdx.x >> 4) * kss + nmax * 16 + (threadIdx.x & 15)] = make_float2(0.f, 0.f);
    // staging of conj(V) of one coil: coalesced 16-column rows, at most FAST_CA_KQ values per thread
    // buffer resources (one 32-bit lane offset, scalar offsets): the staged k-space, S and x_k
    const BufRes rV = buf_res(d.kbuf, 8 * (long)d.F * C * nmax * W);
    const BufRes rS = buf_res(d.S, 8 * (long)C * N);
    const unsigned vl = lane * 8u, vlc = lanec * 8u, rowB = (unsigned)W * 8u;
    float2 pre[FAST_CA_KQ];
    auto stage_load = [&](int c) {
        const unsigned kb = (unsigned)(8 * ((long)f * C + c) * nmax * W);
#pragma unroll
        for (int q = 0; q < FAST_CA_KQ; ++q) {
            const int e = threadIdx.x + OCL_NT * q;
            pre[q] = (e < nst && wv) ? bld2(rV, vl, kb + 16u * q * rowB) : make_float2(0.f, 0.f);
        }
    };
    auto stage_store = [&](int buf) {
#pragma unroll
        for (int q = 0; q < FAST_CA_KQ; ++q) {
            const int e = threadIdx.x + OCL_NT * q;
            if (e < nst) kst[buf * kss + e] = c_conj(pre[q]);
        }
    };
    stage_load(0);
    stage_store(0);

    float2 xk[P];
    if (mode == 0) {
        const BufRes rX = buf_res(__builtin_amdgcn_readfirstlane(st->it_cur) ? d.x1 : d.x0, 8 * (long)d.F * N);
        const unsigned xf = (unsigned)(8 * (long)f * N);
#pragma unroll
        for (int n1 = 0; n1 < P; ++n1) xk[n1] = bld2(rX, vlc, xf + 16u * n1 * rowB);
    }
    float2 acc[P];
#pragma unroll
    for (int i = 0; i < P; ++i) acc[i] = make_float2(0.f, 0.f);
    double r2 = 0.0;
    __syncthreads();

    for (int c = 0; c < C; ++c) {
        const unsigned sc = (unsigned)(8 * (long)c * N);       // byte offset of S_c
        const float2* ks = kst + (c & 1) * kss + s;
        if (c + 1 < C) stage_load(c + 1);
        if (mode == 0) {
            // ---- forward column DFT of S_c x_k (FA phase 1 in registers, twiddle, exchange) ----
            // all P loads of S_c are issued before the first product: left to itself the scheduler may interleave
            // load, wait and multiply one value at a time, which serializes P memory latencies
            float2 v[P];
#pragma unroll
            for (int n1 = 0; n1 < P; ++n1) v[n1] = bld2(rS, vlc, sc + 16u * n1 * rowB);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int n1 = 0; n1 < P; ++n1) v[n1] = c_mul(xk[n1], v[n1]);
            rf::rdft<P>(v);
            const float2* tr = twa + P * t;                      // tr[0] = 1
            float2 tw[P];                                        // twiddles read together, as S_c above
#pragma unroll
            for (int k1 = 1; k1 < P; ++k1) tw[k1] = tr[k1];
            __builtin_amdgcn_sched_barrier(0);
            X[t * 16] = v[0];
#pragma unroll
            for (int k1 = 1; k1 < P; ++k1) X[k1 * FAST_CA_KS + t * 16] = c_mul(v[k1], tw[k1]);
            __syncthreads();
        }
        // ---- per role k1: (forward FA phase 2,) residual at the carried rows, inverse FB phase 1 ----
        float rq = 0.f;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int k1 = t + 16 * q;
            if (k1 < P) {
                int jv[16];
                load_rows16(rpk + 16 * k1, jv);
                float2 u[16];
                if (mode == 0) {
#pragma unroll
                    for (int i = 0; i < 16; ++i) u[i] = X[k1 * FAST_CA_KS + i * 16];
                    rf::rdft<16>(u);                              // K at rows k1 + P k2
                    float mk[16];
                    load_rows16((const int*)(mkf + 16 * k1), (int*)mk);
#pragma unroll
                    for (int k2 = 0; k2 < 16; ++k2) {
                        // conj(rho) = m conj(K) W / sqrt(N) - conj(V); zero at rows that carry nothing
                        const float2 cv = ks[jv[k2] * 16];
                        const float2 rc = make_float2(fmaf(mk[k2], u[k2].x, -cv.x), fmaf(-mk[k2], u[k2].y, -cv.y));
                        rq = fmaf(rc.x, rc.x, fmaf(rc.y, rc.y, rq));
                        u[k2] = rc;
                    }
                } else {
#pragma unroll
                    for (int k2 = 0; k2 < 16; ++k2) u[k2] = ks[jv[k2] * 16];
                }
                rf::rdft<16>(u);
                const float2* tr = twk + 16 * k1;                // tr[0] = 1
                X[k1 * FAST_CA_KS] = u[0];
#pragma unroll
                for (int n2 = 1; n2 < 16; ++n2) X[k1 * FAST_CA_KS + n2 * 16] = c_mul(u[n2], tr[n2]);
            }
        }
        if (wv) r2 += (double)rq;         // a missing column's lanes transform column 0's values: not counted
        // S of this coil for the accumulate, issued before the exchange barrier
        float2 sv[P];
#pragma unroll
        for (int n1 = 0; n1 < P; ++n1) sv[n1] = bld2(rS, vlc, sc + 16u * n1 * rowB);
        __syncthreads();
        // the other slab was last read in the previous coil's role phase, which the barrier above has closed
        if (c + 1 < C) stage_store((c + 1) & 1);
        // ---- inverse FB phase 2 (role n2 = t), conj(S_c) accumulate ----
        float2 v[P];
#pragma unroll
        for (int k1 = 0; k1 < P; ++k1) v[k1] = X[k1 * FAST_CA_KS + t * 16];
        rf::rdft<P>(v);
        // acc += conj(v) conj(S) as four fused multiply-adds
#pragma unroll
        for (int n1 = 0; n1 < P; ++n1) {
            acc[n1].x = fmaf(v[n1].x, sv[n1].x, fmaf(-v[n1].y, sv[n1].y, acc[n1].x));
            acc[n1].y = fmaf(-v[n1].y, sv[n1].x, fmaf(-v[n1].x, sv[n1].y, acc[n1].y));
        }
        if (mode != 0) __syncthreads();       // setup: no forward exchange separates the next coil's role phase
    }
    if (mode == 1) {
#pragma unroll
        for (int n1 = 0; n1 < P; ++n1)
            if (wv) stg2(d.x0 + (long)f * N + (long)16 * n1 * W, lane, c_scale(acc[n1], d.inv_sqrtN));
        return;
    }
    double part[3] = {r2 / (double)W, 0.0, 0.0};
    if (st->it_sstart) gram_update<P, true>(d, st, f, lanec, wv, xk, acc, part + 1);
    else gram_update<P, false>(d, st, f, lanec, wv, xk, acc, part + 1);
    block_sum<3>(part, red);
    if (threadIdx.x == 0) {
        d.partG[3 * local] = part[0];
        d.partG[3 * local + 1] = part[1];
        d.partG[3 * local + 2] = part[2];
    }
}
