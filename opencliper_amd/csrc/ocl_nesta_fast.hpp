// Register four-step bodies of the fused NESTA passes for H = 16 P_H and W = 16 P_W (see ocl_rfft.hpp).
// Included by ocl_nesta.hip only.
#pragma once
#include "ocl_nesta_common.hpp"
#include "ocl_rfft.hpp"

// LDS footprint (float2 elements) of the fast passes
#define FAST_CA_KS 272                          // per-k1 stride of the column exchange: 16 roles x 16 columns + pad
#define FAST_CA_KQ 4                            // staged k-space values per thread and coil (nmax <= 64)
// per-row stride of the row exchange (k1 stride 17 within a row): the least RS >= 17 P with RS = 16 (mod 32) float2,
// so the 4 rows of a wavefront alternate between complementary halves of the 64 LDS banks and every b64 access of the
// exchange takes the 2-cycle minimum
#define FAST_B_RS(P) (17 * (P) + (((16 - 17 * (P)) % 32) + 32) % 32)
// Per-role twiddle tables: [P][TK] by k1 role and [16][TA] by a role (one table when P = 16).  Pass B, whose 16 roles
// of a wavefront read 16 different rows, pads the rows (TK = 17, TA = P + 1) so they start in different LDS banks;
// pass CA, whose wavefronts read 4 rows, keeps them unpadded (TK = 16, TA = P) for 128-bit reads.
__host__ __device__ constexpr int fast_twk_len(int P, int TK) { return TK * P; }
__host__ __device__ constexpr int fast_twa_len(int P, int TA) { return P == 16 ? 0 : 16 * TA; }
// pass CA: column exchange, twiddle tables, row positions [P][16] (ints)
__host__ __device__ constexpr int fast_ca_lds(int P) {
    return P * FAST_CA_KS + fast_twk_len(P, 16) + fast_twa_len(P, P) + 8 * P;
}
// Gram pass G (ocl_nesta_gram.hpp): the column pass's exchange, twiddle and row-slot tables, GRAM_SPARE more for the
// per-row residual factors (H <= 320 floats), and two staging slabs of (nmax + 1) x 16 values of V
#define GRAM_SPARE 160
__host__ __device__ constexpr int gram_lds(int P, int nmax) {
    return fast_ca_lds(P) + GRAM_SPARE + 2 * 16 * (nmax + 1);
}
// pass B: row exchange, twiddle tables
__host__ __device__ constexpr int fast_b_lds(int P) {
    return 16 * FAST_B_RS(P) + fast_twk_len(P, 17) + fast_twa_len(P, P + 1);
}

// ---------------------------------------------------------------------------------------------------------------------
// pass B, rows of length W = 16 P: 16 rows per workgroup, 16 threads per row
// ---------------------------------------------------------------------------------------------------------------------
template <int P>
__device__ __forceinline__ double nesta_B_fast(const SliceDesc& d, int local, int mode, float2* lds) {
    constexpr int W = 16 * P;
    constexpr int RS = FAST_B_RS(P);
    const int a = threadIdx.x & 15, r = threadIdx.x >> 4;
    const int g = local % d.nrow_groups;
    const int fc = local / d.nrow_groups;
    const int C = d.C, H = d.H, nmax = d.nmax;
    const int f = fc / C, c = fc - f * C;
    const int nr = d.nrows[f];
    const int j = g * 16 + r;
    const bool valid = j < nr;
    // per-role twiddle tables: twk[k1][n2] = W_W^{k1 n2} (FB phase 1, row stride 17), twa[a][k1] = W_W^{a k1} (FA phase
    // 1, row stride P + 1; the same table when P = 16), so a thread reads its row at one base with immediate offsets
    float2* twk = lds + 16 * RS;
    float2* twa = P == 16 ? twk : twk + fast_twk_len(P, 17);
    for (int i = threadIdx.x; i < W; i += blockDim.x) {
        twk[(i >> 4) * 17 + (i & 15)] = d.pw.tw[(i >> 4) * (i & 15)];
        if (P != 16) {
            const int ra = i / P, k = i - ra * P;      // i = P a + k1
            twa[ra * (P + 1) + k] = d.pw.tw[ra * k];
        }
    }
    float2* X = lds + r * RS;
    // wave-uniform bases of this (frame, coil) and per-thread 32-bit row offsets
    float2* kb = d.kbuf + ((long)f * C + c) * nmax * W;
    const unsigned kro = (unsigned)((valid ? j : 0) * W);
    const int ky = valid ? d.rows[(long)f * nmax + j] : 0;
    const float2* yfc = d.y + ((long)f * C + c) * H * W;
    const unsigned yro = (unsigned)(ky * W);
    double acc = 0.0;
    // y and mask at the residual positions kx = k1 + P k2 of this thread's roles, fetched ahead of the FFT (the 16
    // mask values of a role are the bits of one precomputed word)
    constexpr int NQ = P > 16 ? 2 : 1;
    float2 yq[NQ][16];
    unsigned mw[NQ];
    if (mode == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int k1 = a + 16 * q;
            const bool ok = valid && k1 < P;
            mw[q] = ok ? (unsigned)d.mbits[((long)f * nmax + j) * P + k1] : 0u;
#pragma unroll
            for (int k2 = 0; k2 < 16; ++k2) yq[q][k2] = ok ? ldg2(yfc + P * k2, yro + k1) : make_float2(0.f, 0.f);
        }
    }
    float2 v[P];
    // ---- FA phase 1 on U (mode 0) or conj(y) (mode 1) ----
    // (the branches stay outside the loops: a branch per value puts each load in a block of its own, waited on there)
    if (valid && mode == 0) {
#pragma unroll
        for (int b = 0; b < P; ++b) v[b] = ldg2(kb + 16 * b, kro + a);
    } else if (valid) {
#pragma unroll
        for (int b = 0; b < P; ++b) v[b] = c_conj(ldg2(yfc + 16 * b, yro + a));
    } else {
#pragma unroll
        for (int b = 0; b < P; ++b) v[b] = make_float2(0.f, 0.f);
    }
    __syncthreads();   // twiddles in LDS
    rf::rdft<P>(v);
    {
        const float2* tr = twa + (P + 1) * a;          // tr[0] = 1
        X[a] = v[0];
#pragma unroll
        for (int k1 = 1; k1 < P; ++k1) X[k1 * 17 + a] = c_mul(v[k1], tr[k1]);
    }
    __syncthreads();
    // ---- FA phase 2: role k1 in {a, a+16} ----
    float2 u[NQ][16];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int k1 = a + 16 * q;
        if (k1 < P) {
#pragma unroll
            for (int i = 0; i < 16; ++i) u[q][i] = X[k1 * 17 + i];
            rf::rdft<16>(u[q]);
        }
    }
    if (mode == 1) {
        // V = IFFT_W(y row) = conj(FFT(conj(y)))  at kx = k1 + P k2
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int k1 = a + 16 * q;
            if (k1 < P && valid)
#pragma unroll
                for (int k2 = 0; k2 < 16; ++k2) stg2(kb + P * k2, kro + k1, c_conj(u[q][k2]));
        }
        return 0.0;
    }
    // ---- residual r = M K / sqrt(N) - y at kx = k1 + P k2, then FB phase 1 on conj(r) ----
    const float s = d.inv_sqrtN;
    __syncthreads();   // X is rewritten below
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int k1 = a + 16 * q;
        if (k1 < P) {
#pragma unroll
            for (int k2 = 0; k2 < 16; ++k2) {
                float2 rv = make_float2(0.f, 0.f);
                if (valid) {
                    const float2 yv = yq[q][k2];
                    const float2 K = u[q][k2];
                    rv = ((mw[q] >> k2) & 1u) ? make_float2(K.x * s - yv.x, K.y * s - yv.y) : make_float2(-yv.x, -yv.y);
                    acc += (double)rv.x * rv.x + (double)rv.y * rv.y;
                }
                u[q][k2] = c_conj(rv);
            }
            rf::rdft<16>(u[q]);
            const float2* tr = twk + 17 * k1;          // tr[0] = 1
            X[k1 * 17] = u[q][0];
#pragma unroll
            for (int n2 = 1; n2 < 16; ++n2) X[k1 * 17 + n2] = c_mul(u[q][n2], tr[n2]);
        }
    }
    __syncthreads();
    // ---- FB phase 2: role n2 = a, out[n2 + 16 n1] ----
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) v[k1] = X[k1 * 17 + a];
    rf::rdft<P>(v);
    if (valid)
#pragma unroll
        for (int n1 = 0; n1 < P; ++n1) stg2(kb + 16 * n1, kro + a, c_conj(v[n1]));
    return acc;
}

// 16 consecutive ints of LDS (16-byte aligned) in four 128-bit reads
__device__ __forceinline__ void load_rows16(const int* src, int* dst) {
    const int4* s4 = (const int4*)src;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int4 v = s4[i];
        dst[4 * i] = v.x; dst[4 * i + 1] = v.y; dst[4 * i + 2] = v.z; dst[4 * i + 3] = v.w;
    }
}

// NESTA update of the P pixels h = t + 16 n1 of this thread's column (lane offset p; wv: the column exists).  The
// loads of four rows are issued together, unconditionally (p is clamped to a valid column), and only the stores are
// predicated, so each group of four waits on memory once.  SST: the first iteration of a stage.
template <int P, bool MC, bool SST>
__device__ __forceinline__ void ca_update(const UpdRow& r, int W, unsigned p, bool wv, const float2* acc, float sc,
                                          float2* xn) {
    constexpr int CH = 4;
#pragma unroll
    for (int n0 = 0; n0 < P; n0 += CH) {
        float2 xf[CH], a[CH], b[CH], xr[CH], wo[CH];
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            if (n0 + j >= P) break;
            const long ub = (long)16 * (n0 + j) * W;
            xf[j] = ldg2(r.xf + ub, p);
            if constexpr (MC) {
                a[j] = ldg2(r.g + ub, p);
                b[j] = make_float2(0.f, 0.f);
            } else {
                a[j] = ldg2(r.xm + ub, p);
                b[j] = ldg2(r.xp + ub, p);
            }
            if constexpr (!SST) {
                xr[j] = ldg2(r.xref_in + ub, p);
                wo[j] = ldg2(r.Wk + ub, p);
            }
        }
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            const int n1 = n0 + j;
            if (n1 >= P) break;
            const long ub = (long)16 * n1 * W;
            if constexpr (SST) {
                xr[j] = xf[j];
                wo[j] = make_float2(0.f, 0.f);
            }
            float2 wn;
            const float2 x1 = nesta_update_math<MC>(r, xf[j], a[j], b[j], xr[j], wo[j], c_scale(acc[n1], sc), wn);
            xn[n1] = wv ? x1 : make_float2(0.f, 0.f);
            if (wv) {
                if constexpr (SST) stg2(r.xref_out + ub, p, xf[j]);
                stg2(r.Wk + ub, p, wn);
                stg2(r.xo + ub, p, x1);
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------------------------------
// pass CA, columns of length H = 16 P: 16 columns per workgroup, 16 roles per column
// ---------------------------------------------------------------------------------------------------------------------
// STAGE = true adds two prefetches that keep global-memory latency off each coil's critical path: the sampled
// k-space rows of the next coil are loaded cooperatively (coalesced, at most FAST_CA_KQ values per thread) while the
// current coil is transformed and parked in one of two LDS slabs, and the S values of the next coil's forward pass
// are issued before the current coil's exchange barrier.  It needs nmax <= 16 FAST_CA_KQ and 2 * 16 * (nmax + 1) more
// float2 of LDS.  STAGE = false gathers the sampled rows straight from global memory.
// MC: the table uses motion compensation (update gradient from mcg; mode 2 = setup from a given x_ref)
template <int P, bool STAGE, bool MC = false>
__device__ __forceinline__ void nesta_CA_fast(const SliceDesc& d, int local, int mode, float2* lds) {
    constexpr int H = 16 * P;
    const int s = threadIdx.x & 15, t = threadIdx.x >> 4;
    float2* X = lds + s;
    const int tile = local % d.ncol_tiles;
    const int f = local / d.ncol_tiles;
    const int F = d.F, C = d.C, W = d.W, nmax = d.nmax;
    const long N = d.N;
    const int w = tile * 16 + s;
    const bool wv = w < W;
    const int wc = wv ? w : 0;
    const unsigned lane = (unsigned)(t * W + w), lanec = (unsigned)(t * W + wc);   // offsets of row t, column w
    // Per-role tables, so a thread reads its 16 (or P) values at one base with immediate offsets, all at once:
    //   twk[k1][n2] = W_H^{k1 n2} (FB phase 1), twa[a][k1] = W_H^{a k1} (FA phase 1; the same table when P = 16),
    //   rpk[k1][k2] = k-space slot of row k1 + P k2 in this frame, or nmax when the row is not sampled.
    float2* twk = lds + P * FAST_CA_KS;
    float2* twa = P == 16 ? twk : twk + fast_twk_len(P, 16);
    int* rpk = (int*)(twk + fast_twk_len(P, 16) + fast_twa_len(P, P));
    // STAGE: slabs [2][nmax + 1][16] of conjugated k-space rows; slot nmax stays zero, so an unsampled row reads zero
    float2* kst = (float2*)(rpk + H);
    const int kss = (nmax + 1) * 16;                   // slab stride
    const int nst = d.nrows[f] * 16;                   // staged values per coil
    for (int i = threadIdx.x; i < H; i += blockDim.x) {
        const int k1 = i >> 4, n2 = i & 15;            // i = 16 k1 + n2 (k1 < P)
        twk[i] = d.ph.tw[k1 * n2];
        if (P != 16) {
            const int a = i / P, k = i - a * P;        // i = P a + k
            twa[i] = d.ph.tw[a * k];
        }
        const int j = k1 + P * n2;                     // the row of role k1, index n2
        const int slot = d.rowpos[(long)f * H + j];
        rpk[i] = slot < 0 ? nmax : slot;
    }
    if constexpr (STAGE) {
        if (threadIdx.x < 32) kst[(threadIdx.x >> 4) * kss + nmax * 16 + (threadIdx.x & 15)] = make_float2(0.f, 0.f);
    }
    float2 pre[FAST_CA_KQ];
    auto stage_load = [&](int c) {
        const float2* kb = d.kbuf + ((long)f * C + c) * nmax * W;
#pragma unroll
        for (int q = 0; q < FAST_CA_KQ; ++q) {
            const int e = threadIdx.x + OCL_NT * q;
            pre[q] = (e < nst && wv) ? ldg2(kb + (long)16 * q * W, lane) : make_float2(0.f, 0.f);
        }
    };
    auto stage_store = [&](int buf) {
#pragma unroll
        for (int q = 0; q < FAST_CA_KQ; ++q) {
            const int e = threadIdx.x + OCL_NT * q;
            if (e < nst) kst[buf * kss + e] = c_conj(pre[q]);
        }
    };
    if constexpr (STAGE) {
        if (!MC || mode != 2) {
            stage_load(0);
            stage_store(0);
        }
    }
    __syncthreads();

    float2 acc[P];
#pragma unroll
    for (int i = 0; i < P; ++i) acc[i] = make_float2(0.f, 0.f);

    // ---- A^H r: per coil, inverse column DFT (FB on conj(V), input pruned to the sampled rows) ----
    constexpr int NQ = P > 16 ? 2 : 1;
    float2 sv[P];
    const int Cinv = (MC && mode == 2) ? 0 : C;            // no A^H r when x_ref is given
    for (int c = 0; c < Cinv; ++c) {
        const float2* kb = d.kbuf + ((long)f * C + c) * nmax * W;
        const float2* ks = kst + (c & 1) * kss + s;
        const float2* Sc = d.S + (long)c * N;
        if constexpr (STAGE) {
            if (c + 1 < C) stage_load(c + 1);
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int k1 = t + 16 * q;
            if (k1 < P) {
                int jv[16];
                load_rows16(rpk + 16 * k1, jv);
                float2 u[16];
#pragma unroll
                for (int k2 = 0; k2 < 16; ++k2) {
                    const int jj = jv[k2];
                    if constexpr (STAGE) {
                        u[k2] = ks[jj * 16];                             // conjugated when staged; zero row at nmax
                    } else {
                        // every load issued (an unsampled row reads row 0, a missing column column 0) and masked by a
                        // multiply, so none waits in a branch of its own
                        const bool ok = jj < nmax;
                        const float m = (ok && wv) ? 1.f : 0.f;
                        const float2 kv = ldg2(kb, (unsigned)(ok ? jj * W + wc : wc));
                        u[k2] = make_float2(kv.x * m, -(kv.y * m));
                    }
                }
                rf::rdft<16>(u);
                const float2* tr = twk + 16 * k1;                // tr[0] = 1
                X[k1 * FAST_CA_KS] = u[0];
#pragma unroll
                for (int n2 = 1; n2 < 16; ++n2) X[k1 * FAST_CA_KS + n2 * 16] = c_mul(u[n2], tr[n2]);
            }
        }
        // S of this coil: issued here, after the first phase, so it is not live across that phase's DFTs
#pragma unroll
        for (int n1 = 0; n1 < P; ++n1) sv[n1] = ldg2(Sc + (long)16 * n1 * W, lanec);
        __syncthreads();
        float2 v[P];
#pragma unroll
        for (int k1 = 0; k1 < P; ++k1) v[k1] = X[k1 * FAST_CA_KS + t * 16];
        if constexpr (STAGE) {
            // the other slab was last read in the previous coil's first phase, which the barrier above has closed
            if (c + 1 < C) stage_store((c + 1) & 1);
        }
        __syncthreads();
        rf::rdft<P>(v);
#pragma unroll
        for (int n1 = 0; n1 < P; ++n1) {
            const float2 z = c_conj(v[n1]);                       // IFFT_H(V) at h = t + 16 n1
            acc[n1] = c_add(acc[n1], c_mulconj(z, sv[n1]));
        }
    }

    // ---- NESTA update (NestaUp.cpp:669-875 fused) at h = t + 16 n1, w ----
    float2 xn[P];
    const float sc = d.inv_sqrtN;
    if (mode == 1) {
#pragma unroll
        for (int n1 = 0; n1 < P; ++n1) {
            xn[n1] = c_scale(acc[n1], sc);
            if (wv) stg2(d.x0 + (long)f * N + (long)16 * n1 * W, lane, xn[n1]);
        }
    } else if (MC && mode == 2) {                          // x_ref given in x0: forward columns only
#pragma unroll
        for (int n1 = 0; n1 < P; ++n1)
            xn[n1] = wv ? ldg2(d.x0 + (long)f * N + (long)16 * n1 * W, lane) : make_float2(0.f, 0.f);
    } else {
        const UpdRow u = upd_row(d, f);
        if (u.sstart) ca_update<P, MC, true>(u, W, lanec, wv, acc, sc, xn);
        else ca_update<P, MC, false>(u, W, lanec, wv, acc, sc, xn);
    }

    // ---- forward column DFTs of S_c x_{k+1} (FA), output pruned to the sampled rows ----
    float2 sn[P];                                          // STAGE: S of the coil about to be transformed
    if constexpr (STAGE) {
#pragma unroll
        for (int n1 = 0; n1 < P; ++n1) sn[n1] = ldg2(d.S + (long)16 * n1 * W, lanec);
    }
    for (int c = 0; c < C; ++c) {
        const float2* Sc = d.S + (long)c * N;
        float2 v[P];
#pragma unroll
        for (int n1 = 0; n1 < P; ++n1) {
            const float2 sc1 = STAGE ? sn[n1] : ldg2(Sc + (long)16 * n1 * W, lanec);
            v[n1] = c_mul(xn[n1], sc1);      // a missing column (!wv) stays in its own lanes and is never stored
        }
        rf::rdft<P>(v);
        const float2* tr = twa + P * t;                          // tr[0] = 1
        X[t * 16] = v[0];
#pragma unroll
        for (int k1 = 1; k1 < P; ++k1) X[k1 * FAST_CA_KS + t * 16] = c_mul(v[k1], tr[k1]);
        if constexpr (STAGE) {
            if (c + 1 < C)
#pragma unroll
                for (int n1 = 0; n1 < P; ++n1) sn[n1] = ldg2(Sc + N + (long)16 * n1 * W, lanec);
        }
        __syncthreads();
        float2* kb = d.kbuf + ((long)f * C + c) * nmax * W;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int k1 = t + 16 * q;
            if (k1 < P) {
                int jv[16];
                load_rows16(rpk + 16 * k1, jv);
                float2 u[16];
#pragma unroll
                for (int i = 0; i < 16; ++i) u[i] = X[k1 * FAST_CA_KS + i * 16];
                rf::rdft<16>(u);
#pragma unroll
                for (int k2 = 0; k2 < 16; ++k2) {
                    const int jj = jv[k2];
                    if (jj < nmax && wv) stg2(kb, (unsigned)(__umul24(jj, W) + w), u[k2]);
                }
            }
        }
        __syncthreads();
    }
}
