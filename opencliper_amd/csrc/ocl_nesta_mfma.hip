// Split-precision NESTA-TV passes for gfx950: the row and column DFTs of E = M F S and its adjoint run as dense
// complex GEMMs on the BF16 matrix cores (v_mfma_f32_16x16x32_bf16) with a 3-term split product, accumulated in FP32.
//
// Every FP32 operand value v is split into hi = bf16(v) and lo = bf16(v - hi); a product is taken as
// a_hi b_hi + a_hi b_lo + a_lo b_hi (the a_lo b_lo term, ~2^-16 relative, is dropped).  A complex 16x16x32 block then
// costs 12 MFMAs: Re += Ar Br - Ai Bi and Im += Ar Bi + Ai Br, three terms each.  The DFT matrices are the A operands
// (precomputed per slice in double and split, k_dft_split); the data are the B operands, split in registers.
//
// The passes have the structure of the FP32 passes (pass B: rows, pass CA: columns + fused NESTA update; the scalar
// step, the state machine and the temporal-TV reductions are shared), with one wavefront per workgroup:
//   B  (frame f, coil c, 16 sampled rows): K = X D_W (forward row DFT), r = M K / sqrt(N) - y, V = r conj(D_W)
//   CA (frame f, 16 columns): for each coil Z = D_inv(f) V (inverse column DFT from the sampled rows), A^H r =
//      sum_c conj(S_c) Z / sqrt(N), NESTA update, then for each coil U = D_fwd(f) (S_c x_{k+1}) onto the sampled rows.
// The output tile of the first GEMM of a pass leaves each lane holding 4 consecutive rows (m = 16 t + 4 g + i) of its
// column; the second GEMM of the pass sums over that same index, so its K order is permuted to
// k = 32 s + 8 g + e  <->  m = 16 (2 s + e / 4) + 4 g + e % 4, which makes its B operand exactly the lane's own
// registers (no LDS exchange).  The matching A matrices (drowI, dcolF) are stored in that permuted K order.
#include "ocl_nesta_common.hpp"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));

namespace {

struct Op8 {           // 8 complex values as bf16 hi/lo planes
    bf16x8 rh, rl, ih, il;
};

__device__ __forceinline__ void split8(const float2 (&v)[8], Op8& o) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const __bf16 hr = (__bf16)v[e].x, hi = (__bf16)v[e].y;
        o.rh[e] = hr;
        o.rl[e] = (__bf16)(v[e].x - (float)hr);
        o.ih[e] = hi;
        o.il[e] = (__bf16)(v[e].y - (float)hi);
    }
}

__device__ __forceinline__ bf16x8 bneg(bf16x8 v) {
    u32x4 u = __builtin_bit_cast(u32x4, v);
    u ^= 0x80008000u;
    return __builtin_bit_cast(bf16x8, u);
}

// A operand: 8 consecutive K entries of row m of a 4-plane matrix (plane stride ps), off = m * K + k0
__device__ __forceinline__ Op8 loadA(const uint16_t* __restrict__ base, long ps, long off) {
    Op8 a;
    a.rh = *(const bf16x8*)(base + off);
    a.rl = *(const bf16x8*)(base + ps + off);
    a.ih = *(const bf16x8*)(base + 2 * ps + off);
    a.il = *(const bf16x8*)(base + 3 * ps + off);
    return a;
}

#define MF(a, b, c) __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0)

// Keeps memory operations of the unrolled loops on their own side of this point (compiler-level: an empty asm with a
// memory clobber, plus a machine-scheduling barrier), so the loads in flight, and their registers, stay bounded.
__device__ __forceinline__ void pipeline_fence() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
}

// (dr, di) += A B for one 16 x 16 x 32 complex block, 3-term split product
__device__ __forceinline__ void cmac(f32x4& dr, f32x4& di, const Op8& a, const Op8& b) {
    const bf16x8 nih = bneg(b.ih), nil = bneg(b.il);
    dr = MF(a.rh, b.rh, dr);
    di = MF(a.rh, b.ih, di);
    dr = MF(a.rh, b.rl, dr);
    di = MF(a.rh, b.il, di);
    dr = MF(a.rl, b.rh, dr);
    di = MF(a.rl, b.ih, di);
    dr = MF(a.ih, nih, dr);
    di = MF(a.ih, b.rh, di);
    dr = MF(a.ih, nil, dr);
    di = MF(a.ih, b.rl, di);
    dr = MF(a.il, nih, dr);
    di = MF(a.il, b.rh, di);
}

// split8 with one conversion per pair of values: hi = RNE bf16 of both, the exact remainders v - hi (FP32), their RNE
// bf16 as lo.  The same values as split8, in 3 instructions per value instead of 4.
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned cvt_pair(float a, float b) {
    f32x2_t v = {a, b};
    return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2_t));
}
__device__ __forceinline__ void split8p(const float2 (&v)[8], Op8& o) {
    u32x4 rh, rl, ih, il;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const float2 a = v[2 * p], b = v[2 * p + 1];
        const unsigned hr = cvt_pair(a.x, b.x), hi = cvt_pair(a.y, b.y);
        rh[p] = hr;
        ih[p] = hi;
        rl[p] = cvt_pair(a.x - __builtin_bit_cast(float, hr << 16), b.x - __builtin_bit_cast(float, hr & 0xffff0000u));
        il[p] = cvt_pair(a.y - __builtin_bit_cast(float, hi << 16), b.y - __builtin_bit_cast(float, hi & 0xffff0000u));
    }
    o.rh = __builtin_bit_cast(bf16x8, rh);
    o.rl = __builtin_bit_cast(bf16x8, rl);
    o.ih = __builtin_bit_cast(bf16x8, ih);
    o.il = __builtin_bit_cast(bf16x8, il);
}

// (dr - dn, di) += A B for one 16 x 16 x 32 complex block, 3-term split product, with the Im A Im B terms summed
// positive into dn (subtracted once at the end) instead of negating each B operand: three accumulator chains
__device__ __forceinline__ void cmac3(f32x4& dr, f32x4& dn, f32x4& di, const Op8& a, const Op8& b) {
    dr = MF(a.rh, b.rh, dr);
    dn = MF(a.ih, b.ih, dn);
    di = MF(a.rh, b.ih, di);
    dr = MF(a.rh, b.rl, dr);
    dn = MF(a.ih, b.il, dn);
    di = MF(a.ih, b.rh, di);
    dr = MF(a.rl, b.rh, dr);
    dn = MF(a.il, b.ih, dn);
    di = MF(a.rh, b.il, di);
    di = MF(a.ih, b.rl, di);
    di = MF(a.rl, b.ih, di);
    di = MF(a.il, b.rh, di);
}

__device__ __forceinline__ int kperm(int k) {   // permuted K index -> row index of the first GEMM's output
    const int s = k >> 5, r = k & 31, g = r >> 3, e = r & 7;
    return 16 * (2 * s + (e >> 2)) + 4 * g + (e & 3);
}

__device__ __forceinline__ void stage_planes(uint16_t* dst, const uint16_t* src, int M, int K) {
    const int ld = mf_lds_ld(K), cpr = K / 8, total = 4 * M * cpr;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int pm = i / cpr, ck = i - pm * cpr;          // pm = plane * M + row
        *(u32x4*)(dst + pm * ld + ck * 8) = *(const u32x4*)(src + (long)pm * K + ck * 8);
    }
}

// A operand from a staged matrix with M rows and row stride ld: 8 consecutive K entries of row m starting at k0
__device__ __forceinline__ Op8 loadA_lds(const uint16_t* base, int M, int ld, int m, int k0) {
    const int ps = M * ld, off = m * ld + k0;
    Op8 a;
    a.rh = *(const bf16x8*)(base + off);
    a.rl = *(const bf16x8*)(base + ps + off);
    a.ih = *(const bf16x8*)(base + 2 * ps + off);
    a.il = *(const bf16x8*)(base + 3 * ps + off);
    return a;
}

// ---------------------------------------------------------------------------------------------------------------------
// pass B: MF_B_TASKS row tasks (16 sampled rows of one (frame, coil) each, one wavefront per task) per workgroup;
// W = 16 MT.  Both row DFTs run k-step-outer: the workgroup stages the 32-column k-slab of the (shared) DFT matrix in
// LDS and every wavefront multiplies its rows against it, so each A fragment leaves L2 once per workgroup.
// ---------------------------------------------------------------------------------------------------------------------
// stages columns [k0, k0 + 32) of a 4-plane [M][K] matrix into LDS rows of mf_lds_ld(32) elements
__device__ __forceinline__ void stage_slab(uint16_t* dst, const uint16_t* src, int M, int K, int k0) {
    const int ld = mf_lds_ld(32), total = 4 * M * 4;       // 4 chunks of 8 elements per row
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int pm = i >> 2, ck = i & 3;
        *(u32x4*)(dst + pm * ld + ck * 8) = *(const u32x4*)(src + (long)pm * K + k0 + ck * 8);
    }
}

template <int MT>
__device__ __forceinline__ double nesta_B_mfma(const SliceDesc& d, int local, int mode, uint16_t* lds) {
    constexpr int W = 16 * MT, KS = MT / 2;
    constexpr int LD = mf_lds_ld(32);
    const int l = threadIdx.x & 63, n = l & 15, g = l >> 4;
    const int G = d.nrow_groups;
    const int task = local * MF_B_TASKS + (threadIdx.x >> 6);
    const int C = d.C, H = d.H, nmax = d.nmax;
    const bool tv = task < d.F * C * G;                      // the last workgroup may hold idle wavefronts
    const int grp = tv ? task % G : 0, fc = tv ? task / G : 0;
    const int f = fc / C, c = fc - f * C;
    const int nr = d.nrows[f];
    const bool act = tv && grp * 16 < nr;
    const int j = grp * 16 + n;
    const bool valid = act && j < nr;
    float2* kb = d.kbuf + (((long)f * C + c) * nmax + (valid ? j : 0)) * W;
    const int ky = valid ? d.rows[(long)f * nmax + j] : 0;
    const float2* yrow = d.y + (((long)f * C + c) * H + ky) * W;
    const uint8_t* mrow = d.M + ((long)f * H + ky) * W;
    f32x4 ar[MT], ai[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t) { ar[t] = f32x4{0.f, 0.f, 0.f, 0.f}; ai[t] = ar[t]; }
    double racc = 0.0;
    if (mode == 0) {
        // forward row DFT: D[kx][row] = sum_w F_W[kx][w] X[row][w]
        for (int s = 0; s < KS; ++s) {
            __syncthreads();                                 // previous slab fully read
            stage_slab(lds, d.drowF, W, W, 32 * s);
            __syncthreads();
            if (!act) continue;
            float2 v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = valid ? kb[32 * s + 8 * g + e] : make_float2(0.f, 0.f);
            Op8 b;
            split8(v, b);
#pragma unroll
            for (int t = 0; t < MT; ++t) cmac(ar[t], ai[t], loadA_lds(lds, W, LD, 16 * t + n, 8 * g), b);
        }
        // residual r = M K / sqrt(N) - y at kx = 16 t + 4 g + i (NestaUp.cpp:716-732)
        const float sc = d.inv_sqrtN;
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int kx = 16 * t + 4 * g + i;
                float2 rv = make_float2(0.f, 0.f);
                if (valid) {
                    const float2 yv = yrow[kx];
                    rv = mrow[kx] ? make_float2(ar[t][i] * sc - yv.x, ai[t][i] * sc - yv.y) : make_float2(-yv.x, -yv.y);
                    racc += (double)rv.x * rv.x + (double)rv.y * rv.y;
                }
                ar[t][i] = rv.x;
                ai[t][i] = rv.y;
            }
    } else {
        // setup: the inverse row DFT of the y row itself
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float2 yv = valid ? yrow[16 * t + 4 * g + i] : make_float2(0.f, 0.f);
                ar[t][i] = yv.x;
                ai[t][i] = yv.y;
            }
    }
    // inverse row DFT (unnormalised): V[w][row] = sum_kx conj(F_W)[w][kx] r[kx][row], K in the permuted order; k-step
    // s consumes the residual of m-tiles 2 s and 2 s + 1
    f32x4 orr[MT], oi[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t) { orr[t] = f32x4{0.f, 0.f, 0.f, 0.f}; oi[t] = orr[t]; }
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        __syncthreads();
        stage_slab(lds, d.drowI, W, W, 32 * s);
        __syncthreads();
        if (act) {
            float2 v[8];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                v[e] = make_float2(ar[2 * s][e], ai[2 * s][e]);
                v[e + 4] = make_float2(ar[2 * s + 1][e], ai[2 * s + 1][e]);
            }
            Op8 rb;
            split8(v, rb);
#pragma unroll
            for (int t = 0; t < MT; ++t) cmac(orr[t], oi[t], loadA_lds(lds, W, LD, 16 * t + n, 8 * g), rb);
        }
    }
    if (valid) {
#pragma unroll
        for (int t = 0; t < MT; ++t) {
            float4* o = (float4*)(kb + 16 * t + 4 * g);
            o[0] = make_float4(orr[t][0], oi[t][0], orr[t][1], oi[t][1]);
            o[1] = make_float4(orr[t][2], oi[t][2], orr[t][3], oi[t][3]);
        }
    }
    return racc;
}

// ---------------------------------------------------------------------------------------------------------------------
// pass CA: MF_CA_TILES column tiles of 16 columns of one frame per workgroup, one wavefront per tile; H = 16 MT,
// sampled-row capacity kp <= 32 KJ.  The frame's DFT matrices do not depend on the coil: the workgroup stages the
// inverse one in LDS for the whole coil loop of A^H r, then the forward one for the coil loop of the next iteration's
// forward columns, so each A fragment leaves L2 once per workgroup instead of once per coil and wavefront.
// ---------------------------------------------------------------------------------------------------------------------
template <int MT, int KJ>
__device__ __forceinline__ void nesta_CA_mfma(const SliceDesc& d, int local, int mode, uint16_t* lds) {
    constexpr int H = 16 * MT, KS = MT / 2;
    const int l = threadIdx.x & 63, n = l & 15, g = l >> 4;
    const int ngrp = (d.ncol_tiles + MF_CA_TILES - 1) / MF_CA_TILES;
    const int tg = local % ngrp, f = local / ngrp;
    const int tile = tg * MF_CA_TILES + (threadIdx.x >> 6);
    const bool act = tile < d.ncol_tiles;                  // the last group of a frame may have idle wavefronts
    const int F = d.F, C = d.C, W = d.W, nmax = d.nmax, kp = d.kp;
    const long N = d.N;
    const int w = (act ? tile : 0) * 16 + n;
    const int nr = d.nrows[f];
    const int ldI = mf_lds_ld(kp), ldF = mf_lds_ld(H);
    f32x4 ar[MT], ai[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t) { ar[t] = f32x4{0.f, 0.f, 0.f, 0.f}; ai[t] = ar[t]; }

    // ---- A^H r: Z = D_inv(f) V per coil (K = sampled rows), conj(S_c) multiply, coil sum ----
    stage_planes(lds, d.dcolI + (long)f * 4 * H * kp, H, kp);
    __syncthreads();
    for (int c = 0; c < C && act; ++c) {
        const float2* kb = d.kbuf + ((long)f * C + c) * nmax * W;
        const float2* Sc = d.S + (long)c * N;
        Op8 bj[KJ];
#pragma unroll
        for (int s = 0; s < KJ; ++s) {
            float2 v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int jj = 32 * s + 8 * g + e;
                v[e] = jj < nr ? ldg2(kb, (unsigned)(jj * W + w)) : make_float2(0.f, 0.f);
            }
            split8(v, bj[s]);
        }
        // one m-tile in flight ahead: its A fragments (LDS) and S values (global) load during this tile's MFMAs
        Op8 an[KJ];
        float2 sn[4];
#pragma unroll
        for (int s = 0; s < KJ; ++s) an[s] = loadA_lds(lds, H, ldI, n, 32 * s + 8 * g);
#pragma unroll
        for (int i = 0; i < 4; ++i) sn[i] = ldg2(Sc + (long)(4 * g + i) * W, (unsigned)w);
#pragma unroll
        for (int t = 0; t < MT; ++t) {
            Op8 a[KJ];
            float2 sv[4];
#pragma unroll
            for (int s = 0; s < KJ; ++s) a[s] = an[s];
#pragma unroll
            for (int i = 0; i < 4; ++i) sv[i] = sn[i];
            if (t + 1 < MT) {
#pragma unroll
                for (int s = 0; s < KJ; ++s) an[s] = loadA_lds(lds, H, ldI, 16 * (t + 1) + n, 32 * s + 8 * g);
#pragma unroll
                for (int i = 0; i < 4; ++i) sn[i] = ldg2(Sc + (long)(16 * (t + 1) + 4 * g + i) * W, (unsigned)w);
            }
            f32x4 zr = {0.f, 0.f, 0.f, 0.f}, zi = zr;
#pragma unroll
            for (int s = 0; s < KJ; ++s)
                if (32 * s < kp) cmac(zr, zi, a[s], bj[s]);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                ar[t][i] += zr[i] * sv[i].x + zi[i] * sv[i].y;
                ai[t][i] += zi[i] * sv[i].x - zr[i] * sv[i].y;
            }
            pipeline_fence();
        }
    }
    __syncthreads();
    stage_planes(lds, d.dcolF + (long)f * 4 * kp * H, kp, H);
    __syncthreads();
    if (!act) return;

    // ---- NESTA update at h = 16 t + 4 g + i (shared with the FP32 passes); x_{k+1} replaces the sums ----
    const float sc = d.inv_sqrtN;
    if (mode == 1) {
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                ar[t][i] *= sc;
                ai[t][i] *= sc;
                stg2(d.x0 + (long)f * N + (long)(16 * t + 4 * g + i) * W, (unsigned)w, make_float2(ar[t][i], ai[t][i]));
            }
    } else {
        const UpdRow u = upd_row(d, f);
#pragma unroll
        for (int t = 0; t < MT; ++t) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float2 xn = nesta_update_at(u, (long)(16 * t + 4 * g + i) * W, (unsigned)w,
                                                  make_float2(ar[t][i] * sc, ai[t][i] * sc));
                ar[t][i] = xn.x;
                ai[t][i] = xn.y;
            }
            pipeline_fence();   // bounds the loads in flight to one m-tile's pixels
        }
    }

    // ---- forward column DFT of S_c x_{k+1} onto the sampled rows, K (= h) in the permuted order ----
    for (int c = 0; c < C; ++c) {
        const float2* Sc = d.S + (long)c * N;
        float2* kb = d.kbuf + ((long)f * C + c) * nmax * W;
        f32x4 orr[2 * KJ], oi[2 * KJ];
#pragma unroll
        for (int q = 0; q < 2 * KJ; ++q) { orr[q] = f32x4{0.f, 0.f, 0.f, 0.f}; oi[q] = orr[q]; }
        // one k-step in flight ahead: its S values load during this step's MFMAs
        float2 sn[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) sn[e] = ldg2(Sc + (long)(16 * (e >> 2) + 4 * g + (e & 3)) * W, (unsigned)w);
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            float2 v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int t = 2 * s + (e >> 2), i = e & 3;
                v[e] = make_float2(ar[t][i] * sn[e].x - ai[t][i] * sn[e].y, ar[t][i] * sn[e].y + ai[t][i] * sn[e].x);
            }
            if (s + 1 < KS) {
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    sn[e] = ldg2(Sc + (long)(16 * (2 * s + 2 + (e >> 2)) + 4 * g + (e & 3)) * W, (unsigned)w);
            }
            Op8 b;
            split8(v, b);
#pragma unroll
            for (int q = 0; q < 2 * KJ; ++q)
                if (16 * q < kp) cmac(orr[q], oi[q], loadA_lds(lds, kp, ldF, 16 * q + n, 32 * s + 8 * g), b);
            pipeline_fence();
        }
#pragma unroll
        for (int q = 0; q < 2 * KJ; ++q)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int jj = 16 * q + 4 * g + i;
                if (jj < nr) stg2(kb, (unsigned)(jj * W + w), make_float2(orr[q][i], oi[q][i]));
            }
    }
}

template <int MT>
__global__ __launch_bounds__(64 * MF_B_TASKS) void k_nesta_B_mfma(const SliceDesc* __restrict__ descs,
                                                                  const int* __restrict__ wg_off, int nslices,
                                                                  int mode) {
    extern __shared__ uint16_t mlds[];
    __shared__ double red[64];
    const int si = find_slice(wg_off, nslices, blockIdx.x);
    const SliceDesc& d = descs[si];
    if (mode == 0 && !d.st->active) return;
    const int local = blockIdx.x - wg_off[si];
    if (local >= d.wg_B) {
        if (mode == 0) nesta_U_task(d, local - d.wg_B, red);
        return;
    }
    double acc[1] = {nesta_B_mfma<MT>(d, local, mode, mlds)};
    if (mode != 0) return;
    block_sum<1>(acc, red);
    if (threadIdx.x == 0) d.partB[local] = acc[0];
}

template <int MT, int KJ>
__global__ __launch_bounds__(64 * MF_CA_TILES) void k_nesta_CA_mfma(const SliceDesc* __restrict__ descs,
                                                                    const int* __restrict__ wg_off, int nslices,
                                                                    int mode) {
    extern __shared__ uint16_t mlds[];
    const int si = find_slice(wg_off, nslices, blockIdx.x);
    const SliceDesc& d = descs[si];
    if (mode == 0 && !d.st->it_pending) return;
    nesta_CA_mfma<MT, KJ>(d, blockIdx.x - wg_off[si], mode, mlds);
}

typedef void (*NestaKernel)(const SliceDesc*, const int*, int, int);

NestaKernel pick_B_mfma(int MT) {
    switch (MT) {
        case 4: return k_nesta_B_mfma<4>;
        case 8: return k_nesta_B_mfma<8>;
        case 12: return k_nesta_B_mfma<12>;
        case 16: return k_nesta_B_mfma<16>;
        default: return nullptr;
    }
}

NestaKernel pick_CA_mfma(int MT, int KJ) {
    if (KJ == 1) {
        switch (MT) {
            case 4: return k_nesta_CA_mfma<4, 1>;
            case 8: return k_nesta_CA_mfma<8, 1>;
            case 12: return k_nesta_CA_mfma<12, 1>;
            case 16: return k_nesta_CA_mfma<16, 1>;
            default: return nullptr;
        }
    }
    if (KJ == 2) {
        switch (MT) {
            case 4: return k_nesta_CA_mfma<4, 2>;
            case 8: return k_nesta_CA_mfma<8, 2>;
            case 12: return k_nesta_CA_mfma<12, 2>;
            case 16: return k_nesta_CA_mfma<16, 2>;
            default: return nullptr;
        }
    }
    return nullptr;
}

// ---------------------------------------------------------------------------------------------------------------------
// Gram pass of the split-precision form (GS): the Gram form of the fused iteration for line masks (the identities are
// stated in ocl_nesta_gram.hpp) with both column DFTs as split-precision GEMMs, so no row transform runs in the
// iteration at all.  One workgroup = MF_CA_TILES column tiles of 16 columns of one frame; each tile is covered by
// NR = MT / RT row-waves of RW = 16 RT rows, and a wavefront keeps the A^H r accumulator of its RW x 16 pixels in
// registers, in the MFMA output layout h = h0 + 16 t + 4 g + i (t < RT, i < 4, column n of its tile).  Both DFT
// matrices of the frame are staged in LDS once per workgroup:
//   DA = dcolF (kind 3) [4][kp][H]: K index = h permuted, so the B operand of k-step s is the lane's own pixels of
//        m-tiles 2 s and 2 s + 1;
//   DB = dcolI (kind 4) [4][H][kp]: K index = j permuted, so the B operand is the lane's own forward outputs.
// Per coil c: K = DA (S_c x_k) over the wave's rows (a partial sum; the two row-waves of a tile exchange their partials
// through LDS), rho = m K W / sqrt(N) - V_c at the carried rows (V_c = W F_x^{-1} y_c in kbuf; m = 1 on sampled mask
// rows), ||rho||^2, Z = DB rho, acc += conj(S_c) Z.  Then the NESTA update of the wave's pixels with the Gram form's
// ping-ponged x and W, and the partial sums of the scalar step (k_nesta_S2).  mode 1 (setup): x0 = x_ref =
// A^H sqrt(N) y = sum_c conj(S_c) DB V_c / sqrt(N).
// Global accesses go through buffer resources: a 32-bit lane offset (in bytes) and a scalar row offset, so no per-lane
// 64-bit addresses are formed or kept.
// ---------------------------------------------------------------------------------------------------------------------
// NESTA update (NestaUp.cpp:696-875) of the wave's pixels from x_k and A^H r (acc, in registers); accumulates |u_f|^2
// and Re(conj(u_f) Ux_f) of x_k.  SST: first iteration of a stage (x_ref := x_k, W_k := 0).  so0: byte offset of row h0
// of frame f; vo: the lane's byte offset.
template <int RT, bool SST>
__device__ __forceinline__ void gs_update(const SliceDesc& d, const NState* st, int f, unsigned so0, unsigned vo,
                                          bool act, const float2 (&acc)[RT][4], double* uterm) {
    const long N = d.N;
    const unsigned W8 = 8u * (unsigned)d.W, fb = (unsigned)(8 * (long)f * N);
    const long bytes = 8 * (long)d.F * N;
    const int fm = f == 0 ? d.F - 1 : f - 1, fp = f == d.F - 1 ? 0 : f + 1;
    // buffer selections as scalars (the state is read with vector loads)
    const int cur = __builtin_amdgcn_readfirstlane(st->it_cur), wcur = __builtin_amdgcn_readfirstlane(st->it_wcur);
    const BufRes rx = buf_res(cur ? d.x1 : d.x0, bytes), rxo = buf_res(cur ? d.x0 : d.x1, bytes);
    const BufRes rwi = buf_res(wcur ? d.Wk2 : d.Wk, bytes), rwo = buf_res(wcur ? d.Wk : d.Wk2, bytes);
    const BufRes rxr = buf_res(d.xref, bytes);
    const unsigned om = so0 - fb + (unsigned)(8 * (long)fm * N), op = so0 - fb + (unsigned)(8 * (long)fp * N);
    const float mu = (float)st->it_mu, imu = 1.0f / mu, lam = (float)st->lambda, L1 = (float)st->it_L1;
    const float apk = (float)st->it_apk, tau = (float)st->it_tau, sc = d.inv_sqrtN;
    float u2 = 0.f, fx = 0.f;
#pragma unroll
    for (int t = 0; t < RT; ++t) {
        float2 a[4], b[4], xr[4], wo[4], xc[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const unsigned ro = (unsigned)(16 * t + i) * W8;
            xc[i] = bld2(rx, vo, so0 + ro);
            a[i] = bld2(rx, vo, om + ro);
            b[i] = bld2(rx, vo, op + ro);
            if constexpr (!SST) {
                xr[i] = bld2(rxr, vo, so0 + ro);
                wo[i] = bld2(rwi, vo, so0 + ro);
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const unsigned ro = so0 + (unsigned)(16 * t + i) * W8;
            const float2 xf = xc[i];
            if constexpr (SST) {
                xr[i] = xf;
                wo[i] = make_float2(0.f, 0.f);
            }
            // u = Ux / max(mu, |Ux|) as Ux * min(1/mu, rsqrt(|Ux|^2)), for frames f and f + 1
            const float2 uxf = c_sub(xf, a[i]), uxp = c_sub(b[i], xf);
            const float2 uf = c_scale(uxf, fminf(imu, __builtin_amdgcn_rsqf(c_abs2(uxf))));
            const float2 up = c_scale(uxp, fminf(imu, __builtin_amdgcn_rsqf(c_abs2(uxp))));
            const float2 ar = c_scale(acc[t][i], sc);
            const float2 df = make_float2(lam * (uf.x - up.x) + ar.x, lam * (uf.y - up.y) + ar.y);
            const float2 yk = make_float2(xf.x - L1 * df.x, xf.y - L1 * df.y);
            const float2 wn = make_float2(wo[i].x + apk * df.x, wo[i].y + apk * df.y);
            const float2 z = make_float2(xr[i].x - L1 * wn.x, xr[i].y - L1 * wn.y);
            const float2 x1 = make_float2(tau * z.x + (1.0f - tau) * yk.x, tau * z.y + (1.0f - tau) * yk.y);
            if (act) {
                u2 += c_abs2(uf);
                fx += uf.x * uxf.x + uf.y * uxf.y;
                if constexpr (SST) bst2(rxr, vo, ro, xf);
                bst2(rwo, vo, ro, wn);
                bst2(rxo, vo, ro, x1);
            }
        }
        pipeline_fence();   // one m-tile's operands in flight at a time
    }
    uterm[0] += (double)u2;
    uterm[1] += (double)fx;
}

// The pass itself.  The two row-waves of a tile exchange their forward partials through LDS (own + peer is the same
// sum in both, as floating-point addition commutes).  S_c and x_k are streamed from L2 one step ahead of each use.
template <int MT, int KJ, int RT>
__device__ __forceinline__ void nesta_GS(const SliceDesc& d, int local, int mode, uint16_t* lds, double* red) {
    constexpr int H = 16 * MT, KP = 32 * KJ, MJ = 2 * KJ, RW = 16 * RT, NR = MT / RT, KS1 = RT / 2;
    constexpr int TC = MF_CA_TILES;
    static_assert(MT % RT == 0 && RT % 2 == 0, "a row-wave covers a multiple of 32 rows of the column");
    static_assert(NR <= 2, "the partial exchange pairs at most two row-waves");
    constexpr int ldA = mf_lds_ld(H), ldB = mf_lds_ld(KP);
    const int l = threadIdx.x & 63, n = l & 15, g = l >> 4;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), tcs = wv % TC, q = wv / TC;
    // consecutive workgroups take consecutive frames of one column-tile group (they share S_c's columns)
    const int ngrp = (d.ncol_tiles + TC - 1) / TC;
    const int f = local % d.F, tg = local / d.F;
    const int tile = tg * TC + tcs;
    const bool act = tile < d.ncol_tiles;        // idle wavefronts of a last group still meet every barrier
    const int C = d.C, W = d.W, nmax = d.nmax;
    const long N = d.N;
    const int nr = d.nrows[f];
    const int h0 = q * RW;
    const NState* st = d.st;
    const unsigned vo = 8u * (unsigned)(4 * g * W + (act ? tile : 0) * 16 + n);
    const unsigned so0 = (unsigned)(8 * ((long)f * N + (long)h0 * W));
    const BufRes rS = buf_res(d.S, 8 * (long)C * N);
    const BufRes rV = buf_res(d.kbuf, 8 * (long)d.F * C * nmax * W);
    uint16_t* DA = lds;
    uint16_t* DB = lds + 4 * KP * ldA;
    float* slab = (float*)(DB + 4 * H * ldB);
    if (mode == 0) stage_planes(DA, d.dcolF + (long)f * 4 * KP * H, KP, H);
    stage_planes(DB, d.dcolI + (long)f * 4 * H * KP, H, KP);
    (void)ngrp;

    // residual factor W / sqrt(N) at the lane's carried-row slots j = 16 mj + 4 g + i that are sampled mask rows,
    // as one bit per slot
    const float sW = (float)W * d.inv_sqrtN;
    unsigned mkb = 0;
#pragma unroll
    for (int mj = 0; mj < MJ; ++mj)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int j = 16 * mj + 4 * g + i;
            if (j < nr && d.rowmask[(long)f * nmax + j]) mkb |= 1u << (4 * mj + i);
        }
    unsigned W8 = 8u * (unsigned)W;
    auto load_v = [&](int c, float2 (&v)[MJ][4]) {
        const unsigned sb = (unsigned)(8 * ((long)f * C + c) * nmax * W);
#pragma unroll
        for (int mj = 0; mj < MJ; ++mj)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int j = 16 * mj + 4 * g + i;
                v[mj][i] = j < nr ? bld2(rV, vo, sb + (unsigned)(16 * mj + i) * W8) : make_float2(0.f, 0.f);
            }
    };
    // S_c at the rows of m-tiles t and t + 1 of this wave
    auto load_s2 = [&](int c, int t, float2 (&sv)[8]) {
        const unsigned sb = (unsigned)(8 * ((long)c * N + (long)h0 * W));
#pragma unroll
        for (int e = 0; e < 8; ++e) sv[e] = bld2(rS, vo, sb + (unsigned)(16 * (t + (e >> 2)) + (e & 3)) * W8);
    };

    // x_k is re-read per coil (an L2 hit: the workgroups of a column-tile group run on one XCD) instead of being held
    // in registers across the coil loop
    const BufRes rX = buf_res(__builtin_amdgcn_readfirstlane(st->it_cur) ? d.x1 : d.x0, 8 * (long)d.F * N);
    auto load_x2 = [&](int t, float2 (&xv)[8]) {
#pragma unroll
        for (int e = 0; e < 8; ++e) xv[e] = bld2(rX, vo, so0 + (unsigned)(16 * (t + (e >> 2)) + (e & 3)) * W8);
    };
    float2 xa[8], xb[8];
    float2 acc[RT][4], vv[MJ][4], sa[8], sb[8];
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[t][i] = make_float2(0.f, 0.f);
    // inverse column DFT of rho (as the B operands rop) from the carried rows, two m-tiles per step, conj(S_c)
    // accumulate; sa holds S_c of m-tiles 0, 1 on entry and S of the next coil's m-tiles 0, 1 on exit
    auto inverse_acc = [&](int c, const Op8 (&rop)[KJ]) {
#pragma unroll
        for (int t = 0; t < RT; t += 2) {
            if (t + 2 < RT) load_s2(c, t + 2, sb);
            else if (c + 1 < C) load_s2(c + 1, 0, sb);
            if (mode == 0 && t + 2 >= RT && c + 1 < C) load_x2(0, xa);   // the next coil's first forward operand
            f32x4 zr[2], zn[2], zi[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                zr[u] = f32x4{0.f, 0.f, 0.f, 0.f};
                zn[u] = zr[u];
                zi[u] = zr[u];
            }
#pragma unroll
            for (int s = 0; s < KJ; ++s)
#pragma unroll
                for (int u = 0; u < 2; ++u)
                    cmac3(zr[u], zn[u], zi[u], loadA_lds(DB, H, ldB, h0 + 16 * (t + u) + n, 32 * s + 8 * g), rop[s]);
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float zx = zr[u][i] - zn[u][i], zy = zi[u][i];
                    const float2 sv_ = sa[4 * u + i];
                    acc[t + u][i].x = fmaf(zx, sv_.x, fmaf(zy, sv_.y, acc[t + u][i].x));
                    acc[t + u][i].y = fmaf(zy, sv_.x, fmaf(-zx, sv_.y, acc[t + u][i].y));
                }
#pragma unroll
            for (int e = 0; e < 8; ++e) sa[e] = sb[e];
            pipeline_fence();
        }
    };
    auto split_rho = [&](const float2 (&rho)[MJ][4], Op8 (&rop)[KJ]) {
#pragma unroll
        for (int s = 0; s < KJ; ++s) {
            float2 v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = rho[2 * s + (e >> 2)][e & 3];
            split8p(v, rop[s]);
        }
    };

    if (mode == 1) {
        // ---- setup: x0 = x_ref = sum_c conj(S_c) DB V_c / sqrt(N) ----
        load_v(0, vv);
        load_s2(0, 0, sa);
        __syncthreads();                          // DB staged
        for (int c = 0; c < C; ++c) {
            asm volatile("" : "+s"(W8) : : "memory");
            Op8 rop[KJ];
            split_rho(vv, rop);
            if (c + 1 < C) load_v(c + 1, vv);
            inverse_acc(c, rop);
        }
        if (act) {
            const BufRes r0 = buf_res(d.x0, 8 * (long)d.F * N);
#pragma unroll
            for (int t = 0; t < RT; ++t)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    bst2(r0, vo, so0 + (unsigned)(16 * t + i) * W8, c_scale(acc[t][i], d.inv_sqrtN));
        }
        return;
    }

    load_s2(0, 0, sa);
    load_x2(0, xa);
    double r2 = 0.0;
    __syncthreads();                              // DA, DB staged

    for (int c = 0; c < C; ++c) {
        asm volatile("" : "+s"(W8) : : "memory");
        // ---- forward column DFT onto the carried rows (partial over this wave's rows); sa, xa hold S and x of
        // k-step s (m-tiles 2 s, 2 s + 1) ----
        f32x4 kr[MJ], kn[MJ], ki[MJ];
#pragma unroll
        for (int mj = 0; mj < MJ; ++mj) {
            kr[mj] = f32x4{0.f, 0.f, 0.f, 0.f};
            kn[mj] = kr[mj];
            ki[mj] = kr[mj];
        }
#pragma unroll
        for (int s = 0; s < KS1; ++s) {
            // the next operands: k-step s + 1, or S_c of m-tiles 0, 1 for the inverse DFT
            load_s2(c, s + 1 < KS1 ? 2 * s + 2 : 0, sb);
            if (s + 1 < KS1) load_x2(2 * s + 2, xb);
            float2 v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = c_mul(xa[e], sa[e]);
            Op8 b;
            split8p(v, b);
#pragma unroll
            for (int mj = 0; mj < MJ; ++mj)
                cmac3(kr[mj], kn[mj], ki[mj], loadA_lds(DA, KP, ldA, 16 * mj + n, h0 + 32 * s + 8 * g), b);
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                sa[e] = sb[e];
                if (s + 1 < KS1) xa[e] = xb[e];
            }
            pipeline_fence();
        }
#pragma unroll
        for (int mj = 0; mj < MJ; ++mj) kr[mj] -= kn[mj];
        load_v(c, vv);                            // V_c, in flight across the partial exchange
        if constexpr (NR > 1) {
            float* own = slab + (long)wv * (MJ * 8 * 64) + l;
            const float* peer = slab + (long)((1 - q) * TC + tcs) * (MJ * 8 * 64) + l;
            __syncthreads();                      // the previous coil's partials have been read
#pragma unroll
            for (int mj = 0; mj < MJ; ++mj)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    own[((mj * 4 + i) * 2) * 64] = kr[mj][i];
                    own[((mj * 4 + i) * 2 + 1) * 64] = ki[mj][i];
                }
            __syncthreads();
#pragma unroll
            for (int mj = 0; mj < MJ; ++mj)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    kr[mj][i] += peer[((mj * 4 + i) * 2) * 64];
                    ki[mj][i] += peer[((mj * 4 + i) * 2 + 1) * 64];
                }
        }
        float2 rho[MJ][4];
        float rq = 0.f;
#pragma unroll
        for (int mj = 0; mj < MJ; ++mj)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float mkv = (mkb >> (4 * mj + i)) & 1u ? sW : 0.f;
                const float2 rv = make_float2(fmaf(mkv, kr[mj][i], -vv[mj][i].x), fmaf(mkv, ki[mj][i], -vv[mj][i].y));
                rq = fmaf(rv.x, rv.x, fmaf(rv.y, rv.y, rq));
                rho[mj][i] = rv;
            }
        if (q == 0 && act) r2 += (double)rq;
        Op8 rop[KJ];
        split_rho(rho, rop);
        inverse_acc(c, rop);
    }
    double part[3] = {r2 / (double)W, 0.0, 0.0};
    if (__builtin_amdgcn_readfirstlane(st->it_sstart)) gs_update<RT, true>(d, st, f, so0, vo, act, acc, part + 1);
    else gs_update<RT, false>(d, st, f, so0, vo, act, acc, part + 1);
    block_sum<3>(part, red);
    if (threadIdx.x == 0) {
        d.partG[3 * local] = part[0];
        d.partG[3 * local + 1] = part[1];
        d.partG[3 * local + 2] = part[2];
    }
}

// Logical workgroup of a launch whose physical workgroups are dealt round-robin over the 8 XCDs: XCD x runs the x-th
// contiguous range of logical workgroups, in order, so workgroups that share data (a column-tile group's frames) share
// that XCD's L2.
__device__ __forceinline__ int xcd_logical(int b, int total) {
    const int q = total >> 3, r = total & 7, x = b & 7, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

// KJM: largest kp / 32 of the table (a slice with kp = 32 in a KJM = 2 launch takes the KJ = 1 path).  Workgroups
// are taken in XCD-contiguous order, and consecutive ones are consecutive frames of one column-tile group, so the
// frames that read the same columns of S_c share an L2.
template <int MT, int RT, int KJM>
__global__ __launch_bounds__(64 * MF_CA_TILES * (MT / RT)) void k_nesta_GS(const SliceDesc* __restrict__ descs,
                                                                           const int* __restrict__ wg_off,
                                                                           int nslices, int mode) {
    extern __shared__ uint16_t mlds[];
    __shared__ double red[64];
    const int b = xcd_logical(blockIdx.x, gridDim.x);
    const int si = find_slice(wg_off, nslices, b);
    const SliceDesc& d = descs[si];
    if (mode == 0 && !d.st->it_pending) return;
    const int local = b - wg_off[si];
    if (KJM == 1 || d.kp == 32) nesta_GS<MT, 1, RT>(d, local, mode, mlds, red);
    else nesta_GS<MT, KJM, RT>(d, local, mode, mlds, red);
}

NestaKernel pick_GS(int MT, int KJ) {
    switch (MT) {
        case 4: return KJ == 1 ? k_nesta_GS<4, mf_gs_rt(4), 1> : k_nesta_GS<4, mf_gs_rt(4), 2>;
        case 8: return KJ == 1 ? k_nesta_GS<8, mf_gs_rt(8), 1> : nullptr;
        case 12: return KJ == 1 ? k_nesta_GS<12, mf_gs_rt(12), 1> : nullptr;
        default: return nullptr;
    }
}

}  // namespace

bool mfma_size_supported(int n) { return n == 64 || n == 128 || n == 192 || n == 256; }

// The Gram pass is instantiated for H = 64 (kp 32 or 64), 128 and 192 (kp 32): the sizes whose two DFT matrices and
// partial slabs fit one workgroup's LDS.
bool mfma_gram_supported(int H, int kp) {
    if (H % 16 != 0 || kp % 32 != 0 || !pick_GS(H / 16, kp / 32)) return false;
    return mf_gs_lds_bytes(H, kp) + 64 * sizeof(double) <= 160 * 1024;
}

hipError_t launch_nesta_G_mfma(const NestaLaunch& L, hipStream_t stream, int mode) {
    NestaKernel k = pick_GS(L.PH, L.kj);
    if (!k) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k, dim3(L.totCA), dim3(64 * MF_CA_TILES * (L.PH / mf_gs_rt(L.PH))),
                       mf_gs_lds_bytes(16 * L.PH, 32 * L.kj), stream, L.dd, L.offCA, L.ns, mode);
    return hipGetLastError();
}

void mfma_kernel_attributes(int lds_bytes) {
    for (int mt : {4, 8, 12, 16})
        (void)hipFuncSetAttribute((const void*)pick_B_mfma(mt), hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
    for (int mt : {4, 8, 12, 16})
        for (int kj : {1, 2})
            (void)hipFuncSetAttribute((const void*)pick_CA_mfma(mt, kj), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      lds_bytes);
    for (int mt : {4, 8, 12})
        for (int kj : {1, 2})
            if (pick_GS(mt, kj))
                (void)hipFuncSetAttribute((const void*)pick_GS(mt, kj), hipFuncAttributeMaxDynamicSharedMemorySize,
                                          lds_bytes);
    (void)hipGetLastError();
}

hipError_t launch_nesta_B_mfma(const NestaLaunch& L, hipStream_t stream, int mode) {
    NestaKernel k = pick_B_mfma(L.PW);
    if (!k) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k, dim3(L.totB), dim3(64 * MF_B_TASKS), L.ldsB, stream, L.dd, L.offB, L.ns, mode);
    return hipGetLastError();
}

hipError_t launch_nesta_CA_mfma(const NestaLaunch& L, hipStream_t stream, int mode) {
    NestaKernel k = pick_CA_mfma(L.PH, L.kj);
    if (!k) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k, dim3(L.totCA), dim3(64 * MF_CA_TILES), L.ldsCA, stream, L.dd, L.offCA, L.ns, mode);
    return hipGetLastError();
}

// DFT matrices of the split-precision passes, computed in double and split into bf16 hi/lo planes
// (plane order re_hi, re_lo, im_hi, im_lo; plane = M x K entries, row-major).  blockIdx.y = frame for the column kinds.
//   kind 0  row forward   M = K = W : A[kx][w]    = exp(-2 pi i kx w / W)
//   kind 1  row inverse   M = K = W : A[w][k]     = exp(+2 pi i kperm(k) w / W)
//   kind 2  col inverse   M = H, K = kp : A[h][j] = exp(+2 pi i rows_j h / H)          (0 for j >= nrows_f)
//   kind 3  col forward   M = kp, K = H : A[j][k] = exp(-2 pi i rows_j kperm(k) / H)   (0 for j >= nrows_f)
//   kind 4  col inverse   M = H, K = kp : A[h][k] = exp(+2 pi i rows_kperm(k) h / H)   (0 for kperm(k) >= nrows_f);
//           the Gram pass's inverse, whose B operand is the lane's own forward output (K index permuted)
__global__ void k_dft_split(uint16_t* out, int kind, int L, int kp, const int16_t* rows, const int* nrows, int nmax) {
    const int f = blockIdx.y;
    const int M = kind == 3 ? kp : L, K = (kind == 2 || kind == 4) ? kp : L;
    const long plane = (long)M * K;
    uint16_t* base = out + (kind >= 2 ? (long)f * 4 * plane : 0);
    const int nr = kind >= 2 ? nrows[f] : 0;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < plane; idx += (long)gridDim.x * blockDim.x) {
        const int m = (int)(idx / K), k = (int)(idx - (long)m * K);
        long prod = 0;
        double sign = 1.0;
        bool zero = false;
        switch (kind) {
            case 0: prod = (long)m * k; sign = -1.0; break;
            case 1: prod = (long)m * kperm(k); break;
            case 2: if (k < nr) prod = (long)rows[(long)f * nmax + k] * m; else zero = true; break;
            case 3: if (m < nr) prod = (long)rows[(long)f * nmax + m] * kperm(k); else zero = true; sign = -1.0; break;
            default: {
                const int j = kperm(k);
                if (j < nr) prod = (long)rows[(long)f * nmax + j] * m; else zero = true;
                break;
            }
        }
        double sn = 0.0, cs = 0.0;
        if (!zero) sincospi(2.0 * (double)(prod % L) / (double)L, &sn, &cs);
        const double re = cs, im = sign * sn;
        const __bf16 rh = (__bf16)(float)re, ih = (__bf16)(float)im;
        const __bf16 rl = (__bf16)(float)(re - (double)(float)rh), il = (__bf16)(float)(im - (double)(float)ih);
        base[idx] = __builtin_bit_cast(uint16_t, rh);
        base[plane + idx] = __builtin_bit_cast(uint16_t, rl);
        base[2 * plane + idx] = __builtin_bit_cast(uint16_t, ih);
        base[3 * plane + idx] = __builtin_bit_cast(uint16_t, il);
    }
}
