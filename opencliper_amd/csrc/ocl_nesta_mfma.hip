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
// Gram pass GS: the Gram form of the fused iteration for line masks (the identities are stated in ocl_nesta_gram.hpp)
// with both column DFTs as split-precision GEMMs, so no row transform runs in the iteration at all.
// One workgroup = MF_CA_TILES column tiles of 16 columns of one frame.  Each tile is covered by NR = MT / RT
// row-waves of RW = 16 RT rows; a wavefront keeps the A^H r accumulator of its RW x 16 pixels in registers, in the
// MFMA output layout h = h0 + 16 t + 4 g + i (t < RT, i < 4, column n of the tile).  Both DFT matrices of the frame
// are staged in LDS once per workgroup:
//   DA = dcolF (kind 3) [4][kp][H]: K index = h permuted, so the B operand of k-step s is the lane's own pixels of
//        m-tiles 2 s and 2 s + 1;
//   DB = dcolI (kind 4) [4][H][kp]: K index = j permuted, so the B operand is the lane's own forward outputs.
// Per coil c: K = DA (S_c x_k) over the wave's rows (a partial sum; the two row-waves of a tile exchange their
// partials through LDS), rho = m K W / sqrt(N) - V_c at the carried rows (V_c = W F_x^{-1} y_c in kbuf; m = 1 on
// sampled mask rows), ||rho||^2, Z = DB rho, acc += conj(S_c) Z.  x_k and S_c are re-read per coil (L2), one step
// ahead of their use.  Then the NESTA update of the wave's pixels with the Gram form's ping-ponged x and W, and the
// partial sums of the scalar step (k_nesta_S2).  mode 1 (setup): x0 = x_ref = A^H sqrt(N) y = sum_c conj(S_c) DB V_c
// / sqrt(N).
// Global accesses go through buffer resources: a 32-bit lane offset (4 g W + w, in bytes) and a scalar row offset,
// so no per-lane 64-bit addresses are formed or kept.
// ---------------------------------------------------------------------------------------------------------------------
typedef __amdgpu_buffer_rsrc_t BufRes;
__device__ __forceinline__ BufRes buf_res(const void* p, long bytes) {
    // raw buffer: stride 0, range in bytes, gfx9 data-format word
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)(unsigned)bytes, 0x00020000);
}
typedef float f32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 bld2(BufRes r, unsigned vo, unsigned so) {
    // the 64-bit load's value is taken as a whole (bit cast), never by element
    const f32x2 v = __builtin_bit_cast(f32x2, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ void bst2(BufRes r, unsigned vo, unsigned so, float2 x) {
    f32x2 v;
    v.x = x.x;
    v.y = x.y;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v),
                                          r, vo, so, 0);
}

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

template <int MT, int KJ, int RT>
__device__ __forceinline__ void nesta_GS(const SliceDesc& d, int local, int mode, uint16_t* lds, double* red) {
    constexpr int H = 16 * MT, KP = 32 * KJ, MJ = 2 * KJ, RW = 16 * RT, NR = MT / RT, KS1 = RT / 2;
    constexpr int TC = MF_CA_TILES;
    static_assert(MT % RT == 0 && RT % 2 == 0, "a row-wave covers a multiple of 32 rows of the column");
    static_assert(NR <= 2, "the partial exchange pairs at most two row-waves");
    constexpr int ldA = mf_lds_ld(H), ldB = mf_lds_ld(KP);
    const int l = threadIdx.x & 63, n = l & 15, g = l >> 4;
    // wavefront index as a scalar, so the row offsets derived from it stay wave-uniform
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), tcs = wv % TC, q = wv / TC;
    const int ngrp = (d.ncol_tiles + TC - 1) / TC;
    const int tg = local % ngrp, f = local / ngrp;
    const int tile = tg * TC + tcs;
    const bool act = tile < d.ncol_tiles;        // idle wavefronts of a frame's last group still meet every barrier
    const int C = d.C, W = d.W, nmax = d.nmax;
    const long N = d.N;
    const int nr = d.nrows[f];
    const int h0 = q * RW;
    const NState* st = d.st;
    const unsigned W8 = 8u * (unsigned)W;
    // byte offset of pixel (h0 + 16 t + 4 g + i, w) = so (scalar: frame, row h0 + 16 t + i) + vo (lane)
    const unsigned vo = 8u * (unsigned)(4 * g * W + (act ? tile : 0) * 16 + n);
    const unsigned so0 = (unsigned)(8 * ((long)f * N + (long)h0 * W));
    const BufRes rS = buf_res(d.S, 8 * (long)C * N);
    const BufRes rX = buf_res(__builtin_amdgcn_readfirstlane(st->it_cur) ? d.x1 : d.x0, 8 * (long)d.F * N);
    const BufRes rV = buf_res(d.kbuf, 8 * (long)d.F * C * nmax * W);
    uint16_t* DA = lds;
    uint16_t* DB = lds + 4 * KP * ldA;
    float* slab = (float*)(DB + 4 * H * ldB);
    if (mode == 0) stage_planes(DA, d.dcolF + (long)f * 4 * KP * H, KP, H);
    stage_planes(DB, d.dcolI + (long)f * 4 * H * KP, H, KP);

    // residual factor per carried row j = 16 mj + 4 g + i of this lane: W / sqrt(N) on a sampled mask row, else 0
    const float sW = (float)W * d.inv_sqrtN;
    float mk[MJ][4];
#pragma unroll
    for (int mj = 0; mj < MJ; ++mj)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int j = 16 * mj + 4 * g + i;
            mk[mj][i] = (j < nr && d.rowmask[(long)f * nmax + j]) ? sW : 0.f;
        }
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
    // S_c and x_k rows of m-tile t
    auto load_s = [&](int c, int t, float2 (&sv)[4]) {
        const unsigned sb = (unsigned)(8 * ((long)c * N + (long)h0 * W)) + (unsigned)(16 * t) * W8;
#pragma unroll
        for (int i = 0; i < 4; ++i) sv[i] = bld2(rS, vo, sb + (unsigned)i * W8);
    };
    auto load_x = [&](int t, float2 (&xv)[4]) {
#pragma unroll
        for (int i = 0; i < 4; ++i) xv[i] = bld2(rX, vo, so0 + (unsigned)(16 * t + i) * W8);
    };

    float2 acc[RT][4], vv[MJ][4], s1[2][4], x1[2][4];
    if (mode == 0) {
        load_s(0, 0, s1[0]);
        load_s(0, 1, s1[1]);
        load_x(0, x1[0]);
        load_x(1, x1[1]);
    }
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[t][i] = make_float2(0.f, 0.f);
    load_v(0, vv);
    double r2 = 0.0;
    __syncthreads();                              // DA, DB staged

    for (int c = 0; c < C; ++c) {
        // the DFT fragments are re-read from LDS per coil instead of being hoisted out of the loop into registers
        asm volatile("" ::: "memory");
        float2 rho[MJ][4];
        float2 s2[4];
        if (mode == 0) {
            // ---- forward column DFT onto the carried rows (partial over this wave's rows) ----
            f32x4 kr[MJ], ki[MJ];
#pragma unroll
            for (int mj = 0; mj < MJ; ++mj) { kr[mj] = f32x4{0.f, 0.f, 0.f, 0.f}; ki[mj] = kr[mj]; }
#pragma unroll
            for (int s = 0; s < KS1; ++s) {
                float2 s1n[2][4], x1n[2][4];
                if (s + 1 < KS1) {
                    load_s(c, 2 * s + 2, s1n[0]);
                    load_s(c, 2 * s + 3, s1n[1]);
                    load_x(2 * s + 2, x1n[0]);
                    load_x(2 * s + 3, x1n[1]);
                } else {
                    load_s(c, 0, s2);
                }
                float2 v[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = c_mul(x1[e >> 2][e & 3], s1[e >> 2][e & 3]);
                Op8 b;
                split8(v, b);
#pragma unroll
                for (int mj = 0; mj < MJ; ++mj)
                    cmac(kr[mj], ki[mj], loadA_lds(DA, KP, ldA, 16 * mj + n, h0 + 32 * s + 8 * g), b);
                if (s + 1 < KS1)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        s1[0][i] = s1n[0][i];
                        s1[1][i] = s1n[1][i];
                        x1[0][i] = x1n[0][i];
                        x1[1][i] = x1n[1][i];
                    }
                pipeline_fence();
            }
            if constexpr (NR > 1) {
                // the two row-waves of the tile exchange their partials; own + peer is the same sum in both
                // (floating-point addition commutes)
                float* own = slab + (long)wv * (MJ * 8 * 64) + l;
                const float* peer = slab + (long)((1 - q) * TC + tcs) * (MJ * 8 * 64) + l;
                __syncthreads();                  // the previous coil's partials have been read
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
            // ---- residual at the carried rows ----
            float rq = 0.f;
#pragma unroll
            for (int mj = 0; mj < MJ; ++mj)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float2 rv = make_float2(fmaf(mk[mj][i], kr[mj][i], -vv[mj][i].x),
                                                  fmaf(mk[mj][i], ki[mj][i], -vv[mj][i].y));
                    rq = fmaf(rv.x, rv.x, fmaf(rv.y, rv.y, rq));
                    rho[mj][i] = rv;
                }
            if (q == 0 && act) r2 += (double)rq;
        } else {
            load_s(c, 0, s2);
#pragma unroll
            for (int mj = 0; mj < MJ; ++mj)
#pragma unroll
                for (int i = 0; i < 4; ++i) rho[mj][i] = vv[mj][i];
        }
        Op8 rop[KJ];
#pragma unroll
        for (int s = 0; s < KJ; ++s) {
            float2 v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = rho[2 * s + (e >> 2)][e & 3];
            split8(v, rop[s]);
        }
        if (c + 1 < C) load_v(c + 1, vv);
        // ---- inverse column DFT from the carried rows, conj(S_c) accumulate ----
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            float2 s2n[4];
            if (t + 1 < RT) {
                load_s(c, t + 1, s2n);
            } else if (mode == 0 && c + 1 < C) {          // the next coil's first forward operands
                load_s(c + 1, 0, s1[0]);
                load_s(c + 1, 1, s1[1]);
                load_x(0, x1[0]);
                load_x(1, x1[1]);
            }
            f32x4 zr = {0.f, 0.f, 0.f, 0.f}, zi = zr;
#pragma unroll
            for (int s = 0; s < KJ; ++s) cmac(zr, zi, loadA_lds(DB, H, ldB, h0 + 16 * t + n, 32 * s + 8 * g), rop[s]);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                acc[t][i].x = fmaf(zr[i], s2[i].x, fmaf(zi[i], s2[i].y, acc[t][i].x));
                acc[t][i].y = fmaf(zi[i], s2[i].x, fmaf(-zr[i], s2[i].y, acc[t][i].y));
            }
            if (t + 1 < RT)
#pragma unroll
                for (int i = 0; i < 4; ++i) s2[i] = s2n[i];
            pipeline_fence();
        }
    }
    if (mode == 1) {
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

// KJM: largest kp / 32 of the table (a slice with kp = 32 in a KJM = 2 launch takes the KJ = 1 path)
template <int MT, int RT, int KJM>
__global__ __launch_bounds__(64 * MF_CA_TILES * (MT / RT)) void k_nesta_GS(const SliceDesc* __restrict__ descs,
                                                                           const int* __restrict__ wg_off,
                                                                           int nslices, int mode) {
    extern __shared__ uint16_t mlds[];
    __shared__ double red[64];
    const int si = find_slice(wg_off, nslices, blockIdx.x);
    const SliceDesc& d = descs[si];
    if (mode == 0 && !d.st->it_pending) return;
    const int local = blockIdx.x - wg_off[si];
    if (KJM == 1 || d.kp == 32) nesta_GS<MT, 1, RT>(d, local, mode, mlds, red);
    else nesta_GS<MT, KJM, RT>(d, local, mode, mlds, red);
}

NestaKernel pick_GS(int MT, int KJ) {
    constexpr int R128 = 8 / OCL_GS_NR, R192 = 12 / OCL_GS_NR;
    switch (MT) {
        case 4: return KJ == 1 ? k_nesta_GS<4, 4, 1> : k_nesta_GS<4, 4, 2>;
        case 8: return KJ == 1 ? k_nesta_GS<8, R128, 1> : nullptr;
        case 12: return KJ == 1 ? k_nesta_GS<12, R192, 1> : nullptr;
        default: return nullptr;
    }
}

}  // namespace

bool mfma_size_supported(int n) { return n == 64 || n == 128 || n == 192 || n == 256; }

// The Gram pass is instantiated for H = 64 (kp 32 or 64), 128 and 192 (kp 32): the sizes whose two DFT matrices and
// partial slabs fit one workgroup's LDS.
bool mfma_gram_supported(int H, int kp) {
    if (!pick_GS(H / 16, kp / 32) || H % 16 != 0 || kp % 32 != 0) return false;
    return mf_gs_lds_bytes(H, kp) + 64 * sizeof(double) <= 160 * 1024;
}

hipError_t launch_nesta_G_mfma(const NestaLaunch& L, hipStream_t stream, int mode) {
    NestaKernel k = pick_GS(L.PH, L.kj);
    if (!k) return hipErrorInvalidValue;
    const int threads = 64 * MF_CA_TILES * mf_gs_nr(16 * L.PH);
    hipLaunchKernelGGL(k, dim3(L.totCA), dim3(threads), L.ldsCA, stream, L.dd, L.offCA, L.ns, mode);
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
