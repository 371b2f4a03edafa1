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

}  // namespace

bool mfma_size_supported(int n) { return n == 64 || n == 128 || n == 192 || n == 256; }

void mfma_kernel_attributes(int lds_bytes) {
    for (int mt : {4, 8, 12, 16})
        (void)hipFuncSetAttribute((const void*)pick_B_mfma(mt), hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
    for (int mt : {4, 8, 12, 16})
        for (int kj : {1, 2})
            (void)hipFuncSetAttribute((const void*)pick_CA_mfma(mt, kj), hipFuncAttributeMaxDynamicSharedMemorySize,
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
__global__ void k_dft_split(uint16_t* out, int kind, int L, int kp, const int16_t* rows, const int* nrows, int nmax) {
    const int f = blockIdx.y;
    const int M = kind == 3 ? kp : L, K = kind == 2 ? kp : L;
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
            default: if (m < nr) prod = (long)rows[(long)f * nmax + m] * kperm(k); else zero = true; sign = -1.0; break;
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
