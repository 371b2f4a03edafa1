// Fused NESTA-TV iteration for gfx950 (src/processes/nesta/NestaUp.cpp:651-889 restated as two passes).
//
// The 2-D FFT of the encoding operator E = M F S is split column-first so that only the k-space rows that carry
// data ever leave the workgroup:
//   pass CA(k)   per (slice, frame, column tile):  for each coil  Z = IFFT_H(V) (input pruned to the sampled rows),
//                A^H r = sum_c conj(S_c) Z / sqrt(N); fused NESTA update of x, W (temporal TV from x_{f-1}, x_f,
//                x_{f+1}); then, for iteration k+1, U = FFT_H(S_c x_{k+1}) keeping only the sampled rows.
//   pass B(k+1)  per (slice, frame, coil, row group):  K = FFT_W(U row); r = M K / sqrt(N) - y; ||r||^2;
//                V = IFFT_W(r) written over U.
// The last workgroup of CA folds the per-workgroup partial sums in index order (deterministic), evaluates
// f = 1/2 ||r||^2 + lambda (Re<u,Ux> - mu/2 ||u||^2), runs the reference's stop test and advances the device state
// (NestaUp.cpp:797-889), so no host round trip is needed between iterations.
#include "ocl_nesta_common.hpp"
#include "ocl_nesta_fast.hpp"
#include "ocl_nesta_gram.hpp"

// Row flags for the sampled-row lists, flag[f][h]: bit 0 if the mask row or any coil's k-space row is non-zero,
// bit 1 if the mask row is not constant along the row, bit 2 if its first mask value is set.
__global__ void k_rowflags(const float2* __restrict__ y, const uint8_t* __restrict__ M, int F, int C, int H, int W,
                           uint8_t* __restrict__ flags) {
    const int fh = blockIdx.x;
    const int f = fh / H, h = fh - f * H;
    const uint8_t* m = M + ((long)f * H + h) * W;
    const int m0 = m[0] != 0;
    int any = 0, varies = 0;
    for (int i = threadIdx.x; i < W; i += blockDim.x) {
        const int mi = m[i] != 0;
        any |= mi;
        varies |= mi != m0;
    }
    varies = __syncthreads_or(varies);
    for (int c = 0; c < C && !__syncthreads_or(any); ++c)
        for (int i = threadIdx.x; i < W; i += blockDim.x) {
            const float2 v = y[(((long)f * C + c) * H + h) * W + i];
            any |= (v.x != 0.f || v.y != 0.f);
        }
    any = __syncthreads_or(any);
    if (threadIdx.x == 0) flags[fh] = (uint8_t)((any != 0) | ((varies != 0) << 1) | (m0 << 2));
}

// Mask bits of the sampled rows for the four-step pass B (W = 16 P): word (f, j, k1) holds, in bit k2, whether the mask
// of frame f at column k1 + P k2 of sampled row j is set (rows j >= nrows[f] get 0).
__global__ void k_mask_bits(const uint8_t* __restrict__ M, const int16_t* __restrict__ rows,
                            const int* __restrict__ nrows, int F, int H, int W, int nmax, int P,
                            uint16_t* __restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)F * nmax * P) return;
    const int k1 = (int)(i % P);
    const long fj = i / P;
    const int j = (int)(fj % nmax), f = (int)(fj / nmax);
    unsigned w = 0;
    if (j < nrows[f]) {
        const uint8_t* m = M + ((long)f * H + rows[(long)f * nmax + j]) * W;
        for (int k2 = 0; k2 < 16; ++k2) w |= (m[k1 + P * k2] != 0 ? 1u : 0u) << k2;
    }
    out[i] = (uint16_t)w;
}

// ------------------------------------------------------------------------------------------------------------------
// scalar step (NestaUp.cpp:797-889), run by the last workgroup of pass B(k) once ||r_k||^2 and the temporal-TV
// terms of x_k are complete.  It fixes the coefficients pass CA(k) applies and advances the state to k+1.
// Like the reference, an iteration that ends its stage through the stop test takes apk and tau with k = maxIter
// (NestaUp.cpp:809-831).
// ------------------------------------------------------------------------------------------------------------------
__device__ void nesta_scalar_step(const SliceDesc& d, double* red) {
    if (!d.st->active) {                 // every stage done: pass CA of this iteration must not run
        if (threadIdx.x == 0) d.st->it_pending = 0;
        return;
    }
    double s[3] = {0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < d.wg_B; i += blockDim.x) s[0] += d.partB[i];
    for (int i = threadIdx.x; i < d.wg_U; i += blockDim.x) { s[1] += d.partU[2 * i]; s[2] += d.partU[2 * i + 1]; }
    block_sum<3>(s, red);
    if (threadIdx.x != 0) return;
    NState* st = d.st;
    const double l2 = s[0], u2 = s[1], fxs = s[2];
    const double l1 = fxs - (st->mu / 2.0) * u2;
    const double fv = 0.5 * l2 + st->lambda * l1;
    st->last_f = fv;
    if (d.trace && st->ntrace < st->trace_cap) {
        d.trace[3 * st->ntrace] = fv;
        d.trace[3 * st->ntrace + 1] = l2;
        d.trace[3 * st->ntrace + 2] = l1;
        st->ntrace++;
    }
    int k = st->k;
    const int mi = st->miniter;
    bool end_stage = false;
    if (st->stoptest == 1) {
        bool skip_write = false;
        if (k >= mi) {
            double mean = 0.0;
            for (int ne = 0; ne < mi; ++ne) mean += st->fmean[mi - ne - 1];
            mean = mean / mi;
            const double qp = (mean - fv) / mean;
            st->last_qp = qp;
            if (qp <= st->tolVar) {
                if (st->qpDone) end_stage = true;
                st->qpDone = 1;
                skip_write = true;
            }
        }
        if (!skip_write) st->fmean[(mi - 1) - ((k + 1) % mi)] = fv;
    }
    // coefficients of iteration k for pass CA
    const double keff = end_stage ? (double)st->maxIter : (double)k;
    st->it_mu = st->mu;
    st->it_L1 = st->Lmu1;
    st->it_apk = 0.5 * (keff + 1.0);
    st->it_tau = 2.0 / (keff + 3.0);
    st->it_cur = st->cur;
    st->it_sstart = st->stage_start;
    st->it_pending = 1;
    // advance to iteration k+1
    st->iters_stage++;
    st->total_iters++;
    st->cur ^= 1;
    st->stage_start = 0;
    k = k + 1;
    if (end_stage || k >= st->maxIter) {
        if (st->stage - 1 < OCL_MAX_NSTAGES) st->iters[st->stage - 1] = st->iters_stage;
        st->stage++;
        st->iters_stage = 0;
        if (st->stage > st->maxIntIter) {
            st->active = 0;
        } else {
            st->mu = st->mu * st->gamma;
            st->tolVar = st->tolVar * st->gammat;
            st->Lmu = st->lambda * (st->normU / st->mu) + st->La;
            st->Lmu1 = 1.0 / st->Lmu;
            st->qpDone = 0;
            st->fmean[mi - 1] = 9223372036854775807.0;
            st->stage_start = 1;
            k = 0;
        }
    }
    st->k = k;
}

// ------------------------------------------------------------------------------------------------------------------
// scalar step of the Gram form, after pass G(k): the objective of x_k from G's partial sums, the stop test
// (NestaUp.cpp:797-831), then either the commit of iteration k with the coefficients of iteration k+1 for the next
// pass G, or, when the stop test ends the stage at k, a recompute of iteration k with apk and tau taken at
// k = maxIter (G read x_k and W_k and wrote x_{k+1}, W_{k+1} to the other buffers, so both inputs are still there).
// ------------------------------------------------------------------------------------------------------------------
__device__ void nesta_gram_step(const SliceDesc& d, double* red) {
    NState* st = d.st;
    if (!st->active || !st->it_pending) {
        if (threadIdx.x == 0) st->it_pending = 0;
        return;
    }
    double s[3] = {0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < d.wg_CA; i += blockDim.x) {
        s[0] += d.partG[3 * i];
        s[1] += d.partG[3 * i + 1];
        s[2] += d.partG[3 * i + 2];
    }
    block_sum<3>(s, red);
    if (threadIdx.x != 0) return;
    int k = st->k;
    const int mi = st->miniter;
    bool end_stage = false;
    if (st->it_redo) {
        st->it_redo = 0;                 // the recompute of a stage's last iteration: commit it, end the stage
        end_stage = true;
    } else {
        const double l2 = s[0], u2 = s[1], fxs = s[2];
        const double l1 = fxs - (st->mu / 2.0) * u2;
        const double fv = 0.5 * l2 + st->lambda * l1;
        st->last_f = fv;
        if (d.trace && st->ntrace < st->trace_cap) {
            d.trace[3 * st->ntrace] = fv;
            d.trace[3 * st->ntrace + 1] = l2;
            d.trace[3 * st->ntrace + 2] = l1;
            st->ntrace++;
        }
        if (st->stoptest == 1) {
            bool skip_write = false;
            if (k >= mi) {
                double mean = 0.0;
                for (int ne = 0; ne < mi; ++ne) mean += st->fmean[mi - ne - 1];
                mean = mean / mi;
                const double qp = (mean - fv) / mean;
                st->last_qp = qp;
                if (qp <= st->tolVar) {
                    if (st->qpDone) end_stage = true;
                    st->qpDone = 1;
                    skip_write = true;
                }
            }
            if (!skip_write) st->fmean[(mi - 1) - ((k + 1) % mi)] = fv;
        }
        st->iters_stage++;
        st->total_iters++;
        if (end_stage) {                 // recompute iteration k from the same x_k, W_k with k = maxIter
            st->it_apk = 0.5 * ((double)st->maxIter + 1.0);
            st->it_tau = 2.0 / ((double)st->maxIter + 3.0);
            st->it_redo = 1;
            st->it_pending = 1;
            return;
        }
    }
    // commit iteration k: x_{k+1} and W_{k+1} become current
    st->cur ^= 1;
    st->wcur ^= 1;
    st->stage_start = 0;
    k = k + 1;
    if (end_stage || k >= st->maxIter) {
        if (st->stage - 1 < OCL_MAX_NSTAGES) st->iters[st->stage - 1] = st->iters_stage;
        st->stage++;
        st->iters_stage = 0;
        if (st->stage > st->maxIntIter) {
            st->active = 0;
        } else {
            st->mu = st->mu * st->gamma;
            st->tolVar = st->tolVar * st->gammat;
            st->Lmu = st->lambda * (st->normU / st->mu) + st->La;
            st->Lmu1 = 1.0 / st->Lmu;
            st->qpDone = 0;
            st->fmean[mi - 1] = 9223372036854775807.0;
            st->stage_start = 1;
            k = 0;
        }
    }
    st->k = k;
    if (!st->active) {
        st->it_pending = 0;
        return;
    }
    // coefficients of iteration k for the next pass G, as if its stop test does not end the stage
    st->it_mu = st->mu;
    st->it_L1 = st->Lmu1;
    st->it_apk = 0.5 * ((double)k + 1.0);
    st->it_tau = 2.0 / ((double)k + 3.0);
    st->it_cur = st->cur;
    st->it_wcur = st->wcur;
    st->it_sstart = st->stage_start;
    st->it_pending = 1;
}

__global__ __launch_bounds__(OCL_NT) void k_nesta_S2(const SliceDesc* __restrict__ descs, int nslices) {
    __shared__ double red[64];
    if ((int)blockIdx.x >= nslices) return;
    nesta_gram_step(descs[blockIdx.x], red);
}

// Pass G of the Gram form over a launch table (PHU > 0: every slice has H = 16 PHU; 0: any four-step H).
template <int PHU>
__global__ __launch_bounds__(OCL_NT) void k_nesta_G_t(const SliceDesc* __restrict__ descs,
                                                      const int* __restrict__ wg_off, int nslices, int mode) {
    extern __shared__ float2 lds[];
    __shared__ double red[64];
    const int si = find_slice(wg_off, nslices, blockIdx.x);
    const SliceDesc& d = descs[si];
    if (mode == 0 && !d.st->it_pending) return;
    const int local = blockIdx.x - wg_off[si];
    switch (PHU > 0 ? PHU : d.fastH) {
        case 2: nesta_G_fast<2>(d, local, mode, lds, red); break;
        case 4: nesta_G_fast<4>(d, local, mode, lds, red); break;
        case 6: nesta_G_fast<6>(d, local, mode, lds, red); break;
        case 8: nesta_G_fast<8>(d, local, mode, lds, red); break;
        case 10: nesta_G_fast<10>(d, local, mode, lds, red); break;
        case 12: nesta_G_fast<12>(d, local, mode, lds, red); break;
        case 14: nesta_G_fast<14>(d, local, mode, lds, red); break;
        case 16: nesta_G_fast<16>(d, local, mode, lds, red); break;
        case 18: nesta_G_fast<18>(d, local, mode, lds, red); break;
        default: nesta_G_fast<20>(d, local, mode, lds, red); break;
    }
}

// One workgroup per slice, between pass B(k) and pass CA(k).  Kernel boundaries order the partial sums written by
// B before this read and the coefficients written here before CA reads them, so no grid-wide fences are needed.
__global__ __launch_bounds__(OCL_NT) void k_nesta_S(const SliceDesc* __restrict__ descs, int nslices) {
    __shared__ double red[64];
    if ((int)blockIdx.x >= nslices) return;
    nesta_scalar_step(descs[blockIdx.x], red);
}

// ------------------------------------------------------------------------------------------------------------------
// pass B.  Workgroups [0, wg_B) of a slice handle k-space rows; [wg_B, wg_B + wg_U) reduce the temporal-TV terms
// of x_k (mode 0 only).  mode 0: NESTA residual + scalar step.  mode 1: setup (V = IFFT_W(y row)).
// ------------------------------------------------------------------------------------------------------------------
template <int PWU>
__global__ __launch_bounds__(OCL_NT) void k_nesta_B_t(const SliceDesc* __restrict__ descs,
                                                      const int* __restrict__ wg_off, int nslices, int mode) {
    extern __shared__ float2 lds[];
    __shared__ double red[64];
    const int si = find_slice(wg_off, nslices, blockIdx.x);
    const SliceDesc& d = descs[si];
    NState* st = d.st;
    if (mode == 0 && !st->active) return;
    const int local = blockIdx.x - wg_off[si];
    const int C = d.C, W = d.W, H = d.H, nmax = d.nmax, RB = d.rows_per_wg;
    if (local >= d.wg_B) {
        // ---- temporal-TV terms of x_k: sum |u|^2 and sum Re(conj(u) Ux), u = Ux / max(mu, |Ux|) ----
        if (mode != 0) return;
        nesta_U_task(d, local - d.wg_B, red);
    } else if (PWU > 0 || d.fastW) {
        double acc[1];
        switch (PWU > 0 ? PWU : d.fastW) {
            case 2: acc[0] = nesta_B_fast<2>(d, local, mode, lds); break;
            case 4: acc[0] = nesta_B_fast<4>(d, local, mode, lds); break;
            case 6: acc[0] = nesta_B_fast<6>(d, local, mode, lds); break;
            case 8: acc[0] = nesta_B_fast<8>(d, local, mode, lds); break;
            case 10: acc[0] = nesta_B_fast<10>(d, local, mode, lds); break;
            case 12: acc[0] = nesta_B_fast<12>(d, local, mode, lds); break;
            case 14: acc[0] = nesta_B_fast<14>(d, local, mode, lds); break;
            case 16: acc[0] = nesta_B_fast<16>(d, local, mode, lds); break;
            case 18: acc[0] = nesta_B_fast<18>(d, local, mode, lds); break;
            default: acc[0] = nesta_B_fast<20>(d, local, mode, lds); break;
        }
        if (mode != 0) return;
        block_sum<1>(acc, red);
        if (threadIdx.x == 0) d.partB[local] = acc[0];
    } else {
        const int g = local % d.nrow_groups;
        const int fc = local / d.nrow_groups;
        const int f = fc / C, c = fc - f * C;
        const int nr = d.nrows[f];
        const int j0 = g * RB;
        double acc[1] = {0.0};
        if (j0 < nr) {
            const int ld = W + 1;
            float2* a = lds;
            float2* b = a + RB * ld;
            float2* tw = b + RB * ld;
            const float invW = 1.0f / (float)W;
            const int16_t* rows = d.rows + (long)f * nmax;
            float2* kb = d.kbuf + ((long)f * C + c) * nmax * W;
            const float2* yfc = d.y + ((long)f * C + c) * H * W;
            load_twiddles(tw, d.pw);
            float2* res;
            if (mode == 0) {
                for (int e = threadIdx.x; e < RB * W; e += blockDim.x) {
                    const int r = fdiv(e, invW), i = e - r * W;
                    a[r * ld + i] = (j0 + r < nr) ? kb[(long)(j0 + r) * W + i] : make_float2(0.f, 0.f);
                }
                __syncthreads();
                res = lds_fft(a, b, tw, d.pw, RB, ld, 1, false);
                const uint8_t* Mf = d.M + (long)f * H * W;
                const float s = d.inv_sqrtN;
                for (int e = threadIdx.x; e < RB * W; e += blockDim.x) {
                    const int r = fdiv(e, invW), i = e - r * W;
                    float2 rv = make_float2(0.f, 0.f);
                    if (j0 + r < nr) {
                        const int ky = rows[j0 + r];
                        const float2 yv = yfc[(long)ky * W + i];
                        const float2 K = res[r * ld + i];
                        rv = Mf[(long)ky * W + i] ? make_float2(K.x * s - yv.x, K.y * s - yv.y)
                                                  : make_float2(-yv.x, -yv.y);
                        acc[0] += (double)rv.x * rv.x + (double)rv.y * rv.y;
                    }
                    res[r * ld + i] = c_conj(rv);
                }
            } else {
                for (int e = threadIdx.x; e < RB * W; e += blockDim.x) {
                    const int r = fdiv(e, invW), i = e - r * W;
                    float2 v = make_float2(0.f, 0.f);
                    if (j0 + r < nr) v = yfc[(long)rows[j0 + r] * W + i];
                    a[r * ld + i] = c_conj(v);
                }
                res = a;
            }
            __syncthreads();
            float2* other = (res == a) ? b : a;
            float2* out = lds_fft(res, other, tw, d.pw, RB, ld, 1, false);
            for (int e = threadIdx.x; e < RB * W; e += blockDim.x) {
                const int r = fdiv(e, invW), i = e - r * W;
                if (j0 + r < nr) kb[(long)(j0 + r) * W + i] = c_conj(out[r * ld + i]);
            }
        }
        if (mode != 0) return;
        block_sum<1>(acc, red);
        if (threadIdx.x == 0) d.partB[local] = acc[0];
    }
}

// ------------------------------------------------------------------------------------------------------------------
// pass CA.  mode 0: NESTA iteration k with the coefficients prepared by B(k), then the forward columns of x_{k+1}.
// mode 1: setup (x_ref = A^H y, forward columns of iteration 0).  mode 2: setup from a given x_ref in x0 (motion
// compensation starts from the previous reconstruction): forward columns only.
// ------------------------------------------------------------------------------------------------------------------
template <int MAXE, bool MC = false>
__device__ __forceinline__ void nesta_CA_generic(const SliceDesc& d, int local, int mode, float2* lds) {
    const int tile = local % d.ncol_tiles;
    const int f = local / d.ncol_tiles;
    const int F = d.F, C = d.C, H = d.H, W = d.W, TW = d.tile_cols, nmax = d.nmax;
    const long N = d.N;
    const int w0 = tile * TW;
    const int ld = TW + 1;
    float2* a = lds;
    float2* b = a + H * ld;
    float2* tw = b + H * ld;
    const int nr = d.nrows[f];
    const int16_t* rows = d.rows + (long)f * nmax;
    const int nel = H * TW;
    const int lsh = __ffs(TW) - 1;   // TW is a power of two
    load_twiddles(tw, d.ph);

    float2 acc[MAXE];
#pragma unroll
    for (int i = 0; i < MAXE; ++i) acc[i] = make_float2(0.f, 0.f);

    // ---- A^H r: inverse column DFTs of the sampled rows, conj(S) multiply, coil sum (not in mode 2) ----
    const int Cinv = (MC && mode == 2) ? 0 : C;
    for (int c = 0; c < Cinv; ++c) {
        for (int e = threadIdx.x; e < nel; e += blockDim.x) a[(e >> lsh) * ld + (e & (TW - 1))] = make_float2(0.f, 0.f);
        __syncthreads();
        const float2* kb = d.kbuf + ((long)f * C + c) * nmax * W;
        for (int e = threadIdx.x; e < nr * TW; e += blockDim.x) {
            const int j = e >> lsh, wl = e & (TW - 1);
            if (w0 + wl < W) a[rows[j] * ld + wl] = c_conj(kb[(long)j * W + w0 + wl]);
        }
        __syncthreads();
        float2* res = lds_fft(a, b, tw, d.ph, TW, 1, ld, true);
        const float2* Sc = d.S + (long)c * N;
#pragma unroll
        for (int i = 0; i < MAXE; ++i) {
            const int e = threadIdx.x + i * OCL_NT;
            if (e < nel) {
                const int h = e >> lsh, wl = e & (TW - 1);
                if (w0 + wl < W) acc[i] = c_add(acc[i], c_mulconj(c_conj(res[h * ld + wl]), Sc[(long)h * W + w0 + wl]));
            }
        }
        __syncthreads();
    }

    // ---- NESTA update (NestaUp.cpp:669-875 fused) ----
    float2 xn[MAXE];
    const float s = d.inv_sqrtN;
    if (mode == 1) {
#pragma unroll
        for (int i = 0; i < MAXE; ++i) {
            const int e = threadIdx.x + i * OCL_NT;
            xn[i] = c_scale(acc[i], s);
            if (e < nel) {
                const int h = e >> lsh, wl = e & (TW - 1);
                if (w0 + wl < W) d.x0[(long)f * N + (long)h * W + w0 + wl] = xn[i];
            }
        }
    } else if (MC && mode == 2) {                          // x_ref given in x0: forward columns only
#pragma unroll
        for (int i = 0; i < MAXE; ++i) {
            const int e = threadIdx.x + i * OCL_NT;
            xn[i] = make_float2(0.f, 0.f);
            if (e < nel) {
                const int h = e >> lsh, wl = e & (TW - 1);
                if (w0 + wl < W) xn[i] = d.x0[(long)f * N + (long)h * W + w0 + wl];
            }
        }
    } else {
        const UpdRow u = upd_row(d, f);
#pragma unroll
        for (int i = 0; i < MAXE; ++i) {
            const int e = threadIdx.x + i * OCL_NT;
            xn[i] = make_float2(0.f, 0.f);
            if (e < nel) {
                const int h = e >> lsh, wl = e & (TW - 1);
                if (w0 + wl < W) xn[i] = nesta_update_at<MC>(u, (unsigned)(h * W + w0 + wl), c_scale(acc[i], s));
            }
        }
    }

    // ---- forward column DFTs of S_c x_{k+1}, keeping only the sampled rows ----
    for (int c = 0; c < C; ++c) {
        const float2* Sc = d.S + (long)c * N;
#pragma unroll
        for (int i = 0; i < MAXE; ++i) {
            const int e = threadIdx.x + i * OCL_NT;
            if (e < nel) {
                const int h = e >> lsh, wl = e & (TW - 1);
                a[h * ld + wl] = (w0 + wl < W) ? c_mul(xn[i], Sc[(long)h * W + w0 + wl]) : make_float2(0.f, 0.f);
            }
        }
        __syncthreads();
        float2* res = lds_fft(a, b, tw, d.ph, TW, 1, ld, true);
        float2* kb = d.kbuf + ((long)f * C + c) * nmax * W;
        for (int e = threadIdx.x; e < nr * TW; e += blockDim.x) {
            const int j = e >> lsh, wl = e & (TW - 1);
            if (w0 + wl < W) kb[(long)j * W + w0 + wl] = res[rows[j] * ld + wl];
        }
        __syncthreads();
    }

}

// STG: 0 / 1 every slice of the table uses the direct / prefetching four-step column pass, 2 either (per slice)
template <int P, int STG, bool MC>
__device__ __forceinline__ void nesta_CA_fast_any(const SliceDesc& d, int local, int mode, float2* lds) {
    if constexpr (STG == 0) nesta_CA_fast<P, false, MC>(d, local, mode, lds);
    else if constexpr (STG == 1) nesta_CA_fast<P, true, MC>(d, local, mode, lds);
    else if (d.ca_stage) nesta_CA_fast<P, true, MC>(d, local, mode, lds);
    else nesta_CA_fast<P, false, MC>(d, local, mode, lds);
}

template <int PHU, int MAXE, int STG, bool MC>
__device__ __forceinline__ void nesta_CA_body(const SliceDesc* __restrict__ descs, const int* __restrict__ wg_off,
                                              int nslices, int mode) {
    extern __shared__ float2 lds[];
    const int si = find_slice(wg_off, nslices, blockIdx.x);
    const SliceDesc& d = descs[si];
    NState* st = d.st;
    if (mode == 0 && !st->it_pending) return;
    const int local = blockIdx.x - wg_off[si];
    if constexpr (PHU > 0) {
        nesta_CA_fast_any<PHU, STG, MC>(d, local, mode, lds);
    } else if (d.fastH) {
        switch (d.fastH) {
            case 2: nesta_CA_fast_any<2, STG, MC>(d, local, mode, lds); break;
            case 4: nesta_CA_fast_any<4, STG, MC>(d, local, mode, lds); break;
            case 6: nesta_CA_fast_any<6, STG, MC>(d, local, mode, lds); break;
            case 8: nesta_CA_fast_any<8, STG, MC>(d, local, mode, lds); break;
            case 10: nesta_CA_fast_any<10, STG, MC>(d, local, mode, lds); break;
            case 12: nesta_CA_fast_any<12, STG, MC>(d, local, mode, lds); break;
            case 14: nesta_CA_fast_any<14, STG, MC>(d, local, mode, lds); break;
            case 16: nesta_CA_fast_any<16, STG, MC>(d, local, mode, lds); break;
            case 18: nesta_CA_fast_any<18, STG, MC>(d, local, mode, lds); break;
            default: nesta_CA_fast_any<20, STG, MC>(d, local, mode, lds); break;
        }
    } else {
        nesta_CA_generic<MAXE, MC>(d, local, mode, lds);
    }
}

template <int PHU, int MAXE, int STG>
__global__ __launch_bounds__(OCL_NT) void k_nesta_CA_t(const SliceDesc* __restrict__ descs,
                                                       const int* __restrict__ wg_off, int nslices, int mode) {
    nesta_CA_body<PHU, MAXE, STG, false>(descs, wg_off, nslices, mode);
}

// The same pass for tables with motion compensation (update gradient from mcg, mode 2 setup); STG is per slice.
template <int PHU, int MAXE>
__global__ __launch_bounds__(OCL_NT) void k_nesta_CA_mc_t(const SliceDesc* __restrict__ descs,
                                                          const int* __restrict__ wg_off, int nslices, int mode) {
    nesta_CA_body<PHU, MAXE, 2, true>(descs, wg_off, nslices, mode);
}

// ---------------------------------------------------------------------------------------------------------------------
// host-side launch helpers (kernel templates are instantiated in this translation unit)
// ---------------------------------------------------------------------------------------------------------------------
typedef void (*NestaKernel)(const SliceDesc*, const int*, int, int);

static NestaKernel pick_B(int PW) {
    switch (PW) {
        case 8: return k_nesta_B_t<8>;
        case 10: return k_nesta_B_t<10>;
        case 12: return k_nesta_B_t<12>;
        case 14: return k_nesta_B_t<14>;
        case 16: return k_nesta_B_t<16>;
        case 18: return k_nesta_B_t<18>;
        case 20: return k_nesta_B_t<20>;
        default: return k_nesta_B_t<0>;
    }
}

static NestaKernel pick_CA_mc(int PH, int maxe) {
    switch (PH) {
        case 8: return k_nesta_CA_mc_t<8, 4>;
        case 10: return k_nesta_CA_mc_t<10, 4>;
        case 12: return k_nesta_CA_mc_t<12, 4>;
        case 14: return k_nesta_CA_mc_t<14, 4>;
        case 16: return k_nesta_CA_mc_t<16, 4>;
        case 18: return k_nesta_CA_mc_t<18, 4>;
        case 20: return k_nesta_CA_mc_t<20, 4>;
        default: break;
    }
    return maxe <= 4 ? k_nesta_CA_mc_t<0, 4> : maxe <= 8 ? k_nesta_CA_mc_t<0, 8> : maxe <= 16 ? k_nesta_CA_mc_t<0, 16>
                                                                                     : k_nesta_CA_mc_t<0, 32>;
}

static NestaKernel pick_CA(int PH, int maxe, int stg) {
#define OCL_CA_P(P)                                                                                   \
    case P: return stg == 0 ? k_nesta_CA_t<P, 4, 0> : stg == 1 ? k_nesta_CA_t<P, 4, 1> : k_nesta_CA_t<P, 4, 2>;
    switch (PH) {
        OCL_CA_P(8) OCL_CA_P(10) OCL_CA_P(12) OCL_CA_P(14) OCL_CA_P(16) OCL_CA_P(18) OCL_CA_P(20)
        default: break;
    }
#undef OCL_CA_P
    return maxe <= 4 ? k_nesta_CA_t<0, 4, 2> : maxe <= 8 ? k_nesta_CA_t<0, 8, 2> : maxe <= 16 ? k_nesta_CA_t<0, 16, 2>
                                                                                         : k_nesta_CA_t<0, 32, 2>;
}

hipError_t launch_nesta_B(const NestaLaunch& L, hipStream_t stream, int mode) {
    if (L.prec) return launch_nesta_B_mfma(L, stream, mode);
    hipLaunchKernelGGL(pick_B(L.PW), dim3(L.totB), dim3(OCL_NT), L.ldsB, stream, L.dd, L.offB, L.ns, mode);
    return hipGetLastError();
}

static NestaKernel pick_G(int PH) {
    switch (PH) {
        case 8: return k_nesta_G_t<8>;
        case 10: return k_nesta_G_t<10>;
        case 12: return k_nesta_G_t<12>;
        case 14: return k_nesta_G_t<14>;
        case 16: return k_nesta_G_t<16>;
        case 18: return k_nesta_G_t<18>;
        case 20: return k_nesta_G_t<20>;
        default: return k_nesta_G_t<0>;
    }
}

hipError_t launch_nesta_G(const NestaLaunch& L, hipStream_t stream, int mode) {
    if (L.prec) return launch_nesta_G_mfma(L, stream, mode);
    hipLaunchKernelGGL(pick_G(L.PH), dim3(L.totCA), dim3(OCL_NT), L.ldsCA, stream, L.dd, L.offCA, L.ns, mode);
    return hipGetLastError();
}

hipError_t launch_nesta_S2(const NestaLaunch& L, hipStream_t stream) {
    hipLaunchKernelGGL(k_nesta_S2, dim3(L.ns), dim3(OCL_NT), 0, stream, L.dd, L.ns);
    return hipGetLastError();
}

hipError_t launch_nesta_S(const NestaLaunch& L, hipStream_t stream) {
    hipLaunchKernelGGL(k_nesta_S, dim3(L.ns), dim3(OCL_NT), 0, stream, L.dd, L.ns);
    return hipGetLastError();
}

hipError_t launch_nesta_CA(const NestaLaunch& L, hipStream_t stream, int mode) {
    if (L.prec) return launch_nesta_CA_mfma(L, stream, mode);
    hipLaunchKernelGGL(L.mc ? pick_CA_mc(L.PH, L.maxe) : pick_CA(L.PH, L.maxe, L.stage), dim3(L.totCA), dim3(OCL_NT),
                       L.ldsCA, stream, L.dd, L.offCA, L.ns, mode);
    return hipGetLastError();
}

void nesta_kernel_attributes(int lds_bytes) {
    const int ps[] = {0, 8, 10, 12, 14, 16, 18, 20};
    for (int p : ps) {
        (void)hipFuncSetAttribute((const void*)pick_B(p), hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
        (void)hipFuncSetAttribute((const void*)pick_G(p), hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
        for (int e : {4, 8, 16, 32}) {
            for (int g : {0, 1, 2})
                (void)hipFuncSetAttribute((const void*)pick_CA(p, e, g), hipFuncAttributeMaxDynamicSharedMemorySize,
                                          lds_bytes);
            (void)hipFuncSetAttribute((const void*)pick_CA_mc(p, e), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      lds_bytes);
        }
    }
    (void)hipGetLastError();
}

// Setup maxima (NestaUp.cpp:521-568): per workgroup max |x_ref| and max |U x_ref| over a chunk of x_ref = x0.
__global__ void k_setup_max(const SliceDesc* __restrict__ descs, const int* __restrict__ wg_off, int nslices,
                            double* unused) {
    __shared__ double red[32];
    const int si = find_slice(wg_off, nslices, blockIdx.x);
    const SliceDesc& d = descs[si];
    const int local = blockIdx.x - wg_off[si];
    const long total = (long)d.F * d.N;
    const long chunk = (total + d.wg_CA - 1) / d.wg_CA;
    const long lo = (long)local * chunk, hi = min(total, lo + chunk);
    double m0 = -1.0, m1 = -1.0;
    const float2* z = d.mcz ? d.mcz : d.x0;              // U x_ref = tTV(MC(x_ref)) with motion compensation
    for (long i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const int f = (int)(i / d.N);
        const long p = i - (long)f * d.N;
        const int fm = f == 0 ? d.F - 1 : f - 1;
        const float2 x = d.x0[i];
        m0 = fmax(m0, (double)sqrtf(c_abs2(x)));
        m1 = fmax(m1, (double)sqrtf(c_abs2(c_sub(z[i], z[(long)fm * d.N + p]))));
    }
    m0 = block_max(m0, red);
    m1 = block_max(m1, red);
    if (threadIdx.x == 0) { d.maxpart[2 * local] = m0; d.maxpart[2 * local + 1] = m1; }
}

// Copies the final x (the buffer the state points at) to each slice's output.
__global__ void k_finalize(const SliceDesc* __restrict__ descs, int nslices, float2* const* __restrict__ outs) {
    const int si = blockIdx.y;
    if (si >= nslices) return;
    const SliceDesc& d = descs[si];
    const float2* x = d.st->cur ? d.x1 : d.x0;
    float2* o = outs[si];
    const long total = (long)d.F * d.N;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) o[i] = x[i];
}
