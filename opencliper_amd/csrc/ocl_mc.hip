// Motion compensation of the temporal-TV sparsity operator (the argsMC branch of NestaUp::operatorU / operatorUt,
// NestaUp.cpp:201-278): MotionCompensation = copy, then bilinear resampling of the bound-box pixels at deformed
// positions (src/kernels/optimizer.cl:832-867); AdjointMotionCompensation = copy with the box cleared
// (initialize.cl:55-69), then the transpose scatter of the box pixels (optimizer.cl:880-930).
//
// Geometry: box rows r0 .. r0+nr-1 and columns c0 .. c0+nc-1 of every frame.  For box pixel (r, c) of frame k the
// output position is (hor, ver) = mesh x at [r + c H] minus 1 (x: int32 [2][W][H], 1-based), the sample position is
// (nhor, nver) = xn at [c + r W + k H W] (and + H W F) minus 1 (xn: float32 [2][F][H][W], 1-based).  The reference
// leaves positions outside the frame undefined (it reads / writes out of bounds); here neighbours outside the frame
// read as zero and scatter targets outside it are dropped, which agrees with the reference whenever it stays inside.
//
// Memory: one thread per (box pixel, frame) with the column index fastest, so xn, the input neighbourhoods and the
// output rows are read and written along contiguous rows.  The scatter uses float atomics (the reference does too),
// so its summation order, and the last bits of overlapping contributions, vary from run to run.
#include "ocl_nesta_common.hpp"

namespace {

struct McPoint {
    long out;        // output index within the frame, or -1
    int ph, pv;      // top-left neighbour
    float nhor, nver;
    bool inside;     // the sample position is near enough to the frame for any neighbour to exist
};

__device__ __forceinline__ McPoint mc_point(const float* __restrict__ xn, const int* __restrict__ x, int F, int H,
                                            int W, int r, int c, int k) {
    const long N = (long)H * W;
    McPoint p;
    const int hor = x[r + (long)c * H] - 1, ver = x[r + (long)c * H + N] - 1;
    p.out = (hor >= 0 && hor < W && ver >= 0 && ver < H) ? hor + (long)ver * W : -1;
    p.nhor = xn[c + (long)r * W + k * N] - 1.0f;
    p.nver = xn[c + (long)r * W + k * N + N * F] - 1.0f;
    // comparisons are false for NaN, so a NaN position reads as outside
    p.inside = p.nhor > -2.0f && p.nhor < (float)(W + 1) && p.nver > -2.0f && p.nver < (float)(H + 1);
    p.ph = p.inside ? (int)floorf(p.nhor) : 0;
    p.pv = p.inside ? (int)floorf(p.nver) : 0;
    return p;
}

__device__ __forceinline__ float2 mc_pixel(const float2* __restrict__ fr, int H, int W, int h, int v) {
    return (h >= 0 && h < W && v >= 0 && v < H) ? fr[h + (long)v * W] : make_float2(0.f, 0.f);
}

// Bilinear sample of frame fr at the deformed position (nhor, nver) (0-based), with the interpolator's arithmetic
__device__ __forceinline__ float2 mc_bilinear(const float2* __restrict__ fr, int H, int W, float nhor, float nver,
                                              int ph, int pv) {
    const float2 i1 = mc_pixel(fr, H, W, ph, pv), i2 = mc_pixel(fr, H, W, ph + 1, pv);
    const float2 i3 = mc_pixel(fr, H, W, ph, pv + 1), i4 = mc_pixel(fr, H, W, ph + 1, pv + 1);
    const float wl = (float)(ph + 1) - nhor, wr = nhor - (float)ph;
    const float wu = (float)(pv + 1) - nver, wd = nver - (float)pv;
    const float2 a1 = make_float2(wl * i1.x + wr * i2.x, wl * i1.y + wr * i2.y);
    const float2 a2 = make_float2(wl * i3.x + wr * i4.x, wl * i3.y + wr * i4.y);
    return make_float2(wu * a1.x + wd * a2.x, wu * a1.y + wd * a2.y);
}

// Scatter of v with the bilinear weights of (nhor, nver) into frame fr (the adjoint interpolator's arithmetic)
__device__ __forceinline__ void mc_scatter4(float2* __restrict__ fr, int H, int W, float nhor, float nver, int ph,
                                            int pv, float2 v) {
    const float Whh = nhor - (float)ph, Wvv = nver - (float)pv;
    float* f = (float*)fr;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int h = ph + (q & 1), vv = pv + (q >> 1);
        if (h < 0 || h >= W || vv < 0 || vv >= H) continue;
        const float hw = (q & 1) ? Whh : 1.0f - Whh, vw = (q >> 1) ? Wvv : 1.0f - Wvv;
        const long o = 2 * (h + (long)vv * W);
        atomicAdd(f + o, v.x * hw * vw);
        atomicAdd(f + o + 1, v.y * hw * vw);
    }
}

// Sample position of pixel p of frame k from xn (0-based), and whether any neighbour can exist
__device__ __forceinline__ bool mc_position(const float* __restrict__ xn, int F, int H, int W, long N, int k, long p,
                                            float* nhor, float* nver, int* ph, int* pv) {
    *nhor = xn[p + k * N] - 1.0f;
    *nver = xn[p + k * N + N * F] - 1.0f;
    const bool inside = *nhor > -2.0f && *nhor < (float)(W + 1) && *nver > -2.0f && *nver < (float)(H + 1);
    *ph = inside ? (int)floorf(*nhor) : 0;
    *pv = inside ? (int)floorf(*nver) : 0;
    return inside;
}

// D^T u at pixel p of frame f for u = D z / max(mu, |D z|) (periodic in f): u_f - u_{f+1}
__device__ __forceinline__ float2 mc_tv_grad(const float2* __restrict__ z, int F, long N, int f, long p, float mu) {
    const int fm = f == 0 ? F - 1 : f - 1, fp = f == F - 1 ? 0 : f + 1;
    const float2 zf = z[(long)f * N + p], zm = z[(long)fm * N + p], zp = z[(long)fp * N + p];
    const float2 uxf = c_sub(zf, zm), uxp = c_sub(zp, zf);
    const float nf = fmaxf(mu, sqrtf(c_abs2(uxf))), np = fmaxf(mu, sqrtf(c_abs2(uxp)));
    return make_float2(uxf.x / nf - uxp.x / np, uxf.y / nf - uxp.y / np);
}

}  // namespace

// out[box pixels] = bilinear sample of in at the deformed positions (out holds a copy of in elsewhere)
__global__ void k_mc_interp(const float2* __restrict__ in, float2* __restrict__ out, const float* __restrict__ xn,
                            const int* __restrict__ x, int F, int H, int W, int r0, int nr, int c0, int nc) {
    const long per = (long)nr * nc, total = per * F, N = (long)H * W;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int k = (int)(t / per);
        const long b = t - (long)k * per;
        const int i = (int)(b / nc), j = (int)(b - (long)i * nc);
        const McPoint p = mc_point(xn, x, F, H, W, r0 + i, c0 + j, k);
        if (p.out < 0) continue;
        out[k * N + p.out] = p.inside ? mc_bilinear(in + k * N, H, W, p.nhor, p.nver, p.ph, p.pv) : make_float2(0.f, 0.f);
    }
}

// out = in with the box rows x columns of every frame cleared (copy + initZeroRect in one pass)
__global__ void k_mc_copy_clear_box(const float2* __restrict__ in, float2* __restrict__ out, int F, int H, int W,
                                    int r0, int nr, int c0, int nc) {
    const long N = (long)H * W, total = N * F;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const long p = t % N;
        const int r = (int)(p / W), c = (int)(p - (long)r * W);
        const bool box = r >= r0 && r < r0 + nr && c >= c0 && c < c0 + nc;
        out[t] = box ? make_float2(0.f, 0.f) : in[t];
    }
}

// out[neighbours] += weight * in[box pixel]: the transpose of k_mc_interp's resampling
__global__ void k_mc_scatter(const float2* __restrict__ in, float2* __restrict__ out, const float* __restrict__ xn,
                             const int* __restrict__ x, int F, int H, int W, int r0, int nr, int c0, int nc) {
    const long per = (long)nr * nc, total = per * F, N = (long)H * W;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int k = (int)(t / per);
        const long b = t - (long)k * per;
        const int i = (int)(b / nc), j = (int)(b - (long)i * nc);
        const McPoint p = mc_point(xn, x, F, H, W, r0 + i, c0 + j, k);
        if (p.out < 0 || !p.inside) continue;
        mc_scatter4(out + k * N, H, W, p.nhor, p.nver, p.ph, p.pv, in[k * N + p.out]);
    }
}

// ---------------------------------------------------------------------------------------------------------------------
// Passes of the fused batched iteration with motion compensation (identity mesh: box pixel (r, c) writes to (r, c)).
// Each workgroup covers OCL_MC_CHUNK consecutive pixel-frames of one slice (pixel passes) or of its box (scatter).
// ---------------------------------------------------------------------------------------------------------------------
namespace {
__device__ __forceinline__ bool in_box(const SliceDesc& d, int r, int c) {
    return r >= d.mc_r0 && r < d.mc_r0 + d.mc_nr && c >= d.mc_c0 && c < d.mc_c0 + d.mc_nc;
}
}  // namespace

// Z: mcz = MC(x_k) (mode 0, x_k where the state points) or MC(x0) (mode 1, setup)
__global__ __launch_bounds__(OCL_NT) void k_nesta_mc_z(const SliceDesc* __restrict__ descs, const int* __restrict__ off,
                                                       int ns, int mode) {
    const int si = find_slice(off, ns, blockIdx.x);
    const SliceDesc& d = descs[si];
    if (mode == 0 && !d.st->active) return;
    const float2* xk = mode ? d.x0 : (d.st->cur ? d.x1 : d.x0);
    const long N = d.N, total = (long)d.F * N;
    const long base = (long)(blockIdx.x - off[si]) * OCL_MC_CHUNK;
#pragma unroll
    for (int q = 0; q < OCL_MC_CHUNK / OCL_NT; ++q) {
        const long i = base + q * OCL_NT + threadIdx.x;
        if (i >= total) continue;
        const int f = (int)(i / N);
        const long p = i - (long)f * N;
        const int r = (int)(p / d.W), c = (int)(p - (long)r * d.W);
        float2 v = xk[i];
        if (in_box(d, r, c)) {
            float nhor, nver;
            int ph, pv;
            v = mc_position(d.mc_xn, d.F, d.H, d.W, N, f, p, &nhor, &nver, &ph, &pv)
                    ? mc_bilinear(xk + (long)f * N, d.H, d.W, nhor, nver, ph, pv)
                    : make_float2(0.f, 0.f);
        }
        d.mcz[i] = v;
    }
}

// Setup: mc_dmax = max over the box of max(|nhor - c|, |nver - r|) for positions with neighbours in the frame.  The
// host turns it into the gather radius of pass G: box pixel p reaches target q only if |q - p| <= floor(D) + 1 per
// axis (the bilinear footprint of a position within D of p spans p - ceil(D) .. p + floor(D) + 1).
__global__ __launch_bounds__(OCL_NT) void k_nesta_mc_dmax(const SliceDesc* __restrict__ descs,
                                                          const int* __restrict__ off, int ns) {
    __shared__ double red[32];
    const int si = find_slice(off, ns, blockIdx.x);
    const SliceDesc& d = descs[si];
    const long N = d.N, per = (long)d.mc_nr * d.mc_nc, total = per * d.F;
    const long base = (long)(blockIdx.x - off[si]) * OCL_MC_CHUNK;
    double m = 0.0;
#pragma unroll
    for (int q = 0; q < OCL_MC_CHUNK / OCL_NT; ++q) {
        const long i = base + q * OCL_NT + threadIdx.x;
        if (i >= total) continue;
        const int f = (int)(i / per);
        const long b = i - (long)f * per;
        const int ii = (int)(b / d.mc_nc), jj = (int)(b - (long)ii * d.mc_nc);
        const int r = d.mc_r0 + ii, c = d.mc_c0 + jj;
        float nhor, nver;
        int ph, pv;
        if (mc_position(d.mc_xn, d.F, d.H, d.W, N, f, (long)r * d.W + c, &nhor, &nver, &ph, &pv))
            m = fmax(m, (double)fmaxf(fabsf(nhor - (float)c), fabsf(nver - (float)r)));
    }
    m = block_max(m, red);
    if (threadIdx.x == 0) atomicMax(d.mc_dmax, __float_as_uint((float)m));   // non-negative floats order as uints
}

namespace {
// Adjoint motion compensation of one OCL_MC_TR x OCL_MC_TC tile of frame f by gathering (deterministic), identity mesh
// (box pixel (r, c) writes to (r, c)): the box pixels of the tile and its R-halo are staged in LDS with their sample
// positions and source values val(f, p); each target pixel then starts from val outside the box (0 inside: copy +
// initZeroRect of the adjoint) and adds, in row-major order of the sources (the order the adjoint's scatter visits
// them), the bilinear weight of each source's footprint on it times the source's value.  The weights are the tent
// form of the interpolator's (equal up to the rounding of 1 - (floor + 1 - nhor) against nhor - floor).  Sources must move by less
// than R pixels.  outf: the output frame.
template <int R, class Val>
__device__ __forceinline__ void mc_gather_tile(const float* __restrict__ xn, int F, int H, int W, int f, int r0,
                                               int c0, int br0, int br1, int bc0, int bc1, const Val& val,
                                               float2* __restrict__ outf, float2* spos, float2* sval) {
    constexpr int SR = OCL_MC_TR + 2 * R, SC = OCL_MC_TC + 2 * R;
    const long N = (long)H * W;
    const bool near = r0 - R < br1 && r0 + OCL_MC_TR + R > br0 && c0 - R < bc1 && c0 + OCL_MC_TC + R > bc0;
    if (near) {
        constexpr int NS = (SR * SC + OCL_NT - 1) / OCL_NT;   // staged cells per thread
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int e = threadIdx.x + k * OCL_NT;
            const int pr = r0 - R + e / SC, pc = c0 - R + e % SC;
            float2 pos = make_float2(__int_as_float(0x7fc00000), 0.f), v = make_float2(0.f, 0.f);
            if (e < SR * SC && pr >= br0 && pr < br1 && pc >= bc0 && pc < bc1) {
                const long pp = (long)pr * W + pc;
                float nhor, nver;
                int ph, pv;
                if (mc_position(xn, F, H, W, N, f, pp, &nhor, &nver, &ph, &pv)) {
                    pos = make_float2(nhor, nver);
                    v = val(f, pp);
                }
            }
            if (e < SR * SC) {
                spos[e] = pos;
                sval[e] = v;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < OCL_MC_TR * OCL_MC_TC / OCL_NT; ++k) {
        const int e = threadIdx.x + k * OCL_NT;
        const int lr = e / OCL_MC_TC, lc = e % OCL_MC_TC;
        const int r = r0 + lr, c = c0 + lc;
        if (r >= H || c >= W) continue;
        const long p = (long)r * W + c;
        const bool inbox = r >= br0 && r < br1 && c >= bc0 && c < bc1;
        float2 g = inbox ? make_float2(0.f, 0.f) : val(f, p);
        const float cf = (float)c, rf = (float)r;
        if (near) {
#pragma unroll
            for (int dr = -R; dr <= R; ++dr)
#pragma unroll
                for (int dc = -R; dc <= R; ++dc) {
                    // Branch-free tent weights: the bilinear weight of target column c for a sample at nhor is
                    // max(0, 1 - |c - nhor|) (1 - frac at floor(nhor), frac at floor(nhor) + 1, 0 elsewhere); a
                    // non-source (NaN position) gets max(0, NaN) = 0.
                    const int cell = (lr + R + dr) * SC + (lc + R + dc);
                    const float2 pos = spos[cell];
                    const float2 t = sval[cell];
                    const float hw = fmaxf(0.f, 1.0f - fabsf(cf - pos.x));
                    const float vw = fmaxf(0.f, 1.0f - fabsf(rf - pos.y));
                    g.x += t.x * hw * vw;                          // the adjoint scatter's association
                    g.y += t.y * hw * vw;
                }
        }
        outf[p] = g;
    }
}
}  // namespace

// G, gather form (deterministic): workgroup = one tile of one frame, mcg = MC^T(D^T u(mcz)).
template <int R>
__global__ __launch_bounds__(OCL_NT) void k_nesta_mc_ggather(const SliceDesc* __restrict__ descs,
                                                             const int* __restrict__ off, int ns) {
    constexpr int SR = OCL_MC_TR + 2 * R, SC = OCL_MC_TC + 2 * R;
    __shared__ float2 spos[SR * SC];   // (nhor, nver) of a source, nhor = NaN if not a source
    __shared__ float2 sval[SR * SC];   // D^T u of a source
    const int si = find_slice(off, ns, blockIdx.x);
    const SliceDesc& d = descs[si];
    if (!d.st->it_pending) return;
    const float mu = (float)d.st->it_mu;
    const long N = d.N;
    const int ntc = (d.W + OCL_MC_TC - 1) / OCL_MC_TC;
    const int local = blockIdx.x - off[si];
    // frames fastest: consecutive workgroups share a tile across adjacent frames, whose z rows (f-1, f, f+1) then
    // meet in L2
    const int tt = local / d.F, f = local - tt * d.F;
    const int r0 = (tt / ntc) * OCL_MC_TR, c0 = (tt % ntc) * OCL_MC_TC;
    const float2* z = d.mcz;
    const int F = d.F;
    auto val = [=](int ff, long p) { return mc_tv_grad(z, F, N, ff, p, mu); };
    mc_gather_tile<R>(d.mc_xn, d.F, d.H, d.W, f, r0, c0, d.mc_r0, d.mc_r0 + d.mc_nr, d.mc_c0, d.mc_c0 + d.mc_nc,
                      val, d.mcg + (long)f * N, spos, sval);
}

// out = MC^T(tTV^T y) for the identity mesh, by gathering: the adjoint sparsity operator of myNormest with motion
// compensation (NestaUp.cpp:238-278), with the temporal-TV adjoint (y_{f+1} - y_f, periodic) evaluated at each
// source.  Workgroup = one tile of one frame, frames fastest.
template <int R>
__global__ __launch_bounds__(OCL_NT) void k_mc_adj_tv_gather(const float2* __restrict__ y, float2* __restrict__ out,
                                                             const float* __restrict__ xn, int F, int H, int W,
                                                             int br0, int br1, int bc0, int bc1) {
    constexpr int SR = OCL_MC_TR + 2 * R, SC = OCL_MC_TC + 2 * R;
    __shared__ float2 spos[SR * SC];
    __shared__ float2 sval[SR * SC];
    const long N = (long)H * W;
    const int ntc = (W + OCL_MC_TC - 1) / OCL_MC_TC;
    const int tt = blockIdx.x / F, f = blockIdx.x - tt * F;
    const int r0 = (tt / ntc) * OCL_MC_TR, c0 = (tt % ntc) * OCL_MC_TC;
    auto val = [=](int ff, long p) {
        const int fp = ff == F - 1 ? 0 : ff + 1;
        return c_sub(y[(long)fp * N + p], y[(long)ff * N + p]);
    };
    mc_gather_tile<R>(xn, F, H, W, f, r0, c0, br0, br1, bc0, bc1, val, out + (long)f * N, spos, sval);
}

// out = tTV(MC(x)) for the identity mesh in one pass: (MC x)_f - (MC x)_{f-1} (periodic), with MC x = x outside the
// box and the bilinear resample inside it (the per-pixel arithmetic of k_mc_interp and k_ttv, so the values are those
// of the two passes).  The sparsity operator of myNormest with motion compensation (NestaUp.cpp:201-237).
__global__ void k_mc_tv_fwd(const float2* __restrict__ x, float2* __restrict__ out, const float* __restrict__ xn,
                            int F, int H, int W, int br0, int br1, int bc0, int bc1) {
    const long N = (long)H * W, total = (long)F * N;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int f = (int)(i / N);
        const long p = i - (long)f * N;
        const int r = (int)(p / W), c = (int)(p - (long)r * W);
        const bool inbox = r >= br0 && r < br1 && c >= bc0 && c < bc1;
        auto mcv = [&](int ff) {
            if (!inbox) return x[(long)ff * N + p];
            float nhor, nver;
            int ph, pv;
            return mc_position(xn, F, H, W, N, ff, p, &nhor, &nver, &ph, &pv)
                       ? mc_bilinear(x + (long)ff * N, H, W, nhor, nver, ph, pv)
                       : make_float2(0.f, 0.f);
        };
        out[i] = c_sub(mcv(f), mcv(f == 0 ? F - 1 : f - 1));
    }
}

hipError_t launch_mc_tv_fwd(const float2* x, float2* out, const float* xn, int F, int H, int W, int r0, int nr, int c0,
                            int nc, hipStream_t stream) {
    const long blocks = ((long)F * H * W + OCL_NT - 1) / OCL_NT;          // grid-stride beyond 8192 workgroups
    hipLaunchKernelGGL(k_mc_tv_fwd, dim3((unsigned)(blocks < 8192 ? blocks : 8192)), dim3(OCL_NT), 0, stream, x, out,
                       xn, F, H, W, r0, r0 + nr, c0, c0 + nc);
    return hipGetLastError();
}

hipError_t launch_mc_adj_tv_gather(const float2* y, float2* out, const float* xn, int F, int H, int W, int r0, int nr,
                                   int c0, int nc, int R, hipStream_t stream) {
    const int tiles = ((H + OCL_MC_TR - 1) / OCL_MC_TR) * ((W + OCL_MC_TC - 1) / OCL_MC_TC);
    void (*k)(const float2*, float2*, const float*, int, int, int, int, int, int, int) =
        R == 1 ? k_mc_adj_tv_gather<1> : R == 2 ? k_mc_adj_tv_gather<2> : R == 3 ? k_mc_adj_tv_gather<3>
                                                                                 : k_mc_adj_tv_gather<4>;
    hipLaunchKernelGGL(k, dim3((unsigned)(tiles * F)), dim3(OCL_NT), 0, stream, y, out, xn, F, H, W, r0, r0 + nr, c0,
                       c0 + nc);
    return hipGetLastError();
}

// G, scatter form (displacements beyond the gather radius), first half: mcg = D^T u(mcz) outside the box, 0 inside
__global__ __launch_bounds__(OCL_NT) void k_nesta_mc_gclear(const SliceDesc* __restrict__ descs,
                                                            const int* __restrict__ off, int ns) {
    const int si = find_slice(off, ns, blockIdx.x);
    const SliceDesc& d = descs[si];
    if (!d.st->it_pending) return;
    const float mu = (float)d.st->it_mu;
    const long N = d.N, total = (long)d.F * N;
    const long base = (long)(blockIdx.x - off[si]) * OCL_MC_CHUNK;
#pragma unroll
    for (int q = 0; q < OCL_MC_CHUNK / OCL_NT; ++q) {
        const long i = base + q * OCL_NT + threadIdx.x;
        if (i >= total) continue;
        const int f = (int)(i / N);
        const long p = i - (long)f * N;
        const int r = (int)(p / d.W), c = (int)(p - (long)r * d.W);
        d.mcg[i] = in_box(d, r, c) ? make_float2(0.f, 0.f) : mc_tv_grad(d.mcz, d.F, N, f, p, mu);
    }
}

// G, scatter form, second half: the box pixels' D^T u(mcz) added into mcg with float atomics (order varies)
__global__ __launch_bounds__(OCL_NT) void k_nesta_mc_gscatter(const SliceDesc* __restrict__ descs,
                                                              const int* __restrict__ off, int ns) {
    const int si = find_slice(off, ns, blockIdx.x);
    const SliceDesc& d = descs[si];
    if (!d.st->it_pending) return;
    const float mu = (float)d.st->it_mu;
    const long N = d.N, per = (long)d.mc_nr * d.mc_nc, total = per * d.F;
    const long base = (long)(blockIdx.x - off[si]) * OCL_MC_CHUNK;
#pragma unroll
    for (int q = 0; q < OCL_MC_CHUNK / OCL_NT; ++q) {
        const long i = base + q * OCL_NT + threadIdx.x;
        if (i >= total) continue;
        const int f = (int)(i / per);
        const long b = i - (long)f * per;
        const int ii = (int)(b / d.mc_nc), jj = (int)(b - (long)ii * d.mc_nc);
        const long p = (long)(d.mc_r0 + ii) * d.W + d.mc_c0 + jj;
        float nhor, nver;
        int ph, pv;
        if (!mc_position(d.mc_xn, d.F, d.H, d.W, N, f, p, &nhor, &nver, &ph, &pv)) continue;
        mc_scatter4(d.mcg + (long)f * N, d.H, d.W, nhor, nver, ph, pv, mc_tv_grad(d.mcz, d.F, N, f, p, mu));
    }
}

hipError_t launch_nesta_mc_dmax(const NestaLaunch& L, hipStream_t stream) {
    if (!L.mc || !L.totX) return hipSuccess;
    hipLaunchKernelGGL(k_nesta_mc_dmax, dim3(L.totX), dim3(OCL_NT), 0, stream, L.dd, L.offX, L.ns);
    return hipGetLastError();
}

hipError_t launch_nesta_mc_z(const NestaLaunch& L, hipStream_t stream, int mode) {
    if (!L.mc || !L.totP) return hipSuccess;
    hipLaunchKernelGGL(k_nesta_mc_z, dim3(L.totP), dim3(OCL_NT), 0, stream, L.dd, L.offP, L.ns, mode);
    return hipGetLastError();
}

hipError_t launch_nesta_mc_g(const NestaLaunch& L, hipStream_t stream) {
    if (!L.mc || !L.totP) return hipSuccess;
    if (L.mcR > 0) {
        void (*k)(const SliceDesc*, const int*, int) =
            L.mcR == 1 ? k_nesta_mc_ggather<1> : L.mcR == 2 ? k_nesta_mc_ggather<2> : L.mcR == 3
                ? k_nesta_mc_ggather<3> : k_nesta_mc_ggather<4>;
        hipLaunchKernelGGL(k, dim3(L.totG), dim3(OCL_NT), 0, stream, L.dd, L.offG, L.ns);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_nesta_mc_gclear, dim3(L.totP), dim3(OCL_NT), 0, stream, L.dd, L.offP, L.ns);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !L.totX) return e;
    hipLaunchKernelGGL(k_nesta_mc_gscatter, dim3(L.totX), dim3(OCL_NT), 0, stream, L.dd, L.offX, L.ns);
    return hipGetLastError();
}
