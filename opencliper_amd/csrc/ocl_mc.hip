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
#include "ocl_kernels.hpp"

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
        float2 v = make_float2(0.f, 0.f);
        if (p.inside) {
            const float2* fr = in + k * N;
            const float2 i1 = mc_pixel(fr, H, W, p.ph, p.pv), i2 = mc_pixel(fr, H, W, p.ph + 1, p.pv);
            const float2 i3 = mc_pixel(fr, H, W, p.ph, p.pv + 1), i4 = mc_pixel(fr, H, W, p.ph + 1, p.pv + 1);
            const float wl = (float)(p.ph + 1) - p.nhor, wr = p.nhor - (float)p.ph;
            const float wu = (float)(p.pv + 1) - p.nver, wd = p.nver - (float)p.pv;
            const float2 a1 = make_float2(wl * i1.x + wr * i2.x, wl * i1.y + wr * i2.y);
            const float2 a2 = make_float2(wl * i3.x + wr * i4.x, wl * i3.y + wr * i4.y);
            v = make_float2(wu * a1.x + wd * a2.x, wu * a1.y + wd * a2.y);
        }
        out[k * N + p.out] = v;
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
        const float2 i0 = in[k * N + p.out];
        const float Whh = p.nhor - (float)p.ph, Wvv = p.nver - (float)p.pv;
        float* fr = (float*)(out + k * N);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int h = p.ph + (q & 1), v = p.pv + (q >> 1);
            if (h < 0 || h >= W || v < 0 || v >= H) continue;
            const float hw = (q & 1) ? Whh : 1.0f - Whh, vw = (q >> 1) ? Wvv : 1.0f - Wvv;
            const long o = 2 * (h + (long)v * W);
            atomicAdd(fr + o, i0.x * hw * vw);
            atomicAdd(fr + o + 1, i0.y * hw * vw);
        }
    }
}
