// gfx950 kernels behind the individual OpenCLIPER Processes of the NESTA-TV path and the BLAS-1 calls NestaUp makes.
// Each kernel is grid-stride over the flattened HBM layout (W fastest), so every access is a coalesced 8-byte
// complex load or store.  Launch wrappers live in ocl_capi.cpp.
#include "ocl_device.hpp"
#include "ocl_kernels.hpp"
#include "ocl_rfft.hpp"

// TemporalTV (src/kernels/temporalTV.cl:49-107).  dir=1: out_f = in_f - in_{f-1 mod F};
// dir=0 (the reference's "adjoint", i.e. -D^T): out_f = in_{f+1 mod F} - in_f.
__global__ void k_ttv(const float2* __restrict__ in, float2* __restrict__ out, int F, long N, int dir) {
    const long total = (long)F * N;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int f = (int)(i / N);
        const long p = i - (long)f * N;
        if (dir) {
            const int g = f == 0 ? F - 1 : f - 1;
            out[i] = c_sub(in[i], in[(long)g * N + p]);
        } else {
            const int g = f == F - 1 ? 0 : f + 1;
            out[i] = c_sub(in[(long)g * N + p], in[i]);
        }
    }
}

// VectorNormalization (src/kernels/vectorNormalization.cl:42-56): out = in / max(mu, |in|)
__global__ void k_vecnorm(const float2* __restrict__ in, float2* __restrict__ out, long n, float mu) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float2 v = in[i];
        const float fac = fmaxf(mu, sqrtf(c_abs2(v)));
        out[i] = make_float2(v.x / fac, v.y / fac);
    }
}

// ComplexElementProd (src/kernels/complexElementProd.cl:42-76): out[f][c] = in[f][(c)] * (conj?) S[c].
// in_coils = 0: the input is an image [F][N] broadcast over the coils.
__global__ void k_cep(const float2* __restrict__ in, const float2* __restrict__ S, float2* __restrict__ out, int F,
                      int C, long N, int in_coils, int conj) {
    const long total = (long)F * C * N;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const long fc = i / N;
        const long p = i - fc * N;
        const int f = (int)(fc / C), c = (int)(fc - (long)f * C);
        const float2 a = in_coils ? in[i] : in[(long)f * N + p];
        const float2 s = S[(long)c * N + p];
        out[i] = conj ? c_mulconj(a, s) : c_mul(a, s);
    }
}

// ApplyMask (src/kernels/applyMask.cl:5-40): zero k-space samples whose mask byte is 0
__global__ void k_apply_mask(float2* __restrict__ d, const uint8_t* __restrict__ M, int F, int C, long N) {
    const long total = (long)F * C * N;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const long fc = i / N;
        const long p = i - fc * N;
        const int f = (int)(fc / C);
        if (M[(long)f * N + p] == 0) d[i] = make_float2(0.f, 0.f);
    }
}

// XImageSum (src/kernels/xImageSum.cl:42-61): x_f = sum_c in_{f,c}, coils summed in order
__global__ void k_ximage_sum(const float2* __restrict__ in, float2* __restrict__ out, int F, int C, long N) {
    const long total = (long)F * N;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int f = (int)(i / N);
        const long p = i - (long)f * N;
        float2 acc = make_float2(0.f, 0.f);
        for (int c = 0; c < C; ++c) acc = c_add(acc, in[((long)f * C + c) * N + p]);
        out[i] = acc;
    }
}

// ComplexElementProd with conj(S) in place followed by XImageSum (SimpleMRIRecon.cpp:59-101) in one pass:
// y[f][c] <- y[f][c] conj(S_c), x[f] = sum_c y[f][c] (coils summed in order, as xImageSum.cl:42-61 does).
__global__ void k_cep_sum(float2* __restrict__ y, const float2* __restrict__ S, float2* __restrict__ x, int F, int C,
                          long N) {
    const long total = (long)F * N;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int f = (int)(i / N);
        const long p = i - (long)f * N;
        float2 acc = make_float2(0.f, 0.f);
        for (int c = 0; c < C; ++c) {
            const long k = ((long)f * C + c) * N + p;
            const float2 v = c_mulconj(y[k], S[(long)c * N + p]);
            y[k] = v;
            acc = c_add(acc, v);
        }
        x[i] = acc;
    }
}

// RSoS (src/kernels/rss.cl): out_f = (sqrt(sum_c |in_{f,c}|^2), 0), the sum taken in coil order.  The reference
// writes only the real part; the imaginary part is set to 0 here.
__global__ void k_rsos(const float2* __restrict__ in, float2* __restrict__ out, int F, int C, long N) {
    const long total = (long)F * N;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int f = (int)(i / N);
        const long p = i - (long)f * N;
        float s = 0.f;
        for (int c = 0; c < C; ++c) s += c_abs2(in[((long)f * C + c) * N + p]);
        out[i] = make_float2(sqrtf(s), 0.f);
    }
}

// ComplexAbs / ComplexAbsPow2 (src/kernels/complexAbs.cl:43-70, complexAbsPow2.cl:43-71): out = (|z| or |z|^2, 0)
__global__ void k_abs(const float2* __restrict__ in, float2* __restrict__ out, long n, int pow2) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float a2 = c_abs2(in[i]);
        out[i] = make_float2(pow2 ? a2 : sqrtf(a2), 0.f);
    }
}

// CopyDataGPU (initialize.cl:76-80) / BLAS copy, scal, axpy with a complex alpha
// Byte copy in 4-byte words (both addresses and n are multiples of 4): the small host <-> device transfers through
// page-locked host memory, with either side mapped host memory.
__global__ void k_copy_words(const unsigned* __restrict__ x, unsigned* __restrict__ y, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = x[i];
}

hipError_t launch_copy_words(const void* src, void* dst, size_t bytes, hipStream_t stream) {
    const long n = (long)(bytes / 4);
    if (n == 0) return hipSuccess;
    const long g = (n + OCL_NT - 1) / OCL_NT;
    hipLaunchKernelGGL(k_copy_words, dim3((unsigned)(g < 64 ? g : 64)), dim3(OCL_NT), 0, stream,
                       (const unsigned*)src, (unsigned*)dst, n);
    return hipGetLastError();
}

__global__ void k_copy(const float2* __restrict__ x, float2* __restrict__ y, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = x[i];
}
__global__ void k_scal(float2* __restrict__ x, long n, float ar, float ai) {
    const float2 a = make_float2(ar, ai);
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] = c_mul(a, x[i]);
}
// x *= 1 / ||x|| with the norm read on the device (the factor 1.0f / (float)norm that the host computes otherwise)
__global__ void k_scal_inv_norm(float2* __restrict__ x, long n, const double* __restrict__ nrm) {
    const float2 a = make_float2(1.0f / (float)nrm[0], 0.f);
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] = c_mul(a, x[i]);
}
__global__ void k_axpy(long n, float ar, float ai, const float2* __restrict__ x, float2* __restrict__ y) {
    const float2 a = make_float2(ar, ai);
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = c_add(y[i], c_mul(a, x[i]));
}
__global__ void k_fill(float2* __restrict__ x, long n, float2 v) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] = v;
}
__global__ void k_fill_bytes(uint8_t* __restrict__ x, long n, uint8_t v) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] = v;
}

// Two-stage deterministic reductions over complex vectors.  Stage 1 writes one partial per workgroup
// (partials[op_width * blockIdx.x + q]); stage 2 (one workgroup) folds them in index order.
//   RED_NRM2: sum |x|^2              RED_DOTC: sum conj(x)*y (2 values)   RED_SUM: sum re+im
//   RED_ASUM: sum |re|+|im|          RED_MAXABS: max |x|
__global__ void k_reduce_partial(const float2* __restrict__ x, const float2* __restrict__ y, long n, int op,
                                 double* __restrict__ partials) {
    __shared__ double red[32];
    double v[2] = {0.0, 0.0};
    if (op == RED_MAXABS) v[0] = -1.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float2 a = x[i];
        switch (op) {
            case RED_NRM2: v[0] += (double)a.x * a.x + (double)a.y * a.y; break;
            case RED_DOTC: {
                const float2 b = y[i];
                v[0] += (double)a.x * b.x + (double)a.y * b.y;
                v[1] += (double)a.x * b.y - (double)a.y * b.x;
                break;
            }
            case RED_SUM: v[0] += (double)a.x + (double)a.y; break;
            case RED_ASUM: v[0] += fabs((double)a.x) + fabs((double)a.y); break;
            default: v[0] = fmax(v[0], (double)sqrtf(c_abs2(a))); break;
        }
    }
    if (op == RED_MAXABS) {
        v[0] = block_max(v[0], red);
    } else {
        block_sum<2>(v, red);
    }
    if (threadIdx.x == 0) { partials[2 * blockIdx.x] = v[0]; partials[2 * blockIdx.x + 1] = v[1]; }
}

__global__ void k_reduce_final(const double* __restrict__ partials, int nparts, int op, double* __restrict__ out) {
    __shared__ double red[32];
    double v[2] = {0.0, 0.0};
    if (op == RED_MAXABS) v[0] = -1.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        if (op == RED_MAXABS) v[0] = fmax(v[0], partials[2 * i]);
        else { v[0] += partials[2 * i]; v[1] += partials[2 * i + 1]; }
    }
    if (op == RED_MAXABS) v[0] = block_max(v[0], red);
    else block_sum<2>(v, red);
    if (threadIdx.x == 0) {
        out[0] = op == RED_NRM2 ? sqrt(v[0]) : v[0];
        out[1] = v[1];
    }
}

// ------------------------------------------------------------------------------------------------------------------
// Standalone batched 2-D FFT (FFT Process, src/processes/FFT.cpp): a row pass and a column pass through LDS.
// conj_io = 1 turns the forward kernel into the unscaled inverse (conj in, conj out); scale multiplies the output.
// ------------------------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(OCL_NT) void k_fft_rows(const float2* __restrict__ in, float2* __restrict__ out,
                                                     long nrows, FftPlan P, int rows_per_wg, int conj_io, float scale) {
    extern __shared__ float2 lds[];
    const int n = P.n, ld = n + 1;
    float2* a = lds;
    float2* b = a + rows_per_wg * ld;
    float2* tw = b + rows_per_wg * ld;
    const long r0 = (long)blockIdx.x * rows_per_wg;
    const float invn = 1.0f / (float)n;
    load_twiddles(tw, P);
    for (int e = threadIdx.x; e < rows_per_wg * n; e += blockDim.x) {
        const int r = fdiv(e, invn), i = e - r * n;
        float2 v = make_float2(0.f, 0.f);
        if (r0 + r < nrows) v = in[(r0 + r) * n + i];
        if (conj_io) v.y = -v.y;
        a[r * ld + i] = v;
    }
    __syncthreads();
    float2* res = lds_fft(a, b, tw, P, rows_per_wg, ld, 1, false);
    for (int e = threadIdx.x; e < rows_per_wg * n; e += blockDim.x) {
        const int r = fdiv(e, invn), i = e - r * n;
        if (r0 + r < nrows) {
            float2 v = res[r * ld + i];
            if (conj_io) v.y = -v.y;
            out[(r0 + r) * n + i] = c_scale(v, scale);
        }
    }
}

__global__ __launch_bounds__(OCL_NT) void k_fft_cols(const float2* __restrict__ in, float2* __restrict__ out, int H,
                                                     int W, FftPlan P, int tw_cols, int conj_io, float scale) {
    extern __shared__ float2 lds[];
    const int ld = tw_cols + 1;
    float2* a = lds;
    float2* b = a + H * ld;
    float2* tw = b + H * ld;
    const long img = blockIdx.y;
    const int w0 = blockIdx.x * tw_cols;
    const float2* I = in + img * (long)H * W;
    float2* O = out + img * (long)H * W;
    load_twiddles(tw, P);
    for (int e = threadIdx.x; e < H * tw_cols; e += blockDim.x) {
        const int h = e / tw_cols, wl = e - h * tw_cols;
        float2 v = make_float2(0.f, 0.f);
        if (w0 + wl < W) v = I[(long)h * W + w0 + wl];
        if (conj_io) v.y = -v.y;
        a[h * ld + wl] = v;
    }
    __syncthreads();
    float2* res = lds_fft(a, b, tw, P, tw_cols, 1, ld, true);
    for (int e = threadIdx.x; e < H * tw_cols; e += blockDim.x) {
        const int h = e / tw_cols, wl = e - h * tw_cols;
        if (w0 + wl < W) {
            float2 v = res[h * ld + wl];
            if (conj_io) v.y = -v.y;
            O[(long)h * W + w0 + wl] = c_scale(v, scale);
        }
    }
}

// Register four-step variants for lengths n = 16 P (P even, 2..20): one LDS transpose per transform.
// Rows: 16 rows of one workgroup, 16 threads per row.  Columns: 16 columns x 16 roles.
template <int P>
__device__ __forceinline__ void fft_rows_fast(const float2* __restrict__ in, float2* __restrict__ out, long nrows,
                                              const float2* __restrict__ twg, int conj_io, float scale) {
    constexpr int n = 16 * P, RS = P * 17 + 1;
    extern __shared__ float2 lds[];
    float2* tw = lds + 16 * RS;
    const int a = threadIdx.x & 15, r = threadIdx.x >> 4;
    const long row = (long)blockIdx.x * 16 + r;
    const bool valid = row < nrows;
    for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = twg[i];
    // loads from a clamped row, zeroed afterwards (a branch per value would put each load in a block of its own,
    // waited on there); twiddles read before the exchange writes, which share the LDS array with them
    const long rowc = valid ? row : 0;
    float2 v[P];
#pragma unroll
    for (int b = 0; b < P; ++b) {
        float2 x = in[rowc * n + a + 16 * b];
        if (!valid) x = make_float2(0.f, 0.f);
        if (conj_io) x.y = -x.y;
        v[b] = x;
    }
    __syncthreads();
    rf::rdft<P>(v);
    float2* X = lds + r * RS;
    float2 twr[P];
#pragma unroll
    for (int k1 = 1; k1 < P; ++k1) twr[k1] = tw[a * k1];
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) X[k1 * 17 + a] = (k1 == 0 || a == 0) ? v[k1] : c_mul(v[k1], twr[k1]);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int k1 = a + 16 * q;
        if (k1 < P) {
            float2 u[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) u[i] = X[k1 * 17 + i];
            rf::rdft<16>(u);
            if (valid)
#pragma unroll
                for (int k2 = 0; k2 < 16; ++k2) {
                    float2 y = u[k2];
                    if (conj_io) y.y = -y.y;
                    out[row * n + k1 + P * k2] = c_scale(y, scale);
                }
        }
    }
}

template <int P>
__device__ __forceinline__ void fft_cols_fast(const float2* __restrict__ in, float2* __restrict__ out, int W,
                                              const float2* __restrict__ twg, int conj_io, float scale) {
    constexpr int H = 16 * P, KS = FAST_COLS_KS;
    extern __shared__ float2 lds[];
    float2* tw = lds + P * KS;
    const int s = threadIdx.x & 15, t = threadIdx.x >> 4;
    const long img = blockIdx.y;
    const int w = blockIdx.x * 16 + s;
    const bool wv = w < W;
    const float2* I = in + img * (long)H * W;
    float2* O = out + img * (long)H * W;
    for (int i = threadIdx.x; i < H; i += blockDim.x) tw[i] = twg[i];
    // as in the row pass: loads from a clamped column, zeroed afterwards; twiddles read before the exchange writes
    const int wc = wv ? w : 0;
    float2 v[P];
#pragma unroll
    for (int b = 0; b < P; ++b) {
        float2 x = I[(long)(t + 16 * b) * W + wc];
        if (!wv) x = make_float2(0.f, 0.f);
        if (conj_io) x.y = -x.y;
        v[b] = x;
    }
    __syncthreads();
    rf::rdft<P>(v);
    float2 twr[P];
#pragma unroll
    for (int k1 = 1; k1 < P; ++k1) twr[k1] = tw[t * k1];
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) lds[k1 * KS + t * 16 + s] = (k1 == 0 || t == 0) ? v[k1] : c_mul(v[k1], twr[k1]);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int k1 = t + 16 * q;
        if (k1 < P) {
            float2 u[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) u[i] = lds[k1 * KS + i * 16 + s];
            rf::rdft<16>(u);
            if (wv)
#pragma unroll
                for (int k2 = 0; k2 < 16; ++k2) {
                    float2 y = u[k2];
                    if (conj_io) y.y = -y.y;
                    O[(long)(k1 + P * k2) * W + w] = c_scale(y, scale);
                }
        }
    }
}

// SimpleMRIRecon after its row transform (src/processes/examples/SimpleMRIRecon.cpp:59-101): for 16 columns of frame
// f, every coil's inverse column DFT (register four-step, FA order), times conj(S_c), is written back in place (the
// state the reference leaves in its input) and summed over the coils in registers, so x is written once.  scale is
// the 1 / (H W) of the inverse 2-D transform.
template <int P>
__device__ __forceinline__ void simple_cols_fast(float2* __restrict__ y, const float2* __restrict__ S,
                                                 float2* __restrict__ x, int C, int W, const float2* __restrict__ twg,
                                                 float scale, int f) {
    constexpr int H = 16 * P, KS = FAST_COLS_KS;
    constexpr int NQ = P > 16 ? 2 : 1;
    extern __shared__ float2 lds[];
    float2* tw = lds + P * KS;
    const int s = threadIdx.x & 15, t = threadIdx.x >> 4;
    const int w = blockIdx.x * 16 + s;
    const bool wv = w < W;
    const int wc = wv ? w : 0;
    const long N = (long)H * W;
    // every access of a coil image goes through a buffer resource over that image (N x 8 bytes < 4 GiB): one 32-bit
    // lane offset (row t, column wc) and scalar row offsets, instead of 64-bit per-lane addresses kept per row
    const unsigned vo = (unsigned)(t * W + wc) * 8u;
    const unsigned rowB = (unsigned)W * 8u;
    for (int i = threadIdx.x; i < H; i += blockDim.x) tw[i] = twg[i];
    float2 acc[NQ][16];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) acc[q][k2] = make_float2(0.f, 0.f);
    for (int c = 0; c < C; ++c) {
        const BufRes rI = buf_res(y + ((long)f * C + c) * N, N * 8);
        const BufRes rS = buf_res(S + (long)c * N, N * 8);
        float2 v[P];
#pragma unroll
        for (int b = 0; b < P; ++b) {
            const float2 a = bld2(rI, vo, 16u * b * rowB);
            v[b] = wv ? make_float2(a.x, -a.y) : make_float2(0.f, 0.f);    // IFFT = conj(FFT(conj(.)))
        }
        __syncthreads();   // twiddles ready / the previous coil's exchange fully read
        rf::rdft<P>(v);
        // reads before writes: the twiddles share the LDS array with the exchange, so a read placed after a write
        // would wait for it, one value at a time
        float2 twr[P];
#pragma unroll
        for (int k1 = 1; k1 < P; ++k1) twr[k1] = tw[t * k1];
#pragma unroll
        for (int k1 = 0; k1 < P; ++k1) lds[k1 * KS + t * 16 + s] = (k1 == 0 || t == 0) ? v[k1] : c_mul(v[k1], twr[k1]);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int k1 = t + 16 * q;
            if (k1 < P) {
                // row k1 + P k2 = t + (16 q + P k2): the lane offset vo and a scalar row offset
                float2 u[16], sv[16];
#pragma unroll
                for (int k2 = 0; k2 < 16; ++k2) sv[k2] = bld2(rS, vo, (unsigned)(16 * q + P * k2) * rowB);
#pragma unroll
                for (int i = 0; i < 16; ++i) u[i] = lds[k1 * KS + i * 16 + s];
                rf::rdft<16>(u);
#pragma unroll
                for (int k2 = 0; k2 < 16; ++k2) {
                    const float2 z = make_float2(u[k2].x * scale, -u[k2].y * scale);
                    const float2 pz = c_mulconj(z, sv[k2]);
                    acc[q][k2] = c_add(acc[q][k2], pz);
                    if (wv) bst2(rI, vo, (unsigned)(16 * q + P * k2) * rowB, pz);
                }
            }
        }
    }
    const BufRes rX = buf_res(x + (long)f * N, N * 8);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int k1 = t + 16 * q;
        if (k1 < P && wv)
#pragma unroll
            for (int k2 = 0; k2 < 16; ++k2) bst2(rX, vo, (unsigned)(16 * q + P * k2) * rowB, acc[q][k2]);
    }
}

// One kernel per column size: a kernel holding every P is allocated the registers of its largest case (P = 20,
// two 16-value accumulators), which held P = 16 at one wavefront per SIMD.
template <int P>
__global__ __launch_bounds__(OCL_NT) void k_simple_cols_t(float2* y, const float2* S, float2* x, int C, int W,
                                                          const float2* tw, float scale) {
    simple_cols_fast<P>(y, S, x, C, W, tw, scale, blockIdx.y);
}

// A batch of reconstructions of one shape: blockIdx.y = recon * F + frame (columns), recon (rows, blockIdx.x the
// 16-row group within it).  The pointers travel in the kernel argument.
template <int P>
__global__ __launch_bounds__(OCL_NT) void k_simple_cols_batch_t(ReconPtrs p, int F, int C, int W, const float2* tw,
                                                                float scale) {
    const int r = blockIdx.y / F, f = blockIdx.y - r * F;
    simple_cols_fast<P>(p.y[r], p.S[r], p.x[r], C, W, tw, scale, f);
}

#define OCL_SIMPLE_COLS_CASES(X) X(2) X(4) X(6) X(8) X(10) X(12) X(14) X(16) X(18) X(20)

hipError_t launch_simple_cols(float2* y, const float2* S, float2* x, int F, int C, int W, int P, const float2* tw,
                              float scale, size_t lds, hipStream_t stream) {
    const dim3 grid((unsigned)((W + 15) / 16), (unsigned)F);
    switch (P) {
#define OCL_SC_CASE(Q) \
        case Q: hipLaunchKernelGGL(k_simple_cols_t<Q>, grid, dim3(OCL_NT), lds, stream, y, S, x, C, W, tw, scale); break;
        OCL_SIMPLE_COLS_CASES(OCL_SC_CASE)
#undef OCL_SC_CASE
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_simple_cols_batch(const ReconPtrs& p, int m, int F, int C, int W, int P, const float2* tw,
                                    float scale, size_t lds, hipStream_t stream) {
    const dim3 grid((unsigned)((W + 15) / 16), (unsigned)(m * F));
    switch (P) {
#define OCL_SC_CASE(Q) \
        case Q: hipLaunchKernelGGL(k_simple_cols_batch_t<Q>, grid, dim3(OCL_NT), lds, stream, p, F, C, W, tw, scale); \
            break;
        OCL_SIMPLE_COLS_CASES(OCL_SC_CASE)
#undef OCL_SC_CASE
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

void simple_cols_attributes(int lds_max) {
#define OCL_SC_CASE(Q)                                                                                              \
    (void)hipFuncSetAttribute((const void*)k_simple_cols_t<Q>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max); \
    (void)hipFuncSetAttribute((const void*)k_simple_cols_batch_t<Q>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
    OCL_SIMPLE_COLS_CASES(OCL_SC_CASE)
#undef OCL_SC_CASE
}

__global__ __launch_bounds__(OCL_NT) void k_fft_rows_fast_batch(ReconPtrs p, long nrows, int P, const float2* tw,
                                                                int conj_io, float scale) {
    float2* y = p.y[blockIdx.y];
    switch (P) {
        case 2: fft_rows_fast<2>(y, y, nrows, tw, conj_io, scale); break;
        case 4: fft_rows_fast<4>(y, y, nrows, tw, conj_io, scale); break;
        case 6: fft_rows_fast<6>(y, y, nrows, tw, conj_io, scale); break;
        case 8: fft_rows_fast<8>(y, y, nrows, tw, conj_io, scale); break;
        case 10: fft_rows_fast<10>(y, y, nrows, tw, conj_io, scale); break;
        case 12: fft_rows_fast<12>(y, y, nrows, tw, conj_io, scale); break;
        case 14: fft_rows_fast<14>(y, y, nrows, tw, conj_io, scale); break;
        case 16: fft_rows_fast<16>(y, y, nrows, tw, conj_io, scale); break;
        case 18: fft_rows_fast<18>(y, y, nrows, tw, conj_io, scale); break;
        default: fft_rows_fast<20>(y, y, nrows, tw, conj_io, scale); break;
    }
}

__global__ __launch_bounds__(OCL_NT) void k_fft_rows_fast(const float2* __restrict__ in, float2* __restrict__ out,
                                                          long nrows, int P, const float2* __restrict__ tw,
                                                          int conj_io, float scale) {
    switch (P) {
        case 2: fft_rows_fast<2>(in, out, nrows, tw, conj_io, scale); break;
        case 4: fft_rows_fast<4>(in, out, nrows, tw, conj_io, scale); break;
        case 6: fft_rows_fast<6>(in, out, nrows, tw, conj_io, scale); break;
        case 8: fft_rows_fast<8>(in, out, nrows, tw, conj_io, scale); break;
        case 10: fft_rows_fast<10>(in, out, nrows, tw, conj_io, scale); break;
        case 12: fft_rows_fast<12>(in, out, nrows, tw, conj_io, scale); break;
        case 14: fft_rows_fast<14>(in, out, nrows, tw, conj_io, scale); break;
        case 16: fft_rows_fast<16>(in, out, nrows, tw, conj_io, scale); break;
        case 18: fft_rows_fast<18>(in, out, nrows, tw, conj_io, scale); break;
        default: fft_rows_fast<20>(in, out, nrows, tw, conj_io, scale); break;
    }
}

__global__ __launch_bounds__(OCL_NT) void k_fft_cols_fast(const float2* __restrict__ in, float2* __restrict__ out,
                                                          int W, int P, const float2* __restrict__ tw, int conj_io,
                                                          float scale) {
    switch (P) {
        case 2: fft_cols_fast<2>(in, out, W, tw, conj_io, scale); break;
        case 4: fft_cols_fast<4>(in, out, W, tw, conj_io, scale); break;
        case 6: fft_cols_fast<6>(in, out, W, tw, conj_io, scale); break;
        case 8: fft_cols_fast<8>(in, out, W, tw, conj_io, scale); break;
        case 10: fft_cols_fast<10>(in, out, W, tw, conj_io, scale); break;
        case 12: fft_cols_fast<12>(in, out, W, tw, conj_io, scale); break;
        case 14: fft_cols_fast<14>(in, out, W, tw, conj_io, scale); break;
        case 16: fft_cols_fast<16>(in, out, W, tw, conj_io, scale); break;
        case 18: fft_cols_fast<18>(in, out, W, tw, conj_io, scale); break;
        default: fft_cols_fast<20>(in, out, W, tw, conj_io, scale); break;
    }
}

// y = M * (y + sigma n), n complex Gaussian from splitmix64(seed ^ index) and the Box-Muller transform
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__global__ void k_add_noise_masked(float2* __restrict__ y, const uint8_t* __restrict__ M, int F, int C, long N,
                                   float sigma, unsigned long long seed) {
    const long total = (long)F * C * N;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const long fc = i / N;
        const long p = i - fc * N;
        const int f = (int)(fc / C);
        if (!M[(long)f * N + p]) { y[i] = make_float2(0.f, 0.f); continue; }
        const unsigned long long r = mix64(seed ^ (unsigned long long)i);
        const double u1 = ((double)(r >> 11) + 1.0) * (1.0 / 9007199254740993.0);   // (0, 1]
        const double u2 = (double)(mix64(r) >> 11) * (1.0 / 9007199254740992.0);
        const double rad = sqrt(-log(u1)) * (double)sigma;                             // variance sigma^2 / 2 per part
        const float2 v = y[i];
        y[i] = make_float2(v.x + (float)(rad * cos(6.283185307179586 * u2)), v.y + (float)(rad * sin(6.283185307179586 * u2)));
    }
}

// drand48 stream from a given 48-bit state, written as (float(value), 0) pairs: out[i] = X_{start+i+1} / 2^48.
// Each thread jumps to its chunk with the LCG's affine power, then steps sequentially.
__global__ void k_drand48_fill(float2* __restrict__ out, long n, unsigned long long x0, long start) {
    const unsigned long long A = 0x5DEECE66DULL, Cc = 0xBULL, MASK = (1ULL << 48) - 1;
    const int CH = 64;
    const long chunk = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long i0 = chunk * CH;
    if (i0 >= n) return;
    // state after (start + i0) steps: x = A^k x0 + C (A^{k-1} + ... + 1)
    unsigned long long k = (unsigned long long)(start + i0);
    unsigned long long ma = 1, mc = 0;          // accumulated map x -> ma*x + mc
    unsigned long long pa = A, pc = Cc;         // map for 2^b steps
    while (k) {
        if (k & 1) { ma = (ma * pa) & MASK; mc = (mc * pa + pc) & MASK; }
        pc = (pc * pa + pc) & MASK;
        pa = (pa * pa) & MASK;
        k >>= 1;
    }
    unsigned long long x = (ma * x0 + mc) & MASK;
    for (int t = 0; t < CH && i0 + t < n; ++t) {
        x = (A * x + Cc) & MASK;
        out[i0 + t] = make_float2((float)((double)x / 281474976710656.0), 0.f);
    }
}
