// Device-side building blocks shared by every gfx950 kernel of the NESTA-TV path:
//  * complex float2 arithmetic,
//  * an LDS-resident mixed-radix Stockham FFT (radices 2,3,4,5,7,8) driven by a host-built plan,
//  * fixed-order block reductions in double (deterministic for a fixed launch shape).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define OCL_FFT_MAX_STAGES 12   // radix stages of one FFT plan
#define OCL_NT 256          // threads per workgroup for every kernel of this library (4 wavefronts of 64)

// Factorisation of one DFT length plus its twiddle table tw[m] = exp(-2*pi*i*m/n) (device memory).
struct FftPlan {
    int n;
    int nst;
    int radix[OCL_FFT_MAX_STAGES];
    const float2* tw;
};

// ------------------------------------------------------------------------------------------------------------------
// global-memory access at a 32-bit element index from a wave-uniform base: the pointer is taken as a global-segment
// pointer and the index as an unsigned 32-bit offset, so the loads and stores use the scalar-base + 32-bit-offset
// global forms instead of flat accesses with 64-bit address arithmetic.  Callers keep base + index inside one array
// of less than 4 GiB (bases advance per frame / coil).
// ------------------------------------------------------------------------------------------------------------------
typedef float ocl_f2v __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const char ocl_gcchar;
typedef __attribute__((address_space(1))) char ocl_gchar;

__device__ __forceinline__ float2 ldg2(const float2* base, unsigned i) {
    const ocl_f2v v = *(__attribute__((address_space(1))) const ocl_f2v*)((ocl_gcchar*)base + i * 8u);
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ void stg2(float2* base, unsigned i, float2 x) {
    ocl_f2v v;
    v.x = x.x;
    v.y = x.y;
    *(__attribute__((address_space(1))) ocl_f2v*)((ocl_gchar*)base + i * 8u) = v;
}
__device__ __forceinline__ uint8_t ldg1(const uint8_t* base, unsigned i) {
    return *(__attribute__((address_space(1))) const uint8_t*)((ocl_gcchar*)base + i);
}

// Buffer-resource access: a raw buffer (stride 0, range in bytes, gfx9 data-format word) addressed by a 32-bit lane
// offset in a vector register and a 32-bit offset in a scalar register, both in bytes.  Inside loops this keeps the
// per-row offsets scalar: pointer-based forms let the compiler hoist one 64-bit per-lane address per row out of the
// loop.  Ranges stay below 4 GiB.
typedef __amdgpu_buffer_rsrc_t BufRes;
__device__ __forceinline__ BufRes buf_res(const void* p, long bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)(unsigned)bytes, 0x00020000);
}
__device__ __forceinline__ float2 bld2(BufRes r, unsigned vo, unsigned so) {
    // the 64-bit load's value is taken as a whole (bit cast), never by element
    const ocl_f2v v = __builtin_bit_cast(ocl_f2v, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ void bst2(BufRes r, unsigned vo, unsigned so, float2 x) {
    ocl_f2v v;
    v.x = x.x;
    v.y = x.y;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v),
                                          r, vo, so, 0);
}

// ------------------------------------------------------------------------------------------------------------------
// complex helpers
// ------------------------------------------------------------------------------------------------------------------
__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 c_mul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 c_mulconj(float2 a, float2 s) {   // a * conj(s)
    return make_float2(a.x * s.x + a.y * s.y, a.y * s.x - a.x * s.y);
}
__device__ __forceinline__ float2 c_scale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 c_conj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 c_mul_negi(float2 a) { return make_float2(a.y, -a.x); }   // -i * a
__device__ __forceinline__ float c_abs2(float2 a) { return a.x * a.x + a.y * a.y; }

// floor(j / d) for 0 <= j < 2^22 and d >= 1, using a precomputed 1/d (exact in that range).
__device__ __forceinline__ int fdiv(int j, float inv_d) { return (int)(((float)j + 0.5f) * inv_d); }

// ------------------------------------------------------------------------------------------------------------------
// small forward DFTs held in registers
// ------------------------------------------------------------------------------------------------------------------
__device__ __forceinline__ void dft2(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = c_add(a, b);
    v[1] = c_sub(a, b);
}

__device__ __forceinline__ void dft4(float2& v0, float2& v1, float2& v2, float2& v3) {
    float2 t0 = c_add(v0, v2), t1 = c_sub(v0, v2), t2 = c_add(v1, v3), t3 = c_mul_negi(c_sub(v1, v3));
    v0 = c_add(t0, t2);
    v2 = c_sub(t0, t2);
    v1 = c_add(t1, t3);
    v3 = c_sub(t1, t3);
}

__device__ __forceinline__ void dft8(float2* v) {
    // radix-2 split into even/odd 4-point DFTs, then combine with w8^q
    float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4(e0, e1, e2, e3);
    dft4(o0, o1, o2, o3);
    const float h = 0.70710678118654752440f;
    o1 = make_float2(h * (o1.x + o1.y), h * (o1.y - o1.x));      // * (1 - i)/sqrt2
    o2 = c_mul_negi(o2);                                         // * -i
    o3 = make_float2(h * (o3.y - o3.x), -h * (o3.x + o3.y));     // * (-1 - i)/sqrt2
    v[0] = c_add(e0, o0); v[4] = c_sub(e0, o0);
    v[1] = c_add(e1, o1); v[5] = c_sub(e1, o1);
    v[2] = c_add(e2, o2); v[6] = c_sub(e2, o2);
    v[3] = c_add(e3, o3); v[7] = c_sub(e3, o3);
}

// Odd prime R in {3,5,7}: X_q = a0 + sum_k cos(2 pi qk/R)(a_k + a_{R-k}) -/+ i sum_k sin(2 pi qk/R)(a_k - a_{R-k})
template <int R> struct OddConst;
template <> struct OddConst<3> {
    static __device__ __forceinline__ float c(int m) { return -0.5f; }
    static __device__ __forceinline__ float s(int m) { return m == 1 ? 0.86602540378443864676f : -0.86602540378443864676f; }
};
template <> struct OddConst<5> {
    static __device__ __forceinline__ float c(int m) {
        return (m == 1 || m == 4) ? 0.30901699437494742410f : -0.80901699437494742410f;
    }
    static __device__ __forceinline__ float s(int m) {
        switch (m) {
            case 1: return 0.95105651629515357212f;
            case 2: return 0.58778525229247312917f;
            case 3: return -0.58778525229247312917f;
            default: return -0.95105651629515357212f;
        }
    }
};
template <> struct OddConst<7> {
    static __device__ __forceinline__ float c(int m) {
        switch (m) {
            case 1: case 6: return 0.62348980185873353053f;
            case 2: case 5: return -0.22252093395631440429f;
            default: return -0.90096886790241912624f;
        }
    }
    static __device__ __forceinline__ float s(int m) {
        switch (m) {
            case 1: return 0.78183148246802980871f;
            case 2: return 0.97492791218182360702f;
            case 3: return 0.43388373911755812048f;
            case 4: return -0.43388373911755812048f;
            case 5: return -0.97492791218182360702f;
            default: return -0.78183148246802980871f;
        }
    }
};

template <int R> __device__ __forceinline__ void dft_odd(float2* v) {
    constexpr int H = (R - 1) / 2;
    float2 b[H], d[H];
#pragma unroll
    for (int k = 1; k <= H; ++k) { b[k - 1] = c_add(v[k], v[R - k]); d[k - 1] = c_sub(v[k], v[R - k]); }
    float2 a0 = v[0];
    float2 s0 = a0;
#pragma unroll
    for (int k = 0; k < H; ++k) s0 = c_add(s0, b[k]);
#pragma unroll
    for (int q = 1; q <= H; ++q) {
        float re_x = a0.x, re_y = a0.y, im_x = 0.f, im_y = 0.f;
#pragma unroll
        for (int k = 1; k <= H; ++k) {
            const int m = (q * k) % R;
            const float cc = OddConst<R>::c(m), ss = OddConst<R>::s(m);
            re_x += cc * b[k - 1].x; re_y += cc * b[k - 1].y;
            im_x += ss * d[k - 1].x; im_y += ss * d[k - 1].y;
        }
        // X_q = re - i*im ; X_{R-q} = re + i*im   (i*(x+iy) = -y + ix)
        v[q] = make_float2(re_x + im_y, re_y - im_x);
        v[R - q] = make_float2(re_x - im_y, re_y + im_x);
    }
    v[0] = s0;
}

template <int R> __device__ __forceinline__ void dft(float2* v) {
    if constexpr (R == 2) dft2(v);
    else if constexpr (R == 4) dft4(v[0], v[1], v[2], v[3]);
    else if constexpr (R == 8) dft8(v);
    else dft_odd<R>(v);
}

// ------------------------------------------------------------------------------------------------------------------
// Stockham stage: every thread handles the butterflies j = lane, lane+TP, ... of sequence s.
// Element e of sequence s sits at buf[s*ss + e*es].
// ------------------------------------------------------------------------------------------------------------------
template <int R>
__device__ __forceinline__ void fft_stage(const float2* __restrict__ src, float2* __restrict__ dst,
                                          const float2* __restrict__ tw, int n, int Ls, int s, int ss, int es,
                                          int lane, int TP) {
    const int m = n / R;
    const int tstride = n / (Ls * R);
    const float invLs = 1.0f / (float)Ls;
    const int sb = s * ss;
    for (int j = lane; j < m; j += TP) {
        const int jh = fdiv(j, invLs);
        const int k = j - jh * Ls;
        float2 v[R];
        v[0] = src[sb + j * es];
        const int t1 = k * tstride;
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = c_mul(src[sb + (j + r * m) * es], tw[r * t1]);
        dft<R>(v);
        const int base = jh * Ls * R + k;
#pragma unroll
        for (int q = 0; q < R; ++q) dst[sb + (base + q * Ls) * es] = v[q];
    }
}

// Forward DFT of `nseq` sequences held in LDS buffer a (ping-pong with b, same layout).  The caller has
// synchronised after filling a.  Returns the buffer that holds the result; ends with a barrier.
// Thread mapping: seqfast -> s = tid % nseq (lanes run along sequences; column tiles), otherwise
// s = tid / (NT/nseq) (lanes run along one sequence; row tiles).  nseq must divide blockDim.x.
__device__ inline float2* lds_fft(float2* a, float2* b, const float2* tw, const FftPlan& P, int nseq, int ss, int es,
                                  bool seqfast) {
    const int NT = blockDim.x;
    const int TP = NT / nseq;
    const int tid = threadIdx.x;
    const int s = seqfast ? (tid % nseq) : (tid / TP);
    const int lane = seqfast ? (tid / nseq) : (tid % TP);
    int Ls = 1;
    for (int st = 0; st < P.nst; ++st) {
        const int R = P.radix[st];
        switch (R) {
            case 2: fft_stage<2>(a, b, tw, P.n, Ls, s, ss, es, lane, TP); break;
            case 3: fft_stage<3>(a, b, tw, P.n, Ls, s, ss, es, lane, TP); break;
            case 4: fft_stage<4>(a, b, tw, P.n, Ls, s, ss, es, lane, TP); break;
            case 5: fft_stage<5>(a, b, tw, P.n, Ls, s, ss, es, lane, TP); break;
            case 7: fft_stage<7>(a, b, tw, P.n, Ls, s, ss, es, lane, TP); break;
            default: fft_stage<8>(a, b, tw, P.n, Ls, s, ss, es, lane, TP); break;
        }
        __syncthreads();
        float2* t = a; a = b; b = t;
        Ls *= R;
    }
    return a;
}

// Copies the twiddle table of plan P into LDS (no barrier).
__device__ __forceinline__ void load_twiddles(float2* tw_lds, const FftPlan& P) {
    for (int i = threadIdx.x; i < P.n; i += blockDim.x) tw_lds[i] = P.tw[i];
}

// ------------------------------------------------------------------------------------------------------------------
// fixed-order block reductions (blockDim.x == OCL_NT); result valid in every thread
// ------------------------------------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double* red /* LDS, >= K*OCL_NT/64 doubles */) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, nw = blockDim.x >> 6;
#pragma unroll
    for (int i = 0; i < K; ++i) v[i] = wave_sum(v[i]);
    __syncthreads();
    if (l == 0)
#pragma unroll
        for (int i = 0; i < K; ++i) red[i * 16 + w] = v[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < K; ++i) {
        double s = 0;
        for (int q = 0; q < nw; ++q) s += red[i * 16 + q];
        v[i] = s;
    }
    __syncthreads();
}

__device__ __forceinline__ double block_max(double v, double* red) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, nw = blockDim.x >> 6;
    v = wave_max(v);
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    double m = red[0];
    for (int q = 1; q < nw; ++q) m = fmax(m, red[q]);
    __syncthreads();
    return m;
}
