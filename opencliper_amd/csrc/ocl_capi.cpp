// Host side of libocl_amd.so: the C ABI declared in include/ocl_amd.h.
//
// One ocl_ctx = one HIP device + one in-order stream (the role CLapp plays in src/CLapp.cpp:139-569).  FFT plans are
// factorisations into radices {8,4,2,3,5,7} with a per-length twiddle table cached in the context.  NESTA-TV runs in
// two forms: PLAIN issues the reference's op sequence (NestaUp.cpp:442-910) with a host stop test after every
// iteration; FUSED advances one or many slices with two fused passes per iteration and a device-side stop test.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "ocl_amd.h"
#include "ocl_kernels.hpp"
#include "ocl_nesta_fast.hpp"

// ---------------------------------------------------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) return fail(OCL_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define CHK(x)                  \
    do {                        \
        int r_ = (x);           \
        if (r_ != OCL_OK) return r_; \
    } while (0)

extern "C" const char* ocl_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------------------------------------------------
struct ocl_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::map<int, float2*> tw;                              // twiddle tables by DFT length
    double* red_part = nullptr;                             // reduction partials [1024][2]
    double* red_out = nullptr;                              // reduction results [8][2]
    double* red_host = nullptr;                             // pinned mirror of red_out
    std::map<std::pair<int, long>, std::pair<double, int>> normU_cache;
    std::mutex mu;
    int num_cus = 256;
};

static const int RED_PARTS = 1024;
static const size_t LDS_MAX = 64 * 1024;   // per-workgroup dynamic LDS budget of every launch

extern "C" int ocl_ctx_create(int device, ocl_ctx** out) {
    if (!out) return fail(OCL_E_INVALID, "ocl_ctx_create: null output pointer");
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(OCL_E_INVALID, "ocl_ctx_create: no such device");
    HIPCHK(hipSetDevice(device));
    ocl_ctx* c = new ocl_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&c->red_part, sizeof(double) * 2 * RED_PARTS);
    if (e == hipSuccess) e = hipMalloc(&c->red_out, sizeof(double) * 16);
    if (e == hipSuccess) e = hipHostMalloc(&c->red_host, sizeof(double) * 16, hipHostMallocDefault);
    const void* lds_kernels[] = {(const void*)k_fft_rows, (const void*)k_fft_cols, (const void*)k_fft_rows_fast,
                                 (const void*)k_fft_cols_fast};
    if (e == hipSuccess) {
        for (const void* k : lds_kernels) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_MAX);
        nesta_kernel_attributes((int)LDS_MAX);
    }
    (void)hipGetLastError();
    if (e != hipSuccess) {
        std::string m = std::string("ocl_ctx_create: ") + hipGetErrorString(e);
        ocl_ctx_destroy(c);
        return fail(OCL_E_HIP, m);
    }
    *out = c;
    return OCL_OK;
}

extern "C" int ocl_ctx_destroy(ocl_ctx* c) {
    if (!c) return OCL_OK;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto& kv : c->tw) hipFree(kv.second);
    if (c->red_part) hipFree(c->red_part);
    if (c->red_out) hipFree(c->red_out);
    if (c->red_host) hipHostFree(c->red_host);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return OCL_OK;
}

extern "C" int ocl_ctx_sync(ocl_ctx* c) {
    if (!c) return fail(OCL_E_INVALID, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return OCL_OK;
}

extern "C" void* ocl_ctx_stream(ocl_ctx* c) { return c ? (void*)c->stream : nullptr; }

extern "C" int ocl_device_info(ocl_ctx* c, char* name, size_t name_len, size_t* total_mem, int* num_cus) {
    if (!c) return fail(OCL_E_INVALID, "null context");
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, c->device));
    if (name && name_len) {
        std::snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (total_mem) *total_mem = p.totalGlobalMem;
    if (num_cus) *num_cus = p.multiProcessorCount;
    return OCL_OK;
}

// ---------------------------------------------------------------------------------------------------------------------
// timing
// ---------------------------------------------------------------------------------------------------------------------
struct ocl_timer {
    hipEvent_t a = nullptr, b = nullptr;
};

extern "C" int ocl_timer_begin(ocl_ctx* c, ocl_timer** t) {
    if (!c || !t) return fail(OCL_E_INVALID, "ocl_timer_begin: null argument");
    ocl_timer* q = new ocl_timer();
    if (hipEventCreate(&q->a) != hipSuccess || hipEventCreate(&q->b) != hipSuccess ||
        hipEventRecord(q->a, c->stream) != hipSuccess) {
        if (q->a) hipEventDestroy(q->a);
        if (q->b) hipEventDestroy(q->b);
        delete q;
        return fail(OCL_E_HIP, "ocl_timer_begin: event creation failed");
    }
    *t = q;
    return OCL_OK;
}

extern "C" int ocl_timer_end(ocl_ctx* c, ocl_timer* t, double* ms) {
    if (!c || !t) return fail(OCL_E_INVALID, "ocl_timer_end: null argument");
    float v = 0.f;
    hipError_t e = hipEventRecord(t->b, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(t->b);
    if (e == hipSuccess) e = hipEventElapsedTime(&v, t->a, t->b);
    hipEventDestroy(t->a);
    hipEventDestroy(t->b);
    delete t;
    if (e != hipSuccess) return fail(OCL_E_HIP, std::string("ocl_timer_end: ") + hipGetErrorString(e));
    if (ms) *ms = v;
    return OCL_OK;
}

// ---------------------------------------------------------------------------------------------------------------------
// memory
// ---------------------------------------------------------------------------------------------------------------------
extern "C" int ocl_malloc(ocl_ctx* c, size_t bytes, void** p) {
    if (!c || !p) return fail(OCL_E_INVALID, "ocl_malloc: null argument");
    HIPCHK(hipSetDevice(c->device));
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e == hipErrorOutOfMemory) return fail(OCL_E_NOMEM, "ocl_malloc: out of device memory");
    HIPCHK(e);
    return OCL_OK;
}
extern "C" int ocl_free(ocl_ctx* c, void* p) {
    if (!c) return fail(OCL_E_INVALID, "ocl_free: null context");
    if (p) {
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(p));
    }
    return OCL_OK;
}
extern "C" int ocl_memcpy_h2d(ocl_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c) return fail(OCL_E_INVALID, "null context");
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return OCL_OK;
}
extern "C" int ocl_memcpy_d2h(ocl_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c) return fail(OCL_E_INVALID, "null context");
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return OCL_OK;
}
extern "C" int ocl_memcpy_d2d(ocl_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c) return fail(OCL_E_INVALID, "null context");
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return OCL_OK;
}

static int grid_for(long n) { return (int)std::max(1L, std::min((n + OCL_NT - 1) / OCL_NT, 8192L)); }

extern "C" int ocl_memset_zero(ocl_ctx* c, void* p, size_t bytes) {
    if (!c) return fail(OCL_E_INVALID, "null context");
    if (!bytes) return OCL_OK;
    if (bytes % sizeof(float2) == 0 && ((uintptr_t)p % sizeof(float2)) == 0) {
        const long n = (long)(bytes / sizeof(float2));
        hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(OCL_NT), 0, c->stream, (float2*)p, n, make_float2(0.f, 0.f));
        HIPCHK(hipGetLastError());
        return OCL_OK;
    }
    hipLaunchKernelGGL(k_fill_bytes, dim3(grid_for((long)bytes)), dim3(OCL_NT), 0, c->stream, (uint8_t*)p,
                       (long)bytes, (uint8_t)0);
    HIPCHK(hipGetLastError());
    return OCL_OK;
}

// ---------------------------------------------------------------------------------------------------------------------
// FFT plans
// ---------------------------------------------------------------------------------------------------------------------
static int make_plan(ocl_ctx* c, int n, FftPlan* P) {
    if (n < 1) return fail(OCL_E_INVALID, "FFT length must be positive");
    std::vector<int> r;
    int m = n, p2 = 0;
    while (m % 2 == 0) { m /= 2; ++p2; }
    while (p2 >= 3) { r.push_back(8); p2 -= 3; }
    if (p2 == 2) r.push_back(4);
    if (p2 == 1) r.push_back(2);
    for (int p : {3, 5, 7})
        while (m % p == 0) { r.push_back(p); m /= p; }
    if (m != 1)
        return fail(OCL_E_UNSUPPORTED,
                    "FFT length " + std::to_string(n) + " is not a product of powers of 2, 3, 5 and 7");
    if ((int)r.size() > OCL_MAX_STAGES) return fail(OCL_E_UNSUPPORTED, "FFT length has too many factors");
    std::memset(P, 0, sizeof(*P));
    P->n = n;
    P->nst = (int)r.size();
    for (size_t i = 0; i < r.size(); ++i) P->radix[i] = r[i];
    std::lock_guard<std::mutex> lk(c->mu);
    auto it = c->tw.find(n);
    if (it == c->tw.end()) {
        std::vector<float2> h(n);
        for (int k = 0; k < n; ++k) {
            const double a = -2.0 * M_PI * (double)k / (double)n;
            h[k] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
        float2* d = nullptr;
        HIPCHK(hipMalloc(&d, sizeof(float2) * n));
        HIPCHK(hipMemcpy(d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
        it = c->tw.emplace(n, d).first;
    }
    P->tw = it->second;
    return OCL_OK;
}

// P when n = 16 P takes the register four-step path (P even, 2..20), else 0
static int fast_p(int n) {
    if (n % 16) return 0;
    const int P = n / 16;
    return (P >= 2 && P <= 20 && P % 2 == 0) ? P : 0;
}

static int pow2_floor(int v) {
    int p = 1;
    while (p * 2 <= v) p *= 2;
    return p;
}

// ---------------------------------------------------------------------------------------------------------------------
// individual Processes
// ---------------------------------------------------------------------------------------------------------------------
#define CTXCHK(c) \
    if (!(c)) return fail(OCL_E_INVALID, "null context")

extern "C" int ocl_ttv(ocl_ctx* c, const void* in, void* out, int F, long N, int dir) {
    CTXCHK(c);
    if (F < 1 || N < 1 || !in || !out) return fail(OCL_E_INVALID, "ocl_ttv: bad arguments");
    if (in == out) return fail(OCL_E_INVALID, "ocl_ttv: in and out must not alias");
    hipLaunchKernelGGL(k_ttv, dim3(grid_for((long)F * N)), dim3(OCL_NT), 0, c->stream, (const float2*)in,
                       (float2*)out, F, N, dir ? 1 : 0);
    HIPCHK(hipGetLastError());
    return OCL_OK;
}

extern "C" int ocl_vecnorm(ocl_ctx* c, const void* in, void* out, long n, float mu) {
    CTXCHK(c);
    if (n < 0 || (n && (!in || !out))) return fail(OCL_E_INVALID, "ocl_vecnorm: bad arguments");
    if (!n) return OCL_OK;
    hipLaunchKernelGGL(k_vecnorm, dim3(grid_for(n)), dim3(OCL_NT), 0, c->stream, (const float2*)in, (float2*)out, n,
                       mu);
    HIPCHK(hipGetLastError());
    return OCL_OK;
}

extern "C" int ocl_cep(ocl_ctx* c, const void* in, const void* S, void* out, int F, int C, long N, int in_coils,
                       int conj) {
    CTXCHK(c);
    if (!S) return fail(OCL_E_INVALID, "ComplexElementProd: non-existing SensitivityMaps");
    if (F < 1 || C < 1 || N < 1 || !in || !out) return fail(OCL_E_INVALID, "ComplexElementProd: inputData size is 0");
    hipLaunchKernelGGL(k_cep, dim3(grid_for((long)F * C * N)), dim3(OCL_NT), 0, c->stream, (const float2*)in,
                       (const float2*)S, (float2*)out, F, C, N, in_coils ? 1 : 0, conj ? 1 : 0);
    HIPCHK(hipGetLastError());
    return OCL_OK;
}

extern "C" int ocl_apply_mask(ocl_ctx* c, void* data, const uint8_t* M, int F, int C, long N) {
    CTXCHK(c);
    if (!M) return fail(OCL_E_INVALID, "ApplyMask: null sampling masks");
    if (F < 1 || C < 1 || N < 1 || !data) return fail(OCL_E_INVALID, "ApplyMask: bad arguments");
    hipLaunchKernelGGL(k_apply_mask, dim3(grid_for((long)F * C * N)), dim3(OCL_NT), 0, c->stream, (float2*)data, M, F,
                       C, N);
    HIPCHK(hipGetLastError());
    return OCL_OK;
}

extern "C" int ocl_ximage_sum(ocl_ctx* c, const void* in, void* out, int F, int C, long N) {
    CTXCHK(c);
    if (F < 1 || C < 1 || N < 1 || !in || !out) return fail(OCL_E_INVALID, "XImageSum: bad arguments");
    hipLaunchKernelGGL(k_ximage_sum, dim3(grid_for((long)F * N)), dim3(OCL_NT), 0, c->stream, (const float2*)in,
                       (float2*)out, F, C, N);
    HIPCHK(hipGetLastError());
    return OCL_OK;
}

extern "C" int ocl_complex_abs(ocl_ctx* c, const void* in, void* out, long n, int pow2) {
    CTXCHK(c);
    if (n < 1 || !in || !out) return fail(OCL_E_INVALID, "ComplexAbs: bad arguments");
    hipLaunchKernelGGL(k_abs, dim3(grid_for(n)), dim3(OCL_NT), 0, c->stream, (const float2*)in, (float2*)out, n,
                       pow2 ? 1 : 0);
    HIPCHK(hipGetLastError());
    return OCL_OK;
}

extern "C" int ocl_copy(ocl_ctx* c, const void* in, void* out, long n) {
    CTXCHK(c);
    if (n < 0 || (n && (!in || !out))) return fail(OCL_E_INVALID, "CopyDataGPU: bad arguments");
    if (!n || in == out) return OCL_OK;
    hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(OCL_NT), 0, c->stream, (const float2*)in, (float2*)out, n);
    HIPCHK(hipGetLastError());
    return OCL_OK;
}

extern "C" int ocl_scal(ocl_ctx* c, void* x, long n, float ar, float ai) {
    CTXCHK(c);
    if (n < 0 || (n && !x)) return fail(OCL_E_INVALID, "scal: bad arguments");
    if (!n) return OCL_OK;
    hipLaunchKernelGGL(k_scal, dim3(grid_for(n)), dim3(OCL_NT), 0, c->stream, (float2*)x, n, ar, ai);
    HIPCHK(hipGetLastError());
    return OCL_OK;
}

extern "C" int ocl_axpy(ocl_ctx* c, long n, float ar, float ai, const void* x, void* y) {
    CTXCHK(c);
    if (n < 0 || (n && (!x || !y))) return fail(OCL_E_INVALID, "axpy: bad arguments");
    if (!n) return OCL_OK;
    hipLaunchKernelGGL(k_axpy, dim3(grid_for(n)), dim3(OCL_NT), 0, c->stream, n, ar, ai, (const float2*)x,
                       (float2*)y);
    HIPCHK(hipGetLastError());
    return OCL_OK;
}

// enqueue a reduction whose two doubles land in dev_out (device)
static int reduce_async(ocl_ctx* c, const void* x, const void* y, long n, int op, double* dev_out) {
    if (op < 0 || op > RED_MAXABS) return fail(OCL_E_INVALID, "reduce: unknown operation");
    if (n < 1 || !x || (op == RED_DOTC && !y)) return fail(OCL_E_INVALID, "reduce: bad arguments");
    const int parts = (int)std::max(1L, std::min((long)RED_PARTS, (n + OCL_NT - 1) / OCL_NT));
    hipLaunchKernelGGL(k_reduce_partial, dim3(parts), dim3(OCL_NT), 0, c->stream, (const float2*)x, (const float2*)y,
                       n, op, c->red_part);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(OCL_NT), 0, c->stream, (const double*)c->red_part, parts, op,
                       dev_out);
    HIPCHK(hipGetLastError());
    return OCL_OK;
}

extern "C" int ocl_reduce(ocl_ctx* c, const void* x, const void* y, long n, int op, double* result) {
    CTXCHK(c);
    if (!result) return fail(OCL_E_INVALID, "reduce: null result");
    CHK(reduce_async(c, x, y, n, op, c->red_out));
    HIPCHK(hipMemcpyAsync(c->red_host, c->red_out, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    result[0] = c->red_host[0];
    result[1] = c->red_host[1];
    return OCL_OK;
}

// ---------------------------------------------------------------------------------------------------------------------
// 2-D FFT
// ---------------------------------------------------------------------------------------------------------------------
static int fft_rows_generic(ocl_ctx* c, const void* in, void* out, long nrows, int W, const FftPlan& pw,
                            int conj_io);
static int fft_cols_generic(ocl_ctx* c, void* out, long batch, int H, int W, const FftPlan& ph, int conj_io,
                            float scale2);

static int fft2d(ocl_ctx* c, const void* in, void* out, long batch, int H, int W, int dir) {
    if (batch < 1 || H < 1 || W < 1 || !in || !out) return fail(OCL_E_INVALID, "FFT: bad arguments");
    if (H > 2048 || W > 2048) return fail(OCL_E_UNSUPPORTED, "FFT: lengths above 2048 are not supported");
    FftPlan pw, ph;
    CHK(make_plan(c, W, &pw));
    CHK(make_plan(c, H, &ph));
    const int conj_io = dir > 0 ? 1 : 0;
    const float scale2 = dir > 0 ? (float)(1.0 / ((double)H * (double)W)) : 1.0f;
    const long nrows = batch * H;
    const int PW = fast_p(W), PH = fast_p(H);
    if (PW) {
        const size_t lds = sizeof(float2) * ((size_t)16 * (PW * 17 + 1) + W);
        hipLaunchKernelGGL(k_fft_rows_fast, dim3((unsigned)((nrows + 15) / 16)), dim3(OCL_NT), lds, c->stream,
                           (const float2*)in, (float2*)out, nrows, PW, pw.tw, conj_io, 1.0f);
        HIPCHK(hipGetLastError());
    }
    if (PH) {
        if (!PW) CHK(fft_rows_generic(c, in, out, nrows, W, pw, conj_io));
        const size_t lds = sizeof(float2) * ((size_t)PH * FAST_COLS_KS + H);
        hipLaunchKernelGGL(k_fft_cols_fast, dim3((unsigned)((W + 15) / 16), (unsigned)batch), dim3(OCL_NT), lds,
                           c->stream, (const float2*)out, (float2*)out, W, PH, ph.tw, conj_io, scale2);
        HIPCHK(hipGetLastError());
        return OCL_OK;
    }
    if (!PW) CHK(fft_rows_generic(c, in, out, nrows, W, pw, conj_io));
    return fft_cols_generic(c, out, batch, H, W, ph, conj_io, scale2);
}

static int fft_rows_generic(ocl_ctx* c, const void* in, void* out, long nrows, int W, const FftPlan& pw,
                            int conj_io) {
    int RB = std::min(16, pow2_floor(std::max(1, 2048 / W)));
    while (RB > 1 && sizeof(float2) * ((size_t)2 * RB * (W + 1) + W) > LDS_MAX) RB >>= 1;
    size_t lds = sizeof(float2) * ((size_t)2 * RB * (W + 1) + W);
    if (lds > LDS_MAX) return fail(OCL_E_UNSUPPORTED, "FFT: row length too large for one LDS tile");
    hipLaunchKernelGGL(k_fft_rows, dim3((unsigned)((nrows + RB - 1) / RB)), dim3(OCL_NT), lds, c->stream,
                       (const float2*)in, (float2*)out, nrows, pw, RB, conj_io, 1.0f);
    HIPCHK(hipGetLastError());
    return OCL_OK;
}

static int fft_cols_generic(ocl_ctx* c, void* out, long batch, int H, int W, const FftPlan& ph, int conj_io,
                            float scale2) {
    int TW = 16;
    while (TW > 1 && sizeof(float2) * ((size_t)2 * H * (TW + 1) + H) > LDS_MAX) TW >>= 1;
    const size_t lds = sizeof(float2) * ((size_t)2 * H * (TW + 1) + H);
    if (lds > LDS_MAX) return fail(OCL_E_UNSUPPORTED, "FFT: column length too large for one LDS tile");
    hipLaunchKernelGGL(k_fft_cols, dim3((unsigned)((W + TW - 1) / TW), (unsigned)batch), dim3(OCL_NT), lds,
                       c->stream, (const float2*)out, (float2*)out, H, W, ph, TW, conj_io, scale2);
    HIPCHK(hipGetLastError());
    return OCL_OK;
}

extern "C" int ocl_fft2d(ocl_ctx* c, const void* in, void* out, long batch, int H, int W, int dir) {
    CTXCHK(c);
    if (dir != -1 && dir != 1) return fail(OCL_E_INVALID, "FFT: direction must be -1 (forward) or +1 (backward)");
    return fft2d(c, in, out, batch, H, W, dir);
}

extern "C" int ocl_simple_mri_recon(ocl_ctx* c, void* y, const void* S, void* x, int F, int C, int H, int W) {
    CTXCHK(c);
    const long N = (long)H * W;
    CHK(fft2d(c, y, y, (long)F * C, H, W, +1));
    CHK(ocl_cep(c, y, S, y, F, C, N, 1, 1));
    CHK(ocl_ximage_sum(c, y, x, F, C, N));
    return OCL_OK;
}

// ---------------------------------------------------------------------------------------------------------------------
// myNormest (NestaUp.cpp:291-439) on the device, FP32 vectors and float scalars as in the reference
// ---------------------------------------------------------------------------------------------------------------------
static const unsigned long long DRAND48_X0 = 0x1234ABCD330EULL;

static int normest_compute(ocl_ctx* c, int F, long N, double* out, int* iters) {
    const long n = (long)F * N;
    float2 *x = nullptr, *Sx = nullptr;
    HIPCHK(hipMalloc(&x, sizeof(float2) * n));
    if (hipMalloc(&Sx, sizeof(float2) * n) != hipSuccess) {
        hipFree(x);
        return fail(OCL_E_NOMEM, "normest: out of device memory");
    }
    int rc = OCL_OK;
    double r[2];
    const float tol = 1e-3f;
    unsigned cnt = 0;
    long drand_pos = 0;
    float e = 0.f, e0 = 0.f;
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(OCL_NT), 0, c->stream, x, n, make_float2(1.f, 0.f));
    if ((rc = ocl_reduce(c, x, nullptr, n, RED_NRM2, r)) != OCL_OK) goto done;
    e = (float)r[0];
    if ((rc = ocl_scal(c, x, n, 1.0f / e, 0.f)) != OCL_OK) goto done;
    while ((std::abs(e - e0) > (tol * e)) && (cnt < 100u)) {
        e0 = e;
        if ((rc = ocl_ttv(c, x, Sx, F, N, 1)) != OCL_OK) goto done;
        if ((rc = ocl_reduce(c, Sx, nullptr, n, RED_ASUM, r)) != OCL_OK) goto done;
        if ((float)r[0] == 0.f) {
            const long chunks = (n + 63) / 64;
            hipLaunchKernelGGL(k_drand48_fill, dim3((unsigned)((chunks + OCL_NT - 1) / OCL_NT)), dim3(OCL_NT), 0,
                               c->stream, Sx, n, DRAND48_X0, drand_pos);
            drand_pos += n;
        }
        if ((rc = ocl_reduce(c, Sx, nullptr, n, RED_NRM2, r)) != OCL_OK) goto done;
        e = (float)r[0];
        if ((rc = ocl_ttv(c, Sx, x, F, N, 0)) != OCL_OK) goto done;
        if ((rc = ocl_reduce(c, x, nullptr, n, RED_NRM2, r)) != OCL_OK) goto done;
        if ((rc = ocl_scal(c, x, n, 1.0f / (float)r[0], 0.f)) != OCL_OK) goto done;
        ++cnt;
    }
    *out = e;
    *iters = (int)cnt;
done:
    hipStreamSynchronize(c->stream);
    hipFree(x);
    hipFree(Sx);
    return rc;
}

extern "C" int ocl_normest(ocl_ctx* c, int F, long N, double* normU, int* iters) {
    CTXCHK(c);
    if (F < 1 || N < 1 || !normU) return fail(OCL_E_INVALID, "normest: bad arguments");
    {
        std::lock_guard<std::mutex> lk(c->mu);
        auto it = c->normU_cache.find({F, N});
        if (it != c->normU_cache.end()) {
            *normU = it->second.first;
            if (iters) *iters = it->second.second;
            return OCL_OK;
        }
    }
    double v = 0;
    int it = 0;
    CHK(normest_compute(c, F, N, &v, &it));
    std::lock_guard<std::mutex> lk(c->mu);
    c->normU_cache[{F, N}] = {v, it};
    *normU = v;
    if (iters) *iters = it;
    return OCL_OK;
}

// ---------------------------------------------------------------------------------------------------------------------
// NESTA — PLAIN form: the reference's op sequence, float scalars, three host reads per iteration
// ---------------------------------------------------------------------------------------------------------------------
static int nesta_plain(ocl_ctx* c, const void* y, const void* S, const uint8_t* M, void* xref, int F, int C, int H,
                       int W, const ocl_nesta_params& prm, ocl_nesta_report* rep) {
    const long N = (long)H * W, nX = (long)F * N, nK = (long)F * C * N;
    std::vector<void*> bufs;
    auto alloc = [&](long n) -> float2* {
        void* p = nullptr;
        if (hipMalloc(&p, sizeof(float2) * std::max(1L, n)) != hipSuccess) return nullptr;
        bufs.push_back(p);
        return (float2*)p;
    };
    struct Cleanup {
        std::vector<void*>* b;
        hipStream_t s;
        ~Cleanup() {
            hipStreamSynchronize(s);
            for (void* p : *b) hipFree(p);
        }
    } cleanup{&bufs, c->stream};
    float2 *yS = alloc(nK), *aux = alloc(nK), *Res = alloc(nK), *AuxRes = alloc(nK);
    float2 *Ux = alloc(nX), *Wk = alloc(nX), *Xk = alloc(nX), *Yk = alloc(nX), *Zk = alloc(nX), *Uk = alloc(nX),
           *AuxFx = alloc(nX), *Df = alloc(nX), *ARes = alloc(nX), *tmp = alloc(nX);
    for (void* p : bufs)
        if (!p) return fail(OCL_E_NOMEM, "NESTA plain: out of device memory");
    if (bufs.size() != 14) return fail(OCL_E_NOMEM, "NESTA plain: out of device memory");
    float2* X = (float2*)xref;

    auto opAt = [&](const float2* r, float2* xo) -> int {            // NestaUp.cpp:117-151
        CHK(fft2d(c, r, aux, (long)F * C, H, W, +1));
        CHK(ocl_cep(c, aux, S, aux, F, C, N, 1, 1));
        if (C != 1) return ocl_ximage_sum(c, aux, xo, F, C, N);
        return ocl_copy(c, aux, xo, nX);
    };
    auto opA = [&](const float2* xi, float2* out) -> int {            // NestaUp.cpp:162-190
        CHK(ocl_cep(c, xi, S, aux, F, C, N, 0, 0));
        CHK(fft2d(c, aux, out, (long)F * C, H, W, -1));
        return ocl_apply_mask(c, out, M, F, C, N);
    };
    double r[2];
    CHK(ocl_copy(c, y, yS, nK));                                      // :468-481
    const float sqrtN = std::sqrt((float)(H * W));
    CHK(ocl_scal(c, yS, nK, sqrtN, 0.f));
    CHK(opAt(yS, X));                                                 // :484-486
    CHK(ocl_ttv(c, X, Ux, F, N, 1));                                  // :491
    CHK(ocl_reduce(c, X, nullptr, nX, RED_MAXABS, r));                // :521-545
    const float maxXref = (float)r[0];
    CHK(ocl_reduce(c, Ux, nullptr, nX, RED_MAXABS, r));               // :547-568
    const float maxUXref = (float)r[0];
    const int miniter = prm.miniter;
    const float lambda = (float)prm.lambda_i * maxXref;               // :570-583
    const float muf = (float)prm.mu_f * maxXref;
    float mu0 = (float)(0.9 * (double)maxUXref);
    double nU = 0;
    int ne_it = 0;
    CHK(ocl_normest(c, F, N, &nU, &ne_it));
    const float normU = (float)nU;
    const float muL = lambda / (float)prm.La;
    if (muL > mu0) mu0 = muL;
    const float gamma = (float)std::pow((double)(muf / mu0), 1.0 / prm.maxIntIter);
    float mu = mu0;
    const float gammat = (float)std::pow((double)(float)prm.tolVar / 0.1, 1.0 / prm.maxIntIter);
    float tolVar = 0.1f;
    if (rep) {
        std::memset(rep, 0, sizeof(*rep));
        rep->normU = normU; rep->maxXref = maxXref; rep->maxUXref = maxUXref; rep->lambda = lambda;
        rep->muf = muf; rep->mu0 = mu0; rep->gamma = gamma; rep->gammat = gammat; rep->normest_iters = ne_it;
    }
    std::vector<float> fmean(std::max(miniter, 1), 0.f);
    const float invSqrtN = (float)(1.0 / std::sqrt((double)(H * W)));
    double* dres = c->red_out;   // three result slots: [0..1] ||u||, [2..3] <u,Ux>, [4..5] sum |r|^2
    for (int nl = 1; nl <= prm.maxIntIter; ++nl) {
        mu = mu * gamma;                                              // :608-609
        tolVar = tolVar * gammat;
        CHK(ocl_memset_zero(c, Wk, sizeof(float2) * nX));             // :615-641
        CHK(ocl_copy(c, X, Xk, nX));
        CHK(ocl_memset_zero(c, Yk, sizeof(float2) * nX));
        CHK(ocl_memset_zero(c, Zk, sizeof(float2) * nX));
        float Ak = 0.f;                                               // :643-649
        float Lmu = normU / mu;
        fmean[miniter - 1] = (float)LONG_MAX;
        Lmu = lambda * Lmu + (float)prm.La;
        const float Lmu1 = 1.f / Lmu;
        bool qpDone = false;
        int done_k = 0;
        for (int k = 0; k < prm.maxIter; ++k) {
            ++done_k;
            CHK(ocl_ttv(c, Xk, Uk, F, N, 1));                         // :669
            CHK(ocl_copy(c, Uk, AuxFx, nX));                          // :675
            CHK(ocl_vecnorm(c, Uk, Uk, nX, mu));                      // :680-683
            CHK(reduce_async(c, Uk, nullptr, nX, RED_NRM2, dres));    // :686
            CHK(reduce_async(c, Uk, AuxFx, nX, RED_DOTC, dres + 2));  // :691
            CHK(ocl_ttv(c, Uk, tmp, F, N, 0));                        // :696
            CHK(ocl_copy(c, tmp, Df, nX));
            CHK(ocl_scal(c, Df, nX, -1.f, 0.f));                      // :698-704
            CHK(opA(Xk, Res));                                        // :710
            CHK(ocl_scal(c, Res, nK, invSqrtN, 0.f));                 // :712-718
            CHK(ocl_axpy(c, nK, -1.f, 0.f, y, Res));                  // :720-727
            CHK(ocl_copy(c, Res, AuxRes, nK));                        // :732
            CHK(ocl_complex_abs(c, Res, Res, nK, 1));                 // :738-741
            CHK(reduce_async(c, Res, nullptr, nK, RED_SUM, dres + 4));// :744
            CHK(ocl_scal(c, AuxRes, nK, sqrtN, 0.f));                 // :749-753
            CHK(opAt(AuxRes, ARes));                                  // :756
            CHK(ocl_scal(c, Df, nX, lambda, 0.f));                    // :759-764
            CHK(ocl_axpy(c, nX, 1.f, 0.f, ARes, Df));                 // :766-772
            CHK(ocl_copy(c, Xk, Yk, nX));                             // :776-779
            CHK(ocl_axpy(c, nX, -Lmu1, 0.f, Df, Yk));                 // :781-786
            HIPCHK(hipMemcpyAsync(c->red_host, dres, 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));                  // :791-795
            const float normUk = (float)c->red_host[0];
            float fx = (float)c->red_host[2];
            const float l2 = (float)c->red_host[4];
            const float l1 = (float)(fx - (mu / 2.0f) * std::pow((double)normUk, 2));
            fx = 0.5f * l2 + lambda * l1;                             // :797-798
            if (rep && rep->ntrace < OCL_TRACE) {
                rep->trace_f[rep->ntrace] = fx; rep->trace_l2[rep->ntrace] = l2; rep->trace_l1[rep->ntrace] = l1;
                rep->ntrace++;
            }
            if (prm.stoptest == 1) {                                  // :799-824
                float mean = 0.f;
                bool skip_write = false;
                if (k >= miniter) {
                    for (int ne = 0; ne < miniter; ++ne) mean += fmean[miniter - ne - 1];
                    mean = mean / (float)miniter;
                    const float qp = (mean - fx) / mean;
                    if (qp <= tolVar) {
                        if (qpDone) k = prm.maxIter;
                        qpDone = true;
                        skip_write = true;
                    }
                }
                if (!skip_write) fmean[(miniter - 1) - ((k + 1) % miniter)] = fx;
            }
            const float apk = 0.5f * ((float)k + 1.0f);               // :829-831
            Ak = Ak + apk;
            const float tauk = 2.0f / ((float)k + 3.0f);
            CHK(ocl_axpy(c, nX, apk, 0.f, Df, Wk));                   // :837
            CHK(ocl_copy(c, X, Zk, nX));                              // :843
            CHK(ocl_axpy(c, nX, -Lmu1, 0.f, Wk, Zk));                 // :850
            CHK(ocl_copy(c, Zk, Xk, nX));                             // :858
            CHK(ocl_scal(c, Xk, nX, tauk, 0.f));                      // :865
            CHK(ocl_axpy(c, nX, 1.0f - tauk, 0.f, Yk, Xk));           // :873
        }
        (void)Ak;
        CHK(ocl_copy(c, Xk, X, nX));                                  // :889
        if (rep && nl <= OCL_MAX_STAGES) { rep->iters[nl - 1] = done_k; rep->stages = nl; }
    }
    if (rep) { rep->final_tolVar = tolVar; rep->final_mu = mu; }
    HIPCHK(hipStreamSynchronize(c->stream));
    return OCL_OK;
}

// ---------------------------------------------------------------------------------------------------------------------
// batched / fused NESTA
// ---------------------------------------------------------------------------------------------------------------------
struct SliceSetup {
    double maxX, maxUX, normU, lambda, muf, mu0, gamma, gammat;
    int ne_iters;
};

struct ocl_batch {
    ocl_ctx* ctx = nullptr;
    std::vector<ocl_slice> sl;
    ocl_nesta_params prm{};
    std::vector<SliceDesc> hd;
    std::vector<SliceSetup> setup;
    std::vector<int> offB, offCA;
    int totB = 0, totCA = 0;
    size_t ldsB = 0, ldsCA = 0;
    int maxe = 4;
    NestaLaunch L{};
    SliceDesc* dd = nullptr;
    int* d_offB = nullptr;
    int* d_offCA = nullptr;
    float2** d_outs = nullptr;
    NState* h_st = nullptr;            // pinned mirror of the states
    std::vector<void*> allocs;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_iters = 0;
    bool ready = false;
    bool allocated = false;
};

static int batch_alloc(ocl_batch* b, size_t bytes, void** p) {
    hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 16));
    if (e == hipErrorOutOfMemory) return fail(OCL_E_NOMEM, "batch: out of device memory");
    HIPCHK(e);
    b->allocs.push_back(*p);
    return OCL_OK;
}

extern "C" int ocl_batch_destroy(ocl_batch* b) {
    if (!b) return OCL_OK;
    hipSetDevice(b->ctx->device);
    hipStreamSynchronize(b->ctx->stream);
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->graph) hipGraphDestroy(b->graph);
    for (void* p : b->allocs) hipFree(p);
    if (b->h_st) hipHostFree(b->h_st);
    delete b;
    return OCL_OK;
}

extern "C" int ocl_batch_create(ocl_ctx* c, int nslices, const ocl_slice* slices, const ocl_nesta_params* prm,
                                ocl_batch** out) {
    CTXCHK(c);
    if (nslices < 1 || !slices || !prm || !out) return fail(OCL_E_INVALID, "batch: bad arguments");
    if (prm->miniter < 1 || prm->miniter > OCL_MAX_MINITER)
        return fail(OCL_E_INVALID, "NESTA: miniter must be in [1, 32]");
    if (prm->maxIntIter < 0 || prm->maxIter < 0) return fail(OCL_E_INVALID, "NESTA: negative iteration counts");
    for (int i = 0; i < nslices; ++i) {
        const ocl_slice& s = slices[i];
        if (s.F < 1 || s.C < 1 || s.H < 1 || s.W < 1 || !s.y || !s.S || !s.M || !s.x_out)
            return fail(OCL_E_INVALID, "batch: slice " + std::to_string(i) + " has bad dimensions or null pointers");
        if (s.H > 512 || s.W > 1024)
            return fail(OCL_E_UNSUPPORTED, "batch: fused NESTA supports H <= 512 and W <= 1024");
    }
    ocl_batch* b = new ocl_batch();
    b->ctx = c;
    b->sl.assign(slices, slices + nslices);
    b->prm = *prm;
    *out = b;
    return OCL_OK;
}

static int choose_tiles(const ocl_batch* b, const ocl_slice& s, int* tile_cols, int* rows_per_wg) {
    long total_cols = 0;
    for (const auto& t : b->sl) total_cols += (long)t.F * t.W;
    int tc = (total_cols / 16 >= 2048) ? 16 : 8;
    while (tc > 1 && ((long)s.H * tc > (long)OCL_MAXE * OCL_NT ||
                      sizeof(float2) * ((size_t)2 * s.H * (tc + 1) + s.H) > LDS_MAX))
        tc >>= 1;
    *tile_cols = tc;
    *rows_per_wg = std::min(16, pow2_floor(std::max(1, 2048 / s.W)));
    return OCL_OK;
}

static int batch_compute_setup(ocl_batch* b);

extern "C" int ocl_batch_setup(ocl_batch* b) {
    if (!b) return fail(OCL_E_INVALID, "null batch");
    if (b->allocated) return batch_compute_setup(b);
    ocl_ctx* c = b->ctx;
    HIPCHK(hipSetDevice(c->device));
    const int ns = (int)b->sl.size();
    b->hd.assign(ns, SliceDesc{});
    b->setup.assign(ns, SliceSetup{});
    b->offB.assign(ns + 1, 0);
    b->offCA.assign(ns + 1, 0);
    b->maxe = 4;
    b->ldsB = b->ldsCA = 0;
    uint8_t* flags_d = nullptr;
    size_t flags_cap = 0;
    std::vector<uint8_t> flags;
    for (int i = 0; i < ns; ++i) {
        const ocl_slice& s = b->sl[i];
        SliceDesc& d = b->hd[i];
        d.F = s.F; d.C = s.C; d.H = s.H; d.W = s.W;
        d.N = (long)s.H * s.W;
        CHK(make_plan(c, s.H, &d.ph));
        CHK(make_plan(c, s.W, &d.pw));
        d.fastH = fast_p(s.H);
        d.fastW = fast_p(s.W);
        CHK(choose_tiles(b, s, &d.tile_cols, &d.rows_per_wg));
        if (d.fastH) d.tile_cols = 16;
        if (d.fastW) d.rows_per_wg = 16;
        // sampled-row lists
        const size_t fb = (size_t)s.F * s.H;
        if (fb > flags_cap) {
            if (flags_d) hipFree(flags_d);
            HIPCHK(hipMalloc(&flags_d, fb));
            flags_cap = fb;
        }
        hipLaunchKernelGGL(k_rowflags, dim3((unsigned)fb), dim3(OCL_NT), 0, c->stream, (const float2*)s.y, s.M, s.F,
                           s.C, s.H, s.W, flags_d);
        HIPCHK(hipGetLastError());
        flags.resize(fb);
        HIPCHK(hipMemcpyAsync(flags.data(), flags_d, fb, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        std::vector<int> nrows(s.F, 0);
        int nmax = 1;
        for (int f = 0; f < s.F; ++f) {
            for (int h = 0; h < s.H; ++h) nrows[f] += flags[(size_t)f * s.H + h];
            nmax = std::max(nmax, nrows[f]);
        }
        std::vector<int16_t> rows((size_t)s.F * nmax, 0), rowpos((size_t)s.F * s.H, -1);
        for (int f = 0; f < s.F; ++f) {
            int j = 0;
            for (int h = 0; h < s.H; ++h)
                if (flags[(size_t)f * s.H + h]) {
                    rowpos[(size_t)f * s.H + h] = (int16_t)j;
                    rows[(size_t)f * nmax + j++] = (int16_t)h;
                }
        }
        d.nmax = nmax;
        d.ncol_tiles = (s.W + d.tile_cols - 1) / d.tile_cols;
        d.nrow_groups = (nmax + d.rows_per_wg - 1) / d.rows_per_wg;
        d.wg_B = s.F * s.C * d.nrow_groups;
        d.wg_CA = s.F * d.ncol_tiles;
        d.u_chunk = (int)std::min<long>(d.N, 4096);
        d.wg_U = s.F * (int)((d.N + d.u_chunk - 1) / d.u_chunk);
        d.inv_sqrtN = (float)(1.0 / std::sqrt((double)d.N));
        d.y = (const float2*)s.y;
        d.S = (const float2*)s.S;
        d.M = s.M;
        const size_t xb = sizeof(float2) * (size_t)d.F * d.N;
        void* p;
        CHK(batch_alloc(b, xb, &p)); d.x0 = (float2*)p;
        CHK(batch_alloc(b, xb, &p)); d.x1 = (float2*)p;
        CHK(batch_alloc(b, xb, &p)); d.xref = (float2*)p;
        CHK(batch_alloc(b, xb, &p)); d.Wk = (float2*)p;
        CHK(batch_alloc(b, sizeof(float2) * (size_t)d.F * d.C * nmax * d.W, &p)); d.kbuf = (float2*)p;
        CHK(batch_alloc(b, sizeof(int16_t) * rows.size(), &p)); d.rows = (const int16_t*)p;
        HIPCHK(hipMemcpyAsync(p, rows.data(), sizeof(int16_t) * rows.size(), hipMemcpyHostToDevice, c->stream));
        CHK(batch_alloc(b, sizeof(int16_t) * rowpos.size(), &p)); d.rowpos = (const int16_t*)p;
        HIPCHK(hipMemcpyAsync(p, rowpos.data(), sizeof(int16_t) * rowpos.size(), hipMemcpyHostToDevice, c->stream));
        CHK(batch_alloc(b, sizeof(int) * s.F, &p)); d.nrows = (const int*)p;
        HIPCHK(hipMemcpyAsync(p, nrows.data(), sizeof(int) * s.F, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));   // rows / nrows are host temporaries
        CHK(batch_alloc(b, sizeof(NState), &p)); d.st = (NState*)p;
        CHK(batch_alloc(b, sizeof(double) * d.wg_B, &p)); d.partB = (double*)p;
        CHK(batch_alloc(b, sizeof(double) * 2 * d.wg_U, &p)); d.partU = (double*)p;
        CHK(batch_alloc(b, sizeof(unsigned) * 4, &p)); d.counter = (unsigned*)p;
        CHK(ocl_memset_zero(c, p, sizeof(unsigned) * 4));
        CHK(batch_alloc(b, sizeof(double) * 3 * OCL_TRACE, &p)); d.trace = (double*)p;
        CHK(batch_alloc(b, sizeof(double) * 2 * d.wg_CA, &p)); d.maxpart = (double*)p;
        b->offB[i + 1] = b->offB[i] + d.wg_B + d.wg_U;
        b->offCA[i + 1] = b->offCA[i] + d.wg_CA;
        b->ldsB = std::max(b->ldsB, d.fastW ? sizeof(float2) * fast_b_lds(d.fastW)
                                            : sizeof(float2) * ((size_t)2 * d.rows_per_wg * (d.W + 1) + d.W));
        b->ldsCA = std::max(b->ldsCA, d.fastH ? sizeof(float2) * fast_ca_lds(d.fastH)
                                              : sizeof(float2) * ((size_t)2 * d.H * (d.tile_cols + 1) + d.H));
        if (!d.fastH) {
            const int need = (d.H * d.tile_cols + OCL_NT - 1) / OCL_NT;
            while (b->maxe < need) b->maxe *= 2;
        }
    }
    if (flags_d) {
        hipStreamSynchronize(c->stream);
        hipFree(flags_d);
    }
    if (b->ldsB > LDS_MAX || b->ldsCA > LDS_MAX) return fail(OCL_E_UNSUPPORTED, "batch: LDS tile too large");
    b->totB = b->offB[ns];
    b->totCA = b->offCA[ns];
    void* p;
    CHK(batch_alloc(b, sizeof(SliceDesc) * ns, &p)); b->dd = (SliceDesc*)p;
    CHK(batch_alloc(b, sizeof(int) * (ns + 1), &p)); b->d_offB = (int*)p;
    CHK(batch_alloc(b, sizeof(int) * (ns + 1), &p)); b->d_offCA = (int*)p;
    CHK(batch_alloc(b, sizeof(float2*) * ns, &p)); b->d_outs = (float2**)p;
    std::vector<float2*> outs(ns);
    for (int i = 0; i < ns; ++i) outs[i] = (float2*)b->sl[i].x_out;
    HIPCHK(hipMemcpyAsync(b->dd, b->hd.data(), sizeof(SliceDesc) * ns, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(b->d_offB, b->offB.data(), sizeof(int) * (ns + 1), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(b->d_offCA, b->offCA.data(), sizeof(int) * (ns + 1), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(b->d_outs, outs.data(), sizeof(float2*) * ns, hipMemcpyHostToDevice, c->stream));
    if (!b->h_st) HIPCHK(hipHostMalloc(&b->h_st, sizeof(NState) * ns, hipHostMallocDefault));
    int PH = b->hd[0].fastH, PW = b->hd[0].fastW;
    for (int i = 1; i < ns; ++i) {
        if (b->hd[i].fastH != PH) PH = 0;
        if (b->hd[i].fastW != PW) PW = 0;
    }
    b->L = NestaLaunch{b->dd, b->d_offB, b->d_offCA, ns, b->totB, b->totCA, b->ldsB, b->ldsCA, b->maxe, PH, PW};
    HIPCHK(hipStreamSynchronize(c->stream));
    b->allocated = true;
    return batch_compute_setup(b);
}

// Per-reconstruction setup on already allocated buffers: E^H y, maxima, normU, initial state (NestaUp.cpp:442-649).
// The sampled-row lists are those found when the batch was allocated.
static int batch_compute_setup(ocl_batch* b) {
    ocl_ctx* c = b->ctx;
    HIPCHK(hipSetDevice(c->device));
    const int ns = (int)b->sl.size();
    for (int i = 0; i < ns; ++i) CHK(ocl_memset_zero(c, b->hd[i].counter, sizeof(unsigned) * 4));

    // x_ref = A^H y (NestaUp.cpp:468-486) and the forward columns of iteration 0
    HIPCHK(launch_nesta_B(b->L, c->stream, 1));
    HIPCHK(launch_nesta_CA(b->L, c->stream, 1));
    hipLaunchKernelGGL(k_setup_max, dim3(b->totCA), dim3(OCL_NT), 0, c->stream, (const SliceDesc*)b->dd,
                       (const int*)b->d_offCA, ns, (double*)nullptr);
    HIPCHK(hipGetLastError());
    std::vector<double> mp;
    std::vector<NState> st(ns);
    for (int i = 0; i < ns; ++i) {
        const SliceDesc& d = b->hd[i];
        mp.resize(2 * d.wg_CA);
        HIPCHK(hipMemcpyAsync(mp.data(), d.maxpart, sizeof(double) * 2 * d.wg_CA, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        double m0 = (double)LONG_MIN, m1 = (double)LONG_MIN;
        for (int q = 0; q < d.wg_CA; ++q) { m0 = std::max(m0, mp[2 * q]); m1 = std::max(m1, mp[2 * q + 1]); }
        SliceSetup& u = b->setup[i];
        u.maxX = (float)m0;
        u.maxUX = (float)m1;
        CHK(ocl_normest(c, d.F, d.N, &u.normU, &u.ne_iters));
        // NestaUp.cpp:570-583 (scalars kept in double)
        const ocl_nesta_params& P = b->prm;
        u.lambda = P.lambda_i * u.maxX;
        u.muf = P.mu_f * u.maxX;
        u.mu0 = 0.9 * u.maxUX;
        const double muL = u.lambda / P.La;
        if (muL > u.mu0) u.mu0 = muL;
        u.gamma = std::pow(u.muf / u.mu0, 1.0 / std::max(1, P.maxIntIter));
        u.gammat = std::pow(P.tolVar / 0.1, 1.0 / std::max(1, P.maxIntIter));
        NState& S = st[i];
        std::memset(&S, 0, sizeof(S));
        S.lambda = u.lambda; S.La = P.La; S.normU = u.normU; S.gamma = u.gamma; S.gammat = u.gammat;
        S.maxIntIter = P.maxIntIter; S.maxIter = P.maxIter; S.stoptest = P.stoptest; S.miniter = P.miniter;
        S.active = (P.maxIntIter >= 1 && P.maxIter >= 1) ? 1 : 0;
        S.stage = 1;
        S.k = 0;
        S.stage_start = 1;
        S.cur = 0;
        S.qpDone = 0;
        S.trace_cap = OCL_TRACE;
        S.mu = u.mu0 * u.gamma;
        S.tolVar = 0.1 * u.gammat;
        S.Lmu = u.lambda * (u.normU / S.mu) + P.La;
        S.Lmu1 = 1.0 / S.Lmu;
        S.fmean[P.miniter - 1] = 9223372036854775807.0;
        HIPCHK(hipMemcpyAsync(d.st, &S, sizeof(NState), hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    b->ready = true;
    return OCL_OK;
}

static int batch_enqueue(ocl_batch* b, int n_iters) {
    ocl_ctx* c = b->ctx;
    const int ns = (int)b->sl.size();
    for (int it = 0; it < n_iters; ++it) {
        HIPCHK(launch_nesta_B(b->L, c->stream, 0));
        HIPCHK(launch_nesta_S(b->L, c->stream));
        HIPCHK(launch_nesta_CA(b->L, c->stream, 0));
    }
    HIPCHK(hipGetLastError());
    return OCL_OK;
}

extern "C" int ocl_batch_time(ocl_batch* b, int n_iters, double* ms_B, double* ms_CA, double* ms_total) {
    if (!b || !b->ready) return fail(OCL_E_INVALID, "batch: not set up");
    if (n_iters < 1) return fail(OCL_E_INVALID, "batch_time: need at least one iteration");
    ocl_ctx* c = b->ctx;
    HIPCHK(hipSetDevice(c->device));
    const int ns = (int)b->sl.size();
    std::vector<hipEvent_t> ev(3 * (size_t)n_iters);
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    for (int it = 0; it < n_iters; ++it) {
        HIPCHK(hipEventRecord(ev[3 * it], c->stream));
        HIPCHK(launch_nesta_B(b->L, c->stream, 0));
        HIPCHK(launch_nesta_S(b->L, c->stream));
        HIPCHK(hipEventRecord(ev[3 * it + 1], c->stream));
        HIPCHK(launch_nesta_CA(b->L, c->stream, 0));
        HIPCHK(hipEventRecord(ev[3 * it + 2], c->stream));
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    double tb = 0, tc = 0;
    for (int it = 0; it < n_iters; ++it) {
        float a = 0, q = 0;
        HIPCHK(hipEventElapsedTime(&a, ev[3 * it], ev[3 * it + 1]));
        HIPCHK(hipEventElapsedTime(&q, ev[3 * it + 1], ev[3 * it + 2]));
        tb += a;
        tc += q;
    }
    float tot = 0;
    HIPCHK(hipEventElapsedTime(&tot, ev[0], ev[3 * (size_t)n_iters - 1]));
    for (auto& e : ev) hipEventDestroy(e);
    if (ms_B) *ms_B = tb / n_iters;
    if (ms_CA) *ms_CA = tc / n_iters;
    if (ms_total) *ms_total = tot;
    return OCL_OK;
}

extern "C" int ocl_batch_iterate(ocl_batch* b, int n_iters) {
    if (!b || !b->ready) return fail(OCL_E_INVALID, "batch: not set up");
    if (n_iters < 0) return fail(OCL_E_INVALID, "batch: negative iteration count");
    HIPCHK(hipSetDevice(b->ctx->device));
    return batch_enqueue(b, n_iters);
}

extern "C" int ocl_batch_capture(ocl_batch* b, int n_iters) {
    if (!b || !b->ready) return fail(OCL_E_INVALID, "batch: not set up");
    if (n_iters < 1) return fail(OCL_E_INVALID, "batch: capture needs at least one iteration");
    HIPCHK(hipSetDevice(b->ctx->device));
    if (b->gexec) { hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
    HIPCHK(hipStreamBeginCapture(b->ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = batch_enqueue(b, n_iters);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->ctx->stream, &g);
    if (rc != OCL_OK) { if (g) hipGraphDestroy(g); return rc; }
    HIPCHK(e);
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    b->graph_iters = n_iters;
    return OCL_OK;
}

extern "C" int ocl_batch_replay(ocl_batch* b) {
    if (!b || !b->gexec) return fail(OCL_E_INVALID, "batch: no captured graph");
    HIPCHK(hipGraphLaunch(b->gexec, b->ctx->stream));
    return OCL_OK;
}

static int batch_pull_states(ocl_batch* b) {
    const int ns = (int)b->sl.size();
    for (int i = 0; i < ns; ++i)
        HIPCHK(hipMemcpyAsync(&b->h_st[i], b->hd[i].st, sizeof(NState), hipMemcpyDeviceToHost, b->ctx->stream));
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    return OCL_OK;
}

extern "C" int ocl_batch_active(ocl_batch* b, int* n_active) {
    if (!b || !b->ready || !n_active) return fail(OCL_E_INVALID, "batch: not set up");
    CHK(batch_pull_states(b));
    int n = 0;
    for (size_t i = 0; i < b->sl.size(); ++i) n += b->h_st[i].active ? 1 : 0;
    *n_active = n;
    return OCL_OK;
}

extern "C" int ocl_batch_finalize(ocl_batch* b) {
    if (!b || !b->ready) return fail(OCL_E_INVALID, "batch: not set up");
    const int ns = (int)b->sl.size();
    hipLaunchKernelGGL(k_finalize, dim3(256, ns), dim3(OCL_NT), 0, b->ctx->stream, (const SliceDesc*)b->dd, ns,
                       (float2* const*)b->d_outs);
    HIPCHK(hipGetLastError());
    return OCL_OK;
}

extern "C" int ocl_batch_report(ocl_batch* b, int slice, ocl_nesta_report* rep) {
    if (!b || !b->ready || !rep || slice < 0 || slice >= (int)b->sl.size())
        return fail(OCL_E_INVALID, "batch_report: bad arguments");
    CHK(batch_pull_states(b));
    const NState& S = b->h_st[slice];
    const SliceSetup& u = b->setup[slice];
    std::memset(rep, 0, sizeof(*rep));
    rep->normU = u.normU; rep->maxXref = u.maxX; rep->maxUXref = u.maxUX; rep->lambda = u.lambda;
    rep->muf = u.muf; rep->mu0 = u.mu0; rep->gamma = u.gamma; rep->gammat = u.gammat;
    rep->normest_iters = u.ne_iters;
    rep->final_tolVar = S.tolVar;
    rep->final_mu = S.mu;
    const int stages_done = S.stage - 1 + (S.active ? 1 : 0);
    rep->stages = std::min(stages_done, OCL_MAX_STAGES);
    for (int q = 0; q < rep->stages; ++q) rep->iters[q] = (q == S.stage - 1 && S.active) ? S.iters_stage : S.iters[q];
    rep->ntrace = std::min(S.ntrace, OCL_TRACE);
    std::vector<double> tr(3 * (size_t)rep->ntrace);
    if (rep->ntrace) {
        HIPCHK(hipMemcpy(tr.data(), b->hd[slice].trace, sizeof(double) * tr.size(), hipMemcpyDeviceToHost));
        for (int q = 0; q < rep->ntrace; ++q) {
            rep->trace_f[q] = tr[3 * q];
            rep->trace_l2[q] = tr[3 * q + 1];
            rep->trace_l1[q] = tr[3 * q + 2];
        }
    }
    return OCL_OK;
}

extern "C" int ocl_batch_run(ocl_batch* b, int chunk) {
    if (!b) return fail(OCL_E_INVALID, "null batch");
    if (!b->ready) CHK(ocl_batch_setup(b));
    const long max_total = (long)std::max(0, b->prm.maxIntIter) * std::max(0, b->prm.maxIter);
    if (b->prm.stoptest != 1) {
        CHK(batch_enqueue(b, (int)max_total));
    } else {
        if (chunk < 1) chunk = 16;
        long done = 0;
        int n_active = 1;
        while (done < max_total && n_active > 0) {
            const int step = (int)std::min<long>(chunk, max_total - done);
            CHK(batch_enqueue(b, step));
            done += step;
            CHK(ocl_batch_active(b, &n_active));
        }
    }
    CHK(ocl_batch_finalize(b));
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    return OCL_OK;
}

// ---------------------------------------------------------------------------------------------------------------------
// NestaUp::launch entry point
// ---------------------------------------------------------------------------------------------------------------------
extern "C" int ocl_nesta_run(ocl_ctx* c, const void* y, const void* S, const uint8_t* M, void* x_out, int F, int C,
                             int H, int W, const ocl_nesta_params* prm, int form, ocl_nesta_report* rep) {
    CTXCHK(c);
    if (!prm) return fail(OCL_E_INVALID, "NESTA: null launch parameters");
    if (!y || !x_out || F < 1 || C < 1 || H < 1 || W < 1) return fail(OCL_E_INVALID, "NESTA: bad input/output");
    if (!S) return fail(OCL_E_INVALID, "NESTA: non-existing SensitivityMaps");
    if (!M) return fail(OCL_E_INVALID, "NESTA: null sampling masks");
    if (prm->miniter < 1 || prm->miniter > OCL_MAX_MINITER) return fail(OCL_E_INVALID, "NESTA: miniter out of range");
    HIPCHK(hipSetDevice(c->device));
    if (form == OCL_FORM_PLAIN) return nesta_plain(c, y, S, M, x_out, F, C, H, W, *prm, rep);
    if (form != OCL_FORM_FUSED) return fail(OCL_E_INVALID, "NESTA: unknown form");
    ocl_slice s{y, S, M, x_out, F, C, H, W};
    ocl_batch* b = nullptr;
    CHK(ocl_batch_create(c, 1, &s, prm, &b));
    int rc = ocl_batch_run(b, 16);
    if (rc == OCL_OK && rep) rc = ocl_batch_report(b, 0, rep);
    ocl_batch_destroy(b);
    return rc;
}
