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
#include <tuple>
#include <vector>

#include "ocl_internal.hpp"

static thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

extern "C" const char* ocl_last_error(void) { return g_err.c_str(); }

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
    // coherent: kernel writes reach host memory without a cache flush, kernel reads see what the host wrote
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->zc_host, ZC_BYTES, hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&c->zc_dev, c->zc_host, 0);
    if (e == hipSuccess && scratch_init_device(device) != OCL_OK) {
        const std::string m = ocl_last_error();
        ocl_ctx_destroy(c);
        return fail(OCL_E_HIP, m);
    }
    const void* lds_kernels[] = {(const void*)k_fft_rows, (const void*)k_fft_cols, (const void*)k_fft_rows_fast,
                                 (const void*)k_fft_cols_fast};
    if (e == hipSuccess) {
        for (const void* k : lds_kernels) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_MAX);
        simple_cols_attributes((int)LDS_MAX);
        nesta_kernel_attributes((int)LDS_MAX_NESTA);
        mfma_kernel_attributes((int)LDS_MAX_NESTA);
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
    if (c->zc_host) hipHostFree(c->zc_host);
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
extern "C" int ocl_host_malloc(ocl_ctx* c, size_t bytes, void** p) {
    if (!c || !p) return fail(OCL_E_INVALID, "ocl_host_malloc: null argument");
    *p = nullptr;
    HIPCHK(hipSetDevice(c->device));
    hipError_t e = hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault);
    if (e == hipErrorOutOfMemory) return fail(OCL_E_NOMEM, "ocl_host_malloc: out of page-locked host memory");
    HIPCHK(e);
    return OCL_OK;
}
extern "C" int ocl_host_free(ocl_ctx* c, void* p) {
    if (p) {
        if (c) HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipHostFree(p));        // waits for the device itself when no context is given
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

int grid_for(long n) { return (int)std::max(1L, std::min((n + OCL_NT - 1) / OCL_NT, 8192L)); }

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
int make_plan(ocl_ctx* c, int n, FftPlan* P) {
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
    if ((int)r.size() > OCL_FFT_MAX_STAGES) return fail(OCL_E_UNSUPPORTED, "FFT length has too many factors");
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
        if (dev_scratch_active()) {
            // an allocation-free call: the table lives in the scratch array for this call only and is not cached
            HIPCHK(dev_malloc((void**)&d, sizeof(float2) * n));
            HIPCHK(hipMemcpy(d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
            P->tw = d;
            return OCL_OK;
        }
        HIPCHK(hipMalloc(&d, sizeof(float2) * n));
        HIPCHK(hipMemcpy(d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
        it = c->tw.emplace(n, d).first;
    }
    P->tw = it->second;
    return OCL_OK;
}

// P when n = 16 P takes the register four-step path (P even, 2..20), else 0
int fast_p(int n) {
    if (n % 16) return 0;
    const int P = n / 16;
    return (P >= 2 && P <= 20 && P % 2 == 0) ? P : 0;
}

int pow2_floor(int v) {
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

extern "C" int ocl_add_noise_masked(ocl_ctx* c, void* y, const uint8_t* M, int F, int C, long N, float sigma,
                                    unsigned long long seed) {
    CTXCHK(c);
    if (!y || !M || F < 1 || C < 1 || N < 1) return fail(OCL_E_INVALID, "add_noise_masked: bad arguments");
    hipLaunchKernelGGL(k_add_noise_masked, dim3(grid_for((long)F * C * N)), dim3(OCL_NT), 0, c->stream, (float2*)y, M,
                       F, C, N, sigma, seed);
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
int reduce_async(ocl_ctx* c, const void* x, const void* y, long n, int op, double* dev_out) {
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

extern "C" int ocl_rsos(ocl_ctx* c, const void* in, void* out, int F, int C, long N) {
    CTXCHK(c);
    if (!in || !out || F < 1 || C < 1 || N < 1) return fail(OCL_E_INVALID, "RSoS: bad arguments");
    hipLaunchKernelGGL(k_rsos, dim3(grid_for((long)F * N)), dim3(OCL_NT), 0, c->stream, (const float2*)in,
                       (float2*)out, F, C, N);
    HIPCHK(hipGetLastError());
    return OCL_OK;
}

extern "C" int ocl_simple_mri_recon_sos(ocl_ctx* c, void* y, void* x, int F, int C, int H, int W) {
    CTXCHK(c);
    if (!y || !x || F < 1 || C < 1) return fail(OCL_E_INVALID, "SimpleMRIReconSOS: bad arguments");
    CHK(fft2d(c, y, y, (long)F * C, H, W, +1));
    return ocl_rsos(c, y, x, F, C, (long)H * W);
}

extern "C" int ocl_simple_mri_recon(ocl_ctx* c, void* y, const void* S, void* x, int F, int C, int H, int W) {
    CTXCHK(c);
    const long N = (long)H * W;
    const int PH = fast_p(H), PW = fast_p(W);
    if (PH && PW && F >= 1 && C >= 1 && y && S && x && (long)F * ((W + 15) / 16) >= 128) {
        // two passes: inverse row DFTs in place, then the fused column DFT + conj(S) + coil sum (enough column
        // tiles x frames to fill the device; otherwise the coil-parallel column transform below)
        FftPlan pw, ph;
        CHK(make_plan(c, W, &pw));
        CHK(make_plan(c, H, &ph));
        const long nrows = (long)F * C * H;
        const size_t ldsr = sizeof(float2) * ((size_t)16 * (PW * 17 + 1) + W);
        hipLaunchKernelGGL(k_fft_rows_fast, dim3((unsigned)((nrows + 15) / 16)), dim3(OCL_NT), ldsr, c->stream,
                           (const float2*)y, (float2*)y, nrows, PW, pw.tw, 1, 1.0f);
        HIPCHK(hipGetLastError());
        const size_t ldsc = sizeof(float2) * ((size_t)PH * FAST_COLS_KS + H);
        HIPCHK(launch_simple_cols((float2*)y, (const float2*)S, (float2*)x, F, C, W, PH, ph.tw,
                                  (float)(1.0 / ((double)H * (double)W)), ldsc, c->stream));
        return OCL_OK;
    }
    CHK(fft2d(c, y, y, (long)F * C, H, W, +1));
    if (!S || !x || F < 1 || C < 1) return fail(OCL_E_INVALID, "SimpleMRIRecon: bad arguments");
    hipLaunchKernelGGL(k_cep_sum, dim3(grid_for((long)F * N)), dim3(OCL_NT), 0, c->stream, (float2*)y,
                       (const float2*)S, (float2*)x, F, C, N);
    HIPCHK(hipGetLastError());
    return OCL_OK;
}

// Many SimpleMRIRecon launches of one shape at once (no reference counterpart; each entry is what
// ocl_simple_mri_recon computes): the inverse row DFTs of every input in one launch, then the fused column DFT +
// conj(S) + coil sum of every frame of every input in another, OCL_RECON_BATCH inputs per launch pair.  Four-step
// sizes (H, W = 16 P) only; other sizes run ocl_simple_mri_recon per input.
extern "C" int ocl_simple_mri_recon_batch(ocl_ctx* c, int n, void* const* y, const void* const* S, void* const* x,
                                          int F, int C, int H, int W) {
    CTXCHK(c);
    if (n < 0 || (n > 0 && (!y || !S || !x)) || F < 1 || C < 1 || H < 1 || W < 1)
        return fail(OCL_E_INVALID, "SimpleMRIRecon batch: bad arguments");
    for (int i = 0; i < n; ++i)
        if (!y[i] || !S[i] || !x[i]) return fail(OCL_E_INVALID, "SimpleMRIRecon batch: null input or output");
    const int PH = fast_p(H), PW = fast_p(W);
    if (!PH || !PW) {
        for (int i = 0; i < n; ++i) CHK(ocl_simple_mri_recon(c, y[i], S[i], x[i], F, C, H, W));
        return OCL_OK;
    }
    FftPlan pw, ph;
    CHK(make_plan(c, W, &pw));
    CHK(make_plan(c, H, &ph));
    const long nrows = (long)F * C * H;
    const size_t ldsr = sizeof(float2) * ((size_t)16 * (PW * 17 + 1) + W);
    const size_t ldsc = sizeof(float2) * ((size_t)PH * FAST_COLS_KS + H);
    for (int i0 = 0; i0 < n; i0 += OCL_RECON_BATCH) {
        const int m = std::min(OCL_RECON_BATCH, n - i0);
        ReconPtrs p{};
        for (int j = 0; j < m; ++j) {
            p.y[j] = (float2*)y[i0 + j];
            p.S[j] = (const float2*)S[i0 + j];
            p.x[j] = (float2*)x[i0 + j];
        }
        hipLaunchKernelGGL(k_fft_rows_fast_batch, dim3((unsigned)((nrows + 15) / 16), (unsigned)m), dim3(OCL_NT),
                           ldsr, c->stream, p, nrows, PW, pw.tw, 1, 1.0f);
        HIPCHK(hipGetLastError());
        HIPCHK(launch_simple_cols_batch(p, m, F, C, W, PH, ph.tw, (float)(1.0 / ((double)H * (double)W)), ldsc,
                                        c->stream));
    }
    return OCL_OK;
}

// ---------------------------------------------------------------------------------------------------------------------
// Motion compensation (MotionCompensation.cpp, AdjointMotionCompensation.cpp; kernels in ocl_mc.hip)
// ---------------------------------------------------------------------------------------------------------------------
struct McGeom {
    int r0, nr, c0, nc;
};

// Box rows / columns as Interpolator.cpp:31-47 builds r1margin / r2margin from the 1-based bound box
static int mc_geometry(int H, int W, const ocl_mc_args* mc, McGeom* g) {
    if (!mc || !mc->xn || !mc->x) return fail(OCL_E_INVALID, "motion compensation: null deformation arrays");
    const long m0 = H / 40, m1 = W / 40;
    const long r0 = (long)mc->bound_box[1] - m0 - 2, nr = (long)mc->bound_box[3] - (long)mc->bound_box[1] + 2 * m0 + 3;
    const long c0 = (long)mc->bound_box[0] - m1 - 2, nc = (long)mc->bound_box[2] - (long)mc->bound_box[0] + 2 * m1 + 3;
    if (nr < 1 || nc < 1 || r0 < 0 || c0 < 0 || r0 + nr > H || c0 + nc > W)
        return fail(OCL_E_INVALID, "motion compensation: the bound box with its margins leaves the frame");
    *g = McGeom{(int)r0, (int)nr, (int)c0, (int)nc};
    return OCL_OK;
}

static int grid_box(const McGeom& g, int F) { return grid_for((long)g.nr * g.nc * F); }

static int mc_apply(ocl_ctx* c, const float2* in, float2* out, int F, int H, int W, const ocl_mc_args* mc,
                    const McGeom& g, bool adjoint) {
    const long n = (long)F * H * W;
    if (!adjoint) {
        hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(OCL_NT), 0, c->stream, in, out, n);
        hipLaunchKernelGGL(k_mc_interp, dim3(grid_box(g, F)), dim3(OCL_NT), 0, c->stream, in, out, mc->xn, mc->x, F,
                           H, W, g.r0, g.nr, g.c0, g.nc);
    } else {
        hipLaunchKernelGGL(k_mc_copy_clear_box, dim3(grid_for(n)), dim3(OCL_NT), 0, c->stream, in, out, F, H, W, g.r0,
                           g.nr, g.c0, g.nc);
        hipLaunchKernelGGL(k_mc_scatter, dim3(grid_box(g, F)), dim3(OCL_NT), 0, c->stream, in, out, mc->xn, mc->x, F,
                           H, W, g.r0, g.nr, g.c0, g.nc);
    }
    HIPCHK(hipGetLastError());
    return OCL_OK;
}

static int mc_entry(ocl_ctx* c, const void* in, void* out, int F, int H, int W, const ocl_mc_args* mc, bool adjoint) {
    CTXCHK(c);
    if (!in || !out || F < 1 || H < 1 || W < 1) return fail(OCL_E_INVALID, "motion compensation: bad arguments");
    if (in == out) return fail(OCL_E_INVALID, "motion compensation: in and out must not alias");
    McGeom g;
    CHK(mc_geometry(H, W, mc, &g));
    HIPCHK(hipSetDevice(c->device));
    return mc_apply(c, (const float2*)in, (float2*)out, F, H, W, mc, g, adjoint);
}

extern "C" int ocl_motion_compensation(ocl_ctx* c, const void* in, void* out, int F, int H, int W,
                                       const ocl_mc_args* mc) {
    return mc_entry(c, in, out, F, H, W, mc, false);
}

extern "C" int ocl_adjoint_motion_compensation(ocl_ctx* c, const void* in, void* out, int F, int H, int W,
                                               const ocl_mc_args* mc) {
    return mc_entry(c, in, out, F, H, W, mc, true);
}

// The sparsity operator and its adjoint (NestaUp::operatorU / operatorUt, NestaUp.cpp:201-278); mc == nullptr is
// the plain temporal TV.  aux: one image-sized scratch buffer.  gatherR > 0: the mesh is the identity and every
// displacement is below gatherR pixels, so the operator runs as one pass (MC and tTV fused) and the adjoint as one
// deterministic gather (tTV^T fused in) instead of tTV^T, copy-and-clear and the atomic scatter.
struct SparsityOp {
    ocl_ctx* c;
    int F, H, W;
    const ocl_mc_args* mc;
    McGeom g;
    float2* aux;
    int gatherR = 0;
    int U(const float2* x, float2* out) const {
        const long N = (long)H * W;
        if (!mc) return ocl_ttv(c, x, out, F, N, 1);
        if (gatherR > 0) {
            HIPCHK(launch_mc_tv_fwd(x, out, mc->xn, F, H, W, g.r0, g.nr, g.c0, g.nc, c->stream));
            return OCL_OK;
        }
        CHK(mc_apply(c, x, aux, F, H, W, mc, g, false));
        return ocl_ttv(c, aux, out, F, N, 1);
    }
    // tTV^T into tmp, then MC^T (or a copy) into out
    int Ut(const float2* u, float2* tmp, float2* out) const {
        const long N = (long)H * W;
        if (mc && gatherR > 0) {
            HIPCHK(launch_mc_adj_tv_gather((const float2*)u, out, mc->xn, F, H, W, g.r0, g.nr, g.c0, g.nc, gatherR,
                                           c->stream));
            return OCL_OK;
        }
        CHK(ocl_ttv(c, u, tmp, F, N, 0));
        if (!mc) return ocl_copy(c, tmp, out, (long)F * N);
        return mc_apply(c, tmp, out, F, H, W, mc, g, true);
    }
};

// ---------------------------------------------------------------------------------------------------------------------
// myNormest (NestaUp.cpp:291-439) on the device, FP32 vectors and float scalars as in the reference
// ---------------------------------------------------------------------------------------------------------------------
static const unsigned long long DRAND48_X0 = 0x1234ABCD330EULL;

// op: the sparsity operator with motion compensation (null: plain temporal TV over F frames of N pixels).
// The first power iteration runs host-driven (U of the constant start vector is zero for the temporal TV, and the
// reference then refills Sx with drand48 values).  Later iterations run in batches of NE_BATCH with the norms and the
// normalisation kept on the device: one host round trip per batch, after which the host replays the stop test on the
// recorded norms and takes the value at the iteration where the sequential loop stops (the batch's later iterations
// only change the discarded vector).  An exact zero inside a batch (where the sequential loop would refill) restarts
// the whole estimate host-driven, so the result is always that of the sequential loop.
// start_pos: draws already taken from the drand48 stream; *end_pos: where the stream is after the estimate.
static int normest_compute(ocl_ctx* c, int F, long N, double* out, int* iters, const SparsityOp* op = nullptr,
                           long long start_pos = 0, long long* end_pos = nullptr) {
    constexpr int NE_BATCH = 8;
    const long n = (long)F * N;
    float2 *x = nullptr, *Sx = nullptr, *tmp = nullptr;
    double* dn = nullptr;                                   // [NE_BATCH + 1][2] norms of a batch, then ||x||
    auto release = [&]() {
        hipStreamSynchronize(c->stream);
        if (x) dev_free(x);
        if (Sx) dev_free(Sx);
        if (tmp) dev_free(tmp);
        if (dn) dev_free(dn);
    };
    if (dev_malloc((void**)&x, sizeof(float2) * n) != hipSuccess ||
        dev_malloc((void**)&Sx, sizeof(float2) * n) != hipSuccess ||
        (op && dev_malloc((void**)&tmp, sizeof(float2) * n) != hipSuccess) ||
        dev_malloc((void**)&dn, sizeof(double) * 2 * (NE_BATCH + 1)) != hipSuccess) {
        (void)hipGetLastError();
        release();
        return fail(OCL_E_NOMEM, "normest: out of device memory");
    }
    const float tol = 1e-3f;
    auto U = [&]() { return op ? op->U(x, Sx) : ocl_ttv(c, x, Sx, F, N, 1); };
    auto Ut = [&]() { return op ? op->Ut(Sx, tmp, x) : ocl_ttv(c, Sx, x, F, N, 0); };
    int rc = OCL_OK;
    for (int host_only = 0; host_only < 2; ++host_only) {
        double r[2];
        unsigned cnt = 0;
        long long drand_pos = start_pos;
        float e = 0.f, e0 = 0.f;
        bool restart = false;
        hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(OCL_NT), 0, c->stream, x, n, make_float2(1.f, 0.f));
        if ((rc = ocl_reduce(c, x, nullptr, n, RED_NRM2, r)) != OCL_OK) break;
        e = (float)r[0];
        if ((rc = ocl_scal(c, x, n, 1.0f / e, 0.f)) != OCL_OK) break;
        while ((std::abs(e - e0) > (tol * e)) && (cnt < 100u) && rc == OCL_OK) {
            if (cnt == 0 || host_only) {
                e0 = e;
                if ((rc = U()) != OCL_OK) break;
                // the reference tests sum |Sx| for zero; the 2-norm (accumulated in double) is zero exactly when
                // that sum is, so one reduction serves both unless the random refill runs
                if ((rc = ocl_reduce(c, Sx, nullptr, n, RED_NRM2, r)) != OCL_OK) break;
                if ((float)r[0] == 0.f) {
                    const long chunks = (n + 63) / 64;
                    hipLaunchKernelGGL(k_drand48_fill, dim3((unsigned)((chunks + OCL_NT - 1) / OCL_NT)), dim3(OCL_NT),
                                       0, c->stream, Sx, n, DRAND48_X0, drand_pos);
                    drand_pos += n;
                    if ((rc = ocl_reduce(c, Sx, nullptr, n, RED_NRM2, r)) != OCL_OK) break;
                }
                e = (float)r[0];
                if ((rc = Ut()) != OCL_OK) break;
                if ((rc = ocl_reduce(c, x, nullptr, n, RED_NRM2, r)) != OCL_OK) break;
                if ((rc = ocl_scal(c, x, n, 1.0f / (float)r[0], 0.f)) != OCL_OK) break;
                ++cnt;
                continue;
            }
            const int K = (int)std::min<unsigned>(NE_BATCH, 100u - cnt);
            for (int j = 0; j < K && rc == OCL_OK; ++j) {
                if ((rc = U()) != OCL_OK) break;
                if ((rc = reduce_async(c, Sx, nullptr, n, RED_NRM2, dn + 2 * j)) != OCL_OK) break;
                if ((rc = Ut()) != OCL_OK) break;
                if ((rc = reduce_async(c, x, nullptr, n, RED_NRM2, dn + 2 * NE_BATCH)) != OCL_OK) break;
                hipLaunchKernelGGL(k_scal_inv_norm, dim3(grid_for(n)), dim3(OCL_NT), 0, c->stream, x, n,
                                   (const double*)(dn + 2 * NE_BATCH));
                if (hipGetLastError() != hipSuccess) rc = fail(OCL_E_HIP, "normest: launch failed");
            }
            if (rc != OCL_OK) break;
            double h[2 * NE_BATCH];
            if (hipMemcpyAsync(h, dn, sizeof(double) * 2 * K, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess) {
                rc = fail(OCL_E_HIP, "normest: copy failed");
                break;
            }
            // replay the sequential loop on the recorded norms
            for (int j = 0; j < K; ++j) {
                if ((float)h[2 * j] == 0.f) {
                    restart = true;
                    break;
                }
                e0 = e;
                e = (float)h[2 * j];
                ++cnt;
                if (!((std::abs(e - e0) > (tol * e)) && (cnt < 100u))) break;
            }
            if (restart) break;
        }
        if (rc != OCL_OK) break;
        if (restart) continue;
        *out = e;
        *iters = (int)cnt;
        if (end_pos) *end_pos = drand_pos;
        break;
    }
    release();
    return rc;
}

int mc_box_geometry(int H, int W, const ocl_mc_args* mc, int* r0, int* nr, int* c0, int* nc) {
    McGeom g;
    CHK(mc_geometry(H, W, mc, &g));
    *r0 = g.r0; *nr = g.nr; *c0 = g.c0; *nc = g.nc;
    return OCL_OK;
}

int normest_mc(ocl_ctx* c, int F, int H, int W, const ocl_mc_args* mc, double* normU, int* iters, int gatherR) {
    const long n = (long)F * H * W;
    float2* aux = nullptr;
    if (dev_malloc((void**)&aux, sizeof(float2) * n) != hipSuccess)
        return fail(OCL_E_NOMEM, "normest: out of device memory");
    SparsityOp sp{c, F, H, W, mc, {}, aux};
    sp.gatherR = gatherR;
    int rc = mc_geometry(H, W, mc, &sp.g);
    long long end = 0;
    const long long start = c->normest_stream ? c->drand_pos : 0;
    if (rc == OCL_OK) rc = normest_compute(c, F, (long)H * W, normU, iters, &sp, start, &end);
    if (rc == OCL_OK && c->normest_stream) c->drand_pos = end;
    hipStreamSynchronize(c->stream);
    dev_free(aux);
    return rc;
}

extern "C" int ocl_ctx_set_normest_stream(ocl_ctx* c, int carry) {
    CTXCHK(c);
    c->normest_stream = carry ? 1 : 0;
    c->drand_pos = 0;
    return OCL_OK;
}

extern "C" int ocl_ctx_normest_stream_position(ocl_ctx* c, long long* position) {
    CTXCHK(c);
    if (!position) return fail(OCL_E_INVALID, "normest stream position: null output");
    *position = c->drand_pos;
    return OCL_OK;
}

// normU from the default drand48 state is a function of (F, N) alone: estimates are shared by every context of a
// device, so that a new context (a server's, another CLapp) does not repeat a power method another has already run
static std::mutex g_normU_mu;
static std::map<std::tuple<int, int, long>, std::pair<double, int>> g_normU;

extern "C" int ocl_normest(ocl_ctx* c, int F, long N, double* normU, int* iters) {
    CTXCHK(c);
    if (F < 1 || N < 1 || !normU) return fail(OCL_E_INVALID, "normest: bad arguments");
    if (c->normest_stream) {             // the carried stream: the value depends on the position, nothing is cached
        double v = 0;
        int it = 0;
        long long end = 0;
        CHK(normest_compute(c, F, N, &v, &it, nullptr, c->drand_pos, &end));
        c->drand_pos = end;
        *normU = v;
        if (iters) *iters = it;
        return OCL_OK;
    }
    {
        std::lock_guard<std::mutex> lk(c->mu);
        auto it = c->normU_cache.find({F, N});
        if (it != c->normU_cache.end()) {
            *normU = it->second.first;
            if (iters) *iters = it->second.second;
            return OCL_OK;
        }
    }
    std::pair<double, int> e{0.0, 0};
    bool have = false;
    {
        std::lock_guard<std::mutex> lk(g_normU_mu);
        auto it = g_normU.find(std::make_tuple(c->device, F, N));
        if (it != g_normU.end()) {
            e = it->second;
            have = true;
        }
    }
    if (!have) {
        CHK(normest_compute(c, F, N, &e.first, &e.second));
        std::lock_guard<std::mutex> lk(g_normU_mu);
        g_normU[std::make_tuple(c->device, F, N)] = e;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    c->normU_cache[{F, N}] = e;
    *normU = e.first;
    if (iters) *iters = e.second;
    return OCL_OK;
}

// ---------------------------------------------------------------------------------------------------------------------
// NESTA — PLAIN form: the reference's op sequence, float scalars, three host reads per iteration
// ---------------------------------------------------------------------------------------------------------------------
// mc != nullptr: the argsMC branch; xref then holds x_ref on entry (the previous reconstruction) and U includes the
// motion compensation.
static int nesta_plain(ocl_ctx* c, const void* y, const void* S, const uint8_t* M, void* xref, int F, int C, int H,
                       int W, const ocl_nesta_params& prm, ocl_nesta_report* rep, const ocl_mc_args* mc = nullptr) {
    const long N = (long)H * W, nX = (long)F * N, nK = (long)F * C * N;
    std::vector<void*> bufs;
    auto alloc = [&](long n) -> float2* {
        void* p = nullptr;
        if (dev_malloc(&p, sizeof(float2) * std::max(1L, n)) != hipSuccess) return nullptr;
        bufs.push_back(p);
        return (float2*)p;
    };
    struct Cleanup {
        std::vector<void*>* b;
        hipStream_t s;
        ~Cleanup() {
            hipStreamSynchronize(s);
            for (void* p : *b) dev_free(p);
        }
    } cleanup{&bufs, c->stream};
    float2 *yS = alloc(nK), *aux = alloc(nK), *Res = alloc(nK), *AuxRes = alloc(nK);
    float2 *Ux = alloc(nX), *Wk = alloc(nX), *Xk = alloc(nX), *Yk = alloc(nX), *Zk = alloc(nX), *Uk = alloc(nX),
           *AuxFx = alloc(nX), *Df = alloc(nX), *ARes = alloc(nX), *tmp = alloc(nX);
    float2* mcAux = mc ? alloc(nX) : nullptr;
    for (void* p : bufs)
        if (!p) return fail(OCL_E_NOMEM, "NESTA plain: out of device memory");
    if (bufs.size() != (mc ? 15u : 14u)) return fail(OCL_E_NOMEM, "NESTA plain: out of device memory");
    float2* X = (float2*)xref;
    SparsityOp sp{c, F, H, W, mc, {}, mcAux};
    if (mc) CHK(mc_geometry(H, W, mc, &sp.g));

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
    if (!mc) CHK(opAt(yS, X));                                        // :483-486
    CHK(sp.U(X, Ux));                                                 // :491
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
    if (mc) {
        long long end = 0;
        CHK(normest_compute(c, F, N, &nU, &ne_it, &sp, c->normest_stream ? c->drand_pos : 0, &end));
        if (c->normest_stream) c->drand_pos = end;
    } else {
        CHK(ocl_normest(c, F, N, &nU, &ne_it));
    }
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
            CHK(sp.U(Xk, Uk));                                        // :669
            CHK(ocl_copy(c, Uk, AuxFx, nX));                          // :675
            CHK(ocl_vecnorm(c, Uk, Uk, nX, mu));                      // :680-683
            CHK(reduce_async(c, Uk, nullptr, nX, RED_NRM2, dres));    // :686
            CHK(reduce_async(c, Uk, AuxFx, nX, RED_DOTC, dres + 2));  // :691
            CHK(sp.Ut(Uk, tmp, Df));                                  // :696
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
    if (form != OCL_FORM_FUSED && form != OCL_FORM_SPLIT) return fail(OCL_E_INVALID, "NESTA: unknown form");
    ocl_slice s{y, S, M, x_out, F, C, H, W};
    ocl_batch* b = nullptr;
    CHK(ocl_batch_create_form(c, 1, &s, prm, form, &b));
    int rc = ocl_batch_run(b, 16);
    if (rc == OCL_OK && rep) rc = ocl_batch_report(b, 0, rep);
    ocl_batch_destroy(b);
    return rc;
}

extern "C" int ocl_nesta_run_mc(ocl_ctx* c, const void* y, const void* S, const uint8_t* M, void* x_io, int F, int C,
                                int H, int W, const ocl_nesta_params* prm, int form, const ocl_mc_args* mc,
                                ocl_nesta_report* rep) {
    CTXCHK(c);
    if (!prm) return fail(OCL_E_INVALID, "NESTA: null launch parameters");
    if (!y || !x_io || F < 1 || C < 1 || H < 1 || W < 1) return fail(OCL_E_INVALID, "NESTA: bad input/output");
    if (!S) return fail(OCL_E_INVALID, "NESTA: non-existing SensitivityMaps");
    if (!M) return fail(OCL_E_INVALID, "NESTA: null sampling masks");
    if (prm->miniter < 1 || prm->miniter > OCL_MAX_MINITER) return fail(OCL_E_INVALID, "NESTA: miniter out of range");
    if (form != OCL_FORM_PLAIN && form != OCL_FORM_FUSED)
        return fail(OCL_E_UNSUPPORTED, "NESTA with motion compensation runs in the plain or the fused form");
    McGeom g;
    CHK(mc_geometry(H, W, mc, &g));
    HIPCHK(hipSetDevice(c->device));
    if (form == OCL_FORM_PLAIN) return nesta_plain(c, y, S, M, x_io, F, C, H, W, *prm, rep, mc);
    ocl_slice s{y, S, M, x_io, F, C, H, W};
    ocl_batch* b = nullptr;
    CHK(ocl_batch_create_mc(c, 1, &s, prm, mc, &b));
    int rc = ocl_batch_run(b, 16);
    if (rc == OCL_OK && rep) rc = ocl_batch_report(b, 0, rep);
    ocl_batch_destroy(b);
    return rc;
}
