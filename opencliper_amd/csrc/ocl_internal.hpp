// Declarations shared by the host-side translation units of libocl_amd.so (ocl_capi.cpp, ocl_batch.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "ocl_amd.h"
#include "ocl_kernels.hpp"
#include "ocl_nesta_fast.hpp"

// Records msg as the calling thread's last error and returns code.
int fail(int code, const std::string& msg);

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) return fail(OCL_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define CHK(x)                       \
    do {                             \
        int r_ = (x);                \
        if (r_ != OCL_OK) return r_; \
    } while (0)

// Tile shape of the fused FP32 NESTA passes for one (H, W): the register four-step kernels (fast) or the LDS
// Stockham kernels with tile_cols columns (pass CA) / rows_per_wg rows (pass B) per workgroup.  ca_fast: 0 Stockham,
// 1 four-step gathering the sampled rows from global memory, 2 four-step prefetching the next coil's sampled rows
// (through LDS) and sensitivities, 3 the Gram form (pass G) for line masks, the prefetching pass otherwise.
struct TileCfg {
    int ca_fast = 2, tile_cols = 16;
    int b_fast = 1, rows_per_wg = 16;
    double ms_CA = 0.0, ms_B = 0.0;   // measured per slice-iteration of the tuning workload (autotuned entries)
};

// One HIP device + one in-order stream (the role CLapp plays in src/CLapp.cpp:139-569 of the reference).
struct ocl_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::map<int, float2*> tw;                              // twiddle tables by DFT length
    double* red_part = nullptr;                             // reduction partials [1024][2]
    double* red_out = nullptr;                              // reduction results [8][2]
    double* red_host = nullptr;                             // pinned mirror of red_out
    std::map<std::pair<int, long>, std::pair<double, int>> normU_cache;
    // drand48 stream of myNormest's restarts: 0 = every normest starts from the POSIX default state (cached per
    // (F, N)); 1 = one stream carried across the context's normests, as the reference's unseeded process-wide
    // drand48 (NestaUp.cpp:375) is carried across the myNormest calls of a process (no cache then)
    int normest_stream = 0;
    long long drand_pos = 0;                                // draws taken from that stream so far
    std::mutex mu;
    int num_cus = 256;
    int autotune = 0;                                       // 1: tune the tile shape of each new (H, W) on first use
    // split-precision form: 1 runs the Gram pass (column GEMMs only) for line masks where it is instantiated, 0 keeps
    // the row and column passes for every slice
    int split_gram = 1;
    int group_min_wg = 0;                                   // launch-group threshold of mixed tables (0: default)
    std::map<std::pair<int, int>, TileCfg> tiles;           // tuned tile shapes by (H, W)
    // page-locked host buffer of ZC_BYTES that kernels read and write directly (zc_dev is its device address): the
    // small transfers of admission and of each chunk go through it with a copy kernel instead of the copy engines,
    // where they would queue behind a caller's large uploads
    char* zc_host = nullptr;
    char* zc_dev = nullptr;
};

static const size_t ZC_BYTES = 1 << 20;

static const int RED_PARTS = 1024;
static const size_t LDS_MAX = 64 * 1024;          // per-workgroup dynamic LDS budget of the standalone FFT launches
static const size_t LDS_MAX_NESTA = 160 * 1024;   // per-workgroup LDS of the fused NESTA passes (gfx950: 160 KB)

// Device memory of the library's own work buffers (slice state, launch tables, normest vectors, FFT twiddles).
// Every such allocation goes through dev_malloc / dev_free.  While an allocation-free call runs on the calling
// thread (ocl_nesta_run_noalloc), dev_malloc carves from that device's static scratch array instead of calling
// hipMalloc, fails with hipErrorOutOfMemory when the array is full, and dev_free of scratch memory does nothing.
hipError_t dev_malloc(void** p, size_t bytes);
hipError_t dev_free(void* p);
bool dev_scratch_active();                 // an allocation-free call is running on this thread
int scratch_init_device(int device);       // records the device's scratch address (ocl_ctx_create, device current)

int grid_for(long n);
int make_plan(ocl_ctx* c, int n, FftPlan* P);
int fast_p(int n);
int pow2_floor(int v);
int reduce_async(ocl_ctx* c, const void* x, const void* y, long n, int op, double* dev_out);
// motion compensation (ocl_capi.cpp): box rows r0..r0+nr-1, columns c0..c0+nc-1 of a bound box (OCL_E_INVALID when
// it leaves the frame), and myNormest over U = tTV o MC
int mc_box_geometry(int H, int W, const ocl_mc_args* mc, int* r0, int* nr, int* c0, int* nc);
// gatherR > 0: identity mesh with displacements below gatherR (<= 4) pixels, adjoint by gathering
int normest_mc(ocl_ctx* c, int F, int H, int W, const ocl_mc_args* mc, double* normU, int* iters, int gatherR = 0);
