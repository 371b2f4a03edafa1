// Types shared by the host side (ocl_capi.cpp) and the gfx950 kernels (ocl_ops.hip, ocl_nesta.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "ocl_device.hpp"

enum { RED_NRM2 = 0, RED_DOTC = 1, RED_SUM = 2, RED_ASUM = 3, RED_MAXABS = 4 };

#define OCL_MAX_MINITER 32
#define OCL_MAX_NSTAGES 32
#ifndef OCL_U_FRAMES
#define OCL_U_FRAMES 32      // frames per temporal-TV reduction workgroup (one group covers F <= 32); B+S of
#endif                       // 16 x (20, 8, 256^2): 0.172 ms at 32, 0.175 at 16, 0.184 at 8
#define OCL_MAXE 32          // most owned elements per thread in pass CA: H * tile_cols <= OCL_MAXE * OCL_NT

// Device-resident NESTA state of one slice (NestaUp.cpp:570-889 scalars).  Scalars are kept in double.
struct NState {
    double lambda, La, normU, gamma, gammat;
    int maxIntIter, maxIter, stoptest, miniter;
    int active;        // 0 once every continuation stage has finished
    int stage;         // 1-based continuation stage of the iteration about to run
    int k;             // inner iteration index of the iteration about to run
    int stage_start;   // 1: the next iteration is the first of its stage (x_ref := x_k, W_k := 0)
    int cur;           // which x buffer holds x_k
    int qpDone;
    int iters_stage, total_iters, ntrace, trace_cap;
    double mu, tolVar, Lmu, Lmu1;
    double last_f, last_qp;
    // coefficients of the iteration prepared by pass B and consumed by pass CA
    int it_pending;    // 1 when the scalar step has prepared an iteration for pass CA
    int it_cur, it_sstart;
    double it_mu, it_L1, it_apk, it_tau;
    // Gram form (line masks): W_k is ping-ponged too, and an iteration that the stop test ends is recomputed
    int wcur;          // which W buffer holds W_k
    int it_wcur;       // W_k buffer of the prepared iteration
    int it_redo;       // 1: the prepared iteration recomputes iteration k with the stage-ending coefficients
    double fmean[OCL_MAX_MINITER];
    int iters[OCL_MAX_NSTAGES];
};

// One slice of a (possibly uneven) batch.  All pointers are device pointers.
struct SliceDesc {
    int F, C, H, W;
    long N;
    int nmax;           // sampled-row capacity per frame in kbuf
    int tile_cols;      // column tile width of pass CA (power of two)
    int rows_per_wg;    // rows per workgroup of pass B (divides OCL_NT)
    int ncol_tiles;     // ceil(W / tile_cols)
    int nrow_groups;    // ceil(nmax / rows_per_wg)
    int wg_B, wg_CA;    // workgroups of this slice in passes B (k-space rows) and CA
    int wg_U;           // extra workgroups of pass B that reduce the temporal-TV terms of x_k
    int u_chunk;        // pixels per U workgroup (each covers OCL_U_FRAMES frames of its pixels)
    float inv_sqrtN;
    int fastH, fastW;   // P with H (W) = 16 P when the register four-step path applies, else 0
    int ca_stage;       // four-step pass CA with the next coil's sampled rows and S prefetched (nmax <= 64)
    FftPlan ph, pw;     // column (length H) and row (length W) plans
    const float2* y;    // [F][C][H][W]
    const float2* S;    // [C][H][W]
    const uint8_t* M;   // [F][H][W]
    float2* x0;         // ping-pong x_k buffers, [F][H][W] each
    float2* x1;
    float2* xref;       // [F][H][W]
    float2* Wk;         // [F][H][W]
    float2* kbuf;       // [F][C][nmax][W]  sampled-row intermediates
    const int16_t* rows;// [F][nmax] sampled rows of each frame
    const int16_t* rowpos; // [F][H] slot of row h in kbuf, or -1
    const int* nrows;   // [F]
    const uint16_t* mbits; // four-step pass B: [F][nmax][P_W] mask bits of the sampled rows, bit k2 of word k1 = M at
                           // column k1 + P_W k2 (the 16 mask values one thread of the pass needs, in one load)
    NState* st;
    double* partB;      // [wg_B]        sum |r|^2 per row workgroup
    double* partU;      // [wg_U][2]     sum |u|^2, sum Re(conj(u) Ux) per U workgroup
    unsigned* counter;  // [4] spare device counters (zeroed at setup)
    double* trace;      // [trace_cap][3] : f, ||r||^2, l1 per inner iteration (may be null)
    double* maxpart;    // [wg_CA][2] partial maxima of |x_ref| and |U x_ref| (setup)
    // split-precision form (prec = 1): DFT matrices as bf16 hi/lo planes (re_hi, re_lo, im_hi, im_lo), see
    // ocl_nesta_mfma.hip for the index order of each
    int prec;           // 0: FP32 FFT passes, 1: BF16-MFMA DFT passes with a 3-term split product
    int kp;             // nmax rounded up to a multiple of 32 (the K extent of the inverse column DFT)
    const uint16_t* drowF;  // [4][W][W]  forward row DFT
    const uint16_t* drowI;  // [4][W][W]  inverse row DFT, K index permuted
    const uint16_t* dcolI;  // [F][4][H][kp]  inverse column DFT from the sampled rows
    const uint16_t* dcolF;  // [F][4][kp][H]  forward column DFT onto the sampled rows, K index permuted
    // motion compensation (null mc_xn: none).  The sparsity operator is then tTV o MC: pass Z writes mcz = MC(x_k)
    // before pass B (whose temporal-TV terms read it), pass G writes mcg = MC^T(D^T u(mcz)) after the scalar step,
    // and pass CA's update reads mcg in place of the local temporal-TV gradient.  The original mesh must be the
    // identity (checked at batch creation), so box pixel (r, c) resamples at xn and writes to (r, c).
    const float* mc_xn;     // [2][F][H][W] deformed sample positions, 1-based ([0] column, [1] row)
    int mc_r0, mc_nr, mc_c0, mc_nc;   // box rows / columns (0-based)
    float2* mcz;            // [F][H][W]
    float2* mcg;            // [F][H][W]
    int wg_P, wg_X, wg_G;   // workgroups of the pixel passes, of the box passes and of the tiled gather
    unsigned* mc_dmax;      // largest displacement |xn - position| over the box (float bits), set at setup
    // Gram form of the iteration (line masks, four-step column size, nmax <= 64): one pass per iteration over column
    // tiles (k_nesta_G_t) and a scalar step after it; kbuf holds V = W F_x^{-1} y of the carried rows throughout
    int gram_ok;            // the slice can run the Gram form (mask constant along every row, sizes supported)
    float2* Wk2;            // second W buffer [F][H][W] (W_k is ping-ponged with Wk)
    const uint8_t* rowmask; // [F][nmax] 1 when the carried row is a sampled (mask) row
    double* partG;          // [wg_CA][3] per column tile: ||r||^2, sum |u|^2, sum Re(conj(u) Ux)
};
#define OCL_MC_GATHER_RMAX 4 // pass G gathers (deterministically) when floor(max displacement) + 1 <= this
#ifndef OCL_MC_TR
#define OCL_MC_TR 16         // tile of the gather form of pass G: rows x columns of one frame (32 x 64 measured
#endif                       // slower: 63 against 66 fused MC recons/s)
#ifndef OCL_MC_TC
#define OCL_MC_TC 64
#endif
#define OCL_MC_CHUNK 1024    // pixel-frames per workgroup of the motion-compensation passes
// The build-time constants above may be overridden with -D; the kernels rely on these properties of them:
static_assert(OCL_U_FRAMES >= 1, "OCL_U_FRAMES: at least one frame per temporal-TV workgroup");
static_assert(OCL_MC_TR >= 1 && OCL_MC_TC >= 1 && (OCL_MC_TR * OCL_MC_TC) % OCL_NT == 0,
              "OCL_MC_TR x OCL_MC_TC: the gather writes TR x TC / OCL_NT pixels per thread, so the tile must be a multiple "
              "of the workgroup size");
static_assert(2 * sizeof(float) * 2 * (OCL_MC_TR + 2 * OCL_MC_GATHER_RMAX) * (OCL_MC_TC + 2 * OCL_MC_GATHER_RMAX) <=
                  64 * 1024,
              "OCL_MC_TR x OCL_MC_TC: the gather's static LDS staging (positions and values of the tile plus its halo at "
              "the largest radius) must fit 64 KB");

// kernels (ocl_ops.hip)
__global__ void k_ttv(const float2* in, float2* out, int F, long N, int dir);
__global__ void k_vecnorm(const float2* in, float2* out, long n, float mu);
__global__ void k_cep(const float2* in, const float2* S, float2* out, int F, int C, long N, int in_coils, int conj);
__global__ void k_apply_mask(float2* d, const uint8_t* M, int F, int C, long N);
__global__ void k_ximage_sum(const float2* in, float2* out, int F, int C, long N);
__global__ void k_cep_sum(float2* y, const float2* S, float2* x, int F, int C, long N);
__global__ void k_rsos(const float2* in, float2* out, int F, int C, long N);
__global__ void k_abs(const float2* in, float2* out, long n, int pow2);
__global__ void k_copy(const float2* x, float2* y, long n);
__global__ void k_scal(float2* x, long n, float ar, float ai);
__global__ void k_scal_inv_norm(float2* x, long n, const double* nrm);
__global__ void k_axpy(long n, float ar, float ai, const float2* x, float2* y);
__global__ void k_fill(float2* x, long n, float2 v);
__global__ void k_fill_bytes(uint8_t* x, long n, uint8_t v);
__global__ void k_reduce_partial(const float2* x, const float2* y, long n, int op, double* partials);
__global__ void k_reduce_final(const double* partials, int nparts, int op, double* out);
__global__ void k_fft_rows(const float2* in, float2* out, long nrows, FftPlan P, int rows_per_wg, int conj_io,
                           float scale);
__global__ void k_fft_cols(const float2* in, float2* out, int H, int W, FftPlan P, int tw_cols, int conj_io,
                           float scale);
__global__ void k_drand48_fill(float2* out, long n, unsigned long long x0, long start);
__global__ void k_add_noise_masked(float2* y, const uint8_t* M, int F, int C, long N, float sigma,
                                   unsigned long long seed);
#define FAST_COLS_KS 272   // per-k1 stride (float2) of the column exchange in the fast FFT kernels
__global__ void k_fft_rows_fast(const float2* in, float2* out, long nrows, int P, const float2* tw, int conj_io,
                                float scale);
__global__ void k_fft_cols_fast(const float2* in, float2* out, int W, int P, const float2* tw, int conj_io,
                                float scale);
// a batch of SimpleMRIRecon inputs of one shape, passed by value in the kernel argument
#define OCL_RECON_BATCH 64
struct ReconPtrs {
    float2* y[OCL_RECON_BATCH];
    const float2* S[OCL_RECON_BATCH];
    float2* x[OCL_RECON_BATCH];
};
// SimpleMRIRecon's fused column DFT + conj(S) + coil sum, one kernel per column size H = 16 P (ocl_ops.hip): F frames
// of one input, or m inputs of F frames each
hipError_t launch_simple_cols(float2* y, const float2* S, float2* x, int F, int C, int W, int P, const float2* tw,
                              float scale, size_t lds, hipStream_t stream);
hipError_t launch_simple_cols_batch(const ReconPtrs& p, int m, int F, int C, int W, int P, const float2* tw,
                                    float scale, size_t lds, hipStream_t stream);
void simple_cols_attributes(int lds_max);
__global__ void k_fft_rows_fast_batch(ReconPtrs p, long nrows, int P, const float2* tw, int conj_io, float scale);

// kernels (ocl_mc.hip): motion compensation over the box rows r0..r0+nr-1, columns c0..c0+nc-1 of every frame
__global__ void k_mc_interp(const float2* in, float2* out, const float* xn, const int* x, int F, int H, int W, int r0,
                            int nr, int c0, int nc);
__global__ void k_mc_copy_clear_box(const float2* in, float2* out, int F, int H, int W, int r0, int nr, int c0, int nc);
__global__ void k_mc_scatter(const float2* in, float2* out, const float* xn, const int* x, int F, int H, int W, int r0,
                             int nr, int c0, int nc);

// kernels (ocl_nesta.hip)
__global__ void k_rowflags(const float2* y, const uint8_t* M, int F, int C, int H, int W, uint8_t* flags);
__global__ void k_mask_bits(const uint8_t* M, const int16_t* rows, const int* nrows, int F, int H, int W, int nmax, int P,
                            uint16_t* out);
// launch shape of the fused passes over one batch; PH / PW > 0 when every slice shares that fast column / row size
struct NestaLaunch {
    const SliceDesc* dd;
    const int* offB;
    const int* offCA;
    int ns, totB, totCA;
    size_t ldsB, ldsCA;
    int maxe, PH, PW;
    int prec;           // 1: every slice uses the split-precision passes (then PH = H/16, PW = W/16)
    int kj;             // split-precision passes: largest kp / 32 of the table
    int stage;          // 0 / 1: every slice uses the direct / prefetching four-step pass CA, 2: mixed
    int mc;             // 1: every slice uses motion compensation (passes Z and G run)
    const int* offP;    // workgroup offsets of the motion-compensation pixel passes
    const int* offX;    // workgroup offsets of the motion-compensation box scatter
    const int* offG;    // workgroup offsets of the tiled gather form of pass G
    int totP, totX, totG;
    int mcR;            // gather radius of pass G (largest over the table), 0: scatter form; set at setup
    int gram;           // 1: every slice runs the Gram form (pass G over column tiles + scalar step S2)
};
// byte copy in 4-byte words by a kernel (addresses and size multiples of 4), for transfers through mapped host memory
hipError_t launch_copy_words(const void* src, void* dst, size_t bytes, hipStream_t stream);
hipError_t launch_nesta_B(const NestaLaunch& L, hipStream_t stream, int mode);
// Gram form: pass G (mode 0 iteration, 1 setup x_ref = A^H sqrt(N) y into x0) and the scalar step after it
hipError_t launch_nesta_G(const NestaLaunch& L, hipStream_t stream, int mode);
hipError_t launch_nesta_S2(const NestaLaunch& L, hipStream_t stream);
hipError_t launch_nesta_S(const NestaLaunch& L, hipStream_t stream);
// mode 0: iteration, 1: setup (x_ref = A^H y), 2: setup from a given x_ref in x0 (forward columns only)
hipError_t launch_nesta_CA(const NestaLaunch& L, hipStream_t stream, int mode);
// motion-compensation passes of the fused iteration (ocl_mc.hip): Z (mode 0: x_k, mode 1: setup from x0), G
hipError_t launch_nesta_mc_z(const NestaLaunch& L, hipStream_t stream, int mode);
hipError_t launch_nesta_mc_g(const NestaLaunch& L, hipStream_t stream);
hipError_t launch_nesta_mc_dmax(const NestaLaunch& L, hipStream_t stream);
// out = tTV(MC(x)) in one pass, identity mesh, box rows r0..r0+nr-1 x columns c0..c0+nc-1
hipError_t launch_mc_tv_fwd(const float2* x, float2* out, const float* xn, int F, int H, int W, int r0, int nr, int c0,
                            int nc, hipStream_t stream);
// out = MC^T(tTV^T y) by gathering, identity mesh, box rows r0..r0+nr-1 x columns c0..c0+nc-1, displacements < R <= 4
hipError_t launch_mc_adj_tv_gather(const float2* y, float2* out, const float* xn, int F, int H, int W, int r0, int nr,
                                   int c0, int nc, int R, hipStream_t stream);
void nesta_kernel_attributes(int lds_bytes);
// split-precision passes (ocl_nesta_mfma.hip)
hipError_t launch_nesta_B_mfma(const NestaLaunch& L, hipStream_t stream, int mode);
hipError_t launch_nesta_CA_mfma(const NestaLaunch& L, hipStream_t stream, int mode);
// split-precision Gram pass (line masks): mode 0 iteration, 1 setup x_ref = A^H sqrt(N) y into x0
hipError_t launch_nesta_G_mfma(const NestaLaunch& L, hipStream_t stream, int mode);
bool mfma_gram_supported(int H, int kp);   // H and sampled-row capacity the split-precision Gram pass runs
bool mfma_size_supported(int n);      // H or W the split-precision passes are instantiated for
void mfma_kernel_attributes(int lds_bytes);
#define MF_CA_TILES 4                  // column tiles (one wavefront each) per workgroup of the split-precision pass CA
#define MF_B_TASKS 4                   // 16-row tasks (one wavefront each) per workgroup of the split-precision pass B
// LDS row stride (bf16 elements) of a staged DFT matrix with K columns: the 8-element pad keeps the 16-byte fragment
// reads of a quarter-wavefront on distinct banks
__host__ __device__ constexpr int mf_lds_ld(int K) { return K + 8; }
__host__ __device__ constexpr size_t mf_b_lds_bytes(int W) { return 2 * 4 * (size_t)W * mf_lds_ld(32); }
__host__ __device__ constexpr size_t mf_ca_lds_bytes(int H, int kp) {
    return 2 * 4 * (size_t)((H * mf_lds_ld(kp)) > (kp * mf_lds_ld(H)) ? H * mf_lds_ld(kp) : kp * mf_lds_ld(H));
}
// Gram pass of the split-precision form (line masks): row-waves of mf_gs_rt(MT) m-tiles (16 rows each) per 16-column
// tile; a tile's two row-waves (H >= 128) exchange their forward partials through LDS
__host__ __device__ constexpr int mf_gs_rt(int MT) { return MT <= 4 ? MT : MT / 2; }
// LDS of the Gram pass: both DFT matrices of a frame (forward [4][kp][H], inverse [4][H][kp], padded rows) and, with
// two row-waves, one slab of forward partials per wavefront (kp / 16 m-tiles x 8 floats x 64 lanes)
__host__ __device__ constexpr size_t mf_gs_lds_bytes(int H, int kp) {
    return 2 * 4 * ((size_t)kp * mf_lds_ld(H) + (size_t)H * mf_lds_ld(kp)) +
           (mf_gs_rt(H / 16) < H / 16 ? (size_t)4 * 2 * MF_CA_TILES * (kp / 16) * 8 * 64 : 0);
}
// fills one DFT matrix of the split-precision passes (kind: 0 row forward, 1 row inverse, 2 column inverse,
// 3 column forward, 4 column inverse with the K index permuted); rows/nrows/nmax describe the sampled rows of each
// frame for the column kinds
__global__ void k_dft_split(uint16_t* out, int kind, int L, int kp, const int16_t* rows, const int* nrows, int nmax);
__global__ void k_setup_max(const SliceDesc* descs, const int* wg_off, int nslices, double* out);
__global__ void k_finalize(const SliceDesc* descs, int nslices, float2* const* outs);
