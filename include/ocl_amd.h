/*
 * C ABI of libocl_amd.so — the MI355X (gfx950) implementation of OpenCLIPER's NESTA-TV hot path.
 *
 * Plain pointers and sizes only.  Every function returns 0 on success or a negative OCL_E* code; the message of the
 * last failure on the calling thread is returned by ocl_last_error().  "Device pointers" are HBM addresses obtained
 * from ocl_malloc (or any hipMalloc'ed memory of the same device).  Complex data is interleaved float (re, im).
 *
 * Layouts (W fastest, no padding, no in-band header):
 *   XData  [F][H][W] complex64        KData  [F][C][H][W] complex64
 *   SensitivityMapsData [C][H][W]     SamplingMasksData [F][H][W] uint8 (non-zero = sampled)
 *
 * Each entry point names the reference interface it replaces (paths relative to the reference checkout).
 */
#ifndef OCL_AMD_H
#define OCL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OCL_OK 0
#define OCL_E_INVALID -1     /* bad argument (std::invalid_argument in the reference) */
#define OCL_E_HIP -2         /* HIP runtime failure (CLError in the reference) */
#define OCL_E_UNSUPPORTED -3 /* e.g. FFT length not 2^a 3^b 5^c 7^d (src/processes/FFT.cpp:374-379) */
#define OCL_E_NOMEM -4       /* device memory exhausted / admission refused */

typedef struct ocl_ctx ocl_ctx;
typedef struct ocl_batch ocl_batch;

/* NestaUp::LaunchParameters (include/OpenCLIPER/processes/nesta/NestaUp.hpp:67-83), same field order.
 * argsMC is passed separately: ocl_nesta_run_mc and ocl_batch_create_mc take it as an ocl_mc_args. */
typedef struct {
    double lambda_i, mu_f, La;
    int maxIntIter;
    double tolVar;
    int verbose, maxIter, stoptest, miniter;
} ocl_nesta_params;

#define OCL_MAX_STAGES 16
#define OCL_TRACE 4096
/* What NestaUp::launch computes along the way (NestaUp.cpp:521-583, per-iteration f at :797-798). */
typedef struct {
    double normU, maxXref, maxUXref, lambda, muf, mu0, gamma, gammat, final_tolVar, final_mu;
    int normest_iters, stages;
    int iters[OCL_MAX_STAGES];
    int ntrace;
    double trace_f[OCL_TRACE];
    double trace_l2[OCL_TRACE];
    double trace_l1[OCL_TRACE];
} ocl_nesta_report;

/* One slice of a batch: device pointers plus sizes. */
typedef struct {
    const void* y;       /* KData [F][C][H][W] */
    const void* S;       /* [C][H][W] */
    const uint8_t* M;    /* [F][H][W] */
    void* x_out;         /* XData [F][H][W] */
    int F, C, H, W;
} ocl_slice;

/* ---- context: replaces CLapp::create/init (src/CLapp.cpp:139-569): one device, one in-order stream ---- */
const char* ocl_last_error(void);
int ocl_ctx_create(int device, ocl_ctx** out);
int ocl_ctx_destroy(ocl_ctx* ctx);   /* destroy the context's batches and registrations first */
int ocl_ctx_sync(ocl_ctx* ctx);
void* ocl_ctx_stream(ocl_ctx* ctx);                       /* CLapp::getCommandQueue -> hipStream_t */
int ocl_device_info(ocl_ctx* ctx, char* name, size_t name_len, size_t* total_mem, int* num_cus);

/* ---- stream-ordered timing: replaces ProcessCore::startProfiling / stopProfiling (src/ProcessCore.cpp:184-241) ---- */
typedef struct ocl_timer ocl_timer;
int ocl_timer_begin(ocl_ctx* ctx, ocl_timer** t);              /* records a start event on the context's stream */
int ocl_timer_end(ocl_ctx* ctx, ocl_timer* t, double* ms);     /* records the stop event, waits, frees the timer */

/* ---- device buffers: replace Data::host2Device / device2Host (src/Data.cpp:1493-1495) ---- */
int ocl_malloc(ocl_ctx* ctx, size_t bytes, void** dptr);
int ocl_free(ocl_ctx* ctx, void* dptr);
int ocl_memcpy_h2d(ocl_ctx* ctx, void* dst, const void* src, size_t bytes);   /* synchronous */
int ocl_memcpy_d2h(ocl_ctx* ctx, void* dst, const void* src, size_t bytes);   /* synchronous */
int ocl_memcpy_d2d(ocl_ctx* ctx, void* dst, const void* src, size_t bytes);   /* stream-ordered */
int ocl_memset_zero(ocl_ctx* ctx, void* dptr, size_t bytes);                  /* stream-ordered fill kernel */
/* page-locked host memory for inputs uploaded again and again (a server's request payloads): replaces the reference's
 * mapped host view of a device buffer (src/DeviceDataProperties.cpp:242, enqueueMapBuffer).  ocl_memcpy_h2d from it
 * runs as one DMA transfer instead of through the driver's pageable staging.  Freed by ocl_host_free only, which
 * waits for the context's stream first (ctx may be NULL: the memory does not belong to a context). */
int ocl_host_malloc(ocl_ctx* ctx, size_t bytes, void** hptr);
int ocl_host_free(ocl_ctx* ctx, void* hptr);

/* ---- individual Processes (stream-ordered; outputs may alias inputs where the reference allows it) ---- */
/* TemporalTV::launch (src/processes/nesta/TemporalTV.cpp:49-86, kernels src/kernels/temporalTV.cl:49-107);
 * dir = 1 FORWARD, 0 ADJOINT (the reference's adjoint kernel, i.e. -D^T).  in and out must not alias. */
int ocl_ttv(ocl_ctx* ctx, const void* in, void* out, int F, long N, int dir);
/* VectorNormalization::launch (src/processes/nesta/VectorNormalization.cpp:50-87): out = in / max(mu, |in|) */
int ocl_vecnorm(ocl_ctx* ctx, const void* in, void* out, long n, float mu);
/* ComplexElementProd::launch (src/processes/ComplexElementProd.cpp:60-131); in_coils = 0 for XData input */
int ocl_cep(ocl_ctx* ctx, const void* in, const void* S, void* out, int F, int C, long N, int in_coils, int conj);
/* ApplyMask::launch (src/processes/ApplyMask.cpp:21-51), in place */
int ocl_apply_mask(ocl_ctx* ctx, void* data, const uint8_t* M, int F, int C, long N);
/* XImageSum::launch (src/processes/XImageSum.cpp:66-96) */
int ocl_ximage_sum(ocl_ctx* ctx, const void* in, void* out, int F, int C, long N);
/* ComplexAbs / ComplexAbsPow2::launch (src/processes/ComplexAbs.cpp:48-78, ComplexAbsPow2.cpp:48-78) */
int ocl_complex_abs(ocl_ctx* ctx, const void* in, void* out, long n, int pow2);
/* CopyDataGPU::launch (src/processes/groupwiseRegistration/initialize/CopyDataGPU.cpp:53-74) */
int ocl_copy(ocl_ctx* ctx, const void* in, void* out, long n);
/* FFT::launch with dim = -1 (src/processes/FFT.cpp:127-146,510-525): batch of 2-D C2C transforms of [H][W];
 * dir = -1 FORWARD (unscaled), +1 BACKWARD (scaled by 1/(H*W)).  in == out is allowed. */
int ocl_fft2d(ocl_ctx* ctx, const void* in, void* out, long batch, int H, int W, int dir);
/* SimpleMRIRecon::launch (src/processes/examples/SimpleMRIRecon.cpp:59-101): x = sum_c conj(S_c) IFFT2(y_c).
 * Like the reference it transforms y in place. */
int ocl_simple_mri_recon(ocl_ctx* ctx, void* y, const void* S, void* x, int F, int C, int H, int W);
/* n SimpleMRIRecon launches of one shape at once (no reference counterpart: for callers with many reconstructions,
 * where one launch per reconstruction would be launch-bound).  y[i], S[i], x[i] are the i-th input's device
 * buffers; each result is what ocl_simple_mri_recon gives, and each y[i] is transformed in place as there. */
int ocl_simple_mri_recon_batch(ocl_ctx* ctx, int n, void* const* y, const void* const* S, void* const* x, int F, int C,
                               int H, int W);

/* RSoS::launch (src/processes/RSoS.cpp, kernel src/kernels/rss.cl): out_f = (sqrt(sum_c |in_{f,c}|^2), 0) */
int ocl_rsos(ocl_ctx* ctx, const void* in, void* out, int F, int C, long N);
/* SimpleMRIReconSOS::launch (src/processes/examples/SimpleMRIReconSOS.cpp): inverse 2-D FFT of y in place, then RSoS */
int ocl_simple_mri_recon_sos(ocl_ctx* ctx, void* y, void* x, int F, int C, int H, int W);

/* Synthetic-input helper (no reference counterpart): y = M * (y + sigma * n) with complex Gaussian n whose real and
 * imaginary parts have variance 1/2, drawn from a counter-based hash of (seed, index).  Used to build large k-space
 * test inputs on the device. */
int ocl_add_noise_masked(ocl_ctx* ctx, void* y, const uint8_t* M, int F, int C, long N, float sigma,
                         unsigned long long seed);

/* ---- BLAS-1 as NestaUp calls it through CLBlast (NestaUp.cpp call sites listed in the docs) ---- */
int ocl_scal(ocl_ctx* ctx, void* x, long n, float ar, float ai);                          /* CLBlastCscal */
int ocl_axpy(ocl_ctx* ctx, long n, float ar, float ai, const void* x, void* y);           /* CLBlastCaxpy */
/* result[0..1] on the host (blocking, like the reference's clEnqueueReadBuffer):
 * op 0 Scnrm2, 1 Cdotc (conj(x).y, re/im), 2 Scsum (sum re+im), 3 Scasum, 4 max |x| */
int ocl_reduce(ocl_ctx* ctx, const void* x, const void* y, long n, int op, double* result);

/* ---- NESTA-TV: NestaUp::launch (src/processes/nesta/NestaUp.cpp:442-910) ---- */
#define OCL_FORM_PLAIN 0   /* the reference's op sequence on the GPU, host stop test after each iteration */
#define OCL_FORM_FUSED 1   /* fused row pass, scalar step, fused column pass; device-side stop test, no host sync */
#define OCL_FORM_SPLIT 2   /* as FUSED, with the row and column DFTs as BF16 MFMA GEMMs (3-term hi/lo split product,
                              FP32 accumulation; NESTA scalars and norms in FP64); H, W in {64, 128, 192, 256} */
/* myNormest (NestaUp.cpp:291-439) for an F-frame image of N pixels.  Its restarts draw from drand48, which the
 * reference never seeds (NestaUp.cpp:375): one process-wide stream from the POSIX default state.  By default each
 * estimate here starts that stream afresh (what the first reconstruction of a process sees; the value is then cached
 * per (F, N)); ocl_ctx_set_normest_stream(ctx, 1) carries one stream across the context's estimates instead, as the
 * reference's process carries it across every NestaUp::launch (e.g. NESTA then MC-NESTA, tests/MRIRecon.cpp:232-272),
 * and resets it to the default state.  Batches and servers on such a context draw in table order. */
int ocl_normest(ocl_ctx* ctx, int F, long N, double* normU, int* iters);
int ocl_ctx_set_normest_stream(ocl_ctx* ctx, int carry);
int ocl_ctx_normest_stream_position(ocl_ctx* ctx, long long* position);   /* draws taken so far */
int ocl_nesta_run(ocl_ctx* ctx, const void* y, const void* S, const uint8_t* M, void* x_out, int F, int C, int H,
                  int W, const ocl_nesta_params* prm, int form, ocl_nesta_report* report);

/* No counterpart in the reference.  ocl_nesta_run (form OCL_FORM_FUSED or OCL_FORM_SPLIT) for callers that cannot
 * allocate device memory: serving loops and harnesses that forbid hipMalloc.  It calls no hipMalloc, hipFree or other
 * allocating routine: every work buffer (slice state, sampled-row tables, launch table, normest vectors) is carved from
 * a static device array of OCL_SCRATCH_BYTES that each GPU holds from the first ocl_ctx_create on it until the process
 * ends.  Inputs and outputs are the caller's, as for ocl_nesta_run.  A reconstruction whose work buffers do not fit
 * returns OCL_E_NOMEM with a message naming the limit (about 80 MiB for 20 frames x 8 coils x 256^2 at acceleration 6,
 * including the first normest of a context).  Tile shapes are never tuned here: a context's tuned or pinned shape is
 * used when it has one, else the heuristic.  All calls on one GPU share its array, which keeps the last call's
 * contents: calls that overlap on one GPU (from two threads or two contexts) race and must be serialized by the
 * caller.  The call is synchronous, like ocl_nesta_run. */
#define OCL_SCRATCH_BYTES ((size_t)128 << 20)
int ocl_nesta_run_noalloc(ocl_ctx* ctx, const void* y, const void* S, const uint8_t* M, void* x_out, int F, int C,
                          int H, int W, const ocl_nesta_params* prm, int form, ocl_nesta_report* report);

/* ---- motion compensation (ArgumentsMotionCompensation, include/OpenCLIPER/processes/GroupwiseRegistration.hpp:95-100;
 * MotionCompensation.cpp, AdjointMotionCompensation.cpp, src/kernels/optimizer.cl:832-930) ----
 * xn: device float [2][F][H][W], deformed sample positions, 1-based ([0] column, [1] row);
 * x:  device int32 [2][W][H], the original mesh, 1-based, element [r + c*H] = (column + 1, row + 1);
 * bound_box: 1-based, as the reference's Interpolator reads it: box rows bound_box[1] - m0 - 2 ..
 * bound_box[3] + m0 and columns bound_box[0] - m1 - 2 .. bound_box[2] + m1 (0-based), m0 = H / 40, m1 = W / 40.
 * A box that leaves the frame is OCL_E_INVALID.  Neighbours outside the frame read as zero. */
typedef struct {
    const float* xn;
    const int* x;
    unsigned bound_box[4];
} ocl_mc_args;
/* MotionCompensation::launch: out = in, then the box pixels resampled bilinearly at xn (in != out) */
int ocl_motion_compensation(ocl_ctx* ctx, const void* in, void* out, int F, int H, int W, const ocl_mc_args* mc);
/* AdjointMotionCompensation::launch: out = in with the box cleared, then the transpose scatter (float atomics) */
int ocl_adjoint_motion_compensation(ocl_ctx* ctx, const void* in, void* out, int F, int H, int W,
                                    const ocl_mc_args* mc);
/* NestaUp::launch with argsMC != nullptr (NestaUp.cpp:201-278,483-491): U = tTV(MC(x)), U^T = MC^T(tTV^T(u)), and
 * x_io holds x_ref on entry (the previous reconstruction) instead of A^H y.  form: OCL_FORM_PLAIN (any mesh) or
 * OCL_FORM_FUSED (the original, identity mesh that GroupwiseRegistration produces); OCL_FORM_SPLIT is unsupported. */
int ocl_nesta_run_mc(ocl_ctx* ctx, const void* y, const void* S, const uint8_t* M, void* x_io, int F, int C, int H,
                     int W, const ocl_nesta_params* prm, int form, const ocl_mc_args* mc, ocl_nesta_report* report);

/* ---- groupwise B-spline registration (GroupwiseRegistration, src/processes/GroupwiseRegistration.cpp:93-615,
 * groupwiseRegistration/{creabspline,optimizer}.cpp): estimates the motion that ocl_mc_args carries ----
 * GroupwiseRegistration::InitParameters (include/OpenCLIPER/processes/GroupwiseRegistration.hpp:145-158), same order:
 * W step weight (Wn = W / ROI area), flagW adaptive Wn, radius of the circular ROI, E spline order, Dp control
 * spacing, nmax iterations, et step-norm and eh cost-decrease thresholds, lambda regularisation weights. */
typedef struct {
    float W;
    int flagW;
    unsigned radius, E;
    int Dp[2];
    unsigned nmax;
    float et, eh;
    float lambda[4];
} ocl_gwr_params;
typedef struct ocl_gwr ocl_gwr;
typedef struct {
    int Cp;                  /* control points per dimension */
    int support;             /* pixels per dimension of a control point's support */
    int box;                 /* side of the ROI's bounding box */
    int areaX;               /* ROI pixels */
    int use_lambda;          /* the regularisation enters the cost (int-accumulated weight sum != 0) */
    unsigned bound_box[4];   /* ArgumentsMotionCompensation::bound_box, 1-based */
    long T_elems;            /* 2 * F * Cp * Cp floats: T [2][F][Cp][Cp] (the reference's TData {Cp1, Cp0, F, 2}) */
    long xn_elems;           /* 2 * F * H * W floats: xn [2][F][H][W] (ArgumentsMotionCompensation::xnData) */
    const int* mesh;         /* device int32 [2][W][H]: ArgumentsMotionCompensation::xData (the identity mesh) */
} ocl_gwr_info_t;
#define OCL_GWR_TRACE 1024
/* What GroupwiseRegistration's optimizer prints per iteration (optimizer.cpp:350-356): the cost H (H[0] before the
 * loop), the step norm Tdif, the cost decrease Hdif and Wn after adaptation, for the first OCL_GWR_TRACE - 1
 * iterations. */
typedef struct {
    int iters, ntrace;
    float H0, Wn, oldmax;
    float H[OCL_GWR_TRACE];
    float Tdif[OCL_GWR_TRACE], Hdif[OCL_GWR_TRACE], Wn_trace[OCL_GWR_TRACE];
} ocl_gwr_report;
/* GroupwiseRegistration::init + creabspline: the ROI, control grid and B-spline tables of an F x H x W series.
 * Supported: a square image, radius > 0, E = 3, Dp[0] == Dp[1], 2 <= F <= 64, and a geometry whose control points,
 * supports and interpolation margin stay inside the frame (OCL_E_UNSUPPORTED otherwise, with the reason). */
int ocl_gwr_create(ocl_ctx* ctx, int F, int H, int W, const ocl_gwr_params* prm, ocl_gwr** out);
int ocl_gwr_info(const ocl_gwr* g, ocl_gwr_info_t* info);
/* GroupwiseRegistration::launch: x device complex64 [F][H][W] (the reconstruction); T device float [2][F][Cp][Cp],
 * read as the starting coefficients when T_in != 0 (LaunchParameters::TData from an earlier motion iteration), else
 * cleared, and holding the estimate on return; xn device float [2][F][H][W] receives the deformed mesh; registered
 * (optional) receives the normalised frames resampled at xn.  Synchronous. */
int ocl_gwr_run(ocl_gwr* g, const void* x, float* T, int T_in, float* xn, float* registered, ocl_gwr_report* report);
int ocl_gwr_destroy(ocl_gwr* g);

/* ---- batched NESTA-TV: many uneven slices advanced together, one launch per pass per iteration ---- */
int ocl_batch_create(ocl_ctx* ctx, int nslices, const ocl_slice* slices, const ocl_nesta_params* prm,
                     ocl_batch** out);
/* form = OCL_FORM_FUSED (what ocl_batch_create uses) or OCL_FORM_SPLIT (every slice of the batch then shares H, W) */
int ocl_batch_create_form(ocl_ctx* ctx, int nslices, const ocl_slice* slices, const ocl_nesta_params* prm, int form,
                          ocl_batch** out);
/* fused batch with motion compensation, mcs[i] for slice i (identity mesh); each slice's x_out holds x_ref on entry */
int ocl_batch_create_mc(ocl_ctx* ctx, int nslices, const ocl_slice* slices, const ocl_nesta_params* prm,
                        const ocl_mc_args* mcs, ocl_batch** out);
int ocl_batch_setup(ocl_batch* b);                    /* E^H y, maxima, normU, state (per slice) */
/* enqueue n_iters iterations (no host sync).  With line masks (a mask constant along every k-space row) the fused FP32
 * form runs one column pass per iteration (the Gram form); there a stage that the stop test ends takes one more
 * iteration slot, so a stop-test schedule can need up to maxIntIter slots beyond maxIntIter x maxIter: iterate until
 * ocl_batch_active reports 0 (ocl_batch_run does). */
int ocl_batch_iterate(ocl_batch* b, int n_iters);
int ocl_batch_capture(ocl_batch* b, int n_iters);     /* record n_iters iterations into a hipGraph */
int ocl_batch_replay(ocl_batch* b);                   /* launch the recorded graph */
int ocl_batch_active(ocl_batch* b, int* n_active);    /* synchronises; number of slices still iterating */
int ocl_batch_finalize(ocl_batch* b);                 /* copies each slice's result to its x_out */
int ocl_batch_report(ocl_batch* b, int slice, ocl_nesta_report* report);
int ocl_batch_run(ocl_batch* b, int chunk);           /* setup + iterate in chunks until done + finalize */
/* enqueue n_iters iterations with HIP events around each launch; average ms of pass B and pass CA and the total
 * span (synchronises).  Used to price the fused iteration against the HBM roofline. */
int ocl_batch_time(ocl_batch* b, int n_iters, double* ms_B, double* ms_CA, double* ms_total);
/* launch shape of a set-up batch (no reference counterpart): the size groups its launch table is split into (0: passes
 * B and CA run as one launch each over the whole table) and the workgroups of passes B and CA per iteration */
int ocl_batch_info(ocl_batch* b, int* launch_groups, long* wg_B, long* wg_CA);
int ocl_batch_destroy(ocl_batch* b);

/* ---- tile-shape autotuner of the FP32 fused passes (no reference counterpart: clFFT bakes its own plans,
 *      src/processes/FFT.cpp:430).  For one H x W it times each admissible column-pass variant (register four-step
 *      with direct and prefetched sampled-row reads, LDS Stockham with 4/8/16 columns per workgroup) and row-pass variant (four-step, Stockham with 4/8/16 rows)
 *      on a synthetic 4-slice batch and caches the fastest per context.  With autotuning on, batches created on the
 *      context tune every new H x W at setup; otherwise tuned shapes are used when present, else a heuristic. ---- */
int ocl_ctx_set_autotune(ocl_ctx* ctx, int on);
/* split-precision form (OCL_FORM_SPLIT, no reference counterpart): 1 (the default) runs slices with line masks (a mask
 * constant along every k-space row) through the Gram pass, whose only transforms are the column DFTs as BF16-MFMA
 * GEMMs, where it is instantiated (H in {64, 128, 192} with at most 32 sampled rows per frame, or H = 64 with 64);
 * 0 keeps the row and column passes for every slice of batches created afterwards */
int ocl_ctx_set_split_gram(ocl_ctx* ctx, int on);
/* launch grouping of mixed-size tables (no reference counterpart): a column size whose slices have at least
 * min_workgroups column-pass workgroups runs as a launch of its own with that size's kernels; the other sizes share
 * one launch of the kernels that hold every size.  Applies to tables built afterwards on this context (its batches;
 * a server keeps the library default, 2048).  min_workgroups >= 1. */
int ocl_ctx_set_group_threshold(ocl_ctx* ctx, int min_workgroups);
/* pins the tile shape of H x W (e.g. a saved tuning): ca_fast 1 / 2 selects the four-step column pass that gathers
 * the sampled rows from global memory / prefetches the next coil's rows and sensitivities, b_fast 1 the four-step
 * row pass; 0 selects the Stockham passes sized by tile_cols / rows_per_wg (powers of two <= 16).  ca_fast 3 selects
 * the Gram form for slices with line masks (one column pass per iteration, no row pass; other masks fall back to the
 * prefetching column pass).  ocl_autotune_shape returns the same codes: 3 when the Gram form won. */
int ocl_ctx_set_tiles(ocl_ctx* ctx, int H, int W, int ca_fast, int tile_cols, int b_fast, int rows_per_wg);
int ocl_autotune_shape(ocl_ctx* ctx, int H, int W, int* ca_fast, int* tile_cols, int* b_fast, int* rows_per_wg,
                       double* ms_CA_per_slice_iter, double* ms_B_per_slice_iter);

/* ---- serving: continuous batching of NESTA-TV requests (no counterpart in the reference, which runs one
 *      reconstruction per NestaUp::launch).  A worker thread with its own stream admits queued requests in arrival
 *      order while their device footprint fits the HBM budget, co-batches the iterations of every running request,
 *      and retires each one as its last continuation stage ends.  A finished request's work buffers stay with the
 *      server and are reused by later requests of the same sizes; buffers in use plus buffers kept never exceed the
 *      budget.  Inputs and outputs are caller-owned device memory of the same device and must stay valid until the
 *      request is done. ---- */
typedef struct ocl_server ocl_server;
#define OCL_SERVE_QUEUED 0
#define OCL_SERVE_RUNNING 1
#define OCL_SERVE_DONE 2
typedef struct {
    long completed, failed, queued, active;
    size_t used_bytes, peak_used_bytes, budget_bytes;
    long slice_iterations, launches;
    double busy_s, elapsed_s;
    long budget_waits;   /* requests that reached the head of the queue and waited there for HBM budget */
    long cache_hits, cache_misses;   /* request buffers reused from the server's cache / newly allocated */
    double admit_s, retire_s;        /* worker time admitting (plan, buffers, setup passes, table rebuilds) and
                                        retiring requests */
} ocl_server_stats;
int ocl_server_create(int device, size_t hbm_budget_bytes, int chunk_iters, ocl_server** out);
/* report may be NULL; when given it is filled when the request completes */
int ocl_server_submit(ocl_server* s, const ocl_slice* slice, const ocl_nesta_params* prm, ocl_nesta_report* report,
                      long* request_id);
int ocl_server_poll(ocl_server* s, long request_id, int* status, double* latency_s);
int ocl_server_wait(ocl_server* s, long request_id, double timeout_s, int* status);
int ocl_server_request_times(ocl_server* s, long request_id, double* t_submit, double* t_admit, double* t_done);
int ocl_server_get_stats(ocl_server* s, ocl_server_stats* stats);
int ocl_server_destroy(ocl_server* s);      /* finishes every submitted request, then stops */
/* failure handling without a failing device: the worker takes `message` as a device error at its next cycle, as it
 * takes a failed launch or copy: every running and queued request fails with OCL_E_HIP, the worker stops, and later
 * submissions are refused with OCL_E_HIP.  For tests of that path. */
int ocl_server_inject_failure(ocl_server* s, const char* message);

#ifdef __cplusplus
}
#endif
#endif /* OCL_AMD_H */
