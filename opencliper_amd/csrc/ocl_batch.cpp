// Batched and served NESTA-TV on one GPU.
//
// Every slice (one reconstruction) owns its device-resident NESTA state and work buffers (SliceRes).  A launch table
// (descriptor array + per-slice workgroup offsets) lists the slices a pass covers; the fused passes B, S and CA then
// advance every slice of the table by one iteration per three launches, whatever their sizes.  The batch API keeps a
// fixed table (and can record it in a hipGraph); the server rebuilds its table whenever requests are admitted or
// finish, which is what lets new reconstructions join while others are mid-solve (continuous batching).
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <map>
#include <condition_variable>
#include <deque>
#include <thread>
#include <tuple>
#include <mutex>

#include "ocl_bufcache.hpp"
#include "ocl_internal.hpp"

// ---------------------------------------------------------------------------------------------------------------------
// Static scratch of the allocation-free entry point (ocl_nesta_run_noalloc).  Each device holds its own copy of the
// array for the life of the process; its address on a device is looked up once, at the first context creation on
// that device, and only that address is used by host code.  Calls of the allocation-free entry point on one device
// share the array (they must not overlap); the array keeps the last call's contents, and every buffer carved from
// it is written before it is read, as with hipMalloc'ed memory.
// ---------------------------------------------------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[OCL_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
std::map<int, char*> g_scratch_ptr;                 // device -> address of that device's scratch array
struct ScratchArena {
    char* base = nullptr;
    size_t used = 0;
    bool overflow = false;                          // a request did not fit
};
thread_local ScratchArena* t_arena = nullptr;       // set while an allocation-free call runs on this thread

char* scratch_of(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    return it == g_scratch_ptr.end() ? nullptr : it->second;
}
}  // namespace

int scratch_init_device(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr.count(device)) return OCL_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(OCL_E_HIP, std::string("scratch address lookup: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = (char*)ptr;
    return OCL_OK;
}

bool dev_scratch_active() { return t_arena != nullptr; }

hipError_t dev_malloc(void** p, size_t bytes) {
    if (!t_arena) return hipMalloc(p, bytes);
    const size_t off = (t_arena->used + 255) & ~(size_t)255;
    if (off > OCL_SCRATCH_BYTES || bytes > OCL_SCRATCH_BYTES - off) {
        *p = nullptr;
        t_arena->overflow = true;
        return hipErrorOutOfMemory;
    }
    *p = t_arena->base + off;
    t_arena->used = off + bytes;
    return hipSuccess;
}

hipError_t dev_free(void* p) {
    if (!p) return hipSuccess;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        for (const auto& kv : g_scratch_ptr)
            if ((char*)p >= kv.second && (char*)p < kv.second + OCL_SCRATCH_BYTES) return hipSuccess;
    }
    return hipFree(p);
}

// Small host <-> device transfers through the context's page-locked buffer (ocl_ctx::zc_host), moved by a copy kernel
// on the context's stream instead of the copy engines, where they would wait behind a caller's large uploads.  Both
// calls return once the copies are done.  A transfer that does not fit the buffer, or whose size or address is not a
// multiple of 4 bytes, goes through hipMemcpyAsync instead.
struct ZcSeg {
    void* dev;         // device address
    void* host;        // host address (source of zc_to_device, destination of zc_from_device)
    size_t bytes;
};

static bool zc_fits(const std::vector<ZcSeg>& segs) {
    size_t off = 0;
    for (const ZcSeg& g : segs) {
        if (g.bytes % 4 || (uintptr_t)g.dev % 4) return false;
        off = ((off + 255) & ~(size_t)255) + g.bytes;
    }
    return off <= ZC_BYTES;
}

int zc_to_device(ocl_ctx* c, const std::vector<ZcSeg>& segs) {
    if (!c->zc_host || !zc_fits(segs)) {
        for (const ZcSeg& g : segs)
            if (g.bytes) HIPCHK(hipMemcpyAsync(g.dev, g.host, g.bytes, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return OCL_OK;
    }
    HIPCHK(hipStreamSynchronize(c->stream));     // the buffer may still feed an earlier copy
    size_t off = 0;
    for (const ZcSeg& g : segs) {
        off = (off + 255) & ~(size_t)255;
        std::memcpy(c->zc_host + off, g.host, g.bytes);
        HIPCHK(launch_copy_words(c->zc_dev + off, g.dev, g.bytes, c->stream));
        off += g.bytes;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return OCL_OK;
}

int zc_from_device(ocl_ctx* c, const std::vector<ZcSeg>& segs) {
    if (!c->zc_host || !zc_fits(segs)) {
        for (const ZcSeg& g : segs)
            if (g.bytes) HIPCHK(hipMemcpyAsync(g.host, g.dev, g.bytes, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return OCL_OK;
    }
    size_t off = 0;
    std::vector<size_t> offs;
    for (const ZcSeg& g : segs) {
        off = (off + 255) & ~(size_t)255;
        offs.push_back(off);
        HIPCHK(launch_copy_words(g.dev, c->zc_dev + off, g.bytes, c->stream));
        off += g.bytes;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < segs.size(); ++i) std::memcpy(segs[i].host, c->zc_host + offs[i], segs[i].bytes);
    return OCL_OK;
}

namespace {

struct SliceSetup {
    double maxX, maxUX, normU, lambda, muf, mu0, gamma, gammat;
    int ne_iters;
};

// The server's block cache (ocl_bufcache.hpp) over this library's device allocation and the server's stream.
struct HipBlockDev {
    hipStream_t stream = nullptr;
    int alloc(void** p, size_t n) {
        const hipError_t e = dev_malloc(p, n);
        return e == hipSuccess ? 0 : e == hipErrorOutOfMemory ? kOom : (int)e;
    }
    int release(void* p) { return (int)dev_free(p); }
    int sync() { return (int)hipStreamSynchronize(stream); }
    static constexpr int kOom = (int)hipErrorOutOfMemory;
};
using BufCache = BlockCache<HipBlockDev>;

// Device resources of one slice.
struct SliceRes {
    ocl_ctx* c = nullptr;
    ocl_slice in{};
    ocl_nesta_params prm{};
    SliceDesc d{};
    SliceSetup setup{};
    std::vector<int16_t> rows, rowpos;
    std::vector<int> nrows;
    std::vector<uint8_t> rowmask;     // [F][nmax] carried row is a mask row (Gram form)
    std::vector<void*> allocs;
    size_t bytes = 0;
    bool allocated = false;
    bool has_mc = false;      // motion compensation (argsMC) for this slice
    ocl_mc_args mc{};
    // a server's slice: its buffers are carved from one block of the server's cache (blk_need bytes asked for, the
    // block may be larger); a buffer that does not fit the block takes a block of its own (`extra`)
    BufCache* cache = nullptr;
    char* blk = nullptr;
    size_t blk_need = 0, blk_size = 0, blk_used = 0;
    std::vector<std::pair<void*, size_t>> extra;

    ~SliceRes() {
        if (cache) {
            if (blk) cache->give(blk, blk_size);
            for (auto& e : extra) cache->give(e.first, e.second);
            return;
        }
        if (allocs.empty()) return;
        hipStreamSynchronize(c->stream);
        for (void* p : allocs) dev_free(p);
    }
    size_t device_bytes() const {
        if (!cache) return bytes;
        size_t n = blk ? blk_size : 0;
        for (const auto& e : extra) n += e.second;
        return n;
    }
    int alloc(size_t n, void** p) {
        n = std::max<size_t>(n, 16);
        if (cache) {
            const size_t a = (n + 255) & ~(size_t)255;
            hipError_t e = hipSuccess;
            if (!blk && blk_need) {
                void* b = nullptr;
                e = (hipError_t)cache->take(&b, std::max(blk_need, a), &blk_size);
                if (e == hipSuccess) blk = (char*)b;
            }
            if (e == hipSuccess && blk && blk_used + a <= blk_size) {
                *p = blk + blk_used;
                blk_used += a;
            } else if (e == hipSuccess) {
                size_t got = 0;
                e = (hipError_t)cache->take(p, a, &got);
                if (e == hipSuccess) extra.push_back({*p, got});
            }
            if (e == hipErrorOutOfMemory) return fail(OCL_E_NOMEM, "NESTA: out of device memory");
            HIPCHK(e);
            bytes += n;
            return OCL_OK;
        }
        hipError_t e = dev_malloc(p, n);
        if (e == hipErrorOutOfMemory) return fail(OCL_E_NOMEM, "NESTA: out of device memory");
        HIPCHK(e);
        allocs.push_back(*p);
        bytes += n;
        return OCL_OK;
    }
};

// Column-pass workgroups from which one (column, row) size of a mixed-size table gets launches of its own.  Measured
// (64 uneven slices / 20 s served trace): 512 gives served p99 203 ms, 1024 179 ms, 2048 148-152 ms with the uneven
// batch at 133-134 recons/s for all three, 4096 132 recons/s.
// Pixels per temporal-TV reduction workgroup of pass B (each covers OCL_U_FRAMES frames of them).  Measured on 16 x
// (20, 8, 256^2): B+S 0.179 ms at 256, 0.171 at 512, 0.180 at 1024 per iteration.
#ifndef OCL_U_CHUNK
#define OCL_U_CHUNK 512
#endif
#ifndef OCL_GROUP_MIN_WG
#define OCL_GROUP_MIN_WG 2048
#endif
static_assert(OCL_U_CHUNK >= 1, "OCL_U_CHUNK: at least one pixel per temporal-TV workgroup");
static_assert(OCL_GROUP_MIN_WG >= 1, "OCL_GROUP_MIN_WG: a launch group needs at least one workgroup");

// Tile shape of the generic (non four-step) passes; the four-step passes always use 16 x 16.
void choose_tiles(long total_cols, const ocl_slice& s, int* tile_cols, int* rows_per_wg) {
    int tc = (total_cols / 16 >= 2048) ? 16 : 8;
    while (tc > 1 && ((long)s.H * tc > (long)OCL_MAXE * OCL_NT ||
                      sizeof(float2) * ((size_t)2 * s.H * (tc + 1) + s.H) > LDS_MAX))
        tc >>= 1;
    *tile_cols = tc;
    *rows_per_wg = std::min(16, pow2_floor(std::max(1, 2048 / s.W)));
}

size_t lds_B(const SliceDesc& d) {
    if (d.prec) return mf_b_lds_bytes(d.W);
    return d.fastW ? sizeof(float2) * fast_b_lds(d.fastW)
                   : sizeof(float2) * ((size_t)2 * d.rows_per_wg * (d.W + 1) + d.W);
}
size_t lds_CA(const SliceDesc& d) {
    if (d.prec) return mf_ca_lds_bytes(d.H, d.kp);
    return d.fastH ? sizeof(float2) * (fast_ca_lds(d.fastH) + (d.ca_stage ? 2 * 16 * ((size_t)d.nmax + 1) : 0))
                   : sizeof(float2) * ((size_t)2 * d.H * (d.tile_cols + 1) + d.H);
}

size_t lds_G(const SliceDesc& d) {
    return d.prec ? mf_gs_lds_bytes(d.H, d.kp) : sizeof(float2) * (size_t)gram_lds(d.fastH, d.nmax);
}

size_t lds_CA_generic(int H, int tc) { return sizeof(float2) * ((size_t)2 * H * (tc + 1) + H); }
size_t lds_B_generic(int W, int rb) { return sizeof(float2) * ((size_t)2 * rb * (W + 1) + W); }

int tune_shape(ocl_ctx* c, int H, int W);

// Tuned tile shapes are shared by every autotuning context of the process on the same device: the first context to
// tune a shape publishes it and later ones adopt it, so two contexts (a server and a solo caller, say) run the same
// kernels for a shape and give bit-identical results.  (Two correct variants, e.g. the Stockham and the four-step
// column pass, round differently, and NESTA's iterations amplify that to ~1e-4 over long schedules.)  Shapes pinned
// with ocl_ctx_set_tiles stay per context.
namespace {
std::mutex g_tuned_mu;
std::map<std::tuple<int, int, int>, TileCfg> g_tuned;    // by (device, H, W)
}

// The tile shape for (H, W): the context's own (pinned or adopted) entry, else, when `tune`, the device's published
// one, tuned and published first if there is none.  *out = nullptr: the heuristic tiles apply.
int tiles_for(ocl_ctx* c, int H, int W, bool tune, const TileCfg** out) {
    *out = nullptr;
    auto it = c->tiles.find({H, W});
    if (it == c->tiles.end()) {
        if (!tune) return OCL_OK;
        std::lock_guard<std::mutex> lk(g_tuned_mu);         // one tuning run per shape and device
        auto g = g_tuned.find(std::make_tuple(c->device, H, W));
        if (g == g_tuned.end()) {
            CHK(tune_shape(c, H, W));
            g = g_tuned.emplace(std::make_tuple(c->device, H, W), c->tiles[{H, W}]).first;
        }
        c->tiles[{H, W}] = g->second;
        it = c->tiles.find({H, W});
    }
    *out = &it->second;
    return OCL_OK;
}

// Step 1 of admission: plans, sampled-row lists (one small device round trip) and the exact device footprint.
// force (tuning runs only) fixes the tile shape; otherwise the context's tuned shape for (H, W) is used when there is
// one (tuned on first use when autotuning is on), else the heuristic of choose_tiles.
int slice_plan(ocl_ctx* c, const ocl_slice& s, const ocl_nesta_params& prm, long total_cols, int prec, SliceRes* r,
               const TileCfg* force = nullptr) {
    if (s.F < 1 || s.C < 1 || s.H < 1 || s.W < 1 || !s.y || !s.S || !s.M || !s.x_out)
        return fail(OCL_E_INVALID, "NESTA: slice has bad dimensions or null pointers");
    if (s.H > 512 || s.W > 1024) return fail(OCL_E_UNSUPPORTED, "NESTA: fused form supports H <= 512 and W <= 1024");
    if (prec && (!mfma_size_supported(s.H) || !mfma_size_supported(s.W)))
        return fail(OCL_E_UNSUPPORTED, "NESTA: the split-precision form supports H, W in {64, 128, 192, 256}");
    r->c = c;
    r->in = s;
    r->prm = prm;
    SliceDesc& d = r->d;
    d.F = s.F; d.C = s.C; d.H = s.H; d.W = s.W;
    d.N = (long)s.H * s.W;
    bool want_gram = false;
    CHK(make_plan(c, s.H, &d.ph));
    CHK(make_plan(c, s.W, &d.pw));
    d.fastH = fast_p(s.H);
    d.fastW = fast_p(s.W);
    choose_tiles(total_cols, s, &d.tile_cols, &d.rows_per_wg);
    if (!prec) {
        const TileCfg* cfg = force;
        // an allocation-free call never tunes (tuning allocates): it takes a tuned or pinned shape when the context
        // has one, else the heuristic
        if (!cfg) CHK(tiles_for(c, s.H, s.W, c->autotune != 0 && !dev_scratch_active(), &cfg));
        if (cfg && !cfg->ca_fast) { d.fastH = 0; d.tile_cols = cfg->tile_cols; }
        // the Gram form when the mask allows it and no tuned or pinned shape asks otherwise; else (and for masks
        // that vary along a row) the prefetching four-step column pass unless the shape asks for the direct one
        want_gram = d.fastH && (!cfg || cfg->ca_fast == 3);
        d.ca_stage = (d.fastH && (!cfg || cfg->ca_fast >= 2)) ? 1 : 0;
        if (cfg && !cfg->b_fast) { d.fastW = 0; d.rows_per_wg = cfg->rows_per_wg; }
    }
    if (d.fastH || prec) d.tile_cols = 16;
    if (d.fastW || prec) d.rows_per_wg = 16;
    d.prec = prec;
    if (!prec && (lds_B(d) > LDS_MAX_NESTA || lds_CA(d) > LDS_MAX_NESTA))
        return fail(OCL_E_UNSUPPORTED, "NESTA: LDS tile too large");
    // rows carrying data: any sampled mask pixel or any non-zero k-space sample in the row
    const size_t fb = (size_t)s.F * s.H;
    std::vector<uint8_t> flags(fb);
    if (c->zc_host && fb <= ZC_BYTES) {
        // straight into the context's page-locked buffer: no device buffer, no copy engine
        hipLaunchKernelGGL(k_rowflags, dim3((unsigned)fb), dim3(OCL_NT), 0, c->stream, (const float2*)s.y, s.M, s.F,
                           s.C, s.H, s.W, (uint8_t*)c->zc_dev);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        std::memcpy(flags.data(), c->zc_host, fb);
    } else {
        uint8_t* flags_d = nullptr;
        HIPCHK(dev_malloc((void**)&flags_d, fb));
        hipLaunchKernelGGL(k_rowflags, dim3((unsigned)fb), dim3(OCL_NT), 0, c->stream, (const float2*)s.y, s.M, s.F,
                           s.C, s.H, s.W, flags_d);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(flags.data(), flags_d, fb, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        dev_free(flags_d);
        HIPCHK(e);
    }
    r->nrows.assign(s.F, 0);
    int nmax = 1;
    bool line_mask = true;               // every mask row constant along the row
    for (int f = 0; f < s.F; ++f) {
        for (int h = 0; h < s.H; ++h) {
            r->nrows[f] += flags[(size_t)f * s.H + h] & 1;
            line_mask &= !(flags[(size_t)f * s.H + h] & 2);
        }
        nmax = std::max(nmax, r->nrows[f]);
    }
    r->rows.assign((size_t)s.F * nmax, 0);
    r->rowpos.assign((size_t)s.F * s.H, -1);
    r->rowmask.assign((size_t)s.F * nmax, 0);
    for (int f = 0; f < s.F; ++f) {
        int j = 0;
        for (int h = 0; h < s.H; ++h)
            if (flags[(size_t)f * s.H + h] & 1) {
                r->rowpos[(size_t)f * s.H + h] = (int16_t)j;
                r->rowmask[(size_t)f * nmax + j] = (flags[(size_t)f * s.H + h] >> 2) & 1;
                r->rows[(size_t)f * nmax + j++] = (int16_t)h;
            }
    }
    d.nmax = nmax;
    if (d.ca_stage && (nmax > 16 * FAST_CA_KQ || lds_CA(d) > LDS_MAX_NESTA)) d.ca_stage = 0;
    d.gram_ok = (want_gram && !prec && line_mask && nmax <= 16 * FAST_CA_KQ &&
                 sizeof(float2) * (size_t)gram_lds(d.fastH, nmax) <= LDS_MAX_NESTA) ? 1 : 0;
    d.kp = (nmax + 31) / 32 * 32;
    if (prec && d.kp > 64)
        return fail(OCL_E_UNSUPPORTED, "NESTA: the split-precision form supports at most 64 sampled rows per frame");
    // the split-precision Gram pass: line masks, sizes it is instantiated for, unless the context pins the row and
    // column passes
    if (prec) {
        // (its buffer resources address kbuf and S with 32-bit byte offsets)
        const bool fits32 = 8.0 * s.F * s.C * ((nmax + 31) / 32 * 32) * s.W < 4294967295.0 &&
                            8.0 * s.C * s.H * s.W < 4294967295.0;
        d.gram_ok = (c->split_gram && line_mask && fits32 && mfma_gram_supported(s.H, d.kp) &&
                     mf_gs_lds_bytes(s.H, d.kp) <= LDS_MAX_NESTA) ? 1 : 0;
    }
    d.ncol_tiles = (s.W + d.tile_cols - 1) / d.tile_cols;
    d.nrow_groups = (nmax + d.rows_per_wg - 1) / d.rows_per_wg;
    d.wg_B = prec ? (s.F * s.C * d.nrow_groups + MF_B_TASKS - 1) / MF_B_TASKS : s.F * s.C * d.nrow_groups;
    d.wg_CA = prec ? s.F * ((d.ncol_tiles + MF_CA_TILES - 1) / MF_CA_TILES) : s.F * d.ncol_tiles;
    d.u_chunk = (int)std::min<long>(d.N, OCL_U_CHUNK);      // pixels per temporal-TV workgroup
    d.wg_U = (int)((d.N + d.u_chunk - 1) / d.u_chunk) * ((s.F + OCL_U_FRAMES - 1) / OCL_U_FRAMES);
    d.inv_sqrtN = (float)(1.0 / std::sqrt((double)d.N));
    d.y = (const float2*)s.y;
    d.S = (const float2*)s.S;
    d.M = s.M;
    return OCL_OK;
}

// Device footprint that slice_alloc will take (the server's admission test).
size_t slice_footprint(const SliceRes& r) {
    const SliceDesc& d = r.d;
    auto rnd = [](size_t n) { return (std::max<size_t>(n, 16) + 255) & ~(size_t)255; };
    const size_t xb = sizeof(float2) * (size_t)d.F * d.N;
    return 4 * rnd(xb) + rnd(sizeof(float2) * (size_t)d.F * d.C * d.nmax * d.W) + rnd(sizeof(int16_t) * r.rows.size()) +
           rnd(sizeof(int16_t) * r.rowpos.size()) + rnd(sizeof(int) * d.F) + rnd(sizeof(NState)) +
           rnd(sizeof(double) * d.wg_B) + rnd(sizeof(double) * 2 * d.wg_U) + rnd(sizeof(unsigned) * 4) +
           rnd(sizeof(double) * 3 * OCL_TRACE) + rnd(sizeof(double) * 2 * d.wg_CA) +
           (d.fastW && !d.prec ? rnd(sizeof(uint16_t) * (size_t)d.F * d.nmax * d.fastW) : 0) +
           (d.prec ? 2 * rnd(sizeof(uint16_t) * 4 * (size_t)d.W * d.W) +
                         2 * rnd(sizeof(uint16_t) * 4 * (size_t)d.F * d.H * d.kp)
                   : 0) +
           (d.gram_ok ? rnd(xb) + rnd(r.rowmask.size()) + rnd(sizeof(double) * 3 * d.wg_CA) : 0);
}

// Step 2 of admission: device buffers and row tables.
// Motion-compensation geometry and pass sizes of a slice (before slice_alloc)
int slice_plan_mc(SliceRes* r) {
    SliceDesc& d = r->d;
    if (!r->has_mc) {
        d.mc_xn = nullptr;
        d.wg_P = d.wg_X = d.wg_G = 0;
        return OCL_OK;
    }
    CHK(mc_box_geometry(d.H, d.W, &r->mc, &d.mc_r0, &d.mc_nr, &d.mc_c0, &d.mc_nc));
    d.mc_xn = r->mc.xn;
    d.wg_P = (int)(((long)d.F * d.N + OCL_MC_CHUNK - 1) / OCL_MC_CHUNK);
    d.wg_X = (int)(((long)d.F * d.mc_nr * d.mc_nc + OCL_MC_CHUNK - 1) / OCL_MC_CHUNK);
    d.wg_G = d.F * ((d.H + OCL_MC_TR - 1) / OCL_MC_TR) * ((d.W + OCL_MC_TC - 1) / OCL_MC_TC);
    return OCL_OK;
}

int slice_alloc(SliceRes* r) {
    if (r->allocated) return OCL_OK;
    ocl_ctx* c = r->c;
    SliceDesc& d = r->d;
    const size_t xb = sizeof(float2) * (size_t)d.F * d.N;
    void* p;
    CHK(r->alloc(xb, &p)); d.x0 = (float2*)p;
    CHK(r->alloc(xb, &p)); d.x1 = (float2*)p;
    CHK(r->alloc(xb, &p)); d.xref = (float2*)p;
    CHK(r->alloc(xb, &p)); d.Wk = (float2*)p;
    CHK(r->alloc(sizeof(float2) * (size_t)d.F * d.C * d.nmax * d.W, &p)); d.kbuf = (float2*)p;
    CHK(r->alloc(sizeof(int16_t) * r->rows.size(), &p)); d.rows = (const int16_t*)p;
    HIPCHK(hipMemcpyAsync(p, r->rows.data(), sizeof(int16_t) * r->rows.size(), hipMemcpyHostToDevice, c->stream));
    CHK(r->alloc(sizeof(int16_t) * r->rowpos.size(), &p)); d.rowpos = (const int16_t*)p;
    HIPCHK(hipMemcpyAsync(p, r->rowpos.data(), sizeof(int16_t) * r->rowpos.size(), hipMemcpyHostToDevice, c->stream));
    CHK(r->alloc(sizeof(int) * d.F, &p)); d.nrows = (const int*)p;
    HIPCHK(hipMemcpyAsync(p, r->nrows.data(), sizeof(int) * d.F, hipMemcpyHostToDevice, c->stream));
    d.mbits = nullptr;
    if (d.fastW && !d.prec) {
        const long nw = (long)d.F * d.nmax * d.fastW;
        CHK(r->alloc(sizeof(uint16_t) * nw, &p)); d.mbits = (const uint16_t*)p;
        hipLaunchKernelGGL(k_mask_bits, dim3((unsigned)((nw + OCL_NT - 1) / OCL_NT)), dim3(OCL_NT), 0, c->stream, d.M,
                           d.rows, d.nrows, d.F, d.H, d.W, d.nmax, d.fastW, (uint16_t*)p);
        HIPCHK(hipGetLastError());
    }
    CHK(r->alloc(sizeof(NState), &p)); d.st = (NState*)p;
    CHK(r->alloc(sizeof(double) * d.wg_B, &p)); d.partB = (double*)p;
    CHK(r->alloc(sizeof(double) * 2 * d.wg_U, &p)); d.partU = (double*)p;
    CHK(r->alloc(sizeof(unsigned) * 4, &p)); d.counter = (unsigned*)p;
    CHK(r->alloc(sizeof(double) * 3 * OCL_TRACE, &p)); d.trace = (double*)p;
    CHK(r->alloc(sizeof(double) * 2 * d.wg_CA, &p)); d.maxpart = (double*)p;
    d.Wk2 = nullptr;
    d.rowmask = nullptr;
    d.partG = nullptr;
    if (d.gram_ok) {
        CHK(r->alloc(xb, &p)); d.Wk2 = (float2*)p;
        CHK(r->alloc(r->rowmask.size(), &p)); d.rowmask = (const uint8_t*)p;
        HIPCHK(hipMemcpyAsync(p, r->rowmask.data(), r->rowmask.size(), hipMemcpyHostToDevice, c->stream));
        CHK(r->alloc(sizeof(double) * 3 * d.wg_CA, &p)); d.partG = (double*)p;
    }
    d.mcz = d.mcg = nullptr;
    d.mc_dmax = nullptr;
    if (d.mc_xn) {
        CHK(r->alloc(xb, &p)); d.mcz = (float2*)p;
        CHK(r->alloc(xb, &p)); d.mcg = (float2*)p;
        CHK(r->alloc(sizeof(unsigned), &p)); d.mc_dmax = (unsigned*)p;
    }
    if (d.prec) {
        // DFT matrices of the split-precision passes (row: forward, inverse; column: inverse, forward per frame)
        const size_t rowb = sizeof(uint16_t) * 4 * (size_t)d.W * d.W;
        const size_t colb = sizeof(uint16_t) * 4 * (size_t)d.F * d.H * d.kp;
        uint16_t* m[4];
        CHK(r->alloc(rowb, &p)); m[0] = (uint16_t*)p;
        CHK(r->alloc(rowb, &p)); m[1] = (uint16_t*)p;
        CHK(r->alloc(colb, &p)); m[2] = (uint16_t*)p;
        CHK(r->alloc(colb, &p)); m[3] = (uint16_t*)p;
        for (int slot = 0; slot < 4; ++slot) {
            // the Gram pass takes its inverse column DFT with the K index permuted (kind 4) in the kind-2 slot
            const int kind = (slot == 2 && d.gram_ok) ? 4 : slot;
            const int L = kind < 2 ? d.W : d.H;
            const long plane = kind < 2 ? (long)L * L : (long)L * d.kp;
            const dim3 grid((unsigned)std::min<long>((plane + OCL_NT - 1) / OCL_NT, 1024), kind < 2 ? 1 : d.F);
            hipLaunchKernelGGL(k_dft_split, grid, dim3(OCL_NT), 0, c->stream, m[slot], kind, L, d.kp, d.rows, d.nrows,
                               d.nmax);
            HIPCHK(hipGetLastError());
        }
        d.drowF = m[0]; d.drowI = m[1]; d.dcolI = m[2]; d.dcolF = m[3];
    }
    HIPCHK(hipStreamSynchronize(c->stream));    // the host-side row tables are reused by later admissions
    r->allocated = true;
    return OCL_OK;
}

// A group of a mixed-size table: the slices sharing a (column, row) four-step size, with their own descriptor copy
// and offsets (sub-ranges of the table's group arrays) so passes B and CA run that size's specialized kernels.
struct LaunchGroup {
    NestaLaunch L{};
};

// Launch table over a set of slices.
struct Table {
    ocl_ctx* c = nullptr;
    std::vector<LaunchGroup> groups;   // empty: passes B and CA run once over the whole table
    SliceDesc* dd = nullptr;
    int* offB = nullptr;
    int* offCA = nullptr;
    int* offP = nullptr;
    int* offX = nullptr;
    int* offG = nullptr;
    SliceDesc* gdd = nullptr;   // [cap] descriptors in group order
    int* goff = nullptr;        // [2 (cap + cap)] per group: B offsets then CA offsets (group size + 1 each)
    int cap = 0;
    NestaLaunch L{};
    // slices of each iteration form: passes B, S, CA (n_ng slices) and the Gram form's G, S2 (n_g slices).  A table
    // holding both runs them as separate launch groups; Lng / Lgr address the descriptors of each form for its
    // scalar step (the non-Gram ones first in group order)
    int n_ng = 0, n_g = 0;
    NestaLaunch Lng{}, Lgr{};
    std::vector<char> isg;             // per slice (table order): runs the Gram form
    ~Table() {
        if (dd) dev_free(dd);
        if (offB) dev_free(offB);
        if (offCA) dev_free(offCA);
        if (offP) dev_free(offP);
        if (offX) dev_free(offX);
        if (offG) dev_free(offG);
        if (gdd) dev_free(gdd);
        if (goff) dev_free(goff);
    }
};

// Passes B and CA over a table: once per launch group of a mixed-size table, else once
// Passes B and CA (G) over a table: once per launch group of its form, else once.  Pass B in setup mode (1) runs for
// both forms (the Gram form keeps its output, V = W F_x^{-1} y); in iteration mode only for the B/S/CA form.
hipError_t table_launch_B(const Table& t, hipStream_t s, int mode) {
    if (t.groups.empty()) return (mode == 0 && t.L.gram) ? hipSuccess : launch_nesta_B(t.L, s, mode);
    for (const LaunchGroup& g : t.groups) {
        if (mode == 0 && g.L.gram) continue;
        const hipError_t e = launch_nesta_B(g.L, s, mode);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t table_launch_CA(const Table& t, hipStream_t s, int mode) {
    if (t.groups.empty()) return t.L.gram ? hipSuccess : launch_nesta_CA(t.L, s, mode);
    for (const LaunchGroup& g : t.groups) {
        if (g.L.gram) continue;
        const hipError_t e = launch_nesta_CA(g.L, s, mode);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t table_launch_G(const Table& t, hipStream_t s, int mode) {
    if (t.groups.empty()) return t.L.gram ? launch_nesta_G(t.L, s, mode) : hipSuccess;
    for (const LaunchGroup& g : t.groups) {
        if (!g.L.gram) continue;
        const hipError_t e = launch_nesta_G(g.L, s, mode);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

int table_build(ocl_ctx* c, Table* t, const std::vector<SliceRes*>& v) {
    const int ns = (int)v.size();
    t->c = c;
    if (ns > t->cap) {
        if (t->dd) {
            hipStreamSynchronize(c->stream);
            dev_free(t->dd); dev_free(t->offB); dev_free(t->offCA); dev_free(t->offP); dev_free(t->offX); dev_free(t->offG);
            dev_free(t->gdd); dev_free(t->goff);
        }
        t->dd = nullptr; t->offB = t->offCA = t->offP = t->offX = t->offG = t->goff = nullptr;
        t->gdd = nullptr;
        const int cap = std::max(ns, 2 * t->cap);
        HIPCHK(dev_malloc((void**)&t->dd, sizeof(SliceDesc) * cap));
        HIPCHK(dev_malloc((void**)&t->offB, sizeof(int) * (cap + 1)));
        HIPCHK(dev_malloc((void**)&t->offCA, sizeof(int) * (cap + 1)));
        HIPCHK(dev_malloc((void**)&t->offP, sizeof(int) * (cap + 1)));
        HIPCHK(dev_malloc((void**)&t->offX, sizeof(int) * (cap + 1)));
        HIPCHK(dev_malloc((void**)&t->offG, sizeof(int) * (cap + 1)));
        HIPCHK(dev_malloc((void**)&t->gdd, sizeof(SliceDesc) * cap));
        HIPCHK(dev_malloc((void**)&t->goff, sizeof(int) * 4 * cap));
        t->cap = cap;
    }
    // the table's own copy of every descriptor: a table may change a slice's column-pass variant (below) without
    // touching the slice, so a slice runs the same kernels in every table it is not grouped differently in
    std::vector<SliceDesc> hd(ns);
    for (int i = 0; i < ns; ++i) hd[i] = v[i]->d;
    std::vector<int> oB(ns + 1, 0), oC(ns + 1, 0), oP(ns + 1, 0), oX(ns + 1, 0), oG(ns + 1, 0);
    NestaLaunch& L = t->L;
    L = NestaLaunch{};
    L.maxe = 4;
    int PH = ns ? v[0]->d.fastH : 0, PW = ns ? v[0]->d.fastW : 0;
    L.prec = ns ? v[0]->d.prec : 0;
    // iteration form of each slice: the Gram form where the slice allows it (motion compensation keeps passes B, S,
    // CA); a table holding both forms runs each as launch groups of its own
    std::vector<char>& isg = t->isg;
    isg.assign(ns, 0);
    int n_g = 0;
    for (int i = 0; i < ns; ++i) {
        isg[i] = (hd[i].gram_ok && !hd[i].mc_xn) ? 1 : 0;
        n_g += isg[i];
    }
    const bool mixed_forms = n_g > 0 && n_g < ns;
    L.gram = (ns > 0 && n_g == ns) ? 1 : 0;
    t->n_g = n_g;
    t->n_ng = ns - n_g;
    auto lds_col = [&](int i) { return isg[i] ? lds_G(hd[i]) : lds_CA(hd[i]); };
    // Launch groups.  In a mixed-size table, the slices of a (column, row) four-step size whose column pass has
    // enough workgroups to fill the GPU several times over (OCL_GROUP_MIN_WG) run passes B and CA as one launch each
    // with that size's specialized kernels; the other slices share one launch of the kernels that hold every size.  Those run
    // at lower occupancy, where the prefetching column pass (STAGE) hides the latency the direct pass exposes, so the
    // four-step slices of a shared mixed launch all take it, whatever their shape was tuned to alone.  (Few, small
    // groups would serialize launches that each under-fill the GPU, as in a server with a handful of requests.)
    // The Gram form's pass G has no row transform, so its slices are grouped by column size alone.
    // Within a launch, slices whose column-pass workgroups run longest (coils x column length) come first.  The
    // dispatcher hands out workgroups in index order, so the long ones then overlap the short ones instead of ending
    // the launch alone; with coil counts from 4 to 32 in one table that tail was a tenth of some launches.
    auto wg_cost = [&](int i) { return (long)(hd[i].C + 2) * hd[i].H; };
    auto longest_first = [&](std::vector<int>& s) {
        std::stable_sort(s.begin(), s.end(), [&](int a, int b) { return wg_cost(a) > wg_cost(b); });
    };
    std::vector<std::vector<int>> sets;
    std::vector<char> set_gram;
    for (int form = 0; form < 2; ++form) {
        std::vector<int> members;
        for (int i = 0; i < ns; ++i)
            if (isg[i] == form) members.push_back(i);
        if (members.empty()) continue;
        std::vector<std::vector<int>> fsets;
        std::map<std::pair<int, int>, std::vector<int>> by_size;
        for (int i : members) by_size[{hd[i].fastH, form ? 0 : hd[i].fastW}].push_back(i);
        std::vector<int> rest;
        if (by_size.size() > 1 && !L.prec) {
            for (const auto& kv : by_size) {
                long wg = 0;
                for (int i : kv.second) wg += hd[i].wg_CA;
                if (wg >= (c->group_min_wg > 0 ? c->group_min_wg : OCL_GROUP_MIN_WG)) fsets.push_back(kv.second);
                else rest.insert(rest.end(), kv.second.begin(), kv.second.end());
            }
        } else {
            rest = members;
        }
        std::sort(rest.begin(), rest.end());
        bool restMixedH = false;
        for (size_t j = 1; j < rest.size(); ++j) restMixedH |= hd[rest[j]].fastH != hd[rest[0]].fastH;
        if (restMixedH && !L.prec && !form)
            for (int i : rest) {
                SliceDesc& d = hd[i];
                if (!d.fastH || d.ca_stage || d.nmax > 16 * FAST_CA_KQ) continue;
                d.ca_stage = 1;
                if (lds_CA(d) > LDS_MAX_NESTA) d.ca_stage = 0;
            }
        if ((!fsets.empty() || mixed_forms) && !rest.empty()) fsets.push_back(rest);
        for (auto& fs : fsets) {
            longest_first(fs);
            sets.push_back(fs);
            set_gram.push_back((char)form);
        }
    }
    // a table of one set runs as one group of its own when the longest-first order differs from the table's (the
    // table's descriptor order is the slices' order, which setup and finalize address by slice)
    if (sets.empty() && ns > 1) {
        std::vector<int> all(ns);
        for (int i = 0; i < ns; ++i) all[i] = i;
        longest_first(all);
        bool moved = false;
        for (int i = 0; i < ns; ++i) moved |= all[i] != i;
        if (moved) {
            sets.push_back(all);
            set_gram.push_back((char)(n_g == ns));
        }
    }
    const bool grouped = !sets.empty();
    L.kj = 1;
    L.stage = ns ? hd[0].ca_stage : 0;
    L.mc = ns && v[0]->d.mc_xn ? 1 : 0;
    for (int i = 0; i < ns; ++i) {
        const SliceDesc& d = hd[i];
        if (d.prec != L.prec) return fail(OCL_E_INVALID, "batch: slices of one table must share the precision form");
        if ((d.mc_xn ? 1 : 0) != L.mc)
            return fail(OCL_E_INVALID, "batch: motion compensation must be on for every slice of a table or none");
        oP[i + 1] = oP[i] + d.wg_P;
        oX[i + 1] = oX[i] + d.wg_X;
        oG[i + 1] = oG[i] + d.wg_G;
        if (L.prec && (d.H != v[0]->d.H || d.W != v[0]->d.W))
            return fail(OCL_E_UNSUPPORTED, "batch: split-precision slices of one table must share H and W");
        L.kj = std::max(L.kj, d.kp / 32);
        if (d.ca_stage != L.stage) L.stage = 2;
        oB[i + 1] = oB[i] + d.wg_B + d.wg_U;
        oC[i + 1] = oC[i] + d.wg_CA;
        L.ldsB = std::max(L.ldsB, lds_B(d));
        L.ldsCA = std::max(L.ldsCA, lds_col(i));
        if (!d.fastH) {
            const int need = (d.H * d.tile_cols + OCL_NT - 1) / OCL_NT;
            while (L.maxe < need) L.maxe *= 2;
        }
        if (d.fastH != PH) PH = 0;
        if (d.fastW != PW) PW = 0;
    }
    L.dd = t->dd;
    L.offB = t->offB;
    L.offCA = t->offCA;
    L.offP = t->offP;
    L.offX = t->offX;
    L.offG = t->offG;
    L.ns = ns;
    L.totB = oB[ns];
    L.totCA = oC[ns];
    L.totP = oP[ns];
    L.totX = oX[ns];
    L.totG = oG[ns];
    L.mcR = 0;
    L.PH = PH;
    L.PW = PW;
    if (L.prec && ns) {
        L.PH = v[0]->d.H / 16;
        L.PW = v[0]->d.W / 16;
        L.ldsB = 0;
        L.ldsCA = 0;
        for (int i = 0; i < ns; ++i) L.ldsB = std::max(L.ldsB, mf_b_lds_bytes(v[i]->d.W));
        for (int i = 0; i < ns; ++i) L.ldsCA = std::max(L.ldsCA, lds_col(i));
        if (L.ldsCA > LDS_MAX_NESTA) return fail(OCL_E_UNSUPPORTED, "batch: split-precision column tile exceeds LDS");
    }
    // launch groups of a mixed-size table: descriptors in group order in gdd, per group its B and CA offsets in goff
    t->groups.clear();
    std::vector<SliceDesc> gd;
    std::vector<int> go;
    if (grouped) {
        for (size_t si = 0; si < sets.size(); ++si) {
            const std::vector<int>& idx = sets[si];
            const int gn = (int)idx.size();
            const size_t d0 = gd.size(), o0 = go.size();
            LaunchGroup g;
            NestaLaunch& G = g.L;
            G = L;
            G.gram = set_gram[si];
            G.maxe = 4;
            G.ldsB = G.ldsCA = 0;
            G.PH = L.prec ? L.PH : hd[idx[0]].fastH;     // the split-precision kernels are sized by H / 16, W / 16
            G.PW = L.prec ? L.PW : hd[idx[0]].fastW;
            G.stage = hd[idx[0]].ca_stage;
            go.resize(o0 + 2 * (gn + 1), 0);
            for (int j = 0; j < gn; ++j) {
                const SliceDesc& d = hd[idx[j]];
                gd.push_back(d);
                go[o0 + j + 1] = go[o0 + j] + d.wg_B + d.wg_U;
                go[o0 + gn + 1 + j + 1] = go[o0 + gn + 1 + j] + d.wg_CA;
                G.ldsB = std::max(G.ldsB, lds_B(d));
                G.ldsCA = std::max(G.ldsCA, lds_col(idx[j]));
                if (d.ca_stage != G.stage) G.stage = 2;
                if (!L.prec && d.fastH != G.PH) G.PH = 0;
                if (!L.prec && d.fastW != G.PW) G.PW = 0;
                if (!d.fastH) {
                    const int need = (d.H * d.tile_cols + OCL_NT - 1) / OCL_NT;
                    while (G.maxe < need) G.maxe *= 2;
                }
            }
            G.dd = t->gdd + d0;
            G.offB = t->goff + o0;
            G.offCA = t->goff + o0 + gn + 1;
            G.ns = gn;
            G.totB = go[o0 + gn];
            G.totCA = go[o0 + 2 * gn + 1];
            t->groups.push_back(g);
        }
    }
    // descriptors addressed by each form's scalar step
    t->Lng = NestaLaunch{};
    t->Lgr = NestaLaunch{};
    if (grouped) {
        t->Lng.dd = t->gdd;
        t->Lng.ns = t->n_ng;
        t->Lgr.dd = t->gdd + t->n_ng;
        t->Lgr.ns = t->n_g;
    } else {
        (L.gram ? t->Lgr : t->Lng) = L;
    }
    if (ns) {
        // hd, gd and the offset vectors are host temporaries: the copies are complete when this returns
        std::vector<ZcSeg> segs = {{t->dd, hd.data(), sizeof(SliceDesc) * ns},
                                   {t->offB, oB.data(), sizeof(int) * (ns + 1)},
                                   {t->offCA, oC.data(), sizeof(int) * (ns + 1)},
                                   {t->offP, oP.data(), sizeof(int) * (ns + 1)},
                                   {t->offX, oX.data(), sizeof(int) * (ns + 1)},
                                   {t->offG, oG.data(), sizeof(int) * (ns + 1)}};
        if (!gd.empty()) {
            segs.push_back({t->gdd, gd.data(), sizeof(SliceDesc) * gd.size()});
            segs.push_back({t->goff, go.data(), sizeof(int) * go.size()});
        }
        CHK(zc_to_device(c, segs));
    }
    return OCL_OK;
}

// Per-reconstruction setup (NestaUp.cpp:442-649): x_ref = A^H y, the forward columns of iteration 0, the maxima of
// |x_ref| and |U x_ref|, normU, and the initial device state.  Scalars are kept in double.
// With motion compensation (NestaUp.cpp:483-491) x_ref is the slice's output on entry (the previous reconstruction):
// it is copied to x0, MC(x_ref) feeds max |U x_ref|, pass CA (mode 2) only takes the forward columns, and normU is
// the norm of tTV o MC.
int slices_setup(ocl_ctx* c, Table& t, std::vector<SliceRes*>& v) {
    if (v.empty()) return OCL_OK;
    for (SliceRes* r : v) CHK(ocl_memset_zero(c, r->d.counter, sizeof(unsigned) * 4));
    if (t.L.mc) {
        for (SliceRes* r : v) {
            CHK(ocl_memcpy_d2d(c, r->d.x0, r->in.x_out, sizeof(float2) * (size_t)r->d.F * r->d.N));
            CHK(ocl_memset_zero(c, r->d.mc_dmax, sizeof(unsigned)));
        }
        HIPCHK(launch_nesta_mc_dmax(t.L, c->stream));
        HIPCHK(launch_nesta_mc_z(t.L, c->stream, 1));
        HIPCHK(table_launch_CA(t, c->stream, 2));
    } else {
        // kbuf = V = W F_x^{-1} y at the carried rows (the Gram form keeps it throughout); x0 = x_ref = A^H sqrt(N) y
        HIPCHK(table_launch_B(t, c->stream, 1));
        HIPCHK(table_launch_CA(t, c->stream, 1));
        HIPCHK(table_launch_G(t, c->stream, 1));
    }
    hipLaunchKernelGGL(k_setup_max, dim3(t.L.totCA), dim3(OCL_NT), 0, c->stream, (const SliceDesc*)t.dd,
                       (const int*)t.offCA, t.L.ns, (double*)nullptr);
    HIPCHK(hipGetLastError());
    std::vector<double> mp;
    std::vector<NState> st(v.size());
    std::vector<int> gatherR(v.size(), 0);     // per slice: gather radius of the adjoint (0: scatter form)
    if (t.L.mc) {
        // gather radius of pass G: floor(largest displacement) + 1 over the table, or the scatter form beyond
        // OCL_MC_GATHER_RMAX (the margin covers rounding of |nhor - c| near integers)
        int R = 1;
        for (size_t i = 0; i < v.size(); ++i) {
            SliceRes* r = v[i];
            unsigned bits = 0;
            HIPCHK(hipMemcpyAsync(&bits, r->d.mc_dmax, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            float D;
            std::memcpy(&D, &bits, sizeof(D));
            const int Ri = (D >= 0.f && D < 1e6f) ? (int)std::floor(D + 1e-3f) + 1 : OCL_MC_GATHER_RMAX + 1;
            R = std::max(R, Ri);
            gatherR[i] = Ri <= OCL_MC_GATHER_RMAX ? Ri : 0;
        }
        t.L.mcR = R <= OCL_MC_GATHER_RMAX ? R : 0;
    }
    std::vector<std::vector<double>> mps(v.size());
    {
        std::vector<ZcSeg> segs;
        for (size_t i = 0; i < v.size(); ++i) {
            mps[i].resize(2 * v[i]->d.wg_CA);
            segs.push_back({v[i]->d.maxpart, mps[i].data(), sizeof(double) * mps[i].size()});
        }
        CHK(zc_from_device(c, segs));
    }
    std::vector<ZcSeg> st_segs;
    for (size_t i = 0; i < v.size(); ++i) {
        SliceRes* r = v[i];
        const SliceDesc& d = r->d;
        mp.swap(mps[i]);
        double m0 = (double)LONG_MIN, m1 = (double)LONG_MIN;
        for (int q = 0; q < d.wg_CA; ++q) { m0 = std::max(m0, mp[2 * q]); m1 = std::max(m1, mp[2 * q + 1]); }
        SliceSetup& u = r->setup;
        u.maxX = (float)m0;
        u.maxUX = (float)m1;
        if (r->has_mc) CHK(normest_mc(c, d.F, d.H, d.W, &r->mc, &u.normU, &u.ne_iters, gatherR[i]));
        else CHK(ocl_normest(c, d.F, d.N, &u.normU, &u.ne_iters));
        const ocl_nesta_params& P = r->prm;
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
        S.stage_start = 1;
        S.trace_cap = OCL_TRACE;
        S.mu = u.mu0 * u.gamma;
        S.tolVar = 0.1 * u.gammat;
        S.Lmu = u.lambda * (u.normU / S.mu) + P.La;
        S.Lmu1 = 1.0 / S.Lmu;
        S.fmean[P.miniter - 1] = 9223372036854775807.0;
        if (t.isg[i]) {
            // the Gram form's first pass G takes iteration 0's coefficients from here (later ones from its scalar step)
            S.it_mu = S.mu; S.it_L1 = S.Lmu1; S.it_apk = 0.5; S.it_tau = 2.0 / 3.0;
            S.it_cur = 0; S.it_wcur = 0; S.it_sstart = 1; S.it_redo = 0;
            S.it_pending = S.active;
        }
        st_segs.push_back({d.st, &S, sizeof(NState)});
    }
    return zc_to_device(c, st_segs);
}

int table_iterate(ocl_ctx* c, const Table& t, int n) {
    if (!t.L.ns) return OCL_OK;
    for (int it = 0; it < n; ++it) {
        if (t.n_ng) {                                      // passes B, S, CA
            HIPCHK(launch_nesta_mc_z(t.L, c->stream, 0));  // motion compensation only
            HIPCHK(table_launch_B(t, c->stream, 0));
            HIPCHK(launch_nesta_S(t.Lng, c->stream));
            HIPCHK(launch_nesta_mc_g(t.L, c->stream));     // motion compensation only
            HIPCHK(table_launch_CA(t, c->stream, 0));
        }
        if (t.n_g) {                                       // the Gram form: pass G, then its scalar step
            HIPCHK(table_launch_G(t, c->stream, 0));
            HIPCHK(launch_nesta_S2(t.Lgr, c->stream));
        }
    }
    return OCL_OK;
}

// n iterations with HIP events around pass B (+ scalar step) and pass CA: average ms of each, and the total span.
int time_table(ocl_ctx* c, const Table& t, int n, double* ms_B, double* ms_CA, double* ms_total) {
    std::vector<hipEvent_t> ev(3 * (size_t)n);
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    int rc = OCL_OK;
    // [ev 3it] passes B + S [ev 3it+1] pass CA, then the Gram form's G + S2 [ev 3it+2]
    for (int it = 0; it < n && rc == OCL_OK; ++it) {
        hipError_t e = hipEventRecord(ev[3 * it], c->stream);
        if (t.n_ng) {
            if (e == hipSuccess) e = launch_nesta_mc_z(t.L, c->stream, 0);
            if (e == hipSuccess) e = table_launch_B(t, c->stream, 0);
            if (e == hipSuccess) e = launch_nesta_S(t.Lng, c->stream);
        }
        if (e == hipSuccess) e = hipEventRecord(ev[3 * it + 1], c->stream);
        if (t.n_ng) {
            if (e == hipSuccess) e = launch_nesta_mc_g(t.L, c->stream);
            if (e == hipSuccess) e = table_launch_CA(t, c->stream, 0);
        }
        if (t.n_g) {
            if (e == hipSuccess) e = table_launch_G(t, c->stream, 0);
            if (e == hipSuccess) e = launch_nesta_S2(t.Lgr, c->stream);
        }
        if (e == hipSuccess) e = hipEventRecord(ev[3 * it + 2], c->stream);
        if (e != hipSuccess) rc = fail(OCL_E_HIP, std::string("timing launch: ") + hipGetErrorString(e));
    }
    double tb = 0, tc = 0;
    float tot = 0;
    if (rc == OCL_OK) {
        hipError_t e = hipStreamSynchronize(c->stream);
        for (int it = 0; it < n && e == hipSuccess; ++it) {
            float a = 0, q = 0;
            e = hipEventElapsedTime(&a, ev[3 * it], ev[3 * it + 1]);
            if (e == hipSuccess) e = hipEventElapsedTime(&q, ev[3 * it + 1], ev[3 * it + 2]);
            tb += a;
            tc += q;
        }
        if (e == hipSuccess) e = hipEventElapsedTime(&tot, ev[0], ev[3 * (size_t)n - 1]);
        if (e != hipSuccess) rc = fail(OCL_E_HIP, std::string("timing events: ") + hipGetErrorString(e));
    }
    for (auto& e : ev) hipEventDestroy(e);
    if (rc != OCL_OK) return rc;
    if (ms_B) *ms_B = tb / n;
    if (ms_CA) *ms_CA = tc / n;
    if (ms_total) *ms_total = tot;
    return OCL_OK;
}

// Autotuner of the FP32 passes' tile shapes.  For one (H, W) it times every admissible pass-CA variant (register
// four-step when H = 16 P, LDS Stockham with 4, 8 or 16 columns per workgroup) and every pass-B variant (four-step
// when W = 16 P, Stockham with 4, 8 or 16 rows per workgroup) on a synthetic batch of 16 slices of 16 frames x 8 coils
// with a ky-line mask at acceleration 6 (H / 6 + 4 sampled rows per frame, within the prefetching column pass's
// 64-row limit up to H = 320, so both four-step column variants are really compared), and keeps the fastest of each
// (the passes are timed separately, so the two choices are independent).  Results are cached per context and
// published to the device (tiles_for).
int tune_shape(ocl_ctx* c, int H, int W) {
    std::vector<std::pair<int, int>> caC, bC;
    if (fast_p(H) && fast_ca_lds(fast_p(H)) * sizeof(float2) <= LDS_MAX_NESTA) {
        caC.push_back({1, 16});
        caC.push_back({2, 16});
    }
    for (int tc : {4, 8, 16})
        if ((long)H * tc <= (long)OCL_MAXE * OCL_NT && lds_CA_generic(H, tc) <= LDS_MAX_NESTA) caC.push_back({0, tc});
    if (fast_p(W) && fast_b_lds(fast_p(W)) * sizeof(float2) <= LDS_MAX_NESTA) bC.push_back({1, 16});
    for (int rb : {4, 8, 16})
        if (lds_B_generic(W, rb) <= LDS_MAX_NESTA) bC.push_back({0, rb});
    if (caC.empty() || bC.empty()) return fail(OCL_E_UNSUPPORTED, "autotune: no tile shape fits");
    // 16 slices: enough workgroups to fill the GPU as a real batch does (a smaller tuning batch favours the
    // higher-occupancy direct column pass, which loses at batch scale)
    const int F = 16, C = 8, NS = 16;
    const long N = (long)H * W;
    std::vector<void*> bufs;
    auto cleanup = [&]() {
        hipStreamSynchronize(c->stream);
        for (void* p : bufs) dev_free(p);
    };
    auto dalloc = [&](size_t n, void** p) -> int {
        if (dev_malloc(p, n) != hipSuccess) { (void)hipGetLastError(); return fail(OCL_E_NOMEM, "autotune: out of memory"); }
        bufs.push_back(*p);
        return OCL_OK;
    };
    void *y, *S, *M;
    std::vector<void*> xs(NS);
    int rc = dalloc(sizeof(float2) * F * C * N, &y);
    if (rc == OCL_OK) rc = dalloc(sizeof(float2) * C * N, &S);
    if (rc == OCL_OK) rc = dalloc((size_t)F * N, &M);
    for (int k = 0; k < NS && rc == OCL_OK; ++k) rc = dalloc(sizeof(float2) * F * N, &xs[k]);
    if (rc != OCL_OK) { cleanup(); return rc; }
    std::vector<uint8_t> hm((size_t)F * N, 0);
    for (int f = 0; f < F; ++f)
        for (int h = 0; h < H; ++h)
            if (h < 2 || h >= H - 2 || (h + f) % 6 == 0) std::memset(&hm[((size_t)f * H + h) * W], 1, W);
    hipError_t e = hipMemcpy(M, hm.data(), hm.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) { cleanup(); return fail(OCL_E_HIP, std::string("autotune: ") + hipGetErrorString(e)); }
    hipLaunchKernelGGL(k_fill, dim3(grid_for(F * C * N)), dim3(OCL_NT), 0, c->stream, (float2*)y, (long)F * C * N,
                       make_float2(0.f, 0.f));
    hipLaunchKernelGGL(k_fill, dim3(grid_for(C * N)), dim3(OCL_NT), 0, c->stream, (float2*)S, (long)C * N,
                       make_float2(1.0f / std::sqrt((float)C), 0.f));
    hipLaunchKernelGGL(k_add_noise_masked, dim3(grid_for(F * C * N)), dim3(OCL_NT), 0, c->stream, (float2*)y,
                       (const uint8_t*)M, F, C, N, 1.0f, 0x5EEDULL);
    ocl_nesta_params prm{1e-2, 3e-4, 1.0, 1, 3e-4, 0, 1 << 20, 0, 7};
    TileCfg best;
    best.ms_CA = best.ms_B = 1e30;
    double best_score = 1e30;
    // runs pair the i-th column-pass variant with the i-th row-pass variant; one more run times the Gram form (its
    // pass G against the column-pass variants plus the fastest pass B, since it runs no pass B per iteration)
    const bool try_gram = fast_p(H) && sizeof(float2) * (size_t)gram_lds(fast_p(H), H / 6 + 4) <= LDS_MAX_NESTA;
    const size_t n = std::max(caC.size(), bC.size());
    double gram_ms = -1.0;
    std::vector<double> ca_ms(caC.size(), 0.0);
    for (size_t i = 0; i < n + (try_gram ? 1 : 0) && rc == OCL_OK; ++i) {
        TileCfg cfg;
        const bool gram_run = i == n;
        const auto& a = gram_run ? std::pair<int, int>(3, 16) : caC[std::min(i, caC.size() - 1)];
        const auto& b = bC[std::min(i, bC.size() - 1)];
        cfg.ca_fast = a.first; cfg.tile_cols = a.second;
        cfg.b_fast = b.first; cfg.rows_per_wg = b.second;
        std::vector<std::unique_ptr<SliceRes>> res;
        std::vector<SliceRes*> view;
        for (int k = 0; k < NS && rc == OCL_OK; ++k) {
            ocl_slice sl{y, S, (const uint8_t*)M, xs[k], F, C, H, W};
            res.emplace_back(new SliceRes());
            rc = slice_plan(c, sl, prm, (long)NS * F * W, 0, res.back().get(), &cfg);
            if (rc == OCL_OK) rc = slice_alloc(res.back().get());
            view.push_back(res.back().get());
        }
        Table t;
        double ms_b = 0, ms_ca = 0;
        if (rc == OCL_OK) rc = table_build(c, &t, view);
        if (rc == OCL_OK) rc = slices_setup(c, t, view);
        if (rc == OCL_OK) rc = table_iterate(c, t, 2);
        if (rc == OCL_OK) rc = time_table(c, t, 6, &ms_b, &ms_ca, nullptr);
        if (rc != OCL_OK) break;
        if (gram_run) {
            if (t.L.gram) gram_ms = (ms_ca + ms_b) / NS;     // pass G + its scalar step
            continue;
        }
        if (i < caC.size()) ca_ms[i] = ms_ca / NS;
        if (i < bC.size() && ms_b / NS < best.ms_B) {
            best.b_fast = cfg.b_fast; best.rows_per_wg = cfg.rows_per_wg; best.ms_B = ms_b / NS;
        }
    }
    cleanup();
    if (rc != OCL_OK) return rc;
    // the direct four-step column pass must win by 5 %: on a tie the prefetching pass, whose latency hiding matters
    // more in the fuller and mixed launches of real batches, is kept.  The Gram form competes with a column pass plus
    // the fastest row pass, the iteration it replaces, and is kept unless that pair wins by 5 %: it moves fewer bytes,
    // and a shape that leaves it splits mixed batches into more launches.
    for (size_t i = 0; i < caC.size(); ++i) {
        const double score = ca_ms[i] * (caC[i].first == 1 ? 1.05 : 1.0);
        if (score < best_score) {
            best.ca_fast = caC[i].first; best.tile_cols = caC[i].second; best.ms_CA = ca_ms[i];
            best_score = score;
        }
    }
    if (gram_ms > 0 && gram_ms < 1.05 * (best_score + best.ms_B)) {
        best.ca_fast = 3; best.tile_cols = 16; best.ms_CA = gram_ms; best.ms_B = 0.0;
    }
    c->tiles[{H, W}] = best;
    return OCL_OK;
}

int pull_states(ocl_ctx* c, const std::vector<SliceRes*>& v, NState* h) {
    std::vector<ZcSeg> segs;
    for (size_t i = 0; i < v.size(); ++i) segs.push_back({v[i]->d.st, &h[i], sizeof(NState)});
    return zc_from_device(c, segs);
}

// Copies the current x of a slice (the buffer its state points at) to the caller's output.
int slice_finalize(ocl_ctx* c, SliceRes* r, const NState& s) {
    const SliceDesc& d = r->d;
    HIPCHK(hipMemcpyAsync(r->in.x_out, s.cur ? d.x1 : d.x0, sizeof(float2) * (size_t)d.F * d.N,
                          hipMemcpyDeviceToDevice, c->stream));
    return OCL_OK;
}

void fill_report(const SliceRes& r, const NState& S, const double* trace, ocl_nesta_report* rep) {
    const SliceSetup& u = r.setup;
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
    for (int q = 0; q < rep->ntrace; ++q) {
        rep->trace_f[q] = trace[3 * q];
        rep->trace_l2[q] = trace[3 * q + 1];
        rep->trace_l1[q] = trace[3 * q + 2];
    }
}

}  // namespace

// =====================================================================================================================
// batch API: a fixed set of slices
// =====================================================================================================================
struct ocl_batch {
    ocl_ctx* ctx = nullptr;
    ocl_nesta_params prm{};
    std::vector<ocl_slice> sl;
    std::vector<std::unique_ptr<SliceRes>> res;
    std::vector<SliceRes*> view;
    Table t;
    std::vector<NState> h_st;
    float2** d_outs = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int prec = 0;
    bool ready = false;
    std::vector<ocl_mc_args> mcs;   // per slice, or empty (no motion compensation)
};

extern "C" int ocl_batch_create(ocl_ctx* c, int nslices, const ocl_slice* slices, const ocl_nesta_params* prm,
                                ocl_batch** out) {
    return ocl_batch_create_form(c, nslices, slices, prm, OCL_FORM_FUSED, out);
}

extern "C" int ocl_batch_create_form(ocl_ctx* c, int nslices, const ocl_slice* slices, const ocl_nesta_params* prm,
                                     int form, ocl_batch** out) {
    if (form != OCL_FORM_FUSED && form != OCL_FORM_SPLIT)
        return fail(OCL_E_INVALID, "batch: form must be OCL_FORM_FUSED or OCL_FORM_SPLIT");
    if (!c) return fail(OCL_E_INVALID, "null context");
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
    b->prec = form == OCL_FORM_SPLIT ? 1 : 0;
    b->sl.assign(slices, slices + nslices);
    b->prm = *prm;
    *out = b;
    return OCL_OK;
}

// True when the mesh is GroupwiseRegistration's original mesh ([d][c][r] = (c + 1, r + 1)), which the fused passes
// rely on (box pixel (r, c) writes to (r, c))
static int mesh_is_identity(ocl_ctx* c, const int* x, int H, int W, bool* ok) {
    std::vector<int> h((size_t)2 * H * W);
    HIPCHK(hipMemcpy(h.data(), x, sizeof(int) * h.size(), hipMemcpyDeviceToHost));
    *ok = true;
    for (int cc = 0; cc < W && *ok; ++cc)
        for (int r = 0; r < H; ++r)
            if (h[r + (size_t)cc * H] != cc + 1 || h[r + (size_t)cc * H + (size_t)H * W] != r + 1) { *ok = false; break; }
    return OCL_OK;
}

extern "C" int ocl_batch_create_mc(ocl_ctx* c, int nslices, const ocl_slice* slices, const ocl_nesta_params* prm,
                                   const ocl_mc_args* mcs, ocl_batch** out) {
    if (!c) return fail(OCL_E_INVALID, "null context");
    if (!mcs || nslices < 1 || !slices || !out) return fail(OCL_E_INVALID, "batch: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    for (int i = 0; i < nslices; ++i) {
        if (slices[i].H < 1 || slices[i].W < 1 || slices[i].F < 1)
            return fail(OCL_E_INVALID, "batch: slice " + std::to_string(i) + " has bad dimensions");
        int r0, nr, c0, nc;
        CHK(mc_box_geometry(slices[i].H, slices[i].W, &mcs[i], &r0, &nr, &c0, &nc));
        bool ok = false;
        CHK(mesh_is_identity(c, mcs[i].x, slices[i].H, slices[i].W, &ok));
        if (!ok)
            return fail(OCL_E_UNSUPPORTED, "batch: the fused motion-compensated form needs the original (identity) mesh; "
                                           "use OCL_FORM_PLAIN for other meshes");
    }
    CHK(ocl_batch_create_form(c, nslices, slices, prm, OCL_FORM_FUSED, out));
    (*out)->mcs.assign(mcs, mcs + nslices);
    return OCL_OK;
}

extern "C" int ocl_batch_destroy(ocl_batch* b) {
    if (!b) return OCL_OK;
    hipSetDevice(b->ctx->device);
    hipStreamSynchronize(b->ctx->stream);
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->d_outs) dev_free(b->d_outs);
    delete b;
    return OCL_OK;
}

extern "C" int ocl_batch_setup(ocl_batch* b) {
    if (!b) return fail(OCL_E_INVALID, "null batch");
    ocl_ctx* c = b->ctx;
    HIPCHK(hipSetDevice(c->device));
    if (b->res.empty()) {
        long total_cols = 0;
        for (const auto& s : b->sl) total_cols += (long)s.F * s.W;
        for (size_t i = 0; i < b->sl.size(); ++i) {
            const ocl_slice& s = b->sl[i];
            b->res.emplace_back(new SliceRes());
            SliceRes* r = b->res.back().get();
            CHK(slice_plan(c, s, b->prm, total_cols, b->prec, r));
            if (!b->mcs.empty()) {
                r->has_mc = true;
                r->mc = b->mcs[i];
            }
            CHK(slice_plan_mc(r));
            CHK(slice_alloc(r));
            b->view.push_back(b->res.back().get());
        }
        CHK(table_build(c, &b->t, b->view));
        b->h_st.resize(b->view.size());
        std::vector<float2*> outs;
        for (SliceRes* r : b->view) outs.push_back((float2*)r->in.x_out);
        HIPCHK(dev_malloc((void**)&b->d_outs, sizeof(float2*) * outs.size()));
        HIPCHK(hipMemcpy(b->d_outs, outs.data(), sizeof(float2*) * outs.size(), hipMemcpyHostToDevice));
    }
    CHK(slices_setup(c, b->t, b->view));
    b->ready = true;
    return OCL_OK;
}

extern "C" int ocl_batch_iterate(ocl_batch* b, int n_iters) {
    if (!b || !b->ready) return fail(OCL_E_INVALID, "batch: not set up");
    if (n_iters < 0) return fail(OCL_E_INVALID, "batch: negative iteration count");
    HIPCHK(hipSetDevice(b->ctx->device));
    return table_iterate(b->ctx, b->t, n_iters);
}

extern "C" int ocl_batch_time(ocl_batch* b, int n_iters, double* ms_B, double* ms_CA, double* ms_total) {
    if (!b || !b->ready) return fail(OCL_E_INVALID, "batch: not set up");
    if (n_iters < 1) return fail(OCL_E_INVALID, "batch_time: need at least one iteration");
    HIPCHK(hipSetDevice(b->ctx->device));
    return time_table(b->ctx, b->t, n_iters, ms_B, ms_CA, ms_total);
}

extern "C" int ocl_ctx_set_autotune(ocl_ctx* c, int on) {
    if (!c) return fail(OCL_E_INVALID, "null context");
    c->autotune = on ? 1 : 0;
    return OCL_OK;
}

extern "C" int ocl_ctx_set_group_threshold(ocl_ctx* c, int min_workgroups) {
    if (!c) return fail(OCL_E_INVALID, "null context");
    if (min_workgroups < 1) return fail(OCL_E_INVALID, "set_group_threshold: at least one workgroup");
    c->group_min_wg = min_workgroups;
    return OCL_OK;
}

extern "C" int ocl_ctx_set_split_gram(ocl_ctx* c, int on) {
    if (!c) return fail(OCL_E_INVALID, "null context");
    c->split_gram = on ? 1 : 0;
    return OCL_OK;
}

extern "C" int ocl_ctx_set_tiles(ocl_ctx* c, int H, int W, int ca_fast, int tile_cols, int b_fast, int rows_per_wg) {
    if (!c) return fail(OCL_E_INVALID, "null context");
    if (H < 2 || W < 2 || H > 512 || W > 1024) return fail(OCL_E_INVALID, "set_tiles: bad size");
    if ((ca_fast && !fast_p(H)) || (b_fast && !fast_p(W)))
        return fail(OCL_E_INVALID, "set_tiles: the four-step passes need H, W = 16 P with P even, P <= 20");
    if (!ca_fast && (tile_cols < 1 || tile_cols > 16 || (tile_cols & (tile_cols - 1)) ||
                     (long)H * tile_cols > (long)OCL_MAXE * OCL_NT || lds_CA_generic(H, tile_cols) > LDS_MAX_NESTA))
        return fail(OCL_E_INVALID, "set_tiles: tile_cols must be a power of two <= 16 that fits the column pass");
    if (!b_fast && (rows_per_wg < 1 || rows_per_wg > 16 || (rows_per_wg & (rows_per_wg - 1)) ||
                    lds_B_generic(W, rows_per_wg) > LDS_MAX_NESTA))
        return fail(OCL_E_INVALID, "set_tiles: rows_per_wg must be a power of two <= 16 that fits the row pass");
    if (ca_fast < 0 || ca_fast > 3) return fail(OCL_E_INVALID, "set_tiles: ca_fast must be 0, 1, 2 or 3");
    TileCfg t;
    t.ca_fast = ca_fast; t.tile_cols = ca_fast ? 16 : tile_cols;
    t.b_fast = b_fast ? 1 : 0; t.rows_per_wg = b_fast ? 16 : rows_per_wg;
    c->tiles[{H, W}] = t;
    return OCL_OK;
}

extern "C" int ocl_autotune_shape(ocl_ctx* c, int H, int W, int* ca_fast, int* tile_cols, int* b_fast,
                                  int* rows_per_wg, double* ms_CA, double* ms_B) {
    if (!c) return fail(OCL_E_INVALID, "null context");
    if (H < 2 || W < 2 || H > 512 || W > 1024) return fail(OCL_E_INVALID, "autotune: bad size");
    HIPCHK(hipSetDevice(c->device));
    const TileCfg* tp = nullptr;
    CHK(tiles_for(c, H, W, true, &tp));
    const TileCfg& t = *tp;
    if (ca_fast) *ca_fast = t.ca_fast;
    if (tile_cols) *tile_cols = t.tile_cols;
    if (b_fast) *b_fast = t.b_fast;
    if (rows_per_wg) *rows_per_wg = t.rows_per_wg;
    if (ms_CA) *ms_CA = t.ms_CA;
    if (ms_B) *ms_B = t.ms_B;
    return OCL_OK;
}

extern "C" int ocl_batch_info(ocl_batch* b, int* launch_groups, long* wg_B, long* wg_CA) {
    if (!b || !b->ready) return fail(OCL_E_INVALID, "batch: not set up");
    if (launch_groups) *launch_groups = (int)b->t.groups.size();
    if (wg_B) *wg_B = b->t.L.totB;
    if (wg_CA) *wg_CA = b->t.L.totCA;
    return OCL_OK;
}

extern "C" int ocl_batch_capture(ocl_batch* b, int n_iters) {
    if (!b || !b->ready) return fail(OCL_E_INVALID, "batch: not set up");
    if (n_iters < 1) return fail(OCL_E_INVALID, "batch: capture needs at least one iteration");
    HIPCHK(hipSetDevice(b->ctx->device));
    if (b->gexec) { hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
    HIPCHK(hipStreamBeginCapture(b->ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = table_iterate(b->ctx, b->t, n_iters);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->ctx->stream, &g);
    if (rc != OCL_OK) { if (g) hipGraphDestroy(g); return rc; }
    HIPCHK(e);
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    return OCL_OK;
}

extern "C" int ocl_batch_replay(ocl_batch* b) {
    if (!b || !b->gexec) return fail(OCL_E_INVALID, "batch: no captured graph");
    HIPCHK(hipGraphLaunch(b->gexec, b->ctx->stream));
    return OCL_OK;
}

extern "C" int ocl_batch_active(ocl_batch* b, int* n_active) {
    if (!b || !b->ready || !n_active) return fail(OCL_E_INVALID, "batch: not set up");
    CHK(pull_states(b->ctx, b->view, b->h_st.data()));
    int n = 0;
    for (const auto& s : b->h_st) n += s.active ? 1 : 0;
    *n_active = n;
    return OCL_OK;
}

extern "C" int ocl_batch_finalize(ocl_batch* b) {
    if (!b || !b->ready) return fail(OCL_E_INVALID, "batch: not set up");
    const int ns = (int)b->view.size();
    hipLaunchKernelGGL(k_finalize, dim3(256, ns), dim3(OCL_NT), 0, b->ctx->stream, (const SliceDesc*)b->t.dd, ns,
                       (float2* const*)b->d_outs);
    HIPCHK(hipGetLastError());
    return OCL_OK;
}

extern "C" int ocl_batch_report(ocl_batch* b, int slice, ocl_nesta_report* rep) {
    if (!b || !b->ready || !rep || slice < 0 || slice >= (int)b->view.size())
        return fail(OCL_E_INVALID, "batch_report: bad arguments");
    CHK(pull_states(b->ctx, b->view, b->h_st.data()));
    const NState& S = b->h_st[slice];
    std::vector<double> tr(3 * (size_t)std::max(1, std::min(S.ntrace, OCL_TRACE)));
    HIPCHK(hipMemcpy(tr.data(), b->view[slice]->d.trace, sizeof(double) * tr.size(), hipMemcpyDeviceToHost));
    fill_report(*b->view[slice], S, tr.data(), rep);
    return OCL_OK;
}

extern "C" int ocl_batch_run(ocl_batch* b, int chunk) {
    if (!b) return fail(OCL_E_INVALID, "null batch");
    CHK(ocl_batch_setup(b));
    const long max_total = (long)std::max(0, b->prm.maxIntIter) * std::max(0, b->prm.maxIter);
    if (b->prm.stoptest != 1) {
        CHK(table_iterate(b->ctx, b->t, (int)max_total));
    } else {
        // a stage that the stop test ends takes one more pass in the Gram form (the recompute of its last iteration)
        if (chunk < 1) chunk = 16;
        long done = 0;
        int n_active = 1;
        const long max_passes = max_total + std::max(0, b->prm.maxIntIter);
        while (done < max_passes && n_active > 0) {
            const int step = (int)std::min<long>(chunk, max_passes - done);
            CHK(table_iterate(b->ctx, b->t, step));
            done += step;
            CHK(ocl_batch_active(b, &n_active));
        }
    }
    CHK(ocl_batch_finalize(b));
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    return OCL_OK;
}

// NestaUp::launch (FUSED or SPLIT form) with every work buffer carved from the device's static scratch array: no
// hipMalloc or hipFree on this path.  The same kernels and tile shapes as ocl_nesta_run, except that tile shapes are
// never tuned here (a context's tuned or pinned shape is used when it has one, else the heuristic).
extern "C" int ocl_nesta_run_noalloc(ocl_ctx* c, const void* y, const void* S, const uint8_t* M, void* x_out, int F,
                                     int C, int H, int W, const ocl_nesta_params* prm, int form,
                                     ocl_nesta_report* rep) {
    if (!c) return fail(OCL_E_INVALID, "null context");
    if (!prm) return fail(OCL_E_INVALID, "NESTA: null launch parameters");
    if (!y || !x_out || F < 1 || C < 1 || H < 1 || W < 1) return fail(OCL_E_INVALID, "NESTA: bad input/output");
    if (!S) return fail(OCL_E_INVALID, "NESTA: non-existing SensitivityMaps");
    if (!M) return fail(OCL_E_INVALID, "NESTA: null sampling masks");
    if (prm->miniter < 1 || prm->miniter > OCL_MAX_MINITER) return fail(OCL_E_INVALID, "NESTA: miniter out of range");
    if (form == OCL_FORM_PLAIN)
        return fail(OCL_E_UNSUPPORTED, "ocl_nesta_run_noalloc: the plain form needs four coil-domain work arrays; "
                                       "use OCL_FORM_FUSED or OCL_FORM_SPLIT");
    if (form != OCL_FORM_FUSED && form != OCL_FORM_SPLIT) return fail(OCL_E_INVALID, "NESTA: unknown form");
    if (t_arena) return fail(OCL_E_INVALID, "ocl_nesta_run_noalloc: nested call on one thread");
    const std::string too_big = "ocl_nesta_run_noalloc: a reconstruction of " + std::to_string(F) + " x " +
                                std::to_string(C) + " x " + std::to_string(H) + " x " + std::to_string(W) +
                                " needs more than the " + std::to_string(OCL_SCRATCH_BYTES >> 20) +
                                " MiB scratch array; use ocl_nesta_run";
    // the four image-sized state arrays alone (x_k twice, x_ref, W_k) are a lower bound of the footprint
    if ((double)F * H * W * sizeof(float2) * 4 > (double)OCL_SCRATCH_BYTES) return fail(OCL_E_NOMEM, too_big);
    char* base = scratch_of(c->device);
    if (!base) return fail(OCL_E_HIP, "ocl_nesta_run_noalloc: no scratch address for this device");
    HIPCHK(hipSetDevice(c->device));
    ScratchArena arena;
    arena.base = base;
    struct Scope {
        explicit Scope(ScratchArena* a) { t_arena = a; }
        ~Scope() { t_arena = nullptr; }
    } scope(&arena);
    ocl_slice s{y, S, M, x_out, F, C, H, W};
    ocl_batch* b = nullptr;
    CHK(ocl_batch_create_form(c, 1, &s, prm, form, &b));
    int rc = ocl_batch_run(b, 16);
    if (rc == OCL_OK && rep) rc = ocl_batch_report(b, 0, rep);
    ocl_batch_destroy(b);
    if (arena.overflow) return fail(OCL_E_NOMEM, too_big);
    return rc;
}

// =====================================================================================================================
// server: continuous batching with admission by HBM budget
// =====================================================================================================================
namespace {
struct ServeReq {
    long id = 0;
    ocl_slice s{};
    ocl_nesta_params prm{};
    double t_submit = 0, t_admit = 0, t_done = 0;
    int status = OCL_SERVE_QUEUED;
    std::unique_ptr<SliceRes> res;
    size_t footprint = 0;
    std::string err;
    ocl_nesta_report* rep = nullptr;   // optional, filled at completion
    bool waited = false;               // waited at the head of the queue for HBM budget
};
}  // namespace

struct ocl_server {
    ocl_ctx* c = nullptr;                    // the server's own context (stream) on the device
    size_t budget = 0;
    int chunk = 8;
    std::mutex mu;
    std::condition_variable cv, done_cv;
    std::thread th;
    bool stop = false;
    bool dead = false;                       // the worker stopped after a device error; no request is accepted
    std::string dead_msg;
    std::string injected;                    // a failure the worker takes as a device error (ocl_server_inject_failure)
    std::deque<ServeReq*> queue;
    std::map<long, std::unique_ptr<ServeReq>> reqs;
    long next_id = 1;
    size_t used = 0, peak_used = 0;
    long completed = 0, failed = 0, iterations = 0, launches = 0, n_active = 0, budget_waits = 0;
    long cache_hits = 0, cache_misses = 0;   // copies of cache.hits / misses, taken under mu
    double admit_s = 0, retire_s = 0;        // worker time spent admitting (plan, buffers, setup) and retiring
    BufCache cache;                          // request buffers kept between requests (worker thread only)
    double busy_s = 0;
    std::chrono::steady_clock::time_point t0;
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

// Marks every request that has not completed as failed (device error in the worker).
static void server_fail_all(ocl_server* sv, std::vector<ServeReq*>& active) {
    std::lock_guard<std::mutex> lk(sv->mu);
    const std::string msg = std::string("server: device error: ") + (sv->injected.empty() ? ocl_last_error()
                                                                                         : sv->injected.c_str());
    for (ServeReq* r : active) {
        r->status = OCL_E_HIP;
        r->err = msg;
        r->res.reset();
        sv->failed++;
    }
    for (ServeReq* r : sv->queue) {
        r->status = OCL_E_HIP;
        r->err = msg;
        r->res.reset();
        sv->failed++;
    }
    sv->queue.clear();
    active.clear();
    sv->used = 0;
    sv->dead = true;
    sv->dead_msg = msg;
    sv->done_cv.notify_all();
}

static void server_loop(ocl_server* sv) {
    ocl_ctx* c = sv->c;
    hipSetDevice(c->device);
    Table table;
    std::vector<ServeReq*> active;
    std::vector<SliceRes*> view;
    std::vector<NState> h;
    bool dirty = false;
    for (;;) {
        bool inject = false;
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            inject = !sv->injected.empty();
        }
        if (inject) {
            server_fail_all(sv, active);
            break;
        }
        // ---- admission: first come first served while the footprint fits the budget ----
        const double ta = sv->now();
        std::vector<ServeReq*> admitted;
        for (;;) {
            ServeReq* r = nullptr;
            {
                std::lock_guard<std::mutex> lk(sv->mu);
                if (sv->queue.empty()) break;
                r = sv->queue.front();
            }
            if (!r->res) {
                r->res.reset(new SliceRes());
                r->res->cache = &sv->cache;
                int rc = slice_plan(c, r->s, r->prm, (long)r->s.F * r->s.W, 0, r->res.get());
                if (rc == OCL_OK) r->footprint = slice_footprint(*r->res);
                if (rc == OCL_OK && r->footprint > sv->budget) {
                    rc = OCL_E_NOMEM;
                    fail(rc, "server: request exceeds the whole HBM budget");
                }
                if (rc != OCL_OK) {
                    std::lock_guard<std::mutex> lk(sv->mu);
                    r->err = ocl_last_error();
                    r->status = rc;
                    r->res.reset();
                    sv->queue.pop_front();
                    sv->failed++;
                    sv->done_cv.notify_all();
                    continue;
                }
            }
            if (sv->used + r->footprint > sv->budget) {       // wait for running requests to free memory
                if (!r->waited) {
                    r->waited = true;
                    std::lock_guard<std::mutex> lk(sv->mu);
                    sv->budget_waits++;
                }
                break;
            }
            r->res->blk_need = r->footprint;
            int rc = slice_alloc(r->res.get());
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->queue.pop_front();
            if (rc != OCL_OK) {
                r->err = ocl_last_error();
                r->status = rc;
                r->res.reset();
                sv->failed++;
                sv->done_cv.notify_all();
                continue;
            }
            r->footprint = r->res->device_bytes();   // a reused block may be larger than the request asked for
            sv->used += r->footprint;
            sv->peak_used = std::max(sv->peak_used, sv->used);
            sv->cache_hits = sv->cache.hits;
            sv->cache_misses = sv->cache.misses;
            r->t_admit = sv->now();
            r->status = OCL_SERVE_RUNNING;
            admitted.push_back(r);
        }
        if (!admitted.empty()) {
            Table tnew;
            std::vector<SliceRes*> nv;
            for (ServeReq* r : admitted) nv.push_back(r->res.get());
            int rc = table_build(c, &tnew, nv);
            if (rc == OCL_OK) rc = slices_setup(c, tnew, nv);
            if (rc != OCL_OK) {
                std::lock_guard<std::mutex> lk(sv->mu);
                for (ServeReq* r : admitted) {
                    r->err = ocl_last_error();
                    r->status = rc;
                    sv->used -= r->footprint;
                    r->res.reset();
                    sv->failed++;
                }
                sv->done_cv.notify_all();
            } else {
                for (ServeReq* r : admitted) active.push_back(r);
                dirty = true;
            }
        }
        if (!admitted.empty()) {
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->admit_s += sv->now() - ta;
        }
        if (active.empty()) {
            std::unique_lock<std::mutex> lk(sv->mu);
            if (sv->stop && sv->queue.empty()) break;
            sv->cv.wait_for(lk, std::chrono::milliseconds(1));
            continue;
        }
        // ---- advance every running request by one chunk of iterations ----
        if (dirty) {
            const double tt = sv->now();
            view.clear();
            for (ServeReq* r : active) view.push_back(r->res.get());
            if (table_build(c, &table, view) != OCL_OK) { server_fail_all(sv, active); break; }
            h.resize(view.size());
            dirty = false;
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->admit_s += sv->now() - tt;     // rebuilding the launch table counts with admission
        }
        const double tb = sv->now();
        if (table_iterate(c, table, sv->chunk) != OCL_OK || pull_states(c, view, h.data()) != OCL_OK) {
            server_fail_all(sv, active);
            break;
        }
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->busy_s += sv->now() - tb;
            sv->iterations += (long)sv->chunk * (long)view.size();
            sv->launches += (table.L.gram ? 2L : 3L) * sv->chunk;
            sv->n_active = (long)view.size();
        }
        // ---- completions ----
        const double tr = sv->now();
        bool retired = false;
        std::vector<ServeReq*> still;
        for (size_t i = 0; i < active.size(); ++i) {
            ServeReq* r = active[i];
            if (h[i].active) {
                still.push_back(r);
                continue;
            }
            slice_finalize(c, r->res.get(), h[i]);
            if (r->rep) {
                std::vector<double> tr(3 * (size_t)std::max(1, std::min(h[i].ntrace, OCL_TRACE)));
                hipMemcpy(tr.data(), r->res->d.trace, sizeof(double) * tr.size(), hipMemcpyDeviceToHost);
                fill_report(*r->res, h[i], tr.data(), r->rep);
            }
            hipStreamSynchronize(c->stream);
            std::lock_guard<std::mutex> lk(sv->mu);
            r->res.reset();
            sv->used -= r->footprint;
            r->t_done = sv->now();
            r->status = OCL_SERVE_DONE;
            sv->completed++;
            dirty = true;
            retired = true;
        }
        if (retired) {
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->retire_s += sv->now() - tr;
        }
        if (still.size() != active.size()) {
            active.swap(still);
            sv->done_cv.notify_all();
        }
    }
}

extern "C" int ocl_server_create(int device, size_t hbm_budget_bytes, int chunk_iters, ocl_server** out) {
    if (!out || hbm_budget_bytes == 0) return fail(OCL_E_INVALID, "server: bad arguments");
    ocl_server* sv = new ocl_server();
    int rc = ocl_ctx_create(device, &sv->c);
    if (rc != OCL_OK) { delete sv; return rc; }
    sv->budget = hbm_budget_bytes;
    sv->cache.limit = hbm_budget_bytes;
    sv->cache.dev.stream = sv->c->stream;
    sv->chunk = chunk_iters > 0 ? chunk_iters : 8;
    sv->t0 = std::chrono::steady_clock::now();
    sv->th = std::thread(server_loop, sv);
    *out = sv;
    return OCL_OK;
}

extern "C" int ocl_server_submit(ocl_server* sv, const ocl_slice* s, const ocl_nesta_params* prm,
                                 ocl_nesta_report* report, long* request_id) {
    if (!sv || !s || !prm || !request_id) return fail(OCL_E_INVALID, "server_submit: bad arguments");
    if (prm->miniter < 1 || prm->miniter > OCL_MAX_MINITER) return fail(OCL_E_INVALID, "NESTA: miniter out of range");
    std::unique_ptr<ServeReq> r(new ServeReq());
    r->s = *s;
    r->prm = *prm;
    r->rep = report;
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->dead) return fail(OCL_E_HIP, "server_submit: the server stopped after a device error (" + sv->dead_msg + ")");
    if (sv->stop) return fail(OCL_E_INVALID, "server_submit: server is stopping");
    r->id = sv->next_id++;
    r->t_submit = sv->now();
    *request_id = r->id;
    sv->queue.push_back(r.get());
    sv->reqs[r->id] = std::move(r);
    sv->cv.notify_all();
    return OCL_OK;
}

extern "C" int ocl_server_poll(ocl_server* sv, long id, int* status, double* latency_s) {
    if (!sv || !status) return fail(OCL_E_INVALID, "server_poll: bad arguments");
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return fail(OCL_E_INVALID, "server_poll: unknown request");
    *status = it->second->status;
    if (latency_s) *latency_s = it->second->status == OCL_SERVE_DONE ? it->second->t_done - it->second->t_submit : -1;
    return OCL_OK;
}

extern "C" int ocl_server_wait(ocl_server* sv, long id, double timeout_s, int* status) {
    if (!sv || !status) return fail(OCL_E_INVALID, "server_wait: bad arguments");
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return fail(OCL_E_INVALID, "server_wait: unknown request");
    ServeReq* r = it->second.get();
    sv->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s),
                         [&] { return r->status == OCL_SERVE_DONE || r->status < 0; });
    *status = r->status;
    if (r->status < 0) return fail(r->status, r->err);
    return OCL_OK;
}

extern "C" int ocl_server_request_times(ocl_server* sv, long id, double* t_submit, double* t_admit, double* t_done) {
    if (!sv) return fail(OCL_E_INVALID, "server: null");
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return fail(OCL_E_INVALID, "server: unknown request");
    if (t_submit) *t_submit = it->second->t_submit;
    if (t_admit) *t_admit = it->second->t_admit;
    if (t_done) *t_done = it->second->t_done;
    return OCL_OK;
}

extern "C" int ocl_server_get_stats(ocl_server* sv, ocl_server_stats* st) {
    if (!sv || !st) return fail(OCL_E_INVALID, "server_stats: bad arguments");
    std::lock_guard<std::mutex> lk(sv->mu);
    st->completed = sv->completed;
    st->failed = sv->failed;
    st->queued = (long)sv->queue.size();
    st->active = sv->n_active;
    st->used_bytes = sv->used;
    st->peak_used_bytes = sv->peak_used;
    st->budget_bytes = sv->budget;
    st->slice_iterations = sv->iterations;
    st->launches = sv->launches;
    st->busy_s = sv->busy_s;
    st->elapsed_s = sv->now();
    st->budget_waits = sv->budget_waits;
    st->cache_hits = sv->cache_hits;
    st->cache_misses = sv->cache_misses;
    st->admit_s = sv->admit_s;
    st->retire_s = sv->retire_s;
    return OCL_OK;
}

extern "C" int ocl_server_inject_failure(ocl_server* sv, const char* message) {
    if (!sv || !message || !*message) return fail(OCL_E_INVALID, "server_inject_failure: bad arguments");
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->injected = message;
    sv->cv.notify_all();
    return OCL_OK;
}

extern "C" int ocl_server_destroy(ocl_server* sv) {
    if (!sv) return OCL_OK;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
        sv->cv.notify_all();
    }
    if (sv->th.joinable()) sv->th.join();
    sv->reqs.clear();
    sv->cache.clear();
    ocl_ctx_destroy(sv->c);
    delete sv;
    return OCL_OK;
}
