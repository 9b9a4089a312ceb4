// dme_engine.hip -- host orchestration and the C ABI (include/dme.h).
//
// A context owns one HIP stream on one device, the registered pictures (with
// their lazily built pyramid levels and up-converted planes, the device
// counterpart of EncPicture's mutable caches, enc_picture.h:146-148) and the
// per-picture MEData being estimated.  Every stage runs as batched launches
// over all jobs of a call; the single-picture entry points are batches of one.
//
// Stage order follows PictureCompressor (picture_compress.cpp:76-298):
//   pel search (coarse to fine, bulk + wavefront per level)
//   -> lambda map (host, long double)
//   -> sub-pel refinement (wavefront) or the pixel-precision doubling
//   -> mode decision (bulk level 2 + superblock wavefront) -> SetDC.

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/dme_learned.h"
#include "dme_device.h"
#include "dme_internal.h"

namespace dme {

hipError_t launch_downconvert(const PlaneTask *pt, int npt, int total, hipStream_t st);
hipError_t launch_upconvert(const PlaneTask *pt, int npt, int total, hipStream_t st);
hipError_t launch_upconvert8(const PlaneTask *pt, int npt, int total, hipStream_t st);
hipError_t launch_mc(const McComp *cs, int ncs, int total, bool pad, hipStream_t st);
hipError_t launch_mc_pack(const McPackJob *jobs, int njobs, int total, hipStream_t st);
hipError_t launch_pel_full(const JobDesc *jobs, const Task *tasks, int ntasks, int total, int pp, int rp, int rows,
                           hipStream_t st);
hipError_t launch_bytes(const PlaneTask *pt, int npt, int total, hipStream_t st);
hipError_t launch_lambda(const JobDesc *jobs, const Task *tasks, int ntasks, int *flags, hipStream_t st);
hipError_t launch_pel_bulk(const JobDesc *jobs, const Task *tasks, int ntasks, int total, int level, int max_blen_x,
                           int max_blen_y, bool has_guide, bool u8, hipStream_t st);
hipError_t launch_combine(const CombTask *ct, int nct, int total, hipStream_t st);
hipError_t launch_subpel_table16(const JobDesc *jobs, const Task *tasks, int ntasks, int total, int max_blen_x,
                                 int max_blen_y, hipStream_t st);
hipError_t launch_mode_tables16(const JobDesc *jobs, const Task *tasks, int ntasks, int total, bool eighth,
                                hipStream_t st);
hipError_t launch_pel_wave(const JobDesc *jobs, const Task *tasks, int ntasks, int level, int max_rows, int max_blen_x,
                           int max_blen_y, int threads, hipStream_t st);
hipError_t launch_subpel_bulk(const JobDesc *jobs, const Task *tasks, int ntasks, int total, int max_blen_x,
                              int max_blen_y, hipStream_t st);
hipError_t launch_build_qp(const PlaneTask *pt, int npt, int total, bool up8, hipStream_t st);
hipError_t launch_check_range(const int16_t *y, size_t n, int *flag, hipStream_t st);
hipError_t launch_lp_features(const JobDesc *jobs, const Task *tasks, int ntasks, int total, uint16_t *feat, float *lab,
                              hipStream_t st);
hipError_t launch_lp_refine(const JobDesc *jobs, const Task *tasks, int ntasks, int total, const float *out4,
                            unsigned long long *stats, int max_blen_x, int max_blen_y, hipStream_t st);
hipError_t launch_lp_pel_costs(const JobDesc *jobs, const Task *tasks, int ntasks, int total, hipStream_t st);
hipError_t launch_lp_frags(const float *w, uint16_t *fwd, uint16_t *bwd, hipStream_t st);
hipError_t launch_lp_conv(const uint16_t *in, int npos, int H, int W, const uint16_t *frag, const float *bias, int relu,
                          uint16_t *out, float *out_f32, hipStream_t st);
hipError_t launch_lp_wgrad(const uint16_t *in, const uint16_t *g, int npos, int H, int W, float *dw, float *db,
                           hipStream_t st);
hipError_t launch_lp_relu_bwd(const uint16_t *act, const float *g, size_t count, uint16_t *out, hipStream_t st);
hipError_t launch_lp_adam(float *p, const float *g, float *m, float *v, int count, float lr, float b1, float b2,
                          float eps, float bc1, float bc2, hipStream_t st);
hipError_t launch_lp_wta(const float *out, const float *lab, int npos, float alpha, float beta, float scale,
                         float *dout, float *loss, hipStream_t st);
hipError_t launch_lp_zero(void *p, size_t bytes, hipStream_t st);
hipError_t launch_subpel_wave(const JobDesc *jobs, const Task *tasks, int ntasks, int max_rows, int threads, int kinds,
                              hipStream_t st);
hipError_t launch_mode_tables(const JobDesc *jobs, const Task *tasks, int ntasks, int total, int max_x0, int max_y0,
                              int max_x1, int max_y1, int kinds, hipStream_t st);
hipError_t launch_shift_vectors(const JobDesc *jobs, const Task *tasks, int ntasks, int total, int left,
                                hipStream_t st);
hipError_t launch_mode_l2(const JobDesc *jobs, const Task *tasks, int ntasks, int total, bool up8, int kinds,
                          hipStream_t st);
hipError_t launch_mode_wave(const JobDesc *jobs, int njobs, int threads, int kinds, hipStream_t st);
hipError_t launch_block_dc(const JobDesc *jobs, const Task *tasks, int ntasks, int total, int max_width,
                           hipStream_t st);
hipError_t launch_fast_subpel(const JobDesc *jobs, const Task *tasks, int ntasks, int total, int max_blen_x,
                              int max_blen_y, hipStream_t st);
hipError_t launch_fast_l2(const JobDesc *jobs, const Task *tasks, int ntasks, int total, hipStream_t st);
hipError_t launch_fast_mode(const JobDesc *jobs, const Task *tasks, int ntasks, int total, int max_x, int max_y,
                            hipStream_t st);
hipError_t launch_fast_eval(const JobDesc *jobs, const Task *tasks, int ntasks, int total, hipStream_t st);
hipError_t launch_dc_smooth(const JobDesc *jobs, const Task *tasks, int ntasks, int total, hipStream_t st);
hipError_t launch_pack(const PackSeg *segs, int nseg, uint8_t *dst, hipStream_t st);
hipError_t launch_intra_count(const JobDesc *jobs, int njobs, hipStream_t st);

#define HIP_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return set_error(DME_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

#define DME_TRY(expr)              \
    do {                           \
        int rc_ = (expr);          \
        if (rc_ != DME_OK) return rc_; \
    } while (0)

// True on a thread while it captures a batch step into a HIP graph: no
// buffer may grow then (an allocation or a stream synchronisation cannot be
// captured), so growth fails instead.
thread_local bool t_capturing = false;

// Device buffer that only grows.  A borrowed buffer points into memory it
// does not own (the no-allocation scratch): it is never freed and never
// grows, so a request beyond it fails instead of allocating.
struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    bool borrowed = false;
    DevBuf() {}
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p && !borrowed) (void)hipFree(p);
        p = nullptr;
        n = 0;
        borrowed = false;
    }
    void borrow(void *q, size_t bytes) {
        release();
        p = q;
        n = bytes;
        borrowed = true;
    }
    int ensure(size_t bytes) {
        if (bytes <= n) return DME_OK;
        if (t_capturing) return set_error(DME_ERR_STATE, "a device buffer would grow during graph capture");
        if (borrowed)
            return set_error(DME_ERR_INVALID, "scratch layout: " + std::to_string(bytes) + " bytes requested of a " +
                                                  std::to_string(n) + "-byte buffer");
        if (p) {
            (void)hipFree(p);
            p = nullptr;
            n = 0;
        }
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) return set_error(DME_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        n = bytes;
        return DME_OK;
    }
    template <class T>
    T *as() const {
        return static_cast<T *>(p);
    }
    void swap(DevBuf &o) {
        std::swap(p, o.p);
        std::swap(n, o.n);
        std::swap(borrowed, o.borrowed);
    }
};

struct Picture {
    int num = 0, W = 0, H = 0;
    uint64_t serial = 0;  // distinct for every registration (batch graphs check it)
    const int16_t *y = nullptr, *u = nullptr, *v = nullptr;  // device planes (owned or bound)
    DevBuf own[3];
    DevBuf pyr[DME_MAX_DEPTH + 1];  // levels 1..
    int pyr_levels = 0;             // levels built so far
    DevBuf up, qp;  // up-converted plane (built on request) and the quarter-pel planes
    bool up_ready = false, qp_ready = false;
    DevBuf mcp[3];  // up-converted Y, U, V for motion compensation (k_upconvert<true>: bytes, samples + 128)
    bool mcp_ready = false;
    DevBuf b8;      // byte plane with margins (k_bytes)
    int bp = 0;     // its row pitch
    bool b8_ready = false;
    DevBuf u8;      // byte up-plane with margins (the learned path's fast form)
    bool u8_ready = false;
    // combined-component ME (EncPicture::CombinedData, enc_picture.cpp:237-
    // 254): the combined picture as a picture of its own (its y in own[0],
    // its pyramid and int16 up plane built as a luma picture's are), made
    // from this picture's Y, U, V when a job needs it
    std::unique_ptr<Picture> cpic;
    bool comb_ready = false;
    bool is_comb = false;  // this is such a combined picture (no byte planes)
    int chroma_ok = -1;    // chroma within [-128, 127]: -1 not yet checked
    const uint8_t *u8_origin() const { return u8.as<uint8_t>() + (size_t)kUp8Margin * up8_pitch(W) + kUp8Margin; }
    const uint8_t *b8_origin() const { return b8.as<uint8_t>() + (size_t)kB8Margin * bp + kB8Margin; }
};

struct JobState {
    int pic_num = -1;
    int nrefs = 0, refs[2] = {0, 0};
    dme_params p;
    Geometry g;
    int stage = 0;  // 1 pel, 2 lambda, 3 subpel, 4 mode
    DevBuf vec[2], pc[2], mode, intra, bi, dc[3], split, sbc, lam, lam1, lam0, ucost, dct[3], icount;
    DevBuf lvec[2][DME_MAX_DEPTH + 1], lpc[2][DME_MAX_DEPTH + 1];
    DevBuf ptab[2], qtab[2], m1tab, m0tab;
    DevBuf lwork;  // transition maps of the device lambda stage
    DevBuf fv0[2], fv2[2], fs2[2], fsb;  // the learned path's fast form (learned jobs only)
    int bpos = 0;  // position in the current batch (device lambda flags)
    bool fast = false;  // the learned path's fast form: byte up-planes instead of quarter-pel planes
    JobDesc desc;
};

// Pinned host staging that grows on demand (freed with the context).  A
// fixed one (the no-allocation lane's, allocated when the context is
// created) never grows.
struct PinnedBuf {
    void *p = nullptr;
    size_t bytes = 0;
    bool fixed = false;
    int ensure(size_t n) {
        if (n <= bytes) return DME_OK;
        if (t_capturing) return set_error(DME_ERR_STATE, "pinned staging would grow during graph capture");
        if (fixed) return set_error(DME_ERR_INVALID, "no-allocation staging of " + std::to_string(bytes) + " bytes is too small");
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
        if (hipHostMalloc(&p, n, hipHostMallocDefault) != hipSuccess)
            return set_error(DME_ERR_NOMEM, "pinned host allocation of " + std::to_string(n) + " bytes failed");
        bytes = n;
        return DME_OK;
    }
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    void swap(PinnedBuf &o) {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
        std::swap(fixed, o.fixed);
    }
};

// A stream with the buffers its queued work reads: descriptors, task
// tables, plane tasks and the pinned staging arena of the current call.
struct Lane {
    hipStream_t stream = nullptr;
    DevBuf d_jobs, d_tasks, d_planes, d_planes16, d_comb;
    PinnedBuf arena;
    size_t arena_used = 0;
};

// Streams of a context: as many as the device's default hardware queues (4;
// more streams share queues and serialise).  A batch runs as up to that many
// parts, one per stream (dme_tuning.parts).  With stagger, each part's planes
// are built after the previous part's, so the parts run out of step: the
// latency-bound wavefront kernels of one part overlap the wide kernels of
// another instead of running at the same time as theirs (one step of 64
// 1080p pictures: 22.9 ms with two parts in step, 21.0 ms with four
// staggered parts).
constexpr int kLanes = 4;

// The built-in launch setting.
inline dme_tuning default_tuning() {
    dme_tuning t;
    t.parts = 4;
    t.stagger = 1;
    t.pel_wave_threads = 1024;
    t.subpel_wave_threads = 512;
    t.mode_wave_threads = 1024;
    t.step_ms = 0.0f;
    return t;
}

// Pinned host staging of the no-allocation lane: descriptors, task tables
// and, for a picture whose device lambda threshold is undecided, its pel
// vectors and host-made lambda maps (about 1.3 MB at 60,000 blocks).
constexpr size_t kNoallocStagingBytes = (size_t)4 << 20;

int noalloc_scratch(int device, void **out);  // dme_noalloc.hip

}  // namespace dme

struct dme_ctx {
    int device = 0;
    hipStream_t stream = nullptr;  // lane 0's stream
    dme_params params;
    bool params_set = false;
    std::map<int, std::unique_ptr<dme::Picture>> pics;
    std::map<int, std::unique_ptr<dme::JobState>> jobs;
    dme::DevBuf d_flag;
    dme::DevBuf d_lflags;  // per batch position: the device lambda maps need the host (k_lambda)
    dme::DevBuf d_mc;      // motion compensation: descriptors, weights, uploaded MvData
    dme::DevBuf d_mc_planes;  // host-array compensation: the arrays on the device
    dme::DevBuf d_mc_units;   // packed per-block MvData records (k_mc_pack)
    // batch export: the packed MEData arrays on the device and in pinned
    // host memory, and the segment table (k_pack)
    dme::DevBuf x_dev, x_segs;
    dme::PinnedBuf x_host;
    // Lanes (stream + buffers); stages queue work on the current lane.
    // Large batches run as parts on the lanes, so that one part's host work
    // and latency-bound wavefronts overlap the other parts' kernels.  The arena of a lane holds the pinned host staging of one
    // call (descriptors, task tables, plane tasks, lambda-stage vectors and
    // maps), bump-allocated so no upload waits for an earlier one; it is
    // reset when a call starts, with the streams idle.
    dme::Lane lane[dme::kLanes];
    dme::Lane *L = &lane[0];
    dme::Lane *L0 = &lane[0];            // the lane a call starts on
    dme::DevBuf *lflags = &d_lflags;     // the device lambda flags of the current call
    // No-allocation calls (dme_motion_estimate_noalloc): their pictures, job
    // and lane, whose device buffers are borrowed from this device's scratch
    // array for the call, and the lane's pinned staging, allocated here.
    void *scratch = nullptr;             // the device's scratch array (hipGetSymbolAddress)
    dme::Picture na_pic[3];
    int na_npics = 0;                    // > 0 while a no-allocation call runs
    dme::JobState na_job;
    dme::Lane na_lane;
    dme::DevBuf na_lflags, na_flag;
    // learned proposal: weights (f32 master), BF16 fragments per layer,
    // activations of a batch, per-job counters
    bool lp_ready = false;
    dme::DevBuf lp_w, lp_frag, lp_act[4], lp_out;
    // refinement counters, a pair per batch position (k_lp_refine), and the
    // totals of counters retired when the array grew
    dme::DevBuf lp_stats;
    size_t lp_slots = 0;
    bool lp_dirty = false;  // a learned batch counted since the last fold
    int lp_form = DME_LP_FORM_FAST;
    long long lp_blocks = 0, lp_fallbacks = 0;
    std::vector<dme::Task> h_tasks;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_fork = nullptr;              // the other lanes after lane 0's shared planes
    hipEvent_t ev_join[dme::kLanes] = {};      // lane 0 after each other lane
    hipEvent_t ev_planes[dme::kLanes] = {};    // a lane's planes are built (staggered starts)
    dme_stats stats;
    // launch settings: the default, the table by batch signature, the one
    // of the call that runs (fixed while dme_autotune_batch times settings)
    dme_tuning tune = dme::default_tuning();
    dme_tuning act = dme::default_tuning();
    bool act_fixed = false;
    std::map<std::string, dme_tuning> tune_table;
    std::set<dme_batch_graph *> graphs;  // live batch graphs (destroyed with the context)
    uint64_t next_serial = 1;            // Picture::serial
    // per-launch timing while profiling is enabled
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, int>> ev_used;  // (kernel kind, pool index of the start event)
};

// A captured batch step (dme_batch_graph_create): its jobs' state and the
// lane buffers and staging the captured copies and kernels read, all owned
// here so that no later call of the context can move or overwrite them.
struct dme_batch_graph {
    dme_ctx *ctx = nullptr;
    std::vector<std::unique_ptr<dme::JobState>> jobs;
    std::vector<std::pair<int, uint64_t>> pics;  // picture numbers and serials at capture
    dme::Lane lanes[dme::kLanes];                // buffers only; the streams are the context's
    dme::DevBuf lflags;
    dme_tuning tuning;
    int launches = 0;  // kernel launches of one replay
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~dme_batch_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

namespace dme {

enum KernelKind {
    KK_DOWNCONVERT, KK_UPCONVERT, KK_PEL_BULK, KK_PEL_WAVE, KK_SUBPEL_BULK, KK_SUBPEL_WAVE, KK_SHIFT,
    KK_MODE_L2, KK_MODE_TABLES, KK_MODE_WAVE, KK_BLOCK_DC, KK_DC_SMOOTH, KK_BUILD_QP, KK_LP_NET, KK_LP_REFINE,
    KK_BYTES, KK_LAMBDA, KK_MC, KK_MC_PAD, KK_MC_PACK, KK_PEL_FULL, KK_FAST_SUBPEL, KK_FAST_L2, KK_FAST_MODE,
    KK_FAST_EVAL, KK_BUILD_UP8, KK_INTRA, KK_COMBINE, KK_SUBPEL_T16, KK_MODE_T16, KK_COUNT
};
static const char *const kKernelNames[DME_NUM_KERNELS] = {
    "k_downconvert", "k_upconvert", "k_pel_bulk", "k_pel_wave", "k_subpel_bulk", "k_subpel_wave",
    "k_shift_vectors", "k_mode_l2", "k_mode_tables", "k_mode_wave", "k_block_dc", "k_dc_smooth", "k_build_qp",
    "k_lp_net", "k_lp_refine", "k_bytes", "k_lambda", "k_mc", "k_mc_pad", "k_mc_pack", "k_pel_full",
    "k_fast_subpel", "k_fast_l2", "k_fast_mode", "k_fast_eval", "k_build_up8", "k_intra_count", "k_combine",
    "k_subpel_table16", "k_mode_tables16"};
static_assert(KK_COUNT <= DME_NUM_KERNELS, "kernel kinds");

// Runs a launch; with profiling on, brackets it with a pair of events.
template <class F>
static int timed_launch(dme_ctx *ctx, int kind, int &launches, F f) {
    size_t idx = 0;
    if (ctx->profiling) {
        idx = ctx->ev_used.size() * 2;
        while (ctx->ev_pool.size() < idx + 2) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            ctx->ev_pool.push_back(e);
        }
        HIP_TRY(hipEventRecord(ctx->ev_pool[idx], ctx->L->stream));
    }
    HIP_TRY(f());
    if (ctx->profiling) {
        HIP_TRY(hipEventRecord(ctx->ev_pool[idx + 1], ctx->L->stream));
        ctx->ev_used.push_back(std::make_pair(kind, (int)idx));
    }
    ++launches;
    return DME_OK;
}

static int set_device(dme_ctx *ctx) {
    ctx->L = ctx->L0;  // entry points start on lane 0
    if (!ctx->act_fixed) ctx->act = ctx->tune;
    HIP_TRY(hipSetDevice(ctx->device));
    return DME_OK;
}

// Staging space for an upload that is queued but not yet done.  Growing the
// buffer waits for the copies already queued from it.
static int arena_alloc(dme_ctx *ctx, size_t bytes, void **out) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (ctx->L->arena_used + bytes > ctx->L->arena.bytes) {
        if (t_capturing) return set_error(DME_ERR_STATE, "the staging arena would grow during graph capture");
        HIP_TRY(hipStreamSynchronize(ctx->L->stream));
        DME_TRY(ctx->L->arena.ensure(std::max(2 * ctx->L->arena.bytes, ctx->L->arena_used + bytes + ((size_t)1 << 20))));
        ctx->L->arena_used = 0;
    }
    *out = static_cast<char *>(ctx->L->arena.p) + ctx->L->arena_used;
    ctx->L->arena_used += bytes;
    return DME_OK;
}

// Copies host data to a device buffer in stream order through the arena;
// the device buffer grows (with the stream synced) when too small.
static int upload(dme_ctx *ctx, DevBuf &dst, const void *src, size_t bytes) {
    if (bytes == 0) return DME_OK;
    if (dst.n < bytes) {
        if (t_capturing) return set_error(DME_ERR_STATE, "a device buffer would grow during graph capture");
        HIP_TRY(hipStreamSynchronize(ctx->L->stream));
        DME_TRY(dst.ensure(bytes));
    }
    void *h = nullptr;
    DME_TRY(arena_alloc(ctx, bytes, &h));
    std::memcpy(h, src, bytes);
    HIP_TRY(hipMemcpyAsync(dst.p, h, bytes, hipMemcpyHostToDevice, ctx->L->stream));
    return DME_OK;
}

// What a set of jobs needs built: pyramid depth per picture (luma pictures
// and combined pictures alike), the references' quarter-pel planes, byte
// up-planes (fast form) and int16 up planes (combined pictures), and the
// pictures whose combined picture must be made first.
struct PlaneNeeds {
    std::map<Picture *, int> levels;
    std::vector<Picture *> ups, ups8, ups16, combs;
};

static void add_once(std::vector<Picture *> &v, Picture *P) {
    if (std::find(v.begin(), v.end(), P) == v.end()) v.push_back(P);
}

// Builds the missing combined pictures (EncPicture::Combine,
// enc_picture.cpp:285-367), pyramid levels (PixelMatcher::MakePicHierarchy,
// pixel_match.cpp:159-183: each level from the previous, sizes from full
// resolution) and up-converted planes (EncPicture::UpOrigData /
// UpCombinedData, enc_picture.cpp:131-156, 256-280) of a set of pictures,
// one launch per level.
static int build_planes(dme_ctx *ctx, const PlaneNeeds &nd, int &launches) {
    const std::map<Picture *, int> &levels = nd.levels;
    const std::vector<Picture *> &ups = nd.ups, &ups8 = nd.ups8;
    {
        std::vector<CombTask> ct;
        int off = 0;
        for (Picture *P : nd.combs) {
            if (P->comb_ready) continue;
            CombTask t;
            t.y = P->y;
            t.u = P->u;
            t.v = P->v;
            t.out = P->cpic->own[0].as<int16_t>();
            t.W = P->W;
            t.H = P->H;
            t.offset = off;
            off += combine_tiles(P->W, P->H);
            ct.push_back(t);
        }
        if (!ct.empty()) {
            DME_TRY(upload(ctx, ctx->L->d_comb, ct.data(), sizeof(CombTask) * ct.size()));
            DME_TRY(timed_launch(ctx, KK_COMBINE, launches, [&] {
                return launch_combine(ctx->L->d_comb.as<CombTask>(), (int)ct.size(), off, ctx->L->stream);
            }));
            for (Picture *P : nd.combs) P->comb_ready = true;
        }
    }
    {
        // int16 up planes of the combined references (k_upconvert; their
        // quarter-pel samples are formed from it on the fly)
        std::vector<PlaneTask> ut;
        int off = 0;
        for (Picture *P : nd.ups16) {
            if (P->up_ready) continue;
            PlaneTask t;
            t.in = P->y;
            t.out = P->up.as<int16_t>();
            t.in_w = P->W;
            t.ow = P->W;
            t.oh = P->H;
            t.offset = off;
            off += upconvert_tiles(P->W, P->H);
            ut.push_back(t);
        }
        if (!ut.empty()) {
            DME_TRY(upload(ctx, ctx->L->d_planes16, ut.data(), sizeof(PlaneTask) * ut.size()));
            DME_TRY(timed_launch(ctx, KK_UPCONVERT, launches, [&] {
                return launch_upconvert(ctx->L->d_planes16.as<PlaneTask>(), (int)ut.size(), off, ctx->L->stream);
            }));
            for (Picture *P : nd.ups16) P->up_ready = true;
        }
    }
    int maxl = 0;
    for (auto &kv : levels) maxl = std::max(maxl, kv.second);
    std::vector<PlaneTask> all;
    std::vector<std::pair<size_t, int>> ranges;  // (first, count) per launch, levels 1..maxl then up
    for (int l = 1; l <= maxl; ++l) {
        const size_t first = all.size();
        int off = 0;
        for (auto &kv : levels) {
            Picture &P = *kv.first;
            if (P.pyr_levels >= l || kv.second < l) continue;
            const int ow = P.W >> l, oh = P.H >> l;
            DME_TRY(P.pyr[l].ensure(sizeof(int16_t) * (size_t)ow * oh));
            PlaneTask t;
            t.in = l == 1 ? P.y : P.pyr[l - 1].as<int16_t>();
            t.out = P.pyr[l].as<int16_t>();
            t.in_w = P.W >> (l - 1);
            t.ow = ow;
            t.oh = oh;
            t.offset = off;
            off += ((ow + kDownTileW - 1) / kDownTileW) * ((oh + kDownTileH - 1) / kDownTileH);
            all.push_back(t);
        }
        ranges.push_back(std::make_pair(first, off));
    }
    {
        // quarter-pel planes of the references, straight from the pictures
        const size_t first = all.size();
        int off = 0;
        for (Picture *P : ups) {
            if (P->qp_ready) continue;
            const size_t n = (size_t)(P->W + 2) * 4 * (P->H + 2);  // dwords
            DME_TRY(P->qp.ensure(sizeof(uint32_t) * n));
            PlaneTask t;
            t.in = P->y;
            t.out = P->qp.as<int16_t>();
            t.in_w = P->W;
            t.ow = P->W;
            t.oh = P->H;
            t.offset = off;
            off += ((P->W + 2 + 63) / 64) * ((P->H + 2 + 15) / 16);  // 64 x 16 positions per tile
            all.push_back(t);
        }
        ranges.push_back(std::make_pair(first, off));
    }
    {
        // byte up-planes of the references of fast-form jobs
        const size_t first = all.size();
        int off = 0;
        for (Picture *P : ups8) {
            if (P->u8_ready) continue;
            DME_TRY(P->u8.ensure((size_t)up8_pitch(P->W) * up8_rows(P->H)));
            PlaneTask t;
            t.in = P->y;
            t.out = P->u8.as<int16_t>();
            t.in_w = P->W;
            t.ow = P->W;
            t.oh = P->H;
            t.offset = off;
            off += up8_tiles_x(P->W) * up8_tiles_y(P->H);
            all.push_back(t);
        }
        ranges.push_back(std::make_pair(first, off));
    }
    {
        // byte planes of every picture of the batch
        const size_t first = all.size();
        int off = 0;
        for (auto &kv : levels) {
            Picture &P = *kv.first;
            if (P.b8_ready || P.is_comb) continue;  // combined pictures are searched on 16-bit samples
            P.bp = (P.W + 2 * kB8Margin + 63) & ~63;
            const size_t dw = (size_t)(P.H + 2 * kB8Margin) * (P.bp / 4);
            DME_TRY(P.b8.ensure(sizeof(uint32_t) * dw));
            PlaneTask t;
            t.in = P.y;
            t.out = P.b8.as<int16_t>();
            t.in_w = P.W;
            t.ow = P.bp;
            t.oh = P.H;
            t.offset = off;
            off += (int)((dw + 255) / 256);
            all.push_back(t);
        }
        ranges.push_back(std::make_pair(first, off));
    }
    if (all.empty()) return DME_OK;
    DME_TRY(upload(ctx, ctx->L->d_planes, all.data(), sizeof(PlaneTask) * all.size()));
    const PlaneTask *dp = ctx->L->d_planes.as<PlaneTask>();
    for (int l = 1; l <= maxl; ++l) {
        const auto &rg = ranges[l - 1];
        const int n = (int)(ranges[l].first - rg.first);  // the next range starts where this one ends
        if (rg.second == 0) continue;
        DME_TRY(timed_launch(ctx, KK_DOWNCONVERT, launches,
                             [&] { return launch_downconvert(dp + rg.first, n, rg.second, ctx->L->stream); }));
    }
    const auto &rq = ranges[ranges.size() - 3], &ru = ranges[ranges.size() - 2], &rb = ranges.back();
    if (rb.second) {
        const int n = (int)(all.size() - rb.first);
        DME_TRY(timed_launch(ctx, KK_BYTES, launches,
                             [&] { return launch_bytes(dp + rb.first, n, rb.second, ctx->L->stream); }));
    }
    if (rq.second) {
        const int n = (int)(ru.first - rq.first);
        DME_TRY(timed_launch(ctx, KK_BUILD_QP, launches,
                             [&] { return launch_build_qp(dp + rq.first, n, rq.second, false, ctx->L->stream); }));
    }
    if (ru.second) {
        const int n = (int)(rb.first - ru.first);
        DME_TRY(timed_launch(ctx, KK_BUILD_UP8, launches,
                             [&] { return launch_build_qp(dp + ru.first, n, ru.second, true, ctx->L->stream); }));
    }

    for (auto &kv : levels) kv.first->pyr_levels = std::max(kv.first->pyr_levels, kv.second);
    for (Picture *P : ups) P->qp_ready = true;
    for (Picture *P : ups8) P->u8_ready = true;
    for (auto &kv : levels) kv.first->b8_ready = true;
    return DME_OK;
}

static Picture *find_pic(dme_ctx *ctx, int num) {
    if (ctx->na_npics) {
        for (int i = 0; i < ctx->na_npics; ++i)
            if (ctx->na_pic[i].num == num) return &ctx->na_pic[i];
        return nullptr;
    }
    auto it = ctx->pics.find(num);
    return it == ctx->pics.end() ? nullptr : it->second.get();
}

// Allocates a job's MEData and scratch, fills its descriptor.
// k_pel_full's LDS for blocks of xl x yl and a window of radius (xr, yr):
// minima and the 5 x 5 table, the block, the window with margins of at least 2.
constexpr size_t kFullLdsMax = 160 * 1024;
static int full_pitch(int xl, int xr) { return (((xl + 2 * std::max(xr, 2) + 3) / 4) + 5) | 1; }
static size_t full_lds_bytes(int xl, int yl, int xr, int yr) {
    const size_t pp = (size_t)((xl + 15) / 16) * 4;
    return sizeof(int) * (8 + 26) + 4 * (pp * 64 + (size_t)(yl + 2 * std::max(yr, 2)) * full_pitch(xl, xr));
}

// Chroma of a picture within 8-bit range (checked once per registration):
// EncPicture::Combine's float squares and sums are then exact, and motion
// compensation's up-converted planes fit in bytes.
static int check_chroma(dme_ctx *ctx, Picture *P, const char *what = "combined-component ME") {
    if (P->chroma_ok < 0) {
        const size_t cn = (size_t)(P->W / 2) * (P->H / 2);
        DME_TRY(ctx->d_flag.ensure(sizeof(int)));
        int bad = 0;
        HIP_TRY(hipMemcpyAsync(ctx->d_flag.p, &bad, sizeof(int), hipMemcpyHostToDevice, ctx->L->stream));
        HIP_TRY(launch_check_range(P->u, cn, ctx->d_flag.as<int>(), ctx->L->stream));
        HIP_TRY(launch_check_range(P->v, cn, ctx->d_flag.as<int>(), ctx->L->stream));
        HIP_TRY(hipMemcpyAsync(&bad, ctx->d_flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->L->stream));
        HIP_TRY(hipStreamSynchronize(ctx->L->stream));
        P->chroma_ok = bad ? 0 : 1;
    }
    if (!P->chroma_ok)
        return set_error(DME_ERR_INVALID, std::string(what) + ": chroma of picture " + std::to_string(P->num) +
                                              " must lie in [-128, 127] (8-bit samples minus 128)");
    return DME_OK;
}

// A combined-component job (EncPicture::DataForME(true)): its pictures need
// chroma; their combined pictures are made (buffers allocated here, so that
// no later stage allocates).  The learned forms, full search and the
// no-allocation entry point search luma only.
static int prepare_combined(dme_ctx *ctx, Picture *P, Picture *const *R, int nrefs, const dme_params &p,
                            bool learned) {
    if (learned)
        return set_error(DME_ERR_INVALID, "the learned forms estimate on luma: combined_me needs the plain search");
    if (p.full_search)
        return set_error(DME_ERR_INVALID, "combined-component ME with full search is not supported");
    if (ctx->na_npics)
        return set_error(DME_ERR_INVALID, "the no-allocation entry point estimates on luma (combined_me = 0)");
    Picture *all[3] = {P, R[0], nrefs > 1 ? R[1] : nullptr};
    for (Picture *Q : all) {
        if (!Q) continue;
        if (!Q->u || !Q->v)
            return set_error(DME_ERR_INVALID, "combined-component ME needs the chroma of picture " +
                                                  std::to_string(Q->num));
        DME_TRY(check_chroma(ctx, Q));
        if (!Q->cpic) {
            std::unique_ptr<Picture> C(new Picture());
            C->num = Q->num;
            C->W = Q->W;
            C->H = Q->H;
            C->is_comb = true;
            DME_TRY(C->own[0].ensure(sizeof(int16_t) * (size_t)Q->W * Q->H));
            C->y = C->own[0].as<int16_t>();
            Q->cpic = std::move(C);
            Q->comb_ready = false;
        }
    }
    for (int r = 0; r < nrefs; ++r) DME_TRY(R[r]->cpic->up.ensure(sizeof(int16_t) * 4 * (size_t)R[r]->W * R[r]->H));
    return DME_OK;
}

static int prepare_job(dme_ctx *ctx, JobState &J, const dme_params &p, int pic_num, const int *refs, int nrefs,
                       bool learned = false) {
    // buffers below may be reallocated: nothing queued may still use them
    HIP_TRY(hipStreamSynchronize(ctx->L->stream));
    Picture *P = find_pic(ctx, pic_num);
    if (!P) return set_error(DME_ERR_INVALID, "picture " + std::to_string(pic_num) + " is not registered");
    if (nrefs < 1 || nrefs > 2 || !refs) return set_error(DME_ERR_INVALID, "nrefs must be 1 or 2");
    if (nrefs == 2 && refs[0] == refs[1]) return set_error(DME_ERR_INVALID, "the two references must differ");
    Picture *R[2] = {nullptr, nullptr};
    for (int r = 0; r < nrefs; ++r) {
        R[r] = find_pic(ctx, refs[r]);
        if (!R[r]) return set_error(DME_ERR_INVALID, "reference " + std::to_string(refs[r]) + " is not registered");
        if (R[r]->W != P->W || R[r]->H != P->H) return set_error(DME_ERR_INVALID, "reference size differs from picture");
        if (refs[r] == pic_num) return set_error(DME_ERR_INVALID, "a picture cannot reference itself");
    }
    DME_TRY(make_geometry(p, P->W, P->H, J.g));
    const Geometry &g = J.g;
    {
        const std::string lim = launch_limits(g, p.full_search != 0, learned, p.combined_me != 0);
        if (!lim.empty()) return set_error(DME_ERR_INVALID, "unsupported picture or block size: " + lim);
    }
    if (p.combined_me) DME_TRY(prepare_combined(ctx, P, R, nrefs, p, learned));
    int fxr[2] = {0, 0}, fyr[2] = {0, 0};
    if (p.full_search) {
        for (int r = 0; r < 2; ++r) {
            // m_tdiff and m_big_xr (pixel_match.cpp:79-80, 249-250); one reference: ref2 = ref1
            const int td = std::abs((r < nrefs ? refs[r] : refs[0]) - pic_num);
            fxr[r] = std::min(td, 3) * p.x_range_me;
            fyr[r] = std::min(td, 3) * p.y_range_me;
            if (full_lds_bytes(g.blen[2][0], g.blen[2][1], fxr[r], fyr[r]) > kFullLdsMax)
                return set_error(DME_ERR_INVALID, "full-search window too large for one workgroup's LDS "
                                                  "(reduce x_range_me / y_range_me)");
        }
    }
    J.pic_num = pic_num;
    J.nrefs = nrefs;
    J.refs[0] = refs[0];
    J.refs[1] = nrefs > 1 ? refs[1] : refs[0];
    J.p = p;
    J.stage = 0;
    J.fast = learned && ctx->lp_form == DME_LP_FORM_FAST;
    const size_t B = (size_t)g.xnb * g.ynb, S = (size_t)g.xsb * g.ysb;
    for (int r = 0; r < 2; ++r) {
        DME_TRY(J.vec[r].ensure(sizeof(dme_mv) * B));
        DME_TRY(J.pc[r].ensure(sizeof(dme_cost) * B));
        for (int l = 1; l <= g.depth; ++l) {
            const size_t n = (size_t)g.lxb[l] * g.lyb[l];
            DME_TRY(J.lvec[r][l].ensure(sizeof(dme_mv) * n));
            DME_TRY(J.lpc[r][l].ensure(sizeof(dme_cost) * n));
        }
        // pel records: 28 ints per block of a level; sub-pel records: 56
        size_t maxt = 0;
        for (int l = 0; l <= g.depth; ++l) {
            maxt = std::max(maxt, (size_t)g.lxb[l] * g.lyb[l] * 28);
        }
        DME_TRY(J.ptab[r].ensure(sizeof(int) * maxt));
        DME_TRY(J.qtab[r].ensure(sizeof(int) * 56 * B));
    }
    DME_TRY(J.m1tab.ensure(sizeof(int) * 24 * (B / 4)));
    DME_TRY(J.m0tab.ensure(sizeof(int) * 32 * (B / 16)));
    DME_TRY(J.mode.ensure(sizeof(int32_t) * B));
    DME_TRY(J.icount.ensure(sizeof(int)));
    DME_TRY(J.intra.ensure(sizeof(float) * B));
    DME_TRY(J.bi.ensure(sizeof(dme_cost) * B));
    for (int c = 0; c < 3; ++c) {
        DME_TRY(J.dc[c].ensure(sizeof(int16_t) * B));
        DME_TRY(J.dct[c].ensure(sizeof(int16_t) * B));
    }
    DME_TRY(J.split.ensure(sizeof(int32_t) * S));
    DME_TRY(J.sbc.ensure(sizeof(float) * S));
    DME_TRY(J.lam.ensure(sizeof(float) * B));
    DME_TRY(J.lam1.ensure(sizeof(float) * B / 4));
    DME_TRY(J.lam0.ensure(sizeof(float) * B / 16));
    DME_TRY(J.ucost.ensure(sizeof(float) * B));
    DME_TRY(J.lwork.ensure(sizeof(int) * 2 * B));
    if (learned) {
        for (int r = 0; r < 2; ++r) {
            DME_TRY(J.fv0[r].ensure(sizeof(dme_mv) * B));
            DME_TRY(J.fv2[r].ensure(sizeof(dme_mv) * B));
            DME_TRY(J.fs2[r].ensure(sizeof(int) * B));
        }
        DME_TRY(J.fsb.ensure(sizeof(int) * kFsbInts * S));
    }

    JobDesc &d = J.desc;
    std::memset(&d, 0, sizeof(d));
    d.W = g.W; d.H = g.H; d.CW = g.CW; d.CH = g.CH;
    d.xnb = g.xnb; d.ynb = g.ynb; d.xsb = g.xsb; d.ysb = g.ysb; d.depth = g.depth; d.nrefs = nrefs;
    d.prec = p.mv_precision;
    d.full = p.full_search;
    for (int r = 0; r < 2; ++r) {
        d.fxr[r] = fxr[r];
        d.fyr[r] = fyr[r];
    }
    d.step = p.mv_precision == 0 ? 1 : p.mv_precision;
    d.fp_sem = p.fp_semantics;
    for (int l = 0; l <= g.depth; ++l) {
        d.lw[l] = g.lw[l]; d.lh[l] = g.lh[l]; d.lxb[l] = g.lxb[l]; d.lyb[l] = g.lyb[l];
    }
    for (int l = 0; l < 3; ++l)
        for (int a = 0; a < 2; ++a) {
            d.blen[l][a] = g.blen[l][a]; d.bsep[l][a] = g.bsep[l][a]; d.boff[l][a] = g.boff[l][a];
        }
    for (int a = 0; a < 2; ++a) { d.cblen[a] = g.cblen[a]; d.cbsep[a] = g.cbsep[a]; d.cboff[a] = g.cboff[a]; }
    for (int l = 0; l < 3; ++l) d.lf[l] = g.level_factor[l];
    d.pic[0] = P->y;
    d.py = P->y;
    d.comb = p.combined_me != 0;
    d.cu = P->u;
    d.cv = P->v;
    for (int r = 0; r < 2; ++r) {
        d.vec[r] = J.vec[r].as<dme_mv>();
        d.pc[r] = J.pc[r].as<dme_cost>();
        d.lvec[r][0] = d.vec[r];
        d.lpc[r][0] = d.pc[r];
        for (int l = 1; l <= g.depth; ++l) {
            d.lvec[r][l] = J.lvec[r][l].as<dme_mv>();
            d.lpc[r][l] = J.lpc[r][l].as<dme_cost>();
        }
        d.ptab[r] = J.ptab[r].as<int>();
        d.qtab[r] = J.qtab[r].as<int>();
    }
    d.m1tab = J.m1tab.as<int>();
    d.m0tab = J.m0tab.as<int>();
    for (int r = 0; r < nrefs; ++r) d.ref[r][0] = R[r]->y;
    d.mode = J.mode.as<int32_t>();
    d.intra = J.intra.as<float>();
    d.bi = J.bi.as<dme_cost>();
    for (int c = 0; c < 3; ++c) {
        d.dc[c] = J.dc[c].as<int16_t>();
        d.dct[c] = J.dct[c].as<int16_t>();
    }
    d.split = J.split.as<int32_t>();
    d.sbc = J.sbc.as<float>();
    d.lam = J.lam.as<float>();
    d.icount = J.icount.as<int>();
    d.lam1 = J.lam1.as<float>();
    d.lam0 = J.lam0.as<float>();
    d.ucost = J.ucost.as<float>();
    d.lwork = J.lwork.as<int>();
    for (int r = 0; r < 2; ++r) {
        d.fv0[r] = J.fv0[r].as<dme_mv>();
        d.fv2[r] = J.fv2[r].as<dme_mv>();
        d.fs2[r] = J.fs2[r].as<int>();
    }
    d.fsb = J.fsb.as<int>();
    d.lambda = p.lambda;
    return DME_OK;
}

// The pictures a job's search reads: the picture and its references, or
// for combined-component ME their combined pictures.
static Picture *me_pic(Picture *P, const JobState *j) { return j->p.combined_me ? P->cpic.get() : P; }

// The pictures a set of jobs reads: pyramid depth needed per picture, and the
// references (whose quarter-pel planes the sub-pel and mode stages read).
static int collect_planes(dme_ctx *ctx, const std::vector<JobState *> &js, bool need_up, PlaneNeeds &nd) {
    for (auto *j : js) {
        Picture *P0 = find_pic(ctx, j->pic_num);
        if (!P0) return set_error(DME_ERR_STATE, "picture was dropped");
        const bool comb = j->p.combined_me != 0;
        if (comb && !P0->cpic) return set_error(DME_ERR_STATE, "combined picture missing");
        if (comb) add_once(nd.combs, P0);
        Picture *P = me_pic(P0, j);
        nd.levels[P] = std::max(nd.levels[P], j->g.depth);
        for (int r = 0; r < j->nrefs; ++r) {
            Picture *R0 = find_pic(ctx, j->refs[r]);
            if (!R0) return set_error(DME_ERR_STATE, "reference was dropped");
            if (comb && !R0->cpic) return set_error(DME_ERR_STATE, "combined picture missing");
            if (comb) add_once(nd.combs, R0);
            Picture *R = me_pic(R0, j);
            nd.levels[R] = std::max(nd.levels[R], j->g.depth);
            if (need_up) add_once(comb ? nd.ups16 : (j->fast ? nd.ups8 : nd.ups), R);
        }
    }
    return DME_OK;
}

// Builds, on the current lane, the planes of the pictures that jobs of two
// or more parts read (all they need of them), so that no lane builds a plane
// another lane reads.
static int build_shared_planes(dme_ctx *ctx, const std::vector<std::vector<JobState *>> &parts, int &launches) {
    std::vector<PlaneNeeds> nd(parts.size());
    std::map<Picture *, int> seen, shared;
    std::map<Picture *, int> comb_parts;  // pictures whose combined picture a part needs: parts counted
    for (size_t h = 0; h < parts.size(); ++h) {
        DME_TRY(collect_planes(ctx, parts[h], true, nd[h]));
        for (auto &kv : nd[h].levels) {
            auto it = seen.find(kv.first);
            if (it == seen.end()) {
                seen[kv.first] = kv.second;
            } else {
                it->second = std::max(it->second, kv.second);
                shared[kv.first] = it->second;
            }
        }
        for (Picture *P : nd[h].combs) ++comb_parts[P];
    }
    PlaneNeeds sn;
    for (auto &kv : comb_parts)
        if (kv.second > 1 || shared.count(kv.first->cpic.get())) sn.combs.push_back(kv.first);
    if (shared.empty() && sn.combs.empty()) return DME_OK;
    for (auto &kv : shared) sn.levels[kv.first] = seen[kv.first];
    for (const auto &n : nd) {
        for (Picture *P : n.ups)
            if (shared.count(P)) add_once(sn.ups, P);
        for (Picture *P : n.ups8)
            if (shared.count(P)) add_once(sn.ups8, P);
        for (Picture *P : n.ups16)
            if (shared.count(P)) add_once(sn.ups16, P);
    }
    return build_planes(ctx, sn, launches);
}

// Builds what a batch's pictures need, then fills the descriptors' plane
// pointers.
static int bind_planes(dme_ctx *ctx, std::vector<JobState *> &js, bool need_up, int &launches) {
    PlaneNeeds nd;
    DME_TRY(collect_planes(ctx, js, need_up, nd));
    DME_TRY(build_planes(ctx, nd, launches));
    for (auto *j : js) {
        Picture *P0 = find_pic(ctx, j->pic_num), *P = me_pic(P0, j);
        const bool comb = j->p.combined_me != 0;
        j->desc.pic[0] = P->y;
        for (int l = 1; l <= j->g.depth; ++l) j->desc.pic[l] = P->pyr[l].as<int16_t>();
        j->desc.py = P0->y;
        j->desc.pic8 = comb ? nullptr : P->b8_origin();
        j->desc.p8 = comb ? 0 : P->bp;
        for (int r = 0; r < j->nrefs; ++r) {
            Picture *R = me_pic(find_pic(ctx, j->refs[r]), j);
            j->desc.ref[r][0] = R->y;
            for (int l = 1; l <= j->g.depth; ++l) j->desc.ref[r][l] = R->pyr[l].as<int16_t>();
            j->desc.ref8[r] = comb ? nullptr : R->b8_origin();
            if (need_up && comb) j->desc.qp[r] = R->up.as<uint32_t>();
            if (need_up && !comb && !j->fast) j->desc.qp[r] = R->qp.as<uint32_t>();
            if (need_up && j->fast) {
                j->desc.up8[r] = R->u8_origin();
                j->desc.u8p = up8_pitch(R->W);
            }
        }
    }
    return DME_OK;
}

// Uploads the descriptors of a batch.
static int upload_jobs(dme_ctx *ctx, const std::vector<JobState *> &js) {
    std::vector<JobDesc> h(js.size());
    for (size_t i = 0; i < js.size(); ++i) h[i] = js[i]->desc;
    return upload(ctx, ctx->L->d_jobs, h.data(), sizeof(JobDesc) * h.size());
}

// Task table builder: groups of tasks uploaded in one copy per stage.
struct TaskTable {
    std::vector<Task> all;
    struct Range { size_t first; int n; int total; };
    Range add(const std::vector<Task> &t) {
        Range r;
        r.first = all.size();
        r.n = (int)t.size();
        r.total = t.empty() ? 0 : t.back().offset + t.back().count;
        all.insert(all.end(), t.begin(), t.end());
        return r;
    }
};

static int upload_tasks(dme_ctx *ctx, const TaskTable &tt) {
    return upload(ctx, ctx->L->d_tasks, tt.all.data(), sizeof(Task) * tt.all.size());
}

static const Task *dtask(dme_ctx *ctx, const TaskTable::Range &r) { return ctx->L->d_tasks.as<Task>() + r.first; }

// ---- stages ---------------------------------------------------------------

// Plain pel search of levels depth .. stop (stop 0: the whole search).
static int stage_pel(dme_ctx *ctx, std::vector<JobState *> &js, int &launches, int stop = 0) {
    int maxd = 0;
    // the up-converted planes are built with the pyramids: one pass over the
    // pictures, and the sub-pel stage then finds them ready
    DME_TRY(bind_planes(ctx, js, true, launches));
    for (auto *j : js) maxd = std::max(maxd, j->g.depth);
    DME_TRY(upload_jobs(ctx, js));
    TaskTable tt;
    // per level: the jobs searched from 16-bit samples, and at level 0 the
    // luma jobs searched from the byte planes
    std::vector<TaskTable::Range> bulk(maxd + 1);
    TaskTable::Range bulk0_u8 = tt.add(std::vector<Task>());
    for (int l = maxd; l >= stop; --l) {
        std::vector<Task> t, t8;
        int off = 0, off8 = 0;
        for (size_t i = 0; i < js.size(); ++i)
            for (int r = 0; r < js[i]->nrefs; ++r) {
                if (l > js[i]->g.depth || js[i]->p.full_search) continue;
                const int cnt = js[i]->g.lxb[l] * js[i]->g.lyb[l];
                if (l == 0 && !js[i]->p.combined_me) {
                    t8.push_back(Task{(int)i, r, cnt, off8});
                    off8 += cnt;
                } else {
                    t.push_back(Task{(int)i, r, cnt, off});
                    off += cnt;
                }
            }
        bulk[l] = tt.add(t);
        if (l == 0) bulk0_u8 = tt.add(t8);
    }
    // full-search jobs: their level-0 tables in k_pel_full
    std::vector<Task> ft;
    int foff = 0, fpp = 4, frp = 1, frows = 1;
    if (stop == 0)
        for (size_t i = 0; i < js.size(); ++i)
            for (int r = 0; r < js[i]->nrefs; ++r) {
                if (!js[i]->p.full_search) continue;
                const Geometry &g = js[i]->g;
                const int cnt = g.lxb[0] * g.lyb[0];
                ft.push_back(Task{(int)i, r, cnt, foff});
                foff += cnt;
                const int xr = js[i]->desc.fxr[r], yr = js[i]->desc.fyr[r];
                fpp = std::max(fpp, ((g.blen[2][0] + 15) / 16) * 4);
                frp = std::max(frp, full_pitch(g.blen[2][0], xr));
                frows = std::max(frows, g.blen[2][1] + 2 * std::max(yr, 2));
            }
    const TaskTable::Range full = tt.add(ft);
    std::vector<Task> wt;
    for (size_t i = 0; i < js.size(); ++i)
        for (int r = 0; r < js[i]->nrefs; ++r) wt.push_back(Task{(int)i, r, 1, (int)wt.size()});
    const TaskTable::Range wave = tt.add(wt);
    DME_TRY(upload_tasks(ctx, tt));
    const JobDesc *dj = ctx->L->d_jobs.as<JobDesc>();
    for (int l = maxd; l >= stop; --l) {
        int rows = 1, bx = 1, by = 1;
        bool guide = false;
        for (auto *j : js)
            if (l <= j->g.depth) {
                rows = std::max(rows, j->g.lyb[l]);
                bx = std::max(bx, j->g.blen[2][0]);
                by = std::max(by, j->g.blen[2][1]);
                guide = guide || l < j->g.depth;
            }
        DME_TRY(timed_launch(ctx, KK_PEL_BULK, launches, [&] {
            if (l == 0) {
                const hipError_t e = launch_pel_bulk(dj, dtask(ctx, bulk0_u8), bulk0_u8.n, bulk0_u8.total, 0, bx, by,
                                                     guide, true, ctx->L->stream);
                if (e != hipSuccess) return e;
            }
            return launch_pel_bulk(dj, dtask(ctx, bulk[l]), bulk[l].n, bulk[l].total, l, bx, by, guide, false,
                                   ctx->L->stream);
        }));
        if (l == 0 && full.total > 0)
            DME_TRY(timed_launch(ctx, KK_PEL_FULL, launches, [&] {
                return launch_pel_full(dj, dtask(ctx, full), full.n, full.total, fpp, frp, frows, ctx->L->stream);
            }));
        DME_TRY(timed_launch(ctx, KK_PEL_WAVE, launches,
                             [&] {
                                 return launch_pel_wave(dj, dtask(ctx, wave), wave.n, l, rows, bx, by,
                                                        ctx->act.pel_wave_threads, ctx->L->stream);
                             }));
    }
    for (auto *j : js) j->stage = 1;
    return DME_OK;
}

// Lambda maps: host computation from the level-0 pel vectors (or a given
// level-2 map), then upload of the level-2 map and the level-1/0 maps of
// ModeDecider.
static int stage_lambda_host(dme_ctx *ctx, std::vector<JobState *> &js, const float *const *overrides) {
    // staging layout per job: vectors of each reference, then the level-2,
    // level-1 and level-0 lambda maps
    std::vector<size_t> off(js.size() + 1, 0);
    for (size_t i = 0; i < js.size(); ++i) {
        const size_t B = (size_t)js[i]->g.xnb * js[i]->g.ynb;
        off[i + 1] = off[i] + sizeof(dme_mv) * B * js[i]->nrefs + sizeof(float) * (B + B / 4 + B / 16);
    }
    void *stage = nullptr;
    DME_TRY(arena_alloc(ctx, off.back(), &stage));
    char *base = static_cast<char *>(stage);
    for (size_t i = 0; i < js.size(); ++i) {
        const size_t B = (size_t)js[i]->g.xnb * js[i]->g.ynb;
        for (int r = 0; r < js[i]->nrefs; ++r)
            HIP_TRY(hipMemcpyAsync(base + off[i] + sizeof(dme_mv) * B * r, js[i]->vec[r].p, sizeof(dme_mv) * B,
                                   hipMemcpyDeviceToHost, ctx->L->stream));
    }
    HIP_TRY(hipStreamSynchronize(ctx->L->stream));
    // MEData::SetLambdaMap per picture, pictures spread over host threads
    auto work = [&](size_t i) {
        JobState &j = *js[i];
        const Geometry &g = j.g;
        const size_t B = (size_t)g.xnb * g.ynb;
        const dme_mv *v = reinterpret_cast<const dme_mv *>(base + off[i]);
        float *lam = reinterpret_cast<float *>(base + off[i] + sizeof(dme_mv) * B * j.nrefs);
        if (overrides && overrides[i])
            std::memcpy(lam, overrides[i], sizeof(float) * B);
        else
            lambda_map_l2(g.xnb, g.ynb, j.nrefs, v, j.nrefs > 1 ? v + B : nullptr, j.p.lambda, lam);
        lambda_map_level(lam, g.xnb, 1, (float)(1.0 / g.level_factor[1]), g.xnb / 2, g.ynb / 2, lam + B);
        lambda_map_level(lam, g.xnb, 0, (float)(1.0 / g.level_factor[0]), g.xnb / 4, g.ynb / 4, lam + B + B / 4);
    };
    const size_t nt = std::min<size_t>(js.size(), std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    if (nt <= 1) {
        for (size_t i = 0; i < js.size(); ++i) work(i);
    } else {
        std::vector<std::thread> th;
        for (size_t t = 0; t < nt; ++t)
            th.emplace_back([&, t] {
                for (size_t i = t; i < js.size(); i += nt) work(i);
            });
        for (auto &x : th) x.join();
    }
    for (size_t i = 0; i < js.size(); ++i) {
        JobState &j = *js[i];
        const size_t B = (size_t)j.g.xnb * j.g.ynb;
        const char *lam = base + off[i] + sizeof(dme_mv) * B * j.nrefs;
        HIP_TRY(hipMemcpyAsync(j.lam.p, lam, sizeof(float) * B, hipMemcpyHostToDevice, ctx->L->stream));
        HIP_TRY(hipMemcpyAsync(j.lam1.p, lam + sizeof(float) * B, sizeof(float) * B / 4, hipMemcpyHostToDevice,
                               ctx->L->stream));
        HIP_TRY(hipMemcpyAsync(j.lam0.p, lam + sizeof(float) * (B + B / 4), sizeof(float) * B / 16,
                               hipMemcpyHostToDevice, ctx->L->stream));
        j.stage = 2;
    }
    return DME_OK;
}

// Jobs of js (device lambda maps) whose flag is set; waits for the stream.
static int read_lambda_flags(dme_ctx *ctx, const std::vector<JobState *> &js, std::vector<JobState *> &flagged) {
    int n = 0;
    for (auto *j : js) n = std::max(n, j->bpos + 1);
    std::vector<int> f(n, 0);
    HIP_TRY(hipMemcpyAsync(f.data(), ctx->lflags->p, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->L->stream));
    HIP_TRY(hipStreamSynchronize(ctx->L->stream));
    for (auto *j : js)
        if (f[j->bpos]) flagged.push_back(j);
    return DME_OK;
}

// Lambda maps of a batch: given maps through the host, the others by
// k_lambda.  A job that k_lambda flags (a threshold too close to an integer
// to decide in double) needs the host's maps: with check_now the flags are
// read here and those jobs redone on the host; otherwise the caller reads
// them after the batch (read_lambda_flags) and redoes the flagged jobs.
static int stage_lambda(dme_ctx *ctx, std::vector<JobState *> &js, const float *const *overrides, bool check_now,
                        int &launches) {
    std::vector<JobState *> host, dev;
    std::vector<const float *> host_over;
    for (size_t i = 0; i < js.size(); ++i) {
        if (overrides && overrides[i]) {
            host.push_back(js[i]);
            host_over.push_back(overrides[i]);
        }
    }
    if (!host.empty()) DME_TRY(stage_lambda_host(ctx, host, host_over.data()));
    TaskTable tt;
    std::vector<Task> t;
    for (size_t i = 0; i < js.size(); ++i)
        if (!(overrides && overrides[i])) {
            t.push_back(Task{(int)i, 0, 1, js[i]->bpos});
            dev.push_back(js[i]);
        }
    if (dev.empty()) return DME_OK;
    const TaskTable::Range rt = tt.add(t);
    // this call's descriptors: the lane may hold another call's
    DME_TRY(upload_jobs(ctx, js));
    DME_TRY(upload_tasks(ctx, tt));
    const JobDesc *dj = ctx->L->d_jobs.as<JobDesc>();
    DME_TRY(timed_launch(ctx, KK_LAMBDA, launches, [&] {
        return launch_lambda(dj, dtask(ctx, rt), rt.n, ctx->lflags->as<int>(), ctx->L->stream);
    }));
    for (auto *j : dev) j->stage = 2;
    if (!check_now) return DME_OK;
    std::vector<JobState *> redo;
    DME_TRY(read_lambda_flags(ctx, dev, redo));
    ctx->stats.host_lambda_jobs = (int32_t)redo.size();
    if (!redo.empty()) DME_TRY(stage_lambda_host(ctx, redo, nullptr));
    return DME_OK;
}

// Which kinds of job (DME_KIND_*) a set of jobs holds.
static int job_kinds(const std::vector<JobState *> &js) {
    int k = 0;
    for (auto *j : js) {
        const bool e = j->p.mv_precision == 3;
        if (j->p.combined_me) k |= e ? DME_KIND_COMB_EIGHTH : DME_KIND_COMB_BELOW_EIGHTH;
        else k |= e ? DME_KIND_EIGHTH : DME_KIND_BELOW_EIGHTH;
    }
    return k;
}

static int stage_subpel(dme_ctx *ctx, std::vector<JobState *> &js, int &launches) {
    DME_TRY(bind_planes(ctx, js, true, launches));
    DME_TRY(upload_jobs(ctx, js));
    TaskTable tt;
    std::vector<Task> wt, sh, bt, b16;
    int off = 0, offb = 0, off16 = 0, rows = 1, bx = 1, by = 1;
    for (size_t i = 0; i < js.size(); ++i)
        for (int r = 0; r < js[i]->nrefs; ++r) {
            const int cnt = js[i]->g.xnb * js[i]->g.ynb;
            if (js[i]->p.mv_precision == 0) {
                sh.push_back(Task{(int)i, r, cnt, off});
                off += cnt;
            } else {
                wt.push_back(Task{(int)i, r, 1, (int)wt.size()});
                if (js[i]->p.combined_me) {
                    b16.push_back(Task{(int)i, r, cnt, off16});
                    off16 += cnt;
                } else {
                    bt.push_back(Task{(int)i, r, cnt, offb});
                    offb += cnt;
                }
                rows = std::max(rows, js[i]->g.ynb);
                bx = std::max(bx, js[i]->g.blen[2][0]);
                by = std::max(by, js[i]->g.blen[2][1]);
            }
        }
    const TaskTable::Range wave = tt.add(wt), shift = tt.add(sh), bulk = tt.add(bt), bulk16 = tt.add(b16);
    DME_TRY(upload_tasks(ctx, tt));
    const JobDesc *dj = ctx->L->d_jobs.as<JobDesc>();
    const int kinds = job_kinds(js);
    if (bulk.n)
        DME_TRY(timed_launch(ctx, KK_SUBPEL_BULK, launches, [&] {
            return launch_subpel_bulk(dj, dtask(ctx, bulk), bulk.n, bulk.total, bx, by, ctx->L->stream);
        }));
    if (bulk16.n)
        DME_TRY(timed_launch(ctx, KK_SUBPEL_T16, launches, [&] {
            return launch_subpel_table16(dj, dtask(ctx, bulk16), bulk16.n, bulk16.total, bx, by, ctx->L->stream);
        }));
    if (wave.n) {
        DME_TRY(timed_launch(ctx, KK_SUBPEL_WAVE, launches,
                             [&] {
                                 return launch_subpel_wave(dj, dtask(ctx, wave), wave.n, rows,
                                                           ctx->act.subpel_wave_threads, kinds, ctx->L->stream);
                             }));
    }
    if (shift.n) {
        DME_TRY(timed_launch(ctx, KK_SHIFT, launches, [&] {
            return launch_shift_vectors(dj, dtask(ctx, shift), shift.n, shift.total, 1, ctx->L->stream);
        }));
    }
    for (auto *j : js) j->stage = 3;
    return DME_OK;
}

static int stage_mode(dme_ctx *ctx, std::vector<JobState *> &js, int &launches) {
    // k_mode_l2 writes every block's zero-initialised MEData entries
    // (motion.cpp:55-126) before the later kernels of the stage
    DME_TRY(bind_planes(ctx, js, true, launches));
    DME_TRY(upload_jobs(ctx, js));
    TaskTable tt;
    std::vector<Task> l2, dc, dr, sh, mt, mt16, mt16e;
    int off = 0, offd = 0, offr = 0, offs = 0, offm = 0, offm16 = 0, offm16e = 0, max_w = 0;
    for (size_t i = 0; i < js.size(); ++i) {
        const int cnt = js[i]->g.xnb * js[i]->g.ynb;
        l2.push_back(Task{(int)i, 0, cnt, off});
        off += cnt;
        // unit tables: luma jobs in k_mode_tables, combined-component jobs in
        // k_mode_tables16 (eighth-pel ones apart)
        std::vector<Task> &m = js[i]->p.combined_me ? (js[i]->p.mv_precision == 3 ? mt16e : mt16) : mt;
        int &om = js[i]->p.combined_me ? (js[i]->p.mv_precision == 3 ? offm16e : offm16) : offm;
        m.push_back(Task{(int)i, 1, cnt / 4, om});
        om += cnt / 4;
        m.push_back(Task{(int)i, 0, cnt / 16, om});
        om += cnt / 16;
        for (int c = 0; c < 3; ++c) {
            dc.push_back(Task{(int)i, c, cnt, offd});
            offd += cnt;
            dr.push_back(Task{(int)i, c, js[i]->g.ynb, offr});  // block rows, for k_block_dc
            offr += js[i]->g.ynb;
        }
        max_w = std::max(max_w, js[i]->g.W);
        if (js[i]->p.mv_precision == 0)
            for (int r = 0; r < js[i]->nrefs; ++r) {
                sh.push_back(Task{(int)i, r, cnt, offs});
                offs += cnt;
            }
    }
    const TaskTable::Range rl2 = tt.add(l2), rdc = tt.add(dc), rdr = tt.add(dr), rsh = tt.add(sh), rmt = tt.add(mt);
    const TaskTable::Range rmt16 = tt.add(mt16), rmt16e = tt.add(mt16e);
    DME_TRY(upload_tasks(ctx, tt));
    const JobDesc *dj = ctx->L->d_jobs.as<JobDesc>();
    hipStream_t st = ctx->L->stream;
    const int kinds = job_kinds(js);
    DME_TRY(timed_launch(ctx, KK_MODE_L2, launches,
                         [&] { return launch_mode_l2(dj, dtask(ctx, rl2), rl2.n, rl2.total, false, kinds, st); }));
    int mx0 = 1, my0 = 1, mx1 = 1, my1 = 1;
    for (auto *j : js) {
        mx0 = std::max(mx0, j->g.blen[0][0]);
        my0 = std::max(my0, j->g.blen[0][1]);
        mx1 = std::max(mx1, j->g.blen[1][0]);
        my1 = std::max(my1, j->g.blen[1][1]);
    }
    if (rmt.n)
        DME_TRY(timed_launch(ctx, KK_MODE_TABLES, launches, [&] {
            return launch_mode_tables(dj, dtask(ctx, rmt), rmt.n, rmt.total, mx0, my0, mx1, my1, kinds, st);
        }));
    if (rmt16.n || rmt16e.n)
        DME_TRY(timed_launch(ctx, KK_MODE_T16, launches, [&] {
            const hipError_t e = launch_mode_tables16(dj, dtask(ctx, rmt16), rmt16.n, rmt16.total, false, st);
            if (e != hipSuccess) return e;
            return launch_mode_tables16(dj, dtask(ctx, rmt16e), rmt16e.n, rmt16e.total, true, st);
        }));
    DME_TRY(timed_launch(ctx, KK_MODE_WAVE, launches,
                         [&] { return launch_mode_wave(dj, (int)js.size(), ctx->act.mode_wave_threads, kinds, st); }));
    DME_TRY(timed_launch(ctx, KK_BLOCK_DC, launches,
                         [&] { return launch_block_dc(dj, dtask(ctx, rdr), rdr.n, rdr.total, max_w, st); }));
    DME_TRY(timed_launch(ctx, KK_DC_SMOOTH, launches,
                         [&] { return launch_dc_smooth(dj, dtask(ctx, rdc), rdc.n, rdc.total, st); }));
    DME_TRY(timed_launch(ctx, KK_INTRA, launches, [&] { return launch_intra_count(dj, (int)js.size(), st); }));
    if (rsh.n) {
        DME_TRY(timed_launch(ctx, KK_SHIFT, launches, [&] {
            return launch_shift_vectors(dj, dtask(ctx, rsh), rsh.n, rsh.total, 0, st);
        }));
    }
    for (auto *j : js) j->stage = 4;
    return DME_OK;
}

// Copies a job's MEData to out's non-NULL arrays: host arrays (kind
// hipMemcpyDeviceToHost), or arrays in either memory (hipMemcpyDefault).
static int export_job(dme_ctx *ctx, JobState &J, dme_medata *out, hipMemcpyKind kind = hipMemcpyDeviceToHost) {
    if (!out) return DME_OK;
    const size_t B = (size_t)J.g.xnb * J.g.ynb, S = (size_t)J.g.xsb * J.g.ysb;
    hipStream_t st = ctx->L->stream;
    for (int r = 0; r < J.nrefs; ++r) {
        if (out->vectors[r]) HIP_TRY(hipMemcpyAsync(out->vectors[r], J.vec[r].p, sizeof(dme_mv) * B, kind, st));
        if (out->pred_costs[r]) HIP_TRY(hipMemcpyAsync(out->pred_costs[r], J.pc[r].p, sizeof(dme_cost) * B, kind, st));
    }
    if (J.stage >= 2 && out->lambda_map)
        HIP_TRY(hipMemcpyAsync(out->lambda_map, J.lam.p, sizeof(float) * B, kind, st));
    if (J.stage >= 4) {
        if (out->mode) HIP_TRY(hipMemcpyAsync(out->mode, J.mode.p, sizeof(int32_t) * B, kind, st));
        if (out->intra_costs) HIP_TRY(hipMemcpyAsync(out->intra_costs, J.intra.p, sizeof(float) * B, kind, st));
        if (out->bipred_costs) HIP_TRY(hipMemcpyAsync(out->bipred_costs, J.bi.p, sizeof(dme_cost) * B, kind, st));
        for (int c = 0; c < 3; ++c)
            if (out->dc[c]) HIP_TRY(hipMemcpyAsync(out->dc[c], J.dc[c].p, sizeof(int16_t) * B, kind, st));
        if (out->sb_split) HIP_TRY(hipMemcpyAsync(out->sb_split, J.split.p, sizeof(int32_t) * S, kind, st));
        if (out->sb_costs) HIP_TRY(hipMemcpyAsync(out->sb_costs, J.sbc.p, sizeof(float) * S, kind, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    if (J.stage >= 4 && out->intra_block_ratio) {
        int cnt = 0;
        HIP_TRY(hipMemcpy(&cnt, J.icount.p, sizeof(int), hipMemcpyDeviceToHost));
        const float ratio = static_cast<float>(static_cast<double>(cnt) / static_cast<double>(B));
        HIP_TRY(hipMemcpy(out->intra_block_ratio, &ratio, sizeof(float), hipMemcpyDefault));
    }
    return DME_OK;
}

// The MEData of a batch to host memory: k_pack gathers every requested
// array into one device buffer, one copy moves it to pinned host memory, and
// host threads scatter it into the caller's arrays (one DMA instead of a
// dozen pageable copies per picture).
static int export_batch(dme_ctx *ctx, const std::vector<JobState *> &js, dme_medata *outs) {
    std::vector<PackSeg> segs;
    std::vector<void *> dsts;
    unsigned long long off = 0;
    auto add = [&](void *dst, const DevBuf &src, size_t n) {
        if (!dst || !n) return;
        segs.push_back(PackSeg{src.p, off, (unsigned long long)n});
        dsts.push_back(dst);
        off += (n + 15) & ~(size_t)15;
    };
    for (size_t i = 0; i < js.size(); ++i) {
        JobState &J = *js[i];
        dme_medata *out = &outs[i];
        const size_t B = (size_t)J.g.xnb * J.g.ynb, S = (size_t)J.g.xsb * J.g.ysb;
        for (int r = 0; r < J.nrefs; ++r) {
            add(out->vectors[r], J.vec[r], sizeof(dme_mv) * B);
            add(out->pred_costs[r], J.pc[r], sizeof(dme_cost) * B);
        }
        if (J.stage >= 2) add(out->lambda_map, J.lam, sizeof(float) * B);
        if (J.stage >= 4) {
            add(out->mode, J.mode, sizeof(int32_t) * B);
            add(out->intra_costs, J.intra, sizeof(float) * B);
            add(out->bipred_costs, J.bi, sizeof(dme_cost) * B);
            for (int c = 0; c < 3; ++c) add(out->dc[c], J.dc[c], sizeof(int16_t) * B);
            add(out->sb_split, J.split, sizeof(int32_t) * S);
            add(out->sb_costs, J.sbc, sizeof(float) * S);
        }
    }
    // the INTRA counts, one int per picture, after the arrays
    std::vector<int> cnt(js.size(), 0);
    std::vector<unsigned long long> coff(js.size(), ~0ull);
    for (size_t i = 0; i < js.size(); ++i) {
        if (js[i]->stage < 4 || !outs[i].intra_block_ratio) continue;
        coff[i] = off;
        segs.push_back(PackSeg{js[i]->icount.p, off, (unsigned long long)sizeof(int)});
        dsts.push_back(&cnt[i]);
        off += 16;
    }
    if (segs.empty()) return DME_OK;
    hipStream_t st = ctx->L->stream;
    HIP_TRY(hipStreamSynchronize(st));
    DME_TRY(ctx->x_dev.ensure(off));
    DME_TRY(ctx->x_host.ensure(off));
    DME_TRY(upload(ctx, ctx->x_segs, segs.data(), sizeof(PackSeg) * segs.size()));
    HIP_TRY(launch_pack(ctx->x_segs.as<PackSeg>(), (int)segs.size(), ctx->x_dev.as<uint8_t>(), st));
    HIP_TRY(hipMemcpyAsync(ctx->x_host.p, ctx->x_dev.p, off, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    const uint8_t *h = static_cast<const uint8_t *>(ctx->x_host.p);
    auto scatter = [&](size_t a, size_t b) {
        for (size_t k = a; k < b; ++k) std::memcpy(dsts[k], h + segs[k].off, segs[k].bytes);
    };
    const size_t nt = std::min<size_t>(8, std::max<size_t>(1, off >> 23));  // a thread per 8 MiB, at most 8
    if (nt == 1) {
        scatter(0, segs.size());
    } else {
        std::vector<std::thread> th;
        const size_t per = (segs.size() + nt - 1) / nt;
        for (size_t t = 0; t < nt; ++t) {
            const size_t a = t * per, b = std::min(segs.size(), a + per);
            if (a < b) th.emplace_back(scatter, a, b);
        }
        for (auto &x : th) x.join();
    }
    for (size_t i = 0; i < js.size(); ++i)
        if (coff[i] != ~0ull)
            *outs[i].intra_block_ratio = static_cast<float>(static_cast<double>(cnt[i]) /
                                                            static_cast<double>((size_t)js[i]->g.xnb * js[i]->g.ynb));
    return DME_OK;
}

static JobState *job_for(dme_ctx *ctx, int pic_num) {
    auto &slot = ctx->jobs[pic_num];
    if (!slot) slot.reset(new JobState());
    return slot.get();
}

// Start of a call that queues work: the streams are idle (every call ends
// synchronised), so the staging arenas can be reused from their start.
static int begin_timing(dme_ctx *ctx) {
    ctx->L = ctx->L0;
    for (auto &ln : ctx->lane) {
        HIP_TRY(hipStreamSynchronize(ln.stream));
        ln.arena_used = 0;
    }
    ctx->L0->arena_used = 0;
    ctx->ev_used.clear();
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->L->stream));
    return DME_OK;
}

static int end_timing(dme_ctx *ctx, int launches) {
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->L->stream));
    HIP_TRY(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.kernel_ms = ms;
    ctx->stats.launches = launches;
    for (int k = 0; k < DME_NUM_KERNELS; ++k) {
        ctx->stats.kernel_ms_by[k] = 0.0;
        ctx->stats.launches_by[k] = 0;
    }
    for (auto &u : ctx->ev_used) {
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, ctx->ev_pool[u.second], ctx->ev_pool[u.second + 1]));
        ctx->stats.kernel_ms_by[u.first] += t;
        ctx->stats.launches_by[u.first] += 1;
    }
    ctx->ev_used.clear();
    return DME_OK;
}

// Batch positions of the jobs and their device lambda flags (streams idle).
static int lambda_flag_slots(dme_ctx *ctx, std::vector<JobState *> &js) {
    for (size_t i = 0; i < js.size(); ++i) js[i]->bpos = (int)i;
    return ctx->lflags->ensure(sizeof(int) * js.size());
}

// After a batch: the jobs whose device lambda maps were flagged run again
// with the host's maps (the pel search repeats exactly).
static int redo_flagged(dme_ctx *ctx, std::vector<JobState *> &js, const float *const *lam_over) {
    std::vector<JobState *> dev, redo;
    for (size_t i = 0; i < js.size(); ++i)
        if (!(lam_over && lam_over[i])) dev.push_back(js[i]);
    ctx->stats.host_lambda_jobs = 0;
    if (dev.empty()) return DME_OK;
    DME_TRY(read_lambda_flags(ctx, dev, redo));
    ctx->stats.host_lambda_jobs = (int32_t)redo.size();
    if (redo.empty()) return DME_OK;
    int launches = 0;
    DME_TRY(stage_pel(ctx, redo, launches));
    DME_TRY(stage_lambda_host(ctx, redo, nullptr));
    DME_TRY(stage_subpel(ctx, redo, launches));
    DME_TRY(stage_mode(ctx, redo, launches));
    HIP_TRY(hipStreamSynchronize(ctx->L->stream));
    return DME_OK;
}

// Signature of a batch for the tuning table: its jobs' sizes, reference
// counts and parameters as a sorted multiset ("count*key;..."); jobs that
// differ only in picture numbers or lambda share a key.
static std::string batch_signature(const std::vector<JobState *> &js) {
    std::map<std::string, int> keys;
    for (auto *j : js) {
        const dme_params &p = j->p;
        char b[192];
        std::snprintf(b, sizeof(b), "%dx%d,r%d,b%dx%d/%dx%d,p%d,d%d,f%d,s%d/%d/%d,c%d", j->g.W, j->g.H, j->nrefs,
                      p.xblen, p.yblen, p.xbsep, p.ybsep, p.mv_precision, j->g.depth, p.fp_semantics, p.full_search,
                      p.full_search ? p.x_range_me : 0, p.full_search ? p.y_range_me : 0, p.combined_me);
        ++keys[b];
    }
    std::string sig;
    for (auto &kv : keys) {
        if (!sig.empty()) sig += ';';
        sig += std::to_string(kv.second) + '*' + kv.first;
    }
    return sig;
}

static bool tuning_ok(const dme_tuning &t) {
    return t.parts >= 1 && t.parts <= kLanes && (t.stagger == 0 || t.stagger == 1) &&
           (t.pel_wave_threads == 256 || t.pel_wave_threads == 512 || t.pel_wave_threads == 1024) &&
           (t.subpel_wave_threads == 256 || t.subpel_wave_threads == 512) &&
           (t.mode_wave_threads == 256 || t.mode_wave_threads == 512 || t.mode_wave_threads == 1024);
}

// The setting a batch runs with: its table entry, else the default.
static void choose_tuning(dme_ctx *ctx, const std::vector<JobState *> &js) {
    if (ctx->act_fixed) return;
    ctx->act = ctx->tune;
    if (ctx->tune_table.empty()) return;
    auto it = ctx->tune_table.find(batch_signature(js));
    if (it != ctx->tune_table.end()) ctx->act = it->second;
}

// Queues the whole live sequence of a batch (streams idle, flag slots set)
// as up to act.parts parts of about equal work, one per lane:
//   lane 0: shared planes, planes(A) pel(A)  lambda(A) subpel(A) mode(A)
//   lane 1:                   planes(B) pel(B)  lambda(B) subpel(B) mode(B)
//   lane 2:                             planes(C) pel(C) ...
// (staggered starts; without stagger every part starts at the fork).
// Planes that two parts read are built once, before the fork, so no kernel
// of one lane reads a plane another lane writes; every other buffer a part
// uses is its own jobs' or its lane's.  Ends with lane 0 waiting for the
// others.  With per-kernel profiling on, the batch runs as one part on
// lane 0 so that each kernel is timed alone.  Only stream operations: a
// graph capture records it as it stands.
static int queue_parts(dme_ctx *ctx, std::vector<JobState *> &js, const float *const *lam_over, int &launches) {
    const int np = ctx->profiling ? 1 : (int)std::min<size_t>((size_t)ctx->act.parts, js.size());
    // parts: largest job first to the lightest part, order kept within each
    std::vector<size_t> order(js.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    auto work = [&](size_t i) { return (double)js[i]->g.W * js[i]->g.H * js[i]->nrefs; };
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return work(a) > work(b); });
    std::vector<int> side(js.size());
    std::vector<double> load(np, 0.0);
    for (size_t i : order) {
        int b = 0;
        for (int h = 1; h < np; ++h)
            if (load[h] < load[b]) b = h;
        side[i] = b;
        load[b] += work(i);
    }
    std::vector<std::vector<JobState *>> part(np);
    std::vector<std::vector<const float *>> over(np);
    for (size_t i = 0; i < js.size(); ++i) {
        part[side[i]].push_back(js[i]);
        over[side[i]].push_back(lam_over ? lam_over[i] : nullptr);
    }
    ctx->L = &ctx->lane[0];
    if (np > 1) {
        // Planes of pictures two parts use are built on lane 0 before the
        // fork; every other picture's planes are built by its own part's
        // lane, so one lane's plane building overlaps another's search.
        DME_TRY(build_shared_planes(ctx, part, launches));
        HIP_TRY(hipEventRecord(ctx->ev_fork, ctx->lane[0].stream));
        for (int h = 1; h < np; ++h) HIP_TRY(hipStreamWaitEvent(ctx->lane[h].stream, ctx->ev_fork, 0));
    }
    for (int h = 0; h < np; ++h) {
        ctx->L = &ctx->lane[h];
        if (np > 1 && ctx->act.stagger) {
            if (h > 0) HIP_TRY(hipStreamWaitEvent(ctx->L->stream, ctx->ev_planes[h - 1], 0));
            DME_TRY(bind_planes(ctx, part[h], true, launches));
            HIP_TRY(hipEventRecord(ctx->ev_planes[h], ctx->L->stream));
        }
        DME_TRY(stage_pel(ctx, part[h], launches));
    }
    for (int h = 0; h < np; ++h) {
        ctx->L = &ctx->lane[h];
        DME_TRY(stage_lambda(ctx, part[h], over[h].data(), false, launches));
        DME_TRY(stage_subpel(ctx, part[h], launches));
        DME_TRY(stage_mode(ctx, part[h], launches));
    }
    ctx->L = &ctx->lane[0];
    for (int h = 1; h < np; ++h) {
        HIP_TRY(hipEventRecord(ctx->ev_join[h], ctx->lane[h].stream));
        HIP_TRY(hipStreamWaitEvent(ctx->L->stream, ctx->ev_join[h], 0));
    }
    return DME_OK;
}

// The whole live sequence for a batch of jobs, then the host's redo of jobs
// whose device lambda maps were flagged.
static int run_jobs_parts(dme_ctx *ctx, std::vector<JobState *> &js, const float *const *lam_over) {
    int launches = 0;
    choose_tuning(ctx, js);
    DME_TRY(begin_timing(ctx));
    DME_TRY(lambda_flag_slots(ctx, js));
    DME_TRY(queue_parts(ctx, js, lam_over, launches));
    DME_TRY(end_timing(ctx, launches));
    return redo_flagged(ctx, js, lam_over);
}

// Waits for every lane: after a call that failed part-way, work queued on
// the other lanes may still read pictures and job buffers that a later call
// frees or reuses.
static void sync_lanes(dme_ctx *ctx) {
    for (auto &ln : ctx->lane) (void)hipStreamSynchronize(ln.stream);
    ctx->L = ctx->L0;
}

static int run_jobs(dme_ctx *ctx, std::vector<JobState *> &js, const float *const *lam_over) {
    const int rc = run_jobs_parts(ctx, js, lam_over);
    if (rc != DME_OK) sync_lanes(ctx);
    return rc;
}

}  // namespace dme

using namespace dme;

#define CTX_CHECK(ctx)                                                             \
    do {                                                                           \
        if (!(ctx)) return set_error(DME_ERR_INVALID, "NULL context");            \
        DME_TRY(set_device(ctx));                                                  \
    } while (0)

extern "C" int dme_create(int device, dme_ctx **out) try {
    if (!out) return set_error(DME_ERR_INVALID, "NULL out");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return set_error(DME_ERR_NO_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) return set_error(DME_ERR_INVALID, "device index out of range");
    std::unique_ptr<dme_ctx> c(new dme_ctx());
    c->device = device;
    HIP_TRY(hipSetDevice(device));
    for (auto &ln : c->lane) HIP_TRY(hipStreamCreateWithFlags(&ln.stream, hipStreamNonBlocking));
    c->stream = c->lane[0].stream;
    c->L = &c->lane[0];
    HIP_TRY(hipEventCreate(&c->ev0));
    HIP_TRY(hipEventCreate(&c->ev1));
    HIP_TRY(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    for (int h = 0; h < kLanes; ++h) {
        HIP_TRY(hipEventCreateWithFlags(&c->ev_join[h], hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&c->ev_planes[h], hipEventDisableTiming));
    }
    std::memset(&c->stats, 0, sizeof(c->stats));
    std::memset(&c->params, 0, sizeof(c->params));
    // the no-allocation entry point: this device's scratch array, and its
    // lane (lane 0's stream) with pinned staging of a fixed size
    DME_TRY(noalloc_scratch(device, &c->scratch));
    c->na_lane.stream = c->lane[0].stream;
    DME_TRY(c->na_lane.arena.ensure(kNoallocStagingBytes));
    c->na_lane.arena.fixed = true;
    *out = c.release();
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_destroy(dme_ctx *ctx) try {
    if (!ctx) return DME_OK;
    (void)hipSetDevice(ctx->device);
    for (auto &ln : ctx->lane) (void)hipStreamSynchronize(ln.stream);
    for (auto *g : ctx->graphs) delete g;
    ctx->graphs.clear();
    ctx->jobs.clear();
    ctx->pics.clear();
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    for (int h = 0; h < kLanes; ++h) {
        if (ctx->ev_join[h]) (void)hipEventDestroy(ctx->ev_join[h]);
        if (ctx->ev_planes[h]) (void)hipEventDestroy(ctx->ev_planes[h]);
    }
    for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
    for (auto &ln : ctx->lane)
        if (ln.stream) (void)hipStreamDestroy(ln.stream);
    delete ctx;
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_set_params(dme_ctx *ctx, const dme_params *params) try {
    if (!ctx || !params) return set_error(DME_ERR_INVALID, "NULL argument");
    // validate everything but the size-dependent depth on a large picture
    Geometry g;
    dme_params q = *params;
    q.depth = -1;
    DME_TRY(make_geometry(q, 3840, 2160, g));
    if (params->depth > DME_MAX_DEPTH) return set_error(DME_ERR_INVALID, "pyramid depth must be in [1, 8]");
    ctx->params = *params;
    ctx->params_set = true;
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

static int register_picture(dme_ctx *ctx, int pic_num, int width, int height, const int16_t *y, const int16_t *u,
                            const int16_t *v, bool device) {
    if (!y) return set_error(DME_ERR_INVALID, "luma plane is NULL");
    if (width < 24 || height < 24 || (width & 1) || (height & 1) || width > 16384 || height > 16384)
        return set_error(DME_ERR_INVALID, "picture size must be even and within [24, 16384]");
    if ((u == nullptr) != (v == nullptr)) return set_error(DME_ERR_INVALID, "give both chroma planes or neither");
    std::unique_ptr<Picture> P(new Picture());
    P->num = pic_num;
    P->W = width;
    P->H = height;
    const size_t n = (size_t)width * height, cn = (size_t)(width / 2) * (height / 2);
    if (device) {
        P->y = y;
        P->u = u;
        P->v = v;
    } else {
        DME_TRY(P->own[0].ensure(sizeof(int16_t) * n));
        HIP_TRY(hipMemcpy(P->own[0].p, y, sizeof(int16_t) * n, hipMemcpyHostToDevice));
        P->y = P->own[0].as<int16_t>();
        if (u) {
            DME_TRY(P->own[1].ensure(sizeof(int16_t) * cn));
            DME_TRY(P->own[2].ensure(sizeof(int16_t) * cn));
            HIP_TRY(hipMemcpy(P->own[1].p, u, sizeof(int16_t) * cn, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(P->own[2].p, v, sizeof(int16_t) * cn, hipMemcpyHostToDevice));
            P->u = P->own[1].as<int16_t>();
            P->v = P->own[2].as<int16_t>();
        }
    }
    // luma must be 8-bit samples minus 128
    DME_TRY(ctx->d_flag.ensure(sizeof(int)));
    int bad = 0;
    HIP_TRY(hipMemcpyAsync(ctx->d_flag.p, &bad, sizeof(int), hipMemcpyHostToDevice, ctx->L->stream));
    HIP_TRY(launch_check_range(P->y, n, ctx->d_flag.as<int>(), ctx->L->stream));
    HIP_TRY(hipMemcpyAsync(&bad, ctx->d_flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->L->stream));
    HIP_TRY(hipStreamSynchronize(ctx->L->stream));
    if (bad) return set_error(DME_ERR_INVALID, "luma samples must lie in [-128, 127] (8-bit samples minus 128)");
    P->serial = ctx->next_serial++;
    ctx->pics[pic_num] = std::move(P);
    return DME_OK;
}

extern "C" int dme_upload_picture(dme_ctx *ctx, int pic_num, int width, int height, const int16_t *y,
                                  const int16_t *u, const int16_t *v) try {
    CTX_CHECK(ctx);
    return register_picture(ctx, pic_num, width, height, y, u, v, false);
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_bind_picture_device(dme_ctx *ctx, int pic_num, int width, int height, const int16_t *d_y,
                                       const int16_t *d_u, const int16_t *d_v) try {
    CTX_CHECK(ctx);
    return register_picture(ctx, pic_num, width, height, d_y, d_u, d_v, true);
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_drop_picture(dme_ctx *ctx, int pic_num) try {
    CTX_CHECK(ctx);
    HIP_TRY(hipStreamSynchronize(ctx->L->stream));
    ctx->pics.erase(pic_num);
    ctx->jobs.erase(pic_num);
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_pixel_search(dme_ctx *ctx, int pic_num, const int *refs, int nrefs, dme_medata *out) try {
    CTX_CHECK(ctx);
    if (!ctx->params_set) return set_error(DME_ERR_STATE, "dme_set_params was not called");
    JobState *J = job_for(ctx, pic_num);
    DME_TRY(prepare_job(ctx, *J, ctx->params, pic_num, refs, nrefs));
    std::vector<JobState *> js{J};
    int launches = 0;
    DME_TRY(begin_timing(ctx));
    DME_TRY(stage_pel(ctx, js, launches));
    DME_TRY(end_timing(ctx, launches));
    return export_job(ctx, *J, out);
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_set_lambda_map(dme_ctx *ctx, int pic_num, const float *lambda_map, dme_medata *out) try {
    CTX_CHECK(ctx);
    auto it = ctx->jobs.find(pic_num);
    if (it == ctx->jobs.end() || it->second->stage < 1) return set_error(DME_ERR_STATE, "run dme_pixel_search first");
    std::vector<JobState *> js{it->second.get()};
    const float *ov[1] = {lambda_map};
    HIP_TRY(hipStreamSynchronize(ctx->L->stream));
    ctx->L->arena_used = 0;
    DME_TRY(lambda_flag_slots(ctx, js));
    int launches = 0;
    DME_TRY(stage_lambda(ctx, js, ov, true, launches));
    return export_job(ctx, *js[0], out);
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_load_medata(dme_ctx *ctx, int pic_num, const int *refs, int nrefs, const dme_medata *in,
                               int stage) try {
    CTX_CHECK(ctx);
    if (!ctx->params_set) return set_error(DME_ERR_STATE, "dme_set_params was not called");
    if (!in || stage < 1 || stage > 3) return set_error(DME_ERR_INVALID, "NULL MEData or stage outside 1..3");
    for (int r = 0; r < nrefs && r < 2; ++r)
        if (!in->vectors[r] || !in->pred_costs[r])
            return set_error(DME_ERR_INVALID, "the vectors and PredCosts of every reference are needed");
    if (stage >= 2 && !in->lambda_map) return set_error(DME_ERR_INVALID, "stages 2 and 3 need the lambda map");
    JobState *J = job_for(ctx, pic_num);
    DME_TRY(prepare_job(ctx, *J, ctx->params, pic_num, refs, nrefs));
    const size_t B = (size_t)J->g.xnb * J->g.ynb;
    for (int r = 0; r < J->nrefs; ++r) {
        HIP_TRY(hipMemcpy(J->vec[r].p, in->vectors[r], sizeof(dme_mv) * B, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(J->pc[r].p, in->pred_costs[r], sizeof(dme_cost) * B, hipMemcpyHostToDevice));
    }
    J->stage = 1;
    if (stage >= 2) {
        std::vector<JobState *> js{J};
        const float *ov[1] = {in->lambda_map};
        ctx->L->arena_used = 0;
        DME_TRY(lambda_flag_slots(ctx, js));
        int launches = 0;
        DME_TRY(stage_lambda(ctx, js, ov, true, launches));
        HIP_TRY(hipStreamSynchronize(ctx->L->stream));
    }
    J->stage = stage;
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_subpel(dme_ctx *ctx, int pic_num, dme_medata *out) try {
    CTX_CHECK(ctx);
    auto it = ctx->jobs.find(pic_num);
    if (it == ctx->jobs.end() || it->second->stage < 2) return set_error(DME_ERR_STATE, "set the lambda map first");
    std::vector<JobState *> js{it->second.get()};
    int launches = 0;
    DME_TRY(begin_timing(ctx));
    DME_TRY(stage_subpel(ctx, js, launches));
    DME_TRY(end_timing(ctx, launches));
    return export_job(ctx, *js[0], out);
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_mode_decide(dme_ctx *ctx, int pic_num, dme_medata *out) try {
    CTX_CHECK(ctx);
    auto it = ctx->jobs.find(pic_num);
    if (it == ctx->jobs.end() || it->second->stage < 3) return set_error(DME_ERR_STATE, "run dme_subpel first");
    std::vector<JobState *> js{it->second.get()};
    int launches = 0;
    DME_TRY(begin_timing(ctx));
    DME_TRY(stage_mode(ctx, js, launches));
    DME_TRY(end_timing(ctx, launches));
    return export_job(ctx, *js[0], out);
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_motion_estimate(dme_ctx *ctx, int pic_num, const int *refs, int nrefs, dme_medata *out) try {
    CTX_CHECK(ctx);
    if (!ctx->params_set) return set_error(DME_ERR_STATE, "dme_set_params was not called");
    JobState *J = job_for(ctx, pic_num);
    DME_TRY(prepare_job(ctx, *J, ctx->params, pic_num, refs, nrefs));
    std::vector<JobState *> js{J};
    DME_TRY(run_jobs(ctx, js, nullptr));
    return export_job(ctx, *J, out);
} catch (...) {
    return dme::exception_status();
}

namespace dme {
static int prepare_batch(dme_ctx *ctx, const dme_job *jobs, int njobs, std::vector<JobState *> &js,
                         std::vector<std::unique_ptr<JobState>> *own);
}

extern "C" int dme_run_batch(dme_ctx *ctx, const dme_job *jobs, int njobs, dme_medata *outs) try {
    CTX_CHECK(ctx);
    std::vector<JobState *> js;
    DME_TRY(prepare_batch(ctx, jobs, njobs, js, nullptr));
    DME_TRY(run_jobs(ctx, js, nullptr));
    if (outs) DME_TRY(export_batch(ctx, js, outs));
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

// ---- launch tuning and captured batch steps -----------------------------------

namespace dme {

// The jobs of a batch entry point, validated and prepared in js (the
// context's job state when own is NULL, else new state kept in own).
static int prepare_batch(dme_ctx *ctx, const dme_job *jobs, int njobs, std::vector<JobState *> &js,
                         std::vector<std::unique_ptr<JobState>> *own) {
    if (!jobs || njobs <= 0) return set_error(DME_ERR_INVALID, "empty batch");
    for (int i = 0; i < njobs; ++i) {
        for (int k = 0; k < i; ++k)
            if (jobs[k].pic_num == jobs[i].pic_num) return set_error(DME_ERR_INVALID, "a picture appears twice in a batch");
        JobState *J = nullptr;
        if (own) {
            own->emplace_back(new JobState());
            J = own->back().get();
        } else {
            J = job_for(ctx, jobs[i].pic_num);
        }
        DME_TRY(prepare_job(ctx, *J, jobs[i].params, jobs[i].pic_num, jobs[i].refs, jobs[i].nrefs));
        js.push_back(J);
    }
    return DME_OK;
}

// Marks the planes of the pictures a batch reads as not built.
static void invalidate_batch_planes(dme_ctx *ctx, const std::vector<JobState *> &js) {
    for (auto *j : js) {
        int nums[3] = {j->pic_num, j->refs[0], j->refs[1]};
        for (int k = 0; k <= j->nrefs; ++k) {
            Picture *P = find_pic(ctx, nums[k]);
            if (!P) continue;
            P->pyr_levels = 0;
            P->qp_ready = P->b8_ready = P->u8_ready = P->up_ready = P->comb_ready = false;
            if (P->cpic) {
                P->cpic->pyr_levels = 0;
                P->cpic->up_ready = false;
            }
        }
    }
}

// Exchanges the context's lane buffers and lambda flags with a graph's.
static void swap_graph_buffers(dme_ctx *ctx, dme_batch_graph *g) {
    for (int h = 0; h < kLanes; ++h) {
        Lane &a = ctx->lane[h], &b = g->lanes[h];
        a.d_jobs.swap(b.d_jobs);
        a.d_tasks.swap(b.d_tasks);
        a.d_planes.swap(b.d_planes);
        a.d_planes16.swap(b.d_planes16);
        a.d_comb.swap(b.d_comb);
        a.arena.swap(b.arena);
        std::swap(a.arena_used, b.arena_used);
    }
    ctx->d_lflags.swap(g->lflags);
}

// With the graph's buffers in the context: one eager step that sizes every
// buffer and builds the planes, then the capture of the same step from
// unbuilt planes.
static int warm_and_capture(dme_ctx *ctx, dme_batch_graph *g, std::vector<JobState *> &js) {
    int launches = 0;
    for (auto &ln : ctx->lane) ln.arena_used = 0;
    DME_TRY(lambda_flag_slots(ctx, js));
    invalidate_batch_planes(ctx, js);
    DME_TRY(queue_parts(ctx, js, nullptr, launches));
    for (auto &ln : ctx->lane) HIP_TRY(hipStreamSynchronize(ln.stream));
    for (auto &ln : ctx->lane) ln.arena_used = 0;
    invalidate_batch_planes(ctx, js);
    launches = 0;
    hipStream_t s0 = ctx->lane[0].stream;
    HIP_TRY(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
    t_capturing = true;
    const int rc = queue_parts(ctx, js, nullptr, launches);
    t_capturing = false;
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s0, &graph);
    if (rc != DME_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return set_error(DME_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    g->graph = graph;
    HIP_TRY(hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0));
    g->launches = launches;
    // the host's plane flags say built again: the warm-up built them from
    // the same contents, and every replay rebuilds them
    return DME_OK;
}

// Wall-clock time of one batch step under the current setting (act).
static int time_step(dme_ctx *ctx, std::vector<JobState *> &js, double &ms) {
    invalidate_batch_planes(ctx, js);
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = run_jobs(ctx, js, nullptr);
    const auto t1 = std::chrono::steady_clock::now();
    DME_TRY(rc);
    ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return DME_OK;
}

static std::string tuning_key(const dme_tuning &t) {
    return std::to_string(t.parts) + "," + std::to_string(t.stagger) + "," + std::to_string(t.pel_wave_threads) + "," +
           std::to_string(t.subpel_wave_threads) + "," + std::to_string(t.mode_wave_threads);
}

}  // namespace dme

extern "C" int dme_set_tuning(dme_ctx *ctx, const dme_tuning *t) try {
    CTX_CHECK(ctx);
    const dme_tuning v = t ? *t : default_tuning();
    if (!tuning_ok(v))
        return set_error(DME_ERR_INVALID, "tuning: parts 1..4, stagger 0/1, pel wave threads 256/512/1024, "
                                          "sub-pel 256/512, mode 256/512/1024");
    ctx->tune = v;
    ctx->tune.step_ms = 0.0f;
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_get_tuning(dme_ctx *ctx, dme_tuning *out) try {
    if (!ctx || !out) return set_error(DME_ERR_INVALID, "NULL argument");
    *out = ctx->tune;
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_autotune_batch(dme_ctx *ctx, const dme_job *jobs, int njobs, int reps, dme_tuning *out) try {
    CTX_CHECK(ctx);
    if (reps < 1 || reps > 100) return set_error(DME_ERR_INVALID, "reps must be in [1, 100]");
    if (ctx->profiling) return set_error(DME_ERR_STATE, "autotuning times whole steps: disable profiling first");
    std::vector<JobState *> js;
    DME_TRY(prepare_batch(ctx, jobs, njobs, js, nullptr));
    std::map<std::string, double> seen;
    // best of reps timed steps after one warm-up step; the setting stays
    // fixed for those steps
    auto measure = [&](const dme_tuning &t, double &best) -> int {
        const std::string k = tuning_key(t);
        auto it = seen.find(k);
        if (it != seen.end()) {
            best = it->second;
            return DME_OK;
        }
        ctx->act = t;
        ctx->act_fixed = true;
        double ms = 0.0;
        int rc = time_step(ctx, js, ms);
        best = 1e30;
        for (int i = 0; i < reps && rc == DME_OK; ++i) {
            rc = time_step(ctx, js, ms);
            best = std::min(best, ms);
        }
        ctx->act_fixed = false;
        DME_TRY(rc);
        seen[k] = best;
        return DME_OK;
    };
    dme_tuning best = default_tuning();
    double best_ms = 0.0;
    DME_TRY(measure(best, best_ms));
    auto consider = [&](const dme_tuning &t) -> int {
        double ms = 0.0;
        DME_TRY(measure(t, ms));
        if (ms < best_ms) {
            best_ms = ms;
            best = t;
        }
        return DME_OK;
    };
    // parts and stagger, then each wavefront kernel's threads
    for (int parts = 1; parts <= kLanes && parts <= njobs; ++parts)
        for (int st = 0; st <= (parts > 1 ? 1 : 0); ++st) {
            dme_tuning t = best;
            t.parts = parts;
            t.stagger = st;
            DME_TRY(consider(t));
        }
    for (int th : {256, 512, 1024}) {
        dme_tuning t = best;
        t.pel_wave_threads = th;
        DME_TRY(consider(t));
    }
    for (int th : {256, 512}) {
        dme_tuning t = best;
        t.subpel_wave_threads = th;
        DME_TRY(consider(t));
    }
    for (int th : {256, 512, 1024}) {
        dme_tuning t = best;
        t.mode_wave_threads = th;
        DME_TRY(consider(t));
    }
    best.step_ms = (float)best_ms;
    ctx->tune_table[batch_signature(js)] = best;
    if (out) *out = best;
    return DME_OK;
} catch (...) {
    if (ctx) ctx->act_fixed = false;
    return dme::exception_status();
}

extern "C" int dme_tuning_save(dme_ctx *ctx, const char *path) try {
    if (!ctx || !path) return set_error(DME_ERR_INVALID, "NULL argument");
    FILE *f = std::fopen(path, "w");
    if (!f) return set_error(DME_ERR_INVALID, std::string("cannot write ") + path);
    std::fprintf(f, "# dme tuning table 1: signature parts stagger pel_wave subpel_wave mode_wave step_ms\n");
    for (auto &kv : ctx->tune_table) {
        const dme_tuning &t = kv.second;
        std::fprintf(f, "%s %d %d %d %d %d %.4f\n", kv.first.c_str(), t.parts, t.stagger, t.pel_wave_threads,
                     t.subpel_wave_threads, t.mode_wave_threads, (double)t.step_ms);
    }
    const bool ok = std::ferror(f) == 0;
    if (std::fclose(f) != 0 || !ok) return set_error(DME_ERR_INVALID, std::string("cannot write ") + path);
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_tuning_load(dme_ctx *ctx, const char *path) try {
    if (!ctx || !path) return set_error(DME_ERR_INVALID, "NULL argument");
    FILE *f = std::fopen(path, "r");
    if (!f) return set_error(DME_ERR_INVALID, std::string("cannot read ") + path);
    std::map<std::string, dme_tuning> add;
    char line[4096];
    int ln = 0;
    bool bad = false;
    while (std::fgets(line, sizeof(line), f)) {
        ++ln;
        if (line[0] == '#' || line[0] == '\n') continue;
        std::istringstream is(line);
        std::string sig;
        dme_tuning t;
        if (!(is >> sig >> t.parts >> t.stagger >> t.pel_wave_threads >> t.subpel_wave_threads >> t.mode_wave_threads >>
              t.step_ms) ||
            !tuning_ok(t)) {
            bad = true;
            break;
        }
        add[sig] = t;
    }
    std::fclose(f);
    if (bad) return set_error(DME_ERR_INVALID, std::string(path) + ": bad tuning line " + std::to_string(ln));
    for (auto &kv : add) ctx->tune_table[kv.first] = kv.second;
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_batch_graph_create(dme_ctx *ctx, const dme_job *jobs, int njobs, dme_batch_graph **out) try {
    CTX_CHECK(ctx);
    if (!out) return set_error(DME_ERR_INVALID, "NULL out");
    *out = nullptr;
    if (ctx->profiling) return set_error(DME_ERR_STATE, "a captured step has no per-kernel events: disable profiling");
    std::unique_ptr<dme_batch_graph> g(new dme_batch_graph());
    g->ctx = ctx;
    std::vector<JobState *> js;
    DME_TRY(prepare_batch(ctx, jobs, njobs, js, &g->jobs));
    std::set<int> nums;
    for (auto *j : js) {
        nums.insert(j->pic_num);
        for (int r = 0; r < j->nrefs; ++r) nums.insert(j->refs[r]);
    }
    for (int n : nums) g->pics.push_back(std::make_pair(n, find_pic(ctx, n)->serial));
    choose_tuning(ctx, js);
    g->tuning = ctx->act;
    for (auto &ln : ctx->lane) HIP_TRY(hipStreamSynchronize(ln.stream));
    swap_graph_buffers(ctx, g.get());
    const int rc = warm_and_capture(ctx, g.get(), js);
    if (rc != DME_OK)
        for (auto &ln : ctx->lane) (void)hipStreamSynchronize(ln.stream);
    swap_graph_buffers(ctx, g.get());
    ctx->L = ctx->L0;
    DME_TRY(rc);
    ctx->graphs.insert(g.get());
    *out = g.release();
    return DME_OK;
} catch (...) {
    dme::t_capturing = false;
    return dme::exception_status();
}

extern "C" int dme_batch_graph_launch(dme_batch_graph *g, dme_medata *outs) try {
    if (!g) return set_error(DME_ERR_INVALID, "NULL graph");
    dme_ctx *ctx = g->ctx;
    DME_TRY(set_device(ctx));
    for (auto &pn : g->pics) {
        const Picture *P = find_pic(ctx, pn.first);
        if (!P || P->serial != pn.second)
            return set_error(DME_ERR_STATE, "picture " + std::to_string(pn.first) +
                                                " was dropped or registered again since the graph was captured");
    }
    std::vector<JobState *> js;
    for (auto &j : g->jobs) js.push_back(j.get());
    hipStream_t s0 = ctx->lane[0].stream;
    for (auto &ln : ctx->lane) ln.arena_used = 0;  // streams idle: every call ends synchronised
    ctx->ev_used.clear();
    HIP_TRY(hipEventRecord(ctx->ev0, s0));
    HIP_TRY(hipGraphLaunch(g->exec, s0));
    DME_TRY(end_timing(ctx, g->launches));
    // the replay built the pictures' planes
    for (auto *j : js) {
        int pn[3] = {j->pic_num, j->refs[0], j->refs[1]};
        for (int k = 0; k <= j->nrefs; ++k) {
            Picture *P0 = find_pic(ctx, pn[k]);
            if (j->p.combined_me) P0->comb_ready = true;
            Picture *P = me_pic(P0, j);
            P->pyr_levels = std::max(P->pyr_levels, j->g.depth);
            if (!j->p.combined_me) P->b8_ready = true;
            if (k > 0) (j->p.combined_me ? P->up_ready : P->qp_ready) = true;
        }
    }
    DevBuf *saved = ctx->lflags;
    ctx->lflags = &g->lflags;
    const int rc = redo_flagged(ctx, js, nullptr);
    ctx->lflags = saved;
    DME_TRY(rc);
    if (outs) DME_TRY(export_batch(ctx, js, outs));
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_batch_graph_destroy(dme_batch_graph *g) try {
    if (!g) return DME_OK;
    dme_ctx *ctx = g->ctx;
    (void)hipSetDevice(ctx->device);
    for (auto &ln : ctx->lane) (void)hipStreamSynchronize(ln.stream);
    ctx->graphs.erase(g);
    delete g;
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

// ---- no-allocation entry point -----------------------------------------------

namespace dme {

// Bump allocation in the scratch array.  Past the end it only counts, so
// that the error can state the size the picture needs.
struct ScratchBump {
    char *base;
    size_t cap, used = 0;
    bool ok = true;
    void give(DevBuf &b, size_t bytes) {
        const size_t off = (used + 255) & ~(size_t)255;
        used = off + bytes;
        if (used > cap) {
            ok = false;
            return;
        }
        b.borrow(base + off, bytes);
    }
};

// Lends every device buffer that one picture's estimation against nrefs
// references can request (prepare_job, build_planes, the stage uploads) out
// of the scratch array, each with the size that is requested of it.
static void noalloc_layout(dme_ctx *ctx, ScratchBump &sb, const Geometry &g, int nrefs) {
    const int W = g.W, H = g.H;
    for (int i = 0; i <= nrefs; ++i) {  // the picture, then its references
        Picture &P = ctx->na_pic[i];
        for (int l = 1; l <= g.depth; ++l) sb.give(P.pyr[l], sizeof(int16_t) * (size_t)(W >> l) * (H >> l));
        const int bp = (W + 2 * kB8Margin + 63) & ~63;
        sb.give(P.b8, sizeof(uint32_t) * (size_t)(H + 2 * kB8Margin) * (bp / 4));
        if (i > 0) sb.give(P.qp, sizeof(uint32_t) * (size_t)(W + 2) * 4 * (H + 2));
    }
    JobState &J = ctx->na_job;
    const size_t B = (size_t)g.xnb * g.ynb, S = (size_t)g.xsb * g.ysb;
    size_t maxt = 0;
    for (int l = 0; l <= g.depth; ++l) maxt = std::max(maxt, (size_t)g.lxb[l] * g.lyb[l] * 28);
    for (int r = 0; r < 2; ++r) {
        sb.give(J.vec[r], sizeof(dme_mv) * B);
        sb.give(J.pc[r], sizeof(dme_cost) * B);
        for (int l = 1; l <= g.depth; ++l) {
            const size_t n = (size_t)g.lxb[l] * g.lyb[l];
            sb.give(J.lvec[r][l], sizeof(dme_mv) * n);
            sb.give(J.lpc[r][l], sizeof(dme_cost) * n);
        }
        sb.give(J.ptab[r], sizeof(int) * maxt);
        sb.give(J.qtab[r], sizeof(int) * 56 * B);
    }
    sb.give(J.m1tab, sizeof(int) * 24 * (B / 4));
    sb.give(J.m0tab, sizeof(int) * 32 * (B / 16));
    sb.give(J.mode, sizeof(int32_t) * B);
    sb.give(J.intra, sizeof(float) * B);
    sb.give(J.bi, sizeof(dme_cost) * B);
    for (int c = 0; c < 3; ++c) {
        sb.give(J.dc[c], sizeof(int16_t) * B);
        sb.give(J.dct[c], sizeof(int16_t) * B);
    }
    sb.give(J.split, sizeof(int32_t) * S);
    sb.give(J.sbc, sizeof(float) * S);
    sb.give(J.lam, sizeof(float) * B);
    sb.give(J.lam1, sizeof(float) * B / 4);
    sb.give(J.lam0, sizeof(float) * B / 16);
    sb.give(J.ucost, sizeof(float) * B);
    sb.give(J.lwork, sizeof(int) * 2 * B);
    sb.give(J.icount, sizeof(int));
    // the lane's descriptor, task and plane-task buffers (one job: at most
    // 22 tasks and 29 plane tasks per upload)
    sb.give(ctx->na_lane.d_jobs, sizeof(JobDesc));
    sb.give(ctx->na_lane.d_tasks, sizeof(Task) * 64);
    sb.give(ctx->na_lane.d_planes, sizeof(PlaneTask) * 64);
    sb.give(ctx->na_lflags, sizeof(int));
    sb.give(ctx->na_flag, sizeof(int));
}

// Restores the context's own lane, flags and pictures when a no-allocation
// call ends.
struct NoallocScope {
    dme_ctx *ctx;
    explicit NoallocScope(dme_ctx *c) : ctx(c) {}
    ~NoallocScope() {
        ctx->na_npics = 0;
        ctx->L0 = &ctx->lane[0];
        ctx->L = &ctx->lane[0];
        ctx->lflags = &ctx->d_lflags;
    }
};

}  // namespace dme

extern "C" int dme_motion_estimate_noalloc(dme_ctx *ctx, const dme_params *params, int width, int height,
                                           int pic_num, const int16_t *d_y, const int16_t *d_u, const int16_t *d_v,
                                           int nrefs, const int *ref_nums, const int16_t *const *d_ref_y,
                                           dme_medata *out) try {
    CTX_CHECK(ctx);
    if (!params || !d_y || !d_u || !d_v || !ref_nums || !d_ref_y)
        return set_error(DME_ERR_INVALID, "NULL argument (the picture needs luma and chroma planes)");
    if (nrefs < 1 || nrefs > 2) return set_error(DME_ERR_INVALID, "nrefs must be 1 or 2");
    for (int r = 0; r < nrefs; ++r)
        if (!d_ref_y[r]) return set_error(DME_ERR_INVALID, "NULL reference plane");
    if (!ctx->scratch) return set_error(DME_ERR_STATE, "no scratch array on this device");
    if (params->combined_me)
        return set_error(DME_ERR_INVALID, "the no-allocation entry point estimates on luma (combined_me = 0)");
    Geometry g;
    DME_TRY(make_geometry(*params, width, height, g));
    {
        const std::string lim = launch_limits(g, params->full_search != 0, false, false);
        if (!lim.empty()) return set_error(DME_ERR_INVALID, "unsupported picture or block size: " + lim);
    }
    // every work buffer from the scratch array; nothing is allocated
    ScratchBump sb{static_cast<char *>(ctx->scratch), (size_t)DME_NOALLOC_SCRATCH_BYTES};
    HIP_TRY(hipStreamSynchronize(ctx->lane[0].stream));  // the previous call's work reads the same array
    noalloc_layout(ctx, sb, g, nrefs);
    if (!sb.ok)
        return set_error(DME_ERR_INVALID, "the picture needs " + std::to_string(sb.used) + " bytes of scratch, more than " +
                                              std::to_string((size_t)DME_NOALLOC_SCRATCH_BYTES));
    for (int i = 0; i <= nrefs; ++i) {
        Picture &P = ctx->na_pic[i];
        P.num = i == 0 ? pic_num : ref_nums[i - 1];
        P.W = width;
        P.H = height;
        P.y = i == 0 ? d_y : d_ref_y[i - 1];
        P.u = i == 0 ? d_u : nullptr;
        P.v = i == 0 ? d_v : nullptr;
        P.pyr_levels = 0;
        P.up_ready = P.qp_ready = P.mcp_ready = P.b8_ready = P.u8_ready = false;
        P.bp = 0;
    }
    NoallocScope scope(ctx);
    ctx->na_npics = 1 + nrefs;
    ctx->L0 = ctx->L = &ctx->na_lane;
    ctx->lflags = &ctx->na_lflags;
    // luma must be 8-bit samples minus 128 (the flag cleared by a fill kernel)
    {
        hipStream_t st = ctx->L->stream;
        int *flag = ctx->na_flag.as<int>();
        HIP_TRY(launch_lp_zero(flag, sizeof(int), st));
        for (int i = 0; i <= nrefs; ++i)
            HIP_TRY(launch_check_range(ctx->na_pic[i].y, (size_t)width * height, flag, st));
        int bad = 0;
        HIP_TRY(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (bad) return set_error(DME_ERR_INVALID, "luma samples must lie in [-128, 127] (8-bit samples minus 128)");
    }
    JobState &J = ctx->na_job;
    DME_TRY(prepare_job(ctx, J, *params, pic_num, ref_nums, nrefs));
    std::vector<JobState *> js{&J};
    DME_TRY(run_jobs(ctx, js, nullptr));
    return export_job(ctx, J, out, hipMemcpyDefault);
} catch (...) {
    return dme::exception_status();
}

// ---- motion compensation ------------------------------------------------------

namespace dme {

// HWts / VWts of MotionCompensator::CreateBlock (mot_comp.cpp:742-806): the
// 1-D OBMC window of length len and separation sep.  `full` ramps the leading
// edge (a unit inside the picture); otherwise the leading edge is flat (a unit
// at the picture's left or top edge).  CalculateWeights (:707-724) builds the
// nine 2-D arrays as products of such windows, the trailing-edge variants by
// mirroring, so the kernels keep the 1-D factors only.
static void mc_window(int len, int sep, bool full, signed char *w) {
    const int off = (len - sep) / 2;
    std::vector<int> h((size_t)len, 8);
    if (off != 1) {
        for (int i = 0; i < 2 * off; ++i) {
            h[i] = 1 + (6 * i + off - 1) / (2 * off - 1);
            h[i + sep] = 8 - h[i];
        }
    } else {
        h[0] = h[1 + sep] = 3;
        h[1] = h[sep] = 5;
    }
    if (!full)
        for (int i = 0; i < 2 * off; ++i) h[i] = 8;
    for (int i = 0; i < len; ++i) w[i] = (signed char)h[i];
}

// Weights of one component: across, per split mode (a unit of 4, 2 or 1
// blocks, ReConfig mot_comp.cpp:197-236) and class (left edge, inside, right
// edge); down, per class (top, inside, bottom).
struct McTables {
    int hpitch = 0, yblen = 0;
    std::vector<signed char> hw;
    std::vector<int> vw;  // int32: the kernels read it with uniform indices
};

static void mc_tables(int xblen, int yblen, int xbsep, int ybsep, McTables &t) {
    const int ov = xblen - xbsep;
    t.hpitch = 4 * xbsep + ov;
    t.hw.assign((size_t)9 * t.hpitch, 0);
    for (int s = 0; s < 3; ++s) {
        const int k = 4 >> s, len = k * xbsep + ov, sep = k * xbsep;
        signed char *row = &t.hw[(size_t)s * 3 * t.hpitch];
        mc_window(len, sep, false, row);
        mc_window(len, sep, true, row + t.hpitch);
        for (int i = 0; i < len; ++i) row[2 * t.hpitch + i] = row[len - 1 - i];
    }
    t.yblen = yblen;
    std::vector<signed char> v((size_t)3 * yblen, 0);
    mc_window(yblen, ybsep, false, &v[0]);
    mc_window(yblen, ybsep, true, &v[yblen]);
    for (int i = 0; i < yblen; ++i) v[2 * (size_t)yblen + i] = v[yblen - 1 - i];
    t.vw.assign(v.begin(), v.end());
}

// NUM_USED_BLKS (mot_comp.cpp:54)
static int mc_used_blocks(int w, int sep, int len) { return (w + sep + (len - sep) / 2 - 1) / sep; }

static size_t mc_align(size_t n) { return (n + 255) & ~(size_t)255; }

constexpr int kMcTileRows = 4 * kMcSegRows;  // rows of a k_mc tile; k_mc_pad tiles have 4

// Validated job: geometry, references and the MvData source.
struct McJob {
    const dme_mc_job *j;
    Geometry g;
    Picture *ref[2];
    JobState *est;  // device MvData of the context's estimate, or null (host MvData)
};

// Queues the compensation of jobs (streams idle, timing started) and waits.
static int mc_run(dme_ctx *ctx, const dme_mc_job *jobs, int njobs) {
    std::vector<McJob> mj((size_t)njobs);
    for (int i = 0; i < njobs; ++i) {
        const dme_mc_job &j = jobs[i];
        McJob &m = mj[i];
        m.j = &j;
        if (j.reserved != 0) return set_error(DME_ERR_INVALID, "dme_mc_job.reserved must be zero");
        if (j.nrefs < 1 || j.nrefs > 2) return set_error(DME_ERR_INVALID, "nrefs must be 1 or 2");
        if (j.direction != DME_MC_ADD && j.direction != DME_MC_SUBTRACT)
            return set_error(DME_ERR_INVALID, "direction must be DME_MC_ADD or DME_MC_SUBTRACT");
        if (j.params.mv_precision < 0 || j.params.mv_precision > 3)
            return set_error(DME_ERR_INVALID, "mv_precision must be 0..3 for motion compensation");
        dme_params q = j.params;
        if (q.mv_precision == 3) q.mv_precision = 2;  // same geometry; eighth pel only changes the sampling
        DME_TRY(make_geometry(q, j.width, j.height, m.g));
        if (m.g.xnb != 4 * m.g.xsb || m.g.ynb != 4 * m.g.ysb) return set_error(DME_ERR_INVALID, "block grid");
        for (int r = 0; r < j.nrefs; ++r) {
            Picture *P = find_pic(ctx, j.refs[r]);
            if (!P) return set_error(DME_ERR_INVALID, "unknown reference picture " + std::to_string(j.refs[r]));
            if (P->W != j.width || P->H != j.height)
                return set_error(DME_ERR_INVALID, "reference picture size differs from the picture's");
            if (!P->u || !P->v) return set_error(DME_ERR_INVALID, "reference picture has no chroma planes");
            m.ref[r] = P;
        }
        if (j.nrefs == 1) m.ref[1] = m.ref[0];
        const int pw[3] = {m.g.W, m.g.CW, m.g.CW}, ph[3] = {m.g.H, m.g.CH, m.g.CH};
        for (int c = 0; c < 3; ++c)
            if (!j.planes[c].data || j.planes[c].width < pw[c] || j.planes[c].height < ph[c])
                return set_error(DME_ERR_INVALID, "plane " + std::to_string(c) + " is NULL or smaller than the picture");
        m.est = nullptr;
        if (j.mv) {
            if (!j.mv->vectors[0] || (j.nrefs == 2 && !j.mv->vectors[1]) || !j.mv->mode || !j.mv->dc[0] ||
                !j.mv->dc[1] || !j.mv->dc[2] || !j.mv->sb_split)
                return set_error(DME_ERR_INVALID, "MvData needs vectors, mode, dc[3] and sb_split");
        } else {
            auto it = ctx->jobs.find(j.pic_num);
            if (it == ctx->jobs.end() || it->second->stage < 4)
                return set_error(DME_ERR_STATE, "no completed estimate for picture " + std::to_string(j.pic_num));
            JobState *J = it->second.get();
            if (J->nrefs != j.nrefs || J->g.xnb != m.g.xnb || J->g.ynb != m.g.ynb || J->g.xsb != m.g.xsb)
                return set_error(DME_ERR_INVALID, "the estimate's references or block grid differ from the job's");
            m.est = J;
        }
    }
    for (int i = 0; i < njobs; ++i)
        for (int k = 0; k < i; ++k)
            for (int c = 0; c < 3; ++c)
                for (int d = 0; d < 3; ++d)
                    if (jobs[i].planes[c].data == jobs[k].planes[d].data)
                        return set_error(DME_ERR_INVALID, "two jobs compensate the same array");

    // up-converted reference planes (Picture::UpData, built once and kept)
    std::vector<PlaneTask> pts;
    int ptotal = 0;
    auto add_task = [&](const int16_t *in, int16_t *out, int w, int h) {
        PlaneTask t;
        t.in = in;
        t.out = out;
        t.in_w = w;
        t.ow = w;
        t.oh = h;
        t.offset = ptotal;
        ptotal += upconvert_tiles(w, h);
        pts.push_back(t);
    };
    std::vector<Picture *> building;  // references whose planes this call builds, each once
    for (auto &m : mj)
        for (int r = 0; r < m.j->nrefs; ++r) {
            Picture *P = m.ref[r];
            if (P->mcp_ready || std::find(building.begin(), building.end(), P) != building.end()) continue;
            // byte planes: the reference's chroma must be 8-bit samples (its luma was checked at registration)
            DME_TRY(check_chroma(ctx, P, "motion compensation"));
            for (int c = 0; c < 3; ++c) {
                const int w = c ? P->W / 2 : P->W, h = c ? P->H / 2 : P->H;
                DME_TRY(P->mcp[c].ensure(4 * (size_t)w * h));
                add_task(c == 0 ? P->y : (c == 1 ? P->u : P->v), P->mcp[c].as<int16_t>(), w, h);
            }
            building.push_back(P);
        }
    int launches = 0;
    if (!pts.empty()) {
        DME_TRY(upload(ctx, ctx->L->d_planes, pts.data(), sizeof(PlaneTask) * pts.size()));
        DME_TRY(timed_launch(ctx, KK_UPCONVERT, launches, [&] {
            return launch_upconvert8(ctx->L->d_planes.as<PlaneTask>(), (int)pts.size(), ptotal, ctx->L->stream);
        }));
        for (Picture *P : building) P->mcp_ready = true;
    }

    // one device blob: component descriptors (main launch, then padding),
    // weight tables, host MvData
    std::vector<McTables> tabs((size_t)njobs * 2);
    size_t ncomp = 3 * (size_t)njobs, npad = 0;
    for (int i = 0; i < njobs; ++i) {
        const Geometry &g = mj[i].g;
        mc_tables(g.blen[2][0], g.blen[2][1], g.bsep[2][0], g.bsep[2][1], tabs[2 * i]);
        mc_tables(g.cblen[0], g.cblen[1], g.cbsep[0], g.cbsep[1], tabs[2 * i + 1]);
        if (jobs[i].direction == DME_MC_ADD)
            for (int c = 0; c < 3; ++c)
                if (jobs[i].planes[c].width > (c ? g.CW : g.W) || jobs[i].planes[c].height > (c ? g.CH : g.H)) ++npad;
    }
    const size_t pack_off = mc_align(sizeof(McComp) * (ncomp + npad));
    size_t bytes = pack_off + mc_align(sizeof(McPackJob) * njobs);
    std::vector<size_t> tab_off(tabs.size()), mv_off((size_t)njobs, 0);
    for (size_t t = 0; t < tabs.size(); ++t) {
        tab_off[t] = bytes;
        bytes += mc_align(tabs[t].hw.size()) + mc_align(sizeof(int) * tabs[t].vw.size());
    }
    for (int i = 0; i < njobs; ++i) {
        if (mj[i].est) continue;
        const size_t B = (size_t)mj[i].g.xnb * mj[i].g.ynb, S = (size_t)mj[i].g.xsb * mj[i].g.ysb;
        mv_off[i] = bytes;
        bytes += mc_align(sizeof(dme_mv) * B) * 2 + mc_align(sizeof(int32_t) * B) + 3 * mc_align(sizeof(int16_t) * B) +
                 mc_align(sizeof(int32_t) * S);
    }
    if (ctx->d_mc.n < bytes) {
        HIP_TRY(hipStreamSynchronize(ctx->L->stream));
        DME_TRY(ctx->d_mc.ensure(bytes));
    }
    size_t nunits = 0;
    for (int i = 0; i < njobs; ++i) nunits += (size_t)mj[i].g.xnb * mj[i].g.ynb;
    if (ctx->d_mc_units.n < sizeof(int4) * 2 * nunits) {
        HIP_TRY(hipStreamSynchronize(ctx->L->stream));
        DME_TRY(ctx->d_mc_units.ensure(sizeof(int4) * 2 * nunits));
    }
    char *dev = ctx->d_mc.as<char>();
    std::vector<char> blob(bytes, 0);
    McPackJob *packs = reinterpret_cast<McPackJob *>(blob.data() + pack_off);
    int units_done = 0;
    for (size_t t = 0; t < tabs.size(); ++t) {
        std::memcpy(&blob[tab_off[t]], tabs[t].hw.data(), tabs[t].hw.size());
        std::memcpy(&blob[tab_off[t] + mc_align(tabs[t].hw.size())], tabs[t].vw.data(), sizeof(int) * tabs[t].vw.size());
    }
    McComp *cs = reinterpret_cast<McComp *>(blob.data());
    int total = 0, ptot = 0;
    size_t pi = ncomp;
    for (int i = 0; i < njobs; ++i) {
        const dme_mc_job &j = jobs[i];
        const McJob &m = mj[i];
        const Geometry &g = m.g;
        const size_t B = (size_t)g.xnb * g.ynb, S = (size_t)g.xsb * g.ysb;
        const dme_mv *vec[2];
        const int *mode, *split;
        const int16_t *dc[3];
        if (m.est) {
            vec[0] = m.est->vec[0].as<dme_mv>();
            vec[1] = (j.nrefs == 2 ? m.est->vec[1] : m.est->vec[0]).as<dme_mv>();
            mode = m.est->mode.as<int>();
            split = m.est->split.as<int>();
            for (int c = 0; c < 3; ++c) dc[c] = m.est->dc[c].as<int16_t>();
        } else {
            size_t o = mv_off[i];
            const size_t vb = mc_align(sizeof(dme_mv) * B);
            std::memcpy(&blob[o], j.mv->vectors[0], sizeof(dme_mv) * B);
            std::memcpy(&blob[o + vb], j.mv->vectors[j.nrefs == 2 ? 1 : 0], sizeof(dme_mv) * B);
            vec[0] = reinterpret_cast<const dme_mv *>(dev + o);
            vec[1] = reinterpret_cast<const dme_mv *>(dev + o + vb);
            o += 2 * vb;
            std::memcpy(&blob[o], j.mv->mode, sizeof(int32_t) * B);
            mode = reinterpret_cast<const int *>(dev + o);
            o += mc_align(sizeof(int32_t) * B);
            for (int c = 0; c < 3; ++c) {
                std::memcpy(&blob[o], j.mv->dc[c], sizeof(int16_t) * B);
                dc[c] = reinterpret_cast<const int16_t *>(dev + o);
                o += mc_align(sizeof(int16_t) * B);
            }
            std::memcpy(&blob[o], j.mv->sb_split, sizeof(int32_t) * S);
            split = reinterpret_cast<const int *>(dev + o);
        }
        McPackJob &pk = packs[i];
        pk.vec[0] = vec[0];
        pk.vec[1] = vec[1];
        pk.mode = mode;
        for (int c = 0; c < 3; ++c) pk.dc[c] = dc[c];
        pk.out = ctx->d_mc_units.as<int4>() + 2 * (size_t)units_done;
        pk.blocks = (int)B;
        pk.offset = units_done;
        units_done += (int)B;
        for (int c = 0; c < 3; ++c) {
            McComp &C = cs[3 * i + c];
            const McTables &T = tabs[2 * i + (c ? 1 : 0)];
            const size_t to = tab_off[2 * i + (c ? 1 : 0)];
            const int xblen = c ? g.cblen[0] : g.blen[2][0], yblen = c ? g.cblen[1] : g.blen[2][1];
            const int xbsep = c ? g.cbsep[0] : g.bsep[2][0], ybsep = c ? g.cbsep[1] : g.bsep[2][1];
            const Picture *R0 = m.ref[0], *R1 = m.ref[1];
            C.up[0] = R0->mcp[c].as<uint8_t>();
            C.up[1] = R1->mcp[c].as<uint8_t>();
            C.out = j.planes[c].data;
            C.unit = pk.out;
            C.split = split;
            C.comp = c;
            C.hw = reinterpret_cast<const signed char *>(dev + to);
            C.vw = reinterpret_cast<const int *>(dev + to + mc_align(T.hw.size()));
            C.lx = j.planes[c].width;
            C.ly = j.planes[c].height;
            C.pw = c ? g.CW : g.W;
            C.ph = c ? g.CH : g.H;
            C.xbsep = xbsep;
            C.ybsep = ybsep;
            C.xoff = (xblen - xbsep) / 2;
            C.yoff = (yblen - ybsep) / 2;
            C.xov = xblen - xbsep;
            C.yblen = yblen;
            C.hpitch = T.hpitch;
            C.nbx = std::min(mc_used_blocks(C.pw, xbsep, xblen), g.xnb);
            C.nby = std::min(mc_used_blocks(C.ph, ybsep, yblen), g.ynb);
            C.xnb = g.xnb;
            C.xsb = g.xsb;
            C.shift = c ? 1 : 0;
            C.prec = j.params.mv_precision;
            C.direction = j.direction;
            C.mmx = j.params.fp_semantics == DME_FP_MMX;
            C.tiles_x = (C.lx + 63) / 64;
            C.offset = total;
            total += C.tiles_x * ((C.ly + kMcTileRows - 1) / kMcTileRows);
            if (j.direction == DME_MC_ADD && (C.lx > C.pw || C.ly > C.ph)) {
                cs[pi] = C;
                cs[pi].offset = ptot;
                ptot += C.tiles_x * ((C.ly + 3) / 4);
                ++pi;
            }
        }
    }
    DME_TRY(upload(ctx, ctx->d_mc, blob.data(), bytes));
    const McComp *dcs = ctx->d_mc.as<McComp>();
    DME_TRY(timed_launch(ctx, KK_MC_PACK, launches, [&] {
        return launch_mc_pack(reinterpret_cast<const McPackJob *>(dev + pack_off), njobs, units_done, ctx->L->stream);
    }));
    DME_TRY(timed_launch(ctx, KK_MC, launches,
                         [&] { return launch_mc(dcs, (int)ncomp, total, false, ctx->L->stream); }));
    if (npad)
        DME_TRY(timed_launch(ctx, KK_MC_PAD, launches,
                             [&] { return launch_mc(dcs + ncomp, (int)npad, ptot, true, ctx->L->stream); }));
    return end_timing(ctx, launches);
}

}  // namespace dme

extern "C" int dme_motion_compensate_batch(dme_ctx *ctx, const dme_mc_job *jobs, int njobs) try {
    CTX_CHECK(ctx);
    if (!jobs || njobs <= 0) return set_error(DME_ERR_INVALID, "empty batch");
    DME_TRY(begin_timing(ctx));
    return mc_run(ctx, jobs, njobs);
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_motion_compensate(dme_ctx *ctx, const dme_params *params, int width, int height,
                                     const dme_medata *mv, const int *refs, int nrefs, int direction,
                                     const dme_mc_plane planes[3]) try {
    CTX_CHECK(ctx);
    if (!params || !mv || !refs || !planes) return set_error(DME_ERR_INVALID, "NULL argument");
    if (nrefs < 1 || nrefs > 2) return set_error(DME_ERR_INVALID, "nrefs must be 1 or 2");
    size_t n[3], total = 0;
    for (int c = 0; c < 3; ++c) {
        if (!planes[c].data || planes[c].width <= 0 || planes[c].height <= 0)
            return set_error(DME_ERR_INVALID, "plane " + std::to_string(c) + " is NULL or empty");
        n[c] = (size_t)planes[c].width * planes[c].height;
        total += mc_align(sizeof(int16_t) * n[c]);
    }
    DME_TRY(begin_timing(ctx));
    DME_TRY(ctx->d_mc_planes.ensure(total));
    dme_mc_job j;
    std::memset(&j, 0, sizeof(j));
    j.pic_num = -1;
    j.width = width;
    j.height = height;
    j.nrefs = nrefs;
    for (int r = 0; r < nrefs; ++r) j.refs[r] = refs[r];
    j.direction = direction;
    j.params = *params;
    j.mv = mv;
    char *d = ctx->d_mc_planes.as<char>();
    for (int c = 0; c < 3; ++c) {
        j.planes[c].data = reinterpret_cast<int16_t *>(d);
        j.planes[c].width = planes[c].width;
        j.planes[c].height = planes[c].height;
        HIP_TRY(hipMemcpyAsync(d, planes[c].data, sizeof(int16_t) * n[c], hipMemcpyHostToDevice, ctx->L->stream));
        d += mc_align(sizeof(int16_t) * n[c]);
    }
    DME_TRY(mc_run(ctx, &j, 1));
    for (int c = 0; c < 3; ++c)
        HIP_TRY(hipMemcpyAsync(planes[c].data, j.planes[c].data, sizeof(int16_t) * n[c], hipMemcpyDeviceToHost,
                               ctx->L->stream));
    HIP_TRY(hipStreamSynchronize(ctx->L->stream));
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_get_plane(dme_ctx *ctx, int pic_num, int kind, int level, int16_t *out, size_t out_elems) try {
    CTX_CHECK(ctx);
    Picture *P = find_pic(ctx, pic_num);
    if (!P || !out) return set_error(DME_ERR_INVALID, "unknown picture or NULL output");
    const void *src = nullptr;
    size_t n = 0;
    if (kind == 0) {
        if (level < 1 || level > P->pyr_levels) return set_error(DME_ERR_STATE, "pyramid level not built");
        src = P->pyr[level].p;
        n = (size_t)(P->W >> level) * (P->H >> level);
    } else if (kind == 1) {
        // the search reads the quarter-pel or byte up-planes only; the
        // up-converted plane itself is built here, on request
        if (!P->qp_ready && !P->u8_ready) return set_error(DME_ERR_STATE, "up-converted plane not built");
        if (!P->up_ready) {
            HIP_TRY(hipStreamSynchronize(ctx->L->stream));
            DME_TRY(P->up.ensure(sizeof(int16_t) * 4 * (size_t)P->W * P->H));
            PlaneTask t;
            t.in = P->y;
            t.out = P->up.as<int16_t>();
            t.in_w = P->W;
            t.ow = P->W;
            t.oh = P->H;
            t.offset = 0;
            DME_TRY(upload(ctx, ctx->L->d_planes, &t, sizeof(t)));
            HIP_TRY(launch_upconvert(ctx->L->d_planes.as<PlaneTask>(), 1, upconvert_tiles(P->W, P->H), ctx->L->stream));
            HIP_TRY(hipStreamSynchronize(ctx->L->stream));
            ctx->L->arena_used = 0;
            P->up_ready = true;
        }
        src = P->up.p;
        n = 4 * (size_t)P->W * P->H;
    } else if (kind == 2) {
        if (!P->u8_ready) return set_error(DME_ERR_STATE, "byte up-plane not built");
        n = (size_t)up8_pitch(P->W) * up8_rows(P->H);
        if (out_elems < n) return set_error(DME_ERR_INVALID, "output buffer too small");
        std::vector<uint8_t> b(n);
        HIP_TRY(hipStreamSynchronize(ctx->L->stream));
        HIP_TRY(hipMemcpy(b.data(), P->u8.p, n, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) out[i] = (int16_t)((int)b[i] - 128);
        return DME_OK;
    } else if (kind >= 3 && kind <= 5) {
        // the combined-component picture, its pyramid levels, its up plane
        const Picture *C = P->cpic.get();
        if (!C || !P->comb_ready) return set_error(DME_ERR_STATE, "combined picture not built");
        if (kind == 3) {
            src = C->y;
            n = (size_t)P->W * P->H;
        } else if (kind == 4) {
            if (level < 1 || level > C->pyr_levels) return set_error(DME_ERR_STATE, "pyramid level not built");
            src = C->pyr[level].p;
            n = (size_t)(P->W >> level) * (P->H >> level);
        } else {
            if (!C->up_ready) return set_error(DME_ERR_STATE, "up-converted plane not built");
            src = C->up.p;
            n = 4 * (size_t)P->W * P->H;
        }
    } else {
        return set_error(DME_ERR_INVALID, "kind must be 0 .. 5");
    }
    if (out_elems < n) return set_error(DME_ERR_INVALID, "output buffer too small");
    HIP_TRY(hipStreamSynchronize(ctx->L->stream));
    HIP_TRY(hipMemcpy(out, src, sizeof(int16_t) * n, hipMemcpyDeviceToHost));
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_synchronize(dme_ctx *ctx) try {
    CTX_CHECK(ctx);
    HIP_TRY(hipStreamSynchronize(ctx->L->stream));
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_invalidate_caches(dme_ctx *ctx, int pic_num) try {
    CTX_CHECK(ctx);
    HIP_TRY(hipStreamSynchronize(ctx->L->stream));
    for (auto &kv : ctx->pics)
        if (pic_num < 0 || kv.first == pic_num) {
            kv.second->pyr_levels = 0;
            kv.second->up_ready = false;
            kv.second->mcp_ready = false;
            kv.second->qp_ready = false;
            kv.second->b8_ready = false;
            kv.second->u8_ready = false;
            kv.second->comb_ready = false;
            if (kv.second->cpic) {
                kv.second->cpic->pyr_levels = 0;
                kv.second->cpic->up_ready = false;
            }
        }
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_set_profiling(dme_ctx *ctx, int enable) try {
    if (!ctx) return set_error(DME_ERR_INVALID, "NULL context");
    ctx->profiling = enable != 0;
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" const char *dme_kernel_name(int k) {
    return (k >= 0 && k < DME_NUM_KERNELS && kKernelNames[k]) ? kKernelNames[k] : "";
}

extern "C" int dme_get_stats(dme_ctx *ctx, dme_stats *out) try {
    if (!ctx || !out) return set_error(DME_ERR_INVALID, "NULL argument");
    *out = ctx->stats;
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

// ---- learned proposal --------------------------------------------------------

namespace dme {

// Pel stage with learned proposals: the plain search down to level 2, then
// the network over every picture's level-0 block grid (pictures with equal
// grids form one NHWC tensor), exact refinement and the pel costs.
// Positions of a batch in the proposal tensors, from `base`, grouped by
// block grid so that every group is one run of a convolution launch.
struct LpGroup {
    int off, n, W, H;
};
static int lp_layout(std::vector<JobState *> &js, int base, std::vector<LpGroup> &groups) {
    std::vector<size_t> order(js.size());
    for (size_t i = 0; i < js.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        return std::make_pair(js[a]->g.xnb, js[a]->g.ynb) < std::make_pair(js[b]->g.xnb, js[b]->g.ynb);
    });
    groups.clear();
    int P = base;
    for (size_t o = 0; o < order.size(); ++o) {
        JobState *j = js[order[o]];
        if (groups.empty() || groups.back().W != j->g.xnb || groups.back().H != j->g.ynb)
            groups.push_back(LpGroup{P, 0, j->g.xnb, j->g.ynb});
        groups.back().n++;
        j->desc.lp_off = P;
        j->desc.lp_slot = j->bpos;
        P += j->g.xnb * j->g.ynb;
    }
    return P;
}

// Sums the refinement counters of the slots (host copy h of lp_stats).
static void lp_sum_stats(const std::vector<unsigned long long> &h, size_t nslots, long long &blocks,
                         long long &fallbacks) {
    for (size_t i = 0; i < nslots; ++i)
        for (int c = 0; c < kLpStatCopies; ++c) {
            blocks += (long long)h[i * kLpStatSlot + c * kLpStatStride];
            fallbacks += (long long)h[i * kLpStatSlot + c * kLpStatStride + 1];
        }
}

// Proposal tensors for P positions and counters for nslots jobs (streams
// idle).  Counters of a smaller array are added to the totals first.
static int lp_reserve(dme_ctx *ctx, int P, size_t nslots) {
    for (int a = 0; a < 4; ++a) DME_TRY(ctx->lp_act[a].ensure(sizeof(uint16_t) * 32 * (size_t)P));
    DME_TRY(ctx->lp_out.ensure(sizeof(float) * 32 * (size_t)P));
    if (nslots > ctx->lp_slots) {
        if (ctx->lp_slots) {
            std::vector<unsigned long long> h(kLpStatSlot * ctx->lp_slots);
            HIP_TRY(hipMemcpy(h.data(), ctx->lp_stats.p, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost));
            lp_sum_stats(h, ctx->lp_slots, ctx->lp_blocks, ctx->lp_fallbacks);
        }
        const size_t n = std::max(nslots, (size_t)64);
        DME_TRY(ctx->lp_stats.ensure(sizeof(unsigned long long) * kLpStatSlot * n));
        HIP_TRY(launch_lp_zero(ctx->lp_stats.p, sizeof(unsigned long long) * kLpStatSlot * n, ctx->L->stream));
        HIP_TRY(hipStreamSynchronize(ctx->L->stream));
        ctx->lp_slots = n;
    }
    return DME_OK;
}

// Adds the refinement counters of the last learned batch to the context's
// totals and zeroes them, so that the slots of a batch hold its own counts
// (streams idle: every call ends with them idle).
static int lp_fold_stats(dme_ctx *ctx) {
    if (!ctx->lp_slots || !ctx->lp_dirty) return DME_OK;
    std::vector<unsigned long long> h(kLpStatSlot * ctx->lp_slots);
    HIP_TRY(hipMemcpy(h.data(), ctx->lp_stats.p, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost));
    lp_sum_stats(h, ctx->lp_slots, ctx->lp_blocks, ctx->lp_fallbacks);
    HIP_TRY(launch_lp_zero(ctx->lp_stats.p, sizeof(unsigned long long) * h.size(), ctx->L->stream));
    HIP_TRY(hipStreamSynchronize(ctx->L->stream));
    ctx->lp_dirty = false;
    return DME_OK;
}

// Learned pel stage of a batch laid out by lp_layout: the plain search of
// the levels down to 2, the network on the level-0 block grids, then the
// exact refinement and the pel costs.  Queues work only.
static int stage_pel_learned(dme_ctx *ctx, std::vector<JobState *> &js, const std::vector<LpGroup> &groups,
                             int &launches) {
    if (!ctx->lp_ready) return set_error(DME_ERR_STATE, "dme_lp_set_weights was not called");
    for (auto *j : js)
        if (j->g.depth < 2) return set_error(DME_ERR_INVALID, "learned proposals need a pyramid depth of at least 2");
    DME_TRY(stage_pel(ctx, js, launches, 2));
    DME_TRY(upload_jobs(ctx, js));
    TaskTable tt;
    std::vector<Task> fb, rb;
    int offf = 0, offr = 0;
    for (size_t i = 0; i < js.size(); ++i) {
        const int cnt = js[i]->g.xnb * js[i]->g.ynb;
        fb.push_back(Task{(int)i, 0, cnt, offf});
        offf += cnt;
        for (int r = 0; r < js[i]->nrefs; ++r) {
            rb.push_back(Task{(int)i, r, cnt, offr});
            offr += cnt;
        }
    }
    const TaskTable::Range feat = tt.add(fb), refine = tt.add(rb);
    DME_TRY(upload_tasks(ctx, tt));
    const JobDesc *dj = ctx->L->d_jobs.as<JobDesc>();
    uint16_t *act[4];
    for (int a = 0; a < 4; ++a) act[a] = ctx->lp_act[a].as<uint16_t>();
    const uint16_t *frag = ctx->lp_frag.as<uint16_t>();
    const float *w = ctx->lp_w.as<float>();
    DME_TRY(timed_launch(ctx, KK_LP_NET, launches, [&]() -> hipError_t {
        hipError_t e = launch_lp_features(dj, dtask(ctx, feat), feat.n, feat.total, act[0], nullptr, ctx->L->stream);
        for (const LpGroup &g : groups) {
            const size_t o = (size_t)g.off * 32;
            const int npos = g.n * g.W * g.H;
            for (int layer = 0; layer < 4 && e == hipSuccess; ++layer) {
                const uint16_t *in = act[layer] + o;
                const uint16_t *fr = frag + (size_t)layer * 2 * DME_LP_FRAG_ELEMS;  // forward fragments
                const float *bias = w + (size_t)layer * DME_LP_LAYER_PARAMS + 9 * 32 * 32;
                if (layer < 3)
                    e = launch_lp_conv(in, npos, g.H, g.W, fr, bias, 1, act[layer + 1] + o, nullptr, ctx->L->stream);
                else
                    e = launch_lp_conv(in, npos, g.H, g.W, fr, bias, 0, nullptr, ctx->lp_out.as<float>() + o,
                                       ctx->L->stream);
            }
        }
        return e;
    }));
    int rbx = 1, rby = 1;
    for (auto *j : js) {
        rbx = std::max(rbx, j->g.blen[2][0]);
        rby = std::max(rby, j->g.blen[2][1]);
    }
    ctx->lp_dirty = true;
    DME_TRY(timed_launch(ctx, KK_LP_REFINE, launches, [&]() -> hipError_t {
        hipError_t e = launch_lp_refine(dj, dtask(ctx, refine), refine.n, refine.total, ctx->lp_out.as<float>(),
                                        ctx->lp_stats.as<unsigned long long>(), rbx, rby, ctx->L->stream);
        if (e == hipSuccess) e = launch_lp_pel_costs(dj, dtask(ctx, refine), refine.n, refine.total, ctx->L->stream);
        return e;
    }));
    for (auto *j : js) j->stage = 1;
    return DME_OK;
}

// The learned pel stage of the fast form: the plain search of the coarse
// levels down to 2 (the network's inputs are that search's level-2 results),
// the network on the level-0 block grids and the exact refinement of its
// proposals.  The level-0 vectors keep their SAD; their mvcost and total are
// not needed (the sub-pel stage recomputes both).
static int stage_pel_fast(dme_ctx *ctx, std::vector<JobState *> &js, const std::vector<LpGroup> &groups,
                          int &launches) {
    if (!ctx->lp_ready) return set_error(DME_ERR_STATE, "dme_lp_set_weights was not called");
    int bx = 1, by = 1;
    for (auto *j : js) {
        if (j->g.depth < 2) return set_error(DME_ERR_INVALID, "learned proposals need a pyramid depth of at least 2");
        bx = std::max(bx, j->g.blen[2][0]);
        by = std::max(by, j->g.blen[2][1]);
    }
    DME_TRY(stage_pel(ctx, js, launches, 2));
    DME_TRY(upload_jobs(ctx, js));
    TaskTable tt;
    std::vector<Task> fb, rb;
    int offf = 0, offr = 0;
    for (size_t i = 0; i < js.size(); ++i) {
        const int cnt = js[i]->g.xnb * js[i]->g.ynb;
        fb.push_back(Task{(int)i, 0, cnt, offf});
        offf += cnt;
        for (int r = 0; r < js[i]->nrefs; ++r) {
            rb.push_back(Task{(int)i, r, cnt, offr});
            offr += cnt;
        }
    }
    const TaskTable::Range feat = tt.add(fb), refine = tt.add(rb);
    DME_TRY(upload_tasks(ctx, tt));
    const JobDesc *dj = ctx->L->d_jobs.as<JobDesc>();
    hipStream_t st = ctx->L->stream;
    uint16_t *act[4];
    for (int a = 0; a < 4; ++a) act[a] = ctx->lp_act[a].as<uint16_t>();
    const uint16_t *frag = ctx->lp_frag.as<uint16_t>();
    const float *w = ctx->lp_w.as<float>();
    DME_TRY(timed_launch(ctx, KK_LP_NET, launches, [&]() -> hipError_t {
        hipError_t e = launch_lp_features(dj, dtask(ctx, feat), feat.n, feat.total, act[0], nullptr, st);
        for (const LpGroup &g : groups) {
            const size_t o = (size_t)g.off * 32;
            const int npos = g.n * g.W * g.H;
            for (int layer = 0; layer < 4 && e == hipSuccess; ++layer) {
                const uint16_t *in = act[layer] + o;
                const uint16_t *fr = frag + (size_t)layer * 2 * DME_LP_FRAG_ELEMS;
                const float *bias = w + (size_t)layer * DME_LP_LAYER_PARAMS + 9 * 32 * 32;
                if (layer < 3)
                    e = launch_lp_conv(in, npos, g.H, g.W, fr, bias, 1, act[layer + 1] + o, nullptr, st);
                else
                    e = launch_lp_conv(in, npos, g.H, g.W, fr, bias, 0, nullptr, ctx->lp_out.as<float>() + o, st);
            }
        }
        return e;
    }));
    ctx->lp_dirty = true;
    DME_TRY(timed_launch(ctx, KK_LP_REFINE, launches, [&] {
        return launch_lp_refine(dj, dtask(ctx, refine), refine.n, refine.total, ctx->lp_out.as<float>(),
                                ctx->lp_stats.as<unsigned long long>(), bx, by, st);
    }));
    for (auto *j : js) j->stage = 1;
    return DME_OK;
}

// Sub-pel stage of the fast form: every block refined at once with the
// predictor of the integer vectors (k_fast_subpel), then the level-2 costs of
// the final vectors with the reference's predictor (k_fast_l2).
static int stage_subpel_fast(dme_ctx *ctx, std::vector<JobState *> &js, int &launches) {
    DME_TRY(bind_planes(ctx, js, true, launches));
    DME_TRY(upload_jobs(ctx, js));
    TaskTable tt;
    std::vector<Task> bt;
    int off = 0, bx = 1, by = 1;
    for (size_t i = 0; i < js.size(); ++i)
        for (int r = 0; r < js[i]->nrefs; ++r) {
            const int cnt = js[i]->g.xnb * js[i]->g.ynb;
            bt.push_back(Task{(int)i, r, cnt, off});
            off += cnt;
            bx = std::max(bx, js[i]->g.blen[2][0]);
            by = std::max(by, js[i]->g.blen[2][1]);
        }
    const TaskTable::Range blk = tt.add(bt);
    DME_TRY(upload_tasks(ctx, tt));
    const JobDesc *dj = ctx->L->d_jobs.as<JobDesc>();
    hipStream_t st = ctx->L->stream;
    DME_TRY(timed_launch(ctx, KK_FAST_SUBPEL, launches,
                         [&] { return launch_fast_subpel(dj, dtask(ctx, blk), blk.n, blk.total, bx, by, st); }));
    DME_TRY(timed_launch(ctx, KK_FAST_L2, launches,
                         [&] { return launch_fast_l2(dj, dtask(ctx, blk), blk.n, blk.total, st); }));
    for (auto *j : js) j->stage = 3;
    return DME_OK;
}

// Mode decision of the fast form: the level-2 unit decisions (k_mode_l2, as
// the plain path), levels 1 and 0 of every superblock at once (k_fast_mode),
// the reference's costs of the chosen split (k_fast_eval), SetDC.
static int stage_mode_fast(dme_ctx *ctx, std::vector<JobState *> &js, int &launches) {
    DME_TRY(bind_planes(ctx, js, true, launches));
    DME_TRY(upload_jobs(ctx, js));
    TaskTable tt;
    std::vector<Task> l2, sb, dc, dr;
    int off = 0, offs = 0, offd = 0, offr = 0, max_w = 0, ux = 1, uy = 1;
    for (size_t i = 0; i < js.size(); ++i) {
        const Geometry &g = js[i]->g;
        const int cnt = g.xnb * g.ynb;
        l2.push_back(Task{(int)i, 0, cnt, off});
        off += cnt;
        sb.push_back(Task{(int)i, 0, g.xsb * g.ysb, offs});
        offs += g.xsb * g.ysb;
        for (int c = 0; c < 3; ++c) {
            dc.push_back(Task{(int)i, c, cnt, offd});
            offd += cnt;
            dr.push_back(Task{(int)i, c, g.ynb, offr});
            offr += g.ynb;
        }
        max_w = std::max(max_w, g.W);
        ux = std::max(ux, std::max(g.blen[0][0], g.blen[1][0]));
        uy = std::max(uy, std::max(g.blen[0][1], g.blen[1][1]));
    }
    const TaskTable::Range rl2 = tt.add(l2), rsb = tt.add(sb), rdc = tt.add(dc), rdr = tt.add(dr);
    DME_TRY(upload_tasks(ctx, tt));
    const JobDesc *dj = ctx->L->d_jobs.as<JobDesc>();
    hipStream_t st = ctx->L->stream;
    DME_TRY(timed_launch(ctx, KK_MODE_L2, launches,
                         [&] { return launch_mode_l2(dj, dtask(ctx, rl2), rl2.n, rl2.total, true, DME_KIND_BELOW_EIGHTH, st); }));
    DME_TRY(timed_launch(ctx, KK_FAST_MODE, launches,
                         [&] { return launch_fast_mode(dj, dtask(ctx, rsb), rsb.n, rsb.total, ux, uy, st); }));
    DME_TRY(timed_launch(ctx, KK_FAST_EVAL, launches,
                         [&] { return launch_fast_eval(dj, dtask(ctx, rsb), rsb.n, rsb.total, st); }));
    DME_TRY(timed_launch(ctx, KK_BLOCK_DC, launches,
                         [&] { return launch_block_dc(dj, dtask(ctx, rdr), rdr.n, rdr.total, max_w, st); }));
    DME_TRY(timed_launch(ctx, KK_DC_SMOOTH, launches,
                         [&] { return launch_dc_smooth(dj, dtask(ctx, rdc), rdc.n, rdc.total, st); }));
    DME_TRY(timed_launch(ctx, KK_INTRA, launches, [&] { return launch_intra_count(dj, (int)js.size(), st); }));
    for (auto *j : js) j->stage = 4;
    return DME_OK;
}

// The stages of a part of a learned batch, in the form the context selects.
static int learned_pel(dme_ctx *ctx, std::vector<JobState *> &js, const std::vector<LpGroup> &g, int &launches) {
    return ctx->lp_form == DME_LP_FORM_FAST ? stage_pel_fast(ctx, js, g, launches)
                                            : stage_pel_learned(ctx, js, g, launches);
}
static int learned_rest(dme_ctx *ctx, std::vector<JobState *> &js, int &launches) {
    if (ctx->lp_form == DME_LP_FORM_FAST) {
        DME_TRY(stage_subpel_fast(ctx, js, launches));
        return stage_mode_fast(ctx, js, launches);
    }
    DME_TRY(stage_subpel(ctx, js, launches));
    return stage_mode(ctx, js, launches);
}

// The live sequence with learned proposals, in parts on the lanes as
// run_jobs does.
// Jobs whose device lambda maps were flagged run again (learned pel stage,
// host maps).
static int run_jobs_learned_parts(dme_ctx *ctx, std::vector<JobState *> &js) {
    int launches = 0;
    DME_TRY(lp_fold_stats(ctx));
    DME_TRY(begin_timing(ctx));
    DME_TRY(lambda_flag_slots(ctx, js));
    const int np = (js.size() < 2 || ctx->profiling) ? 1 : (int)std::min<size_t>((size_t)ctx->tune.parts, js.size());
    std::vector<std::vector<JobState *>> part(np);
    for (size_t i = 0; i < js.size(); ++i) part[i % np].push_back(js[i]);
    std::vector<std::vector<LpGroup>> groups(np);
    int P = 0;
    for (int h = 0; h < np; ++h) P = lp_layout(part[h], P, groups[h]);
    DME_TRY(lp_reserve(ctx, P, js.size()));
    if (np == 1) {
        DME_TRY(learned_pel(ctx, js, groups[0], launches));
        DME_TRY(stage_lambda(ctx, js, nullptr, false, launches));
        DME_TRY(learned_rest(ctx, js, launches));
    } else {
        // as run_jobs: shared planes before the fork, each part's own planes
        // after the previous part's
        DME_TRY(build_shared_planes(ctx, part, launches));
        HIP_TRY(hipEventRecord(ctx->ev_fork, ctx->lane[0].stream));
        for (int h = 1; h < np; ++h) HIP_TRY(hipStreamWaitEvent(ctx->lane[h].stream, ctx->ev_fork, 0));
        for (int h = 0; h < np; ++h) {
            ctx->L = &ctx->lane[h];
            if (ctx->tune.stagger) {
                if (h > 0) HIP_TRY(hipStreamWaitEvent(ctx->L->stream, ctx->ev_planes[h - 1], 0));
                DME_TRY(bind_planes(ctx, part[h], true, launches));
                HIP_TRY(hipEventRecord(ctx->ev_planes[h], ctx->L->stream));
            }
            DME_TRY(learned_pel(ctx, part[h], groups[h], launches));
        }
        for (int h = 0; h < np; ++h) {
            ctx->L = &ctx->lane[h];
            DME_TRY(stage_lambda(ctx, part[h], nullptr, false, launches));
            DME_TRY(learned_rest(ctx, part[h], launches));
        }
        ctx->L = &ctx->lane[0];
        for (int h = 1; h < np; ++h) {
            HIP_TRY(hipEventRecord(ctx->ev_join[h], ctx->lane[h].stream));
            HIP_TRY(hipStreamWaitEvent(ctx->L->stream, ctx->ev_join[h], 0));
        }
    }
    DME_TRY(end_timing(ctx, launches));
    std::vector<JobState *> redo;
    DME_TRY(read_lambda_flags(ctx, js, redo));
    ctx->stats.host_lambda_jobs = (int32_t)redo.size();
    if (!redo.empty()) {
        std::vector<LpGroup> g;
        lp_layout(redo, 0, g);
        // the slots hold this batch's counts only (lp_fold_stats): the
        // redone jobs count anew
        for (auto *j : redo)
            HIP_TRY(launch_lp_zero(ctx->lp_stats.as<unsigned long long>() + (size_t)kLpStatSlot * j->bpos,
                                   kLpStatSlot * sizeof(unsigned long long), ctx->L->stream));
        DME_TRY(learned_pel(ctx, redo, g, launches));
        DME_TRY(stage_lambda_host(ctx, redo, nullptr));
        DME_TRY(learned_rest(ctx, redo, launches));
        HIP_TRY(hipStreamSynchronize(ctx->L->stream));
    }
    return DME_OK;
}

// The fast form decides sub-pel vectors, so it needs half or quarter pel,
// and costs them with the MMX build's float semantics.
static int learned_params_ok(dme_ctx *ctx, const dme_params &p) {
    if (ctx->lp_form != DME_LP_FORM_FAST) return DME_OK;
    if (p.mv_precision < 1 || p.mv_precision > 2)
        return set_error(DME_ERR_INVALID, "the fast learned form refines to half or quarter pel (mv_precision 1 or 2)");
    if (p.fp_semantics != DME_FP_MMX)
        return set_error(DME_ERR_INVALID, "the fast learned form uses the MMX build's float semantics");
    return DME_OK;
}

static int run_jobs_learned(dme_ctx *ctx, std::vector<JobState *> &js) {
    const int rc = run_jobs_learned_parts(ctx, js);
    if (rc != DME_OK) sync_lanes(ctx);
    return rc;
}

}  // namespace dme

extern "C" int dme_lp_set_weights(dme_ctx *ctx, const float *params, size_t count) try {
    CTX_CHECK(ctx);
    if (!params || count != (size_t)DME_LP_PARAMS)
        return set_error(DME_ERR_INVALID, "the network has " + std::to_string(DME_LP_PARAMS) + " parameters");
    for (size_t i = 0; i < count; ++i)
        if (!std::isfinite(params[i])) return set_error(DME_ERR_INVALID, "non-finite network parameter");
    HIP_TRY(hipStreamSynchronize(ctx->L->stream));
    DME_TRY(ctx->lp_w.ensure(sizeof(float) * count));
    DME_TRY(ctx->lp_frag.ensure(sizeof(uint16_t) * 4 * 2 * DME_LP_FRAG_ELEMS));
    HIP_TRY(hipMemcpy(ctx->lp_w.p, params, sizeof(float) * count, hipMemcpyHostToDevice));
    for (int layer = 0; layer < 4; ++layer) {
        uint16_t *f = ctx->lp_frag.as<uint16_t>() + (size_t)layer * 2 * DME_LP_FRAG_ELEMS;
        HIP_TRY(launch_lp_frags(ctx->lp_w.as<float>() + (size_t)layer * DME_LP_LAYER_PARAMS, f, f + DME_LP_FRAG_ELEMS,
                                ctx->L->stream));
    }
    HIP_TRY(hipStreamSynchronize(ctx->L->stream));
    ctx->lp_ready = true;
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_motion_estimate_learned(dme_ctx *ctx, int pic_num, const int *refs, int nrefs, dme_medata *out) try {
    CTX_CHECK(ctx);
    if (!ctx->params_set) return set_error(DME_ERR_STATE, "dme_set_params was not called");
    JobState *J = job_for(ctx, pic_num);
    if (ctx->params.full_search)
        return set_error(DME_ERR_INVALID, "learned proposals replace the hierarchical search; use full search without them");
    DME_TRY(learned_params_ok(ctx, ctx->params));
    DME_TRY(prepare_job(ctx, *J, ctx->params, pic_num, refs, nrefs, true));
    if (J->p.full_search)
        return set_error(DME_ERR_INVALID, "learned proposals replace the hierarchical search; use full search without them");
    std::vector<JobState *> js{J};
    DME_TRY(run_jobs_learned(ctx, js));
    return export_job(ctx, *J, out);
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_run_batch_learned(dme_ctx *ctx, const dme_job *jobs, int njobs, dme_medata *outs) try {
    CTX_CHECK(ctx);
    if (!jobs || njobs <= 0) return set_error(DME_ERR_INVALID, "empty batch");
    std::vector<JobState *> js;
    for (int i = 0; i < njobs; ++i) {
        for (int k = 0; k < i; ++k)
            if (jobs[k].pic_num == jobs[i].pic_num) return set_error(DME_ERR_INVALID, "a picture appears twice in a batch");
        JobState *J = job_for(ctx, jobs[i].pic_num);
        if (jobs[i].params.full_search)
            return set_error(DME_ERR_INVALID, "learned proposals replace the hierarchical search; use full search without them");
        DME_TRY(learned_params_ok(ctx, jobs[i].params));
        DME_TRY(prepare_job(ctx, *J, jobs[i].params, jobs[i].pic_num, jobs[i].refs, jobs[i].nrefs, true));
        if (J->p.full_search)
            return set_error(DME_ERR_INVALID, "learned proposals replace the hierarchical search; use full search without them");
        js.push_back(J);
    }
    DME_TRY(run_jobs_learned(ctx, js));
    if (outs)
        DME_TRY(export_batch(ctx, js, outs));
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_lp_set_form(dme_ctx *ctx, int form) try {
    CTX_CHECK(ctx);
    if (form != DME_LP_FORM_FAST && form != DME_LP_FORM_EXACT_STAGES)
        return set_error(DME_ERR_INVALID, "form must be DME_LP_FORM_FAST or DME_LP_FORM_EXACT_STAGES");
    ctx->lp_form = form;
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_lp_get_vectors(dme_ctx *ctx, int pic_num, int which, dme_mv *out1, dme_mv *out2) try {
    CTX_CHECK(ctx);
    auto it = ctx->jobs.find(pic_num);
    if (it == ctx->jobs.end() || it->second->stage < 4 || !it->second->fv0[0].p)
        return set_error(DME_ERR_STATE, "no fast learned estimate of picture " + std::to_string(pic_num));
    if (which != 0 && which != 1) return set_error(DME_ERR_INVALID, "which must be 0 or 1");
    JobState &J = *it->second;
    const size_t B = (size_t)J.g.xnb * J.g.ynb;
    dme_mv *out[2] = {out1, out2};
    for (int r = 0; r < J.nrefs; ++r)
        if (out[r])
            HIP_TRY(hipMemcpy(out[r], which ? J.fv2[r].p : J.fv0[r].p, sizeof(dme_mv) * B, hipMemcpyDeviceToHost));
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_lp_get_outputs(dme_ctx *ctx, int pic_num, float *out) try {
    CTX_CHECK(ctx);
    if (!out) return set_error(DME_ERR_INVALID, "NULL output");
    auto it = ctx->jobs.find(pic_num);
    if (it == ctx->jobs.end() || it->second->stage < 4 || !it->second->fv0[0].p)
        return set_error(DME_ERR_STATE, "no fast learned estimate of picture " + std::to_string(pic_num));
    JobState &J = *it->second;
    const size_t B = (size_t)J.g.xnb * J.g.ynb;
    if (ctx->lp_out.n < sizeof(float) * 32 * ((size_t)J.desc.lp_off + B))
        return set_error(DME_ERR_STATE, "the network outputs of picture " + std::to_string(pic_num) + " are gone");
    for (auto &ln : ctx->lane) HIP_TRY(hipStreamSynchronize(ln.stream));
    HIP_TRY(hipMemcpy(out, ctx->lp_out.as<float>() + (size_t)32 * J.desc.lp_off, sizeof(float) * 32 * B,
                      hipMemcpyDeviceToHost));
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_lp_get_stats(dme_ctx *ctx, long long *blocks, long long *fallbacks) try {
    CTX_CHECK(ctx);
    long long b = ctx->lp_blocks, f = ctx->lp_fallbacks;
    if (ctx->lp_slots) {
        for (auto &ln : ctx->lane) HIP_TRY(hipStreamSynchronize(ln.stream));
        std::vector<unsigned long long> h(kLpStatSlot * ctx->lp_slots);
        HIP_TRY(hipMemcpy(h.data(), ctx->lp_stats.p, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost));
        lp_sum_stats(h, ctx->lp_slots, b, f);
    }
    if (blocks) *blocks = b;
    if (fallbacks) *fallbacks = f;
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_lp_training_pair(dme_ctx *ctx, int pic_num, uint16_t *d_features, float *d_labels) try {
    CTX_CHECK(ctx);
    if (!d_features || !d_labels) return set_error(DME_ERR_INVALID, "NULL output");
    auto it = ctx->jobs.find(pic_num);
    if (it == ctx->jobs.end() || it->second->stage < 1)
        return set_error(DME_ERR_STATE, "no pixel search has run for picture " + std::to_string(pic_num));
    JobState *J = it->second.get();
    if (J->g.depth < 2) return set_error(DME_ERR_INVALID, "learned proposals need a pyramid depth of at least 2");
    std::vector<JobState *> js{J};
    int launches = 0;
    HIP_TRY(hipStreamSynchronize(ctx->L->stream));
    ctx->L->arena_used = 0;
    DME_TRY(bind_planes(ctx, js, true, launches));
    J->desc.lp_off = 0;
    DME_TRY(upload_jobs(ctx, js));
    TaskTable tt;
    const int cnt = J->g.xnb * J->g.ynb;
    const TaskTable::Range feat = tt.add(std::vector<Task>{Task{0, 0, cnt, 0}});
    DME_TRY(upload_tasks(ctx, tt));
    HIP_TRY(launch_lp_features(ctx->L->d_jobs.as<JobDesc>(), dtask(ctx, feat), feat.n, feat.total, d_features, d_labels,
                               ctx->L->stream));
    HIP_TRY(hipStreamSynchronize(ctx->L->stream));
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_lp_prepare(dme_ctx *ctx, const float *d_w, uint16_t *d_frag_fwd, uint16_t *d_frag_bwd) try {
    CTX_CHECK(ctx);
    if (!d_w || !d_frag_fwd || !d_frag_bwd) return set_error(DME_ERR_INVALID, "NULL argument");
    HIP_TRY(launch_lp_frags(d_w, d_frag_fwd, d_frag_bwd, ctx->L->stream));
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_lp_conv(dme_ctx *ctx, const uint16_t *d_in, int n, int h, int w, const uint16_t *d_frag,
                           const float *d_bias, int relu, uint16_t *d_out, float *d_out_f32) try {
    CTX_CHECK(ctx);
    if (!d_in || !d_frag || (!d_out && !d_out_f32) || n <= 0 || h <= 0 || w <= 0)
        return set_error(DME_ERR_INVALID, "bad convolution arguments");
    HIP_TRY(launch_lp_conv(d_in, n * h * w, h, w, d_frag, d_bias, relu, d_out, d_out_f32, ctx->L->stream));
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_lp_wgrad(dme_ctx *ctx, const uint16_t *d_in, const uint16_t *d_g, int n, int h, int w, float *d_dw,
                            float *d_db) try {
    CTX_CHECK(ctx);
    if (!d_in || !d_g || !d_dw || !d_db || n <= 0 || h <= 0 || w <= 0)
        return set_error(DME_ERR_INVALID, "bad gradient arguments");
    HIP_TRY(launch_lp_wgrad(d_in, d_g, n * h * w, h, w, d_dw, d_db, ctx->L->stream));
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_lp_relu_backward(dme_ctx *ctx, const uint16_t *d_act, const float *d_g, size_t count,
                                    uint16_t *d_out) try {
    CTX_CHECK(ctx);
    if (!d_g || !d_out) return set_error(DME_ERR_INVALID, "NULL argument");
    HIP_TRY(launch_lp_relu_bwd(d_act, d_g, count, d_out, ctx->L->stream));
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_lp_wta_loss(dme_ctx *ctx, const float *d_out, const float *d_labels, int npos, float alpha,
                               float beta, float scale, float *d_dout, float *d_loss) try {
    CTX_CHECK(ctx);
    if (!d_out || !d_labels || !d_dout || !d_loss || npos <= 0) return set_error(DME_ERR_INVALID, "bad loss arguments");
    HIP_TRY(launch_lp_wta(d_out, d_labels, npos, alpha, beta, scale, d_dout, d_loss, ctx->L->stream));
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_lp_adam(dme_ctx *ctx, float *d_param, const float *d_grad, float *d_m, float *d_v, int count,
                           float lr, float beta1, float beta2, float eps, int t) try {
    CTX_CHECK(ctx);
    if (!d_param || !d_grad || !d_m || !d_v || count <= 0 || t < 1) return set_error(DME_ERR_INVALID, "bad Adam arguments");
    const float bc1 = 1.0f - std::pow(beta1, (float)t), bc2 = 1.0f - std::pow(beta2, (float)t);
    HIP_TRY(launch_lp_adam(d_param, d_grad, d_m, d_v, count, lr, beta1, beta2, eps, bc1, bc2, ctx->L->stream));
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_lp_zero(dme_ctx *ctx, void *d_ptr, size_t bytes) try {
    CTX_CHECK(ctx);
    if (!d_ptr || (bytes & 3)) return set_error(DME_ERR_INVALID, "zeroing needs a pointer and a multiple of 4 bytes");
    HIP_TRY(launch_lp_zero(d_ptr, bytes, ctx->L->stream));
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_lp_alloc(dme_ctx *ctx, size_t bytes, void **d_ptr) try {
    CTX_CHECK(ctx);
    if (!d_ptr || bytes == 0) return set_error(DME_ERR_INVALID, "bad allocation arguments");
    *d_ptr = nullptr;
    if (hipMalloc(d_ptr, bytes) != hipSuccess) {
        *d_ptr = nullptr;
        (void)hipGetLastError();
        return set_error(DME_ERR_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_lp_free(dme_ctx *ctx, void *d_ptr) try {
    CTX_CHECK(ctx);
    HIP_TRY(hipStreamSynchronize(ctx->L->stream));
    if (d_ptr) HIP_TRY(hipFree(d_ptr));
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_lp_copy(dme_ctx *ctx, void *dst, const void *src, size_t bytes, int kind) try {
    CTX_CHECK(ctx);
    if (!dst || !src || kind < 1 || kind > 3) return set_error(DME_ERR_INVALID, "bad copy arguments");
    const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : (kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, k, ctx->L->stream));
    HIP_TRY(hipStreamSynchronize(ctx->L->stream));
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}
