/*
 * dme.h -- C ABI of the MI355X (gfx950) motion-estimation engine for the
 * Dirac encoder.
 *
 * The reference exposes this path only as C++ classes in libdirac_motionest
 * (there is no plugin registry or FFI in the reference).  Each entry point
 * below replaces one of those class methods; the citation on each names the
 * method and the file:line of its declaration in the reference tree.  All
 * array layouts are those of the reference's MEData (libdirac_common/motion.h)
 * flattened row-major, because TwoDArray is contiguous (arrays.h:502-542):
 * a caller can pass &array[0][0] of its own MEData and get the results
 * written in place.
 *
 * Conventions
 *   - every function returns an int status: DME_OK (0) or a negative code;
 *     dme_last_error() returns a message for the calling thread's last error.
 *   - no C++ exceptions cross this boundary.
 *   - one dme_ctx per host thread; a context is bound to one HIP device.
 *   - results are deterministic and bit-exact against the reference for the
 *     same inputs and parameters (MMX build semantics by default).
 *   - picture samples are int16, 8-bit video stored minus 128, as the
 *     reference stores them (libdirac_common/pic_io.cpp:396-401).
 */
#ifndef DME_H
#define DME_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define DME_OK               0
#define DME_ERR_INVALID     -1  /* bad argument or unsupported parameter set  */
#define DME_ERR_HIP         -2  /* a HIP runtime call failed                  */
#define DME_ERR_NOMEM       -3  /* device or host allocation failed           */
#define DME_ERR_STATE       -4  /* call out of order (e.g. subpel before pel)  */
#define DME_ERR_NO_DEVICE   -5  /* no HIP device / kernels missing            */

/* ---- prediction modes: values of enum PredMode (common.h:84) ------------ */
#define DME_INTRA      0
#define DME_REF1_ONLY  1
#define DME_REF2_ONLY  2
#define DME_REF1AND2   3

/* ---- float semantics of the sub-pel cost functions ---------------------- */
/* The reference's default build defines HAVE_MMX (configure.ac:303-326); its
 * in-bounds up-sampled SAD adds the integer SAD to the start cost once
 * (me_utils_mmx.cpp:270-278).  The plain C build accumulates per pixel in
 * float (me_utils.cpp:532-541).  Both are supported.                        */
#define DME_FP_MMX  0
#define DME_FP_C    1

/* MVector (motion.h:57-183, MotionVector<int>) */
typedef struct dme_mv {
    int32_t x, y;
} dme_mv;

/* MvCostData (motion.h:192-211) */
typedef struct dme_cost {
    float SAD, mvcost, total;
} dme_cost;

/* Motion-estimation parameters.  Block parameters are the raw level-2 values
 * handed to PicturePredParams::SetBlockSizes (common.cpp:207-320), which
 * rationalises them exactly as the reference does; level-1/0 parameters are
 * derived from them there.                                                   */
typedef struct dme_params {
    int32_t xblen, yblen, xbsep, ybsep;
    int32_t mv_precision;  /* 0 pel, 1 half, 2 quarter, 3 eighth
                              (MVPrecisionType); the learned path's fast
                              form takes 1 or 2                              */
    int32_t depth;         /* pyramid depth; < 0 selects the reference's own
                              formula floor(min(log2(W/12), log2(H/12)))
                              (pixel_match.cpp:94-95)                          */
    float   lambda;        /* ME Lagrangian (L1MELambda / L2MELambda)        */
    int32_t fp_semantics;  /* DME_FP_MMX (default) or DME_FP_C               */
    int32_t chroma_format; /* 0 = 4:2:0 (only value supported)               */
    int32_t full_search;   /* EncoderParams::FullSearch (common.h:1308): no
                              pyramid (depth is ignored); the zero list is
                              the whole window of radius min(tdiff, 3) times
                              the ranges below, tdiff = |reference - picture|
                              picture numbers (pixel_match.cpp:79-80,
                              249-264), and the prediction list has radius 1
                              (pixel_match.cpp:365-370)                      */
    int32_t x_range_me;    /* XRangeME / YRangeME (common.h:1311-1314), used  */
    int32_t y_range_me;    /* only with full_search; the encoder default is 32 */
    int32_t combined_me;   /* EncoderParams::CombinedME (common.h:1317): 1 =
                              ME on the combined-component picture of every
                              picture (EncPicture::DataForME/UpDataForME,
                              enc_picture.cpp:114-128; Combine, :285-367),
                              whose samples need 9 bits; the picture and its
                              references must have chroma in [-128, 127];
                              SetDC still reads luma.  Not with full search,
                              the learned forms or the no-allocation entry
                              point (DME_ERR_INVALID).  0 = ME on luma, the
                              encoder default                               */
    int32_t reserved[3];   /* must be zero                                   */
} dme_params;

/* Derived geometry for a picture of a given size (SetMotionParameters,
 * seq_compress.cpp:107-133; MakeMEDataHierarchy, pixel_match.cpp:185-220).   */
#define DME_MAX_DEPTH 8
typedef struct dme_geometry {
    int32_t width, height, chroma_width, chroma_height;
    int32_t xnum_sb, ynum_sb, xnum_blocks, ynum_blocks;
    int32_t depth;
    int32_t level_width[DME_MAX_DEPTH + 1];   /* pyramid picture sizes        */
    int32_t level_height[DME_MAX_DEPTH + 1];
    int32_t level_xblocks[DME_MAX_DEPTH + 1]; /* pel-search block grid       */
    int32_t level_yblocks[DME_MAX_DEPTH + 1];
    int32_t luma_blen[3][2], luma_bsep[3][2]; /* [level][x/y] after rationalising */
    int32_t chroma_blen[3][2], chroma_bsep[3][2];
    float   level_factor[3];                  /* ModeDecider level factors   */
} dme_geometry;

/* MEData of one picture (motion.h:219-410), caller-owned, any pointer may be
 * NULL to skip that array.  Shapes: B = ynum_blocks*xnum_blocks elements in
 * raster order, S = ynum_sb*xnum_sb.  Index [r] is reference r+1 (the
 * reference indexes Vectors/PredCosts with Range(1,num_refs)).               */
typedef struct dme_medata {
    dme_mv   *vectors[2];      /* B  Vectors(1), Vectors(2)                   */
    dme_cost *pred_costs[2];   /* B  PredCosts(1), PredCosts(2)               */
    int32_t  *mode;            /* B  Mode (PredMode values)                   */
    float    *intra_costs;     /* B  IntraCosts                               */
    dme_cost *bipred_costs;    /* B  BiPredCosts                              */
    int16_t  *dc[3];           /* B  DC(Y_COMP), DC(U_COMP), DC(V_COMP)       */
    int32_t  *sb_split;        /* S  SBSplit                                  */
    float    *sb_costs;        /* S  SBCosts                                  */
    float    *lambda_map;      /* B  LambdaMap                                */
    float    *intra_block_ratio; /* 1  MEData::IntraBlockRatio (a float,
                                  * motion.h:356-359): INTRA blocks over all
                                  * blocks of the final Mode array, divided
                                  * in double (PictureCompressor::
                                  * IntraModeAnalyse, picture_compress.cpp:
                                  * 300-320)                                 */
} dme_medata;

/* ---- MV data wire format (host only) ----------------------------------- *
 * The bytes of an inter picture's motion data in the bitstream: what
 * PictureCompressor::CodeMVData writes (picture_compress.cpp:463-530) and
 * PictureDecompressor::DecompressMVData reads (picture_decompress.cpp:
 * 266-357) -- MvDataByteIO's header (mvdata_byteio.cpp: block parameters,
 * MV precision, no global motion, default prediction mode, picture
 * weights), then the arithmetic-coded split modes, prediction modes, vector
 * elements of reference 1 (x, y) and 2 (x, y) and DC values of Y, U, V
 * (SplitModeCodec, PredModeCodec, VectorElementCodec, DCCodec, mv_codec.cpp:
 * 49-641), each preceded by its byte count (mvdataelement_byteio.cpp).  The
 * arrays are the estimator's, in the dme_medata layout; only sb_split, mode,
 * vectors and dc are read or written.                                       */
typedef struct dme_mvdata_info {
    int32_t xnum_sb, ynum_sb;    /* superblock grid; blocks are 4x each way   */
    int32_t num_refs;            /* PictureParams::NumRefs(): 1 or 2          */
    int32_t xblen, yblen, xbsep, ybsep; /* PicturePredParams::LumaBParams(2):
                                    dme_geometry luma_blen[2], luma_bsep[2]  */
    int32_t mv_precision;        /* 0 pel .. 3 eighth pel                     */
    int32_t weights_bits, ref1_weight, ref2_weight; /* picture weights; the
                                    encoder's defaults are 1, 1, 1           */
} dme_mvdata_info;

/* Writes one picture's MV data.  Sets *len to its byte count; with buf ==
 * NULL only sizes it, with cap < *len writes nothing and fails.
 * Replaces PictureCompressor::CodeMVData (picture_compress.cpp:463).        */
int dme_mvdata_encode(const dme_mvdata_info *info, const dme_medata *md,
                      uint8_t *buf, size_t cap, size_t *len);
/* Reads one picture's MV data from buf[0..len): info gives the grid and
 * num_refs and receives the header's fields; md's sb_split, mode, vectors
 * and dc receive the data, each unit's values in all its blocks and zero
 * where nothing is coded (vectors of references a block does not use, DC of
 * inter blocks).  *consumed (may be NULL) = bytes read.  Malformed data
 * fails (naming the element) and leaves the arrays untouched.
 * Replaces PictureDecompressor::DecompressMVData (picture_decompress.cpp:266). */
int dme_mvdata_decode(const uint8_t *buf, size_t len, dme_mvdata_info *info,
                      dme_medata *md, size_t *consumed);
/* dme_mvdata_encode of n pictures on up to `threads` host threads (<= 0: one
 * per core); lens[i] receives each byte count.  No counterpart: the
 * reference codes one picture at a time.                                    */
int dme_mvdata_encode_batch(int n, const dme_mvdata_info *info,
                            const dme_medata *md, uint8_t *const *bufs,
                            const size_t *caps, size_t *lens, int threads);

typedef struct dme_ctx dme_ctx;

/* ---- host-only helpers (no GPU needed) --------------------------------- */

/* Returns the message of the calling thread's last failed call ("" if none). */
const char *dme_last_error(void);

/* Library/ABI version: major*10000 + minor*100 + patch. */
int dme_version(void);

/* Geometry for a width x height 4:2:0 picture under params.
 * Replaces SequenceCompressor::SetMotionParameters (seq_compress.cpp:107-133)
 * + PicturePredParams::SetBlockSizes (common.cpp:207-320)
 * + PixelMatcher::MakeMEDataHierarchy (pixel_match.cpp:185-220).           */
int dme_geometry_for(const dme_params *params, int width, int height,
                     dme_geometry *out);

/* Host restatement of MEData::SetLambdaMap(num_refs, lambda) with
 * FindTransitions (motion.cpp:128-171, 201-333), computed in x87 long double
 * exactly as the reference does.  vectors[r] are the pel-search vectors of
 * reference r+1 (B elements each), out_lambda_map receives B floats.        */
int dme_lambda_map_host(int xnum_blocks, int ynum_blocks, int num_refs,
                        const dme_mv *vectors1, const dme_mv *vectors2,
                        float lambda, float *out_lambda_map);

/* ---- .imt diagnostics text (host only) ---------------------------------- *
 * The encoder's on-disk MEData text: a sequence header, then one record per
 * coded picture.  Numbers are written as the reference's iostreams write them
 * (ints in decimal, floats as printf "%.6g" of the value), one array row per
 * line, each value followed by a space.  Writers return DME_OK and set *len
 * to the record's byte count; with buf == NULL they only size it, and with
 * cap < *len they write nothing and fail.  No terminating NUL is written.
 * A record's size is at most 13 bytes per number (a float or int32 takes at
 * most 12 characters, plus its space), one per row, and 256 for its first
 * lines and blank lines: a buffer of that bound needs no sizing call.
 * Readers parse one header or record at text[0..len), set *consumed to the
 * offset after it and the whitespace that follows, and fail (naming the
 * field) on malformed or out-of-range input.                              */

typedef struct dme_imt_sequence {
    int32_t chroma;               /* ChromaFormat value (format420 = 2)        */
    int32_t width, height;
    int32_t source_sampling;
    int32_t topfieldfirst;        /* 0 or 1                                    */
    int32_t frame_rate_num, frame_rate_den;
    int32_t pix_asr_num, pix_asr_den;
    int32_t picture_coding_mode;  /* 0 frames, 1 fields                        */
} dme_imt_sequence;

typedef struct dme_imt_picture {
    int32_t pic_num;
    int32_t intra;                /* 1: an ">intra" record, no MEData          */
    int32_t num_refs;             /* 1 or 2 (0 for an intra record)            */
    int32_t refs[2];
    int32_t ybsep, xbsep;         /* luma block separation, dme_geometry
                                     luma_bsep[2][1] and [2][0]                */
    int32_t sb_ylen, sb_xlen;     /* ynum_sb, xnum_sb                          */
    int32_t mv_ylen, mv_xlen;     /* ynum_blocks, xnum_blocks                  */
} dme_imt_picture;

/* Sequence header.  Replaces WriteDiagnosticsHeader (encoder/encmain.cpp:
 * 165-193).                                                                */
int dme_imt_write_sequence(const dme_imt_sequence *seq, char *buf, size_t cap,
                           size_t *len);

/* One picture record from MEData arrays laid out as for the estimator
 * (lambda_map is not part of the format and is ignored; an intra record
 * reads no arrays, me may be NULL).  Replaces DiracEncoder::
 * GetInstrumentationData (libdirac_encoder/dirac_encoder.cpp:338-410) with
 * WriteDiagnosticsData (encoder/encmain.cpp:195-330); the text after the
 * record's first line is what operator<<(ostream&, MEData&) writes
 * (libdirac_common/motion.cpp:452-485).                                    */
int dme_imt_write_picture(const dme_imt_picture *pic, const dme_medata *me,
                          char *buf, size_t cap, size_t *len);

/* Sequence header.  Replaces ReadInstrumentationHeader
 * (util/instrumentation/instrmain.cpp:84-118).                             */
int dme_imt_read_sequence(const char *text, size_t len, dme_imt_sequence *seq,
                          size_t *consumed);

/* One picture record.  Replaces ProcessSequence::AddPictureEntry
 * (util/instrumentation/process_sequence.cpp:119-210) with
 * operator>>(istream&, MEData&) (libdirac_common/motion.cpp:414-449).
 * With me == NULL, pic is output only and the arrays are parsed and
 * dropped: call so to learn the shapes.  With me != NULL, pic's four
 * lengths on entry are the caller's array shapes and must equal the
 * record's; NULL array pointers in me skip that array.  pred_costs and
 * bipred_costs get SAD and mvcost; their total is not in the text and is
 * left unchanged.                                                          */
int dme_imt_read_picture(const char *text, size_t len, dme_imt_picture *pic,
                         dme_medata *me, size_t *consumed);

/* ---- device context ---------------------------------------------------- */

/* Creates a context on HIP device `device`, with its own stream. */
int dme_create(int device, dme_ctx **out);
int dme_destroy(dme_ctx *ctx);

/* Sets the parameters for subsequent calls (PixelMatcher/SubpelRefine/
 * ModeDecider constructors take EncoderParams, pixel_match.h:71,
 * me_subpel.h:66, me_mode_decn.h:79).                                       */
int dme_set_params(dme_ctx *ctx, const dme_params *params);

/* Registers picture `pic_num` (EncQueue::PushPicture + SetOrigData,
 * enc_queue.cpp:179, enc_picture.cpp:100).  Planes are host pointers with
 * stride = width (luma) and width/2 (chroma, 4:2:0); u and v may be NULL if
 * the picture is only ever used as a reference.  The planes are copied to
 * device memory.  Luma samples must lie in [-128, 127] (8-bit video stored
 * minus 128, pic_io.cpp:396-401); other values give DME_ERR_INVALID.       */
int dme_upload_picture(dme_ctx *ctx, int pic_num, int width, int height,
                       const int16_t *y, const int16_t *u, const int16_t *v);

/* As dme_upload_picture, but the planes are already resident in device
 * memory of the context's device and are used in place (not copied).  They
 * must stay valid and unchanged until dme_drop_picture.                     */
int dme_bind_picture_device(dme_ctx *ctx, int pic_num, int width, int height,
                            const int16_t *d_y, const int16_t *d_u,
                            const int16_t *d_v);

/* Releases a picture and its cached pyramid / up-converted planes
 * (EncQueue::ClearAll / EncPicture dtor).                                   */
int dme_drop_picture(dme_ctx *ctx, int pic_num);

/* PixelMatcher::DoSearch(EncQueue&, pic_num)  (pixel_match.h:81,
 * pixel_match.cpp:59-157): hierarchical integer search of pic_num against
 * refs[0..nrefs-1] (nrefs 1 or 2).  Fills out->vectors / out->pred_costs
 * (pel units) if non-NULL.  Also initialises the picture's MEData
 * (EncPicture::InitMEData, enc_picture.cpp:106-112).                       */
int dme_pixel_search(dme_ctx *ctx, int pic_num, const int *refs, int nrefs,
                     dme_medata *out);

/* MEData::SetLambdaMap(num_refs, lambda) (motion.h:362, motion.cpp:128-171),
 * computed on the host from the pel-search vectors and uploaded.  If
 * lambda_map is non-NULL it is used as given instead (B floats).            */
int dme_set_lambda_map(dme_ctx *ctx, int pic_num, const float *lambda_map,
                       dme_medata *out);

/* The MEData state the next stage reads, from the caller: the reference's
 * classes read it from the MEData the caller owns (PictureCompressor edits
 * it between stages: SetLambdaMap, and the pixel-precision doubling of
 * picture_compress.cpp:240-259), so a stage-wise caller can hand it back.
 * Prepares pic_num's estimation against refs under the context's current
 * parameters (dme_set_params), then loads, from `in`, the vectors and
 * PredCosts of every reference (stage 1: after PixelMatcher::DoSearch),
 * and the lambda map (stage 2: after MEData::SetLambdaMap), or the same
 * after sub-pel refinement (stage 3: ModeDecider::DoModeDecn follows).
 * Vectors are in the units of the current mv_precision (pel at stages 1-2). */
int dme_load_medata(dme_ctx *ctx, int pic_num, const int *refs, int nrefs,
                    const dme_medata *in, int stage);

/* SubpelRefine::DoSubpel(EncQueue&, pic_num) (me_subpel.h:77,
 * me_subpel.cpp:65-99), or the pixel-precision vector doubling of
 * PictureCompressor::SubPixelME (picture_compress.cpp:240-259).            */
int dme_subpel(dme_ctx *ctx, int pic_num, dme_medata *out);

/* ModeDecider::DoModeDecn(EncQueue&, pic_num) (me_mode_decn.h:93,
 * me_mode_decn.cpp:66-177) including SetDC, and the pixel-precision undo of
 * PictureCompressor::ModeDecisionME (picture_compress.cpp:274-296).  Writes
 * every non-NULL array of out.                                              */
int dme_mode_decide(dme_ctx *ctx, int pic_num, dme_medata *out);

/* The whole live ME sequence of one inter picture, in the order
 * PictureCompressor runs it (picture_compress.cpp:76-80, 208-298):
 * PixelME -> SetLambdaMap -> SubPixelME -> ModeDecisionME.  This is what the
 * reference's (dead) MotionEstimator::DoME (motion_estimate.h:79) declares.  */
int dme_motion_estimate(dme_ctx *ctx, int pic_num, const int *refs, int nrefs,
                        dme_medata *out);

/* Batched form: runs the live sequence for njobs pictures in one pass, all
 * stages of all pictures sharing launches.  Pictures may differ in size,
 * reference count and block parameters (per-job params).  outs[i] receives
 * job i's MEData.  No counterpart in the reference (which handles one
 * picture at a time, seq_compress.cpp:256-310).                              */
typedef struct dme_job {
    int32_t pic_num;
    int32_t nrefs;
    int32_t refs[2];
    dme_params params;
} dme_job;
int dme_run_batch(dme_ctx *ctx, const dme_job *jobs, int njobs,
                  dme_medata *outs);

/* ---- launch tuning and the captured batch step ------------------------- *
 * No counterpart in the reference.  A batch runs as `parts` parts of about
 * equal work, one per stream, each part's planes built after the previous
 * part's when `stagger` is set; the three anti-diagonal wavefront kernels run
 * one workgroup per (picture, reference) or picture with the given thread
 * counts.  No setting changes any result.  A batch whose signature (the
 * multiset of its jobs' sizes, reference counts and parameters) is in the
 * context's tuning table runs with the table's setting; any other with the
 * context's default setting (dme_set_tuning).                               */
typedef struct dme_tuning {
    int32_t parts;               /* 1..4                                      */
    int32_t stagger;             /* 0 or 1                                    */
    int32_t pel_wave_threads;    /* k_pel_wave: 256, 512 or 1024              */
    int32_t subpel_wave_threads; /* k_subpel_wave: 256 or 512                 */
    int32_t mode_wave_threads;   /* k_mode_wave: 256, 512 or 1024             */
    float   step_ms;             /* dme_autotune_batch: the setting's best
                                    measured step (0 elsewhere)               */
} dme_tuning;
/* Default setting of batches not in the tuning table (NULL: the built-in
 * default 4 parts, staggered, 1024 / 512 / 1024 threads).                   */
int dme_set_tuning(dme_ctx *ctx, const dme_tuning *t);
int dme_get_tuning(dme_ctx *ctx, dme_tuning *out);
/* Times the batch under candidate settings (reps timed steps each, after one
 * warm-up step; coordinate search: parts and stagger, then each wavefront
 * kernel's threads), stores the fastest in the tuning table under the
 * batch's signature and returns it in *out (may be NULL).  The pictures'
 * planes are rebuilt in every step, as dme_invalidate_caches leaves them.    */
int dme_autotune_batch(dme_ctx *ctx, const dme_job *jobs, int njobs, int reps,
                       dme_tuning *out);
/* The tuning table as text, one line per signature; load adds to (and
 * replaces entries of) the context's table.                                 */
int dme_tuning_save(dme_ctx *ctx, const char *path);
int dme_tuning_load(dme_ctx *ctx, const char *path);

/* A batch step captured once as a HIP graph and replayed.  create runs the
 * batch once (sizing every buffer), then captures the whole live sequence of
 * dme_run_batch for these jobs -- the pictures' pyramids, byte and
 * quarter-pel planes built from their current luma, pel search, lambda maps,
 * sub-pel refinement, mode decision, DC, the INTRA counts -- across the
 * streams of its tuning setting.  The graph owns the MEData buffers of its
 * jobs (not shared with the context's other entry points) and its own
 * staging.  launch replays it, then does on the host what a graph cannot:
 * jobs whose device lambda threshold was undecided are redone with the
 * host's maps, and outs (may be NULL; else njobs entries) receive the
 * MEData.  Replays read the pictures' current contents: the planes of a
 * picture bound with dme_bind_picture_device may be rewritten between
 * replays.  Re-registering or dropping a picture of the graph makes launch
 * fail with DME_ERR_STATE; the graph must then be created again.  Results
 * equal dme_run_batch's bit for bit.                                         */
typedef struct dme_batch_graph dme_batch_graph;
int dme_batch_graph_create(dme_ctx *ctx, const dme_job *jobs, int njobs,
                           dme_batch_graph **out);
int dme_batch_graph_launch(dme_batch_graph *g, dme_medata *outs);
int dme_batch_graph_destroy(dme_batch_graph *g);

/* ---- no-allocation motion estimation ----------------------------------- *
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation.  dme_motion_estimate for one picture whose planes are
 * given in device memory of the context's device (luma and chroma of the
 * picture, luma of its nrefs references, row-major at stride = width); the
 * pictures need not be registered.  Every intermediate buffer (pyramids, byte
 * and quarter-pel planes, search records, the MEData under estimation) lives
 * in the library's static device scratch array, DME_NOALLOC_SCRATCH_BYTES per
 * GPU, held from library load until the process ends.  The call allocates no
 * device or pinned host memory (the context reserved its small pinned
 * staging at dme_create).  A picture whose buffers exceed the scratch gives
 * DME_ERR_INVALID and names the size it needs (1920x1080 with 2 references
 * and 16/12 blocks needs about 90 MB).  pic_num and ref_nums are picture
 * numbers (full search's temporal distance).  out's non-NULL arrays may be
 * host or device memory; the call returns when they hold the results.
 * Calls on one GPU share its scratch array, which keeps the last call's
 * contents: calls that overlap on one GPU (two contexts, two threads) race. */
#define DME_NOALLOC_SCRATCH_BYTES (128u << 20)
int dme_motion_estimate_noalloc(dme_ctx *ctx, const dme_params *params,
                                int width, int height, int pic_num,
                                const int16_t *d_y, const int16_t *d_u,
                                const int16_t *d_v, int nrefs,
                                const int *ref_nums,
                                const int16_t *const *d_ref_y,
                                dme_medata *out);

/* ---- motion compensation ------------------------------------------------ */

/* AddOrSub (common.h:90) */
#define DME_MC_ADD       0   /* residual + prediction (decoder side)       */
#define DME_MC_SUBTRACT  1   /* picture - prediction (encoder residual)    */

/* One component array of the picture being compensated: row-major int16,
 * `width` samples per row and `height` rows, at least the component's own
 * size (an encoder picture may be padded beyond it).                      */
typedef struct dme_mc_plane {
    int16_t *data;
    int32_t  width;
    int32_t  height;
} dme_mc_plane;

/* MotionCompensator::CompensatePicture (mot_comp.cpp:63-99, OBMC per
 * CompensateComponent/CompensateBlock :238-603): compensates the Y, U and V
 * arrays of a width x height 4:2:0 picture in place, in `direction`, from
 * the MvData in `mv` (vectors[0..nrefs), mode, dc[3], sb_split; the layouts
 * dme_motion_estimate writes) and the registered pictures refs[0..nrefs)
 * (luma and chroma planes required) as references, with the block
 * parameters, vector precision (0..3: pel, half, quarter, eighth) and float
 * semantics (the MMX build's 16-bit rounding when adding) of `params`, and
 * default reference weights.  Padded samples are handled as the reference
 * does: zeroed when subtracting, replicated from the last column and row
 * when adding.  planes[] are host arrays; the call returns when they hold
 * the result.  The references' up-converted planes (Picture::UpData,
 * picture.cpp:141-166) are built on first use and kept.                   */
int dme_motion_compensate(dme_ctx *ctx, const dme_params *params, int width,
                          int height, const dme_medata *mv, const int *refs,
                          int nrefs, int direction,
                          const dme_mc_plane planes[3]);

/* Batched, device-resident form: compensates njobs pictures in one launch,
 * each as dme_motion_compensate does, with planes[].data in device memory of
 * the context's device (compensated in place; arrays of different jobs must
 * not overlap).  mv == NULL takes the MvData the context estimated for
 * pic_num (dme_motion_estimate / dme_run_batch / the learned forms, which
 * keep it on the device): motion estimation followed by compensation then
 * never leaves the GPU.  params.mv_precision may be 3 (eighth pel).  The call
 * returns when the arrays hold the results.  No counterpart in the reference,
 * which compensates one picture at a time (PictureCompressor::
 * MotionCompensate, picture_compress.cpp:322-338).                         */
typedef struct dme_mc_job {
    int32_t pic_num;        /* MvData source when mv is NULL                */
    int32_t width, height;  /* picture size (luma)                           */
    int32_t nrefs;
    int32_t refs[2];        /* registered reference pictures                */
    int32_t direction;      /* DME_MC_ADD or DME_MC_SUBTRACT                */
    int32_t reserved;       /* must be zero                                  */
    dme_params params;
    const dme_medata *mv;   /* host MvData, or NULL (see above)             */
    dme_mc_plane planes[3]; /* device arrays                                 */
} dme_mc_job;
int dme_motion_compensate_batch(dme_ctx *ctx, const dme_mc_job *jobs,
                                int njobs);

/* Diagnostics: copies a cached device plane of picture pic_num to host.
 * kind 0 = pyramid level `level` (1..depth built so far), (W>>level) x
 * (H>>level) samples (PixelMatcher::MakePicHierarchy's down_data,
 * pixel_match.cpp:159-183, which the reference keeps private);
 * kind 1 = the 2W x 2H up-converted plane (EncPicture::UpOrigData,
 * enc_picture.cpp:131-156), built once the picture served as a reference.
 * kind 2 = the byte up-plane the learned path's fast form reads, built once
 * the picture served it as a reference: the up-converted samples of up rows
 * -64 .. R - 65 and columns -64 .. P - 65 (R x P samples, R = 32 *
 * ceil((H + 64) / 16), P = 128 * ceil((W + 64) / 64)), each the sample at
 * the clamped coordinates (the reference's bounds-checked reads).
 * kinds 3, 4, 5 = the combined-component picture (W x H; EncPicture::
 * CombinedData, enc_picture.cpp:237-254), its pyramid level `level` and its
 * 2W x 2H up-converted plane (UpCombinedData, :256-280), built once a
 * combined-component job read the picture.
 * out_elems must be at least the plane's sample count.                    */
int dme_get_plane(dme_ctx *ctx, int pic_num, int kind, int level,
                  int16_t *out, size_t out_elems);

/* Waits for all work queued on the context's stream. */
int dme_synchronize(dme_ctx *ctx);

/* Marks the cached pyramid levels and up-converted plane of pic_num stale
 * (its samples changed in place, e.g. a recycled device frame buffer); they
 * are rebuilt into the same buffers on next use.  pic_num < 0: all pictures. */
int dme_invalidate_caches(dme_ctx *ctx, int pic_num);

/* Counters of the last completed pipeline call on this context.  Per-kernel
 * times are recorded only while profiling is enabled (events around every
 * launch on the context's stream).                                          */
#define DME_NUM_KERNELS 32
typedef struct dme_stats {
    double  sad_terms;        /* algorithmic SAD terms of the last call       */
    double  kernel_ms;        /* stream time of the last call (events)        */
    int32_t launches;         /* kernel launches of the last call             */
    int32_t host_lambda_jobs; /* jobs of the last call whose lambda maps the
                                 host computed (device threshold undecided) */
    int32_t reserved[6];
    double  kernel_ms_by[DME_NUM_KERNELS];  /* summed time per kernel kind   */
    int32_t launches_by[DME_NUM_KERNELS];   /* launches per kernel kind      */
} dme_stats;
int dme_get_stats(dme_ctx *ctx, dme_stats *out);
int dme_set_profiling(dme_ctx *ctx, int enable);
/* Name of kernel kind k (0 <= k < DME_NUM_KERNELS), "" if unused. */
const char *dme_kernel_name(int k);

/* ---- synthetic input generator (deterministic, own PRNG) --------------- */

/* Texture of Gaussian-filtered uniform noise plus flat-with-noise rectangles,
 * 8-bit values stored minus 128.  Writes width*height luma samples. */
int dme_synth_texture(uint64_t seed, int width, int height, int nrects,
                      int16_t *out_y);

/* Motion field description for dme_synth_warp.  Vectors are in pel units
 * (fractions allowed, quantised to quarter pel by the warp).                 */
typedef struct dme_synth_object {
    float x0, y0, x1, y1;   /* object rectangle in the *current* picture     */
    float dx, dy;           /* displacement of the object                    */
} dme_synth_object;
typedef struct dme_synth_motion {
    float zoom;             /* affine zoom about the picture centre (1 = none) */
    float pan_x, pan_y;     /* global translation                            */
    int32_t nobjects;
    const dme_synth_object *objects;
    float noise_sigma;      /* additive Gaussian noise after warping         */
    uint64_t noise_seed;
} dme_synth_motion;

/* ref(x,y) = cur(x + d(x,y)) sampled bilinearly on the quarter-pel grid with
 * clamp-to-edge, plus noise; d is the motion field.  8-bit range, minus 128. */
int dme_synth_warp(const int16_t *cur, int width, int height,
                   const dme_synth_motion *motion, int16_t *out_ref);

/* 4:2:0 chroma planes derived from a luma plane (2x2 mean plus a seeded
 * offset), width/2 x height/2 each.                                          */
int dme_synth_chroma(const int16_t *y, int width, int height, uint64_t seed,
                     int16_t *out_u, int16_t *out_v);

#ifdef __cplusplus
}
#endif

#endif /* DME_H */
