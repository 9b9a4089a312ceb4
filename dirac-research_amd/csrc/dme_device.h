// dme_device.h -- device-side job descriptors shared by the kernels
// (dme_kernels.hip, dme_learned.hip) and the host orchestration
// (dme_engine.cpp).
#ifndef DME_DEVICE_H
#define DME_DEVICE_H

#include <cstdint>

#include "../../include/dme.h"

// Device code reads every descriptor pointer as a global-address-space
// pointer, so its loads are global (not flat) loads, which need no ordering
// against LDS accesses.  Same 64-bit representation on host and device.
#if defined(__HIP_DEVICE_COMPILE__)
#define DME_G __attribute__((address_space(1)))
#else
#define DME_G
#endif
// helpers shared by host and device code
#if defined(__HIPCC__)
#define DME_HD __host__ __device__
#else
#define DME_HD
#endif

namespace dme {

// Everything one picture's motion estimation needs on the device.  Built on
// the host, copied to device memory once per call, read by every kernel.
struct JobDesc {
    // geometry (host-validated)
    int W, H, CW, CH;
    int xnb, ynb, xsb, ysb, depth, nrefs;
    int prec;      // MVPrecision of the refinement (0, 1, 2)
    int step;      // sampling precision of mode decision (1 half, 2 quarter)
    int fp_sem;    // DME_FP_MMX or DME_FP_C
    int lw[DME_MAX_DEPTH + 1], lh[DME_MAX_DEPTH + 1];
    int lxb[DME_MAX_DEPTH + 1], lyb[DME_MAX_DEPTH + 1];
    int blen[3][2], bsep[3][2], boff[3][2];
    int cblen[2], cbsep[2], cboff[2];
    float lf[3];
    // planes
    const int16_t DME_G *pic[DME_MAX_DEPTH + 1];      // [0] = Y of the picture
    const int16_t DME_G *ref[2][DME_MAX_DEPTH + 1];   // reference pyramids
    const uint32_t DME_G *qp[2];                      // quarter-pel planes of the refs (k_build_qp); with comb,
                                                      // the refs' int16 up planes (2W x 2H) instead
    const uint8_t DME_G *pic8, *ref8[2];              // byte planes of picture and refs (k_bytes): sample (0, 0)
    const uint8_t DME_G *up8[2];                      // byte up-planes of the refs (fast form): up sample (0, 0)
    int u8p;                                          // their row pitch in bytes
    int p8;                                           // their row pitch in bytes
    const int16_t DME_G *cu, *cv;                     // chroma of the picture
    const int16_t DME_G *py;                          // luma of the picture (SetDC reads it; pic[0] may be
                                                      // the combined-component picture)
    int comb;                                         // combined-component ME: pic and ref are the combined
                                                      // pictures' pyramids and qp[r] their int16 up planes
    int lp_off;                                       // first position of the picture in the proposal tensors
    int lp_slot;                                      // its pair of refinement counters (k_lp_refine)
    int full;                                         // full search: depth 0, zero window of fxr x fyr (k_pel_full)
    int fxr[2], fyr[2];                               // its radius per reference, min(tdiff, 3) * range
    // pel search, per reference and level ([0] aliases the MEData arrays)
    dme_mv DME_G *lvec[2][DME_MAX_DEPTH + 1];
    dme_cost DME_G *lpc[2][DME_MAX_DEPTH + 1];
    int DME_G *ptab[2];                               // pel block records, wavefront order (k_pel_bulk)
    int DME_G *qtab[2];                               // sub-pel block records, wavefront order (k_subpel_bulk)
    int DME_G *m1tab;                                 // level-1 unit SAD tables (k_mode_tables)
    int DME_G *m0tab;                                 // level-0 unit SAD tables
    // MEData of the picture
    dme_mv DME_G *vec[2];
    dme_cost DME_G *pc[2];
    int32_t DME_G *mode;
    float DME_G *intra;
    dme_cost DME_G *bi;
    int16_t DME_G *dc[3];
    int32_t DME_G *split;
    float DME_G *sbc;
    float DME_G *lam;
    int DME_G *icount;                                // INTRA blocks of the final modes (k_intra_count)
    float lambda;                                     // the picture's lambda (SetLambdaMap)
    int DME_G *lwork;                                 // 2 x blocks: transition maps (k_lambda)
    // mode decision temporaries (levels 1 and 0)
    float DME_G *lam1, *lam0;
    float DME_G *ucost;                               // level-2 unit costs
    // the fast form of the learned path (k_fast_*): per reference the
    // level-0 integer vectors, the pre-merge sub-pel vectors and their SADs;
    // per superblock the SADs of the chosen units (FSB_INTS each)
    dme_mv DME_G *fv0[2], *fv2[2];
    int DME_G *fs2[2];
    int DME_G *fsb;
    int16_t DME_G *dct[3];                            // block DCs before smoothing
};

// Replicated samples on every side of a byte plane (k_bytes).
constexpr int kB8Margin = 32;

// Byte up-planes of the fast form (k_build_qp<true>): the 2W x 2H
// up-converted reference as biased bytes with a replicated margin of
// kUp8Margin samples; made by tiles of 64 x 16 positions (128 x 32 up
// samples), so the plane is whole tiles across and down.
constexpr int kUp8Margin = 64;
DME_HD inline int up8_tiles_x(int W) { return (W + kUp8Margin + 63) / 64; }
DME_HD inline int up8_tiles_y(int H) { return (H + kUp8Margin + 15) / 16; }
DME_HD inline int up8_pitch(int W) { return 128 * up8_tiles_x(W); }
DME_HD inline int up8_rows(int H) { return 32 * up8_tiles_y(H); }

// Per superblock of the fast form: for each of up to four units of its
// split level, the integer SADs of its two vectors and of their
// bi-prediction (k_fast_mode), read by the exact-cost pass (k_fast_eval).
constexpr int kFsbInts = 12;

// A plane conversion of a batched pyramid / up-conversion launch: `in` of
// width in_w, `out` of ow x oh samples (the source size for up-conversion),
// covering tiles [offset, offset + tiles) of the launch.
struct PlaneTask {
    const int16_t DME_G *in;
    int16_t DME_G *out;
    int in_w, ow, oh, offset;
};

// The combined-component picture of a 4:2:0 picture (k_combine): W x H luma,
// W/2 x H/2 chroma, covering 256-sample tiles [offset, offset + tiles).
struct CombTask {
    const int16_t DME_G *y, *u, *v;
    int16_t DME_G *out;
    int W, H, offset;
};
inline int combine_tiles(int W, int H) { return (int)(((size_t)W * H + 255) / 256); }

// k_upconvert tiles: 64 x 16 source samples.
constexpr int kUpTileW = 64, kUpTileH = 16;
inline int upconvert_tiles(int w, int h) { return ((w + kUpTileW - 1) / kUpTileW) * ((h + kUpTileH - 1) / kUpTileH); }

// Output tile of k_downconvert (a workgroup filters kDownTileW x kDownTileH
// samples of the coarser level).
constexpr int kDownTileW = 64, kDownTileH = 16;

// Rows of the column segment one k_mc thread owns; a k_mc tile is 64 columns
// by 4 segments (k_mc_pad tiles are 64 x 4).
constexpr int kMcSegRows = 8;

// One component of one picture being motion compensated (k_mc / k_mc_pad).
// Built on the host, read by every workgroup of the component's tiles.
struct McComp {
    const uint8_t DME_G *up[2];     // up-converted reference planes, 2pw x 2ph bytes (samples + 128)
    int16_t DME_G *out;             // array compensated in place, lx x ly
    const int4 DME_G *unit;         // per block: {v1.x, v1.y, v2.x, v2.y}, {mode, dc Y, dc U, dc V} (k_mc_pack)
    const int DME_G *split;         // xsb x ysb superblocks
    const signed char DME_G *hw;    // OBMC weights across: [split][class][hpitch]
    const int DME_G *vw;            // OBMC weights down: [class][yblen]
    int lx, ly, pw, ph;
    int xbsep, ybsep, xoff, yoff, xov, yblen, hpitch;
    int nbx, nby;                   // blocks in use across / down (NUM_USED_BLKS)
    int xnb, xsb;
    int comp;                       // 0 Y, 1 U, 2 V
    int shift;                      // vector shift (1 for 4:2:0 chroma)
    int prec;                       // 0 pel, 1 half, 2 quarter, 3 eighth
    int direction;                  // 0 add, 1 subtract
    int mmx;                        // 16-bit rounding of the MMX add path
    int tiles_x, offset;            // tiles across; first tile in the launch
};

// MvData of one picture to pack into k_mc's per-block records.
struct McPackJob {
    const dme_mv DME_G *vec[2];
    const int DME_G *mode;
    const int16_t DME_G *dc[3];
    int4 DME_G *out;  // 2 records per block
    int blocks, offset;
};

// Kinds of job a launch of the sub-pel and mode-decision kernels covers
// (a bit mask): their eighth-pel forms are separate instantiations, so that
// the others keep their register budget.
constexpr int DME_KIND_BELOW_EIGHTH = 1, DME_KIND_EIGHTH = 2;
// and the combined-component jobs (int16 up planes) of each
constexpr int DME_KIND_COMB_BELOW_EIGHTH = 4, DME_KIND_COMB_EIGHTH = 8;

// One unit of a batched launch: a (job, reference) pair covering `count`
// consecutive work items starting at `offset` of the launch's index space.
// Refinement counters of k_lp_refine: per batch position (slot)
// kLpStatCopies copies of the pair {blocks, fallbacks}, each on a 128-byte
// line of its own, a wave adding to copy (its index mod kLpStatCopies), so
// that the atomic adds of a batch spread over many lines.
constexpr int kLpStatCopies = 64;
constexpr int kLpStatStride = 16;  // unsigned long long per copy
constexpr int kLpStatSlot = kLpStatCopies * kLpStatStride;

struct Task {
    int job, ref, count, offset;
};

// One array of a batch's MEData export: `bytes` (a multiple of 2) from src
// to byte `off` (a multiple of 16) of the packed export buffer.
struct PackSeg {
    const void *src;
    unsigned long long off, bytes;
};

}  // namespace dme

#endif
